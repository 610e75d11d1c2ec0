// tq_adjoint.hip -- differentiation of recorded frames (tq_tape_begin .. tq_backward).
//
// Recording.  While a tape is open the two operations a run is made of, the faucet
// (tq_emit_faucet, principal.js:301-321) and the frame step (tq_step, fisica-cpu.js:155-159),
// are logged; every `every` steps the whole device state of the tank (the buffers alloc_all
// made) and its host fields are copied to a checkpoint.  With every = 0 each substep's record is
// kept as the frames run instead.
//
// The reverse pass.  tq_backward restores each checkpoint from the last to the first, re-runs
// its frames with substep records open (the substep's hooks copy the particles at P2G, the
// P2G velocities and weights, the density, the types and bubble labels, the post-extrapolation
// velocities and G2P's per-particle decision words), then walks the records backwards:
//   k_adj_g2p      G2P of one particle re-run from its record and reversed: clamp, micro-steps
//                  with grid velocity lookups, SDF push-out and velocity projection
//                  (fisica-cpu.js:560-640, cena.js:102-117); face adjoints and SDF adjoints by
//                  FP64 atomics
//   k_adj_extrap   the two extrapolation layers reversed (fisica-cpu.js:660-678)
//   k_adj_gradq    the gradient (fisica-cpu.js:522-549) reversed to pressure adjoints; bubble
//                  cells' pressure adjoints summed per bubble
//   pressure_solve the adjoint solve: the same MGPCG on the same operator (A is symmetric), its
//                  inputs (types, labels, bubble face and cell counts) put back from the record
//   k_adj_faces    the divergence and bubble face sums (fisica-cpu.js:430-483) reversed, then
//                  gravity/walls and the P2G normalisation u = sum(w v)/sum(w)
//   k_adj_dens     the density-excess term (fisica-cpu.js:465-470) reversed
//   k_ten_*        surface tension (fisica-cpu.js:306-403) re-run from the record's density
//                  and reversed: the air pressures' adjoints (from the air cells' q, the water
//                  cells' divergence and the bubbles' face terms) back through the clamp, the
//                  curvature, the normals, the blurs and wall extensions to the density
//   k_adj_p2g      P2G of one particle (fisica-cpu.js:198-267) reversed, with its density splat
//   k_adj_emit     the emitted particles' velocity adjoints summed per faucet call
// Every discrete decision (types, labels, micro-step counts, push-outs, projections, removals)
// is the one the recorded run took: the result is the derivative of the run with those
// decisions held, which is smooth in positions, velocities, the jet velocity and the SDF
// (surface tension included; the bubbles' volume-return terms are functions of the labels).
// Particle adjoints travel between substeps indexed by insertion number (seq), since the
// counting sort reorders particles every substep.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "tq_device.hpp"
#include "tq_internal.hpp"

namespace tq {

constexpr int FL_COLL = 2, FL_PROJ = 6;
constexpr uint32_t FL_KILL = 1u << 10;

struct SubRec {
  double dt = 0;
  int cap = 0;                   // particle capacity of the buffers below
  int n_up = 0;                  // host upper bound of the particle count
  float2 *pos = nullptr, *vel = nullptr;
  float4* cm = nullptr;
  uint32_t* seq = nullptr;
  uint16_t* flags = nullptr;
  uint32_t* n_dev = nullptr;     // particle count (slots after the sort)
  float *u1 = nullptr, *v1 = nullptr, *uw = nullptr, *vw = nullptr, *dens = nullptr;
  float *u3 = nullptr, *v3 = nullptr, *qAr = nullptr;
  uint8_t* tipo = nullptr;
  uint32_t *rot = nullptr, *facesB = nullptr, *cellsB = nullptr, *volB = nullptr;
  float *volAlvo = nullptr, *yB = nullptr;
  int emit = -1;                 // faucet call whose particles entered just before this substep
};

struct Snapshot {
  std::vector<void*> dev;
  int cur = 0, n_upper = 0;
  uint32_t seq_next = 0;
  bool tipos_prontos = false, has_vmax = false, air_reset_pending = false;
  DevScalars sc{};
};

struct TapeOp {
  int kind = 0;                  // 0 faucet, 1 step
  tq_faucet f{};
  double dt = 0;
  uint32_t seed = 0;             // faucet: the seed before the call
  int substeps = 0;
  uint32_t seq0 = 0, seq1 = 0;   // faucet: insertion numbers of the accepted particles
  int emit = -1;                 // faucet: its index among the recorded faucet calls
};

struct Tape {
  int every = 0;
  std::vector<TapeOp> ops;
  std::vector<std::pair<size_t, Snapshot*>> cps;  // (first op of the segment, state before it)
  int steps = 0, n_emit = 0;
  int pending_emit = -1;
  std::vector<uint32_t> seq_begin;  // live particles' insertion numbers at the start, ascending
  std::vector<SubRec*> recs;
  size_t recs_used = 0;
  bool replaying = false;
  // reverse-pass scratch
  int seq_cap = 0;
  double* adjS = nullptr;        // 8 doubles per insertion number
  int xg_cap = 0;
  double2* xg = nullptr;         // G2P's start-position adjoint of each particle of a record
  double *ub = nullptr, *vb = nullptr, *ub2 = nullptr, *vb2 = nullptr;
  double *ua = nullptr, *uwb = nullptr, *va = nullptr, *vwb = nullptr, *db = nullptr;
  double *ybar = nullptr, *gsdf = nullptr, *gemit = nullptr;
  double *qarb = nullptr;        // air-pressure adjoint per cell
  double* tf = nullptr;          // surface tension: the water fraction before each of its 13 steps, and after
  uint8_t* tfull = nullptr;      // the filled cells before each of the 9 extension layers
  double *tnx = nullptr, *tny = nullptr, *tg = nullptr, *tcoef = nullptr;
  double *tkb = nullptr, *tnxb = nullptr, *tnyb = nullptr, *tfb = nullptr, *tfb2 = nullptr;
  uint8_t *val0u = nullptr, *val1u = nullptr, *val0v = nullptr, *val1v = nullptr;
};

// ------------------------------------------------------------------ device helpers
struct DBil {
  double v, dfx, dfy;
  int k;
  double w0, w1, w2, w3;
};
// Cena #bilinear (cena.js:156-165) with its derivatives and node weights
__device__ inline DBil dbil(const float* __restrict__ sdf, int w, int hh, double fx, double fy) {
#pragma clang fp contract(off)
  DBil b;
  bool cx = false, cy = false;
  if (fx != fx) fx = 0;
  if (fy != fy) fy = 0;
  if (fx < 0) { fx = 0; cx = true; } else if (fx > w - 1) { fx = w - 1; cx = true; }
  if (fy < 0) { fy = 0; cy = true; } else if (fy > hh - 1) { fy = hh - 1; cy = true; }
  const int x0 = min((int)floor(fx), w - 2), y0 = min((int)floor(fy), hh - 2);
  const double tx = fx - x0, ty = fy - y0;
  b.k = y0 * w + x0;
  const double d00 = sdf[b.k], d10 = sdf[b.k + 1], d01 = sdf[b.k + w], d11 = sdf[b.k + w + 1];
  b.v = (d00 * (1 - tx) + d10 * tx) * (1 - ty) + (d01 * (1 - tx) + d11 * tx) * ty;
  b.dfx = cx ? 0.0 : (d10 - d00) * (1 - ty) + (d11 - d01) * ty;
  b.dfy = cy ? 0.0 : (d01 * (1 - tx) + d11 * tx) - (d00 * (1 - tx) + d10 * tx);
  b.w0 = (1 - tx) * (1 - ty); b.w1 = tx * (1 - ty); b.w2 = (1 - tx) * ty; b.w3 = tx * ty;
  return b;
}
__device__ inline void dbil_adj(double* __restrict__ gsdf, int w, const DBil& b, double vbar) {
  if (vbar == 0) return;
  atomicAdd(&gsdf[b.k], vbar * b.w0);
  atomicAdd(&gsdf[b.k + 1], vbar * b.w1);
  atomicAdd(&gsdf[b.k + w], vbar * b.w2);
  atomicAdd(&gsdf[b.k + w + 1], vbar * b.w3);
}

struct DWall {
  DBil d, xp, xm, yp, ym;
  double g, nx, ny;
  bool deg;
};
// Cena.distancia and Cena.normal (cena.js:102-117) at (x, y) metres
__device__ inline DWall dwall(const float* __restrict__ sdf, int w, int hh, double h, double x, double y) {
#pragma clang fp contract(off)
  DWall W;
  const double fx = (x / h) * 2 - 0.5, fy = (y / h) * 2 - 0.5;
  W.d = dbil(sdf, w, hh, fx, fy);
  W.xp = dbil(sdf, w, hh, fx + 1, fy);
  W.xm = dbil(sdf, w, hh, fx - 1, fy);
  W.yp = dbil(sdf, w, hh, fx, fy + 1);
  W.ym = dbil(sdf, w, hh, fx, fy - 1);
  const double gx = W.xp.v - W.xm.v, gy = W.yp.v - W.ym.v;
  W.g = js_hypot(gx, gy);
  W.deg = !(W.g > 1e-9);
  if (W.deg) { W.nx = 0; W.ny = 1; } else { W.nx = gx / W.g; W.ny = gy / W.g; }
  return W;
}
__device__ inline void dwall_adj(double* __restrict__ gsdf, int w, double h, const DWall& W, double dbar, double nxbar,
                                 double nybar, double* xbar, double* ybar) {
#pragma clang fp contract(off)
  double fxbar = dbar * W.d.dfx, fybar = dbar * W.d.dfy;
  dbil_adj(gsdf, w, W.d, dbar);
  if (!W.deg) {
    const double dot = nxbar * W.nx + nybar * W.ny;
    const double gxbar = (nxbar - dot * W.nx) / W.g, gybar = (nybar - dot * W.ny) / W.g;
    fxbar += gxbar * (W.xp.dfx - W.xm.dfx) + gybar * (W.yp.dfx - W.ym.dfx);
    fybar += gxbar * (W.xp.dfy - W.xm.dfy) + gybar * (W.yp.dfy - W.ym.dfy);
    dbil_adj(gsdf, w, W.xp, gxbar);
    dbil_adj(gsdf, w, W.xm, -gxbar);
    dbil_adj(gsdf, w, W.yp, gybar);
    dbil_adj(gsdf, w, W.ym, -gybar);
  }
  *xbar += fxbar * 2 / h;
  *ybar += fybar * 2 / h;
}

struct DSten {
  int i0, j0;
  double tx, ty;
};
// stencils of the u faces, v faces and cell centres (fisica-cpu.js:211-250)
__device__ inline DSten dsten(double fx, double fy, int imax, int jmax) {
  DSten s;
  if (fx != fx) fx = 0;
  if (fy != fy) fy = 0;
  fx = fmin(fmax(fx, -1.0), (double)imax + 2);
  fy = fmin(fmax(fy, -1.0), (double)jmax + 2);
  s.i0 = min(max((int)floor(fx), 0), imax);
  s.j0 = min(max((int)floor(fy), 0), jmax);
  s.tx = fx - s.i0;
  s.ty = fy - s.j0;
  return s;
}
__device__ inline DSten sten_u(int nx, int ny, double inv, double x, double y) { return dsten(x * inv, y * inv - 0.5, nx - 1, ny - 2); }
__device__ inline DSten sten_v(int nx, int ny, double inv, double x, double y) { return dsten(x * inv - 0.5, y * inv, nx - 2, ny - 1); }
__device__ inline DSten sten_c(int nx, int ny, double inv, double x, double y) { return dsten(x * inv - 0.5, y * inv - 0.5, nx - 2, ny - 2); }
__device__ inline double wa(double t, int a) { return a ? t : 1 - t; }
__device__ inline double dwa(int a) { return a ? 1.0 : -1.0; }

// value and the two affine terms of G2P (fisica-cpu.js:571-599) of face field f (row width w)
__device__ inline void dinterp(const float* __restrict__ f, int w, const DSten& s, double inv, double* v, double* cx,
                               double* cy) {
#pragma clang fp contract(off)
  double a0 = 0, a1 = 0, a2 = 0;
  for (int b = 0; b < 2; b++)
    for (int a = 0; a < 2; a++) {
      const double val = f[(s.i0 + a) + (s.j0 + b) * w];
      const double wx = wa(s.tx, a), wy = wa(s.ty, b);
      a0 += wx * wy * val;
      a1 += dwa(a) * wy * inv * val;
      a2 += wx * dwa(b) * inv * val;
    }
  *v = a0; *cx = a1; *cy = a2;
}
__device__ inline void dinterp_adj(const float* __restrict__ f, int w, const DSten& s, double inv, double vbar, double cxbar,
                                   double cybar, double* __restrict__ fbar, double* xbar, double* ybar) {
#pragma clang fp contract(off)
  for (int b = 0; b < 2; b++)
    for (int a = 0; a < 2; a++) {
      const int k = (s.i0 + a) + (s.j0 + b) * w;
      const double val = f[k];
      const double wx = wa(s.tx, a), wy = wa(s.ty, b);
      const double g = vbar * wx * wy + cxbar * dwa(a) * wy * inv + cybar * wx * dwa(b) * inv;
      if (g != 0) atomicAdd(&fbar[k], g);
      *xbar += (vbar * dwa(a) * wy * val + cybar * dwa(a) * dwa(b) * inv * val) * inv;
      *ybar += (vbar * wx * dwa(b) * val + cxbar * dwa(a) * dwa(b) * inv * val) * inv;
    }
}

// ------------------------------------------------------------------ reverse kernels
// G2P of one particle re-run from its record, then reversed (fisica-cpu.js:560-640)
__global__ __launch_bounds__(256) void k_adj_g2p(const float2* __restrict__ pos, const uint32_t* __restrict__ seq,
                                                 const uint16_t* __restrict__ flags, const uint32_t* __restrict__ n_dev,
                                                 const float* __restrict__ u, const float* __restrict__ v,
                                                 const float* __restrict__ sdf, int nx, int ny, double h, double dt,
                                                 const double* __restrict__ adjS, double* __restrict__ ub,
                                                 double* __restrict__ vb, double* __restrict__ gsdf,
                                                 double2* __restrict__ xg) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *n_dev) return;
  const uint32_t fl = flags[i];
  if (fl & FL_KILL) { xg[i] = make_double2(0, 0); return; }
  const double* ob = adjS + 8 * (size_t)seq[i];
  const int w = 2 * nx, hh = 2 * ny;
  const double inv = 1 / h;
  const double vmax = (MICROPASSO * MAX_MICROPASSOS * h) / dt;
  double x = pos[i].x, y = pos[i].y;
  const DSten su = sten_u(nx, ny, inv, x, y), sv = sten_v(nx, ny, inv, x, y);
  double vx, cux, cuy, vy, cvx, cvy;
  dinterp(u, nx + 1, su, inv, &vx, &cux, &cuy);
  dinterp(v, nx, sv, inv, &vy, &cvx, &cvy);
  const double ivx = vx, ivy = vy;
  const double vel = js_hypot(vx, vy);
  const bool clamp = vel > vmax;
  if (clamp) { vx *= vmax / vel; vy *= vmax / vel; }
  const int m = (int)(fl & 3) + 1;
  const double dm = dt / m;
  double X0[MAX_MICROPASSOS], Y0[MAX_MICROPASSOS], X1[MAX_MICROPASSOS], Y1[MAX_MICROPASSOS];
  double VX0[MAX_MICROPASSOS], VY0[MAX_MICROPASSOS];
  for (int k = 0; k < m; k++) {
    X0[k] = x; Y0[k] = y;
    double wx = vx, wy = vy;
    if (m > 1) {
      const DSten gu = sten_u(nx, ny, inv, x, y), gv = sten_v(nx, ny, inv, x, y);
      double t0, t1;
      dinterp(u, nx + 1, gu, inv, &wx, &t0, &t1);
      dinterp(v, nx, gv, inv, &wy, &t0, &t1);
      const double g = js_hypot(wx, wy);
      if (g > vmax) { wx = wx * vmax / g; wy = wy * vmax / g; }
    }
    x += wx * dm;
    y += wy * dm;
    X1[k] = x; Y1[k] = y;
    VX0[k] = vx; VY0[k] = vy;
    if ((fl >> (FL_COLL + k)) & 1) {
      const DWall W = dwall(sdf, w, hh, h, x, y);
      const double push = (RAIO_PARTICULA - W.d.v) * h;
      x += W.nx * push;
      y += W.ny * push;
      if ((fl >> (FL_PROJ + k)) & 1) {
        const double vn = vx * W.nx + vy * W.ny;
        vx -= vn * W.nx;
        vy -= vn * W.ny;
      }
    }
  }
  // reverse
  double xbar = ob[0], ybar = ob[1], vxbar = ob[2], vybar = ob[3];
  for (int k = m - 1; k >= 0; k--) {
    if ((fl >> (FL_COLL + k)) & 1) {
      const DWall W = dwall(sdf, w, hh, h, X1[k], Y1[k]);
      double nxbar = 0, nybar = 0;
      if ((fl >> (FL_PROJ + k)) & 1) {
        const double vn = VX0[k] * W.nx + VY0[k] * W.ny;
        const double bn = vxbar * W.nx + vybar * W.ny;
        nxbar -= vn * vxbar + bn * VX0[k];
        nybar -= vn * vybar + bn * VY0[k];
        vxbar -= bn * W.nx;
        vybar -= bn * W.ny;
      }
      const double push = (RAIO_PARTICULA - W.d.v) * h;
      nxbar += xbar * push;
      nybar += ybar * push;
      const double dbar = -h * (xbar * W.nx + ybar * W.ny);
      dwall_adj(gsdf, w, h, W, dbar, nxbar, nybar, &xbar, &ybar);
    }
    const double wxbar = xbar * dm, wybar = ybar * dm;
    if (m > 1) {
      const DSten gu = sten_u(nx, ny, inv, X0[k], Y0[k]), gv = sten_v(nx, ny, inv, X0[k], Y0[k]);
      double gx, gy, t0, t1;
      dinterp(u, nx + 1, gu, inv, &gx, &t0, &t1);
      dinterp(v, nx, gv, inv, &gy, &t0, &t1);
      double gxbar = wxbar, gybar = wybar;
      const double g = js_hypot(gx, gy);
      if (g > vmax) {
        const double ux = gx / g, uy = gy / g;
        const double dot = wxbar * ux + wybar * uy;
        gxbar = (vmax / g) * (wxbar - dot * ux);
        gybar = (vmax / g) * (wybar - dot * uy);
      }
      dinterp_adj(u, nx + 1, gu, inv, gxbar, 0, 0, ub, &xbar, &ybar);
      dinterp_adj(v, nx, gv, inv, gybar, 0, 0, vb, &xbar, &ybar);
    } else {
      vxbar += wxbar;
      vybar += wybar;
    }
  }
  if (clamp) {
    const double ux = ivx / vel, uy = ivy / vel;
    const double dot = vxbar * ux + vybar * uy;
    vxbar = (vmax / vel) * (vxbar - dot * ux);
    vybar = (vmax / vel) * (vybar - dot * uy);
  }
  dinterp_adj(u, nx + 1, su, inv, vxbar, ob[4], ob[5], ub, &xbar, &ybar);
  dinterp_adj(v, nx, sv, inv, vybar, ob[6], ob[7], vb, &xbar, &ybar);
  xg[i] = make_double2(xbar, ybar);
}

// gradient classes of a face (fisica-cpu.js:522-549): 0 wall (zeroed), 1 next to water
// (pressure difference), 2 between air cells (kept, invalid)
__device__ inline int u_class(const uint8_t* __restrict__ t, int nx, int i, int j) {
  const uint8_t tl = i > 0 ? t[i - 1 + j * nx] : SOLIDO, tr = i < nx ? t[i + j * nx] : SOLIDO;
  if (tl == SOLIDO || tr == SOLIDO) return 0;
  return (tl == AGUA || tr == AGUA) ? 1 : 2;
}
__device__ inline int v_class(const uint8_t* __restrict__ t, int nx, int ny, int i, int j) {
  const uint8_t tb = j > 0 ? t[i + (j - 1) * nx] : SOLIDO, tt = j < ny ? t[i + j * nx] : SOLIDO;
  if (tb == SOLIDO || tt == SOLIDO) return 0;
  return (tb == AGUA || tt == AGUA) ? 1 : 2;
}

// validity before the first extrapolation layer (set by the gradient) and before the second
__global__ __launch_bounds__(256) void k_adj_valid0(const uint8_t* __restrict__ tipo, int nx, int ny,
                                                    uint8_t* __restrict__ vu, uint8_t* __restrict__ vv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int nu = (nx + 1) * ny, nv = nx * (ny + 1);
  if (k < nu) vu[k] = u_class(tipo, nx, k % (nx + 1), k / (nx + 1)) != 2;
  if (k < nv) vv[k] = v_class(tipo, nx, ny, k % nx, k / nx) != 2;
}
__device__ inline int valid_count(const uint8_t* __restrict__ val, int w, int hh, int i, int j) {
  const int k = i + j * w;
  return (i > 0 && val[k - 1]) + (i < w - 1 && val[k + 1]) + (j > 0 && val[k - w]) + (j < hh - 1 && val[k + w]);
}
__global__ __launch_bounds__(256) void k_adj_valid_next(const uint8_t* __restrict__ val, int w, int hh,
                                                        uint8_t* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= w * hh) return;
  out[k] = val[k] || valid_count(val, w, hh, k % w, k / w) > 0;
}
// one extrapolation layer reversed; val = validity before the layer
__global__ __launch_bounds__(256) void k_adj_extrap(const double* __restrict__ ab, const uint8_t* __restrict__ val, int w,
                                                    int hh, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= w * hh) return;
  const int i = k % w, j = k / w;
  if (!val[k]) {
    out[k] = valid_count(val, w, hh, i, j) > 0 ? 0.0 : ab[k];
    return;
  }
  double s = ab[k];
  const int nb[4] = {k - 1, k + 1, k - w, k + w};
  const bool in[4] = {i > 0, i < w - 1, j > 0, j < hh - 1};
  for (int q = 0; q < 4; q++) {
    if (!in[q]) continue;
    const int m = nb[q];
    if (val[m]) continue;
    const int c = valid_count(val, w, hh, m % w, m / w);
    if (c > 0) s += ab[m] / c;
  }
  out[k] = s;
}

// pressure adjoints from the gradient; the adjoint solve's inputs: div = -qbar on water cells
// (the solver solves A q = -div), per-bubble sums of the bubble cells' qbar
__global__ __launch_bounds__(256) void k_adj_gradq(const uint8_t* __restrict__ tipo, const uint32_t* __restrict__ rot,
                                                   const uint32_t* __restrict__ facesB, int ar, int nx, int ny,
                                                   const double* __restrict__ ub, const double* __restrict__ vb,
                                                   float* __restrict__ div, double* __restrict__ ybar,
                                                   double* __restrict__ qarb) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  double qb = 0;
  if (u_class(tipo, nx, i, j) == 1) qb -= ub[i + j * (nx + 1)];
  if (u_class(tipo, nx, i + 1, j) == 1) qb += ub[i + 1 + j * (nx + 1)];
  if (v_class(tipo, nx, ny, i, j) == 1) qb -= vb[i + j * nx];
  if (v_class(tipo, nx, ny, i, j + 1) == 1) qb += vb[i + (j + 1) * nx];
  const uint8_t t = tipo[c];
  div[c] = t == AGUA ? (float)(-qb) : 0.f;
  qarb[c] = t == AR ? qb : 0.0;  // an air cell's q is qAr + its bubble's y (#pressao)
  if (ar && t == AR && qb != 0) {
    const uint32_t r = rot[c];
    if (r >= 2 && r < (uint32_t)MAX_BOLHAS && facesB[r] > 0) atomicAdd(&ybar[r], qb);
  }
}

__device__ inline bool bubble_cell(const uint8_t* __restrict__ tipo, const uint32_t* __restrict__ rot,
                                   const uint32_t* __restrict__ facesB, int nx, int ny, int i, int j, uint32_t* r) {
  if (i < 1 || j < 1 || i > nx - 2 || j > ny - 2) return false;
  const int c = i + j * nx;
  if (tipo[c] != AR) return false;
  *r = rot[c];
  return *r >= 2 && *r < (uint32_t)MAX_BOLHAS && facesB[*r] > 0;
}

// face adjoints of the divergence and bubble sums, gravity/walls and the P2G normalisation:
// ua = adjoint of sum(w val), uwb = adjoint of sum(w), per face (u faces first, then v faces)
__global__ __launch_bounds__(256) void k_adj_faces(const uint8_t* __restrict__ tipo, const uint32_t* __restrict__ rot,
                                                   const uint32_t* __restrict__ facesB, int ar, int nx, int ny,
                                                   const float* __restrict__ q, const float* __restrict__ yB,
                                                   const double* __restrict__ ub, const double* __restrict__ vb,
                                                   const float* __restrict__ u1, const float* __restrict__ uw,
                                                   const float* __restrict__ v1, const float* __restrict__ vw, double gxdt,
                                                   double gydt, double* __restrict__ ua, double* __restrict__ uwb,
                                                   double* __restrict__ va, double* __restrict__ vwb) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int nu = (nx + 1) * ny, nv = nx * (ny + 1);
  if (k < nu) {
    const int i = k % (nx + 1), j = k / (nx + 1);
    double a = 0, b = 0;
    if (u_class(tipo, nx, i, j) != 0 && uw[k] > 0) {
      double g = ub[k];
      const int L = i - 1 + j * nx, R = i + j * nx;
      if (tipo[L] == AGUA) g -= q[L];
      if (tipo[R] == AGUA) g += q[R];
      uint32_t r;
      if (ar && tipo[L] == AGUA && bubble_cell(tipo, rot, facesB, nx, ny, i, j, &r)) g += yB[r];
      if (ar && tipo[R] == AGUA && bubble_cell(tipo, rot, facesB, nx, ny, i - 1, j, &r)) g -= yB[r];
      const double W = uw[k];
      const double u0 = (double)u1[k] - gxdt;
      a = g / W;
      b = -g * u0 / W;
    }
    ua[k] = a;
    uwb[k] = b;
  }
  if (k < nv) {
    const int i = k % nx, j = k / nx;
    double a = 0, b = 0;
    if (v_class(tipo, nx, ny, i, j) != 0 && vw[k] > 0) {
      double g = vb[k];
      const int B = i + (j - 1) * nx, T = i + j * nx;
      if (tipo[B] == AGUA) g -= q[B];
      if (tipo[T] == AGUA) g += q[T];
      uint32_t r;
      if (ar && tipo[B] == AGUA && bubble_cell(tipo, rot, facesB, nx, ny, i, j, &r)) g += yB[r];
      if (ar && tipo[T] == AGUA && bubble_cell(tipo, rot, facesB, nx, ny, i, j - 1, &r)) g -= yB[r];
      const double W = vw[k];
      const double v0 = (double)v1[k] - gydt;
      a = g / W;
      b = -g * v0 / W;
    }
    va[k] = a;
    vwb[k] = b;
  }
}

// density adjoint of the excess term: div has -alvo * excess(dens), so dens-bar = alvo * d(excess) * q
__global__ __launch_bounds__(256) void k_adj_dens(const uint8_t* __restrict__ tipo, const float* __restrict__ dens,
                                                  const float* __restrict__ q, int nx, int ny, double alvo,
                                                  double* __restrict__ db) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  double g = 0;
  if (tipo[c] == AGUA) {
    const int i = c % nx, j = c / nx;
    const double e = (double)dens[c] / PPC - 1;
    bool live = e >= EXCESSO_MIN && e <= EXCESSO_MAX;
    if (live && e < 0) {
      // #miolo (fisica-cpu.js:486-496): all 9 cells water
      bool m = i >= 1 && j >= 1 && i < nx - 1 && j < ny - 1;
      for (int bb = -1; m && bb <= 1; bb++)
        for (int aa = -1; aa <= 1; aa++)
          if (tipo[c + aa + bb * nx] != AGUA) { m = false; break; }
      live = m;
    }
    if (live) g = alvo * (1.0 / PPC) * (double)q[c];
  }
  db[c] = g;
}

// P2G of one particle reversed (fisica-cpu.js:198-267), plus G2P's start-position adjoint
__global__ __launch_bounds__(256) void k_adj_p2g(const float2* __restrict__ pos, const float2* __restrict__ vel,
                                                 const float4* __restrict__ cm, const uint32_t* __restrict__ seq,
                                                 const uint32_t* __restrict__ n_dev, int nx, int ny, double h,
                                                 const double* __restrict__ ua, const double* __restrict__ uwb,
                                                 const double* __restrict__ va, const double* __restrict__ vwb,
                                                 const double* __restrict__ db, const double2* __restrict__ xg,
                                                 double* __restrict__ adjS) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *n_dev) return;
  const double inv = 1 / h;
  const double x = pos[i].x, y = pos[i].y;
  const double pvx = vel[i].x, pvy = vel[i].y;
  const float4 C = cm[i];
  const double cux = C.x, cuy = C.y, cvx = C.z, cvy = C.w;
  double pb[8] = {xg[i].x, xg[i].y, 0, 0, 0, 0, 0, 0};
  {
    const DSten st = sten_u(nx, ny, inv, x, y);
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++) {
        const int k = st.i0 + a + (st.j0 + b) * (nx + 1);
        const double wx = wa(st.tx, a), wy = wa(st.ty, b), w = wx * wy;
        const double dx = (st.i0 + a) * h - x, dy = (st.j0 + b + 0.5) * h - y;
        const double val = pvx + cux * dx + cuy * dy;
        const double g = ua[k] * val + uwb[k];
        pb[0] += g * dwa(a) * wy * inv - w * ua[k] * cux;
        pb[1] += g * wx * dwa(b) * inv - w * ua[k] * cuy;
        pb[2] += w * ua[k];
        pb[4] += w * ua[k] * dx;
        pb[5] += w * ua[k] * dy;
      }
  }
  {
    const DSten st = sten_v(nx, ny, inv, x, y);
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++) {
        const int k = st.i0 + a + (st.j0 + b) * nx;
        const double wx = wa(st.tx, a), wy = wa(st.ty, b), w = wx * wy;
        const double dx = (st.i0 + a + 0.5) * h - x, dy = (st.j0 + b) * h - y;
        const double val = pvy + cvx * dx + cvy * dy;
        const double g = va[k] * val + vwb[k];
        pb[0] += g * dwa(a) * wy * inv - w * va[k] * cvx;
        pb[1] += g * wx * dwa(b) * inv - w * va[k] * cvy;
        pb[3] += w * va[k];
        pb[6] += w * va[k] * dx;
        pb[7] += w * va[k] * dy;
      }
  }
  {
    const DSten st = sten_c(nx, ny, inv, x, y);
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++) {
        const double g = db[st.i0 + a + (st.j0 + b) * nx];
        pb[0] += g * dwa(a) * wa(st.ty, b) * inv;
        pb[1] += g * wa(st.tx, a) * dwa(b) * inv;
      }
  }
  double* o = adjS + 8 * (size_t)seq[i];
  for (int q = 0; q < 8; q++) o[q] = pb[q];
}

// ------------------------------------------------------------------ surface tension (fisica-cpu.js:306-403)
// The water fraction's chain: 3 passes of (wall extension, 3 layers), the first two followed by a
// (1 2 1) blur in x then in y; then normals, curvature and the air pressures.  13 steps.
constexpr int TEN_LAYERS = 9, TEN_STEPS = 13, TEN_STEPS_P1 = 14;

__global__ __launch_bounds__(256) void k_ten_init(const uint8_t* __restrict__ tipo, const float* __restrict__ dens, int C,
                                                  double* __restrict__ f, uint8_t* __restrict__ full) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  f[c] = tipo[c] == SOLIDO ? 0.0 : fmin(1.0, (double)dens[c] / PPC);
  full[c] = tipo[c] != SOLIDO;
}
__global__ __launch_bounds__(256) void k_ten_full(const uint8_t* __restrict__ tipo, int C, uint8_t* __restrict__ full) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C) full[c] = tipo[c] != SOLIDO;
}
// one extension layer (#estenderNaParede, fisica-cpu.js:351-370): fin/full before, fout/fullout after
__global__ __launch_bounds__(256) void k_ten_ext(const double* __restrict__ fin, const uint8_t* __restrict__ full, int nx,
                                                 int ny, double* __restrict__ fout, uint8_t* __restrict__ fullout) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  double v = fin[c];
  uint8_t fl = full[c];
  if (!fl) {
    double sum = 0;
    int n = 0;
    if (i > 0 && full[c - 1]) { sum += fin[c - 1]; n++; }
    if (i < nx - 1 && full[c + 1]) { sum += fin[c + 1]; n++; }
    if (j > 0 && full[c - nx]) { sum += fin[c - nx]; n++; }
    if (j < ny - 1 && full[c + nx]) { sum += fin[c + nx]; n++; }
    if (n) { v = sum / n; fl = 1; }
  }
  fout[c] = v;
  fullout[c] = fl;
}
// the (1 2 1)/4 blur along one axis with the edge cell standing in for its missing neighbour; it
// is symmetric, so the same kernel is its own transpose
__global__ __launch_bounds__(256) void k_ten_blur(const double* __restrict__ fin, int nx, int ny, int axis,
                                                  double* __restrict__ fout) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  int a, b;
  if (axis == 0) { a = i > 0 ? c - 1 : c; b = i < nx - 1 ? c + 1 : c; }
  else { a = j > 0 ? c - nx : c; b = j < ny - 1 ? c + nx : c; }
  fout[c] = (fin[a] + 2 * fin[c] + fin[b]) * 0.25;
}
__device__ inline bool interior(int i, int j, int nx, int ny) { return i >= 1 && j >= 1 && i < nx - 1 && j < ny - 1; }
__global__ __launch_bounds__(256) void k_ten_normals(const double* __restrict__ f, int nx, int ny, double* __restrict__ tnx,
                                                     double* __restrict__ tny, double* __restrict__ tg) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  double x = 0, y = 0, g = 0;
  if (interior(c % nx, c / nx, nx, ny)) {
    const double gx = (f[c + 1] - f[c - 1]) * 0.5, gy = (f[c + nx] - f[c - nx]) * 0.5;
    g = js_hypot(gx, gy);
    if (g > 1e-3) { x = gx / g; y = gy / g; }
  }
  tnx[c] = x; tny[c] = y; tg[c] = g;
}
// per air cell, d(qAr)/d(kappa) of each of its terms: sigma * escala / (terms) when unclamped
__global__ __launch_bounds__(256) void k_ten_coef(const uint8_t* __restrict__ tipo, const double* __restrict__ tnx,
                                                  const double* __restrict__ tny, int nx, int ny, double h, double sigma,
                                                  double escala, double* __restrict__ coef) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  double r = 0;
  const int i = c % nx, j = c / nx;
  if (sigma > 0 && tipo[c] == AR && interior(i, j, nx, ny)) {
    auto kap = [&](int k) {
      const int ki = k % nx, kj = k / nx;
      if (!interior(ki, kj, nx, ny)) return 0.0;
      return -((tnx[k + 1] - tnx[k - 1]) + (tny[k + nx] - tny[k - nx])) / (2 * h);
    };
    double sum = kap(c);
    int n = 1;
    const int nb[4] = {c - 1, c + 1, c - nx, c + nx};
    for (int q = 0; q < 4; q++)
      if (tipo[nb[q]] == AGUA) { sum += kap(nb[q]); n++; }
    const double m = sum / n, kmax = 1 / h;
    if (n > 1 && m > -kmax && m < kmax) r = sigma * escala / n;
  }
  coef[c] = r;
}
// air-pressure adjoints from the adjoint solve: each water cell's rhs has + qAr of its air
// neighbours, each active bubble's rhs - qAr per water face of its (interior) cells
__global__ __launch_bounds__(256) void k_adj_qar_rhs(const uint8_t* __restrict__ tipo, const uint32_t* __restrict__ rot,
                                                     const uint32_t* __restrict__ facesB, int ar, int nx, int ny,
                                                     const float* __restrict__ q, const float* __restrict__ yB,
                                                     double* __restrict__ qarb) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny || tipo[c] != AR) return;
  const int i = c % nx, j = c / nx;
  double g = 0;
  int nf = 0;
  if (i > 0 && tipo[c - 1] == AGUA) { g += q[c - 1]; nf++; }
  if (i < nx - 1 && tipo[c + 1] == AGUA) { g += q[c + 1]; nf++; }
  if (j > 0 && tipo[c - nx] == AGUA) { g += q[c - nx]; nf++; }
  if (j < ny - 1 && tipo[c + nx] == AGUA) { g += q[c + nx]; nf++; }
  uint32_t r;
  if (ar && bubble_cell(tipo, rot, facesB, nx, ny, i, j, &r)) g -= nf * (double)yB[r];
  qarb[c] += g;
}
// curvature adjoints: an air cell's own term and its water neighbours' terms
__global__ __launch_bounds__(256) void k_ten_kb(const uint8_t* __restrict__ tipo, const double* __restrict__ coef,
                                                const double* __restrict__ qarb, int nx, int ny, double* __restrict__ kb) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  double g = coef[c] * qarb[c];
  if (tipo[c] == AGUA) {
    const int nb[4] = {c - 1, c + 1, c - nx, c + nx};
    const bool in[4] = {i > 0, i < nx - 1, j > 0, j < ny - 1};
    for (int q = 0; q < 4; q++)
      if (in[q]) g += coef[nb[q]] * qarb[nb[q]];
  }
  kb[c] = interior(i, j, nx, ny) ? g : 0.0;
}
// normal adjoints from kappa_k = -(nx[k+1] - nx[k-1] + ny[k+nx] - ny[k-nx]) / 2h, then the
// normalisation's adjoint (n = g/|g| above 1e-3)
__global__ __launch_bounds__(256) void k_ten_nb(const double* __restrict__ kb, const double* __restrict__ tnx,
                                                const double* __restrict__ tny, const double* __restrict__ tg, int nx,
                                                int ny, double h, double* __restrict__ gxb, double* __restrict__ gyb) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  double bx = 0, by = 0;
  if (i > 0 && interior(i - 1, j, nx, ny)) bx -= kb[c - 1];
  if (i < nx - 1 && interior(i + 1, j, nx, ny)) bx += kb[c + 1];
  if (j > 0 && interior(i, j - 1, nx, ny)) by -= kb[c - nx];
  if (j < ny - 1 && interior(i, j + 1, nx, ny)) by += kb[c + nx];
  bx /= 2 * h;
  by /= 2 * h;
  double rx = 0, ry = 0;
  if (interior(i, j, nx, ny) && tg[c] > 1e-3) {
    const double dot = bx * tnx[c] + by * tny[c];
    rx = (bx - dot * tnx[c]) / tg[c];
    ry = (by - dot * tny[c]) / tg[c];
  }
  gxb[c] = rx;
  gyb[c] = ry;
}
// fraction adjoints from the central differences g = (f[k+1] - f[k-1], f[k+nx] - f[k-nx]) / 2
__global__ __launch_bounds__(256) void k_ten_fb(const double* __restrict__ gxb, const double* __restrict__ gyb, int nx,
                                                int ny, double* __restrict__ fb) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  double g = 0;
  if (i > 0 && interior(i - 1, j, nx, ny)) g += 0.5 * gxb[c - 1];
  if (i < nx - 1 && interior(i + 1, j, nx, ny)) g -= 0.5 * gxb[c + 1];
  if (j > 0 && interior(i, j - 1, nx, ny)) g += 0.5 * gyb[c - nx];
  if (j < ny - 1 && interior(i, j + 1, nx, ny)) g -= 0.5 * gyb[c + nx];
  fb[c] = g;
}
__global__ __launch_bounds__(256) void k_ten_dens(const uint8_t* __restrict__ tipo, const float* __restrict__ dens,
                                                  const double* __restrict__ fb, int C, double* __restrict__ db) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  if (tipo[c] != SOLIDO && (double)dens[c] / PPC < 1) db[c] += fb[c] / PPC;
}

// sum of the velocity adjoints of insertion numbers [s0, s1) (one faucet call's particles)
__global__ __launch_bounds__(256) void k_adj_emit(const double* __restrict__ adjS, uint32_t s0, uint32_t s1,
                                                  double* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double sx[256], sy[256];
  double ax = 0, ay = 0;
  for (uint32_t q = s0 + threadIdx.x; q < s1; q += 256) {
    ax += adjS[8 * (size_t)q + 2];
    ay += adjS[8 * (size_t)q + 3];
  }
  sx[threadIdx.x] = ax;
  sy[threadIdx.x] = ay;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) { sx[threadIdx.x] += sx[threadIdx.x + o]; sy[threadIdx.x] += sy[threadIdx.x + o]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = sx[0]; out[1] = sy[0]; }
}

// ------------------------------------------------------------------ water fraction of a box
__global__ __launch_bounds__(256) void k_frac_dens(const float2* __restrict__ pos, const DevScalars* __restrict__ sc,
                                                   int nx, int ny, double h, double* __restrict__ dens) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc->n_slots) return;
  const float2 P = pos[i];
  if (P.x != P.x) return;
  const DSten st = sten_c(nx, ny, 1 / h, P.x, P.y);
  for (int b = 0; b < 2; b++)
    for (int a = 0; a < 2; a++) atomicAdd(&dens[st.i0 + a + (st.j0 + b) * nx], wa(st.tx, a) * wa(st.ty, b));
}
// d(fraction)/d(x, y) of each particle: the fraction is the box mean of min(1, dens / PPC)
__global__ __launch_bounds__(256) void k_frac_grad(const float2* __restrict__ pos, const DevScalars* __restrict__ sc,
                                                   int nx, int ny, double h, const double* __restrict__ dens, int i0,
                                                   int j0, int i1, int j1, double coef, float2* __restrict__ out) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc->n_slots) return;
  const float2 P = pos[i];
  if (P.x != P.x) { out[i] = make_float2(0.f, 0.f); return; }
  const double inv = 1 / h;
  const DSten st = sten_c(nx, ny, inv, P.x, P.y);
  double gx = 0, gy = 0;
  for (int b = 0; b < 2; b++)
    for (int a = 0; a < 2; a++) {
      const int ci = st.i0 + a, cj = st.j0 + b;
      if (ci < i0 || ci >= i1 || cj < j0 || cj >= j1) continue;
      if (!(dens[ci + cj * nx] / PPC < 1)) continue;
      gx += coef * dwa(a) * wa(st.ty, b) * inv;
      gy += coef * wa(st.tx, a) * dwa(b) * inv;
    }
  out[i] = make_float2((float)gx, (float)gy);
}

// ------------------------------------------------------------------ water fraction without allocation
// tq_water_fraction_noalloc keeps its working memory in `scratch`, a file-scope device array
// (each GPU the library runs on holds its own copy until the process ends): the density of the
// box's cells (FP64, box-relative index), one partial per 256-cell block and the result.  Host
// code never passes `scratch` itself to a kernel: it passes the address hipGetSymbolAddress gives
// for the current device.
constexpr int WF_MAX_CELLS = 1 << 20;            // largest box (cells)
constexpr int WF_MAX_PARTS = WF_MAX_CELLS / 256; // block partials of the box sum
constexpr size_t WF_SCRATCH = (size_t)WF_MAX_CELLS + WF_MAX_PARTS + 8;
static __device__ __align__(256) double scratch[WF_SCRATCH];  // 8.4 MB per device

__global__ __launch_bounds__(256) void k_wf_clear(double* __restrict__ box, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) box[i] = 0.0;
}
// k_frac_dens restricted to the box: the same splat weights, added only to the box's cells
__global__ __launch_bounds__(256) void k_wf_dens(const float2* __restrict__ pos, const DevScalars* __restrict__ sc,
                                                 int nx, int ny, double h, int i0, int j0, int i1, int j1,
                                                 double* __restrict__ box) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc->n_slots) return;
  const float2 P = pos[i];
  if (P.x != P.x) return;
  const DSten st = sten_c(nx, ny, 1 / h, P.x, P.y);
  const int bw = i1 - i0;
  for (int b = 0; b < 2; b++)
    for (int a = 0; a < 2; a++) {
      const int ci = st.i0 + a, cj = st.j0 + b;
      if (ci < i0 || ci >= i1 || cj < j0 || cj >= j1) continue;
      atomicAdd(&box[(ci - i0) + (cj - j0) * bw], wa(st.tx, a) * wa(st.ty, b));
    }
}
// per 256 box cells: the sum of min(1, dens / PPC)
__global__ __launch_bounds__(256) void k_wf_sum(const double* __restrict__ box, int n, double* __restrict__ part) {
  __shared__ double sh[16];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {i < n ? fmin(1.0, box[i] / PPC) : 0.0};
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}
// the box mean from the partials, summed in block order (one block)
__global__ __launch_bounds__(256) void k_wf_final(const double* __restrict__ part, int nb, double cells,
                                                  double* __restrict__ out) {
  __shared__ double sh[16];
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v[0] += part[b];
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) *out = v[0] / cells;
}

// ------------------------------------------------------------------ host: records
static int blocks(long n) { return (int)std::max(1L, (n + 255) / 256); }

template <class T>
static int grow(T** p, int* cap, int n) {
  if (*cap >= n) return 0;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  const int c = std::max(n, 1024);
  TQ_HIP(hipMalloc((void**)p, sizeof(T) * (size_t)c));
  *cap = c;
  return 0;
}

static void rec_free(SubRec* r) {
  auto f = [](void* p) { if (p) (void)hipFree(p); };
  f(r->pos); f(r->vel); f(r->cm); f(r->seq); f(r->flags); f(r->n_dev);
  f(r->u1); f(r->v1); f(r->uw); f(r->vw); f(r->dens); f(r->u3); f(r->v3); f(r->qAr);
  f(r->tipo); f(r->rot); f(r->facesB); f(r->cellsB); f(r->volB); f(r->volAlvo); f(r->yB);
  delete r;
}

static int rec_alloc_grid(tq_sim* s, SubRec* r) {
  if (r->u1) return 0;
  const size_t C = s->C, nu = s->nu, nv = s->nv;
  TQ_HIP(hipMalloc((void**)&r->n_dev, sizeof(uint32_t)));
  TQ_HIP(hipMalloc((void**)&r->u1, 4 * nu)); TQ_HIP(hipMalloc((void**)&r->v1, 4 * nv));
  TQ_HIP(hipMalloc((void**)&r->uw, 4 * nu)); TQ_HIP(hipMalloc((void**)&r->vw, 4 * nv));
  TQ_HIP(hipMalloc((void**)&r->u3, 4 * nu)); TQ_HIP(hipMalloc((void**)&r->v3, 4 * nv));
  TQ_HIP(hipMalloc((void**)&r->dens, 4 * C)); TQ_HIP(hipMalloc((void**)&r->qAr, 4 * C));
  TQ_HIP(hipMalloc((void**)&r->tipo, C)); TQ_HIP(hipMalloc((void**)&r->rot, 4 * C));
  TQ_HIP(hipMalloc((void**)&r->facesB, 4 * (size_t)MAX_BOLHAS)); TQ_HIP(hipMalloc((void**)&r->volB, 4 * (size_t)MAX_BOLHAS));
  TQ_HIP(hipMalloc((void**)&r->cellsB, 4 * (size_t)MAX_BOLHAS));
  TQ_HIP(hipMalloc((void**)&r->volAlvo, 4 * (size_t)MAX_BOLHAS)); TQ_HIP(hipMalloc((void**)&r->yB, 4 * (size_t)MAX_BOLHAS));
  return 0;
}

static int rec_alloc_particles(SubRec* r, int n) {
  if (r->cap >= n) return 0;
  auto f = [](void* p) { if (p) (void)hipFree(p); };
  f(r->pos); f(r->vel); f(r->cm); f(r->seq); f(r->flags);
  const int c = std::max(n + n / 4, 1024);
  TQ_HIP(hipMalloc((void**)&r->pos, sizeof(float2) * (size_t)c));
  TQ_HIP(hipMalloc((void**)&r->vel, sizeof(float2) * (size_t)c));
  TQ_HIP(hipMalloc((void**)&r->cm, sizeof(float4) * (size_t)c));
  TQ_HIP(hipMalloc((void**)&r->seq, sizeof(uint32_t) * (size_t)c));
  TQ_HIP(hipMalloc((void**)&r->flags, sizeof(uint16_t) * (size_t)c));
  r->cap = c;
  return 0;
}

// copies of what the reverse pass needs, at four points of the substep
int rec_hook(tq_sim* s, int point, double dt) {
  SubRec* r = s->rec;
  const size_t C = s->C, nu = s->nu, nv = s->nv;
  auto cp = [&](void* dst, const void* src, size_t bytes) { return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s->st); };
  if (point == REC_P2G) {
    r->dt = dt;
    if (int e = rec_alloc_grid(s, r)) return e;
    const int n = s->n_upper;
    if (int e = rec_alloc_particles(r, n)) return e;
    r->n_up = n;
    const int o = s->cur;
    TQ_HIP(cp(r->n_dev, &s->sc->n_slots, sizeof(uint32_t)));
    if (n) {
      TQ_HIP(cp(r->pos, s->pos[o], sizeof(float2) * (size_t)n));
      TQ_HIP(cp(r->vel, s->vel[o], sizeof(float2) * (size_t)n));
      TQ_HIP(cp(r->cm, s->cm[o], sizeof(float4) * (size_t)n));
      TQ_HIP(cp(r->seq, s->seq[o], sizeof(uint32_t) * (size_t)n));
    }
    TQ_HIP(cp(r->u1, s->u, 4 * nu)); TQ_HIP(cp(r->v1, s->v, 4 * nv));
    TQ_HIP(cp(r->uw, s->uw, 4 * nu)); TQ_HIP(cp(r->vw, s->vw, 4 * nv));
    TQ_HIP(cp(r->dens, s->dens, 4 * C));
    s->g2p_flags = r->flags;
  } else if (point == REC_SOLVE) {
    TQ_HIP(cp(r->tipo, s->tipo, C)); TQ_HIP(cp(r->rot, s->rot, 4 * C)); TQ_HIP(cp(r->qAr, s->qAr, 4 * C));
    TQ_HIP(cp(r->facesB, s->facesB, 4 * (size_t)MAX_BOLHAS)); TQ_HIP(cp(r->volB, s->volB, 4 * (size_t)MAX_BOLHAS));
    TQ_HIP(cp(r->cellsB, s->cellsB, 4 * (size_t)MAX_BOLHAS));
    TQ_HIP(cp(r->volAlvo, s->volAlvo, 4 * (size_t)MAX_BOLHAS)); TQ_HIP(cp(r->yB, s->yB32, 4 * (size_t)MAX_BOLHAS));
  } else if (point == REC_EXTRAP) {
    TQ_HIP(cp(r->u3, s->u, 4 * nu)); TQ_HIP(cp(r->v3, s->v, 4 * nv));
  } else if (point == REC_G2P) {
    s->g2p_flags = nullptr;
    Tape* T = s->tape;
    T->recs_used++;
    if (T->recs_used >= T->recs.size()) T->recs.push_back(new SubRec());
    SubRec* nx = T->recs[T->recs_used];
    nx->emit = -1;
    s->rec = nx;
  }
  return 0;
}

// ------------------------------------------------------------------ host: checkpoints
static int snap_take(tq_sim* s, Snapshot* S) {
  if (S->dev.empty()) {
    S->dev.resize(s->bufs.size(), nullptr);
    for (size_t k = 0; k < s->bufs.size(); k++) TQ_HIP(hipMalloc(&S->dev[k], s->bufs[k].second));
  }
  for (size_t k = 0; k < s->bufs.size(); k++)
    TQ_HIP(hipMemcpyAsync(S->dev[k], s->bufs[k].first, s->bufs[k].second, hipMemcpyDeviceToDevice, s->st));
  S->cur = s->cur; S->n_upper = s->n_upper; S->seq_next = s->seq_next;
  S->tipos_prontos = s->tipos_prontos; S->has_vmax = s->has_vmax; S->air_reset_pending = s->air_reset_pending;
  std::memcpy(&S->sc, s->sc_host, sizeof(DevScalars));
  TQ_HIP(hipStreamSynchronize(s->st));
  return 0;
}
static int snap_restore(tq_sim* s, const Snapshot* S) {
  for (size_t k = 0; k < s->bufs.size(); k++)
    TQ_HIP(hipMemcpyAsync(s->bufs[k].first, S->dev[k], s->bufs[k].second, hipMemcpyDeviceToDevice, s->st));
  s->cur = S->cur; s->n_upper = S->n_upper; s->seq_next = S->seq_next;
  s->tipos_prontos = S->tipos_prontos; s->has_vmax = S->has_vmax; s->air_reset_pending = S->air_reset_pending;
  std::memcpy(s->sc_host, &S->sc, sizeof(DevScalars));
  TQ_HIP(hipStreamSynchronize(s->st));
  return 0;
}
static void snap_free(Snapshot* S) {
  for (void* p : S->dev) if (p) (void)hipFree(p);
  delete S;
}

void tape_free(tq_sim* s) {
  Tape* T = s->tape;
  if (!T) return;
  (void)hipStreamSynchronize(s->st);
  for (auto& c : T->cps) snap_free(c.second);
  for (SubRec* r : T->recs) rec_free(r);
  auto f = [](void* p) { if (p) (void)hipFree(p); };
  f(T->adjS); f(T->xg); f(T->ub); f(T->vb); f(T->ub2); f(T->vb2); f(T->ua); f(T->uwb); f(T->va); f(T->vwb);
  f(T->db); f(T->ybar); f(T->gsdf); f(T->gemit); f(T->val0u); f(T->val1u); f(T->val0v); f(T->val1v);
  f(T->qarb); f(T->tf); f(T->tfull); f(T->tnx); f(T->tny); f(T->tg); f(T->tcoef); f(T->tkb); f(T->tnxb); f(T->tnyb);
  f(T->tfb); f(T->tfb2);
  delete T;
  s->tape = nullptr;
  s->rec = nullptr;
  s->g2p_flags = nullptr;
}

static void open_records(tq_sim* s) {
  Tape* T = s->tape;
  T->recs_used = 0;
  if (T->recs.empty()) T->recs.push_back(new SubRec());
  T->recs[0]->emit = -1;
  s->rec = T->recs[0];
}

int tape_log_emit(tq_sim* s, const tq_faucet* f, double dt, uint32_t seed, uint32_t seq0, uint32_t seq1) {
  Tape* T = s->tape;
  if (T->replaying) return 0;
  TapeOp op;
  op.kind = 0;
  op.f = *f;
  op.dt = dt;
  op.seed = seed;
  op.seq0 = seq0;
  op.seq1 = seq1;
  op.emit = T->n_emit++;
  T->ops.push_back(op);
  if (s->rec) s->rec->emit = op.emit;  // keep-all mode: the next substep record takes it
  return 0;
}

int tape_log_step(tq_sim* s, double dt, int substeps) {
  Tape* T = s->tape;
  if (T->replaying) return 0;
  TapeOp op;
  op.kind = 1;
  op.dt = dt;
  op.substeps = substeps;
  T->ops.push_back(op);
  T->steps++;
  if (T->every > 0 && T->steps % T->every == 0) {
    Snapshot* S = new Snapshot();
    if (int e = snap_take(s, S)) { snap_free(S); return e; }
    T->cps.emplace_back(T->ops.size(), S);
  }
  return 0;
}

// re-runs ops [a, b) with substep records open
static int replay(tq_sim* s, size_t a, size_t b) {
  Tape* T = s->tape;
  T->replaying = true;
  open_records(s);
  int rc = 0;
  for (size_t k = a; k < b && !rc; k++) {
    const TapeOp& op = T->ops[k];
    if (op.kind == 0) {
      uint32_t seed = op.seed;
      rc = emit_faucet_impl(s, &op.f, op.dt, &seed, nullptr, nullptr);
      if (!rc && s->seq_next != op.seq1) { set_error("tq_backward: faucet replay differs from the recording"); rc = TQ_E_STATE; }
      if (!rc) s->rec->emit = op.emit;
    } else {
      rc = step_enqueue(s, op.dt, op.substeps);
      if (!rc) rc = stream_wait(s);
    }
  }
  s->rec = nullptr;
  s->g2p_flags = nullptr;
  T->replaying = false;
  return rc;
}

// ------------------------------------------------------------------ host: the reverse of one substep
static int reverse_substep(tq_sim* s, SubRec* r, double tol) {
  Tape* T = s->tape;
  const int nx = s->nx, ny = s->ny, C = s->C, nu = s->nu, nv = s->nv;
  if (int e = grow(&T->xg, &T->xg_cap, std::max(r->n_up, 1))) return e;
  TQ_HIP(hipMemsetAsync(T->ub, 0, 8 * (size_t)nu, s->st));
  TQ_HIP(hipMemsetAsync(T->vb, 0, 8 * (size_t)nv, s->st));
  if (r->n_up > 0)
    hipLaunchKernelGGL(k_adj_g2p, dim3(blocks(r->n_up)), dim3(256), 0, s->st, r->pos, r->seq, r->flags, r->n_dev, r->u3,
                       r->v3, s->sdf, nx, ny, s->h, r->dt, T->adjS, T->ub, T->vb, T->gsdf, T->xg);
  // extrapolation layers, the second first
  hipLaunchKernelGGL(k_adj_valid0, dim3(blocks(std::max(nu, nv))), dim3(256), 0, s->st, r->tipo, nx, ny, T->val0u, T->val0v);
  hipLaunchKernelGGL(k_adj_valid_next, dim3(blocks(nu)), dim3(256), 0, s->st, T->val0u, nx + 1, ny, T->val1u);
  hipLaunchKernelGGL(k_adj_valid_next, dim3(blocks(nv)), dim3(256), 0, s->st, T->val0v, nx, ny + 1, T->val1v);
  hipLaunchKernelGGL(k_adj_extrap, dim3(blocks(nu)), dim3(256), 0, s->st, T->ub, T->val1u, nx + 1, ny, T->ub2);
  hipLaunchKernelGGL(k_adj_extrap, dim3(blocks(nv)), dim3(256), 0, s->st, T->vb, T->val1v, nx, ny + 1, T->vb2);
  hipLaunchKernelGGL(k_adj_extrap, dim3(blocks(nu)), dim3(256), 0, s->st, T->ub2, T->val0u, nx + 1, ny, T->ub);
  hipLaunchKernelGGL(k_adj_extrap, dim3(blocks(nv)), dim3(256), 0, s->st, T->vb2, T->val0v, nx, ny + 1, T->vb);
  // the adjoint solve on this substep's operator
  TQ_HIP(hipMemcpyAsync(s->tipo, r->tipo, C, hipMemcpyDeviceToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(s->rot, r->rot, 4 * (size_t)C, hipMemcpyDeviceToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(s->facesB, r->facesB, 4 * (size_t)MAX_BOLHAS, hipMemcpyDeviceToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(s->cellsB, r->cellsB, 4 * (size_t)MAX_BOLHAS, hipMemcpyDeviceToDevice, s->st));
  TQ_HIP(hipMemsetAsync(T->ybar, 0, 8 * (size_t)MAX_BOLHAS, s->st));
  hipLaunchKernelGGL(k_adj_gradq, dim3(blocks(C)), dim3(256), 0, s->st, r->tipo, r->rot, r->facesB, s->d.ar, nx, ny, T->ub,
                     T->vb, s->div, T->ybar, T->qarb);
  TQ_LAUNCH_CHECK();
  TQ_HIP(hipMemcpyAsync(s->rhsB, T->ybar, 8 * (size_t)MAX_BOLHAS, hipMemcpyDeviceToDevice, s->st));
  TQ_HIP(hipMemsetAsync(s->q, 0, 4 * (size_t)C, s->st));
  TQ_HIP(hipMemsetAsync(s->yB32, 0, 4 * (size_t)MAX_BOLHAS, s->st));
  const double tol0 = s->d.tol;
  const int it0 = s->d.iteracoes;
  s->d.tol = tol;
  s->d.iteracoes = std::max(200, it0);
  const int rc = pressure_solve(s);
  s->d.tol = tol0;
  s->d.iteracoes = it0;
  if (rc) return rc;
  const double alvo = s->h / std::max(s->d.tau_densidade, r->dt);
  hipLaunchKernelGGL(k_adj_faces, dim3(blocks(std::max(nu, nv))), dim3(256), 0, s->st, r->tipo, r->rot, r->facesB, s->d.ar,
                     nx, ny, s->q, s->yB32, T->ub, T->vb, r->u1, r->uw, r->v1, r->vw, s->d.gx * r->dt, s->d.gy * r->dt,
                     T->ua, T->uwb, T->va, T->vwb);
  hipLaunchKernelGGL(k_adj_dens, dim3(blocks(C)), dim3(256), 0, s->st, r->tipo, r->dens, s->q, nx, ny, alvo, T->db);
  if (s->d.tensao > 0) {
    // surface tension: the air pressures' adjoints back to the density
    hipLaunchKernelGGL(k_adj_qar_rhs, dim3(blocks(C)), dim3(256), 0, s->st, r->tipo, r->rot, r->facesB, s->d.ar, nx, ny,
                       s->q, s->yB32, T->qarb);
    double* F = T->tf;
    uint8_t* FU = T->tfull;
    auto f_at = [&](int k) { return F + (size_t)k * C; };
    auto m_at = [&](int k) { return FU + (size_t)k * C; };
    // the chain forward, keeping every step's input
    hipLaunchKernelGGL(k_ten_init, dim3(blocks(C)), dim3(256), 0, s->st, r->tipo, r->dens, C, f_at(0), m_at(0));
    int step = 0;
    int kinds[TEN_STEPS];  // -1 - layer index for extension layers, 0/1 blur axis
    for (int pass = 0; pass <= 2; pass++) {
      if (pass > 0) hipLaunchKernelGGL(k_ten_full, dim3(blocks(C)), dim3(256), 0, s->st, r->tipo, C, m_at(3 * pass));
      for (int layer = 0; layer < 3; layer++, step++) {
        const int L = 3 * pass + layer;
        uint8_t* mout = layer < 2 ? m_at(L + 1) : m_at(TEN_LAYERS);
        hipLaunchKernelGGL(k_ten_ext, dim3(blocks(C)), dim3(256), 0, s->st, f_at(step), m_at(L), nx, ny, f_at(step + 1), mout);
        kinds[step] = -1 - L;
      }
      if (pass == 2) break;
      for (int axis = 0; axis < 2; axis++, step++) {
        hipLaunchKernelGGL(k_ten_blur, dim3(blocks(C)), dim3(256), 0, s->st, f_at(step), nx, ny, axis, f_at(step + 1));
        kinds[step] = axis;
      }
    }
    const double escala = r->dt / (DENSIDADE_AGUA * s->h);
    hipLaunchKernelGGL(k_ten_normals, dim3(blocks(C)), dim3(256), 0, s->st, f_at(TEN_STEPS), nx, ny, T->tnx, T->tny, T->tg);
    hipLaunchKernelGGL(k_ten_coef, dim3(blocks(C)), dim3(256), 0, s->st, r->tipo, T->tnx, T->tny, nx, ny, s->h,
                       s->d.tensao, escala, T->tcoef);
    // and backwards
    hipLaunchKernelGGL(k_ten_kb, dim3(blocks(C)), dim3(256), 0, s->st, r->tipo, T->tcoef, T->qarb, nx, ny, T->tkb);
    hipLaunchKernelGGL(k_ten_nb, dim3(blocks(C)), dim3(256), 0, s->st, T->tkb, T->tnx, T->tny, T->tg, nx, ny, s->h, T->tnxb,
                       T->tnyb);
    hipLaunchKernelGGL(k_ten_fb, dim3(blocks(C)), dim3(256), 0, s->st, T->tnxb, T->tnyb, nx, ny, T->tfb);
    double *a = T->tfb, *b = T->tfb2;
    for (int k = TEN_STEPS - 1; k >= 0; k--) {
      if (kinds[k] < 0)
        hipLaunchKernelGGL(k_adj_extrap, dim3(blocks(C)), dim3(256), 0, s->st, a, m_at(-1 - kinds[k]), nx, ny, b);
      else
        hipLaunchKernelGGL(k_ten_blur, dim3(blocks(C)), dim3(256), 0, s->st, a, nx, ny, kinds[k], b);
      std::swap(a, b);
    }
    hipLaunchKernelGGL(k_ten_dens, dim3(blocks(C)), dim3(256), 0, s->st, r->tipo, r->dens, a, C, T->db);
    TQ_LAUNCH_CHECK();
  }
  if (r->n_up > 0)
    hipLaunchKernelGGL(k_adj_p2g, dim3(blocks(r->n_up)), dim3(256), 0, s->st, r->pos, r->vel, r->cm, r->seq, r->n_dev, nx,
                       ny, s->h, T->ua, T->uwb, T->va, T->vwb, T->db, T->xg, T->adjS);
  TQ_LAUNCH_CHECK();
  return 0;
}

static int tape_scratch(tq_sim* s) {
  Tape* T = s->tape;
  if (T->ub) return 0;
  const size_t C = s->C, nu = s->nu, nv = s->nv;
  TQ_HIP(hipMalloc((void**)&T->ub, 8 * nu)); TQ_HIP(hipMalloc((void**)&T->vb, 8 * nv));
  TQ_HIP(hipMalloc((void**)&T->ub2, 8 * nu)); TQ_HIP(hipMalloc((void**)&T->vb2, 8 * nv));
  TQ_HIP(hipMalloc((void**)&T->ua, 8 * nu)); TQ_HIP(hipMalloc((void**)&T->uwb, 8 * nu));
  TQ_HIP(hipMalloc((void**)&T->va, 8 * nv)); TQ_HIP(hipMalloc((void**)&T->vwb, 8 * nv));
  TQ_HIP(hipMalloc((void**)&T->db, 8 * C)); TQ_HIP(hipMalloc((void**)&T->ybar, 8 * (size_t)MAX_BOLHAS));
  TQ_HIP(hipMalloc((void**)&T->qarb, 8 * C));
  TQ_HIP(hipMalloc((void**)&T->tf, 8 * C * TEN_STEPS_P1)); TQ_HIP(hipMalloc((void**)&T->tfull, C * (TEN_LAYERS + 1)));
  for (double** p : {&T->tnx, &T->tny, &T->tg, &T->tcoef, &T->tkb, &T->tnxb, &T->tnyb, &T->tfb, &T->tfb2})
    TQ_HIP(hipMalloc((void**)p, 8 * C));
  TQ_HIP(hipMalloc((void**)&T->gsdf, 8 * 4 * C));
  TQ_HIP(hipMalloc((void**)&T->val0u, nu)); TQ_HIP(hipMalloc((void**)&T->val1u, nu));
  TQ_HIP(hipMalloc((void**)&T->val0v, nv)); TQ_HIP(hipMalloc((void**)&T->val1v, nv));
  return 0;
}

}  // namespace tq

using namespace tq;

extern "C" {

int tq_tape_begin(tq_sim* s, int checkpoint_every) {
  if (!s) { set_error("null handle"); return TQ_E_ARG; }
  if (s->tape) { set_error("tq_tape_begin: frames are already being recorded"); return TQ_E_STATE; }
  if (checkpoint_every < 0) { set_error("tq_tape_begin: checkpoint_every must be >= 0"); return TQ_E_ARG; }
  Tape* T = new Tape();
  T->every = checkpoint_every;
  s->tape = T;
  std::vector<float> parts;
  if (int e = download_ordered(s, parts, &T->seq_begin)) { tape_free(s); return e; }
  if (checkpoint_every > 0) {
    Snapshot* S = new Snapshot();
    if (int e = snap_take(s, S)) { snap_free(S); tape_free(s); return e; }
    T->cps.emplace_back(0, S);
  } else {
    open_records(s);
  }
  return 0;
}

int tq_tape_end(tq_sim* s) {
  if (!s) { set_error("null handle"); return TQ_E_ARG; }
  tape_free(s);
  return 0;
}

int tq_tape_info(tq_sim* s, int* frames, int* faucet_calls, int* checkpoints, int* n_begin) {
  if (!s) { set_error("null handle"); return TQ_E_ARG; }
  if (!s->tape) { set_error("tq_tape_info: no frames recorded"); return TQ_E_STATE; }
  if (frames) *frames = s->tape->steps;
  if (faucet_calls) *faucet_calls = s->tape->n_emit;
  if (checkpoints) *checkpoints = (int)s->tape->cps.size();
  if (n_begin) *n_begin = (int)s->tape->seq_begin.size();
  return 0;
}

// test hook: fields of substep record k of a keep-all tape (checkpoint_every = 0)
//   0 particle count (uint32), 1 types (C bytes), 2 labels (C uint32), 3 qAr (C floats),
//   4 volumeAlvo (MAX_BOLHAS floats), 5 yBolha at the solve (MAX_BOLHAS floats),
//   6 decision words (count uint16, P2G order), 7 insertion numbers (count uint32), 8 dt (double)
int tq_tape_record(tq_sim* s, int k, int field, void* dst, size_t bytes) {
  if (!s || !dst) { set_error("tq_tape_record: null argument"); return TQ_E_ARG; }
  Tape* T = s->tape;
  if (!T || T->every != 0) { set_error("tq_tape_record: no keep-all tape"); return TQ_E_STATE; }
  if (k < 0 || (size_t)k >= T->recs_used) { set_error("tq_tape_record: record out of range"); return TQ_E_ARG; }
  const SubRec* r = T->recs[k];
  uint32_t n = 0;
  TQ_HIP(hipMemcpy(&n, r->n_dev, 4, hipMemcpyDeviceToHost));
  const size_t C = s->C, M = MAX_BOLHAS;
  const void* src = nullptr;
  size_t want = 0;
  switch (field) {
    case 0: want = 4; src = r->n_dev; break;
    case 1: want = C; src = r->tipo; break;
    case 2: want = 4 * C; src = r->rot; break;
    case 3: want = 4 * C; src = r->qAr; break;
    case 4: want = 4 * M; src = r->volAlvo; break;
    case 5: want = 4 * M; src = r->yB; break;
    case 6: want = 2 * (size_t)n; src = r->flags; break;
    case 7: want = 4 * (size_t)n; src = r->seq; break;
    case 8:
      if (bytes != 8) { set_error("tq_tape_record: size mismatch"); return TQ_E_ARG; }
      std::memcpy(dst, &r->dt, 8);
      return 0;
    default: set_error("tq_tape_record: unknown field"); return TQ_E_ARG;
  }
  if (bytes != want) { set_error("tq_tape_record: size mismatch"); return TQ_E_ARG; }
  if (want) TQ_HIP(hipMemcpy(dst, src, want, hipMemcpyDeviceToHost));
  return 0;
}

int tq_tape_records(tq_sim* s, int* n) {
  if (!s || !n) { set_error("tq_tape_records: null argument"); return TQ_E_ARG; }
  if (!s->tape) { set_error("tq_tape_records: no tape"); return TQ_E_STATE; }
  *n = (int)s->tape->recs_used;
  return 0;
}

int tq_water_fraction(tq_sim* s, int i0, int j0, int i1, int j1, double* frac, float* grad, int n) {
  if (!s) { set_error("null handle"); return TQ_E_ARG; }
  if (!frac || i0 < 0 || j0 < 0 || i1 > s->nx || j1 > s->ny || i0 >= i1 || j0 >= j1) {
    set_error("tq_water_fraction: bad box");
    return TQ_E_ARG;
  }
  const size_t C = s->C;
  double* dens = nullptr;
  float2* g = nullptr;
  int rc = 0;
  std::vector<double> hd(C);
  std::vector<float> parts;
  std::vector<uint32_t> seqs;
  auto fail = [&](hipError_t e, const char* what) { return hip_fail(e, what, __FILE__, __LINE__); };
  hipError_t he = hipMalloc((void**)&dens, 8 * C);
  if (he != hipSuccess) return fail(he, "hipMalloc");
  const int slots = std::max(s->n_upper, 1);
  he = hipMalloc((void**)&g, sizeof(float2) * (size_t)slots);
  if (he != hipSuccess) { (void)hipFree(dens); return fail(he, "hipMalloc"); }
  (void)hipMemsetAsync(dens, 0, 8 * C, s->st);
  const int o = s->cur;
  if (s->n_upper > 0)
    hipLaunchKernelGGL(k_frac_dens, dim3(blocks(s->n_upper)), dim3(256), 0, s->st, s->pos[o], s->sc, s->nx, s->ny, s->h, dens);
  he = hipMemcpyAsync(hd.data(), dens, 8 * C, hipMemcpyDeviceToHost, s->st);
  if (he == hipSuccess) he = hipStreamSynchronize(s->st);
  if (he != hipSuccess) rc = fail(he, "tq_water_fraction");
  double sum = 0;
  const long cells = (long)(i1 - i0) * (j1 - j0);
  for (int j = j0; j < j1 && !rc; j++)
    for (int i = i0; i < i1; i++) sum += std::min(1.0, hd[(size_t)i + (size_t)j * s->nx] / PPC);
  if (!rc) *frac = sum / cells;
  if (!rc && grad) {
    if (s->n_upper > 0)
      hipLaunchKernelGGL(k_frac_grad, dim3(blocks(s->n_upper)), dim3(256), 0, s->st, s->pos[o], s->sc, s->nx, s->ny, s->h,
                         dens, i0, j0, i1, j1, 1.0 / ((double)cells * PPC), g);
    // the gradient in the order tq_read_particles gives (ascending insertion number)
    rc = download_ordered(s, parts, &seqs);
    if (!rc && (int)seqs.size() != n) { set_error("tq_water_fraction: n is not the live particle count"); rc = TQ_E_ARG; }
    if (!rc) {
      std::vector<float2> hg(s->n_upper > 0 ? s->n_upper : 1);
      std::vector<uint32_t> sq(hg.size());
      he = hipMemcpyAsync(hg.data(), g, sizeof(float2) * hg.size(), hipMemcpyDeviceToHost, s->st);
      if (he == hipSuccess) he = hipMemcpyAsync(sq.data(), s->seq[o], 4 * sq.size(), hipMemcpyDeviceToHost, s->st);
      if (he == hipSuccess) he = hipStreamSynchronize(s->st);
      if (he != hipSuccess) rc = fail(he, "tq_water_fraction");
      std::vector<int> at;
      if (!rc) {
        const uint32_t top = seqs.empty() ? 0 : seqs.back() + 1;
        at.assign(top, -1);
        for (size_t k = 0; k < seqs.size(); k++) at[seqs[k]] = (int)k;
        std::fill(grad, grad + 8 * (size_t)n, 0.f);
        for (size_t k = 0; k < sq.size() && k < (size_t)s->sc_host->n_slots; k++) {
          if (sq[k] >= top || at[sq[k]] < 0) continue;
          grad[8 * (size_t)at[sq[k]]] = hg[k].x;
          grad[8 * (size_t)at[sq[k]] + 1] = hg[k].y;
        }
      }
    }
  }
  (void)hipFree(dens);
  (void)hipFree(g);
  return rc;
}

// The address of `scratch` on the current device, looked up once per device (each GPU holds its
// own copy at its own address); null (with the error set) when the lookup fails.
static double* wf_scratch_ptr() {
  static std::mutex mu;
  static double* ptrs[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) { (void)hip_fail(e, "hipGetDevice", __FILE__, __LINE__); return nullptr; }
  if (dev < 0 || dev >= 64) { set_error("tq_water_fraction_noalloc: device index out of range"); return nullptr; }
  std::lock_guard<std::mutex> lk(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(tq::scratch));
    if (e != hipSuccess || !p) {
      (void)hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
      return nullptr;
    }
    ptrs[dev] = (double*)p;
  }
  return ptrs[dev];
}

int tq_water_fraction_noalloc(tq_sim* s, int i0, int j0, int i1, int j1, double* frac) {
  if (!s) { set_error("null handle"); return TQ_E_ARG; }
  if (!frac || i0 < 0 || j0 < 0 || i1 > s->nx || j1 > s->ny || i0 >= i1 || j0 >= j1) {
    set_error("tq_water_fraction_noalloc: bad box");
    return TQ_E_ARG;
  }
  const long cells = (long)(i1 - i0) * (j1 - j0);
  if (cells > WF_MAX_CELLS) {
    set_error("tq_water_fraction_noalloc: the box exceeds the 1048576-cell scratch (use tq_water_fraction)");
    return TQ_E_CAPACITY;
  }
  double* base = wf_scratch_ptr();
  if (!base) return TQ_E_HIP;
  double* box = base;
  double* part = base + WF_MAX_CELLS;
  double* out = part + WF_MAX_PARTS;
  const int nb = blocks(cells);
  // the box is cleared by a kernel (no memset node if a caller captures this into a graph)
  hipLaunchKernelGGL(k_wf_clear, dim3(nb), dim3(256), 0, s->st, box, (int)cells);
  if (s->n_upper > 0)
    hipLaunchKernelGGL(k_wf_dens, dim3(blocks(s->n_upper)), dim3(256), 0, s->st, s->pos[s->cur], s->sc, s->nx, s->ny,
                       s->h, i0, j0, i1, j1, box);
  hipLaunchKernelGGL(k_wf_sum, dim3(nb), dim3(256), 0, s->st, (const double*)box, (int)cells, part);
  hipLaunchKernelGGL(k_wf_final, dim3(1), dim3(256), 0, s->st, (const double*)part, nb, (double)cells, out);
  TQ_LAUNCH_CHECK();
  TQ_HIP(hipMemcpyAsync(frac, out, sizeof(double), hipMemcpyDeviceToHost, s->st));
  TQ_HIP(hipStreamSynchronize(s->st));
  return 0;
}

int tq_backward(tq_sim* s, const float* grad_state, int n, double tol, double* grad_faucet, int n_faucet,
                float* grad_sdf, float* grad_state0, int n0) {
  if (!s) { set_error("null handle"); return TQ_E_ARG; }
  Tape* T = s->tape;
  if (!T) { set_error("tq_backward: no frames recorded (tq_tape_begin)"); return TQ_E_STATE; }
  if (!grad_state && n > 0) { set_error("tq_backward: null grad_state"); return TQ_E_ARG; }
  if (n_faucet != T->n_emit || (n_faucet > 0 && !grad_faucet)) { set_error("tq_backward: n_faucet is not the recorded faucet call count"); return TQ_E_ARG; }
  if (grad_state0 && n0 != (int)T->seq_begin.size()) { set_error("tq_backward: n0 is not the particle count at tq_tape_begin"); return TQ_E_ARG; }
  if (!(tol > 0)) { set_error("tq_backward: tol must be > 0"); return TQ_E_ARG; }
  std::vector<float> parts;
  std::vector<uint32_t> seqs;
  if (int e = download_ordered(s, parts, &seqs)) return e;
  if ((int)seqs.size() != n) { set_error("tq_backward: n is not the live particle count"); return TQ_E_ARG; }
  if (int e = tape_scratch(s)) return e;
  // the end state, put back when done
  Snapshot* end = new Snapshot();
  int rc = snap_take(s, end);
  // adjoints by insertion number
  const int top = (int)s->seq_next + 1;
  if (!rc && T->seq_cap < top) {
    if (T->adjS) (void)hipFree(T->adjS);
    T->adjS = nullptr;
    hipError_t he = hipMalloc((void**)&T->adjS, 8 * 8 * (size_t)top);
    if (he != hipSuccess) rc = hip_fail(he, "hipMalloc", __FILE__, __LINE__);
    else T->seq_cap = top;
  }
  if (!rc) {
    std::vector<double> A(8 * (size_t)top, 0.0);
    for (int k = 0; k < n; k++)
      for (int q = 0; q < 8; q++) A[8 * (size_t)seqs[k] + q] = grad_state[8 * (size_t)k + q];
    hipError_t he = hipMemcpyAsync(T->adjS, A.data(), 8 * A.size(), hipMemcpyHostToDevice, s->st);
    if (he == hipSuccess) he = hipMemsetAsync(T->gsdf, 0, 8 * 4 * (size_t)s->C, s->st);
    if (he == hipSuccess && T->gemit) (void)hipFree(T->gemit), T->gemit = nullptr;
    if (he == hipSuccess) he = hipMalloc((void**)&T->gemit, 16 * (size_t)std::max(1, T->n_emit));
    if (he == hipSuccess) he = hipMemsetAsync(T->gemit, 0, 16 * (size_t)std::max(1, T->n_emit), s->st);
    if (he != hipSuccess) rc = hip_fail(he, "tq_backward", __FILE__, __LINE__);
  }
  // segments from the last to the first (keep-all mode: the one recorded segment)
  std::vector<std::pair<size_t, size_t>> segs;
  if (T->every > 0) {
    for (size_t k = 0; k < T->cps.size(); k++) {
      const size_t a = T->cps[k].first, b = k + 1 < T->cps.size() ? T->cps[k + 1].first : T->ops.size();
      if (a < b) segs.emplace_back(k, b);
    }
  }
  auto reverse_records = [&]() -> int {
    for (size_t q = T->recs_used; q-- > 0;) {
      SubRec* r = T->recs[q];
      if (int e = reverse_substep(s, r, tol)) return e;
      if (r->emit >= 0) {
        const TapeOp* op = nullptr;
        for (const TapeOp& o : T->ops) if (o.kind == 0 && o.emit == r->emit) { op = &o; break; }
        if (op && op->seq1 > op->seq0)
          hipLaunchKernelGGL(k_adj_emit, dim3(1), dim3(256), 0, s->st, T->adjS, op->seq0, op->seq1, T->gemit + 2 * r->emit);
        TQ_LAUNCH_CHECK();
      }
    }
    return 0;
  };
  if (!rc) {
    if (T->every > 0) {
      for (size_t k = segs.size(); k-- > 0 && !rc;) {
        const size_t cp = segs[k].first;
        rc = snap_restore(s, T->cps[cp].second);
        if (!rc) rc = replay(s, T->cps[cp].first, segs[k].second);
        if (!rc) rc = reverse_records();
      }
    } else {
      // keep-all: the open record is the unfilled one after the last substep
      rc = reverse_records();
    }
  }
  if (!rc) {
    hipError_t he = hipStreamSynchronize(s->st);
    if (he != hipSuccess) rc = hip_fail(he, "tq_backward", __FILE__, __LINE__);
  }
  if (!rc && grad_faucet && n_faucet > 0) {
    hipError_t he = hipMemcpy(grad_faucet, T->gemit, 16 * (size_t)n_faucet, hipMemcpyDeviceToHost);
    if (he != hipSuccess) rc = hip_fail(he, "tq_backward", __FILE__, __LINE__);
  }
  if (!rc && grad_sdf) {
    std::vector<double> g(4 * (size_t)s->C);
    hipError_t he = hipMemcpy(g.data(), T->gsdf, 8 * g.size(), hipMemcpyDeviceToHost);
    if (he != hipSuccess) rc = hip_fail(he, "tq_backward", __FILE__, __LINE__);
    else for (size_t k = 0; k < g.size(); k++) grad_sdf[k] = (float)g[k];
  }
  if (!rc && grad_state0 && n0 > 0) {
    std::vector<double> A(8 * (size_t)top);
    hipError_t he = hipMemcpy(A.data(), T->adjS, 8 * A.size(), hipMemcpyDeviceToHost);
    if (he != hipSuccess) rc = hip_fail(he, "tq_backward", __FILE__, __LINE__);
    else
      for (int k = 0; k < n0; k++)
        for (int q = 0; q < 8; q++) grad_state0[8 * (size_t)k + q] = (float)A[8 * (size_t)T->seq_begin[k] + q];
  }
  // back to the end state (keep-all mode: the records stay; the open record stays open)
  const int rc2 = snap_restore(s, end);
  snap_free(end);
  if (T->every == 0) {
    // a further tq_step would append to the records: keep the last one open
    s->rec = T->recs[T->recs_used];
  }
  return rc ? rc : rc2;
}

}  // extern "C"
