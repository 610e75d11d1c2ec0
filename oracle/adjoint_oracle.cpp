// adjoint_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
//
// An FP64 CPU restatement of the substep (js/fisica-cpu.js:178-191) with every discrete
// decision taken from a tape, and its hand-derived reverse-mode adjoint.  It is the checker of
// the HIP adjoint and is itself pinned by central finite differences of the same replay.
//
// The replay.  Given a tape, the substep is a smooth function of the particle state (x, y, vx,
// vy, cux, cuy, cvx, cvy), of the faucet jet velocity given to the emitted particles
// (js/principal.js:301-321) and of the wall SDF:
//   frozen (read from the tape)  cell types incl. solid cells, the active bubbles of each cell,
//       the frozen pressure of bubbles without water faces, the volume-return part of each
//       bubble's right-hand side (fisica-cpu.js:445-450), every
//       particle's micro-step count, collision steps, velocity projections (vn < 0: a particle
//       sliding along a wall sits at vn = 0, where the projection has a kink) and removal
//       (fisica-cpu.js:595-620), and
//       the emitted particles' positions and count;
//   live  P2G with its weights and the affine term (:198-267), gravity (:271-291), surface
//       tension (the water fraction, its extension into walls and blurs, the normals, the
//       curvature and the air pressures: :306-403), the divergence with the density-excess and
//       air-pressure terms (:458-483), the bubble face sums with their air pressures, the
//       pressure solve (solved exactly, FP64 CG: the converged MGPCG of pressao-cpu.js:296-402
//       is the same A^-1), the gradient (:522-549), the extrapolation (:551-556, 660-678) and
//       G2P with its clamp, micro-steps, grid velocity lookups and the SDF push-out and
//       velocity projection (:560-640, cena.js:102-117).
// Foam and dye do not feed back into the motion and are not carried.
//
// The loss is L = (F - target)^2 with F the water fraction of a cell box: the mean over its
// cells of min(1, densidade / PPC) (the reference's own water fraction, fisica-cpu.js:305),
// densidade being the P2G density splat of the final particles.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace adj {

constexpr int PPC = 4;                  // constantes.js:44
constexpr int RES = 2;                  // :47
constexpr double SDF_MAX = 10;          // :50
constexpr double RAIO = 0.3;            // :56
constexpr double MICROPASSO = 0.8;      // :66
constexpr int MAX_MICRO = 4;            // :68
constexpr double EXC_MIN = -0.5, EXC_MAX = 1.0;  // :79-80
constexpr uint8_t SOLIDO = 0, AGUA = 1, AR = 2;

// flag word of one particle in one substep
constexpr int FL_MICRO = 3;        // micro-step count - 1
constexpr int FL_COLL = 2;         // bit 2 + k: the SDF push-out ran at micro-step k
constexpr int FL_PROJ = 6;         // bit 6 + k: ... and removed the velocity's inward normal part
constexpr int FL_KILL = 1 << 10;   // removed (deep inside a wall)

static inline int clampi(int v, int a, int b) { return v < a ? a : v > b ? b : v; }

struct Ctx {
  int nx, ny, C, nu, nv;
  double h, inv, gx, gy, tau, sigma;
  int mw, mh;
  const double* sdf;
  double* gsdf;  // SDF adjoint (may be null)
};

// ------------------------------------------------------------------ SDF lookups (cena.js)
// bilinear on the (2nx)x(2ny) lattice with the reference's clamps (cena.js:156-165); returns the
// value and its derivatives in fx, fy, and the four node weights
struct Bil {
  double v, dfx, dfy;
  int k;            // node (x0, y0)
  double w[4];      // weights of k, k+1, k+mw, k+mw+1
};
static Bil bil(const Ctx& c, double fx, double fy) {
  Bil b;
  const int w = c.mw, hh = c.mh;
  bool cx = false, cy = false;
  if (fx < 0) { fx = 0; cx = true; } else if (fx > w - 1) { fx = w - 1; cx = true; }
  if (fy < 0) { fy = 0; cy = true; } else if (fy > hh - 1) { fy = hh - 1; cy = true; }
  const int x0 = std::min((int)std::floor(fx), w - 2), y0 = std::min((int)std::floor(fy), hh - 2);
  const double tx = fx - x0, ty = fy - y0;
  b.k = y0 * w + x0;
  const double d00 = c.sdf[b.k], d10 = c.sdf[b.k + 1], d01 = c.sdf[b.k + w], d11 = c.sdf[b.k + w + 1];
  b.v = (d00 * (1 - tx) + d10 * tx) * (1 - ty) + (d01 * (1 - tx) + d11 * tx) * ty;
  b.dfx = cx ? 0 : (d10 - d00) * (1 - ty) + (d11 - d01) * ty;
  b.dfy = cy ? 0 : (d01 * (1 - tx) + d11 * tx) - (d00 * (1 - tx) + d10 * tx);
  b.w[0] = (1 - tx) * (1 - ty); b.w[1] = tx * (1 - ty); b.w[2] = (1 - tx) * ty; b.w[3] = tx * ty;
  return b;
}
static void bil_adj(const Ctx& c, const Bil& b, double vbar) {
  if (!c.gsdf || vbar == 0) return;
  c.gsdf[b.k] += vbar * b.w[0];
  c.gsdf[b.k + 1] += vbar * b.w[1];
  c.gsdf[b.k + c.mw] += vbar * b.w[2];
  c.gsdf[b.k + c.mw + 1] += vbar * b.w[3];
}

// distance and normal at (x, y) metres with everything the reverse needs
struct Wall {
  double fx, fy;
  Bil d, xp, xm, yp, ym;
  double gx, gy, g;   // gradient samples and their norm
  double nx, ny;
  bool deg;           // |g| <= 1e-9: the constant normal (0, 1)
};
static Wall wall_at(const Ctx& c, double x, double y) {
  Wall W;
  W.fx = (x / c.h) * RES - 0.5;
  W.fy = (y / c.h) * RES - 0.5;
  W.d = bil(c, W.fx, W.fy);
  W.xp = bil(c, W.fx + 1, W.fy); W.xm = bil(c, W.fx - 1, W.fy);
  W.yp = bil(c, W.fx, W.fy + 1); W.ym = bil(c, W.fx, W.fy - 1);
  W.gx = W.xp.v - W.xm.v;
  W.gy = W.yp.v - W.ym.v;
  W.g = std::hypot(W.gx, W.gy);
  W.deg = !(W.g > 1e-9);
  if (W.deg) { W.nx = 0; W.ny = 1; } else { W.nx = W.gx / W.g; W.ny = W.gy / W.g; }
  return W;
}
// adjoints of d and n -> x, y and the SDF
static void wall_adj(const Ctx& c, const Wall& W, double dbar, double nxbar, double nybar, double* xbar, double* ybar) {
  double fxbar = dbar * W.d.dfx, fybar = dbar * W.d.dfy;
  bil_adj(c, W.d, dbar);
  if (!W.deg) {
    const double dot = nxbar * W.nx + nybar * W.ny;
    const double gxbar = (nxbar - dot * W.nx) / W.g, gybar = (nybar - dot * W.ny) / W.g;
    fxbar += gxbar * (W.xp.dfx - W.xm.dfx) + gybar * (W.yp.dfx - W.ym.dfx);
    fybar += gxbar * (W.xp.dfy - W.xm.dfy) + gybar * (W.yp.dfy - W.ym.dfy);
    bil_adj(c, W.xp, gxbar); bil_adj(c, W.xm, -gxbar);
    bil_adj(c, W.yp, gybar); bil_adj(c, W.ym, -gybar);
  }
  *xbar += fxbar * RES / c.h;
  *ybar += fybar * RES / c.h;
}

// ------------------------------------------------------------------ MAC stencils
// u faces at (i h, (j+1/2) h); v faces at ((i+1/2) h, j h); cell centres at ((i+1/2) h, (j+1/2) h)
struct Sten {
  int i0, j0;
  double tx, ty;
};
static Sten sten_u(const Ctx& c, double x, double y) {
  const double fx = x * c.inv, fy = y * c.inv - 0.5;
  Sten s;
  s.i0 = clampi((int)std::floor(fx), 0, c.nx - 1);
  s.j0 = clampi((int)std::floor(fy), 0, c.ny - 2);
  s.tx = fx - s.i0; s.ty = fy - s.j0;
  return s;
}
static Sten sten_v(const Ctx& c, double x, double y) {
  const double fx = x * c.inv - 0.5, fy = y * c.inv;
  Sten s;
  s.i0 = clampi((int)std::floor(fx), 0, c.nx - 2);
  s.j0 = clampi((int)std::floor(fy), 0, c.ny - 1);
  s.tx = fx - s.i0; s.ty = fy - s.j0;
  return s;
}
static Sten sten_c(const Ctx& c, double x, double y) {
  const double fx = x * c.inv - 0.5, fy = y * c.inv - 0.5;
  Sten s;
  s.i0 = clampi((int)std::floor(fx), 0, c.nx - 2);
  s.j0 = clampi((int)std::floor(fy), 0, c.ny - 2);
  s.tx = fx - s.i0; s.ty = fy - s.j0;
  return s;
}
static inline double wa(double t, int a) { return a ? t : 1 - t; }
static inline double dwa(int a) { return a ? 1.0 : -1.0; }

// interpolation of a face field at (x, y): value, and the affine terms of G2P
struct Interp {
  double v, cx, cy;  // sum w f, sum s_a w_y f / h, sum w_x s_b f / h
};
static Interp interp(const Ctx& c, const Sten& s, const double* f, int w) {
  Interp r{0, 0, 0};
  for (int b = 0; b < 2; b++)
    for (int a = 0; a < 2; a++) {
      const double val = f[(s.i0 + a) + (s.j0 + b) * w];
      const double wx = wa(s.tx, a), wy = wa(s.ty, b);
      r.v += wx * wy * val;
      r.cx += dwa(a) * wy * c.inv * val;
      r.cy += wx * dwa(b) * c.inv * val;
    }
  return r;
}
// reverse of interp: adjoints of (v, cx, cy) -> field (scatter) and x, y
static void interp_adj(const Ctx& c, const Sten& s, const double* f, int w, double vbar, double cxbar, double cybar,
                       double* fbar, double* xbar, double* ybar) {
  for (int b = 0; b < 2; b++)
    for (int a = 0; a < 2; a++) {
      const int k = (s.i0 + a) + (s.j0 + b) * w;
      const double val = f[k];
      const double wx = wa(s.tx, a), wy = wa(s.ty, b);
      fbar[k] += vbar * wx * wy + cxbar * dwa(a) * wy * c.inv + cybar * wx * dwa(b) * c.inv;
      // d/dtx and d/dty (dtx/dx = dty/dy = 1/h)
      const double dtx = vbar * dwa(a) * wy * val + cybar * dwa(a) * dwa(b) * c.inv * val;
      const double dty = vbar * wx * dwa(b) * val + cxbar * dwa(a) * dwa(b) * c.inv * val;
      *xbar += dtx * c.inv;
      *ybar += dty * c.inv;
    }
}

// ------------------------------------------------------------------ G2P of one particle
struct Micro {
  double x0, y0;        // position before the move
  bool grid;            // velocity from the grid (micro > 1)
  Sten su, sv;          // its stencils
  double gvx, gvy;      // grid velocity before the clamp
  bool gclamp;
  double wx, wy;        // velocity of the move
  bool coll;
  double x1, y1;        // position after the move
  Wall W;
  double vx0, vy0;      // particle velocity before the projection
  bool proj;            // vn < 0
};
struct G2P {
  Sten su, sv;
  Interp iu, iv;
  bool clamp;
  double vel;
  int m;
  Micro mi[MAX_MICRO];
  bool kill;
  double out[8];
};

// forward; in record mode the micro count, collision steps and removal are decided here and
// written to *flag, otherwise they are read from it
static void g2p_fwd(const Ctx& c, const double* P, const double* U, const double* V, double dt, uint16_t* flag,
                    bool record, G2P& g) {
  double x = P[0], y = P[1];
  g.su = sten_u(c, x, y);
  g.sv = sten_v(c, x, y);
  g.iu = interp(c, g.su, U, c.nx + 1);
  g.iv = interp(c, g.sv, V, c.nx);
  double vx = g.iu.v, vy = g.iv.v;
  const double vmax = (MICROPASSO * MAX_MICRO * c.h) / dt;
  g.vel = std::hypot(vx, vy);
  g.clamp = g.vel > vmax;
  if (g.clamp) { vx *= vmax / g.vel; vy *= vmax / g.vel; }
  if (record) {
    g.m = (int)std::min((double)MAX_MICRO, std::max(1.0, std::ceil((std::min(g.vel, vmax) * dt) / (MICROPASSO * c.h))));
    *flag = (uint16_t)(g.m - 1);
  } else {
    g.m = (*flag & FL_MICRO) + 1;
  }
  g.kill = !record && (*flag & FL_KILL);
  if (g.kill) return;
  const double dm = dt / g.m;
  for (int k = 0; k < g.m; k++) {
    Micro& M = g.mi[k];
    M.x0 = x; M.y0 = y;
    M.grid = g.m > 1;
    double wx = vx, wy = vy;
    if (M.grid) {
      M.su = sten_u(c, x, y);
      M.sv = sten_v(c, x, y);
      M.gvx = interp(c, M.su, U, c.nx + 1).v;
      M.gvy = interp(c, M.sv, V, c.nx).v;
      const double gv = std::hypot(M.gvx, M.gvy);
      M.gclamp = gv > vmax;
      wx = M.gclamp ? M.gvx * vmax / gv : M.gvx;
      wy = M.gclamp ? M.gvy * vmax / gv : M.gvy;
    }
    M.wx = wx; M.wy = wy;
    x += wx * dm;
    y += wy * dm;
    M.x1 = x; M.y1 = y;
    if (record) {
      const double d = bil(c, (x / c.h) * RES - 0.5, (y / c.h) * RES - 0.5).v;
      M.coll = d < RAIO;
      if (M.coll) {
        *flag |= (uint16_t)(1 << (FL_COLL + k));
        if (d < -SDF_MAX + 0.5) { *flag |= FL_KILL; g.kill = true; return; }
      }
    } else {
      M.coll = (*flag >> (FL_COLL + k)) & 1;
    }
    M.vx0 = vx; M.vy0 = vy;
    M.proj = false;
    if (M.coll) {
      M.W = wall_at(c, x, y);
      const double push = (RAIO - M.W.d.v) * c.h;
      x += M.W.nx * push;
      y += M.W.ny * push;
      const double vn = vx * M.W.nx + vy * M.W.ny;
      if (record) {
        M.proj = vn < 0;
        if (M.proj) *flag |= (uint16_t)(1 << (FL_PROJ + k));
      } else {
        M.proj = (*flag >> (FL_PROJ + k)) & 1;
      }
      if (M.proj) { vx -= vn * M.W.nx; vy -= vn * M.W.ny; }
    }
  }
  g.out[0] = x; g.out[1] = y; g.out[2] = vx; g.out[3] = vy;
  g.out[4] = g.iu.cx; g.out[5] = g.iu.cy; g.out[6] = g.iv.cx; g.out[7] = g.iv.cy;
}

// reverse of g2p_fwd: ob = adjoint of the 8 outputs; adds to xbar[2] of the start position,
// the face adjoints Ub/Vb and the SDF adjoint
static void g2p_adj(const Ctx& c, const double* U, const double* V, double dt, const G2P& g, const double* ob,
                    double* Ub, double* Vb, double* xb, double* yb) {
  const double vmax = (MICROPASSO * MAX_MICRO * c.h) / dt;
  const double dm = dt / g.m;
  double xbar = ob[0], ybar = ob[1], vxbar = ob[2], vybar = ob[3];
  for (int k = g.m - 1; k >= 0; k--) {
    const Micro& M = g.mi[k];
    if (M.coll) {
      const Wall& W = M.W;
      double nxbar = 0, nybar = 0;
      // velocity projection v' = v - (v.n) n
      if (M.proj) {
        const double vn = M.vx0 * W.nx + M.vy0 * W.ny;
        const double bn = vxbar * W.nx + vybar * W.ny;
        nxbar -= vn * vxbar + bn * M.vx0;
        nybar -= vn * vybar + bn * M.vy0;
        vxbar -= bn * W.nx;
        vybar -= bn * W.ny;
      }
      // push-out x' = x1 + n (RAIO - d) h
      const double push = (RAIO - W.d.v) * c.h;
      nxbar += xbar * push;
      nybar += ybar * push;
      const double pbar = xbar * W.nx + ybar * W.ny;
      const double dbar = -c.h * pbar;
      wall_adj(c, W, dbar, nxbar, nybar, &xbar, &ybar);
    }
    // the move x1 = x0 + w dm
    const double wxbar = xbar * dm, wybar = ybar * dm;
    if (M.grid) {
      double gxbar = wxbar, gybar = wybar;
      if (M.gclamp) {
        const double gv = std::hypot(M.gvx, M.gvy);
        const double ux = M.gvx / gv, uy = M.gvy / gv;
        const double dot = wxbar * ux + wybar * uy;
        gxbar = (vmax / gv) * (wxbar - dot * ux);
        gybar = (vmax / gv) * (wybar - dot * uy);
      }
      interp_adj(c, M.su, U, c.nx + 1, gxbar, 0, 0, Ub, &xbar, &ybar);
      interp_adj(c, M.sv, V, c.nx, gybar, 0, 0, Vb, &xbar, &ybar);
    } else {
      vxbar += wxbar;
      vybar += wybar;
    }
  }
  // the G2P clamp
  if (g.clamp) {
    const double ux = g.iu.v / g.vel, uy = g.iv.v / g.vel;
    const double dot = vxbar * ux + vybar * uy;
    const double s = vmax / g.vel;
    vxbar = s * (vxbar - dot * ux);
    vybar = s * (vybar - dot * uy);
  }
  interp_adj(c, g.su, U, c.nx + 1, vxbar, ob[4], ob[5], Ub, &xbar, &ybar);
  interp_adj(c, g.sv, V, c.nx, vybar, ob[6], ob[7], Vb, &xbar, &ybar);
  *xb += xbar;
  *yb += ybar;
}

// ------------------------------------------------------------------ P2G
struct P2GOut {
  std::vector<double> Un, W, Vn, VW, dens;
};
static void p2g_fwd(const Ctx& c, const double* P, int n, P2GOut& o) {
  o.Un.assign(c.nu, 0); o.W.assign(c.nu, 0); o.Vn.assign(c.nv, 0); o.VW.assign(c.nv, 0); o.dens.assign(c.C, 0);
  for (int p = 0; p < n; p++) {
    const double* s = P + (size_t)p * 8;
    const double x = s[0], y = s[1];
    {
      Sten st = sten_u(c, x, y);
      for (int b = 0; b < 2; b++)
        for (int a = 0; a < 2; a++) {
          const double w = wa(st.tx, a) * wa(st.ty, b);
          const double dx = (st.i0 + a) * c.h - x, dy = (st.j0 + b + 0.5) * c.h - y;
          const int k = st.i0 + a + (st.j0 + b) * (c.nx + 1);
          o.Un[k] += w * (s[2] + s[4] * dx + s[5] * dy);
          o.W[k] += w;
        }
    }
    {
      Sten st = sten_v(c, x, y);
      for (int b = 0; b < 2; b++)
        for (int a = 0; a < 2; a++) {
          const double w = wa(st.tx, a) * wa(st.ty, b);
          const double dx = (st.i0 + a + 0.5) * c.h - x, dy = (st.j0 + b) * c.h - y;
          const int k = st.i0 + a + (st.j0 + b) * c.nx;
          o.Vn[k] += w * (s[3] + s[6] * dx + s[7] * dy);
          o.VW[k] += w;
        }
    }
    {
      Sten st = sten_c(c, x, y);
      for (int b = 0; b < 2; b++)
        for (int a = 0; a < 2; a++) o.dens[st.i0 + a + (st.j0 + b) * c.nx] += wa(st.tx, a) * wa(st.ty, b);
    }
  }
}
// reverse: per face, ua = adjoint of sum(w val), uw = adjoint of sum(w); per cell, db = adjoint of dens
static void p2g_adj_particle(const Ctx& c, const double* s, const double* ua, const double* uw, const double* va,
                             const double* vw, const double* db, double* pb) {
  const double x = s[0], y = s[1];
  {
    Sten st = sten_u(c, x, y);
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++) {
        const int k = st.i0 + a + (st.j0 + b) * (c.nx + 1);
        const double wx = wa(st.tx, a), wy = wa(st.ty, b), w = wx * wy;
        const double dx = (st.i0 + a) * c.h - x, dy = (st.j0 + b + 0.5) * c.h - y;
        const double val = s[2] + s[4] * dx + s[5] * dy;
        const double g = ua[k] * val + uw[k];
        pb[0] += g * dwa(a) * wy * c.inv - w * ua[k] * s[4];
        pb[1] += g * wx * dwa(b) * c.inv - w * ua[k] * s[5];
        pb[2] += w * ua[k];
        pb[4] += w * ua[k] * dx;
        pb[5] += w * ua[k] * dy;
      }
  }
  {
    Sten st = sten_v(c, x, y);
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++) {
        const int k = st.i0 + a + (st.j0 + b) * c.nx;
        const double wx = wa(st.tx, a), wy = wa(st.ty, b), w = wx * wy;
        const double dx = (st.i0 + a + 0.5) * c.h - x, dy = (st.j0 + b) * c.h - y;
        const double val = s[3] + s[6] * dx + s[7] * dy;
        const double g = va[k] * val + vw[k];
        pb[0] += g * dwa(a) * wy * c.inv - w * va[k] * s[6];
        pb[1] += g * wx * dwa(b) * c.inv - w * va[k] * s[7];
        pb[3] += w * va[k];
        pb[6] += w * va[k] * dx;
        pb[7] += w * va[k] * dy;
      }
  }
  if (db) {
    Sten st = sten_c(c, x, y);
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++) {
        const double g = db[st.i0 + a + (st.j0 + b) * c.nx];
        pb[0] += g * dwa(a) * wa(st.ty, b) * c.inv;
        pb[1] += g * wa(st.tx, a) * dwa(b) * c.inv;
      }
  }
}

// ------------------------------------------------------------------ pressure system
// unknowns: q of water cells, y of the active bubbles; A as resolverPCG builds it
// (pressao-cpu.js:84-93, 271-288): water rows cnt q - sum over free neighbours (water q, active
// bubble y, 0 otherwise); bubble rows faces y - sum over its (bubble cell, water cell) pairs q.
struct System {
  const Ctx* c;
  const uint8_t* tipo;
  const int32_t* bub;
  int nb;
  std::vector<int> faces;
  std::vector<std::pair<int, int>> pairs;  // (bubble, water cell)
  void build() {
    faces.assign(nb, 0);
    pairs.clear();
    const int nx = c->nx, ny = c->ny;
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        const int a = i + j * nx;
        if (tipo[a] != AR || bub[a] < 0) continue;
        const int nbh[4] = {a - 1, a + 1, a - nx, a + nx};
        for (int q = 0; q < 4; q++)
          if (tipo[nbh[q]] == AGUA) { pairs.emplace_back(bub[a], nbh[q]); faces[bub[a]]++; }
      }
  }
  double val(const std::vector<double>& z, int d) const {
    if (tipo[d] == AGUA) return z[d];
    if (tipo[d] == AR && bub[d] >= 0) return z[c->C + bub[d]];
    return 0;
  }
  void apply(const std::vector<double>& z, std::vector<double>& out) const {
    const int nx = c->nx, ny = c->ny, C = c->C;
    std::fill(out.begin(), out.end(), 0.0);
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int k = i + j * nx;
        if (tipo[k] != AGUA) continue;
        double s = 0;
        int cnt = 0;
        if (i > 0 && tipo[k - 1] != SOLIDO) { s += val(z, k - 1); cnt++; }
        if (i < nx - 1 && tipo[k + 1] != SOLIDO) { s += val(z, k + 1); cnt++; }
        if (j > 0 && tipo[k - nx] != SOLIDO) { s += val(z, k - nx); cnt++; }
        if (j < ny - 1 && tipo[k + nx] != SOLIDO) { s += val(z, k + nx); cnt++; }
        out[k] = cnt * z[k] - s;
      }
    for (int b = 0; b < nb; b++) out[C + b] = faces[b] * z[C + b];
    for (auto& pr : pairs) out[C + pr.first] -= z[pr.second];
  }
  double dot(const std::vector<double>& a, const std::vector<double>& b) const {
    double s = 0;
    for (int k = 0; k < c->C; k++) if (tipo[k] == AGUA) s += a[k] * b[k];
    for (int q = 0; q < nb; q++) s += a[c->C + q] * b[c->C + q];
    return s;
  }
  // plain CG to a relative residual; returns 0, or 1 if it did not converge
  int solve(const std::vector<double>& rhs, std::vector<double>& x, double tol) const {
    const size_t N = (size_t)c->C + nb;
    x.assign(N, 0.0);
    std::vector<double> r(rhs), p(N), ap(N);
    for (int k = 0; k < c->C; k++) if (tipo[k] != AGUA) r[k] = 0;
    const double bn = std::sqrt(dot(r, r));
    if (bn == 0) return 0;
    p = r;
    double rr = dot(r, r);
    for (int it = 0; it < 40000; it++) {
      apply(p, ap);
      const double pap = dot(p, ap);
      if (!(pap > 0)) return 1;
      const double al = rr / pap;
      for (size_t k = 0; k < N; k++) { x[k] += al * p[k]; r[k] -= al * ap[k]; }
      const double rr2 = dot(r, r);
      if (std::sqrt(rr2) <= tol * bn) return 0;
      const double be = rr2 / rr;
      rr = rr2;
      for (size_t k = 0; k < N; k++) p[k] = r[k] + be * p[k];
    }
    return 1;
  }
};

// ------------------------------------------------------------------ grid phases
// face (i, j) of the u grid: solid or domain-boundary neighbour (gravidadeEFronteira's rule)
static bool u_wall(const Ctx& c, const uint8_t* t, int i, int j) {
  return i == 0 || i == c.nx || t[i - 1 + j * c.nx] == SOLIDO || t[i + j * c.nx] == SOLIDO;
}
static bool v_wall(const Ctx& c, const uint8_t* t, int i, int j) {
  return j == 0 || j == c.ny || t[i + (j - 1) * c.nx] == SOLIDO || t[i + j * c.nx] == SOLIDO;
}
// gradient classes of a face: 0 solid (zeroed), 1 water-adjacent (q difference), 2 air-air (kept, invalid)
static int u_class(const Ctx& c, const uint8_t* t, int i, int j) {
  const uint8_t tl = i > 0 ? t[i - 1 + j * c.nx] : SOLIDO, tr = i < c.nx ? t[i + j * c.nx] : SOLIDO;
  if (tl == SOLIDO || tr == SOLIDO) return 0;
  return (tl == AGUA || tr == AGUA) ? 1 : 2;
}
static int v_class(const Ctx& c, const uint8_t* t, int i, int j) {
  const uint8_t tb = j > 0 ? t[i + (j - 1) * c.nx] : SOLIDO, tt = j < c.ny ? t[i + j * c.nx] : SOLIDO;
  if (tb == SOLIDO || tt == SOLIDO) return 0;
  return (tb == AGUA || tt == AGUA) ? 1 : 2;
}
static bool miolo(const Ctx& c, const uint8_t* t, int i, int j) {
  if (i < 1 || j < 1 || i >= c.nx - 1 || j >= c.ny - 1) return false;
  for (int b = -1; b <= 1; b++)
    for (int a = -1; a <= 1; a++)
      if (t[i + a + (j + b) * c.nx] != AGUA) return false;
  return true;
}
// density-excess term of a water cell: value and d/d(dens) (fisica-cpu.js:465-470)
static double excess(const Ctx& c, const uint8_t* t, int i, int j, double dens, double* dd) {
  double e = dens / PPC - 1;
  *dd = 1.0 / PPC;
  if (e < 0 && !miolo(c, t, i, j)) { e = 0; *dd = 0; }
  if (e < EXC_MIN) { e = EXC_MIN; *dd = 0; }
  if (e > EXC_MAX) { e = EXC_MAX; *dd = 0; }
  return e;
}

// one extrapolation layer (extrapolarComponente, fisica-cpu.js:660-678); fresh[k] = neighbour count
// of a newly filled entry, 0 otherwise
static void extrap_layer(std::vector<double>& a, std::vector<uint8_t>& val, int w, int hh, std::vector<int>& fresh) {
  std::vector<double> tmp(a);
  fresh.assign(a.size(), 0);
  for (int j = 0; j < hh; j++)
    for (int i = 0; i < w; i++) {
      const int k = i + j * w;
      if (val[k]) continue;
      double s = 0;
      int n = 0;
      if (i > 0 && val[k - 1]) { s += a[k - 1]; n++; }
      if (i < w - 1 && val[k + 1]) { s += a[k + 1]; n++; }
      if (j > 0 && val[k - w]) { s += a[k - w]; n++; }
      if (j < hh - 1 && val[k + w]) { s += a[k + w]; n++; }
      if (n > 0) { tmp[k] = s / n; fresh[k] = n; }
    }
  a = tmp;
  for (size_t k = 0; k < a.size(); k++) if (fresh[k]) val[k] = 1;
}
// reverse of one layer: val = validity BEFORE the layer
static void extrap_layer_adj(std::vector<double>& ab, const std::vector<uint8_t>& val, int w, int hh,
                             const std::vector<int>& fresh) {
  std::vector<double> out(ab);
  for (int j = 0; j < hh; j++)
    for (int i = 0; i < w; i++) {
      const int k = i + j * w;
      if (!fresh[k]) continue;
      const double g = ab[k] / fresh[k];
      if (i > 0 && val[k - 1]) out[k - 1] += g;
      if (i < w - 1 && val[k + 1]) out[k + 1] += g;
      if (j > 0 && val[k - w]) out[k - w] += g;
      if (j < hh - 1 && val[k + w]) out[k + w] += g;
      out[k] = 0;
    }
  ab = out;
}

// ------------------------------------------------------------------ surface tension (fisica-cpu.js:306-403)
// The chain water fraction -> wall extension (3 layers) -> blur x, blur y -> extension -> blur x,
// blur y -> extension -> normals -> curvature -> air pressures, every step linear in its input
// except the normal's normalisation and the pressure clamp; the types (solid cells, water, air)
// are held.
struct Tension {
  std::vector<std::vector<double>> fs;      // the fraction before each step of the chain, and after the last
  std::vector<int> ops;                     // 0 extension layer, 1 blur x, 2 blur y
  std::vector<std::vector<uint8_t>> full;   // extension layers: the cells filled before the layer
  std::vector<double> nx, ny, g, kap;
  std::vector<uint8_t> clamped;             // air pressure clamped to +-1/h (no derivative)
  std::vector<double> qar;
};

static void ext_layer(const Ctx& c, const std::vector<double>& f, std::vector<uint8_t>& full, std::vector<double>& out) {
  const int nx = c.nx, ny = c.ny;
  out = f;
  std::vector<int> fresh;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = i + j * nx;
      if (full[k]) continue;
      double sum = 0;
      int n = 0;
      if (i > 0 && full[k - 1]) { sum += f[k - 1]; n++; }
      if (i < nx - 1 && full[k + 1]) { sum += f[k + 1]; n++; }
      if (j > 0 && full[k - nx]) { sum += f[k - nx]; n++; }
      if (j < ny - 1 && full[k + nx]) { sum += f[k + nx]; n++; }
      if (n) { out[k] = sum / n; fresh.push_back(k); }
    }
  for (int k : fresh) full[k] = 1;
}

static void tension_fwd(const Ctx& c, const uint8_t* t, const std::vector<double>& dens, double sigma, double dt,
                        Tension& T) {
  const int nx = c.nx, ny = c.ny, C = c.C;
  std::vector<double> f(C);
  for (int k = 0; k < C; k++) f[k] = t[k] == SOLIDO ? 0 : std::min(1.0, dens[k] / PPC);
  T.fs.clear(); T.ops.clear(); T.full.clear();
  for (int pass = 0; pass <= 2; pass++) {
    std::vector<uint8_t> full(C);
    for (int k = 0; k < C; k++) full[k] = t[k] == SOLIDO ? 0 : 1;
    for (int layer = 0; layer < 3; layer++) {
      T.fs.push_back(f); T.ops.push_back(0); T.full.push_back(full);
      std::vector<double> o;
      ext_layer(c, f, full, o);
      f.swap(o);
    }
    if (pass == 2) break;
    for (int axis = 0; axis < 2; axis++) {
      T.fs.push_back(f); T.ops.push_back(1 + axis); T.full.emplace_back();
      std::vector<double> o(C);
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          const int k = i + j * nx;
          int a, b;
          if (axis == 0) { a = i > 0 ? k - 1 : k; b = i < nx - 1 ? k + 1 : k; }
          else { a = j > 0 ? k - nx : k; b = j < ny - 1 ? k + nx : k; }
          o[k] = (f[a] + 2 * f[k] + f[b]) * 0.25;
        }
      f.swap(o);
    }
  }
  T.fs.push_back(f);
  T.nx.assign(C, 0); T.ny.assign(C, 0); T.g.assign(C, 0); T.kap.assign(C, 0);
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const int k = i + j * nx;
      const double gx = (f[k + 1] - f[k - 1]) * 0.5, gy = (f[k + nx] - f[k - nx]) * 0.5;
      const double g = std::hypot(gx, gy);
      T.g[k] = g;
      if (g > 1e-3) { T.nx[k] = gx / g; T.ny[k] = gy / g; }
    }
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const int k = i + j * nx;
      T.kap[k] = -((T.nx[k + 1] - T.nx[k - 1]) + (T.ny[k + nx] - T.ny[k - nx])) / (2 * c.h);
    }
  const double escala = dt / (1000.0 * c.h), kmax = 1 / c.h;
  T.qar.assign(C, 0);
  T.clamped.assign(C, 1);
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const int k = i + j * nx;
      if (t[k] != AR || !(sigma > 0)) continue;
      double sum = T.kap[k];
      int n = 1;
      const int nb[4] = {k - 1, k + 1, k - nx, k + nx};
      for (int q = 0; q < 4; q++)
        if (t[nb[q]] == AGUA) { sum += T.kap[nb[q]]; n++; }
      if (n > 1) {
        const double m = sum / n;
        T.clamped[k] = !(m > -kmax && m < kmax);
        T.qar[k] = sigma * std::max(-kmax, std::min(kmax, m)) * escala;
      }
    }
}

// qbar (per cell) -> the density adjoint, added to db
static void tension_adj(const Ctx& c, const uint8_t* t, const std::vector<double>& dens, const Tension& T, double sigma,
                        double dt, const std::vector<double>& qbar, std::vector<double>& db) {
  const int nx = c.nx, ny = c.ny, C = c.C;
  const double escala = dt / (1000.0 * c.h);
  std::vector<double> kb(C, 0.0), nxb(C, 0.0), nyb(C, 0.0), fb(C, 0.0);
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const int k = i + j * nx;
      if (t[k] != AR || !(sigma > 0) || T.clamped[k] || qbar[k] == 0) continue;
      const int nb[4] = {k - 1, k + 1, k - nx, k + nx};
      int n = 1;
      for (int q = 0; q < 4; q++) n += t[nb[q]] == AGUA;
      const double g = qbar[k] * sigma * escala / n;
      kb[k] += g;
      for (int q = 0; q < 4; q++)
        if (t[nb[q]] == AGUA) kb[nb[q]] += g;
    }
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const int k = i + j * nx;
      const double g = kb[k] / (2 * c.h);
      nxb[k + 1] -= g; nxb[k - 1] += g; nyb[k + nx] -= g; nyb[k - nx] += g;
    }
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const int k = i + j * nx;
      if (!(T.g[k] > 1e-3)) continue;
      const double dot = nxb[k] * T.nx[k] + nyb[k] * T.ny[k];
      const double gxb = (nxb[k] - dot * T.nx[k]) / T.g[k], gyb = (nyb[k] - dot * T.ny[k]) / T.g[k];
      fb[k + 1] += 0.5 * gxb; fb[k - 1] -= 0.5 * gxb;
      fb[k + nx] += 0.5 * gyb; fb[k - nx] -= 0.5 * gyb;
    }
  for (int s = (int)T.ops.size() - 1; s >= 0; s--) {
    std::vector<double> o(C, 0.0);
    if (T.ops[s] == 0) {
      const std::vector<uint8_t>& full = T.full[s];
      o = fb;
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          const int k = i + j * nx;
          if (full[k]) continue;
          int n = (i > 0 && full[k - 1]) + (i < nx - 1 && full[k + 1]) + (j > 0 && full[k - nx]) + (j < ny - 1 && full[k + nx]);
          if (!n) continue;
          const double g = fb[k] / n;
          if (i > 0 && full[k - 1]) o[k - 1] += g;
          if (i < nx - 1 && full[k + 1]) o[k + 1] += g;
          if (j > 0 && full[k - nx]) o[k - nx] += g;
          if (j < ny - 1 && full[k + nx]) o[k + nx] += g;
          o[k] -= fb[k];
        }
    } else {
      const int axis = T.ops[s] - 1;
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          const int k = i + j * nx;
          int a, b;
          if (axis == 0) { a = i > 0 ? k - 1 : k; b = i < nx - 1 ? k + 1 : k; }
          else { a = j > 0 ? k - nx : k; b = j < ny - 1 ? k + nx : k; }
          o[a] += 0.25 * fb[k]; o[k] += 0.5 * fb[k]; o[b] += 0.25 * fb[k];
        }
    }
    fb.swap(o);
  }
  for (int k = 0; k < C; k++)
    if (t[k] != SOLIDO && dens[k] / PPC < 1) db[k] += fb[k] / PPC;
}

// ------------------------------------------------------------------ one substep: record and reverse
struct Sub {
  double dt;
  const uint8_t* tipo;
  const int32_t* bub;
  const double* qextra;
  int nb;
  const double* rhsfix;
  uint16_t* flags;
  int n;                                 // particles at the start
  std::vector<double> P;                 // their state
  P2GOut pg;
  std::vector<double> U0, V0;            // P2G velocities (sum w val / sum w)
  std::vector<double> U3, V3;            // after gradient and extrapolation
  std::vector<uint8_t> uval0, vval0, uval1, vval1;  // validity before extrapolation layers 1, 2
  std::vector<int> ufr[2], vfr[2];
  System sys;
  Tension tn;
};

static int substep_fwd(const Ctx& c, Sub& S, std::vector<double>& P, bool record, double cg_tol) {
  const int nx = c.nx, ny = c.ny, C = c.C;
  const uint8_t* t = S.tipo;
  S.n = (int)(P.size() / 8);
  S.P = P;
  p2g_fwd(c, P.data(), S.n, S.pg);
  std::vector<double> U(c.nu), V(c.nv);
  S.U0.assign(c.nu, 0); S.V0.assign(c.nv, 0);
  for (int k = 0; k < c.nu; k++) if (S.pg.W[k] > 0) S.U0[k] = S.pg.Un[k] / S.pg.W[k];
  for (int k = 0; k < c.nv; k++) if (S.pg.VW[k] > 0) S.V0[k] = S.pg.Vn[k] / S.pg.VW[k];
  // gravity and walls (fisica-cpu.js:271-291)
  for (int j = 0; j < ny; j++)
    for (int i = 0; i <= nx; i++) {
      const int k = i + j * (nx + 1);
      U[k] = S.U0[k] + (S.pg.W[k] > 0 ? c.gx * S.dt : 0);
      if (u_wall(c, t, i, j)) U[k] = 0;
    }
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = i + j * nx;
      V[k] = S.V0[k] + (S.pg.VW[k] > 0 ? c.gy * S.dt : 0);
      if (v_wall(c, t, i, j)) V[k] = 0;
    }
  tension_fwd(c, t, S.pg.dens, c.sigma, S.dt, S.tn);
  const std::vector<double>& qar = S.tn.qar;
  // right-hand side [-div ; bubble rhs]
  const double alvo = c.h / std::max(c.tau, S.dt);
  S.sys.c = &c; S.sys.tipo = t; S.sys.bub = S.bub; S.sys.nb = S.nb;
  S.sys.build();
  std::vector<double> rhs((size_t)C + S.nb, 0.0);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = i + j * nx;
      if (t[k] != AGUA) continue;
      const double d = U[i + 1 + j * (nx + 1)] - U[i + j * (nx + 1)] + V[i + (j + 1) * nx] - V[i + j * nx];
      double dd;
      const double e = excess(c, t, i, j, S.pg.dens[k], &dd);
      double ar = 0;
      if (t[k - 1] == AR) ar += qar[k - 1];
      if (t[k + 1] == AR) ar += qar[k + 1];
      if (t[k - nx] == AR) ar += qar[k - nx];
      if (t[k + nx] == AR) ar += qar[k + nx];
      rhs[k] = -(d - alvo * e - ar);
    }
  for (int b = 0; b < S.nb; b++) rhs[C + b] = S.rhsfix[b];
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const int a = i + j * nx;
      if (t[a] != AR || S.bub[a] < 0) continue;
      double& r = rhs[C + S.bub[a]];
      if (t[a - 1] == AGUA) r += U[i + j * (nx + 1)] - qar[a];
      if (t[a + 1] == AGUA) r -= U[i + 1 + j * (nx + 1)] + qar[a];
      if (t[a - nx] == AGUA) r += V[i + j * nx] - qar[a];
      if (t[a + nx] == AGUA) r -= V[i + (j + 1) * nx] + qar[a];
    }
  std::vector<double> z;
  if (S.sys.solve(rhs, z, cg_tol)) return 2;
  std::vector<double> q(C, 0.0);
  for (int k = 0; k < C; k++) {
    if (t[k] == AGUA) q[k] = z[k];
    else if (t[k] == AR) q[k] = qar[k] + S.qextra[k] + (S.bub[k] >= 0 ? z[C + S.bub[k]] : 0);
  }
  // gradient (fisica-cpu.js:522-549)
  S.uval0.assign(c.nu, 0); S.vval0.assign(c.nv, 0);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i <= nx; i++) {
      const int k = i + j * (nx + 1);
      const int cl = u_class(c, t, i, j);
      if (cl == 0) { U[k] = 0; S.uval0[k] = 1; }
      else if (cl == 1) { U[k] -= q[i + j * nx] - q[i - 1 + j * nx]; S.uval0[k] = 1; }
    }
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = i + j * nx;
      const int cl = v_class(c, t, i, j);
      if (cl == 0) { V[k] = 0; S.vval0[k] = 1; }
      else if (cl == 1) { V[k] -= q[i + j * nx] - q[i + (j - 1) * nx]; S.vval0[k] = 1; }
    }
  // extrapolation, two layers (fisica-cpu.js:551-556)
  {
    std::vector<uint8_t> uv = S.uval0, vv = S.vval0;
    extrap_layer(U, uv, nx + 1, ny, S.ufr[0]);
    extrap_layer(V, vv, nx, ny + 1, S.vfr[0]);
    S.uval1 = uv; S.vval1 = vv;
    extrap_layer(U, uv, nx + 1, ny, S.ufr[1]);
    extrap_layer(V, vv, nx, ny + 1, S.vfr[1]);
  }
  S.U3 = U; S.V3 = V;
  // G2P (fisica-cpu.js:560-640); removed particles leave the list in order
  std::vector<double> out;
  out.reserve(P.size());
  for (int p = 0; p < S.n; p++) {
    G2P g;
    g2p_fwd(c, &S.P[(size_t)p * 8], S.U3.data(), S.V3.data(), S.dt, &S.flags[p], record, g);
    if (g.kill) continue;
    out.insert(out.end(), g.out, g.out + 8);
  }
  P.swap(out);
  return 0;
}

// Pb: adjoint of the state after the substep (surviving particles, in order); on return, the
// adjoint of the state at its start
static int substep_adj(const Ctx& c, Sub& S, std::vector<double>& Pb, double cg_tol) {
  const int nx = c.nx, ny = c.ny, C = c.C;
  const uint8_t* t = S.tipo;
  std::vector<double> Ub(c.nu, 0.0), Vb(c.nv, 0.0);
  std::vector<double> Pstart((size_t)S.n * 8, 0.0);
  int o = 0;
  for (int p = 0; p < S.n; p++) {
    G2P g;
    uint16_t fl = S.flags[p];
    g2p_fwd(c, &S.P[(size_t)p * 8], S.U3.data(), S.V3.data(), S.dt, &fl, false, g);
    if (g.kill) continue;
    g2p_adj(c, S.U3.data(), S.V3.data(), S.dt, g, &Pb[(size_t)o * 8], Ub.data(), Vb.data(), &Pstart[(size_t)p * 8],
            &Pstart[(size_t)p * 8 + 1]);
    o++;
  }
  // extrapolation, layers in reverse
  extrap_layer_adj(Ub, S.uval1, nx + 1, ny, S.ufr[1]);
  extrap_layer_adj(Vb, S.vval1, nx, ny + 1, S.vfr[1]);
  extrap_layer_adj(Ub, S.uval0, nx + 1, ny, S.ufr[0]);
  extrap_layer_adj(Vb, S.vval0, nx, ny + 1, S.vfr[0]);
  // gradient
  std::vector<double> qb(C, 0.0);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i <= nx; i++) {
      const int k = i + j * (nx + 1);
      const int cl = u_class(c, t, i, j);
      if (cl == 0) Ub[k] = 0;
      else if (cl == 1) { qb[i + j * nx] -= Ub[k]; qb[i - 1 + j * nx] += Ub[k]; }
    }
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = i + j * nx;
      const int cl = v_class(c, t, i, j);
      if (cl == 0) Vb[k] = 0;
      else if (cl == 1) { qb[i + j * nx] -= Vb[k]; qb[i + (j - 1) * nx] += Vb[k]; }
    }
  // q -> solve: s = A^-1 [qbar_water ; ybar]; rhs adjoint = s
  std::vector<double> rb((size_t)C + S.nb, 0.0);
  for (int k = 0; k < C; k++) {
    if (t[k] == AGUA) rb[k] = qb[k];
    else if (t[k] == AR && S.bub[k] >= 0) rb[C + S.bub[k]] += qb[k];
  }
  std::vector<double> s;
  if (S.sys.solve(rb, s, cg_tol)) return 2;
  // air pressures: the air cells' q (#pressao), the water cells' div (air neighbours), the
  // bubbles' face terms
  std::vector<double> qarb(C, 0.0);
  for (int k = 0; k < C; k++) if (t[k] == AR) qarb[k] += qb[k];
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = i + j * nx;
      if (t[k] != AGUA) continue;
      if (t[k - 1] == AR) qarb[k - 1] += s[k];
      if (t[k + 1] == AR) qarb[k + 1] += s[k];
      if (t[k - nx] == AR) qarb[k - nx] += s[k];
      if (t[k + nx] == AR) qarb[k + nx] += s[k];
    }
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const int a = i + j * nx;
      if (t[a] != AR || S.bub[a] < 0) continue;
      const int nf = (t[a - 1] == AGUA) + (t[a + 1] == AGUA) + (t[a - nx] == AGUA) + (t[a + nx] == AGUA);
      qarb[a] -= nf * s[C + S.bub[a]];
    }
  // rhs[k] = -(d - alvo e - ar) for water cells; bubble rhs face sums
  const double alvo = c.h / std::max(c.tau, S.dt);
  std::vector<double> db(C, 0.0);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = i + j * nx;
      if (t[k] != AGUA) continue;
      const double g = -s[k];  // adjoint of d
      Ub[i + 1 + j * (nx + 1)] += g;
      Ub[i + j * (nx + 1)] -= g;
      Vb[i + (j + 1) * nx] += g;
      Vb[i + j * nx] -= g;
      double dd;
      excess(c, t, i, j, S.pg.dens[k], &dd);
      db[k] += -alvo * dd * g;
    }
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const int a = i + j * nx;
      if (t[a] != AR || S.bub[a] < 0) continue;
      const double g = s[C + S.bub[a]];
      if (t[a - 1] == AGUA) Ub[i + j * (nx + 1)] += g;
      if (t[a + 1] == AGUA) Ub[i + 1 + j * (nx + 1)] -= g;
      if (t[a - nx] == AGUA) Vb[i + j * nx] += g;
      if (t[a + nx] == AGUA) Vb[i + (j + 1) * nx] -= g;
    }
  tension_adj(c, t, S.pg.dens, S.tn, c.sigma, S.dt, qarb, db);
  // gravity and walls, then the normalisation U0 = Un / W
  std::vector<double> ua(c.nu, 0.0), uw(c.nu, 0.0), va(c.nv, 0.0), vw(c.nv, 0.0);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i <= nx; i++) {
      const int k = i + j * (nx + 1);
      if (u_wall(c, t, i, j) || !(S.pg.W[k] > 0)) continue;
      ua[k] = Ub[k] / S.pg.W[k];
      uw[k] = -Ub[k] * S.U0[k] / S.pg.W[k];
    }
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = i + j * nx;
      if (v_wall(c, t, i, j) || !(S.pg.VW[k] > 0)) continue;
      va[k] = Vb[k] / S.pg.VW[k];
      vw[k] = -Vb[k] * S.V0[k] / S.pg.VW[k];
    }
  for (int p = 0; p < S.n; p++)
    p2g_adj_particle(c, &S.P[(size_t)p * 8], ua.data(), uw.data(), va.data(), vw.data(), db.data(),
                     &Pstart[(size_t)p * 8]);
  Pb.swap(Pstart);
  return 0;
}

// water fraction of a cell box and the adjoint of its density splat
static double fraction(const Ctx& c, const std::vector<double>& P, const int* box, std::vector<double>* dbar,
                       double fbar) {
  P2GOut o;
  p2g_fwd(c, P.data(), (int)(P.size() / 8), o);
  const int n = (box[2] - box[0]) * (box[3] - box[1]);
  double s = 0;
  if (dbar) dbar->assign(c.C, 0.0);
  for (int j = box[1]; j < box[3]; j++)
    for (int i = box[0]; i < box[2]; i++) {
      const double d = o.dens[i + j * c.nx] / PPC;
      s += std::min(1.0, d);
      if (dbar && d < 1) (*dbar)[i + j * c.nx] = fbar / n / PPC;
    }
  return s / n;
}

}  // namespace adj

using namespace adj;

extern "C" {

// Runs the replay forward over the frames and, if want_grad, the reverse pass.
//   p0: n0 x 8 initial state; frames: f_nemit emitted particles (positions in emit_xy, velocity
//   vf) then f_nsub substeps of f_dt / f_nsub; per substep s: tipo/bub/qextra (C each),
//   nb[s] bubbles with rhsfix; np[s] particles at its start and flags (concatenated).
//   record != 0: the particle decisions are made here and written to np/flags (flags_cap words).
//   out: [loss, fraction].  Returns 0, 1 (flags_cap too small), 2 (CG did not converge),
//   3 (tape mismatch).
int adj_run(int nx, int ny, double h, double gx, double gy, double tau, const double* sdf, int n0, const double* p0,
            int F, const int* f_nemit, const int* f_nsub, const double* f_dt, const double* emit_xy, const double* vf,
            const uint8_t* tipo, const int32_t* bub, double sigma, const double* qextra, const int* nb,
            const double* rhsfix, int record, long flags_cap, int* np, uint16_t* flags, const int* box,
            double target, double cg_tol, double* out, int want_grad, double* g_vf, double* g_sdf, double* g_p0) {
  Ctx c;
  c.nx = nx; c.ny = ny; c.C = nx * ny; c.nu = (nx + 1) * ny; c.nv = nx * (ny + 1);
  c.h = h; c.inv = 1 / h; c.gx = gx; c.gy = gy; c.tau = tau; c.sigma = sigma;
  c.mw = nx * RES; c.mh = ny * RES;
  c.sdf = sdf;
  c.gsdf = nullptr;
  int S = 0;
  for (int f = 0; f < F; f++) S += f_nsub[f];
  std::vector<Sub> subs(S);
  std::vector<double> P(p0, p0 + (size_t)n0 * 8);
  long fo = 0, eo = 0;
  int s = 0, bo = 0;
  std::vector<int> first_sub(F);
  for (int f = 0; f < F; f++) {
    for (int e = 0; e < f_nemit[f]; e++) {
      const double rec[8] = {emit_xy[2 * (eo + e)], emit_xy[2 * (eo + e) + 1], vf[0], vf[1], 0, 0, 0, 0};
      P.insert(P.end(), rec, rec + 8);
    }
    eo += f_nemit[f];
    first_sub[f] = s;
    for (int k = 0; k < f_nsub[f]; k++, s++) {
      Sub& B = subs[s];
      B.dt = f_dt[f] / f_nsub[f];
      const size_t off = (size_t)s * c.C;
      B.tipo = tipo + off; B.bub = bub + off; B.qextra = qextra + off;
      B.nb = nb[s]; B.rhsfix = rhsfix + bo;
      bo += nb[s];
      const int n = (int)(P.size() / 8);
      if (record) {
        if (fo + n > flags_cap) return 1;
        np[s] = n;
      } else if (np[s] != n) {
        return 3;
      }
      B.flags = flags + fo;
      fo += n;
      if (int e = substep_fwd(c, B, P, record != 0, cg_tol)) return e;
    }
  }
  std::vector<double> db;
  const double frac = fraction(c, P, box, nullptr, 0);
  const double L = (frac - target) * (frac - target);
  out[0] = L;
  out[1] = frac;
  if (!want_grad) return 0;
  // loss adjoint: dL/dF = 2 (F - target), through the density splat of the final particles
  fraction(c, P, box, &db, 2 * (frac - target));
  std::vector<double> Pb(P.size(), 0.0);
  for (size_t p = 0; p < P.size() / 8; p++) {
    Sten st = sten_c(c, P[p * 8], P[p * 8 + 1]);
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++) {
        const double g = db[st.i0 + a + (st.j0 + b) * c.nx];
        Pb[p * 8] += g * dwa(a) * wa(st.ty, b) * c.inv;
        Pb[p * 8 + 1] += g * wa(st.tx, a) * dwa(b) * c.inv;
      }
  }
  if (g_sdf) std::fill(g_sdf, g_sdf + (size_t)c.mw * c.mh, 0.0);
  c.gsdf = g_sdf;
  g_vf[0] = g_vf[1] = 0;
  for (int f = F - 1; f >= 0; f--) {
    for (int k = f_nsub[f] - 1; k >= 0; k--) {
      if (int e = substep_adj(c, subs[first_sub[f] + k], Pb, cg_tol)) return e;
    }
    // the particles emitted at this frame's start are the last f_nemit[f] of the list
    const size_t n = Pb.size() / 8;
    const size_t ne = (size_t)f_nemit[f];
    for (size_t p = n - ne; p < n; p++) { g_vf[0] += Pb[p * 8 + 2]; g_vf[1] += Pb[p * 8 + 3]; }
    Pb.resize((n - ne) * 8);
  }
  if (g_p0) std::copy(Pb.begin(), Pb.end(), g_p0);
  return 0;
}

// The pressure operator of one substep's tape applied to z (C + nb entries); for symmetry checks.
void adj_apply_operator(int nx, int ny, const uint8_t* tipo, const int32_t* bub, int nb, const double* z, double* out) {
  Ctx c;
  c.nx = nx; c.ny = ny; c.C = nx * ny;
  System S;
  S.c = &c; S.tipo = tipo; S.bub = bub; S.nb = nb;
  S.build();
  std::vector<double> zz(z, z + c.C + nb), o(c.C + nb);
  S.apply(zz, o);
  std::copy(o.begin(), o.end(), out);
}

}  // extern "C"
