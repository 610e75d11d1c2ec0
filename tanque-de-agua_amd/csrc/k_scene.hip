// k_scene.hip -- the tank's walls built on the device: the frame and the capsule strokes on the
// (2nx) x (2ny) SDF lattice (js/cena.js:43-72 tracar, 132-144 #moldura/#zerar), one thread per
// lattice point applying every stroke in order with the reference's binary32 store after each,
// then the solid cells (cena.js:119-127); and the chain rule from SDF values back to the
// strokes' ends and radii (the adjoint's wall parameters).
#include <algorithm>
#include <cmath>
#include <vector>

#include "tq_device.hpp"
#include "tq_internal.hpp"

namespace tq {

// one stroke in lattice coordinates, with the point range tracar visits
struct DevStroke {
  double ax, ay, ex, ey, len2, rp;
  int xlo, xhi, ylo, yhi;
  int apagar;
};

__device__ inline double frame_dist_dev(double x, double y, int mw, int mh) {
#pragma clang fp contract(off)
  const double edge = MOLDURA * 2 - 0.5;
  const double d = fmin(fmin(x - edge, mw - 1 - edge - x), fmin(y - edge, mh - 1 - edge - y)) / 2;
  return d > SDF_MAX ? SDF_MAX : d < -SDF_MAX ? -SDF_MAX : d;
}

// distance of lattice point (x, y) to stroke S in cells, clamped; q = nearest point - (x, y),
// t its parameter on the segment
__device__ inline double stroke_dist(const DevStroke& S, double x, double y, double* qx, double* qy, double* t,
                                     double* nrm) {
#pragma clang fp contract(off)
  double tt = S.len2 > 0 ? ((x - S.ax) * S.ex + (y - S.ay) * S.ey) / S.len2 : 0;
  tt = tt < 0 ? 0 : tt > 1 ? 1 : tt;
  *qx = S.ax + tt * S.ex - x;
  *qy = S.ay + tt * S.ey - y;
  *t = tt;
  *nrm = sqrt(*qx * *qx + *qy * *qy);
  const double d = (*nrm - S.rp) / 2;
  return d > SDF_MAX ? SDF_MAX : d < -SDF_MAX ? -SDF_MAX : d;
}

__global__ __launch_bounds__(256) void k_sdf_strokes(const DevStroke* __restrict__ st, int n, int mw, int mh,
                                                     float* __restrict__ sdf) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= mw * mh) return;
  const int x = k % mw, y = k / mw;
  const double fr = frame_dist_dev(x, y, mw, mh);
  float v = (float)fr;
  for (int s = 0; s < n; s++) {
    const DevStroke S = st[s];
    if (x < S.xlo || x > S.xhi || y < S.ylo || y > S.yhi) continue;
    double qx, qy, t, nrm;
    const double d = stroke_dist(S, x, y, &qx, &qy, &t, &nrm);
    const double cur = v;
    v = (float)(S.apagar ? fmin(fmax(cur, -d), fr) : fmin(cur, d));
  }
  sdf[k] = v;
}

// d(sdf)/d(x0, y0, x1, y1, raio) of the stroke that set each lattice point, weighted by g and
// summed per stroke: the point's value is that of the last stroke that changed it (none: the
// frame, no derivative); an erase stroke sets -d, so its derivative changes sign.
__global__ __launch_bounds__(256) void k_sdf_strokes_vjp(const DevStroke* __restrict__ st, int n, int mw, int mh,
                                                         double h, const float* __restrict__ g,
                                                         double* __restrict__ out) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= mw * mh) return;
  const double gk = g[k];
  if (gk == 0) return;
  const int x = k % mw, y = k / mw;
  const double fr = frame_dist_dev(x, y, mw, mh);
  float v = (float)fr;
  int owner = -1;
  double sign = 0, dqx = 0, dqy = 0, dt = 0, dn = 0;
  for (int s = 0; s < n; s++) {
    const DevStroke S = st[s];
    if (x < S.xlo || x > S.xhi || y < S.ylo || y > S.yhi) continue;
    double qx, qy, t, nrm;
    const double d = stroke_dist(S, x, y, &qx, &qy, &t, &nrm);
    const double cur = v;
    const float nv = (float)(S.apagar ? fmin(fmax(cur, -d), fr) : fmin(cur, d));
    if (nv != v) {
      // the new value is this stroke's (clamped) distance, or, for an erase stroke, the frame
      // when the frame is below max(cur, -d)
      const bool from_frame = S.apagar && fr < fmax(cur, -d);
      const bool clamped = !(fabs((nrm - S.rp) / 2) < SDF_MAX);
      if (from_frame || clamped || !(nrm > 0)) {
        owner = -1;
      } else {
        owner = s; sign = S.apagar ? -1.0 : 1.0; dqx = qx; dqy = qy; dt = t; dn = nrm;
      }
    }
    v = nv;
  }
  if (owner < 0) return;
  // d = (|q| - rp)/2 in cells, q = A + t (B - A) - P with A = (x0, y0) R/h - 1/2: by the envelope
  // theorem d|q|/dA = (q/|q|)(1 - t), d|q|/dB = (q/|q|) t; d/d(raio) = -(R/h)/2 = -1/h
  const double c = sign * gk / h;
  const double ux = dqx / dn, uy = dqy / dn;
  atomicAdd(&out[5 * owner + 0], c * ux * (1 - dt));
  atomicAdd(&out[5 * owner + 1], c * uy * (1 - dt));
  atomicAdd(&out[5 * owner + 2], c * ux * dt);
  atomicAdd(&out[5 * owner + 3], c * uy * dt);
  atomicAdd(&out[5 * owner + 4], -c);
}

// the strokes' lattice geometry and point ranges, exactly as tracar computes them
static std::vector<DevStroke> plan_strokes(const tq_sim* s, const double* strokes, int n) {
  const int mw = 2 * s->nx, mh = 2 * s->ny;
  const double k = 2 / s->h;
  std::vector<DevStroke> out;
  out.reserve(n);
  for (int i = 0; i < n; i++) {
    const double* p = strokes + 6 * (size_t)i;
    DevStroke S;
    S.ax = p[0] * k - 0.5; S.ay = p[1] * k - 0.5;
    const double bx = p[2] * k - 0.5, by = p[3] * k - 0.5;
    S.rp = p[4] * k;
    S.ex = bx - S.ax; S.ey = by - S.ay;
    S.len2 = S.ex * S.ex + S.ey * S.ey;
    S.apagar = p[5] != 0;
    const double reach = S.rp + SDF_MAX * 2 + 1;
    const double xlo = std::max(0.0, std::floor(std::min(S.ax, bx) - reach));
    const double xhi = std::min((double)(mw - 1), std::ceil(std::max(S.ax, bx) + reach));
    const double ylo = std::max(0.0, std::floor(std::min(S.ay, by) - reach));
    const double yhi = std::min((double)(mh - 1), std::ceil(std::max(S.ay, by) + reach));
    if (!(xlo <= xhi && ylo <= yhi)) continue;  // tracar returns without touching the lattice
    S.xlo = (int)xlo; S.xhi = (int)xhi; S.ylo = (int)ylo; S.yhi = (int)yhi;
    out.push_back(S);
  }
  return out;
}

}  // namespace tq

using namespace tq;

extern "C" {

int tq_set_strokes(tq_sim* s, const double* strokes, int n) {
  if (!s) { set_error("null handle"); return TQ_E_ARG; }
  if (s->tape) { set_error("tq_set_strokes: not allowed while frames are recorded"); return TQ_E_STATE; }
  if (n < 0 || (n > 0 && !strokes)) { set_error("tq_set_strokes: bad stroke list"); return TQ_E_ARG; }
  for (int i = 0; i < n; i++)
    for (int q = 0; q < 5; q++)
      if (!std::isfinite(strokes[6 * (size_t)i + q])) { set_error("tq_set_strokes: non-finite stroke"); return TQ_E_ARG; }
  const std::vector<DevStroke> plan = plan_strokes(s, strokes, n);
  DevStroke* d = nullptr;
  if (!plan.empty()) {
    TQ_HIP(hipMalloc((void**)&d, sizeof(DevStroke) * plan.size()));
    hipError_t e = hipMemcpyAsync(d, plan.data(), sizeof(DevStroke) * plan.size(), hipMemcpyHostToDevice, s->st);
    if (e != hipSuccess) { (void)hipFree(d); return hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__); }
  }
  const int mw = 2 * s->nx, mh = 2 * s->ny;
  hipLaunchKernelGGL(k_sdf_strokes, dim3((mw * mh + 255) / 256), dim3(256), 0, s->st, d, (int)plan.size(), mw, mh, s->sdf);
  hipError_t e = hipGetLastError();
  int rc = e == hipSuccess ? solid_cells(s) : hip_fail(e, "k_sdf_strokes", __FILE__, __LINE__);
  e = hipStreamSynchronize(s->st);
  if (!rc && e != hipSuccess) rc = hip_fail(e, "tq_set_strokes", __FILE__, __LINE__);
  if (d) (void)hipFree(d);
  return rc;
}

int tq_read_sdf(tq_sim* s, float* dst, size_t bytes) {
  if (!s || !dst) { set_error("tq_read_sdf: null argument"); return TQ_E_ARG; }
  if (bytes != sizeof(float) * 4 * (size_t)s->C) { set_error("tq_read_sdf: size mismatch"); return TQ_E_ARG; }
  TQ_HIP(hipMemcpyAsync(dst, s->sdf, bytes, hipMemcpyDeviceToHost, s->st));
  TQ_HIP(hipStreamSynchronize(s->st));
  return 0;
}

int tq_strokes_vjp(tq_sim* s, const double* strokes, int n, const float* gsdf, double* out) {
  if (!s) { set_error("null handle"); return TQ_E_ARG; }
  if (n < 0 || (n > 0 && (!strokes || !out)) || !gsdf) { set_error("tq_strokes_vjp: bad arguments"); return TQ_E_ARG; }
  // plan_strokes drops strokes that touch no lattice point: keep the caller's indexing
  std::vector<int> index;
  std::vector<double> kept;
  for (int i = 0; i < n; i++) {
    const std::vector<DevStroke> one = plan_strokes(s, strokes + 6 * (size_t)i, 1);
    if (one.empty()) continue;
    index.push_back(i);
    kept.insert(kept.end(), strokes + 6 * (size_t)i, strokes + 6 * (size_t)i + 6);
  }
  std::fill(out, out + 5 * (size_t)n, 0.0);
  if (index.empty()) return 0;
  const std::vector<DevStroke> plan = plan_strokes(s, kept.data(), (int)index.size());
  const int mw = 2 * s->nx, mh = 2 * s->ny;
  const size_t pts = (size_t)mw * mh;
  DevStroke* d = nullptr;
  float* g = nullptr;
  double* o = nullptr;
  int rc = 0;
  auto chk = [&](hipError_t e, const char* w) { if (!rc && e != hipSuccess) rc = hip_fail(e, w, __FILE__, __LINE__); };
  chk(hipMalloc((void**)&d, sizeof(DevStroke) * plan.size()), "hipMalloc");
  chk(hipMalloc((void**)&g, 4 * pts), "hipMalloc");
  chk(hipMalloc((void**)&o, 8 * 5 * plan.size()), "hipMalloc");
  if (!rc) {
    chk(hipMemcpyAsync(d, plan.data(), sizeof(DevStroke) * plan.size(), hipMemcpyHostToDevice, s->st), "hipMemcpyAsync");
    chk(hipMemcpyAsync(g, gsdf, 4 * pts, hipMemcpyHostToDevice, s->st), "hipMemcpyAsync");
    chk(hipMemsetAsync(o, 0, 8 * 5 * plan.size(), s->st), "hipMemsetAsync");
  }
  if (!rc) {
    hipLaunchKernelGGL(k_sdf_strokes_vjp, dim3((unsigned)((pts + 255) / 256)), dim3(256), 0, s->st, d, (int)plan.size(), mw,
                       mh, s->h, g, o);
    chk(hipGetLastError(), "k_sdf_strokes_vjp");
  }
  std::vector<double> h(5 * plan.size());
  if (!rc) chk(hipMemcpyAsync(h.data(), o, 8 * h.size(), hipMemcpyDeviceToHost, s->st), "hipMemcpyAsync");
  if (!rc) chk(hipStreamSynchronize(s->st), "tq_strokes_vjp");
  if (!rc)
    for (size_t i = 0; i < index.size(); i++)
      for (int q = 0; q < 5; q++) out[5 * (size_t)index[i] + q] = h[5 * i + q];
  if (d) (void)hipFree(d);
  if (g) (void)hipFree(g);
  if (o) (void)hipFree(o);
  return rc;
}

}  // extern "C"
