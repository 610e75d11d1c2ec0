// k_render.hip -- the render hand-off: the water layer the reference's 2D renderer computes from
// the particle buffer (js/render-2d.js:110-162), on the device.  At 3 pixels per cell every
// particle is a (1 - r^2)^2 splat of radius RAIO_MANCHA cells into coverage, foam and dye fields;
// each pixel then gets the Beer-Lambert tint of the water (RGBA, multiplied over the panel) and
// the foam colour with its alpha.  The scatter of the reference becomes a gather: the particles
// are sorted by cell for their current positions, and one thread per pixel sums the particles of
// the cells within the splat radius (binary64 sums, no atomics).
#include <algorithm>
#include <cmath>

#include "tq_device.hpp"
#include "tq_internal.hpp"

namespace tq {

constexpr int PX_POR_CELULA = 3;       // render-2d.js:10
constexpr double RAIO_MANCHA = 2.1;    // splat radius in cells (constantes.js:120)
constexpr double PROFUNDIDADE = 0.05;  // water depth for the absorption (constantes.js:41)
__constant__ double ABSORCAO_AGUA[3] = {9.0, 2.6, 1.6};  // constantes.js:138

__device__ inline double suave_px(double a, double b, double x) {
#pragma clang fp contract(off)
  const double t = fmin(1.0, fmax(0.0, (x - a) / (b - a)));
  return t * t * (3 - 2 * t);
}

// a Uint8ClampedArray store: round half to even, clamp to [0, 255]
__device__ inline uint8_t clamped_u8(double v) {
  const double r = rint(v);
  return r <= 0 ? 0 : r >= 255 ? 255 : (uint8_t)r;
}

__global__ __launch_bounds__(256) void k_render_water(const float2* __restrict__ pos, const float4* __restrict__ ext,
                                                      const uint32_t* __restrict__ startKey, int nx, int ny,
                                                      int ntx, double h, int w, int hh, uchar4* __restrict__ tint,
                                                      uchar4* __restrict__ foam) {
#pragma clang fp contract(off)
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= w * hh) return;
  const int x = pix % w, y = pix / w;
  const double escala = PX_POR_CELULA / h;
  const double R = RAIO_MANCHA * PX_POR_CELULA, R2 = R * R;
  // cells whose particles can reach this pixel (one cell of margin on each side)
  const double X = (x + 0.5) / escala / h, Y = (hh - (y + 0.5)) / escala / h;
  int i0 = (int)floor(X - RAIO_MANCHA) - 1, i1 = (int)floor(X + RAIO_MANCHA) + 1;
  int j0 = (int)floor(Y - RAIO_MANCHA) - 1, j1 = (int)floor(Y + RAIO_MANCHA) + 1;
  i0 = i0 < 0 ? 0 : i0;
  j0 = j0 < 0 ? 0 : j0;
  i1 = i1 > nx - 1 ? nx - 1 : i1;
  j1 = j1 > ny - 1 ? ny - 1 : j1;
  double campo = 0, espuma = 0, c0 = 0, c1 = 0, c2 = 0;
  for (int j = j0; j <= j1; j++)
    for (int i = i0; i <= i1; i++) {
      const uint32_t key = cell_key(i, j, ntx);
      const uint32_t k1 = startKey[key + 1];
      for (uint32_t k = startKey[key]; k < k1; k++) {
        const float2 P = pos[k];
        const double cx = P.x * escala, cy = hh - P.y * escala;
        // the splat's pixel box and disc (render-2d.js:123-133)
        if (x < ceil(cx - R - 0.5) || x > floor(cx + R - 0.5) || y < ceil(cy - R - 0.5) || y > floor(cy + R - 0.5))
          continue;
        const double dx = x + 0.5 - cx, dy = y + 0.5 - cy;
        const double r2 = (dx * dx + dy * dy) / R2;
        if (r2 >= 1) continue;
        const double peso = (1 - r2) * (1 - r2);
        const float4 E = ext[k];  // foam, then the dye absorption
        campo += peso;
        espuma += peso * E.x;
        c0 += peso * E.y;
        c1 += peso * E.z;
        c2 += peso * E.w;
      }
    }
  // shading (render-2d.js:140-160)
  const double phi0 = (PPC * M_PI * RAIO_MANCHA * RAIO_MANCHA) / 3;
  const double phi = campo / phi0;
  const double cobre = suave_px(0.26, 0.5, phi);
  if (cobre < 0.002) {
    tint[pix] = make_uchar4(255, 255, 255, 255);
    foam[pix] = make_uchar4(0, 0, 0, 0);
    return;
  }
  const double esp = suave_px(0.15, 0.95, phi) * PROFUNDIDADE;
  const double inv = 1 / fmax(campo, 1e-4);
  const double borda = 1 - cobre * (1 - cobre) * 4 * 0.45;
  const double tr = exp(-(ABSORCAO_AGUA[0] + c0 * inv) * esp) * borda;
  const double tg = exp(-(ABSORCAO_AGUA[1] + c1 * inv) * esp) * borda;
  const double tb = exp(-(ABSORCAO_AGUA[2] + c2 * inv) * esp) * borda;
  tint[pix] = make_uchar4(clamped_u8((1 - cobre + cobre * tr) * 255), clamped_u8((1 - cobre + cobre * tg) * 255),
                          clamped_u8((1 - cobre + cobre * tb) * 255), 255);
  const double e = suave_px(0.05, 0.75, fmin(1.0, espuma * inv)) * 0.85 * cobre;
  foam[pix] = make_uchar4(248, 251, 252, clamped_u8(e * 255));
}

}  // namespace tq

using namespace tq;

extern "C" {

int tq_render_water(tq_sim* s, int w, int h, uint8_t* tint_rgba, uint8_t* foam_rgba) {
  if (!s) { set_error("null handle"); return TQ_E_ARG; }
  if (s->tape) { set_error("tq_render_water: not allowed while frames are recorded"); return TQ_E_STATE; }
  if (w != PX_POR_CELULA * s->nx || h != PX_POR_CELULA * s->ny || !tint_rgba || !foam_rgba) {
    set_error("tq_render_water: the images are (3 nx) x (3 ny) RGBA, both non-null");
    return TQ_E_ARG;
  }
  const size_t px = (size_t)w * h;
  if (s->render_bytes < 8 * px) {
    if (s->render_img) (void)hipFree(s->render_img);
    s->render_img = nullptr;
    s->render_bytes = 0;
    TQ_HIP(hipMalloc((void**)&s->render_img, 8 * px));
    s->render_bytes = 8 * px;
  }
  // the particles sorted by cell for their current positions (P2G sorts them again anyway)
  if (int e = bin_and_sort(s)) return e;
  uchar4* tint = (uchar4*)s->render_img;
  uchar4* foam = tint + px;
  const int o = s->cur;
  hipLaunchKernelGGL(k_render_water, dim3((unsigned)((px + 255) / 256)), dim3(256), 0, s->st, s->pos[o], s->ext[o],
                     s->startKey, s->nx, s->ny, s->ntx, s->h, w, h, tint, foam);
  TQ_LAUNCH_CHECK();
  TQ_HIP(hipMemcpyAsync(tint_rgba, tint, 4 * px, hipMemcpyDeviceToHost, s->st));
  TQ_HIP(hipMemcpyAsync(foam_rgba, foam, 4 * px, hipMemcpyDeviceToHost, s->st));
  TQ_HIP(hipStreamSynchronize(s->st));
  return 0;
}

}  // extern "C"
