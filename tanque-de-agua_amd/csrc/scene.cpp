// scene.cpp -- host-side scene construction for callers without the reference's Cena:
// the tank frame and capsule strokes on the (2nx) x (2ny) SDF lattice (js/cena.js:43-72,
// 132-144), seeded rectangles of particles (fisica-cpu.js:705-727 with the LCG of
// testes/apoio.js:4-9) and the adaptive substep bound (constantes.js:167-175).
// The arithmetic is binary64 with binary32 stores, as in the JavaScript it replaces.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/tanque.h"

namespace {

constexpr int R = 2;             // SDF samples per cell per axis (RES_MASCARA)
constexpr double DMAX = 10;      // SDF clamp in cells (SDF_MAX)
constexpr int FRAME = 2;         // frame thickness in cells (MOLDURA)

inline double clampv(double d, double lo, double hi) { return d > hi ? hi : d < lo ? lo : d; }

// signed distance (cells) to the inner edge of the frame at lattice point (x, y)
inline double frame_dist(double x, double y, int mw, int mh) {
  const double edge = FRAME * R - 0.5;
  double d = std::min(std::min(x - edge, mw - 1 - edge - x), std::min(y - edge, mh - 1 - edge - y)) / R;
  return clampv(d, -DMAX, DMAX);
}

inline double js_round(double x) {
  double f = std::floor(x);
  return (x - f >= 0.5) ? f + 1 : f;
}

inline double lcg(uint32_t* s) {
  *s = *s * 1664525u + 1013904223u;
  return (double)*s / 4294967296.0;
}

}  // namespace

extern "C" {

void tq_scene_frame(float* sdf, int nx, int ny) {
  const int mw = nx * R, mh = ny * R;
  for (int y = 0; y < mh; y++)
    for (int x = 0; x < mw; x++) sdf[(size_t)y * mw + x] = (float)frame_dist(x, y, mw, mh);
}

void tq_scene_stroke(float* sdf, int nx, int ny, double h, double x0, double y0, double x1, double y1, double raio,
                     int apagar) {
  const int mw = nx * R, mh = ny * R;
  const double k = R / h;  // lattice points per metre
  const double ax = x0 * k - 0.5, ay = y0 * k - 0.5, bx = x1 * k - 0.5, by = y1 * k - 0.5;
  const double rp = raio * k;
  const double reach = rp + DMAX * R + 1;
  const double xlo = std::max(0.0, std::floor(std::min(ax, bx) - reach));
  const double xhi = std::min((double)(mw - 1), std::ceil(std::max(ax, bx) + reach));
  const double ylo = std::max(0.0, std::floor(std::min(ay, by) - reach));
  const double yhi = std::min((double)(mh - 1), std::ceil(std::max(ay, by) + reach));
  if (xlo > xhi || ylo > yhi) return;
  const double ex = bx - ax, ey = by - ay;
  const double len2 = ex * ex + ey * ey;
  for (int y = (int)ylo; y <= (int)yhi; y++) {
    for (int x = (int)xlo; x <= (int)xhi; x++) {
      double t = len2 > 0 ? ((x - ax) * ex + (y - ay) * ey) / len2 : 0;
      t = t < 0 ? 0 : t > 1 ? 1 : t;
      const double px = ax + t * ex - x, py = ay + t * ey - y;
      const double d = clampv((std::sqrt(px * px + py * py) - rp) / R, -DMAX, DMAX);
      float& cell = sdf[(size_t)y * mw + x];
      const double cur = cell;
      double v;
      if (apagar) v = std::min(std::max(cur, -d), frame_dist(x, y, mw, mh));
      else v = std::min(cur, d);
      cell = (float)v;
    }
  }
}

int tq_seed_rect(double x0, double y0, double x1, double y1, double h, const double* absorcao, uint32_t* seed,
                 float* out, int max_out) {
  const double side = js_round(std::sqrt(4.0));  // 2x2 per cell (PARTICULAS_POR_CELULA = 4)
  const double step = h / side;
  const double xa = std::min(x0, x1), xb = std::max(x0, x1), ya = std::min(y0, y1), yb = std::max(y0, y1);
  const double ia = std::ceil(xa / step - 0.5), ib = std::floor(xb / step - 0.5);
  const double ja = std::ceil(ya / step - 0.5), jb = std::floor(yb / step - 0.5);
  const double n = std::max(0.0, ib - ia + 1) * std::max(0.0, jb - ja + 1);
  if (!out) return (int)n;
  if (n > max_out) return -1;
  const double a0 = absorcao ? absorcao[0] : 0, a1 = absorcao ? absorcao[1] : 0, a2 = absorcao ? absorcao[2] : 0;
  size_t k = 0;
  for (double j = ja; j <= jb; j++)
    for (double i = ia; i <= ib; i++, k++) {
      float* p = out + k * TQ_CAMPOS;
      std::memset(p, 0, sizeof(float) * TQ_CAMPOS);
      const double rx = lcg(seed), ry = lcg(seed);
      p[0] = (float)((i + 0.5 + (rx - 0.5) * 0.5) * step);
      p[1] = (float)((j + 0.5 + (ry - 0.5) * 0.5) * step);
      p[9] = (float)a0; p[10] = (float)a1; p[11] = (float)a2;
    }
  return (int)k;
}

double tq_passo_maximo(double h, double altura, double vmax, int has_vmax, double dt_quadro) {
  const double g = 9.81, rho = 1000, sigma = 0.0728;
  const double fall = std::sqrt(2 * g * altura) * 1.15 + 0.5;
  const double v = has_vmax ? std::min(fall, vmax * 1.2 + g * dt_quadro + 0.25) : fall;
  const double capillary = 0.4 * std::sqrt((rho * h * h * h) / (2 * M_PI * sigma));
  return std::min(std::min((2.0 * h) / std::max(v, 0.3), capillary), 0.008);
}

}  // extern "C"
