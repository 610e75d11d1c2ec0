// tanque_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
//
// A single-threaded CPU restatement of the reference's CPU engine:
//   js/constantes.js, js/cena.js, js/fisica-cpu.js, js/pressao-cpu.js, js/ar.js
//   and the seeded LCG of testes/apoio.js.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Arithmetic model.  JavaScript computes every expression in IEEE binary64 and rounds
// to binary32 on every Float32Array store.  Sim<float> ("JS-exact" mode) does exactly
// that: all intermediates are double, every array store is a float store.  Sim<double>
// ("FP64" mode) keeps the same operation order but stores in binary64; it is the
// reference against which both the JS-exact oracle and the GPU path state their error.
// Math.hypot follows V8's published algorithm (scaled, Kahan-compensated sum of squares);
// Math.exp is the C library's exp (V8 uses an fdlibm derivative; both are < 1 ulp).
//
// Additions that the reference lacks (flagged where they appear): a residual-terminated
// PCG mode (tol > 0) and reporting of the true residual, used for the 1e-6 workloads.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <utility>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- constants (js/constantes.js)
constexpr double GRAVIDADE = 9.81;              // constantes.js:8
constexpr double DENSIDADE_AGUA = 1000;         // :11
constexpr double TENSAO_SUPERFICIAL = 0.0728;   // :18
constexpr int MAX_BOLHAS = 4096;                // :20
constexpr double VOLTA_VOLUME_BOLHA = 0.05;     // :29
constexpr double PRESSAO_BOLHA_MAX = 20000;     // :30
constexpr double FATOR_CAPILAR = 0.4;           // :36
constexpr double PASSO_MAX = 0.008;             // :38
constexpr int PARTICULAS_POR_CELULA = 4;        // :44
constexpr int RES_MASCARA = 2;                  // :47
constexpr double SDF_MAX = 10;                  // :50
constexpr int MOLDURA = 2;                      // :53
constexpr double RAIO_PARTICULA = 0.3;          // :56
constexpr double CFL = 2.0;                     // :65
constexpr double MICROPASSO = 0.8;              // :66
constexpr int MAX_MICROPASSOS = 4;              // :68
constexpr double TAU_DENSIDADE = 0.1;           // :77
constexpr double EXCESSO_MIN = -0.5;            // :79
constexpr double EXCESSO_MAX = 1.0;             // :80
constexpr int ITERACOES_PCG = 3;                // :87
constexpr int CAMADAS_EXTRAPOLACAO = 2;         // :91
constexpr double ESPUMA_V0 = 1.1, ESPUMA_V1 = 2.8, ESPUMA_TAXA = 5, ESPUMA_VIDA = 1.2;  // :100-104
constexpr double DIFUSAO_CORANTE = 3;           // :126
constexpr int CAMPOS = 12;                      // fisica-cpu.js:26
constexpr uint8_t SOLIDO = 0, AGUA = 1, AR = 2; // fisica-cpu.js:24
constexpr uint32_t ATMOSFERA = 1;               // ar.js:18
constexpr int MIN_BOLHA_CELULAS = 4;            // ar.js:20
constexpr int BLOCO = 16, VARREDURAS = 4, VARREDURAS_GROSSO = 64;  // pressao-cpu.js:17-20

// ---------------------------------------------------------------- JS number helpers
// Math.round: nearest integer, ties toward +infinity (ECMA-262 Math.round).
static double js_round(double x) {
  double f = std::floor(x);
  return (x - f >= 0.5) ? f + 1 : f;
}
// Math.hypot(a, b) as V8 computes it: scale by the largest magnitude, sum the squares with
// Kahan compensation, sqrt, rescale (V8 builtins MathHypot).
static double js_hypot(double a, double b) {
  double xs[2] = {std::fabs(a), std::fabs(b)};
  if (std::isinf(xs[0]) || std::isinf(xs[1])) return INFINITY;
  if (std::isnan(xs[0]) || std::isnan(xs[1])) return NAN;
  double mx = std::max(xs[0], xs[1]);
  if (mx == 0) return 0;
  double sum = 0, comp = 0;
  for (double x : xs) {
    double n = x / mx;
    double summand = n * n - comp;
    double prelim = sum + summand;
    comp = (prelim - sum) - summand;
    sum = prelim;
  }
  return std::sqrt(sum) * mx;
}
// clampi (fisica-cpu.js:697-699): `v < a ? a : v > b ? b : v` on a double (NaN passes through)
static double clampi(double v, double a, double b) { return v < a ? a : v > b ? b : v; }
// suave (fisica-cpu.js:692-695)
static double suave(double a, double b, double x) {
  double t = std::min(1.0, std::max(0.0, (x - a) / (b - a)));
  return t * t * (3 - 2 * t);
}

// passoMaximo (constantes.js:167-175); has_vmax = 0 means `vMax == null`.
static double passo_maximo(double h, double alturaM, double vMax, int has_vmax, double dtQuadro) {
  double queda = std::sqrt(2 * GRAVIDADE * alturaM) * 1.15 + 0.5;
  double v = !has_vmax ? queda : std::min(queda, vMax * 1.2 + GRAVIDADE * dtQuadro + 0.25);
  double capilar = FATOR_CAPILAR * std::sqrt((DENSIDADE_AGUA * h * h * h) / (2 * M_PI * TENSAO_SUPERFICIAL));
  return std::min(std::min((CFL * h) / std::max(v, 0.3), capilar), PASSO_MAX);
}

// seeded LCG (testes/apoio.js:4-9): s = (s*1664525 + 1013904223) >>> 0; return s / 2^32
static double lcg_next(uint32_t* s) {
  *s = *s * 1664525u + 1013904223u;
  return (double)*s / 4294967296.0;
}

// ---------------------------------------------------------------- Cena (js/cena.js)
struct Cena {
  int nx, ny, r, mw, mh;
  double h, largura, altura;
  std::vector<float> sdf;
  Cena(int nx_, int ny_, double h_) : nx(nx_), ny(ny_), r(RES_MASCARA), h(h_) {
    mw = nx * r; mh = ny * r; largura = nx * h; altura = ny * h;
    sdf.assign((size_t)mw * mh, 0.f);
    zerar();
  }
  // #moldura (cena.js:132-137)
  double moldura(double x, double y) const {
    double m = MOLDURA * r - 0.5;
    double d = std::min(std::min(x - m, mw - 1 - m - x), std::min(y - m, mh - 1 - m - y)) / r;
    return d > SDF_MAX ? SDF_MAX : d < -SDF_MAX ? -SDF_MAX : d;
  }
  // #zerar (cena.js:139-144)
  void zerar() {
    for (int y = 0; y < mh; y++)
      for (int x = 0; x < mw; x++) sdf[(size_t)y * mw + x] = (float)moldura(x, y);
  }
  // tracar (cena.js:43-72)
  void tracar(double x0, double y0, double x1, double y1, double raio, bool apagar) {
    double s = r / h;
    double ax = x0 * s - 0.5, ay = y0 * s - 0.5, bx = x1 * s - 0.5, by = y1 * s - 0.5;
    double rp = raio * s;
    double alc = rp + SDF_MAX * r + 1;
    double xmin = std::max(0.0, std::floor(std::min(ax, bx) - alc));
    double xmax = std::min((double)(mw - 1), std::ceil(std::max(ax, bx) + alc));
    double ymin = std::max(0.0, std::floor(std::min(ay, by) - alc));
    double ymax = std::min((double)(mh - 1), std::ceil(std::max(ay, by) + alc));
    if (xmin > xmax || ymin > ymax) return;
    double dx = bx - ax, dy = by - ay;
    double len2 = dx * dx + dy * dy;
    for (int y = (int)ymin; y <= (int)ymax; y++) {
      for (int x = (int)xmin; x <= (int)xmax; x++) {
        double t = len2 > 0 ? ((x - ax) * dx + (y - ay) * dy) / len2 : 0;
        t = t < 0 ? 0 : t > 1 ? 1 : t;
        double px = ax + t * dx - x, py = ay + t * dy - y;
        double d = (std::sqrt(px * px + py * py) - rp) / r;
        if (d > SDF_MAX) d = SDF_MAX; else if (d < -SDF_MAX) d = -SDF_MAX;
        size_t k = (size_t)y * mw + x;
        double cur = sdf[k];
        double v = apagar ? std::max(cur, -d) : std::min(cur, d);
        if (apagar) v = std::min(v, moldura(x, y));
        sdf[k] = (float)v;
      }
    }
  }
  // #bilinear (cena.js:156-165)
  double bilinear(double fx, double fy) const {
    int w = mw, hh = mh;
    if (fx < 0) fx = 0; else if (fx > w - 1) fx = w - 1;
    if (fy < 0) fy = 0; else if (fy > hh - 1) fy = hh - 1;
    double x0 = std::min(std::floor(fx), (double)(w - 2)), y0 = std::min(std::floor(fy), (double)(hh - 2));
    if (std::isnan(x0) || std::isnan(y0)) return NAN;
    double tx = fx - x0, ty = fy - y0;
    size_t k = (size_t)y0 * w + (size_t)x0;
    const float* d = sdf.data();
    return ((double)d[k] * (1 - tx) + (double)d[k + 1] * tx) * (1 - ty) +
           ((double)d[k + w] * (1 - tx) + (double)d[k + w + 1] * tx) * ty;
  }
  // distancia (cena.js:102-106)
  double distancia(double x, double y) const {
    double fx = (x / h) * r - 0.5, fy = (y / h) * r - 0.5;
    return bilinear(fx, fy);
  }
  // normal (cena.js:109-117)
  void normal(double x, double y, double* nxo, double* nyo) const {
    double fx = (x / h) * r - 0.5, fy = (y / h) * r - 0.5;
    double gx = bilinear(fx + 1, fy) - bilinear(fx - 1, fy);
    double gy = bilinear(fx, fy + 1) - bilinear(fx, fy - 1);
    double n = js_hypot(gx, gy);
    if (n > 1e-9) { *nxo = gx / n; *nyo = gy / n; } else { *nxo = 0; *nyo = 1; }
  }
  // celulaSolida (cena.js:119-127): the 4-sample sum is a binary64 sum of binary32 values
  bool celula_solida(int i, int j) const {
    if (i < 0 || j < 0 || i >= nx || j >= ny) return true;
    int w = mw, a = i * r, b = j * r;
    double s = (double)sdf[(size_t)b * w + a] + (double)sdf[(size_t)b * w + a + 1];
    s = s + (double)sdf[(size_t)(b + 1) * w + a];
    s = s + (double)sdf[(size_t)(b + 1) * w + a + 1];
    return s < 0;
  }
};

// semearRetangulo (fisica-cpu.js:705-727); returns the particle count, writes n*12 floats
static int semear_retangulo(double x0, double y0, double x1, double y1, double h, const double* abs3,
                            uint32_t* seed, float* out, int max_out) {
  double lado = js_round(std::sqrt((double)PARTICULAS_POR_CELULA));
  double passo = h / lado;
  double xa = std::min(x0, x1), xb = std::max(x0, x1);
  double ya = std::min(y0, y1), yb = std::max(y0, y1);
  double ia = std::ceil(xa / passo - 0.5), ib = std::floor(xb / passo - 0.5);
  double ja = std::ceil(ya / passo - 0.5), jb = std::floor(yb / passo - 0.5);
  double n = std::max(0.0, ib - ia + 1) * std::max(0.0, jb - ja + 1);
  if (!out) return (int)n;
  if (n > max_out) return -1;
  int k = 0;
  for (double j = ja; j <= jb; j++) {
    for (double i = ia; i <= ib; i++) {
      float* o = out + (size_t)k * CAMPOS;
      for (int q = 0; q < CAMPOS; q++) o[q] = 0.f;
      o[0] = (float)((i + 0.5 + (lcg_next(seed) - 0.5) * 0.5) * passo);
      o[1] = (float)((j + 0.5 + (lcg_next(seed) - 0.5) * 0.5) * passo);
      o[9] = (float)abs3[0]; o[10] = (float)abs3[1]; o[11] = (float)abs3[2];
      k++;
    }
  }
  return k;
}

// ---------------------------------------------------------------- RastreadorDeAr (js/ar.js)
template <class T>
struct Rastreador {
  int nx, ny;
  std::vector<uint32_t> rotulo;
  std::vector<int32_t> pai, componente, numero;
  std::vector<T> volumeAlvo;
  std::vector<uint8_t> vivo;
  bool primeira = true;
  Rastreador(int nx_, int ny_) : nx(nx_), ny(ny_) {
    size_t n = (size_t)nx * ny;
    rotulo.assign(n, 0); pai.assign(n, 0); componente.assign(n, 0); numero.assign(n, -1);
    volumeAlvo.assign(MAX_BOLHAS, 0); vivo.assign(MAX_BOLHAS, 0);
  }
  // reiniciar (ar.js:37-42)
  void reiniciar() {
    std::fill(rotulo.begin(), rotulo.end(), 0);
    std::fill(volumeAlvo.begin(), volumeAlvo.end(), (T)0);
    std::fill(vivo.begin(), vivo.end(), 0);
    primeira = true;
  }
  // #raiz with path halving (ar.js:44-48)
  int raiz(int a) {
    int32_t* p = pai.data();
    while (p[a] != a) { p[a] = p[p[a]]; a = p[a]; }
    return a;
  }
  // #unir (ar.js:155-158)
  void unir(int a, int b) {
    int ra = raiz(a), rb = raiz(b);
    if (ra != rb) pai[std::max(ra, rb)] = std::min(ra, rb);
  }
  struct Info {
    int celulas = 0; bool topo = false; int atm = 0;
    std::vector<std::pair<uint32_t, int>> gas;  // insertion-ordered Map(rotuloAntigo -> células)
  };
  // atualizar (ar.js:56-153).  `pressao_anterior` holds the old per-label pressures;
  // born bubbles are returned as (label, initial pressure).
  std::vector<std::pair<int, double>> atualizar(const uint8_t* tipo, const T* pressao_anterior) {
    int n = nx * ny;
    for (int c = 0; c < n; c++) pai[c] = tipo[c] == AR ? c : -1;
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int c = i + j * nx;
        if (pai[c] < 0) continue;
        if (i > 0 && pai[c - 1] >= 0) unir(c, c - 1);
        if (j > 0 && pai[c - nx] >= 0) unir(c, c - nx);
      }
    int topo = ny - MOLDURA - 1;
    std::vector<int> raizes;
    std::vector<Info> info;
    for (int c = 0; c < n; c++) {
      if (pai[c] < 0) { componente[c] = -1; continue; }
      int r = raiz(c);
      int k = numero[r];
      if (k < 0) { k = (int)info.size(); numero[r] = k; raizes.push_back(r); info.emplace_back(); }
      componente[c] = k;
      Info& d = info[k];
      d.celulas++;
      if (c / nx >= topo) d.topo = true;
      uint32_t antigo = rotulo[c];
      if (antigo == ATMOSFERA) d.atm++;
      else if (antigo >= 2) {
        bool found = false;
        for (auto& g : d.gas) if (g.first == antigo) { g.second++; found = true; break; }
        if (!found) d.gas.emplace_back(antigo, 1);
      }
    }
    for (int r : raizes) numero[r] = -1;
    auto fica = [](const Info& d) { return d.topo || d.celulas >= MIN_BOLHA_CELULAS; };
    std::map<uint32_t, double> total;  // sobreposicaoTotal; only looked up, order irrelevant
    for (auto& d : info) {
      if (d.gas.empty() || !fica(d)) continue;
      for (auto& g : d.gas) total[g.first] += g.second;
    }
    int K = (int)info.size();
    std::vector<T> novoAlvo(K, (T)0);
    std::vector<int> dono(K, -1);
    for (int k = 0; k < K; k++) {
      Info& d = info[k];
      if (!fica(d) || d.topo) continue;
      double v = 0; int melhor = 0;
      if (primeira) v = d.celulas;
      else {
        v = d.atm;
        for (auto& g : d.gas) {
          v += (double)volumeAlvo[g.first] * ((double)g.second / total[g.first]);
          if (g.second > melhor) { melhor = g.second; dono[k] = (int)g.first; }
        }
      }
      novoAlvo[k] = (T)v;
    }
    std::vector<uint32_t> rotuloDaComp(K, 0);
    std::vector<uint8_t> vivoNovo(MAX_BOLHAS, 0);
    std::vector<int> ordem;
    for (int k = 0; k < K; k++) if (!info[k].topo && (double)novoAlvo[k] >= MIN_BOLHA_CELULAS) ordem.push_back(k);
    std::stable_sort(ordem.begin(), ordem.end(), [&](int a, int b) { return (double)novoAlvo[b] < (double)novoAlvo[a]; });
    for (int k : ordem) {
      int r = dono[k];
      if (r >= 2 && !vivoNovo[r]) { rotuloDaComp[k] = r; vivoNovo[r] = 1; }
    }
    std::vector<std::pair<int, double>> nascidas;
    int livre = 2;
    for (int k : ordem) {
      if (rotuloDaComp[k]) continue;
      while (livre < MAX_BOLHAS && (vivoNovo[livre] || vivo[livre])) livre++;
      if (livre >= MAX_BOLHAS) break;
      rotuloDaComp[k] = livre;
      vivoNovo[livre] = 1;
      nascidas.emplace_back(livre, dono[k] >= 2 ? (double)pressao_anterior[dono[k]] : 0.0);
    }
    for (int k = 0; k < K; k++) if (info[k].topo) rotuloDaComp[k] = ATMOSFERA;
    std::fill(volumeAlvo.begin(), volumeAlvo.end(), (T)0);
    for (int k = 0; k < K; k++) if (rotuloDaComp[k] >= 2) volumeAlvo[rotuloDaComp[k]] = novoAlvo[k];
    for (int c = 0; c < n; c++) rotulo[c] = componente[c] >= 0 ? rotuloDaComp[componente[c]] : 0;
    vivo = vivoNovo;
    primeira = false;
    ncomp_last = K;
    return nascidas;
  }
  int ncomp_last = 0;
};

// ---------------------------------------------------------------- MGPCG (js/pressao-cpu.js)
template <class T>
struct Nivel {
  int nx, ny;
  std::vector<uint8_t> tipo, viz, cnt;
  std::vector<T> b, x, tmp, r;
  void alloc() {
    size_t k = (size_t)nx * ny;
    tipo.assign(k, 0); viz.assign(k, 0); cnt.assign(k, 0);
    b.assign(k, 0); x.assign(k, 0); tmp.assign(k, 0); r.assign(k, 0);
  }
};

// montarNiveis (pressao-cpu.js:24-41)
template <class T>
static std::vector<Nivel<T>> montar_niveis(int nx, int ny) {
  std::vector<Nivel<T>> nv;
  { Nivel<T> a; a.nx = nx; a.ny = ny; nv.push_back(a); }
  while (std::min(nv.back().nx, nv.back().ny) > 12) {
    Nivel<T> a; a.nx = (nv.back().nx + 1) / 2; a.ny = (nv.back().ny + 1) / 2;
    nv.push_back(a);
  }
  for (auto& n : nv) n.alloc();
  return nv;
}

// prepararVizinhos (pressao-cpu.js:44-60)
template <class T>
static void preparar_vizinhos(Nivel<T>& n) {
  int nx = n.nx, ny = n.ny;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int c = i + j * nx;
      int m = 0;
      if (i > 0 && n.tipo[c - 1] != SOLIDO) m |= 1;
      if (i < nx - 1 && n.tipo[c + 1] != SOLIDO) m |= 2;
      if (j > 0 && n.tipo[c - nx] != SOLIDO) m |= 4;
      if (j < ny - 1 && n.tipo[c + nx] != SOLIDO) m |= 8;
      n.viz[c] = (uint8_t)m;
      n.cnt[c] = (uint8_t)((m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1) + ((m >> 3) & 1));
    }
}

// tiposGrossos (pressao-cpu.js:63-81)
template <class T>
static void tipos_grossos(std::vector<Nivel<T>>& nv) {
  for (size_t l = 1; l < nv.size(); l++) {
    Nivel<T>& f = nv[l - 1]; Nivel<T>& g = nv[l];
    for (int J = 0; J < g.ny; J++)
      for (int I = 0; I < g.nx; I++) {
        bool temAr = false, temAgua = false;
        for (int b = 0; b < 2; b++)
          for (int a = 0; a < 2; a++) {
            int i = 2 * I + a, j = 2 * J + b;
            if (i >= f.nx || j >= f.ny) continue;
            uint8_t t = f.tipo[i + j * f.nx];
            if (t == AR) temAr = true; else if (t == AGUA) temAgua = true;
          }
        g.tipo[I + J * g.nx] = temAr ? AR : temAgua ? AGUA : SOLIDO;
      }
  }
}

// aplicar (pressao-cpu.js:84-93)
template <class T, class V>
static double aplicar(const Nivel<T>& n, const V* v, int i, int j) {
  int c = i + j * n.nx;
  int m = n.viz[c];
  double s = 0;
  if (m & 1) s += (double)v[c - 1];
  if (m & 2) s += (double)v[c + 1];
  if (m & 4) s += (double)v[c - n.nx];
  if (m & 8) s += (double)v[c + n.nx];
  return n.cnt[c] * (double)v[c] - s;
}

// suavizar (pressao-cpu.js:99-150): snapshot x -> tmp, then swap
template <class T>
static void suavizar(Nivel<T>& n, int desl, int c0, int c1, double omega) {
  int nx = n.nx, ny = n.ny;
  const T* src = n.x.data();
  T* dst = n.tmp.data();
  std::copy(n.x.begin(), n.x.end(), n.tmp.begin());
  const int L = BLOCO + 2;
  T local[L * L];
  int ordem[2] = {c0, c1};
  for (int oy = -desl; oy < ny; oy += BLOCO)
    for (int ox = -desl; ox < nx; ox += BLOCO) {
      for (int ly = 0; ly < L; ly++)
        for (int lx = 0; lx < L; lx++) {
          int i = ox + lx - 1, j = oy + ly - 1;
          local[lx + ly * L] = (i >= 0 && j >= 0 && i < nx && j < ny) ? src[i + j * nx] : (T)0;
        }
      for (int it = 0; it < VARREDURAS; it++)
        for (int o = 0; o < 2; o++) {
          int cor = ordem[o];
          for (int ly = 1; ly <= BLOCO; ly++)
            for (int lx = 1; lx <= BLOCO; lx++) {
              int i = ox + lx - 1, j = oy + ly - 1;
              if (i < 0 || j < 0 || i >= nx || j >= ny) continue;
              if (((i + j + 1) & 1) != cor) continue;
              int c = i + j * nx;
              if (n.tipo[c] != AGUA) continue;
              int k = n.cnt[c];
              if (k == 0) continue;
              int m = n.viz[c], e = lx + ly * L;
              double s = 0;
              if (m & 1) s += (double)local[e - 1];
              if (m & 2) s += (double)local[e + 1];
              if (m & 4) s += (double)local[e - L];
              if (m & 8) s += (double)local[e + L];
              double q = local[e];
              local[e] = (T)(q + omega * ((s + (double)n.b[c]) / k - q));
            }
        }
      for (int ly = 1; ly <= BLOCO; ly++)
        for (int lx = 1; lx <= BLOCO; lx++) {
          int i = ox + lx - 1, j = oy + ly - 1;
          if (i < 0 || j < 0 || i >= nx || j >= ny) continue;
          int c = i + j * nx;
          dst[c] = n.tipo[c] == AGUA ? local[lx + ly * L] : (T)0;
        }
    }
  std::swap(n.x, n.tmp);
}

static inline int grosso_base(int i) { return (i & 1) == 0 ? (i >> 1) - 1 : (i - 1) >> 1; }  // :155
static inline double peso_base(int i) { return (i & 1) == 0 ? 0.25 : 0.75; }               // :156

// restringir (pressao-cpu.js:159-188)
template <class T>
static void restringir(Nivel<T>& f, Nivel<T>& g) {
  int fnx = f.nx, fny = f.ny, gnx = g.nx, gny = g.ny;
  for (int j = 0; j < fny; j++)
    for (int i = 0; i < fnx; i++) {
      int c = i + j * fnx;
      f.r[c] = f.tipo[c] == AGUA ? (T)((double)f.b[c] - aplicar(f, f.x.data(), i, j)) : (T)0;
    }
  std::fill(g.b.begin(), g.b.end(), (T)0);
  std::fill(g.x.begin(), g.x.end(), (T)0);
  for (int j = 0; j < fny; j++) {
    int J0 = grosso_base(j); double wy0 = peso_base(j), wy1 = 1 - wy0;
    for (int i = 0; i < fnx; i++) {
      double r = f.r[i + j * fnx];
      if (r == 0) continue;
      int I0 = grosso_base(i); double wx0 = peso_base(i), wx1 = 1 - wx0;
      for (int b = 0; b < 2; b++) {
        int J = J0 + b;
        if (J < 0 || J >= gny) continue;
        double wy = b ? wy1 : wy0;
        for (int a = 0; a < 2; a++) {
          int I = I0 + a;
          if (I < 0 || I >= gnx) continue;
          int C = I + J * gnx;
          if (g.tipo[C] == AGUA) g.b[C] = (T)((double)g.b[C] + r * (a ? wx1 : wx0) * wy);
        }
      }
    }
  }
}

// prolongar (pressao-cpu.js:191-214)
template <class T>
static void prolongar(Nivel<T>& g, Nivel<T>& f) {
  int fnx = f.nx, fny = f.ny, gnx = g.nx, gny = g.ny;
  for (int j = 0; j < fny; j++) {
    int J0 = grosso_base(j); double wy0 = peso_base(j), wy1 = 1 - wy0;
    for (int i = 0; i < fnx; i++) {
      int c = i + j * fnx;
      if (f.tipo[c] != AGUA) continue;
      int I0 = grosso_base(i); double wx0 = peso_base(i), wx1 = 1 - wx0;
      double e = 0;
      for (int b = 0; b < 2; b++) {
        int J = J0 + b;
        if (J < 0 || J >= gny) continue;
        double wy = b ? wy1 : wy0;
        for (int a = 0; a < 2; a++) {
          int I = I0 + a;
          if (I < 0 || I >= gnx) continue;
          int C = I + J * gnx;
          if (g.tipo[C] == AGUA) e += (double)g.x[C] * (a ? wx1 : wx0) * wy;
        }
      }
      f.x[c] = (T)((double)f.x[c] + e);
    }
  }
}

// resolverGrosso (pressao-cpu.js:217-238)
template <class T>
static void resolver_grosso(Nivel<T>& n) {
  int nx = n.nx, ny = n.ny;
  T* x = n.x.data();
  for (int k = 0; k < VARREDURAS_GROSSO; k++)
    for (int o = 0; o < 2; o++) {
      int cor = (k & 1) ? 1 - o : o;
      for (int c = 0; c < nx * ny; c++) {
        int i = c % nx, j = c / nx;
        if (n.tipo[c] != AGUA || ((i + j + 1) & 1) != cor) continue;
        int q = n.cnt[c];
        if (!q) continue;
        int m = n.viz[c];
        double s = 0;
        if (m & 1) s += (double)x[c - 1];
        if (m & 2) s += (double)x[c + 1];
        if (m & 4) s += (double)x[c - nx];
        if (m & 8) s += (double)x[c + nx];
        x[c] = (T)((s + (double)n.b[c]) / q);
      }
    }
}

// cicloV (pressao-cpu.js:241-255)
template <class T>
static void ciclo_v(std::vector<Nivel<T>>& nv, double omega = 1) {
  int L = (int)nv.size();
  std::fill(nv[0].x.begin(), nv[0].x.end(), (T)0);
  for (int l = 0; l < L - 1; l++) {
    suavizar(nv[l], 0, 0, 1, omega);
    suavizar(nv[l], 8, 0, 1, omega);
    restringir(nv[l], nv[l + 1]);
  }
  resolver_grosso(nv[L - 1]);
  for (int l = L - 2; l >= 0; l--) {
    prolongar(nv[l + 1], nv[l]);
    suavizar(nv[l], 8, 1, 0, omega);
    suavizar(nv[l], 0, 1, 0, omega);
  }
}

// produto (pressao-cpu.js:257-261)
template <class T>
static double produto(const std::vector<T>& a, const std::vector<T>& b, const uint8_t* tipo) {
  double s = 0;
  for (size_t k = 0; k < a.size(); k++) if (tipo[k] == AGUA) s += (double)a[k] * (double)b[k];
  return s;
}

template <class T>
struct Solver {
  std::vector<Nivel<T>> niveis;
  Nivel<T> op;
  std::vector<T> r, z, zVelho, p, ap;
  std::vector<double> normas;
  int iters_done = 0;
  double norm_b = 0, true_resid = 0, proj_mean = 0;
  void init(int nx, int ny) {
    niveis = montar_niveis<T>(nx, ny);
    op = Nivel<T>();
    op.nx = nx; op.ny = ny;
    op.alloc();
    size_t k = (size_t)nx * ny;
    r.assign(k, 0); z.assign(k, 0); zVelho.assign(k, 0); p.assign(k, 0); ap.assign(k, 0);
  }
};

// Insertion-ordered map of bubble label -> double (a JS Map whose keys are labels).
struct BMap {
  std::vector<double> val = std::vector<double>(MAX_BOLHAS, 0.0);
  std::vector<uint8_t> has = std::vector<uint8_t>(MAX_BOLHAS, 0);
  void set(int b, double v) { val[b] = v; has[b] = 1; }
  double get(int b) const { return val[b]; }
};

// resolverPCG (pressao-cpu.js:296-402) with montarBolhas (271-288).  tol > 0 switches to a
// residual-terminated loop (stop when ||[r;rb]|| <= tol*||[-div;rhs]||, at most maxit
// iterations), which the reference lacks; tol <= 0 is the reference's fixed-count loop.
template <class T>
static void resolver_pcg(Solver<T>& S, const uint8_t* tipo, const T* div, T* q, int iteracoes,
                         const uint32_t* rot, const T* rhsB, T* yB, double tol) {
  Nivel<T>& n0 = S.niveis[0];
  int nx = n0.nx, ny = n0.ny, k = nx * ny;
  Nivel<T>& op = S.op;
  std::copy(tipo, tipo + k, op.tipo.begin());
  preparar_vizinhos(op);
  std::copy(tipo, tipo + k, n0.tipo.begin());
  if (rot) for (int c = 0; c < k; c++) if (tipo[c] == AR && rot[c] >= 2) n0.tipo[c] = AGUA;
  tipos_grossos(S.niveis);
  for (auto& n : S.niveis) preparar_vizinhos(n);
  std::vector<T>& r = S.r; std::vector<T>& z = S.z; std::vector<T>& zVelho = S.zVelho;
  std::vector<T>& p = S.p; std::vector<T>& ap = S.ap;
  // montarBolhas: pares in raster order, faces counted per label, ativos in first-seen order
  std::vector<std::pair<int, int>> pares;
  std::vector<int> faces(MAX_BOLHAS, 0);
  std::vector<int> ativos;
  if (rot) {
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        int c = i + j * nx;
        uint32_t b = rot[c];
        if (b < 2 || tipo[c] == AGUA || tipo[c] == SOLIDO) continue;
        if (tipo[c - 1] == AGUA) pares.emplace_back(b, c - 1);
        if (tipo[c + 1] == AGUA) pares.emplace_back(b, c + 1);
        if (tipo[c - nx] == AGUA) pares.emplace_back(b, c - nx);
        if (tipo[c + nx] == AGUA) pares.emplace_back(b, c + nx);
      }
    for (auto& pr : pares) { if (faces[pr.first] == 0) ativos.push_back(pr.first); faces[pr.first]++; }
  }
  BMap yb, rb, zb, zvb, pb, apb;
  for (int b : ativos) yb.set(b, (double)yB[b]);
  auto difundir = [&](std::vector<T>& grade, const BMap& valores) {
    if (!rot) return;
    for (int c = 0; c < k; c++) {
      uint32_t b = rot[c];
      if (b >= 2 && tipo[c] == AR) grade[c] = (T)(valores.has[b] ? valores.get(b) : 0.0);
    }
  };
  auto difundir_ptr = [&](T* grade, const BMap& valores) {
    if (!rot) return;
    for (int c = 0; c < k; c++) {
      uint32_t b = rot[c];
      if (b >= 2 && tipo[c] == AR) grade[c] = (T)(valores.has[b] ? valores.get(b) : 0.0);
    }
  };
  auto aplicar_bolhas = [&](const T* grade, const BMap& valores, BMap& saida) {
    for (int b : ativos) saida.set(b, faces[b] * valores.get(b));
    for (auto& pr : pares) saida.set(pr.first, saida.get(pr.first) - (double)grade[pr.second]);
  };
  auto produto_b = [&](const BMap& a, const BMap& b) {
    double s = 0;
    for (int x : ativos) s += a.get(x) * b.get(x);
    return s;
  };
  for (int c = 0; c < k; c++) if (tipo[c] != AGUA) q[c] = 0;
  difundir_ptr(q, yb);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int c = i + j * nx;
      r[c] = tipo[c] == AGUA ? (T)(-(double)div[c] - aplicar(op, q, i, j)) : (T)0;
    }
  BMap aq;
  aplicar_bolhas(q, yb, aq);
  for (int b : ativos) rb.set(b, (double)rhsB[b] - aq.get(b));
  auto norma = [&]() { return std::sqrt(produto(r, r, tipo) + produto_b(rb, rb)); };
  // Residual-terminated mode only (an addition): when no water cell touches Dirichlet air
  // (atmosphere, vacuum or an untracked pocket), the operator is singular with the constant
  // over all unknowns as null vector; its component is removed from b so the tolerance is
  // reachable.  S.proj_mean records it.
  S.proj_mean = 0;
  if (tol > 0) {
    bool anchored = false;
    double sb = 0, nunk = 0;
    for (int j = 0; j < ny && !anchored; j++)
      for (int i = 0; i < nx; i++) {
        int c = i + j * nx;
        if (tipo[c] != AGUA) continue;
        int m = op.viz[c];
        const int nb[4] = {c - 1, c + 1, c - nx, c + nx};
        for (int q2 = 0; q2 < 4; q2++) {
          if (!(m & (1 << q2))) continue;
          int d = nb[q2];
          if (tipo[d] != AR) continue;
          bool bub = rot && rot[d] >= 2 && faces[rot[d]] > 0;
          if (!bub) anchored = true;
        }
      }
    if (!anchored) {
      for (int c = 0; c < k; c++) if (tipo[c] == AGUA) { sb += -(double)div[c]; nunk += 1; }
      for (int b : ativos) { sb += (double)rhsB[b]; nunk += 1; }
      if (nunk > 0) {
        double mean = sb / nunk;
        S.proj_mean = mean;
        for (int c = 0; c < k; c++) if (tipo[c] == AGUA) r[c] = (T)((double)r[c] - mean);
        for (int b : ativos) rb.set(b, rb.get(b) - mean);
      }
    }
  }
  S.normas.clear();
  S.normas.push_back(norma());
  // ||b|| with b = [-div ; rhs] (projected as above; reported and used by the tolerance mode)
  {
    double s = 0;
    for (int c = 0; c < k; c++) if (tipo[c] == AGUA) { double e = -(double)div[c] - S.proj_mean; s += e * e; }
    for (int b : ativos) { double e = (double)rhsB[b] - S.proj_mean; s += e * e; }
    S.norm_b = std::sqrt(s);
  }
  std::vector<int> celulasDe(MAX_BOLHAS, 0);
  if (rot)
    for (int c = 0; c < k; c++) {
      uint32_t b = rot[c];
      if (b >= 2 && tipo[c] == AR && faces[b] > 0) celulasDe[b]++;
    }
  auto precondicionar = [&](std::vector<T>& destino, BMap& destinoB) {
    std::copy(r.begin(), r.end(), n0.b.begin());
    if (rot)
      for (int c = 0; c < k; c++) {
        uint32_t b = rot[c];
        if (b >= 2 && celulasDe[b] && tipo[c] == AR) n0.b[c] = (T)(rb.get(b) / celulasDe[b]);
      }
    ciclo_v(S.niveis);
    std::copy(n0.x.begin(), n0.x.end(), destino.begin());
    for (int b : ativos) destinoB.set(b, 0);
    if (rot)
      for (int c = 0; c < k; c++) {
        uint32_t b = rot[c];
        if (b >= 2 && celulasDe[b] && tipo[c] == AR) {
          destinoB.set(b, destinoB.get(b) + (double)n0.x[c] / celulasDe[b]);
          destino[c] = 0;
        }
      }
  };
  bool tol_mode = tol > 0;
  int maxit = iteracoes;
  S.iters_done = 0;
  bool skip = tol_mode && (S.normas[0] == 0 || S.normas[0] <= tol * S.norm_b);
  if (!skip && maxit > 0) {
    precondicionar(z, zb);
    std::copy(z.begin(), z.end(), p.begin());
    for (int b : ativos) pb.set(b, zb.get(b));
    double rz = produto(r, z, tipo) + produto_b(rb, zb);
    for (int it = 0; it < maxit; it++) {
      difundir(p, pb);
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          int c = i + j * nx;
          ap[c] = tipo[c] == AGUA ? (T)aplicar(op, p.data(), i, j) : (T)0;
        }
      aplicar_bolhas(p.data(), pb, apb);
      double pap = produto(p, ap, tipo) + produto_b(pb, apb);
      double alfa = std::fabs(pap) < 1e-30 ? 0 : rz / pap;
      for (int c = 0; c < k; c++) {
        if (tipo[c] != AGUA) continue;
        q[c] = (T)((double)q[c] + alfa * (double)p[c]);
        r[c] = (T)((double)r[c] - alfa * (double)ap[c]);
      }
      for (int b : ativos) {
        yb.set(b, yb.get(b) + alfa * pb.get(b));
        rb.set(b, rb.get(b) - alfa * apb.get(b));
      }
      double nr = norma();
      S.normas.push_back(nr);
      S.iters_done = it + 1;
      if (it == maxit - 1) break;
      if (tol_mode && (nr == 0 || nr <= tol * S.norm_b)) break;
      std::copy(z.begin(), z.end(), zVelho.begin());
      for (int b : ativos) zvb.set(b, zb.get(b));
      precondicionar(z, zb);
      double rzNovo = produto(r, z, tipo) + produto_b(rb, zb);
      double cruzado = produto(r, zVelho, tipo) + produto_b(rb, zvb);
      double beta = std::fabs(rz) < 1e-30 ? 0 : std::max(0.0, (rzNovo - cruzado) / rz);
      rz = rzNovo;
      for (int c = 0; c < k; c++) p[c] = tipo[c] == AGUA ? (T)((double)z[c] + beta * (double)p[c]) : (T)0;
      for (int b : ativos) pb.set(b, zb.get(b) + beta * pb.get(b));
    }
  }
  // true residual ||b - A[q;y]|| (reporting only; the reference keeps the recursive one)
  {
    std::vector<T> qq(q, q + k);
    difundir(qq, yb);
    double s = 0;
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int c = i + j * nx;
        if (tipo[c] != AGUA) continue;
        double e = (-(double)div[c] - S.proj_mean) - aplicar(op, qq.data(), i, j);
        s += e * e;
      }
    BMap aq2;
    aplicar_bolhas(qq.data(), yb, aq2);
    for (int b : ativos) { double e = ((double)rhsB[b] - S.proj_mean) - aq2.get(b); s += e * e; }
    S.true_resid = std::sqrt(s);
  }
  for (int c = 0; c < k; c++) if (tipo[c] != AGUA) q[c] = 0;
  if (rot) for (int b : ativos) yB[b] = (T)yb.get(b);
}

// Test hooks (no counterpart in the reference): the level setup resolverPCG performs
// (pressao-cpu.js:303-312) on the current types and labels, then one V-cycle on a given
// level-0 right-hand side, or the operator A (aplicar, pressao-cpu.js:84-93) on a vector.
template <class T>
static void mg_setup(Solver<T>& S, const uint8_t* tipo, const uint32_t* rot) {
  Nivel<T>& n0 = S.niveis[0];
  const int k = n0.nx * n0.ny;
  std::copy(tipo, tipo + k, S.op.tipo.begin());
  preparar_vizinhos(S.op);
  std::copy(tipo, tipo + k, n0.tipo.begin());
  if (rot) for (int c = 0; c < k; c++) if (tipo[c] == AR && rot[c] >= 2) n0.tipo[c] = AGUA;
  tipos_grossos(S.niveis);
  for (auto& n : S.niveis) preparar_vizinhos(n);
}
template <class T>
static void mg_vcycle_hook(Solver<T>& S, const uint8_t* tipo, const uint32_t* rot, const double* in, double* out) {
  mg_setup(S, tipo, rot);
  Nivel<T>& n0 = S.niveis[0];
  const int k = n0.nx * n0.ny;
  for (int c = 0; c < k; c++) n0.b[c] = (T)in[c];
  ciclo_v(S.niveis);
  for (int c = 0; c < k; c++) out[c] = (double)n0.x[c];
}
template <class T>
static void op_apply_hook(Solver<T>& S, const uint8_t* tipo, const double* in, double* out) {
  mg_setup(S, tipo, nullptr);
  const int nx = S.op.nx, ny = S.op.ny;
  std::vector<T> v(in, in + (size_t)nx * ny);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int c = i + j * nx;
      out[c] = tipo[c] == AGUA ? aplicar(S.op, v.data(), i, j) : 0.0;
    }
}

// ---------------------------------------------------------------- SimulacaoCPU (js/fisica-cpu.js)
template <class T>
struct Sim {
  Cena* cena;
  int nx, ny, nu, nv;
  double h;
  double gravidade[2];
  int iteracoes;
  double tauDensidade, tensao;
  bool arAtivo;
  double tol = 0;  // > 0: residual-terminated PCG (addition, see resolver_pcg)
  std::vector<T> u, v, uPeso, vPeso, uTmp, vTmp, densidade, corCelula, div, q;
  std::vector<T> fase, faseTmp, normalX, normalY, curvatura, qAr;
  std::vector<uint8_t> uValida, vValida, tipo, solidoCelula, preenchido;
  std::vector<int32_t> novas;
  std::vector<uint32_t> contagemCelula;
  Rastreador<T> rastreador;
  std::vector<T> pressaoBolha, volumeBolha, rhsBolha, yBolha;
  Solver<T> solver;
  int n = 0, capacidade = 0;
  std::vector<T> p;
  double vMax = 0; bool hasVMax = false;
  bool tiposProntos = false;

  Sim(Cena* c, double gx, double gy, int it, double tau, double ten, bool ar)
      : cena(c), nx(c->nx), ny(c->ny), h(c->h), rastreador(c->nx, c->ny) {
    gravidade[0] = gx; gravidade[1] = gy;
    iteracoes = it; tauDensidade = tau; tensao = ten; arAtivo = ar;
    nu = (nx + 1) * ny; nv = nx * (ny + 1);
    size_t C = (size_t)nx * ny;
    u.assign(nu, 0); v.assign(nv, 0); uPeso.assign(nu, 0); vPeso.assign(nv, 0);
    uValida.assign(nu, 0); vValida.assign(nv, 0); uTmp.assign(nu, 0); vTmp.assign(nv, 0);
    densidade.assign(C, 0); contagemCelula.assign(C, 0); corCelula.assign(3 * C, 0);
    tipo.assign(C, 0); solidoCelula.assign(C, 0); div.assign(C, 0); q.assign(C, 0);
    fase.assign(C, 0); faseTmp.assign(C, 0); normalX.assign(C, 0); normalY.assign(C, 0);
    curvatura.assign(C, 0); qAr.assign(C, 0); preenchido.assign(C, 0); novas.assign(C, 0);
    pressaoBolha.assign(MAX_BOLHAS, 0); volumeBolha.assign(MAX_BOLHAS, 0);
    rhsBolha.assign(MAX_BOLHAS, 0); yBolha.assign(MAX_BOLHAS, 0);
    solver.init(nx, ny);
    garantir(1024);
    atualizar_paredes();
  }
  // atualizarParedes (fisica-cpu.js:80-85)
  void atualizar_paredes() {
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) solidoCelula[i + j * nx] = cena->celula_solida(i, j) ? 1 : 0;
  }
  // #garantir (fisica-cpu.js:663-670)
  void garantir(int m) {
    if (m <= capacidade) return;
    int cap = std::max(m, capacidade * 2);
    p.resize((size_t)cap * CAMPOS, (T)0);
    capacidade = cap;
  }
  // adicionar (fisica-cpu.js:92-111)
  int adicionar(const float* lista, int total, bool checarCheia) {
    garantir(n + total);
    int aceitas = 0;
    for (int k = 0; k < total; k++) {
      const float* o = lista + (size_t)k * CAMPOS;
      double x = o[0], y = o[1];
      if (!(cena->distancia(x, y) > RAIO_PARTICULA + 0.05)) continue;
      if (checarCheia) {
        double i = std::floor(x / h), j = std::floor(y / h);
        if (i < 0 || j < 0 || i >= nx || j >= ny) continue;
        if (contagemCelula[(size_t)i + (size_t)j * nx] >= (uint32_t)(PARTICULAS_POR_CELULA - 1)) continue;
      }
      for (int q2 = 0; q2 < CAMPOS; q2++) p[(size_t)n * CAMPOS + q2] = (T)o[q2];
      n++;
      aceitas++;
    }
    return aceitas;
  }
  // emitirTorneira (principal.js:301-321) with the seeded LCG (testes/apoio.js:4-9) in place of
  // Math.random: the list the app builds (velocity from the pointer's, clamped to 4 m/s; count
  // from the flow; uniform in the disc: angle then radius), then adicionar(lista, true)
  int emitir_torneira(double px, double py, double pvx, double pvy, double raio, const double* abs3, double dt,
                      uint32_t* seed, int* emitted) {
    double vx = pvx * 0.5, vy = pvy * 0.5 - 0.7;
    const double v = js_hypot(vx, vy);
    if (v > 4) {
      vx *= 4 / v;
      vy *= 4 / v;
    }
    const double vazao = 2 * raio * std::max(0.7, js_hypot(vx, vy));
    const double nd = std::min(40000.0, js_round((vazao * dt) / (h * h) * PARTICULAS_POR_CELULA));
    const int nn = nd > 0 ? (int)nd : 0;
    *emitted = nn;
    if (nn <= 0) return 0;
    std::vector<float> lista((size_t)nn * CAMPOS, 0.f);
    for (int k = 0; k < nn; k++) {
      const double a = lcg_next(seed) * M_PI * 2, r = std::sqrt(lcg_next(seed)) * raio;
      float* o = lista.data() + (size_t)k * CAMPOS;
      o[0] = (float)(px + std::cos(a) * r);
      o[1] = (float)(py + std::sin(a) * r);
      o[2] = (float)vx;
      o[3] = (float)vy;
      o[9] = (float)abs3[0];
      o[10] = (float)abs3[1];
      o[11] = (float)abs3[2];
    }
    return adicionar(lista.data(), nn, true);
  }
  // remover (fisica-cpu.js:114-126) with a predicate on (x, y, index)
  template <class F>
  int remover(F fn) {
    int w = 0;
    for (int k = 0; k < n; k++) {
      size_t o = (size_t)k * CAMPOS;
      if (fn((double)p[o], (double)p[o + 1], k)) continue;
      if (w != k) std::copy(p.begin() + o, p.begin() + o + CAMPOS, p.begin() + (size_t)w * CAMPOS);
      w++;
    }
    int rem = n - w;
    n = w;
    return rem;
  }
  // aplicarFerramenta (fisica-cpu.js:129-152); tipo: 1 esponja, 2 mao, 3 corante
  void aplicar_ferramenta(int tipoF, double fx, double fy, double raio, double fvx, double fvy, const double* abs3) {
    if (tipoF == 0) return;
    double r2 = raio * raio;
    if (tipoF == 1) {
      remover([&](double x, double y, int) { return (x - fx) * (x - fx) + (y - fy) * (y - fy) < r2; });
      return;
    }
    for (int k = 0; k < n; k++) {
      size_t o = (size_t)k * CAMPOS;
      double dx = (double)p[o] - fx, dy = (double)p[o + 1] - fy;
      double d2 = dx * dx + dy * dy;
      if (d2 >= r2) continue;
      double peso = 1 - std::sqrt(d2 / r2);
      if (tipoF == 2) {
        double a = std::min(1.0, peso * 0.9);
        p[o + 2] = (T)((double)p[o + 2] + (fvx - (double)p[o + 2]) * a);
        p[o + 3] = (T)((double)p[o + 3] + (fvy - (double)p[o + 3]) * a);
      } else if (tipoF == 3) {
        p[o + 9] = (T)abs3[0]; p[o + 10] = (T)abs3[1]; p[o + 11] = (T)abs3[2];
      }
    }
  }
  // avancar (fisica-cpu.js:155-159)
  void avancar(double dt, int subpassos) {
    if (arAtivo) atualizar_ar();
    double d = dt / subpassos;
    for (int s = 0; s < subpassos; s++) subpasso(d);
  }
  // atualizarAr (fisica-cpu.js:162-167)
  void atualizar_ar() {
    if (!tiposProntos) { para_grade(); tipos(); }
    auto nascidas = rastreador.atualizar(tipo.data(), pressaoBolha.data());
    for (auto& nb : nascidas) pressaoBolha[nb.first] = (T)nb.second;
  }
  // reiniciarAr (fisica-cpu.js:170-174)
  void reiniciar_ar() {
    tiposProntos = false;
    rastreador.reiniciar();
    std::fill(pressaoBolha.begin(), pressaoBolha.end(), (T)0);
  }
  // subpasso (fisica-cpu.js:176-189)
  void subpasso(double dt) {
    para_grade();
    gravidade_e_fronteira(dt);
    tipos();
    if (arAtivo) propagar_rotulos();
    superficie();
    pressao_do_ar(dt);
    classificar(dt);
    if (arAtivo) lado_direito_das_bolhas(dt);
    pressao(dt);
    gradiente();
    extrapolar();
    para_particulas(dt);
  }
  void phase(int ph, double dt) {
    switch (ph) {
      case 0: para_grade(); break;
      case 1: gravidade_e_fronteira(dt); break;
      case 2: tipos(); break;
      case 3: propagar_rotulos(); break;
      case 4: superficie(); break;
      case 5: pressao_do_ar(dt); break;
      case 6: classificar(dt); break;
      case 7: lado_direito_das_bolhas(dt); break;
      case 8: pressao(dt); break;
      case 9: gradiente(); break;
      case 10: extrapolar(); break;
      case 11: para_particulas(dt); break;
      case 12: atualizar_ar(); break;
      default: break;
    }
  }

  // #paraGrade (fisica-cpu.js:198-267)
  void para_grade() {
    std::fill(u.begin(), u.end(), (T)0); std::fill(v.begin(), v.end(), (T)0);
    std::fill(uPeso.begin(), uPeso.end(), (T)0); std::fill(vPeso.begin(), vPeso.end(), (T)0);
    std::fill(densidade.begin(), densidade.end(), (T)0);
    std::fill(contagemCelula.begin(), contagemCelula.end(), 0u);
    std::fill(corCelula.begin(), corCelula.end(), (T)0);
    double inv = 1 / h;
    for (int k = 0; k < n; k++) {
      const T* o = p.data() + (size_t)k * CAMPOS;
      double x = o[0], y = o[1], vx = o[2], vy = o[3];
      double cux = o[4], cuy = o[5], cvx = o[6], cvy = o[7];
      {
        double fx = x * inv, fy = y * inv - 0.5;
        double i0 = clampi(std::floor(fx), 0, nx - 1), j0 = clampi(std::floor(fy), 0, ny - 2);
        double tx = fx - i0, ty = fy - j0;
        for (int b = 0; b < 2; b++) {
          double wy = b ? ty : 1 - ty;
          double dy = (j0 + b + 0.5) * h - y;
          for (int a = 0; a < 2; a++) {
            double w = (a ? tx : 1 - tx) * wy;
            double dx = (i0 + a) * h - x;
            size_t idx = (size_t)(i0 + a + (j0 + b) * (nx + 1));
            u[idx] = (T)((double)u[idx] + w * (vx + cux * dx + cuy * dy));
            uPeso[idx] = (T)((double)uPeso[idx] + w);
          }
        }
      }
      {
        double fx = x * inv - 0.5, fy = y * inv;
        double i0 = clampi(std::floor(fx), 0, nx - 2), j0 = clampi(std::floor(fy), 0, ny - 1);
        double tx = fx - i0, ty = fy - j0;
        for (int b = 0; b < 2; b++) {
          double wy = b ? ty : 1 - ty;
          double dy = (j0 + b) * h - y;
          for (int a = 0; a < 2; a++) {
            double w = (a ? tx : 1 - tx) * wy;
            double dx = (i0 + a + 0.5) * h - x;
            size_t idx = (size_t)(i0 + a + (j0 + b) * nx);
            v[idx] = (T)((double)v[idx] + w * (vy + cvx * dx + cvy * dy));
            vPeso[idx] = (T)((double)vPeso[idx] + w);
          }
        }
      }
      {
        double fx = x * inv - 0.5, fy = y * inv - 0.5;
        double i0 = clampi(std::floor(fx), 0, nx - 2), j0 = clampi(std::floor(fy), 0, ny - 2);
        double tx = fx - i0, ty = fy - j0;
        size_t c = (size_t)(i0 + j0 * nx);
        T* d = densidade.data();
        d[c] = (T)((double)d[c] + (1 - tx) * (1 - ty));
        d[c + 1] = (T)((double)d[c + 1] + tx * (1 - ty));
        d[c + nx] = (T)((double)d[c + nx] + (1 - tx) * ty);
        d[c + nx + 1] = (T)((double)d[c + nx + 1] + tx * ty);
      }
      double ci = clampi(std::floor(x * inv), 0, nx - 1), cj = clampi(std::floor(y * inv), 0, ny - 1);
      size_t cc = (size_t)(ci + cj * nx);
      contagemCelula[cc]++;
      corCelula[cc * 3] = (T)((double)corCelula[cc * 3] + (double)o[9]);
      corCelula[cc * 3 + 1] = (T)((double)corCelula[cc * 3 + 1] + (double)o[10]);
      corCelula[cc * 3 + 2] = (T)((double)corCelula[cc * 3 + 2] + (double)o[11]);
    }
    for (int k = 0; k < nu; k++) if (uPeso[k] > 0) u[k] = (T)((double)u[k] / (double)uPeso[k]);
    for (int k = 0; k < nv; k++) if (vPeso[k] > 0) v[k] = (T)((double)v[k] / (double)vPeso[k]);
  }
  // #gravidadeEFronteira (fisica-cpu.js:271-291)
  void gravidade_e_fronteira(double dt) {
    double gx = gravidade[0], gy = gravidade[1];
    const uint8_t* sol = solidoCelula.data();
    for (int j = 0; j < ny; j++)
      for (int i = 0; i <= nx; i++) {
        int k = i + j * (nx + 1);
        if (uPeso[k] > 0) u[k] = (T)((double)u[k] + gx * dt);
        bool s = i == 0 || i == nx || sol[i - 1 + j * nx] || sol[i + j * nx];
        if (s) u[k] = 0;
      }
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i < nx; i++) {
        int k = i + j * nx;
        if (vPeso[k] > 0) v[k] = (T)((double)v[k] + gy * dt);
        bool s = j == 0 || j == ny || sol[i + (j - 1) * nx] || sol[i + j * nx];
        if (s) v[k] = 0;
      }
  }
  // #tipos (fisica-cpu.js:293-299)
  void tipos() {
    tiposProntos = true;
    int C = nx * ny;
    for (int c = 0; c < C; c++) tipo[c] = solidoCelula[c] ? SOLIDO : contagemCelula[c] > 0 ? AGUA : AR;
  }
  // #superficie (fisica-cpu.js:306-346)
  void superficie() {
    T* f = fase.data(); T* t = faseTmp.data();
    for (int c = 0; c < nx * ny; c++)
      f[c] = tipo[c] == SOLIDO ? (T)0 : (T)std::min(1.0, (double)densidade[c] / PARTICULAS_POR_CELULA);
    const int passadas = 2;
    for (int passada = 0; passada <= passadas; passada++) {
      estender_na_parede(f);
      if (passada == passadas) break;
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          int c = i + j * nx;
          int im = i > 0 ? c - 1 : c, ip = i < nx - 1 ? c + 1 : c;
          t[c] = (T)(((double)f[im] + 2 * (double)f[c] + (double)f[ip]) * 0.25);
        }
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          int c = i + j * nx;
          int jm = j > 0 ? c - nx : c, jp = j < ny - 1 ? c + nx : c;
          f[c] = (T)(((double)t[jm] + 2 * (double)t[c] + (double)t[jp]) * 0.25);
        }
    }
    T* nX = normalX.data(); T* nY = normalY.data(); T* kap = curvatura.data();
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        int c = i + j * nx;
        double gx = ((double)f[c + 1] - (double)f[c - 1]) * 0.5, gy = ((double)f[c + nx] - (double)f[c - nx]) * 0.5;
        double g = js_hypot(gx, gy);
        nX[c] = g > 1e-3 ? (T)(gx / g) : (T)0;
        nY[c] = g > 1e-3 ? (T)(gy / g) : (T)0;
      }
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        int c = i + j * nx;
        kap[c] = (T)(-(((double)nX[c + 1] - (double)nX[c - 1]) + ((double)nY[c + nx] - (double)nY[c - nx])) / (2 * h));
      }
  }
  // #estenderNaParede (fisica-cpu.js:353-374)
  void estender_na_parede(T* f) {
    uint8_t* cheio = preenchido.data();
    for (int c = 0; c < nx * ny; c++) cheio[c] = tipo[c] == SOLIDO ? 0 : 1;
    for (int camada = 0; camada < 3; camada++) {
      int m = 0;
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          int c = i + j * nx;
          if (cheio[c]) continue;
          double s = 0; int k = 0;
          if (i > 0 && cheio[c - 1]) { s += (double)f[c - 1]; k++; }
          if (i < nx - 1 && cheio[c + 1]) { s += (double)f[c + 1]; k++; }
          if (j > 0 && cheio[c - nx]) { s += (double)f[c - nx]; k++; }
          if (j < ny - 1 && cheio[c + nx]) { s += (double)f[c + nx]; k++; }
          if (k) { faseTmp[c] = (T)(s / k); novas[m++] = c; }
        }
      for (int qq = 0; qq < m; qq++) { int c = novas[qq]; f[c] = faseTmp[c]; cheio[c] = 1; }
    }
  }
  // #pressaoDoAr (fisica-cpu.js:381-403)
  void pressao_do_ar(double dt) {
    double escala = dt / (DENSIDADE_AGUA * h);
    double kMax = 1 / h;
    std::fill(qAr.begin(), qAr.end(), (T)0);
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        int c = i + j * nx;
        if (tipo[c] != AR) continue;
        double pp = 0;
        if (tensao > 0) {
          double s = curvatura[c]; int k = 1;
          if (tipo[c - 1] == AGUA) { s += (double)curvatura[c - 1]; k++; }
          if (tipo[c + 1] == AGUA) { s += (double)curvatura[c + 1]; k++; }
          if (tipo[c - nx] == AGUA) { s += (double)curvatura[c - nx]; k++; }
          if (tipo[c + nx] == AGUA) { s += (double)curvatura[c + nx]; k++; }
          if (k > 1) pp += tensao * std::max(-kMax, std::min(kMax, s / k));
        }
        qAr[c] = (T)(pp * escala);
      }
  }
  // #propagarRotulos (fisica-cpu.js:409-422)
  void propagar_rotulos() {
    uint32_t* rot = rastreador.rotulo.data();
    for (int c = 0; c < nx * ny; c++) if (tipo[c] != AR) rot[c] = 0;
    for (int passada = 0; passada < 2; passada++)
      for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
          int c = i + j * nx;
          if (tipo[c] != AR || rot[c]) continue;
          uint32_t r = rot[c - 1];
          if (!r) r = rot[c + 1];
          if (!r) r = rot[c - nx];
          if (!r) r = rot[c + nx];
          rot[c] = r;
        }
  }
  // #ladoDireitoDasBolhas (fisica-cpu.js:429-454)
  void lado_direito_das_bolhas(double dt) {
    const uint32_t* rot = rastreador.rotulo.data();
    std::fill(rhsBolha.begin(), rhsBolha.end(), (T)0);
    std::fill(volumeBolha.begin(), volumeBolha.end(), (T)0);
    T* rhs = rhsBolha.data(); T* vol = volumeBolha.data();
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        int c = i + j * nx;
        uint32_t r = rot[c];
        if (tipo[c] != AR || r < 2) continue;
        vol[r] = (T)((double)vol[r] + 1);
        double t = qAr[c];
        if (tipo[c - 1] == AGUA) rhs[r] = (T)((double)rhs[r] + ((double)u[i + j * (nx + 1)] - t));
        if (tipo[c + 1] == AGUA) rhs[r] = (T)((double)rhs[r] + (-(double)u[i + 1 + j * (nx + 1)] - t));
        if (tipo[c - nx] == AGUA) rhs[r] = (T)((double)rhs[r] + ((double)v[i + j * nx] - t));
        if (tipo[c + nx] == AGUA) rhs[r] = (T)((double)rhs[r] + (-(double)v[i + (j + 1) * nx] - t));
      }
    double paraQ = dt / (DENSIDADE_AGUA * h);
    for (int r = 2; r < MAX_BOLHAS; r++) {
      if (!vol[r]) continue;
      rhs[r] = (T)((double)rhs[r] + VOLTA_VOLUME_BOLHA * ((double)rastreador.volumeAlvo[r] - (double)vol[r]) * h / dt);
      yBolha[r] = (T)((double)pressaoBolha[r] * paraQ);
    }
  }
  // #classificar + #miolo (fisica-cpu.js:456-494)
  bool miolo(int i, int j) const {
    if (i < 1 || j < 1 || i >= nx - 1 || j >= ny - 1) return false;
    for (int b = -1; b <= 1; b++)
      for (int a = -1; a <= 1; a++) {
        int c = i + a + (j + b) * nx;
        if (solidoCelula[c] || contagemCelula[c] == 0) return false;
      }
    return true;
  }
  void classificar(double dt) {
    double alvo = h / std::max(tauDensidade, dt);
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int c = i + j * nx;
        if (tipo[c] != AGUA) { div[c] = 0; q[c] = 0; continue; }
        double d = (double)u[i + 1 + j * (nx + 1)] - (double)u[i + j * (nx + 1)] +
                   (double)v[i + (j + 1) * nx] - (double)v[i + j * nx];
        double excesso = (double)densidade[c] / PARTICULAS_POR_CELULA - 1;
        if (excesso < 0 && !miolo(i, j)) excesso = 0;
        excesso = std::min(EXCESSO_MAX, std::max(EXCESSO_MIN, excesso));
        double ar = 0;
        if (tipo[c - 1] == AR) ar += (double)qAr[c - 1];
        if (tipo[c + 1] == AR) ar += (double)qAr[c + 1];
        if (tipo[c - nx] == AR) ar += (double)qAr[c - nx];
        if (tipo[c + nx] == AR) ar += (double)qAr[c + nx];
        div[c] = (T)(d - alvo * excesso - ar);
      }
  }
  // #pressao (fisica-cpu.js:498-515)
  void pressao(double dt) {
    bool bolhas = arAtivo;
    int maxit = iteracoes;
    resolver_pcg(solver, tipo.data(), div.data(), q.data(), maxit,
                 bolhas ? rastreador.rotulo.data() : nullptr, rhsBolha.data(), yBolha.data(), tol);
    double paraPa = (DENSIDADE_AGUA * h) / dt;
    const uint32_t* rot = rastreador.rotulo.data();
    for (size_t c = 0; c < q.size(); c++) {
      if (tipo[c] != AR) continue;
      uint32_t r = rot[c];
      q[c] = (T)((double)qAr[c] + ((bolhas && r >= 2 && (double)volumeBolha[r] != 0) ? (double)yBolha[r] : 0.0));
    }
    if (bolhas)
      for (int r = 2; r < MAX_BOLHAS; r++)
        if ((double)volumeBolha[r] != 0)
          pressaoBolha[r] = (T)std::max(-PRESSAO_BOLHA_MAX, std::min(PRESSAO_BOLHA_MAX, (double)yBolha[r] * paraPa));
  }
  // #gradiente (fisica-cpu.js:519-545)
  void gradiente() {
    for (int j = 0; j < ny; j++)
      for (int i = 0; i <= nx; i++) {
        int k = i + j * (nx + 1);
        uint8_t tl = i > 0 ? tipo[i - 1 + j * nx] : SOLIDO;
        uint8_t tr = i < nx ? tipo[i + j * nx] : SOLIDO;
        if (tl == SOLIDO || tr == SOLIDO) { u[k] = 0; uValida[k] = 1; continue; }
        if (tl == AGUA || tr == AGUA) {
          u[k] = (T)((double)u[k] - ((double)q[i + j * nx] - (double)q[i - 1 + j * nx]));
          uValida[k] = 1;
        } else uValida[k] = 0;
      }
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i < nx; i++) {
        int k = i + j * nx;
        uint8_t tb = j > 0 ? tipo[i + (j - 1) * nx] : SOLIDO;
        uint8_t tt = j < ny ? tipo[i + j * nx] : SOLIDO;
        if (tb == SOLIDO || tt == SOLIDO) { v[k] = 0; vValida[k] = 1; continue; }
        if (tb == AGUA || tt == AGUA) {
          v[k] = (T)((double)v[k] - ((double)q[i + j * nx] - (double)q[i + (j - 1) * nx]));
          vValida[k] = 1;
        } else vValida[k] = 0;
      }
  }
  // extrapolarComponente (fisica-cpu.js:673-690)
  static void extrapolar_componente(std::vector<T>& a, std::vector<T>& tmp, std::vector<uint8_t>& valida, int w, int hh) {
    tmp = a;
    std::vector<int> nov;
    for (int j = 0; j < hh; j++)
      for (int i = 0; i < w; i++) {
        int k = i + j * w;
        if (valida[k]) continue;
        double soma = 0; int m = 0;
        if (i > 0 && valida[k - 1]) { soma += (double)a[k - 1]; m++; }
        if (i < w - 1 && valida[k + 1]) { soma += (double)a[k + 1]; m++; }
        if (j > 0 && valida[k - w]) { soma += (double)a[k - w]; m++; }
        if (j < hh - 1 && valida[k + w]) { soma += (double)a[k + w]; m++; }
        if (m > 0) { tmp[k] = (T)(soma / m); nov.push_back(k); }
      }
    a = tmp;
    for (int k : nov) valida[k] = 1;
  }
  // #extrapolar (fisica-cpu.js:547-552)
  void extrapolar() {
    for (int camada = 0; camada < CAMADAS_EXTRAPOLACAO; camada++) {
      extrapolar_componente(u, uTmp, uValida, nx + 1, ny);
      extrapolar_componente(v, vTmp, vValida, nx, ny + 1);
    }
  }
  // #velocidadeEm (fisica-cpu.js:641-661)
  void velocidade_em(double x, double y, double vmax, double* wx, double* wy) const {
    double inv = 1 / h;
    double vx, vy;
    {
      double fx = x * inv, fy = y * inv - 0.5;
      double i0 = clampi(std::floor(fx), 0, nx - 1), j0 = clampi(std::floor(fy), 0, ny - 2);
      double tx = fx - i0, ty = fy - j0;
      size_t k = (size_t)(i0 + j0 * (nx + 1));
      vx = ((double)u[k] * (1 - tx) + (double)u[k + 1] * tx) * (1 - ty) +
           ((double)u[k + nx + 1] * (1 - tx) + (double)u[k + nx + 2] * tx) * ty;
    }
    {
      double fx = x * inv - 0.5, fy = y * inv;
      double i0 = clampi(std::floor(fx), 0, nx - 2), j0 = clampi(std::floor(fy), 0, ny - 1);
      double tx = fx - i0, ty = fy - j0;
      size_t k = (size_t)(i0 + j0 * nx);
      vy = ((double)v[k] * (1 - tx) + (double)v[k + 1] * tx) * (1 - ty) +
           ((double)v[k + nx] * (1 - tx) + (double)v[k + nx + 1] * tx) * ty;
    }
    double vv = js_hypot(vx, vy);
    if (vv > vmax) { *wx = vx * vmax / vv; *wy = vy * vmax / vv; } else { *wx = vx; *wy = vy; }
  }
  // #paraParticulas (fisica-cpu.js:556-638)
  void para_particulas(double dt) {
    double inv = 1 / h;
    double vmax = (MICROPASSO * MAX_MICROPASSOS * h) / dt;
    double decai = std::exp(-dt / ESPUMA_VIDA);
    double mistura = std::min(1.0, dt / DIFUSAO_CORANTE);
    bool mortas = false;
    double maior = 0;
    for (int k = 0; k < n; k++) {
      T* o = p.data() + (size_t)k * CAMPOS;
      double x = o[0], y = o[1];
      double vx = 0, vy = 0, cux = 0, cuy = 0, cvx = 0, cvy = 0;
      {
        double fx = x * inv, fy = y * inv - 0.5;
        double i0 = clampi(std::floor(fx), 0, nx - 1), j0 = clampi(std::floor(fy), 0, ny - 2);
        double tx = fx - i0, ty = fy - j0;
        for (int b = 0; b < 2; b++)
          for (int a = 0; a < 2; a++) {
            double val = u[(size_t)(i0 + a + (j0 + b) * (nx + 1))];
            double wx = a ? tx : 1 - tx, wy = b ? ty : 1 - ty;
            vx += wx * wy * val;
            cux += (a ? 1 : -1) * wy * inv * val;
            cuy += wx * (b ? 1 : -1) * inv * val;
          }
      }
      {
        double fx = x * inv - 0.5, fy = y * inv;
        double i0 = clampi(std::floor(fx), 0, nx - 2), j0 = clampi(std::floor(fy), 0, ny - 1);
        double tx = fx - i0, ty = fy - j0;
        for (int b = 0; b < 2; b++)
          for (int a = 0; a < 2; a++) {
            double val = v[(size_t)(i0 + a + (j0 + b) * nx)];
            double wx = a ? tx : 1 - tx, wy = b ? ty : 1 - ty;
            vy += wx * wy * val;
            cvx += (a ? 1 : -1) * wy * inv * val;
            cvy += wx * (b ? 1 : -1) * inv * val;
          }
      }
      double vel = js_hypot(vx, vy);
      double ci = clampi(std::floor(x * inv), 0, nx - 1), cj = clampi(std::floor(y * inv), 0, ny - 1);
      size_t cidx = (size_t)(ci + cj * nx);
      double rarefeita = suave(0.95, 0.45, (double)densidade[cidx] / PARTICULAS_POR_CELULA);
      double fonte = suave(ESPUMA_V0, ESPUMA_V1, vel) * rarefeita;
      double espuma = std::min(1.0, (double)o[8] * decai + fonte * ESPUMA_TAXA * dt);
      uint32_t nCel = contagemCelula[cidx];
      if (nCel > 1) {
        size_t cc = cidx * 3;
        for (int qq = 0; qq < 3; qq++)
          o[9 + qq] = (T)((double)o[9 + qq] + ((double)corCelula[cc + qq] / nCel - (double)o[9 + qq]) * mistura);
      }
      if (vel > vmax) { vx *= vmax / vel; vy *= vmax / vel; }
      if (vel > maior) maior = vel;
      int micro = (int)std::min((double)MAX_MICROPASSOS, std::max(1.0, std::ceil((std::min(vel, vmax) * dt) / (MICROPASSO * h))));
      double dm = dt / micro;
      bool morta = false;
      for (int m = 0; m < micro; m++) {
        double wx = vx, wy = vy;
        if (micro > 1) velocidade_em(x, y, vmax, &wx, &wy);
        x += wx * dm;
        y += wy * dm;
        double d = cena->distancia(x, y);
        if (d < RAIO_PARTICULA) {
          if (d < -SDF_MAX + 0.5) { morta = true; break; }
          double nX, nY;
          cena->normal(x, y, &nX, &nY);
          double empurra = (RAIO_PARTICULA - d) * h;
          x += nX * empurra;
          y += nY * empurra;
          double vn = vx * nX + vy * nY;
          if (vn < 0) { vx -= vn * nX; vy -= vn * nY; }
        }
      }
      if (morta) { o[0] = (T)NAN; mortas = true; continue; }
      o[0] = (T)x; o[1] = (T)y; o[2] = (T)vx; o[3] = (T)vy;
      o[4] = (T)cux; o[5] = (T)cuy; o[6] = (T)cvx; o[7] = (T)cvy;
      o[8] = (T)espuma;
    }
    if (mortas) remover([](double x, double, int) { return std::isnan(x); });
    vMax = maior; hasVMax = true;
  }
  double area() const { return ((double)n / PARTICULAS_POR_CELULA) * h * h; }
};

}  // namespace orc

// ================================================================ C ABI for the test harness
using namespace orc;

struct OrSim {
  int mode;  // 0 = JS-exact (binary32 storage), 1 = FP64 storage
  Sim<float>* f = nullptr;
  Sim<double>* d = nullptr;
};

// Field access.  Field ids are shared with the product's debug hooks:
//  0 u, 1 v, 2 uPeso, 3 vPeso, 4 uValida(u8), 5 vValida(u8), 6 densidade, 7 contagemCelula(u32),
//  8 corCelula(3C), 9 tipo(u8), 10 solidoCelula(u8), 11 div, 12 q, 13 fase, 14 normalX, 15 normalY,
//  16 curvatura, 17 qAr, 18 rotulo(u32), 19 pressaoBolha, 20 volumeBolha, 21 rhsBolha, 22 yBolha,
//  23 volumeAlvo, 24 vivo(u8), 25 particles (n*12)
// Real-valued fields are transferred as binary64 (`as_double`) or binary32.
template <class T>
static long field_info(Sim<T>* S, int f, void** ptr, int* kind) {
  // kind: 0 = real (T), 1 = u8, 2 = u32
  switch (f) {
    case 0: *ptr = S->u.data(); *kind = 0; return S->nu;
    case 1: *ptr = S->v.data(); *kind = 0; return S->nv;
    case 2: *ptr = S->uPeso.data(); *kind = 0; return S->nu;
    case 3: *ptr = S->vPeso.data(); *kind = 0; return S->nv;
    case 4: *ptr = S->uValida.data(); *kind = 1; return S->nu;
    case 5: *ptr = S->vValida.data(); *kind = 1; return S->nv;
    case 6: *ptr = S->densidade.data(); *kind = 0; return (long)S->densidade.size();
    case 7: *ptr = S->contagemCelula.data(); *kind = 2; return (long)S->contagemCelula.size();
    case 8: *ptr = S->corCelula.data(); *kind = 0; return (long)S->corCelula.size();
    case 9: *ptr = S->tipo.data(); *kind = 1; return (long)S->tipo.size();
    case 10: *ptr = S->solidoCelula.data(); *kind = 1; return (long)S->solidoCelula.size();
    case 11: *ptr = S->div.data(); *kind = 0; return (long)S->div.size();
    case 12: *ptr = S->q.data(); *kind = 0; return (long)S->q.size();
    case 13: *ptr = S->fase.data(); *kind = 0; return (long)S->fase.size();
    case 14: *ptr = S->normalX.data(); *kind = 0; return (long)S->normalX.size();
    case 15: *ptr = S->normalY.data(); *kind = 0; return (long)S->normalY.size();
    case 16: *ptr = S->curvatura.data(); *kind = 0; return (long)S->curvatura.size();
    case 17: *ptr = S->qAr.data(); *kind = 0; return (long)S->qAr.size();
    case 18: *ptr = S->rastreador.rotulo.data(); *kind = 2; return (long)S->rastreador.rotulo.size();
    case 19: *ptr = S->pressaoBolha.data(); *kind = 0; return MAX_BOLHAS;
    case 20: *ptr = S->volumeBolha.data(); *kind = 0; return MAX_BOLHAS;
    case 21: *ptr = S->rhsBolha.data(); *kind = 0; return MAX_BOLHAS;
    case 22: *ptr = S->yBolha.data(); *kind = 0; return MAX_BOLHAS;
    case 23: *ptr = S->rastreador.volumeAlvo.data(); *kind = 0; return MAX_BOLHAS;
    case 24: *ptr = S->rastreador.vivo.data(); *kind = 1; return MAX_BOLHAS;
    case 25: *ptr = S->p.data(); *kind = 0; return (long)S->n * CAMPOS;
    default: *ptr = nullptr; *kind = -1; return -1;
  }
}

#define DISPATCH(s, expr) ((s)->mode == 0 ? ([&](auto* S) { return expr; })((s)->f) : ([&](auto* S) { return expr; })((s)->d))

extern "C" {

double or_passo_maximo(double h, double alturaM, double vMax, int has_vmax, double dtQuadro) {
  return passo_maximo(h, alturaM, vMax, has_vmax, dtQuadro);
}
double or_lcg_next(uint32_t* s) { return lcg_next(s); }
double or_js_hypot(double a, double b) { return js_hypot(a, b); }
double or_js_round(double x) { return js_round(x); }

Cena* or_scene_new(int nx, int ny, double h) { return new Cena(nx, ny, h); }
void or_scene_free(Cena* c) { delete c; }
void or_scene_stroke(Cena* c, double x0, double y0, double x1, double y1, double raio, int apagar) {
  c->tracar(x0, y0, x1, y1, raio, apagar != 0);
}
void or_scene_clear(Cena* c) { c->zerar(); }
const float* or_scene_sdf(Cena* c) { return c->sdf.data(); }
void or_scene_set_sdf(Cena* c, const float* src) { std::memcpy(c->sdf.data(), src, c->sdf.size() * sizeof(float)); }
double or_scene_distancia(Cena* c, double x, double y) { return c->distancia(x, y); }
void or_scene_normal(Cena* c, double x, double y, double* out) { c->normal(x, y, out, out + 1); }
int or_scene_solida(Cena* c, int i, int j) { return c->celula_solida(i, j) ? 1 : 0; }

int or_seed_rect(double x0, double y0, double x1, double y1, double h, const double* abs3, uint32_t* seed,
                 float* out, int max_out) {
  return semear_retangulo(x0, y0, x1, y1, h, abs3, seed, out, max_out);
}

OrSim* or_sim_new(Cena* c, int mode, double gx, double gy, int iteracoes, double tau, double tensao, int ar) {
  OrSim* s = new OrSim();
  s->mode = mode;
  if (mode == 0) s->f = new Sim<float>(c, gx, gy, iteracoes, tau, tensao, ar != 0);
  else s->d = new Sim<double>(c, gx, gy, iteracoes, tau, tensao, ar != 0);
  return s;
}
void or_sim_free(OrSim* s) { delete s->f; delete s->d; delete s; }
void or_sim_set_tol(OrSim* s, double tol, int maxit) {
  DISPATCH(s, (S->tol = tol, S->iteracoes = maxit, 0));
}
void or_sim_set_gravity(OrSim* s, double gx, double gy) {
  DISPATCH(s, (S->gravidade[0] = gx, S->gravidade[1] = gy, 0));
}
int or_sim_add(OrSim* s, const float* lista, int total, int check) { return DISPATCH(s, S->adicionar(lista, total, check != 0)); }
int or_sim_emit_faucet(OrSim* s, double x, double y, double pvx, double pvy, double raio, const double* abs3,
                       double dt, uint32_t* seed, int* emitted) {
  return DISPATCH(s, S->emitir_torneira(x, y, pvx, pvy, raio, abs3, dt, seed, emitted));
}
void or_sim_step(OrSim* s, double dt, int sub) { DISPATCH(s, (S->avancar(dt, sub), 0)); }
void or_sim_substep(OrSim* s, double dt) { DISPATCH(s, (S->subpasso(dt), 0)); }
void or_sim_phase(OrSim* s, int ph, double dt) { DISPATCH(s, (S->phase(ph, dt), 0)); }
void or_sim_update_walls(OrSim* s) { DISPATCH(s, (S->atualizar_paredes(), 0)); }
void or_sim_reset_air(OrSim* s) { DISPATCH(s, (S->reiniciar_ar(), 0)); }
void or_sim_clear_water(OrSim* s) {
  DISPATCH(s, (S->n = 0, std::fill(S->contagemCelula.begin(), S->contagemCelula.end(), 0u), S->reiniciar_ar(), 0));
}
void or_sim_tool(OrSim* s, int tipo, double x, double y, double raio, double vx, double vy, const double* abs3) {
  DISPATCH(s, (S->aplicar_ferramenta(tipo, x, y, raio, vx, vy, abs3), 0));
}
int or_sim_n(OrSim* s) { return DISPATCH(s, S->n); }
void or_sim_set_n(OrSim* s, int n) { DISPATCH(s, (S->n = std::min(n, S->n), 0)); }
double or_sim_vmax(OrSim* s, int* has) { return DISPATCH(s, (*has = S->hasVMax ? 1 : 0, S->vMax)); }
double or_sim_area(OrSim* s) { return DISPATCH(s, S->area()); }
int or_sim_ncomp(OrSim* s) { return DISPATCH(s, S->rastreador.ncomp_last); }
int or_sim_solver_stats(OrSim* s, double* norms, int max, double* normb, double* true_res, int* iters) {
  return DISPATCH(s, ([&]() {
    auto& sv = S->solver;
    int m = std::min(max, (int)sv.normas.size());
    for (int i = 0; i < m; i++) norms[i] = sv.normas[i];
    *normb = sv.norm_b; *true_res = sv.true_resid; *iters = sv.iters_done;
    return (int)sv.normas.size();
  })());
}

long or_sim_field_len(OrSim* s, int f) {
  void* p; int kind;
  return DISPATCH(s, field_info(S, f, &p, &kind));
}
// Copies field f into dst (binary64 when as_double, else binary32 for real fields; raw ints otherwise).
long or_sim_read(OrSim* s, int f, void* dst, int as_double) {
  return DISPATCH(s, ([&]() -> long {
    void* p; int kind;
    long n = field_info(S, f, &p, &kind);
    if (n < 0) return -1;
    using TT = std::remove_pointer_t<decltype(S->u.data())>;
    if (kind == 0) {
      TT* src = (TT*)p;
      if (as_double) for (long i = 0; i < n; i++) ((double*)dst)[i] = (double)src[i];
      else for (long i = 0; i < n; i++) ((float*)dst)[i] = (float)src[i];
    } else if (kind == 1) std::memcpy(dst, p, n);
    else std::memcpy(dst, p, n * 4);
    return n;
  })());
}
// Overwrites field f from src (same encoding as or_sim_read).  For particles (25) `count`
// gives the new particle count; for other fields it is ignored.
long or_sim_write(OrSim* s, int f, const void* src, int as_double, int count) {
  return DISPATCH(s, ([&]() -> long {
    if (f == 25) { S->garantir(count); S->n = count; }
    void* p; int kind;
    long n = field_info(S, f, &p, &kind);
    if (n < 0) return -1;
    using TT = std::remove_pointer_t<decltype(S->u.data())>;
    if (kind == 0) {
      TT* d = (TT*)p;
      if (as_double) for (long i = 0; i < n; i++) d[i] = (TT)((const double*)src)[i];
      else for (long i = 0; i < n; i++) d[i] = (TT)((const float*)src)[i];
    } else if (kind == 1) std::memcpy(p, src, n);
    else std::memcpy(p, src, n * 4);
    if (f == 9) S->tiposProntos = true;
    return n;
  })());
}
void or_sim_set_tracker_first(OrSim* s, int primeira) { DISPATCH(s, (S->rastreador.primeira = primeira != 0, 0)); }
int or_sim_tracker_first(OrSim* s) { return DISPATCH(s, S->rastreador.primeira ? 1 : 0); }
// test hooks: one V-cycle / the operator on the sim's current types (labels for the V-cycle)
void or_sim_mg_vcycle(OrSim* s, const double* in, double* out) {
  DISPATCH(s, (mg_vcycle_hook(S->solver, S->tipo.data(), S->arAtivo ? S->rastreador.rotulo.data() : nullptr, in, out), 0));
}
// test hook: treat the stored types as current (atualizarAr then skips the P2G it would run)
void or_sim_set_types_ready(OrSim* s, int ready) { DISPATCH(s, (S->tiposProntos = ready != 0, 0)); }
void or_sim_op_apply(OrSim* s, const double* in, double* out) {
  DISPATCH(s, (op_apply_hook(S->solver, S->tipo.data(), in, out), 0));
}

}  // extern "C"
