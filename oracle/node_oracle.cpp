// node_oracle.cpp -- Node-API binding of the oracle (tanque_oracle.cpp, the single-threaded C++
// restatement of the reference CPU engine) for timing it from Node, the reference's own host,
// beside the device path.  Test infrastructure: only the CPU-baseline leg of bench.py runs it
// (through oracle/bench_oracle.js); nothing shipped links it.
//
// Surface (handles are external values):
//   cena(nx, ny, h) -> scene            tracar(scene, x0, y0, x1, y1, raio, apagar)
//   simulacao(scene, tol, iteracoes) -> sim   (JS-exact mode, gravity (0, -9.81), tau 0.1,
//                                              surface tension 0.0728, air on: SimulacaoCPU's defaults)
//   adicionar(sim, Float32Array, checar) -> accepted
//   avancar(sim, dt, subpassos)          vmax(sim) -> top speed or null
//   passoMaximo(h, altura, vmax or null, dt) -> the reference's step limit (constantes.js:167-175)
#define NAPI_VERSION 4
#include <node_api.h>

#include <cstdint>
#include <cstring>

extern "C" {
double or_passo_maximo(double h, double alturaM, double vMax, int has_vmax, double dtQuadro);
void* or_scene_new(int nx, int ny, double h);
void or_scene_free(void* c);
void or_scene_stroke(void* c, double x0, double y0, double x1, double y1, double raio, int apagar);
void* or_sim_new(void* c, int mode, double gx, double gy, int iteracoes, double tau, double tensao, int ar);
void or_sim_free(void* s);
void or_sim_set_tol(void* s, double tol, int maxit);
int or_sim_add(void* s, const float* lista, int total, int check);
void or_sim_step(void* s, double dt, int sub);
double or_sim_vmax(void* s, int* has);
}

namespace {

bool args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
  if (argc < want) {
    napi_throw_type_error(env, nullptr, "too few arguments");
    return false;
  }
  return true;
}

bool number(napi_env env, napi_value v, double* out) {
  if (napi_get_value_double(env, v, out) == napi_ok) return true;
  napi_throw_type_error(env, nullptr, "number expected");
  return false;
}

void* handle(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) napi_throw_type_error(env, nullptr, "handle expected");
  return p;
}

napi_value make_double(napi_env env, double d) {
  napi_value r;
  napi_create_double(env, d, &r);
  return r;
}

void free_scene(napi_env, void* p, void*) { or_scene_free(p); }
void free_sim(napi_env, void* p, void*) { or_sim_free(p); }

napi_value Cena(napi_env env, napi_callback_info info) {
  napi_value a[3];
  double nx, ny, h;
  if (!args(env, info, 3, a) || !number(env, a[0], &nx) || !number(env, a[1], &ny) || !number(env, a[2], &h))
    return nullptr;
  if (nx < 1 || ny < 1 || !(h > 0)) {
    napi_throw_range_error(env, nullptr, "cena: bad size");
    return nullptr;
  }
  napi_value r;
  napi_create_external(env, or_scene_new((int)nx, (int)ny, h), free_scene, nullptr, &r);
  return r;
}

napi_value Tracar(napi_env env, napi_callback_info info) {
  napi_value a[7];
  double v[5];
  bool apagar = false;
  if (!args(env, info, 7, a)) return nullptr;
  void* c = handle(env, a[0]);
  if (!c) return nullptr;
  for (int k = 0; k < 5; k++)
    if (!number(env, a[1 + k], &v[k])) return nullptr;
  napi_get_value_bool(env, a[6], &apagar);
  or_scene_stroke(c, v[0], v[1], v[2], v[3], v[4], apagar ? 1 : 0);
  return nullptr;
}

napi_value Simulacao(napi_env env, napi_callback_info info) {
  napi_value a[3];
  double tol, it;
  if (!args(env, info, 3, a)) return nullptr;
  void* c = handle(env, a[0]);
  if (!c || !number(env, a[1], &tol) || !number(env, a[2], &it)) return nullptr;
  void* s = or_sim_new(c, 0, 0.0, -9.81, (int)it, 0.1, 0.0728, 1);
  if (tol > 0) or_sim_set_tol(s, tol, (int)it);
  napi_value r;
  napi_create_external(env, s, free_sim, nullptr, &r);
  return r;
}

napi_value Adicionar(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!args(env, info, 3, a)) return nullptr;
  void* s = handle(env, a[0]);
  if (!s) return nullptr;
  napi_typedarray_type t;
  size_t len = 0;
  void* data = nullptr;
  if (napi_get_typedarray_info(env, a[1], &t, &len, &data, nullptr, nullptr) != napi_ok || t != napi_float32_array ||
      len % 12) {
    napi_throw_type_error(env, nullptr, "adicionar: Float32Array of 12 floats per particle expected");
    return nullptr;
  }
  bool checar = false;
  napi_get_value_bool(env, a[2], &checar);
  return make_double(env, or_sim_add(s, static_cast<const float*>(data), (int)(len / 12), checar ? 1 : 0));
}

napi_value Avancar(napi_env env, napi_callback_info info) {
  napi_value a[3];
  double dt, sub;
  if (!args(env, info, 3, a)) return nullptr;
  void* s = handle(env, a[0]);
  if (!s || !number(env, a[1], &dt) || !number(env, a[2], &sub)) return nullptr;
  or_sim_step(s, dt, (int)sub);
  return nullptr;
}

napi_value Vmax(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  void* s = handle(env, a[0]);
  if (!s) return nullptr;
  int has = 0;
  const double v = or_sim_vmax(s, &has);
  if (has) return make_double(env, v);
  napi_value n;
  napi_get_null(env, &n);
  return n;
}

napi_value PassoMaximo(napi_env env, napi_callback_info info) {
  napi_value a[4];
  double h, altura, dt, vmax = 0;
  if (!args(env, info, 4, a) || !number(env, a[0], &h) || !number(env, a[1], &altura) || !number(env, a[3], &dt))
    return nullptr;
  napi_valuetype vt;
  napi_typeof(env, a[2], &vt);
  const int has = vt == napi_number;
  if (has) napi_get_value_double(env, a[2], &vmax);
  return make_double(env, or_passo_maximo(h, altura, vmax, has, dt));
}

napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor fns[] = {
      {"cena", nullptr, Cena, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"tracar", nullptr, Tracar, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"simulacao", nullptr, Simulacao, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"adicionar", nullptr, Adicionar, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"avancar", nullptr, Avancar, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"vmax", nullptr, Vmax, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"passoMaximo", nullptr, PassoMaximo, nullptr, nullptr, nullptr, napi_default, nullptr},
  };
  napi_define_properties(env, exports, sizeof fns / sizeof fns[0], fns);
  return exports;
}

}  // namespace

NAPI_MODULE(oracle_node, init)
