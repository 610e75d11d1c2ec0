// napi_addon.cpp -- Node-API binding of libtanque (include/tanque.h) for the reference's
// JavaScript host.  It exposes one class, SimulacaoHIP, whose methods are the step/state surface
// of SimulacaoCPU (js/fisica-cpu.js:28-194) that MotorCPU drives (js/motor.js:79-120), plus the
// scene and seeding helpers (js/cena.js, fisica-cpu.js:705-727).  Every failure of the library
// becomes a thrown JS Error carrying tq_last_error(); typed arrays are copied in and out.
#define NAPI_VERSION 4
#include <node_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/tanque.h"

namespace {

struct Native {
  tq_sim* s = nullptr;
  int nx = 0, ny = 0;
};

bool throw_tq(napi_env env, int rc, const char* what) {
  if (rc == TQ_OK) return true;
  std::string msg = std::string(what) + ": " + tq_last_error();
  napi_throw_error(env, std::to_string(rc).c_str(), msg.c_str());
  return false;
}

void throw_arg(napi_env env, const char* msg) { napi_throw_type_error(env, nullptr, msg); }

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv, napi_value* self,
              size_t* got = nullptr) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, self, nullptr) != napi_ok) return false;
  for (size_t k = argc; k < want; k++) napi_get_undefined(env, &argv[k]);
  if (got) *got = argc;
  return true;
}

bool is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_undefined || t == napi_null;
}

bool num(napi_env env, napi_value v, double* out) {
  if (napi_get_value_double(env, v, out) != napi_ok) { throw_arg(env, "number expected"); return false; }
  return true;
}

double num_or(napi_env env, napi_value obj, const char* key, double dflt) {
  bool has = false;
  if (is_nullish(env, obj) || napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  double d;
  if (napi_get_value_double(env, v, &d) != napi_ok) {
    bool b;
    if (napi_get_value_bool(env, v, &b) == napi_ok) return b ? 1 : 0;
    return dflt;
  }
  return d;
}

// mgDense: true/1 the dense coarse matrix, false/0 the one-workgroup V-cycle, -1 (or absent) auto
int dense_opt(napi_env env, napi_value o, int dflt) {
  const double v = num_or(env, o, "mgDense", dflt);
  return v < 0 ? -1 : (v != 0 ? 1 : 0);
}

// a Float32Array argument: data and element count
bool f32_array(napi_env env, napi_value v, float** data, size_t* n) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) { throw_arg(env, "Float32Array expected"); return false; }
  napi_typedarray_type t;
  void* p = nullptr;
  napi_value ab;
  size_t off = 0;
  napi_get_typedarray_info(env, v, &t, n, &p, &ab, &off);
  if (t != napi_float32_array) { throw_arg(env, "Float32Array expected"); return false; }
  *data = static_cast<float*>(p);
  return true;
}

napi_value make_typed(napi_env env, napi_typedarray_type t, size_t elems, size_t elem_bytes, void** data) {
  napi_value ab, arr;
  napi_create_arraybuffer(env, elems * elem_bytes, data, &ab);
  napi_create_typedarray(env, t, elems, ab, 0, &arr);
  return arr;
}

napi_value undefined(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

napi_value number(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}

Native* unwrap(napi_env env, napi_value self) {
  Native* n = nullptr;
  napi_unwrap(env, self, reinterpret_cast<void**>(&n));
  if (!n || !n->s) {
    napi_throw_error(env, nullptr, "SimulacaoHIP: destroyed");
    return nullptr;
  }
  return n;
}

void finalize(napi_env, void* data, void*) {
  Native* n = static_cast<Native*>(data);
  if (n->s) tq_destroy(n->s);
  delete n;
}

// new SimulacaoHIP(nx, ny, h, opcoes) -- opcoes as SimulacaoCPU's (fisica-cpu.js:29-37) plus
// {precision, tol, useGraph, maxParticulas}
napi_value construct(napi_env env, napi_callback_info info) {
  napi_value argv[4], self;
  if (!get_args(env, info, 4, argv, &self)) return nullptr;
  double nx, ny, h;
  if (!num(env, argv[0], &nx) || !num(env, argv[1], &ny) || !num(env, argv[2], &h)) return nullptr;
  tq_desc d;
  tq_desc_default(&d, (int)nx, (int)ny, h);
  napi_value o = argv[3];
  if (!is_nullish(env, o)) {
    bool has = false;
    napi_has_named_property(env, o, "gravidade", &has);
    if (has) {
      napi_value g, gx, gy;
      napi_get_named_property(env, o, "gravidade", &g);
      napi_get_element(env, g, 0, &gx);
      napi_get_element(env, g, 1, &gy);
      napi_get_value_double(env, gx, &d.gx);
      napi_get_value_double(env, gy, &d.gy);
    }
    d.iteracoes = (int)num_or(env, o, "iteracoes", d.iteracoes);
    d.tau_densidade = num_or(env, o, "tauDensidade", d.tau_densidade);
    d.tensao = num_or(env, o, "tensao", d.tensao);
    d.ar = num_or(env, o, "ar", d.ar) != 0;
    d.precision = (int)num_or(env, o, "precision", d.precision);
    d.tol = num_or(env, o, "tol", d.tol);
    d.use_graph = num_or(env, o, "useGraph", d.use_graph) != 0;
    d.mg_dense = dense_opt(env, o, d.mg_dense);
    d.max_particles = (int)num_or(env, o, "maxParticulas", d.max_particles);
  }
  Native* n = new Native();
  n->nx = d.nx;
  n->ny = d.ny;
  if (!throw_tq(env, tq_create(&d, &n->s), "tq_create")) { delete n; return nullptr; }
  napi_wrap(env, self, n, finalize, nullptr, nullptr);
  return self;
}

// definirSdf(sdf Float32Array (2nx*2ny), row0, row1): atualizarParedes (fisica-cpu.js:80-85)
napi_value m_set_sdf(napi_env env, napi_callback_info info) {
  napi_value argv[3], self;
  if (!get_args(env, info, 3, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  float* sdf;
  size_t len;
  if (!f32_array(env, argv[0], &sdf, &len)) return nullptr;
  if (len != (size_t)4 * n->nx * n->ny) { throw_arg(env, "sdf must hold (2nx)*(2ny) values"); return nullptr; }
  double r0 = 0, r1 = 2 * n->ny - 1;
  if (!is_nullish(env, argv[1]) && !num(env, argv[1], &r0)) return nullptr;
  if (!is_nullish(env, argv[2]) && !num(env, argv[2], &r1)) return nullptr;
  throw_tq(env, tq_set_sdf(n->s, sdf, (int)r0, (int)r1), "tq_set_sdf");
  return undefined(env);
}

// adicionar(lista Float32Array, checarCheia) -> accepted (fisica-cpu.js:92-111)
napi_value m_add(napi_env env, napi_callback_info info) {
  napi_value argv[2], self;
  if (!get_args(env, info, 2, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  float* lista;
  size_t len;
  if (!f32_array(env, argv[0], &lista, &len)) return nullptr;
  bool check = true;
  if (!is_nullish(env, argv[1])) napi_get_value_bool(env, argv[1], &check);
  int acc = 0;
  if (!throw_tq(env, tq_add_particles(n->s, lista, (int)(len / TQ_CAMPOS), check ? 1 : 0, &acc), "tq_add_particles"))
    return nullptr;
  return number(env, acc);
}

// aplicarFerramenta({tipo, x, y, raio, vx, vy, absorcao}) (fisica-cpu.js:129-152)
napi_value m_tool(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 1, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  napi_value o = argv[0];
  if (is_nullish(env, o)) return undefined(env);
  tq_tool t{};
  napi_value tipo;
  napi_get_named_property(env, o, "tipo", &tipo);
  char name[32] = {0};
  size_t nl = 0;
  napi_get_value_string_utf8(env, tipo, name, sizeof name, &nl);
  const std::string k(name);
  t.kind = k == "esponja" ? 1 : k == "mao" ? 2 : k == "corante" ? 3 : 0;
  if (t.kind == 0) return undefined(env);
  t.x = num_or(env, o, "x", 0);
  t.y = num_or(env, o, "y", 0);
  t.raio = num_or(env, o, "raio", 0);
  t.vx = num_or(env, o, "vx", 0);
  t.vy = num_or(env, o, "vy", 0);
  bool has = false;
  napi_has_named_property(env, o, "absorcao", &has);
  if (has) {
    napi_value a;
    napi_get_named_property(env, o, "absorcao", &a);
    for (uint32_t c = 0; c < 3; c++) {
      napi_value e;
      if (napi_get_element(env, a, c, &e) == napi_ok) napi_get_value_double(env, e, &t.absorcao[c]);
    }
  }
  throw_tq(env, tq_apply_tool(n->s, &t), "tq_apply_tool");
  return undefined(env);
}

// emitirTorneira (principal.js:301-321) with the seeded LCG: ({x, y, vx, vy, raio, absorcao, dt,
// semente}) -> {emitidas, aceitas, semente}; vx, vy are the pointer's velocity
napi_value m_faucet(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 1, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  napi_value o = argv[0];
  tq_faucet f{};
  f.x = num_or(env, o, "x", 0);
  f.y = num_or(env, o, "y", 0);
  f.pvx = num_or(env, o, "vx", 0);
  f.pvy = num_or(env, o, "vy", 0);
  f.raio = num_or(env, o, "raio", 0);
  bool has = false;
  napi_has_named_property(env, o, "absorcao", &has);
  if (has) {
    napi_value a;
    napi_get_named_property(env, o, "absorcao", &a);
    for (uint32_t c = 0; c < 3; c++) {
      napi_value e;
      if (napi_get_element(env, a, c, &e) == napi_ok) napi_get_value_double(env, e, &f.absorcao[c]);
    }
  }
  const double dt = num_or(env, o, "dt", 0);
  uint32_t seed = (uint32_t)num_or(env, o, "semente", 1);
  int emitted = 0, accepted = 0;
  if (!throw_tq(env, tq_emit_faucet(n->s, &f, dt, &seed, &emitted, &accepted), "tq_emit_faucet")) return nullptr;
  napi_value r, v;
  napi_create_object(env, &r);
  napi_create_int32(env, emitted, &v);
  napi_set_named_property(env, r, "emitidas", v);
  napi_create_int32(env, accepted, &v);
  napi_set_named_property(env, r, "aceitas", v);
  napi_create_uint32(env, seed, &v);
  napi_set_named_property(env, r, "semente", v);
  return r;
}

// a Float64Array argument: data and element count
bool f64_array(napi_env env, napi_value v, double** data, size_t* n) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) { throw_arg(env, "Float64Array expected"); return false; }
  napi_typedarray_type t;
  void* p = nullptr;
  napi_value ab;
  size_t off = 0;
  napi_get_typedarray_info(env, v, &t, n, &p, &ab, &off);
  if (t != napi_float64_array) { throw_arg(env, "Float64Array expected"); return false; }
  *data = static_cast<double*>(p);
  return true;
}

// definirTracos(Float64Array n*6: x0 y0 x1 y1 raio apagar): the walls built on the device
// (Cena.refazer + atualizarParedes, tq_set_strokes)
napi_value m_set_strokes(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 1, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  double* d = nullptr;
  size_t len = 0;
  if (!f64_array(env, argv[0], &d, &len)) return nullptr;
  if (len % 6) { throw_arg(env, "6 numbers per stroke expected"); return nullptr; }
  if (!throw_tq(env, tq_set_strokes(n->s, d, (int)(len / 6)), "tq_set_strokes")) return nullptr;
  return undefined(env);
}

// renderizarAgua() -> {largura, altura, tinta, espuma}: the water layer of the 2D renderer
// (render-2d.js:110-162) as two Uint8ClampedArray RGBA images of 3 nx x 3 ny pixels (tq_render_water)
napi_value m_render_water(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 0, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  const int w = 3 * n->nx, h = 3 * n->ny;
  const size_t bytes = (size_t)w * h * 4;
  void *tp = nullptr, *ep = nullptr;
  napi_value tint = make_typed(env, napi_uint8_clamped_array, bytes, 1, &tp);
  napi_value foam = make_typed(env, napi_uint8_clamped_array, bytes, 1, &ep);
  if (!tint || !foam) return nullptr;
  if (!throw_tq(env, tq_render_water(n->s, w, h, static_cast<uint8_t*>(tp), static_cast<uint8_t*>(ep)),
                "tq_render_water"))
    return nullptr;
  napi_value out, vw, vh;
  napi_create_object(env, &out);
  napi_create_int32(env, w, &vw);
  napi_create_int32(env, h, &vh);
  napi_set_named_property(env, out, "largura", vw);
  napi_set_named_property(env, out, "altura", vh);
  napi_set_named_property(env, out, "tinta", tint);
  napi_set_named_property(env, out, "espuma", foam);
  return out;
}

// tracosVjp(Float64Array strokes, Float32Array gsdf) -> Float64Array n*5 (tq_strokes_vjp)
napi_value m_strokes_vjp(napi_env env, napi_callback_info info) {
  napi_value argv[2], self;
  if (!get_args(env, info, 2, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  double* d = nullptr;
  size_t len = 0, glen = 0;
  float* g = nullptr;
  if (!f64_array(env, argv[0], &d, &len) || !f32_array(env, argv[1], &g, &glen)) return nullptr;
  if (len % 6 || glen != (size_t)4 * n->nx * n->ny) { throw_arg(env, "strokes n*6 and gsdf 4*nx*ny expected"); return nullptr; }
  void* out = nullptr;
  napi_value arr = make_typed(env, napi_float64_array, (len / 6) * 5, 8, &out);
  if (!throw_tq(env, tq_strokes_vjp(n->s, d, (int)(len / 6), g, static_cast<double*>(out)), "tq_strokes_vjp"))
    return nullptr;
  return arr;
}

// gravar(checkpointACada): start recording frames for retropropagar (tq_tape_begin); fimGravacao()
napi_value m_tape_begin(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 1, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  double every = 10;
  if (!is_nullish(env, argv[0]) && !num(env, argv[0], &every)) return nullptr;
  if (!throw_tq(env, tq_tape_begin(n->s, (int)every), "tq_tape_begin")) return nullptr;
  return undefined(env);
}

// fracaoAgua(i0, j0, i1, j1, comGradiente) -> {fracao, gradiente: Float32Array n*8 | undefined}
napi_value m_fraction(napi_env env, napi_callback_info info) {
  napi_value argv[5], self;
  if (!get_args(env, info, 5, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  double b[4];
  for (int k = 0; k < 4; k++)
    if (!num(env, argv[k], &b[k])) return nullptr;
  bool want = false;
  if (!is_nullish(env, argv[4])) napi_get_value_bool(env, argv[4], &want);
  double frac = 0;
  napi_value r;
  napi_create_object(env, &r);
  if (want) {
    int c = 0;
    if (!throw_tq(env, tq_count(n->s, &c), "tq_count")) return nullptr;
    void* data = nullptr;
    napi_value g = make_typed(env, napi_float32_array, (size_t)c * 8, 4, &data);
    if (!throw_tq(env, tq_water_fraction(n->s, (int)b[0], (int)b[1], (int)b[2], (int)b[3], &frac, static_cast<float*>(data), c),
                  "tq_water_fraction"))
      return nullptr;
    napi_set_named_property(env, r, "gradiente", g);
  } else if (!throw_tq(env, tq_water_fraction(n->s, (int)b[0], (int)b[1], (int)b[2], (int)b[3], &frac, nullptr, 0),
                       "tq_water_fraction")) {
    return nullptr;
  }
  napi_set_named_property(env, r, "fracao", number(env, frac));
  return r;
}

// retropropagar(gradEstado: Float32Array n*8, tol) -> {torneira: Float64Array 2 per faucet call,
// sdf: Float32Array, estadoInicial: Float32Array n0*8} (tq_backward)
napi_value m_backward(napi_env env, napi_callback_info info) {
  napi_value argv[2], self;
  if (!get_args(env, info, 2, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  float* g = nullptr;
  size_t len = 0;
  if (!f32_array(env, argv[0], &g, &len)) return nullptr;
  double tol = 1e-6;
  if (!is_nullish(env, argv[1]) && !num(env, argv[1], &tol)) return nullptr;
  int frames = 0, calls = 0, cps = 0, n0 = 0;
  if (!throw_tq(env, tq_tape_info(n->s, &frames, &calls, &cps, &n0), "tq_tape_info")) return nullptr;
  void *pf = nullptr, *ps = nullptr, *p0 = nullptr;
  napi_value tf = make_typed(env, napi_float64_array, (size_t)std::max(calls, 0) * 2, 8, &pf);
  napi_value ts = make_typed(env, napi_float32_array, (size_t)4 * n->nx * n->ny, 4, &ps);
  napi_value t0 = make_typed(env, napi_float32_array, (size_t)n0 * 8, 4, &p0);
  if (!throw_tq(env, tq_backward(n->s, g, (int)(len / 8), tol, static_cast<double*>(pf), calls, static_cast<float*>(ps),
                                 static_cast<float*>(p0), n0),
                "tq_backward"))
    return nullptr;
  napi_value r;
  napi_create_object(env, &r);
  napi_set_named_property(env, r, "torneira", tf);
  napi_set_named_property(env, r, "sdf", ts);
  napi_set_named_property(env, r, "estadoInicial", t0);
  return r;
}

// avancar(dt, subpassos) (fisica-cpu.js:155-159)
napi_value m_step(napi_env env, napi_callback_info info) {
  napi_value argv[2], self;
  if (!get_args(env, info, 2, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  double dt, sub = 1;
  if (!num(env, argv[0], &dt)) return nullptr;
  if (!is_nullish(env, argv[1]) && !num(env, argv[1], &sub)) return nullptr;
  throw_tq(env, tq_step(n->s, dt, (int)sub), "tq_step");
  return undefined(env);
}

// avancarAssincrono(dt, subpassos): the frame queued without waiting, its status readback behind it
napi_value m_step_async(napi_env env, napi_callback_info info) {
  napi_value argv[2], self;
  if (!get_args(env, info, 2, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  double dt, sub = 1;
  if (!num(env, argv[0], &dt)) return nullptr;
  if (!is_nullish(env, argv[1]) && !num(env, argv[1], &sub)) return nullptr;
  throw_tq(env, tq_step_async(n->s, dt, (int)sub), "tq_step_async");
  return undefined(env);
}

// pedirEstado(): a status readback queued behind the queued work
napi_value m_status_request(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 0, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  throw_tq(env, tq_status_request(n->s), "tq_status_request");
  return undefined(env);
}

// estado() -> null while the device has not reached the last request, else {n, vMax, msGpu}
napi_value m_status_poll(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 0, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  int ready = 0, c = 0, has = 0;
  double v = 0, ms = -1;
  if (!throw_tq(env, tq_status_poll(n->s, &ready, &c, &v, &has, &ms), "tq_status_poll")) return nullptr;
  napi_value out;
  if (!ready) {
    napi_get_null(env, &out);
    return out;
  }
  napi_create_object(env, &out);
  napi_set_named_property(env, out, "n", number(env, c));
  napi_value vm;
  if (has) vm = number(env, v);
  else napi_get_null(env, &vm);
  napi_set_named_property(env, out, "vMax", vm);
  napi_value mv;
  if (ms >= 0) mv = number(env, ms);
  else napi_get_null(env, &mv);
  napi_set_named_property(env, out, "msGpu", mv);
  return out;
}

// subpasso(dt) (fisica-cpu.js:176-189)
napi_value m_substep(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 1, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  double dt;
  if (!num(env, argv[0], &dt)) return nullptr;
  throw_tq(env, tq_substep(n->s, dt), "tq_substep");
  return undefined(env);
}

template <int (*F)(tq_sim*)>
napi_value m_simple(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 0, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  throw_tq(env, F(n->s), "tanque");
  return undefined(env);
}

// contagem() -> n
napi_value m_count(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 0, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  int c = 0;
  if (!throw_tq(env, tq_count(n->s, &c), "tq_count")) return nullptr;
  return number(env, c);
}

// truncar(m): sim.n = m (fisica-cpu.js: n is a plain field the callers lower)
napi_value m_truncate(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 1, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  double m;
  if (!num(env, argv[0], &m)) return nullptr;
  throw_tq(env, tq_truncate(n->s, (int)m), "tq_truncate");
  return undefined(env);
}

// particulas() -> Float32Array copy of the first n records (motor.js:117)
napi_value m_particles(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 0, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  int c = 0;
  if (!throw_tq(env, tq_count(n->s, &c), "tq_count")) return nullptr;
  void* data = nullptr;
  napi_value arr = make_typed(env, napi_float32_array, (size_t)c * TQ_CAMPOS, 4, &data);
  int got = 0;
  if (!throw_tq(env, tq_read_particles(n->s, static_cast<float*>(data), c, &got), "tq_read_particles")) return nullptr;
  return arr;
}

// vMax() -> number | null (fisica-cpu.js:637)
napi_value m_vmax(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 0, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  double v = 0;
  int has = 0;
  if (!throw_tq(env, tq_vmax(n->s, &v, &has), "tq_vmax")) return nullptr;
  if (!has) {
    napi_value nul;
    napi_get_null(env, &nul);
    return nul;
  }
  return number(env, v);
}

// area() (fisica-cpu.js:192-194)
napi_value m_area(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 0, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  double a = 0;
  if (!throw_tq(env, tq_area(n->s, &a), "tq_area")) return nullptr;
  return number(env, a);
}

// lerCampo(id) -> typed array of the state the reference exposes (ids as tq_read_field)
napi_value m_field(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 1, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  double fd;
  if (!num(env, argv[0], &fd)) return nullptr;
  const int f = (int)fd;
  const size_t nx = n->nx, ny = n->ny, C = nx * ny;
  size_t elems = C;
  napi_typedarray_type t = napi_float32_array;
  size_t eb = 4;
  if (f == 0 || f == 2 || f == 4) elems = (nx + 1) * ny;
  if (f == 1 || f == 3 || f == 5) elems = nx * (ny + 1);
  if (f == 8) elems = 3 * C;
  if (f >= 19) elems = TQ_MAX_BOLHAS;
  if (f == 4 || f == 5 || f == 9 || f == 10 || f == 24) { t = napi_uint8_array; eb = 1; }
  if (f == 7 || f == 18) t = napi_uint32_array;
  void* data = nullptr;
  napi_value arr = make_typed(env, t, elems, eb, &data);
  if (!throw_tq(env, tq_read_field(n->s, f, data, elems * eb), "tq_read_field")) return nullptr;
  return arr;
}

// opcoes({gravidade, iteracoes, tauDensidade, tensao, ar, precision, tol, useGraph})
napi_value m_options(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 1, argv, &self)) return nullptr;
  Native* n = unwrap(env, self);
  if (!n) return nullptr;
  tq_desc d;
  tq_desc_default(&d, n->nx, n->ny, 1.0);
  napi_value o = argv[0];
  bool has = false;
  napi_has_named_property(env, o, "gravidade", &has);
  if (has) {
    napi_value g, gx, gy;
    napi_get_named_property(env, o, "gravidade", &g);
    napi_get_element(env, g, 0, &gx);
    napi_get_element(env, g, 1, &gy);
    napi_get_value_double(env, gx, &d.gx);
    napi_get_value_double(env, gy, &d.gy);
  }
  d.iteracoes = (int)num_or(env, o, "iteracoes", d.iteracoes);
  d.tau_densidade = num_or(env, o, "tauDensidade", d.tau_densidade);
  d.tensao = num_or(env, o, "tensao", d.tensao);
  d.ar = num_or(env, o, "ar", d.ar) != 0;
  d.precision = (int)num_or(env, o, "precision", d.precision);
  d.tol = num_or(env, o, "tol", d.tol);
  d.use_graph = num_or(env, o, "useGraph", d.use_graph) != 0;
  d.mg_dense = dense_opt(env, o, d.mg_dense);
  throw_tq(env, tq_set_options(n->s, &d), "tq_set_options");
  return undefined(env);
}

napi_value m_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1], self;
  if (!get_args(env, info, 0, argv, &self)) return nullptr;
  Native* n = nullptr;
  napi_unwrap(env, self, reinterpret_cast<void**>(&n));
  if (n && n->s) {
    tq_destroy(n->s);
    n->s = nullptr;
  }
  return undefined(env);
}

// ---------------------------------------------------------------- scene and seeding helpers
// cenaMoldura(nx, ny) -> Float32Array SDF of the empty tank (cena.js:139-144)
napi_value f_frame(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv, nullptr)) return nullptr;
  double nx, ny;
  if (!num(env, argv[0], &nx) || !num(env, argv[1], &ny)) return nullptr;
  void* data = nullptr;
  napi_value arr = make_typed(env, napi_float32_array, (size_t)4 * (size_t)nx * (size_t)ny, 4, &data);
  tq_scene_frame(static_cast<float*>(data), (int)nx, (int)ny);
  return arr;
}

// tracar(sdf, nx, ny, h, x0, y0, x1, y1, raio, apagar) (cena.js:43-72), in place
napi_value f_stroke(napi_env env, napi_callback_info info) {
  napi_value argv[10];
  if (!get_args(env, info, 10, argv, nullptr)) return nullptr;
  float* sdf;
  size_t len;
  if (!f32_array(env, argv[0], &sdf, &len)) return nullptr;
  double v[8];
  for (int k = 0; k < 8; k++)
    if (!num(env, argv[1 + k], &v[k])) return nullptr;
  if (len != (size_t)4 * (size_t)v[0] * (size_t)v[1]) { throw_arg(env, "sdf size"); return nullptr; }
  bool apagar = false;
  if (!is_nullish(env, argv[9])) napi_get_value_bool(env, argv[9], &apagar);
  tq_scene_stroke(sdf, (int)v[0], (int)v[1], v[2], v[3], v[4], v[5], v[6], v[7], apagar ? 1 : 0);
  return undefined(env);
}

// semearRetangulo(x0, y0, x1, y1, h, [a, b, c], semente) -> {particulas, semente}
// (fisica-cpu.js:705-727 with the seeded LCG of testes/apoio.js:4-9; semente = LCG state)
napi_value f_seed(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv, nullptr)) return nullptr;
  double v[5];
  for (int k = 0; k < 5; k++)
    if (!num(env, argv[k], &v[k])) return nullptr;
  double absorcao[3] = {0, 0, 0};
  if (!is_nullish(env, argv[5]))
    for (uint32_t c = 0; c < 3; c++) {
      napi_value e;
      if (napi_get_element(env, argv[5], c, &e) == napi_ok) napi_get_value_double(env, e, &absorcao[c]);
    }
  double sd = 1;
  if (!is_nullish(env, argv[6]) && !num(env, argv[6], &sd)) return nullptr;
  uint32_t s0 = (uint32_t)(int64_t)sd, s = s0;
  const int n = tq_seed_rect(v[0], v[1], v[2], v[3], v[4], absorcao, &s, nullptr, 0);
  void* data = nullptr;
  napi_value arr = make_typed(env, napi_float32_array, (size_t)(n > 0 ? n : 0) * TQ_CAMPOS, 4, &data);
  s = s0;
  if (n > 0 && tq_seed_rect(v[0], v[1], v[2], v[3], v[4], absorcao, &s, static_cast<float*>(data), n) != n) {
    napi_throw_error(env, nullptr, "semearRetangulo: count changed");
    return nullptr;
  }
  napi_value res;
  napi_create_object(env, &res);
  napi_set_named_property(env, res, "particulas", arr);
  napi_set_named_property(env, res, "semente", number(env, s));
  return res;
}

// passoMaximo(h, altura, vMax | null, dtQuadro) (constantes.js:167-175)
napi_value f_step_bound(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv, nullptr)) return nullptr;
  double h, alt, dq, vm = 0;
  if (!num(env, argv[0], &h) || !num(env, argv[1], &alt) || !num(env, argv[3], &dq)) return nullptr;
  const bool has = !is_nullish(env, argv[2]);
  if (has && !num(env, argv[2], &vm)) return nullptr;
  return number(env, tq_passo_maximo(h, alt, vm, has ? 1 : 0, dq));
}

napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor methods[] = {
      {"definirSdf", nullptr, m_set_sdf, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"adicionar", nullptr, m_add, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"aplicarFerramenta", nullptr, m_tool, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"emitirTorneira", nullptr, m_faucet, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"avancar", nullptr, m_step, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"avancarAssincrono", nullptr, m_step_async, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"pedirEstado", nullptr, m_status_request, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"estado", nullptr, m_status_poll, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"gravar", nullptr, m_tape_begin, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"definirTracos", nullptr, m_set_strokes, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"tracosVjp", nullptr, m_strokes_vjp, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"renderizarAgua", nullptr, m_render_water, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fimGravacao", nullptr, m_simple<tq_tape_end>, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fracaoAgua", nullptr, m_fraction, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"retropropagar", nullptr, m_backward, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"subpasso", nullptr, m_substep, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"atualizarAr", nullptr, m_simple<tq_update_air>, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"reiniciarAr", nullptr, m_simple<tq_reset_air>, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"limparAgua", nullptr, m_simple<tq_clear_water>, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"esperar", nullptr, m_simple<tq_sync>, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"contagem", nullptr, m_count, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"truncar", nullptr, m_truncate, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"particulas", nullptr, m_particles, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"lerVMax", nullptr, m_vmax, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"area", nullptr, m_area, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"lerCampo", nullptr, m_field, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"opcoes", nullptr, m_options, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"destruir", nullptr, m_destroy, nullptr, nullptr, nullptr, napi_default, nullptr},
  };
  napi_value cls;
  if (napi_define_class(env, "SimulacaoHIP", NAPI_AUTO_LENGTH, construct, nullptr,
                        sizeof methods / sizeof methods[0], methods, &cls) != napi_ok)
    return nullptr;
  napi_set_named_property(env, exports, "SimulacaoHIP", cls);
  napi_property_descriptor fns[] = {
      {"cenaMoldura", nullptr, f_frame, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"tracar", nullptr, f_stroke, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"semearRetangulo", nullptr, f_seed, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"passoMaximo", nullptr, f_step_bound, nullptr, nullptr, nullptr, napi_default, nullptr},
  };
  napi_define_properties(env, exports, sizeof fns / sizeof fns[0], fns);
  return exports;
}

}  // namespace

NAPI_MODULE(tanque_node, init)
