"""Python host binding of libtanque.so (the C ABI in include/tanque.h).

`SimulacaoHIP` mirrors the reference's SimulacaoCPU (js/fisica-cpu.js:28-671): same method names,
argument meaning and state accessors, so code written against the reference's CPU engine (its
tests, MotorCPU) drives the MI355X path unchanged.  There is no CPU fallback: importing this
module fails if the HIP library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libtanque.so")

CAMPOS = 12
MAX_BOLHAS = 4096
PREC_FP32, PREC_MIXED, PREC_MIXED_B, PREC_MIXED_X = 0, 1, 2, 3

F_U, F_V, F_UPESO, F_VPESO, F_UVAL, F_VVAL, F_DENS, F_CONT, F_COR, F_TIPO, F_SOLIDO = range(11)
F_DIV, F_Q, F_FASE, F_NX, F_NY, F_KAPPA, F_QAR, F_ROT = range(11, 19)
F_PBOLHA, F_VOLBOLHA, F_RHSBOLHA, F_YBOLHA, F_VALVO, F_VIVO = range(19, 25)
FIELD_DTYPE = {F_UVAL: np.uint8, F_VVAL: np.uint8, F_TIPO: np.uint8, F_SOLIDO: np.uint8, F_VIVO: np.uint8,
               F_CONT: np.uint32, F_ROT: np.uint32}


class TqError(RuntimeError):
    pass


class Desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("h", C.c_double), ("gx", C.c_double), ("gy", C.c_double),
                ("iteracoes", C.c_int), ("tau_densidade", C.c_double), ("tensao", C.c_double), ("ar", C.c_int),
                ("max_particles", C.c_int), ("precision", C.c_int), ("tol", C.c_double),
                ("mg_min_lds_level", C.c_int), ("mg_half_level", C.c_int), ("use_graph", C.c_int),
                ("mg_dense", C.c_int)]


class TuneResult(C.Structure):
    _fields_ = [("spmv_grid", C.c_int), ("update_grid", C.c_int), ("split_cells", C.c_int),
                ("pass_fold_grid", C.c_int), ("mg_dense", C.c_int), ("default_us", C.c_double),
                ("tuned_us", C.c_double), ("trials", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("steps", C.c_int), ("launches", C.c_int), ("upload_bytes", C.c_double), ("record_s", C.c_double),
                ("plan_s", C.c_double), ("stage_s", C.c_double), ("issue_s", C.c_double), ("wait_s", C.c_double),
                ("graph_builds", C.c_int), ("graph_replays", C.c_int)]


class Tool(C.Structure):
    _fields_ = [("kind", C.c_int), ("x", C.c_double), ("y", C.c_double), ("raio", C.c_double),
                ("vx", C.c_double), ("vy", C.c_double), ("absorcao", C.c_double * 3)]


class SolverStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("norm0", C.c_double), ("norm", C.c_double), ("norm_b", C.c_double),
                ("true_resid", C.c_double), ("levels", C.c_int), ("active_bubbles", C.c_int),
                ("total_iterations", C.c_longlong), ("solves", C.c_longlong), ("operator_builds", C.c_longlong),
                ("true_resid_iterate", C.c_double), ("first_iterations", C.c_int), ("restarts", C.c_int),
                ("air_components", C.c_int), ("overlap_pairs", C.c_int), ("bubble_candidates", C.c_int)]


class Faucet(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("pvx", C.c_double), ("pvy", C.c_double),
                ("raio", C.c_double), ("absorcao", C.c_double * 3)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("total_ms", C.c_double), ("launches", C.c_long),
                ("bytes_per_launch", C.c_double)]


# every symbol include/tanque.h declares, with its ctypes signature
SIGNATURES = {
    "tq_desc_default": (None, [C.POINTER(Desc), C.c_int, C.c_int, C.c_double]),
    "tq_create": (C.c_int, [C.POINTER(Desc), C.POINTER(C.c_void_p)]),
    "tq_destroy": (C.c_int, [C.c_void_p]),
    "tq_set_sdf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
    "tq_add_particles": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "tq_apply_tool": (C.c_int, [C.c_void_p, C.POINTER(Tool)]),
    "tq_emit_faucet": (C.c_int, [C.c_void_p, C.POINTER(Faucet), C.c_double, C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "tq_batch_create_table": (C.c_int, [C.POINTER(Desc), C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "tq_batch_create": (C.c_int, [C.POINTER(Desc), C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "tq_batch_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "tq_batch_tank": (C.c_void_p, [C.c_void_p, C.c_int]),
    "tq_batch_step": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_int)]),
    "tq_batch_vmax": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "tq_batch_stats_get": (C.c_int, [C.c_void_p, C.POINTER(BatchStats), C.c_int]),
    "tq_batch_use_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "tq_batch_destroy": (C.c_int, [C.c_void_p]),
    "tq_step": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "tq_substep": (C.c_int, [C.c_void_p, C.c_double]),
    "tq_update_air": (C.c_int, [C.c_void_p]),
    "tq_reset_air": (C.c_int, [C.c_void_p]),
    "tq_clear_water": (C.c_int, [C.c_void_p]),
    "tq_truncate": (C.c_int, [C.c_void_p, C.c_int]),
    "tq_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "tq_read_particles": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "tq_vmax": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "tq_area": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "tq_set_options": (C.c_int, [C.c_void_p, C.POINTER(Desc)]),
    "tq_read_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
    "tq_write_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
    "tq_write_particles": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "tq_set_tracker_first": (C.c_int, [C.c_void_p, C.c_int]),
    "tq_run_phase": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "tq_debug_vcycle": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tq_tape_begin": (C.c_int, [C.c_void_p, C.c_int]),
    "tq_tape_end": (C.c_int, [C.c_void_p]),
    "tq_tape_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                               C.POINTER(C.c_int)]),
    "tq_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p,
                              C.c_void_p, C.c_int]),
    "tq_water_fraction": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                    C.c_void_p, C.c_int]),
    "tq_water_fraction_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(C.c_double)]),
    "tq_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(TuneResult)]),
    "tq_step_async": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "tq_status_request": (C.c_int, [C.c_void_p]),
    "tq_status_poll": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                                 C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "tq_tape_records": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "tq_set_strokes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "tq_read_sdf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "tq_strokes_vjp": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "tq_tape_record": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_size_t]),
    "tq_solver_stats_get": (C.c_int, [C.c_void_p, C.POINTER(SolverStats), C.c_int]),
    "tq_timing_enable": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p]),
    "tq_timing_read": (C.c_int, [C.c_void_p, C.POINTER(KernelTime), C.c_int, C.POINTER(C.c_int)]),
    "tq_render_water": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "tq_set_launch": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "tq_get_launch": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int)]),
    "tq_sync": (C.c_int, [C.c_void_p]),
    "tq_last_error": (C.c_char_p, []),
    "tq_scene_frame": (None, [C.c_void_p, C.c_int, C.c_int]),
    "tq_scene_stroke": (None, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_double, C.c_int]),
    "tq_seed_rect": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_void_p, C.c_int]),
    "tq_passo_maximo": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_int, C.c_double]),
}

_lib = None


def load():
    """Loads libtanque.so; raises if it is not built (there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TqError("libtanque.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _chk(rc):
    if rc != 0:
        raise TqError("libtanque error %d: %s" % (rc, load().tq_last_error().decode(errors="replace")))


# ------------------------------------------------------------------ host helpers (scene, seeding)
class Cena:
    """Walls as an SDF on a 2x2-per-cell lattice (the reference's Cena, js/cena.js), built by the
    library's host helpers."""

    def __init__(self, nx, ny, h):
        self.nx, self.ny, self.h = nx, ny, h
        self.largura, self.altura = nx * h, ny * h
        self.sdf = np.zeros(4 * nx * ny, np.float32)
        self.tracos = []
        load().tq_scene_frame(self.sdf.ctypes.data, nx, ny)

    def tracar(self, x0, y0, x1, y1, raio, apagar=False):
        self.tracos.append((x0, y0, x1, y1, raio, apagar))
        load().tq_scene_stroke(self.sdf.ctypes.data, self.nx, self.ny, self.h, x0, y0, x1, y1, raio, 1 if apagar else 0)


class Semente:
    """Seeded LCG (testes/apoio.js:4-9) shared with seed_rect."""

    def __init__(self, s=1):
        self.state = C.c_uint32(s & 0xFFFFFFFF)


def semear_retangulo(x0, y0, x1, y1, h, absorcao=(0, 0, 0), semente=None):
    """semearRetangulo (fisica-cpu.js:705-727)."""
    semente = semente or Semente(1)
    L = load()
    a3 = (C.c_double * 3)(*absorcao)
    n = L.tq_seed_rect(x0, y0, x1, y1, h, a3, C.byref(semente.state), None, 0)
    out = np.zeros(max(n, 0) * CAMPOS, np.float32)
    if n > 0:
        got = L.tq_seed_rect(x0, y0, x1, y1, h, a3, C.byref(semente.state), out.ctypes.data, n)
        assert got == n
    return out


def passo_maximo(h, altura, vmax=None, dt_quadro=1 / 60):
    return load().tq_passo_maximo(h, altura, 0.0 if vmax is None else vmax, 0 if vmax is None else 1, dt_quadro)


# ------------------------------------------------------------------ the simulation
class SimulacaoHIP:
    """SimulacaoCPU's interface on the MI355X path (js/fisica-cpu.js:28-671)."""

    def __init__(self, cena, gravidade=(0.0, -9.81), iteracoes=3, tauDensidade=0.1, tensao=0.0728, ar=True,
                 precision=PREC_FP32, tol=0.0, max_particles=0, mg_min_lds_level=0, mg_half_level=-1,
                 use_graph=False, mg_dense=None):
        L = load()
        d = Desc()
        L.tq_desc_default(C.byref(d), cena.nx, cena.ny, cena.h)
        d.gx, d.gy = gravidade
        d.iteracoes, d.tau_densidade, d.tensao, d.ar = iteracoes, tauDensidade, tensao, 1 if ar else 0
        d.precision, d.tol, d.max_particles = precision, tol, max_particles
        d.mg_min_lds_level = mg_min_lds_level
        d.mg_half_level = mg_half_level
        d.use_graph = 1 if use_graph else 0
        d.mg_dense = -1 if mg_dense is None else (1 if mg_dense else 0)
        self._desc = d
        h = C.c_void_p()
        _chk(L.tq_create(C.byref(d), C.byref(h)))
        self._h = h
        self.cena, self.nx, self.ny, self.h = cena, cena.nx, cena.ny, cena.h
        self.arAtivo = bool(ar)
        self.atualizarParedes()

    @classmethod
    def _view(cls, handle, cena, desc, ar=True):
        """A SimulacaoHIP on a handle owned elsewhere (a batch's tank); closing it does nothing."""
        self = cls.__new__(cls)
        self._desc = desc
        self._h = C.c_void_p(handle)
        self._owned = False
        self.cena, self.nx, self.ny, self.h = cena, cena.nx, cena.ny, cena.h
        self.arAtivo = bool(ar)
        self.atualizarParedes()
        return self

    def close(self):
        if getattr(self, "_h", None) and _lib is not None and getattr(self, "_owned", True):
            _lib.tq_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    # ---- options
    def set_options(self, **kw):
        d = self._desc
        for k, v in kw.items():
            if k == "gravidade":
                d.gx, d.gy = v
            elif k == "ar":
                d.ar = 1 if v else 0
                self.arAtivo = bool(v)
            else:
                setattr(d, {"tauDensidade": "tau_densidade"}.get(k, k), v)
        _chk(load().tq_set_options(self._h, C.byref(d)))

    @property
    def gravidade(self):
        return (self._desc.gx, self._desc.gy)

    @gravidade.setter
    def gravidade(self, g):
        self.set_options(gravidade=g)

    # ---- the reference's methods
    def atualizarParedes(self, row0=0, row1=None):
        sdf = np.ascontiguousarray(self.cena.sdf, np.float32)
        row1 = 2 * self.ny - 1 if row1 is None else row1
        _chk(load().tq_set_sdf(self._h, sdf.ctypes.data, row0, row1))

    def adicionar(self, lista, checarCheia=True):
        lista = np.ascontiguousarray(lista, np.float32)
        acc = C.c_int(0)
        _chk(load().tq_add_particles(self._h, lista.ctypes.data, lista.size // CAMPOS, 1 if checarCheia else 0,
                                     C.byref(acc)))
        return acc.value

    def aplicarFerramenta(self, f):
        if not f or f.get("tipo", "nenhuma") == "nenhuma":
            return
        t = Tool()
        t.kind = {"esponja": 1, "mao": 2, "corante": 3}[f["tipo"]]
        t.x, t.y, t.raio = f["x"], f["y"], f["raio"]
        t.vx, t.vy = f.get("vx", 0.0), f.get("vy", 0.0)
        for k, a in enumerate(f.get("absorcao", (0, 0, 0))):
            t.absorcao[k] = a
        _chk(load().tq_apply_tool(self._h, C.byref(t)))

    def emitirTorneira(self, x, y, pvx, pvy, raio, absorcao, dt, semente):
        """The faucet tool (emitirTorneira, principal.js:301-321) with the seeded LCG in place of
        Math.random; `semente` advances.  Returns (emitted, accepted)."""
        f = Faucet(x, y, pvx, pvy, raio, (C.c_double * 3)(*absorcao))
        n, acc = C.c_int(), C.c_int()
        _chk(load().tq_emit_faucet(self._h, C.byref(f), dt, C.byref(semente.state), C.byref(n), C.byref(acc)))
        return n.value, acc.value

    def avancar(self, dt, subpassos=1):
        _chk(load().tq_step(self._h, dt, subpassos))

    def autotune(self, iterations=10):
        """tq_autotune: the solver's launch shapes tuned on the current pressure system; returns
        the choice and the solve times of the defaults and of the choice."""
        r = TuneResult()
        _chk(load().tq_autotune(self._h, iterations, C.byref(r)))
        return {k: getattr(r, k) for k, _ in TuneResult._fields_}

    def avancar_assincrono(self, dt, subpassos=1):
        """The frame queued without waiting, a status readback behind it (tq_step_async)."""
        _chk(load().tq_step_async(self._h, dt, subpassos))

    def pedir_estado(self):
        _chk(load().tq_status_request(self._h))

    def estado(self):
        """None while the device has not reached the last status request, else a dict with n,
        vMax (None before the first G2P) and ms_gpu (the last queued frame's device time)."""
        ready, n, has = C.c_int(), C.c_int(), C.c_int()
        v, ms = C.c_double(), C.c_double()
        _chk(load().tq_status_poll(self._h, C.byref(ready), C.byref(n), C.byref(v), C.byref(has), C.byref(ms)))
        if not ready.value:
            return None
        return {"n": n.value, "vMax": v.value if has.value else None, "ms_gpu": ms.value if ms.value >= 0 else None}

    # ---- walls built on the device
    def definir_tracos(self, tracos):
        """The walls from capsule strokes [(x0, y0, x1, y1, raio, apagar), ...] built on the
        device (Cena.refazer + atualizarParedes); the host Cena is not updated."""
        a = np.ascontiguousarray(np.asarray([[float(v) for v in t] for t in tracos], np.float64).reshape(-1, 6))
        _chk(load().tq_set_strokes(self._h, a.ctypes.data if len(a) else None, len(a)))

    def tracos_vjp(self, tracos, gsdf):
        """sum(gsdf * d(sdf)/d(stroke)) per stroke: (n, 5) for x0 y0 x1 y1 raio."""
        a = np.ascontiguousarray(np.asarray([[float(v) for v in t] for t in tracos], np.float64).reshape(-1, 6))
        g = np.ascontiguousarray(np.asarray(gsdf, np.float32).reshape(-1))
        out = np.zeros((max(len(a), 1), 5))
        _chk(load().tq_strokes_vjp(self._h, a.ctypes.data if len(a) else None, len(a), g.ctypes.data, out.ctypes.data))
        return out[: len(a)]

    def ler_sdf(self):
        """The device's SDF lattice (4 nx ny floats)."""
        L = load()
        out = np.zeros(4 * self.nx * self.ny, np.float32)
        _chk(L.tq_read_sdf(self._h, out.ctypes.data, out.nbytes))
        return out

    # ---- differentiation of recorded frames
    def tape_begin(self, checkpoint_every=10):
        """Records the frames run from now on (emitirTorneira and avancar calls) for backward();
        checkpoint_every = 0 keeps every substep's record instead of checkpoints."""
        _chk(load().tq_tape_begin(self._h, checkpoint_every))

    def tape_end(self):
        _chk(load().tq_tape_end(self._h))

    def tape_info(self):
        f, e, c, n = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _chk(load().tq_tape_info(self._h, C.byref(f), C.byref(e), C.byref(c), C.byref(n)))
        return {"frames": f.value, "faucet_calls": e.value, "checkpoints": c.value, "n_begin": n.value}

    def backward(self, grad_state, tol=1e-6, sdf=True, state0=True):
        """dL/d(inputs) of the recorded frames from dL/d(final state) (n x 8: x y vx vy cux cuy cvx
        cvy per live particle, particulas() order): returns dict(faucet=(calls x 2) jet velocity
        gradients, sdf=(4 nx ny) or None, state0=(n0 x 8) or None)."""
        g = np.ascontiguousarray(np.asarray(grad_state, np.float32).reshape(-1, 8))
        info = self.tape_info()
        gf = np.zeros(2 * max(1, info["faucet_calls"]))
        gs = np.zeros(4 * self.cena.nx * self.cena.ny, np.float32) if sdf else None
        g0 = np.zeros(8 * max(1, info["n_begin"]), np.float32) if state0 else None
        _chk(load().tq_backward(self._h, g.ctypes.data, len(g), tol, gf.ctypes.data, info["faucet_calls"],
                                gs.ctypes.data if sdf else None, g0.ctypes.data if state0 else None,
                                info["n_begin"]))
        return {"faucet": gf[: 2 * info["faucet_calls"]].reshape(-1, 2), "sdf": gs,
                "state0": g0[: 8 * info["n_begin"]].reshape(-1, 8) if state0 else None}

    def water_fraction(self, box, grad=False):
        """Mean of min(1, densidade/PPC) over the cell box (i0, j0, i1, j1); with grad, also its
        gradient with respect to the state (n x 8, particulas() order)."""
        f = C.c_double()
        if not grad:
            _chk(load().tq_water_fraction(self._h, *[int(v) for v in box], C.byref(f), None, 0))
            return f.value
        n = self.n
        g = np.zeros((max(n, 1), 8), np.float32)
        _chk(load().tq_water_fraction(self._h, *[int(v) for v in box], C.byref(f), g.ctypes.data, n))
        return f.value, g[:n]

    def tape_records(self):
        n = C.c_int()
        _chk(load().tq_tape_records(self._h, C.byref(n)))
        return n.value

    def tape_record(self, k):
        """Test hook: record k of a keep-all tape as a dict of arrays."""
        L = load()
        nx, ny = self.cena.nx, self.cena.ny
        out = {}
        cnt = np.zeros(1, np.uint32)
        _chk(L.tq_tape_record(self._h, k, 0, cnt.ctypes.data, 4))
        n = int(cnt[0])
        spec = {"tipo": (1, np.uint8, nx * ny), "rot": (2, np.uint32, nx * ny), "qAr": (3, np.float32, nx * ny),
                "volAlvo": (4, np.float32, 4096), "yB": (5, np.float32, 4096), "flags": (6, np.uint16, n),
                "seq": (7, np.uint32, n)}
        for name, (fid, dt, m) in spec.items():
            a = np.zeros(max(m, 1), dt)
            if m:
                _chk(L.tq_tape_record(self._h, k, fid, a.ctypes.data, a.itemsize * m))
            out[name] = a[:m]
        d = np.zeros(1, np.float64)
        _chk(L.tq_tape_record(self._h, k, 8, d.ctypes.data, 8))
        out["dt"] = float(d[0])
        out["n"] = n
        return out

    def subpasso(self, dt):
        _chk(load().tq_substep(self._h, dt))

    def atualizarAr(self):
        _chk(load().tq_update_air(self._h))

    def reiniciarAr(self):
        _chk(load().tq_reset_air(self._h))

    def limparAgua(self):
        _chk(load().tq_clear_water(self._h))

    def area(self):
        a = C.c_double()
        _chk(load().tq_area(self._h, C.byref(a)))
        return a.value

    @property
    def n(self):
        n = C.c_int()
        _chk(load().tq_count(self._h, C.byref(n)))
        return n.value

    @n.setter
    def n(self, m):
        _chk(load().tq_truncate(self._h, m))

    @property
    def vMax(self):
        v, has = C.c_double(), C.c_int()
        _chk(load().tq_vmax(self._h, C.byref(v), C.byref(has)))
        return v.value if has.value else None

    def particulas(self):
        n = C.c_int()
        _chk(load().tq_read_particles(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value * CAMPOS, np.float32)
        _chk(load().tq_read_particles(self._h, out.ctypes.data, n.value, C.byref(n)))
        return out.reshape(-1, CAMPOS)

    @property
    def p(self):
        return self.particulas().reshape(-1)

    # ---- state (typed arrays of the reference)
    def field_len(self, f):
        nx, ny = self.nx, self.ny
        if f in (F_U, F_UPESO, F_UVAL):
            return (nx + 1) * ny
        if f in (F_V, F_VPESO, F_VVAL):
            return nx * (ny + 1)
        if f == F_COR:
            return 3 * nx * ny
        if f >= F_PBOLHA:
            return MAX_BOLHAS
        return nx * ny

    def read(self, f):
        out = np.zeros(self.field_len(f), FIELD_DTYPE.get(f, np.float32))
        _chk(load().tq_read_field(self._h, f, out.ctypes.data, out.nbytes))
        return out

    def write(self, f, arr):
        arr = np.ascontiguousarray(arr, FIELD_DTYPE.get(f, np.float32))
        assert arr.size == self.field_len(f)
        _chk(load().tq_write_field(self._h, f, arr.ctypes.data, arr.nbytes))

    def write_particles(self, P):
        P = np.ascontiguousarray(P, np.float32).reshape(-1)
        _chk(load().tq_write_particles(self._h, P.ctypes.data, P.size // CAMPOS))

    def set_primeira(self, v):
        _chk(load().tq_set_tracker_first(self._h, 1 if v else 0))

    def fase(self, ph, dt=1 / 240):
        _chk(load().tq_run_phase(self._h, ph, dt))

    def vcycle(self, b):
        """z = M b: one application of the multigrid preconditioner on the current types (test hook)."""
        b = np.ascontiguousarray(b, np.float32)
        z = np.empty_like(b)
        _chk(load().tq_debug_vcycle(self._h, b.ctypes.data, z.ctypes.data))
        return z

    @property
    def tipo(self):
        return self.read(F_TIPO)

    @property
    def rotulo(self):
        return self.read(F_ROT)

    @property
    def contagemCelula(self):
        return self.read(F_CONT)

    @property
    def solidoCelula(self):
        return self.read(F_SOLIDO)

    def solver_stats(self, true_residual=False):
        st = SolverStats()
        _chk(load().tq_solver_stats_get(self._h, C.byref(st), 1 if true_residual else 0))
        return {k: getattr(st, k) for k, _ in SolverStats._fields_}

    def timing_enable(self, on=True, only=None):
        _chk(load().tq_timing_enable(self._h, 1 if on else 0, only.encode() if only else None))

    def render_water(self):
        """The water layer of the reference's 2D renderer (js/render-2d.js:110-162) at 3 pixels per
        cell: (tint, foam), two uint8 RGBA images of shape (3 ny, 3 nx, 4), rows from the top."""
        w, h = 3 * self.nx, 3 * self.ny
        tint = np.empty((h, w, 4), np.uint8)
        foam = np.empty((h, w, 4), np.uint8)
        _chk(load().tq_render_water(self._h, w, h, tint.ctypes.data, foam.ctypes.data))
        return tint, foam

    def set_launch(self, name, value):
        """Launch cap of a solver kernel ("spmv_grid", "update_grid"; 0 = default)."""
        _chk(load().tq_set_launch(self._h, name.encode(), int(value)))

    def get_launch(self, name):
        v = C.c_int()
        _chk(load().tq_get_launch(self._h, name.encode(), C.byref(v)))
        return v.value

    def timing_read(self):
        n = C.c_int()
        _chk(load().tq_timing_read(self._h, None, 0, C.byref(n)))
        arr = (KernelTime * max(n.value, 1))()
        _chk(load().tq_timing_read(self._h, arr, n.value, C.byref(n)))
        return {arr[k].name.decode(): {"ms": arr[k].total_ms, "launches": arr[k].launches,
                                       "bytes_per_launch": arr[k].bytes_per_launch} for k in range(n.value)}

    def sync(self):
        _chk(load().tq_sync(self._h))


class Batch:
    """Independent tanks stepped together (tq_batch_*): `cenas` with one options dict each
    (SimulacaoHIP's keywords); `tanks[k]` is a SimulacaoHIP view of tank k for set-up and reads;
    step(dt, substeps) runs avancar(dt, substeps[k]) on every tank.  mode "table": one launch per
    kernel of the step over all tanks (tq_batch_create_table; tol must be 0), the frames recorded
    on `threads` host threads; mode "threads": every tank on its own stream, enqueued from
    `threads` host threads (tq_batch_create)."""

    def __init__(self, cenas, opcoes, threads=8, mode="table"):
        L = load()
        n = len(cenas)
        descs = (Desc * n)()
        for k, (cena, o) in enumerate(zip(cenas, opcoes)):
            L.tq_desc_default(C.byref(descs[k]), cena.nx, cena.ny, cena.h)
            d = descs[k]
            d.gx, d.gy = o.get("gravidade", (0.0, -9.81))
            d.iteracoes = o.get("iteracoes", 3)
            d.tau_densidade = o.get("tauDensidade", 0.1)
            d.tensao = o.get("tensao", 0.0728)
            d.ar = 1 if o.get("ar", True) else 0
            d.precision = o.get("precision", PREC_FP32)
            d.tol = o.get("tol", 0.0)
            d.use_graph = 1 if o.get("use_graph", False) else 0
            if o.get("mg_dense") is not None:
                d.mg_dense = 1 if o["mg_dense"] else 0
        if mode not in ("table", "threads"):
            raise ValueError("Batch: mode is 'table' or 'threads'")
        h = C.c_void_p()
        create = L.tq_batch_create_table if mode == "table" else L.tq_batch_create
        _chk(create(descs, n, threads, C.byref(h)))
        self._h = h
        self.mode = mode
        self.tanks = []
        for k, cena in enumerate(cenas):
            dk = Desc()
            C.memmove(C.byref(dk), C.byref(descs[k]), C.sizeof(Desc))
            self.tanks.append(SimulacaoHIP._view(L.tq_batch_tank(h, k), cena, dk, bool(dk.ar)))

    def step(self, dt, substeps):
        arr = (C.c_int * len(self.tanks))(*[int(x) for x in substeps])
        _chk(load().tq_batch_step(self._h, dt, arr))

    def vmax(self):
        """every tank's vMax (a list; None where not measured yet), read after a table step
        without a device wait (tq_batch_vmax)"""
        n = len(self.tanks)
        v = (C.c_double * n)()
        has = (C.c_int * n)()
        _chk(load().tq_batch_vmax(self._h, v, has))
        return [v[k] if has[k] else None for k in range(n)]

    def use_graphs(self, on):
        """table mode: capture each frame's table launches as a HIP graph and replay it while the
        launch list repeats (tq_batch_use_graphs; on by default)"""
        _chk(load().tq_batch_use_graphs(self._h, 1 if on else 0))

    def stats(self, reset=False):
        """host seconds of the table steps by part, steps, the last step's launches and upload
        bytes (tq_batch_stats_get)"""
        st = BatchStats()
        _chk(load().tq_batch_stats_get(self._h, C.byref(st), 1 if reset else 0))
        return {f: getattr(st, f) for f, _ in BatchStats._fields_}

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            for t in self.tanks:
                t._h = None
            _lib.tq_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def rodar(sim, segundos, a_cada=None, quadro=1 / 60):
    """The reference tests' adaptive-substep driver (testes/fisica.test.js:12-19)."""
    import math
    quadros = int(math.floor(segundos / quadro + 0.5))
    for f in range(quadros):
        sub = math.ceil(quadro / passo_maximo(sim.h, sim.cena.altura, sim.vMax, quadro))
        sim.avancar(quadro, sub)
        if a_cada:
            a_cada(f * quadro)
