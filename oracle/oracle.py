"""ctypes binding of the CPU oracle (oracle/tanque_oracle.cpp).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

CAMPOS = 12
MAX_BOLHAS = 4096
SOLIDO, AGUA, AR = 0, 1, 2

# field ids (shared with the product's debug hooks)
F_U, F_V, F_UPESO, F_VPESO, F_UVAL, F_VVAL, F_DENS, F_CONT, F_COR, F_TIPO, F_SOLIDO = range(11)
F_DIV, F_Q, F_FASE, F_NX, F_NY, F_KAPPA, F_QAR, F_ROT = range(11, 19)
F_PBOLHA, F_VOLBOLHA, F_RHSBOLHA, F_YBOLHA, F_VALVO, F_VIVO, F_PART = range(19, 26)
FIELD_KIND = {F_UVAL: "u8", F_VVAL: "u8", F_TIPO: "u8", F_SOLIDO: "u8", F_VIVO: "u8",
              F_CONT: "u32", F_ROT: "u32"}

# phase ids (shared with the product's debug hooks)
PH_P2G, PH_GRAV, PH_TIPOS, PH_ROTULOS, PH_SUPERFICIE, PH_PAR, PH_CLASSIFICAR = range(7)
PH_BOLHAS, PH_PRESSAO, PH_GRADIENTE, PH_EXTRAPOLAR, PH_G2P, PH_AR = range(7, 13)

_lib = None


def build(force=False):
    """Compile the oracle with g++ (no fast-math, no FMA contraction: JS is binary64 round-to-nearest)."""
    import subprocess
    src = os.path.join(HERE, "tanque_oracle.cpp")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                               "-fPIC", "-shared", src, "-o", LIB_PATH + ".tmp"])
        os.replace(LIB_PATH + ".tmp", LIB_PATH)
    build_node(force)
    return LIB_PATH


NODE_ADDON = os.path.join(HERE, "oracle_node.node")


def build_node(force=False):
    """The Node-API addon over this library (node_oracle.cpp), for timing the oracle from Node
    (bench_oracle.js); skipped when node's headers are absent."""
    import subprocess
    src = os.path.join(HERE, "node_oracle.cpp")
    if not os.path.exists("/usr/include/node/node_api.h"):
        return None
    if force or not os.path.exists(NODE_ADDON) or os.path.getmtime(NODE_ADDON) < max(
            os.path.getmtime(src), os.path.getmtime(LIB_PATH)):
        subprocess.check_call(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-I/usr/include/node", src, "-o",
                               NODE_ADDON + ".tmp", "-L" + HERE, "-loracle", "-Wl,-rpath,$ORIGIN",
                               "-Wl,--unresolved-symbols=ignore-all"])
        os.replace(NODE_ADDON + ".tmp", NODE_ADDON)
    return NODE_ADDON


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, d, i, f = C.c_void_p, C.c_double, C.c_int, C.c_float
        sig = {
            "or_passo_maximo": (d, [d, d, d, i, d]),
            "or_lcg_next": (d, [C.POINTER(C.c_uint32)]),
            "or_js_hypot": (d, [d, d]),
            "or_js_round": (d, [d]),
            "or_scene_new": (vp, [i, i, d]),
            "or_scene_free": (None, [vp]),
            "or_scene_stroke": (None, [vp, d, d, d, d, d, i]),
            "or_scene_clear": (None, [vp]),
            "or_scene_sdf": (C.POINTER(f), [vp]),
            "or_scene_set_sdf": (None, [vp, vp]),
            "or_scene_distancia": (d, [vp, d, d]),
            "or_scene_normal": (None, [vp, d, d, C.POINTER(d)]),
            "or_scene_solida": (i, [vp, i, i]),
            "or_seed_rect": (i, [d, d, d, d, d, C.POINTER(d), C.POINTER(C.c_uint32), vp, i]),
            "or_sim_new": (vp, [vp, i, d, d, i, d, d, i]),
            "or_sim_free": (None, [vp]),
            "or_sim_set_tol": (None, [vp, d, i]),
            "or_sim_set_gravity": (None, [vp, d, d]),
            "or_sim_add": (i, [vp, vp, i, i]),
            "or_sim_emit_faucet": (i, [vp, d, d, d, d, d, C.POINTER(d), d, C.POINTER(C.c_uint32), C.POINTER(C.c_int)]),
            "or_sim_step": (None, [vp, d, i]),
            "or_sim_substep": (None, [vp, d]),
            "or_sim_phase": (None, [vp, i, d]),
            "or_sim_update_walls": (None, [vp]),
            "or_sim_reset_air": (None, [vp]),
            "or_sim_clear_water": (None, [vp]),
            "or_sim_tool": (None, [vp, i, d, d, d, d, d, C.POINTER(d)]),
            "or_sim_n": (i, [vp]),
            "or_sim_set_n": (None, [vp, i]),
            "or_sim_vmax": (d, [vp, C.POINTER(i)]),
            "or_sim_area": (d, [vp]),
            "or_sim_ncomp": (i, [vp]),
            "or_sim_solver_stats": (i, [vp, C.POINTER(d), i, C.POINTER(d), C.POINTER(d), C.POINTER(i)]),
            "or_sim_field_len": (C.c_long, [vp, i]),
            "or_sim_read": (C.c_long, [vp, i, vp, i]),
            "or_sim_write": (C.c_long, [vp, i, vp, i, i]),
            "or_sim_set_tracker_first": (None, [vp, i]),
            "or_sim_tracker_first": (i, [vp]),
            "or_sim_mg_vcycle": (None, [vp, vp, vp]),
            "or_sim_op_apply": (None, [vp, vp, vp]),
            "or_sim_set_types_ready": (None, [vp, i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def passo_maximo(h, altura, vmax=None, dt_quadro=1 / 60):
    """passoMaximo (js/constantes.js:167-175)."""
    return lib().or_passo_maximo(h, altura, 0.0 if vmax is None else vmax, 0 if vmax is None else 1, dt_quadro)


class Semente:
    """The seeded LCG of testes/apoio.js:4-9."""

    def __init__(self, s=1):
        self.state = C.c_uint32(s & 0xFFFFFFFF)

    def __call__(self):
        return lib().or_lcg_next(C.byref(self.state))


def semear_retangulo(x0, y0, x1, y1, h, absorcao=(0, 0, 0), aleatorio=None):
    """semearRetangulo (js/fisica-cpu.js:705-727) with the LCG state shared with `aleatorio`."""
    if aleatorio is None:
        aleatorio = Semente(1)
    L = lib()
    a3 = (C.c_double * 3)(*absorcao)
    n = L.or_seed_rect(x0, y0, x1, y1, h, a3, C.byref(aleatorio.state), None, 0)
    out = np.zeros(n * CAMPOS, dtype=np.float32)
    if n:
        got = L.or_seed_rect(x0, y0, x1, y1, h, a3, C.byref(aleatorio.state), out.ctypes.data, n)
        assert got == n
    return out


class Cena:
    """Cena (js/cena.js) -- walls as an SDF on a 2x2-per-cell lattice."""

    def __init__(self, nx, ny, h):
        self.nx, self.ny, self.h = nx, ny, h
        self.r = 2
        self.mw, self.mh = nx * 2, ny * 2
        self.largura, self.altura = nx * h, ny * h
        self.tracos = []
        self._p = lib().or_scene_new(nx, ny, h)

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.or_scene_free(self._p)
            self._p = None

    def tracar(self, x0, y0, x1, y1, raio, apagar=False):
        self.tracos.append((x0, y0, x1, y1, raio, apagar))
        lib().or_scene_stroke(self._p, x0, y0, x1, y1, raio, 1 if apagar else 0)

    def refazer(self, tracos):
        lib().or_scene_clear(self._p)
        self.tracos = []
        for t in tracos:
            self.tracar(*t)

    @property
    def sdf(self):
        p = lib().or_scene_sdf(self._p)
        return np.ctypeslib.as_array(p, shape=(self.mw * self.mh,)).copy()

    def set_sdf(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float32)
        assert arr.size == self.mw * self.mh
        lib().or_scene_set_sdf(self._p, arr.ctypes.data)

    @property
    def solido(self):
        return (self.sdf < 0).astype(np.uint8)

    def distancia(self, x, y):
        return lib().or_scene_distancia(self._p, x, y)

    def normal(self, x, y):
        out = (C.c_double * 2)()
        lib().or_scene_normal(self._p, x, y, out)
        return out[0], out[1]

    def celula_solida(self, i, j):
        return bool(lib().or_scene_solida(self._p, i, j))


class Simulacao:
    """SimulacaoCPU (js/fisica-cpu.js:28-671).  mode 0 = JS-exact (binary32 storage),
    mode 1 = FP64 storage."""

    def __init__(self, cena, gravidade=(0.0, -9.81), iteracoes=3, tau=0.1, tensao=0.0728, ar=True, mode=0,
                 tol=0.0):
        self.cena = cena
        self.nx, self.ny, self.h = cena.nx, cena.ny, cena.h
        self.mode = mode
        self._p = lib().or_sim_new(cena._p, mode, gravidade[0], gravidade[1], iteracoes, tau, tensao, 1 if ar else 0)
        if tol:
            lib().or_sim_set_tol(self._p, tol, iteracoes)

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.or_sim_free(self._p)
            self._p = None

    def set_tol(self, tol, maxit):
        lib().or_sim_set_tol(self._p, tol, maxit)

    def set_gravidade(self, gx, gy):
        lib().or_sim_set_gravity(self._p, gx, gy)

    def adicionar(self, lista, checar=True):
        lista = np.ascontiguousarray(lista, dtype=np.float32)
        return lib().or_sim_add(self._p, lista.ctypes.data, lista.size // CAMPOS, 1 if checar else 0)

    def emitir_torneira(self, x, y, pvx, pvy, raio, absorcao, dt, semente):
        """emitirTorneira (principal.js:301-321) with the seeded LCG: returns (emitted, accepted)."""
        a = (C.c_double * 3)(*absorcao)
        n = C.c_int()
        acc = lib().or_sim_emit_faucet(self._p, x, y, pvx, pvy, raio, a, dt, C.byref(semente.state), C.byref(n))
        return n.value, acc

    def avancar(self, dt, subpassos=1):
        lib().or_sim_step(self._p, dt, subpassos)

    def subpasso(self, dt):
        lib().or_sim_substep(self._p, dt)

    def fase(self, ph, dt=1 / 240):
        lib().or_sim_phase(self._p, ph, dt)

    def atualizar_paredes(self):
        lib().or_sim_update_walls(self._p)

    def reiniciar_ar(self):
        lib().or_sim_reset_air(self._p)

    def limpar_agua(self):
        lib().or_sim_clear_water(self._p)

    def ferramenta(self, tipo, x, y, raio, vx=0.0, vy=0.0, absorcao=(0, 0, 0)):
        t = {"nenhuma": 0, "esponja": 1, "mao": 2, "corante": 3}[tipo]
        lib().or_sim_tool(self._p, t, x, y, raio, vx, vy, (C.c_double * 3)(*absorcao))

    @property
    def n(self):
        return lib().or_sim_n(self._p)

    @n.setter
    def n(self, v):
        lib().or_sim_set_n(self._p, v)

    @property
    def vmax(self):
        has = C.c_int(0)
        v = lib().or_sim_vmax(self._p, C.byref(has))
        return v if has.value else None

    def area(self):
        return lib().or_sim_area(self._p)

    def solver_stats(self):
        norms = (C.c_double * 4096)()
        nb, tr, it = C.c_double(), C.c_double(), C.c_int()
        m = lib().or_sim_solver_stats(self._p, norms, 4096, C.byref(nb), C.byref(tr), C.byref(it))
        return {"normas": list(norms[:min(m, 4096)]), "norm_b": nb.value, "true_resid": tr.value,
                "iters": it.value}

    def read(self, f, as_double=False):
        n = lib().or_sim_field_len(self._p, f)
        kind = FIELD_KIND.get(f)
        if kind == "u8":
            out = np.zeros(n, np.uint8)
        elif kind == "u32":
            out = np.zeros(n, np.uint32)
        else:
            out = np.zeros(n, np.float64 if as_double else np.float32)
        if n:
            lib().or_sim_read(self._p, f, out.ctypes.data, 1 if as_double else 0)
        return out

    def write(self, f, arr):
        kind = FIELD_KIND.get(f)
        if kind == "u8":
            arr = np.ascontiguousarray(arr, np.uint8); as_d = 0
        elif kind == "u32":
            arr = np.ascontiguousarray(arr, np.uint32); as_d = 0
        else:
            as_d = 1 if np.asarray(arr).dtype == np.float64 else 0
            arr = np.ascontiguousarray(arr, np.float64 if as_d else np.float32)
        count = arr.size // CAMPOS if f == F_PART else 0
        lib().or_sim_write(self._p, f, arr.ctypes.data, as_d, count)

    @property
    def p(self):
        return self.read(F_PART)

    def particulas(self):
        return self.read(F_PART).reshape(-1, CAMPOS)

    @property
    def tipo(self):
        return self.read(F_TIPO)

    @property
    def rotulo(self):
        return self.read(F_ROT)

    @property
    def contagem(self):
        return self.read(F_CONT)

    @property
    def solido(self):
        return self.read(F_SOLIDO)

    def vcycle(self, b):
        """One V-cycle of the preconditioner on level-0 input b (test hook; current types)."""
        b = np.ascontiguousarray(b, np.float64)
        out = np.zeros_like(b)
        lib().or_sim_mg_vcycle(self._p, b.ctypes.data, out.ctypes.data)
        return out

    def aplicar(self, v):
        """The pressure operator A on the water cells (test hook; current types)."""
        v = np.ascontiguousarray(v, np.float64)
        out = np.zeros_like(v)
        lib().or_sim_op_apply(self._p, v.ctypes.data, out.ctypes.data)
        return out

    def set_tipos_prontos(self, v=True):
        """Test hook: the stored types count as current for the next air update."""
        lib().or_sim_set_types_ready(self._p, 1 if v else 0)

    def set_primeira(self, v):
        lib().or_sim_set_tracker_first(self._p, 1 if v else 0)


def rodar(sim, segundos, a_cada=None, quadro=1 / 60):
    """The tests' adaptive-substep driver (testes/fisica.test.js:12-19)."""
    import math
    quadros = round_js(segundos / quadro)
    for f in range(quadros):
        sub = math.ceil(quadro / passo_maximo(sim.h, sim.cena.altura, sim.vmax, quadro))
        sim.avancar(quadro, sub)
        if a_cada:
            a_cada(f * quadro)


def round_js(x):
    return int(lib().or_js_round(x))
