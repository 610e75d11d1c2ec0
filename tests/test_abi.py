"""The C ABI boundary, checked without a GPU: the library loads, exports every entry point
include/tanque.h declares, and its host-side helpers (scene SDF, seeding, substep bound) agree
bit-for-bit with the oracle's restatement of js/cena.js, fisica-cpu.js:705-727 and
constantes.js:167-175."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "tanque.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tq_[a-z_0-9]+)\s*\(", src)))


@pytest.fixture(scope="module")
def tq():
    import tanque_hip
    return tanque_hip


def test_library_exports_every_declared_symbol(tq):
    lib = tq.load()
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding binds exactly the declared set
    assert sorted(tq.SIGNATURES) == names


def test_scene_frame_matches_oracle(tq, oracle):
    for nx, ny in ((40, 30), (64, 48), (17, 23)):
        c = tq.Cena(nx, ny, 0.01)
        o = oracle.Cena(nx, ny, 0.01)
        assert np.array_equal(c.sdf.view(np.uint32), o.sdf.view(np.uint32))


def test_scene_strokes_match_oracle_bitwise(tq, oracle):
    rnd = oracle.Semente(7)
    nx, ny, h = 70, 50, 0.01
    c = tq.Cena(nx, ny, h)
    o = oracle.Cena(nx, ny, h)
    for _ in range(40):
        t = (rnd() * 0.7, rnd() * 0.5, rnd() * 0.7, rnd() * 0.5, 0.013 + rnd() * 0.03)
        apagar = rnd() < 0.3
        c.tracar(*t, apagar)
        o.tracar(*t, apagar)
    assert np.array_equal(c.sdf.view(np.uint32), o.sdf.view(np.uint32))


def test_seed_rect_matches_oracle_bitwise(tq, oracle):
    for args in ((0.02, 0.02, 0.3, 0.42, 0.01), (0.1, 0.05, 0.02, 0.3, 0.0037), (0.5, 0.5, 0.5, 0.5, 0.01)):
        a = tq.semear_retangulo(*args, absorcao=(1.0, 2.5, 60.0), semente=tq.Semente(9))
        b = oracle.semear_retangulo(*args, absorcao=(1.0, 2.5, 60.0), aleatorio=oracle.Semente(9))
        assert a.shape == b.shape
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_lcg_known_answers(oracle):
    # s = (s*1664525 + 1013904223) mod 2^32, value s/2^32 (testes/apoio.js:4-9); the first
    # states from s = 1 follow directly from the recurrence.
    r = oracle.Semente(1)
    s = 1
    for _ in range(5):
        s = (s * 1664525 + 1013904223) % 2 ** 32
        assert r() == s / 2 ** 32


def test_passo_maximo_matches_oracle(tq, oracle):
    for h in (0.01, 0.002963, 0.0015873):
        for vmax in (None, 0.0, 0.3, 2.0, 7.5):
            assert tq.passo_maximo(h, 0.8, vmax, 1 / 240) == oracle.passo_maximo(h, 0.8, vmax, 1 / 240)


def test_errors_cross_the_boundary_as_codes(tq):
    lib = tq.load()
    d = tq.Desc()
    lib.tq_desc_default(C.byref(d), 2, 2, 0.01)  # too small: rejected before any device call
    h = C.c_void_p()
    rc = lib.tq_create(C.byref(d), C.byref(h))
    assert rc == 1 and not h.value
    assert b"grid" in lib.tq_last_error()
    assert lib.tq_step(None, 0.01, 1) == 1
