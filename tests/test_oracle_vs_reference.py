"""Pins the oracle against the reference's own Fortran for every routine that
builds without FFTW (oracle/build_ref.sh -> oracle/_ref/libref.so):
MT19937 (mod_mt.f90), ludcmp/lubksb and Mixing_Simple/Lambda/Anderson
(mod_iterate.f90, mixing_dim = 2 as shipped) and the Flory-Huggins fields and
energy (mod_FloryHuggins.f90) on the Main_Diblock_Melt grid (64 x 1 x 300).
The FFT and propagator routines need FFTW, which is absent, so they are pinned
by the closed forms in test_oracle.py instead."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="oracle/_ref/libref.so not built (needs flang + reference tree)")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


@pytest.fixture(scope="module")
def ref():
    lib = C.CDLL(LIB)
    lib.ref_grnd.restype = C.c_double
    lib.ref_mixing_dim.restype = C.c_int
    lib.ref_interaction_energy.restype = C.c_double
    lib.ref_interaction_energy.argtypes = [_dp]
    lib.ref_set_box.argtypes = [C.c_double] * 3
    for f in ("ref_mixing_anderson", "ref_mixing_lambda", "ref_mixing_simple"):
        getattr(lib, f).argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_int]
    lib.ref_ludcmp.argtypes = [_dp, C.c_int, _ip, _dp]
    lib.ref_lubksb.argtypes = [_dp, C.c_int, _ip, _dp]
    return lib


def p(a):
    return a.ctypes.data_as(_dp)


def test_mt19937_stream(ref, orc):
    for seed in (2007, 4357, 1, 123456789):
        ref.ref_sgrnd(seed)
        got = np.array([ref.ref_grnd() for _ in range(1500)])
        mt = orc.MT(seed)
        assert np.array_equal(got, mt.grnd(1500))


def test_ludcmp_lubksb_bitwise(ref, orc):
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 7, 20, 40):
        a = rng.standard_normal((n, n))
        if n >= 3:
            a[1] = a[0]  # duplicate row: exercises the tiny-pivot path on some columns
        af = np.asfortranarray(a).copy(order="F")
        buf = np.ascontiguousarray(af.T)  # column-major bytes
        indx = np.zeros(n, dtype=np.int32)
        d = C.c_double()
        ref.ref_ludcmp(p(buf), n, indx.ctypes.data_as(_ip), C.byref(d))
        lu_o, indx_o, d_o, _ = orc.ludcmp(a)
        assert np.array_equal(buf.reshape(n, n).T, lu_o) and np.array_equal(indx, indx_o) and d.value == d_o
        b = rng.standard_normal(n)
        b1 = b.copy()
        ref.ref_lubksb(p(buf), n, indx.ctypes.data_as(_ip), p(b1))
        b2 = b.copy()
        orc.lib().orc_lubksb(np.ascontiguousarray(lu_o.T).ctypes.data_as(_dp), n, indx_o.ctypes.data_as(_ip),
                             b2.ctypes.data_as(_dp))
        assert np.array_equal(b1, b2)


@pytest.mark.parametrize("it", [1, 2, 3, 4, 7])
def test_mixing_routines(ref, orc, it):
    m = ref.ref_mixing_dim()
    assert m == 2
    rng = np.random.default_rng(it)
    n = 777
    F = rng.standard_normal((m + 1, n))
    D = rng.standard_normal((m + 1, n))
    for name, fn in (("anderson", orc.mixing_anderson), ("lambda", orc.mixing_lambda)):
        f1, d1 = F.copy(), D.copy()
        getattr(ref, f"ref_mixing_{name}")(p(f1), p(d1), n, 0.5, it)
        f2, d2 = F.copy(), D.copy()
        fn(f2, d2, 0.5, it)
        assert np.array_equal(d1, d2)
        assert np.allclose(f1, f2, rtol=1e-14, atol=1e-15), name
    f1, d1 = F.copy(), D.copy()
    ref.ref_mixing_simple(p(f1), p(d1), n, 0.3, it)
    f2, d2 = F.copy(), D.copy()
    orc.mixing_simple(f2, d2, 0.3)
    assert np.array_equal(f1, f2)


def test_flory_huggins(ref, orc, tmp_path, monkeypatch):
    dims = np.zeros(4, dtype=np.int32)
    ref.ref_grid(dims.ctypes.data_as(_ip))
    nx, ny, nz, nt = (int(v) for v in dims)
    (tmp_path / "input_interactions").write_text("25.0d0  ! kappa\n\n17.5d0  ! chi\n\n1.2d0\n0.8d0\n")
    monkeypatch.chdir(tmp_path)
    ref.ref_initialize_interactions()
    ref.ref_set_box(3.6, 1.0, 3.6)
    g = orc.Grid(nx, ny, nz, 3.6, 1.0, 3.6, 0.001)
    fh = orc.FH(25.0, [[0.0, 17.5], [17.5, 0.0]], [1.2, 0.8])
    den = np.random.default_rng(0).uniform(0.05, 0.95, (nt, nz, ny, nx))
    f_ref = np.empty_like(den)
    ref.ref_mean_fields(p(den), p(f_ref))
    assert np.array_equal(f_ref, orc.mean_fields(g, fh, den))
    e_ref = ref.ref_interaction_energy(p(den))
    assert abs(e_ref - orc.interaction_energy(g, fh, den)) <= 1e-12 * abs(e_ref)
