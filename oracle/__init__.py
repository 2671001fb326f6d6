"""CPU FP64 parity oracle for the Bulk_3D_FFT SCFT path -- TEST INFRASTRUCTURE ONLY.

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker.  The product
library never routes through it.

``ORACLE`` wraps ``liboracle.so`` (plain C, built by ``oracle/Makefile``) with
numpy-facing helpers.  Arrays follow the reference's Fortran layout: a field
``(nx,ny,nz)`` is held here as a C-ordered numpy array of shape
``(nz, ny, nx)`` (x fastest), a field set ``(nx,ny,nz,k)`` as ``(k, nz, ny, nx)``
and a half spectrum ``complex(8)(nx/2+1,ny,nz)`` as ``(nz, ny, nx//2+1)``
complex128.

Parity pins (each function is checked by the tests named here):
  * FFT, Laplace operator: numpy.fft (pocketfft) cross-check, brute-force DFT
    sums at 4^3..6^3, delta / single-mode known answers (tests/test_oracle.py).
  * Propagate step and chain: w=0 / constant-w closed forms, single-mode decay,
    separable-field factorisation into 1-D schemes, translation equivariance,
    the contour invariant dvol*sum q(s)q+(Ns-s) = Q, mass and swap symmetry,
    and the independent numpy restatement in ``oracle/np_ref.py``.
  * Flory-Huggins fields/energy: per-point closed forms.
  * Mixing (simple/lambda/Anderson) and ludcmp/lubksb: numpy restatement,
    LAPACK solve, the iter<=m known answer, and -- when it builds -- the
    reference's own mod_iterate.f90 compiled by ``oracle/build_ref.sh``.
  * MT19937: numpy's MT19937 driven from the same sgrnd state, and the
    reference's mod_mt.f90 compiled into ``oracle/_ref`` when it builds.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("ds", C.c_double)]


class _MT(C.Structure):
    _fields_ = [("mt", C.c_uint * 624), ("mti", C.c_int)]


class _Chain(C.Structure):
    _fields_ = [("nblocks", C.c_int), ("monomer_type", _ip), ("block_segments", _ip),
                ("ensemble", C.c_int), ("rho", C.c_double), ("mu", C.c_double),
                ("Qsum", C.c_double), ("energy", C.c_double)]


class _FH(C.Structure):
    _fields_ = [("ncomp", C.c_int), ("kappa", C.c_double), ("chi", _dp), ("volume", _dp)]


class _Blend(C.Structure):
    _fields_ = [("npol", C.c_int), ("types", _ip), ("segs", _ip), ("rho", _dp), ("mu", C.c_double * 8),
                ("energy", C.c_double * 8)]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    G = C.POINTER(_Grid)
    sig = {
        "orc_mt_sgrnd": (None, [C.POINTER(_MT), C.c_int]),
        "orc_mt_next32": (C.c_uint, [C.POINTER(_MT)]),
        "orc_mt_grnd": (C.c_double, [C.POINTER(_MT)]),
        "orc_real2fourier": (None, [G, _dp, _dp]),
        "orc_fourier2real": (None, [G, _dp, _dp]),
        "orc_calculate_laplace": (None, [G, _dp, _dp, _dp]),
        "orc_volume_parameters": (None, [G, _dp, _dp]),
        "orc_dft_line": (None, [C.c_int, _dp, _dp, C.c_int]),
        "orc_calculate_expfield": (None, [G, C.c_int, _dp, _dp]),
        "orc_propagate_step": (None, [G, _dp, _dp]),
        "orc_propagate_chain": (C.c_int, [G, C.POINTER(_Chain), _dp, _dp, _dp, _dp, _dp]),
        "orc_block_segments": (C.c_int, [C.c_double, C.c_double]),
        "orc_set_contour_rule": (None, [C.c_int]),
        "orc_contour_weight": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_double]),
        "orc_propagate_homopolymer": (C.c_int, [G, C.POINTER(_Chain), _dp, _dp, _dp]),
        "orc_update_solvent": (C.c_int, [G, C.POINTER(_Chain), _dp, _dp, _dp]),
        "orc_mean_fields": (None, [G, C.POINTER(_FH), _dp, _dp]),
        "orc_interaction_energy": (C.c_double, [G, C.POINTER(_FH), _dp]),
        "orc_average_volume_fraction": (C.c_double, [G, C.POINTER(_FH), _dp]),
        "orc_mixing_simple": (None, [_dp, _dp, C.c_long, C.c_int, C.c_double]),
        "orc_mixing_lambda": (None, [_dp, _dp, C.c_long, C.c_int, C.c_double, C.c_int]),
        "orc_mixing_anderson": (None, [_dp, _dp, C.c_long, C.c_int, C.c_double, C.c_int]),
        "orc_ludcmp": (C.c_int, [_dp, C.c_int, _ip, _dp]),
        "orc_lubksb": (None, [_dp, C.c_int, _ip, _dp]),
        "orc_solve_lu": (None, [_dp, _dp, _dp, C.c_int]),
        "orc_solve_cholesky": (None, [_dp, _dp, _dp, C.c_int]),
        "orc_species_densities": (C.c_int, [G, C.POINTER(_Chain), C.c_int, _dp, _dp]),
        "orc_total_energy": (C.c_double, [G, C.POINTER(_Chain), C.POINTER(_FH), C.c_int, _dp, _dp]),
        "orc_scf_anderson": (C.c_int, [G, C.POINTER(_Chain), C.POINTER(_FH), C.c_int, _dp, _dp,
                                       C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp]),
        "orc_mobility_qq": (None, [G, C.c_int, C.c_int, C.c_double, _dp, C.c_double, _dp]),
        "orc_conjugate_fields": (C.c_int, [G, C.POINTER(_Chain), C.c_int, _dp, _dp, C.c_int, C.c_double,
                                           C.c_double, C.c_int, C.c_int, _dp]),
        "orc_gcfstat": (C.c_int, [C.c_double, C.c_double]),
        "orc_ddft_step": (C.c_int, [G, C.POINTER(_Chain), C.POINTER(_FH), C.c_int, _dp, _dp, _dp, C.c_double,
                                    C.c_int, C.c_double, C.c_int, C.c_int, _ip, _ip]),
        "orc_blend_densities": (C.c_int, [G, C.POINTER(_Blend), C.c_int, _dp, _dp]),
        "orc_blend_conjugate_fields": (C.c_int, [G, C.POINTER(_Blend), C.c_int, _dp, _dp, C.c_int, C.c_double,
                                                 C.c_double, C.c_int, _dp]),
        "orc_field_derivative": (C.c_int, [G, C.POINTER(_Blend), C.POINTER(_FH), C.c_int, _dp, _dp]),
        "orc_epd_timestep_euler": (C.c_int, [G, C.POINTER(_Blend), C.POINTER(_FH), C.c_int, _dp, C.c_double]),
        "orc_epd_timestep_rk4": (C.c_int, [G, C.POINTER(_Blend), C.POINTER(_FH), C.c_int, _dp, C.c_double]),
        "orc_epd_step": (C.c_int, [G, C.POINTER(_Blend), C.POINTER(_FH), C.c_int, _dp, _dp, C.c_double, C.c_int,
                                   C.c_double, C.c_int, _ip]),
        "orc_blend_scf": (C.c_int, [G, C.POINTER(_Blend), C.POINTER(_FH), C.c_int, _dp, _dp, C.c_int, C.c_int,
                                    C.c_double, C.c_double, C.c_int, _dp]),
        "orc_diblock_saddle": (C.c_int, [G, C.POINTER(_Chain), C.POINTER(_FH), C.c_int, _dp, _dp, C.c_int, C.c_int,
                                         C.c_double, C.c_double, C.c_int, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _pi(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


class Grid:
    """Box geometry: grid (nx,ny,nz), side lengths (lx,ly,lz) and contour step ds."""

    def __init__(self, nx, ny, nz, lx, ly, lz, ds):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.lx, self.ly, self.lz = float(lx), float(ly), float(lz)
        self.ds = float(ds)
        self._g = _Grid(self.nx, self.ny, self.nz, self.lx, self.ly, self.lz, self.ds)

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def kshape(self):
        return (self.nz, self.ny, self.nx // 2 + 1)

    @property
    def ng(self):
        return self.nx * self.ny * self.nz

    def volume(self):
        v, dv = C.c_double(), C.c_double()
        lib().orc_volume_parameters(C.byref(self._g), C.byref(v), C.byref(dv))
        return v.value, dv.value


class MT:
    """The reference's MT19937 stream (mod_mt.f90): sgrnd seeding, grnd in [0,1]."""

    def __init__(self, seed):
        self._s = _MT()
        lib().orc_mt_sgrnd(C.byref(self._s), int(seed))

    def next32(self, n=None):
        if n is None:
            return lib().orc_mt_next32(C.byref(self._s))
        return np.array([lib().orc_mt_next32(C.byref(self._s)) for _ in range(n)], dtype=np.uint32)

    def grnd(self, n=None):
        if n is None:
            return lib().orc_mt_grnd(C.byref(self._s))
        return np.array([lib().orc_mt_grnd(C.byref(self._s)) for _ in range(n)])


def dft_line(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().orc_dft_line(len(x), x.view(np.float64).ctypes.data_as(_dp), out.view(np.float64).ctypes.data_as(_dp), int(sign))
    return out


def real2fourier(g: Grid, prop):
    prop = np.ascontiguousarray(prop, dtype=np.float64).reshape(g.shape)
    out = np.empty(g.kshape, dtype=np.complex128)
    lib().orc_real2fourier(C.byref(g._g), _p(prop), out.view(np.float64).ctypes.data_as(_dp))
    return out


def fourier2real(g: Grid, spec):
    spec = np.ascontiguousarray(spec, dtype=np.complex128).reshape(g.kshape)
    out = np.empty(g.shape, dtype=np.float64)
    lib().orc_fourier2real(C.byref(g._g), spec.view(np.float64).ctypes.data_as(_dp), _p(out))
    return out


def calculate_laplace(g: Grid):
    """Returns (laplace, explaplace, nabla[3]) as complex arrays of shape kshape."""
    lap = np.empty(g.kshape, dtype=np.complex128)
    el = np.empty(g.kshape, dtype=np.complex128)
    nab = np.empty((3,) + g.kshape, dtype=np.complex128)
    lib().orc_calculate_laplace(C.byref(g._g), lap.view(np.float64).ctypes.data_as(_dp),
                                el.view(np.float64).ctypes.data_as(_dp), nab.view(np.float64).ctypes.data_as(_dp))
    return lap, el, nab


def expfield(g: Grid, field):
    field = np.ascontiguousarray(field, dtype=np.float64)
    nt = field.size // g.ng
    out = np.empty_like(field)
    lib().orc_calculate_expfield(C.byref(g._g), nt, _p(field), _p(out))
    return out


def propagate_step(g: Grid, prop, expfield_type):
    q = np.array(prop, dtype=np.float64, order="C").reshape(g.shape)
    e = np.ascontiguousarray(expfield_type, dtype=np.float64).reshape(g.shape)
    lib().orc_propagate_step(C.byref(g._g), _p(q), _p(e))
    return q


def block_segments(block_length, ds):
    return lib().orc_block_segments(float(block_length), float(ds))


RULES = {"trapezoid": 0, "simpson": 1}


def set_contour_rule(rule):
    """Contour rule of propagate_chain / propagate_homopolymer and everything built on them
    (process-wide): "trapezoid" (the reference's, default) or "simpson"."""
    lib().orc_set_contour_rule(RULES.get(rule, rule))


def contour_weight(rule, j, nb, ds):
    return lib().orc_contour_weight(RULES.get(rule, rule), int(j), int(nb), float(ds))


class ChainResult:
    def __init__(self, **kw):
        self.__dict__.update(kw)


def propagate_chain(g: Grid, block_types, block_segs, expfield, ensemble=1, rho=1.0, mu=0.0, keep_q=False):
    """Propagate_Multiblock.  expfield: (ntypes, nz, ny, nx).  Returns a ChainResult with
    distribution, density (nblocks, nz, ny, nx), Qsum, mu, rho, energy [, q, qdag]."""
    bt = np.ascontiguousarray(block_types, dtype=np.int32)
    bs = np.ascontiguousarray(block_segs, dtype=np.int32)
    nb = len(bt)
    ns = int(bs.sum())
    e = np.ascontiguousarray(expfield, dtype=np.float64)
    dist = np.empty((nb,) + g.shape)
    den = np.empty((nb,) + g.shape)
    q = np.empty((ns + 1,) + g.shape) if keep_q else None
    qd = np.empty((ns + 1,) + g.shape) if keep_q else None
    ch = _Chain(nb, _pi(bt), _pi(bs), int(ensemble), float(rho), float(mu), 0.0, 0.0)
    rc = lib().orc_propagate_chain(C.byref(g._g), C.byref(ch), _p(e), _p(dist), _p(den),
                                   _p(q) if keep_q else None, _p(qd) if keep_q else None)
    if rc != 0:
        raise ValueError("Invalid ensemble; Cannot calculate Density")
    return ChainResult(distribution=dist, density=den, Qsum=ch.Qsum, mu=ch.mu, rho=ch.rho,
                       energy=ch.energy, q=q, qdag=qd)


def propagate_homopolymer(g: Grid, monomer_type, segments, expfield, ensemble=1, rho=1.0, mu=0.0):
    """Propagate_Homopolymer (mod_homopolymer.f90:90-141).  Returns a ChainResult with
    distribution, density (nz, ny, nx), Qsum, mu, rho, energy."""
    bt = np.array([monomer_type], dtype=np.int32)
    bs = np.array([segments], dtype=np.int32)
    e = np.ascontiguousarray(expfield, dtype=np.float64)
    dist = np.empty(g.shape)
    den = np.empty(g.shape)
    ch = _Chain(1, _pi(bt), _pi(bs), int(ensemble), float(rho), float(mu), 0.0, 0.0)
    rc = lib().orc_propagate_homopolymer(C.byref(g._g), C.byref(ch), _p(e), _p(dist), _p(den))
    if rc != 0:
        raise ValueError("Invalid ensemble; Cannot calculate Density")
    return ChainResult(distribution=dist, density=den, Qsum=ch.Qsum, mu=ch.mu, rho=ch.rho, energy=ch.energy)


def update_solvent(g: Grid, field, ensemble=1, rho=1.0, mu=0.0):
    """Update_Solvent (mod_solvent.f90:85-108) for one field (nz, ny, nx)."""
    bt = np.array([1], dtype=np.int32)
    bs = np.array([0], dtype=np.int32)
    f = np.ascontiguousarray(field, dtype=np.float64)
    dist = np.empty(g.shape)
    den = np.empty(g.shape)
    ch = _Chain(1, _pi(bt), _pi(bs), int(ensemble), float(rho), float(mu), 0.0, 0.0)
    rc = lib().orc_update_solvent(C.byref(g._g), C.byref(ch), _p(f), _p(dist), _p(den))
    if rc != 0:
        raise ValueError("Invalid ensemble; Cannot calculate Density")
    return ChainResult(distribution=dist, density=den, Qsum=ch.Qsum, mu=ch.mu, rho=ch.rho, energy=ch.energy)


class FH:
    """Flory-Huggins parameters: kappa, chi (ncomp x ncomp symmetric), monomer volumes."""

    def __init__(self, kappa, chi, volume):
        self.chi = np.asfortranarray(np.array(chi, dtype=np.float64))
        self.vol = np.ascontiguousarray(volume, dtype=np.float64)
        self.kappa = float(kappa)
        self._chi_f = np.ascontiguousarray(self.chi.T)  # column-major flat
        self._fh = _FH(len(self.vol), self.kappa, _p(self._chi_f), _p(self.vol))


def mean_fields(g: Grid, fh: FH, density):
    d = np.ascontiguousarray(density, dtype=np.float64)
    out = np.empty_like(d)
    lib().orc_mean_fields(C.byref(g._g), C.byref(fh._fh), _p(d), _p(out))
    return out


def interaction_energy(g: Grid, fh: FH, density):
    d = np.ascontiguousarray(density, dtype=np.float64)
    return lib().orc_interaction_energy(C.byref(g._g), C.byref(fh._fh), _p(d))


def average_volume_fraction(g: Grid, fh: FH, density):
    d = np.ascontiguousarray(density, dtype=np.float64)
    return lib().orc_average_volume_fraction(C.byref(g._g), C.byref(fh._fh), _p(d))


def _hist(a):
    # (m+1, fieldsize) C-order == (fieldsize, 0:m) column-major
    assert a.ndim == 2 and a.dtype == np.float64 and a.flags.c_contiguous
    return a


def mixing_simple(field, dfield, lam):
    _hist(field); _hist(dfield)
    lib().orc_mixing_simple(_p(field), _p(dfield), field.shape[1], field.shape[0] - 1, float(lam))


def mixing_lambda(field, dfield, lammax, it):
    _hist(field); _hist(dfield)
    lib().orc_mixing_lambda(_p(field), _p(dfield), field.shape[1], field.shape[0] - 1, float(lammax), int(it))


def mixing_anderson(field, dfield, lam, it):
    """In-place Mixing_Anderson on history arrays of shape (m+1, fieldsize)."""
    _hist(field); _hist(dfield)
    lib().orc_mixing_anderson(_p(field), _p(dfield), field.shape[1], field.shape[0] - 1, float(lam), int(it))


def ludcmp(a):
    """a: (n,n) numpy matrix (row i, column j).  Returns (lu, indx 1-based, d, singular)."""
    n = a.shape[0]
    af = np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)  # column-major
    indx = np.zeros(n, dtype=np.int32)
    d = C.c_double()
    sing = lib().orc_ludcmp(_p(af), n, _pi(indx), C.byref(d))
    return af.T.copy(), indx, d.value, bool(sing)


def solve_lu(a, b):
    n = a.shape[0]
    af = np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.empty(n)
    lib().orc_solve_lu(_p(af), _p(bb), _p(x), n)
    return x


def solve_cholesky(a, b):
    n = a.shape[0]
    af = np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.empty(n)
    lib().orc_solve_cholesky(_p(af), _p(bb), _p(x), n)
    return x


def species_densities(g: Grid, block_types, block_segs, field, ensemble=1, rho=1.0, mu=0.0):
    bt = np.ascontiguousarray(block_types, dtype=np.int32)
    bs = np.ascontiguousarray(block_segs, dtype=np.int32)
    f = np.ascontiguousarray(field, dtype=np.float64)
    nt = f.shape[0]
    den = np.empty_like(f)
    ch = _Chain(len(bt), _pi(bt), _pi(bs), int(ensemble), float(rho), float(mu), 0.0, 0.0)
    rc = lib().orc_species_densities(C.byref(g._g), C.byref(ch), nt, _p(f), _p(den))
    if rc != 0:
        raise ValueError("Invalid ensemble")
    return den, ChainResult(Qsum=ch.Qsum, mu=ch.mu, rho=ch.rho, energy=ch.energy, chain=ch, bt=bt, bs=bs)


def total_energy(g: Grid, chain_info, fh: FH, density, field):
    d = np.ascontiguousarray(density, dtype=np.float64)
    f = np.ascontiguousarray(field, dtype=np.float64)
    return lib().orc_total_energy(C.byref(g._g), C.byref(chain_info.chain), C.byref(fh._fh), d.shape[0], _p(d), _p(f))


def scf_anderson(g: Grid, block_types, block_segs, fh: FH, field, m, lam=0.5, tol=1e-8, maxit=100,
                 ensemble=1, rho=1.0, mu=0.0):
    """Find_SCF_Saddle_Point with Anderson mixing.  Returns dict(field, density, iters, accuracy,
    residuals, Qsum, mu, energy) where field is the mixed field the loop ends with."""
    bt = np.ascontiguousarray(block_types, dtype=np.int32)
    bs = np.ascontiguousarray(block_segs, dtype=np.int32)
    f = np.array(field, dtype=np.float64, order="C")
    den = np.empty_like(f)
    res = np.zeros(maxit)
    acc = C.c_double()
    ch = _Chain(len(bt), _pi(bt), _pi(bs), int(ensemble), float(rho), float(mu), 0.0, 0.0)
    it = lib().orc_scf_anderson(C.byref(g._g), C.byref(ch), C.byref(fh._fh), f.shape[0], _p(f), _p(den),
                                int(m), float(lam), float(tol), int(maxit), C.byref(acc), _p(res))
    n = min(it, maxit)
    return dict(field=f, density=den, iters=it, accuracy=acc.value, residuals=res[:n],
                Qsum=ch.Qsum, mu=ch.mu, rho=ch.rho, energy=ch.energy)


# ---- DDFT (Main_Multiblock/main_multiblock.f90) ------------------------------

def mobility_qq(g: Grid, table, dq, wcutoff):
    """Read_Mobilities' table part (main_multiblock.f90:486-504).  table: (nt, nt, nq) in C
    order, i.e. Fortran mobility_table(nq, nt, nt) with table[t2, t1, iq-1] =
    mobility_table(iq, t1, t2).  Returns mobility_qq as (nt, nt, nz, ny, nx/2+1), the C view
    of Fortran mobility_qq(nx/2+1, ny, nz, nt, nt)."""
    t = np.ascontiguousarray(table, dtype=np.float64)
    nt, nq = t.shape[0], t.shape[2]
    out = np.empty((nt, nt) + g.kshape, dtype=np.float64)
    lib().orc_mobility_qq(C.byref(g._g), nt, nq, float(dq), _p(t), float(wcutoff), _p(out))
    return out


def conjugate_fields(g: Grid, block_types, block_segs, density, field, m, lam=0.5, thresh=5e-4, maxit=2000,
                     measure=0, ensemble=1, rho=1.0, mu=0.0):
    """Get_Conjugate_Fields (main_multiblock.f90:566-640 for measure 0, main_blend.f90:376-428
    for measure 1).  Returns dict(field, iters, accuracy)."""
    bt = np.ascontiguousarray(block_types, dtype=np.int32)
    bs = np.ascontiguousarray(block_segs, dtype=np.int32)
    d = np.ascontiguousarray(density, dtype=np.float64)
    f = np.array(field, dtype=np.float64, order="C")
    acc = C.c_double()
    ch = _Chain(len(bt), _pi(bt), _pi(bs), int(ensemble), float(rho), float(mu), 0.0, 0.0)
    it = lib().orc_conjugate_fields(C.byref(g._g), C.byref(ch), f.shape[0], _p(d), _p(f), int(m), float(lam),
                                    float(thresh), int(maxit), int(measure), C.byref(acc))
    if it < 0:
        raise ValueError("Invalid ensemble")
    return dict(field=f, iters=it, accuracy=acc.value)


def gcfstat(accuracy, accuracy_scf):
    return lib().orc_gcfstat(float(accuracy), float(accuracy_scf))


def ddft_step(g: Grid, block_types, block_segs, fh: FH, density, field, mob_qq, timestep, qdmin=True,
              accuracy_scf=1e-3, itermax_scf=2000, m=10, ensemble=1, rho=1.0, mu=0.0):
    """DDFT_Timestep_Euler (main_multiblock.f90:794-880).  density and field are updated in
    place (C-contiguous float64).  Returns dict(rc, gcfstat, iters, mu, rho, energy) with the
    chain scalars of the last density evaluation."""
    bt = np.ascontiguousarray(block_types, dtype=np.int32)
    bs = np.ascontiguousarray(block_segs, dtype=np.int32)
    mq = np.ascontiguousarray(mob_qq, dtype=np.float64)
    st = np.zeros(2, dtype=np.int32)
    its = np.zeros(2, dtype=np.int32)
    ch = _Chain(len(bt), _pi(bt), _pi(bs), int(ensemble), float(rho), float(mu), 0.0, 0.0)
    rc = lib().orc_ddft_step(C.byref(g._g), C.byref(ch), C.byref(fh._fh), density.shape[0], _p(density),
                             _p(field), _p(mq), float(timestep), int(bool(qdmin)), float(accuracy_scf),
                             int(itermax_scf), int(m), _pi(st), _pi(its))
    return dict(rc=rc, gcfstat=st.tolist(), iters=its.tolist(), mu=ch.mu, rho=ch.rho, energy=ch.energy,
                Qsum=ch.Qsum)


# ---- homopolymer blend and EPD (Main_Blend/main_blend.f90) ----------------------

# single-precision literals assigned to doubles in Main_Blend
EPD_EPS = float(np.float32(0.0001))          # main_blend.f90:630
BLEND_CONJ_ACCURACY = float(np.float32(0.01))  # main_blend.f90:389


class Blend:
    """npol homopolymers: polymer i of monomer type types[i] with segs[i] segments,
    canonical at concentration rho[i]."""

    def __init__(self, types, segs, rho):
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        self.segs = np.ascontiguousarray(segs, dtype=np.int32)
        self.rho = np.ascontiguousarray(rho, dtype=np.float64)
        self._b = _Blend(len(self.types), _pi(self.types), _pi(self.segs), _p(self.rho))

    @property
    def mu(self):
        return list(self._b.mu[:len(self.types)])

    @property
    def energy(self):
        return list(self._b.energy[:len(self.types)])


def blend_densities(g: Grid, blend: Blend, field):
    f = np.ascontiguousarray(field, dtype=np.float64)
    den = np.empty_like(f)
    if lib().orc_blend_densities(C.byref(g._g), C.byref(blend._b), f.shape[0], _p(f), _p(den)) != 0:
        raise ValueError("Invalid ensemble")
    return den


def blend_conjugate_fields(g: Grid, blend: Blend, density, field, m=2, lam=0.5, thresh=BLEND_CONJ_ACCURACY,
                           maxit=100000):
    d = np.ascontiguousarray(density, dtype=np.float64)
    f = np.array(field, dtype=np.float64, order="C")
    acc = C.c_double()
    it = lib().orc_blend_conjugate_fields(C.byref(g._g), C.byref(blend._b), f.shape[0], _p(d), _p(f), int(m),
                                          float(lam), float(thresh), int(maxit), C.byref(acc))
    return dict(field=f, iters=it, accuracy=acc.value)


def epd_step(g: Grid, blend: Blend, fh: FH, density, field, timestep, m=2, accuracy_scf=BLEND_CONJ_ACCURACY,
             itermax=100000):
    """Chain_Dynamics_Timestep_Euler (main_blend.f90:605-667); density and field (C-contiguous
    float64) are updated in place.  Returns the conjugate solve's loop index."""
    it = C.c_int()
    rc = lib().orc_epd_step(C.byref(g._g), C.byref(blend._b), C.byref(fh._fh), density.shape[0], _p(density),
                            _p(field), float(timestep), int(m), float(accuracy_scf), int(itermax), C.byref(it))
    if rc != 0:
        raise ValueError("Invalid ensemble")
    return it.value


def field_derivative(g: Grid, blend: Blend, fh: FH, field):
    """Calculate_Field_Derivative (main_blend.f90:709-740): -laplace * (newfield - field)."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    out = np.empty_like(f)
    if lib().orc_field_derivative(C.byref(g._g), C.byref(blend._b), C.byref(fh._fh), f.shape[0], _p(f), _p(out)):
        raise ValueError("Invalid ensemble")
    return out


def epd_timestep(g: Grid, blend: Blend, fh: FH, field, timestep, method="euler"):
    """EPD_Timestep_Euler (main_blend.f90:580-596) or EPD_Timestep_RungeKutta (:678-701);
    returns the new field."""
    f = np.array(field, dtype=np.float64, order="C")
    fn = lib().orc_epd_timestep_euler if method == "euler" else lib().orc_epd_timestep_rk4
    if fn(C.byref(g._g), C.byref(blend._b), C.byref(fh._fh), f.shape[0], _p(f), float(timestep)):
        raise ValueError("Invalid ensemble")
    return f


BLEND_SCF_ACCURACY = float(np.float32(0.0000001))  # main_blend.f90:451


def blend_scf(g: Grid, blend: Blend, fh: FH, field, m=2, mixing_type=2, lam=0.5, thresh=BLEND_SCF_ACCURACY,
              maxit=100000):
    """Find_SCF_Saddle_Point of Main_Blend (main_blend.f90:437-486).  Returns dict(field,
    density, iters, accuracy)."""
    f = np.array(field, dtype=np.float64, order="C")
    d = np.empty_like(f)
    acc = C.c_double()
    it = lib().orc_blend_scf(C.byref(g._g), C.byref(blend._b), C.byref(fh._fh), f.shape[0], _p(f), _p(d), int(m),
                             int(mixing_type), float(lam), float(thresh), int(maxit), C.byref(acc))
    return dict(field=f, density=d, iters=it, accuracy=acc.value)


def diblock_saddle(g: Grid, block_types, block_segs, fh: FH, field, density, m=2, mixing_type=2, lam=0.1,
                   tol=1e-12, maxit=100000, ensemble=2, rho=1.0, mu=0.0):
    """Calculate_Saddle_Point of Main_Diblock_Melt (main_diblock.f90:398-451).  Returns
    dict(field (the last newfield), mixed_field, density (the last newdensity), iters, accuracy,
    mu, rho, energy)."""
    bt = np.ascontiguousarray(block_types, dtype=np.int32)
    bs = np.ascontiguousarray(block_segs, dtype=np.int32)
    f = np.array(field, dtype=np.float64, order="C")
    d = np.array(density, dtype=np.float64, order="C")
    mf = np.empty_like(f)
    acc = C.c_double()
    ch = _Chain(len(bt), _pi(bt), _pi(bs), int(ensemble), float(rho), float(mu), 0.0, 0.0)
    it = lib().orc_diblock_saddle(C.byref(g._g), C.byref(ch), C.byref(fh._fh), f.shape[0], _p(f), _p(d), int(m),
                                  int(mixing_type), float(lam), float(tol), int(maxit), C.byref(acc), _p(mf))
    return dict(field=f, mixed_field=mf, density=d, iters=it, accuracy=acc.value, mu=ch.mu, rho=ch.rho,
                energy=ch.energy)
