"""adscf_amd: MI355X (gfx950) pseudo-spectral SCFT propagator path of ADSCF's Bulk_3D_FFT.

Python host binding of ``libadscf_hip.so`` (C-ABI declared in ``include/adscf.h``).
Every call runs on the GPU; there is no CPU fallback: importing works without a
GPU, but creating a :class:`Context` raises when the library or the device is
missing.

Arrays use the reference's Fortran layout expressed as C-ordered numpy arrays:
a field ``real(8)(nx,ny,nz)`` is ``(nz, ny, nx)``, a field set
``(nx,ny,nz,k)`` is ``(k, nz, ny, nx)``, a half spectrum
``complex(8)(nx/2+1,ny,nz)`` is ``(nz, ny, nx//2+1)`` complex128, and an
Anderson history ``(fieldsize, 0:m)`` is ``(m+1, fieldsize)``.

The reference's procedure names are available as thin aliases
(:func:`Calculate_Expfield`, :func:`Propagate_Step`, :func:`Propagate_Diblock`,
...), so parity tests read like calls into the reference modules
(Bulk_3D_FFT/General_Routines/mod_propagate.f90, Polymer_Modules/mod_diblock.f90,
Interaction_Modules/mod_FloryHuggins.f90, General_Routines/mod_iterate.f90).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field as dc_field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libadscf_hip.so")

FP64 = 0
FP32 = 1

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_ENSEMBLE = -3
ERR_STATE = -4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_ctxp = C.c_void_p

# name -> (restype, argtypes); the single source of truth for the exported surface
SIGNATURES = {
    "adscf_init": (C.c_int, [C.POINTER(_ctxp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                             C.c_double, C.c_double, C.c_int]),
    "adscf_finalize": (C.c_int, [_ctxp]),
    "adscf_set_box": (C.c_int, [_ctxp, C.c_double, C.c_double, C.c_double]),
    "adscf_set_ds": (C.c_int, [_ctxp, C.c_double]),
    "adscf_set_precision": (C.c_int, [_ctxp, C.c_int]),
    "adscf_volume": (C.c_int, [_ctxp, _dp, _dp]),
    "adscf_last_error": (C.c_char_p, [_ctxp]),
    "adscf_real2fourier": (C.c_int, [_ctxp, _dp, _dp]),
    "adscf_fourier2real": (C.c_int, [_ctxp, _dp, _dp]),
    "adscf_laplace": (C.c_int, [_ctxp, _dp, _dp, _dp]),
    "adscf_expfield": (C.c_int, [_ctxp, _dp, _dp]),
    "adscf_propagate_step": (C.c_int, [_ctxp, _dp, _dp]),
    "adscf_propagate_chain": (C.c_int, [_ctxp, C.c_int, _ip, _ip, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "adscf_propagate_chain_noalloc": (C.c_int, [_ctxp, C.c_int, _ip, _ip, C.c_int, C.c_double, C.c_double,
                                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "adscf_propagate_homopolymer": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                              _dp]),
    "adscf_update_solvent": (C.c_int, [_ctxp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "adscf_set_interactions": (C.c_int, [_ctxp, C.c_double, _dp, _dp]),
    "adscf_mean_fields": (C.c_int, [_ctxp, _dp, _dp]),
    "adscf_interaction_energy": (C.c_int, [_ctxp, _dp, _dp]),
    "adscf_average_volume_fraction": (C.c_int, [_ctxp, _dp, _dp]),
    "adscf_mixing_simple": (C.c_int, [_ctxp, _dp, _dp, C.c_long, C.c_int, C.c_double]),
    "adscf_mixing_lambda": (C.c_int, [_ctxp, _dp, _dp, C.c_long, C.c_int, C.c_double, C.c_int]),
    "adscf_mixing_anderson": (C.c_int, [_ctxp, _dp, _dp, C.c_long, C.c_int, C.c_double, C.c_int]),
    "adscf_ludcmp": (C.c_int, [_dp, C.c_int, _ip, _dp]),
    "adscf_device_lu_solve": (C.c_int, [_ctxp, _dp, C.c_int, _dp, _dp, _ip, _ip]),
    "adscf_lubksb": (C.c_int, [_dp, C.c_int, _ip, _dp]),
    "adscf_solve_lu": (C.c_int, [_dp, _dp, _dp, C.c_int]),
    "adscf_solve_cholesky": (C.c_int, [_dp, _dp, _dp, C.c_int]),
    "adscf_scf_setup": (C.c_int, [_ctxp, C.c_int, _ip, _ip, C.c_int, C.c_double, C.c_double, C.c_int,
                                  C.c_double]),
    "adscf_scf_set_field": (C.c_int, [_ctxp, _dp]),
    "adscf_scf_iterate": (C.c_int, [_ctxp, C.c_int, C.c_double, _ip, _dp, _ip]),
    "adscf_scf_evaluate": (C.c_int, [_ctxp]),
    "adscf_scf_get": (C.c_int, [_ctxp, _dp, _dp]),
    "adscf_scf_energy": (C.c_int, [_ctxp, _dp, _dp, _dp, _dp, _dp]),
    "adscf_scf_residuals": (C.c_int, [_ctxp, _dp, C.c_int, _ip]),
    "adscf_scf_timing": (C.c_int, [_ctxp, _dp, _dp]),
    "adscf_scf_set_residual": (C.c_int, [_ctxp, C.c_int, C.c_int, _dp]),
    "adscf_scf_set_mixing": (C.c_int, [_ctxp, C.c_int]),
    "adscf_set_contour_rule": (C.c_int, [_ctxp, C.c_int]),
    "adscf_scf_set_density": (C.c_int, [_ctxp, _dp]),
    "adscf_conjugate_fields": (C.c_int, [_ctxp, _dp, C.c_double, C.c_int, C.c_int, _ip, _dp]),
    "adscf_mobility_qq": (C.c_int, [_ctxp, C.c_int, C.c_double, _dp, C.c_double, _dp]),
    "adscf_ddft_set_mobility": (C.c_int, [_ctxp, _dp]),
    "adscf_ddft_set_density": (C.c_int, [_ctxp, _dp]),
    "adscf_ddft_get_density": (C.c_int, [_ctxp, _dp]),
    "adscf_ddft_step": (C.c_int, [_ctxp, C.c_double, C.c_int, C.c_double, C.c_int, _ip, _ip]),
    "adscf_ddft_energy": (C.c_int, [_ctxp, _dp, _dp, _dp]),
    "adscf_blend_setup": (C.c_int, [_ctxp, C.c_int, _ip, _ip, _dp, C.c_int, C.c_double]),
    "adscf_blend_scalars": (C.c_int, [_ctxp, _dp, _dp, _dp]),
    "adscf_epd_step": (C.c_int, [_ctxp, C.c_double, C.c_double, C.c_int, _ip]),
    "adscf_field_derivative": (C.c_int, [_ctxp, _dp]),
    "adscf_epd_field_step": (C.c_int, [_ctxp, C.c_double, C.c_int]),
    "adscf_autotune": (C.c_int, [_ctxp, C.c_int, C.c_char_p, C.c_int]),
    "adscf_get_schedule": (C.c_int, [_ctxp, C.c_int, _ip, _ip, _ip, _ip]),
    "adscf_set_schedule": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "adscf_set_pass_tiles": (C.c_int, [_ctxp, C.c_int, C.c_int]),
    "adscf_get_pass_tiles": (C.c_int, [_ctxp, C.c_int, _ip]),
    "adscf_set_plane_passes": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int]),
    "adscf_get_plane_passes": (C.c_int, [_ctxp, _ip, _ip, _ip]),
    "adscf_set_middle_axis": (C.c_int, [_ctxp, C.c_int]),
    "adscf_get_middle_axis": (C.c_int, [_ctxp, _ip]),
    "adscf_batch_create": (C.c_int, [C.POINTER(_ctxp), C.c_int, _ip, _dp, _dp, _ip, _ip, _ip, _dp, _dp, _dp,
                                     C.c_int, C.c_double, C.c_int, C.c_double, C.c_int]),
    "adscf_batch_destroy": (C.c_int, [_ctxp]),
    "adscf_batch_last_error": (C.c_char_p, [_ctxp]),
    "adscf_batch_set_fields": (C.c_int, [_ctxp, _dp]),
    "adscf_batch_set_precision": (C.c_int, [_ctxp, C.c_int]),
    "adscf_batch_set_persistent": (C.c_int, [_ctxp, C.c_int, _ip]),
    "adscf_batch_set_plane_passes": (C.c_int, [_ctxp, C.c_int, _ip]),
    "adscf_batch_set_schedule": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int]),
    "adscf_batch_iterate": (C.c_int, [_ctxp, C.c_int, C.c_double, _ip, _dp, _ip]),
    "adscf_batch_energy": (C.c_int, [_ctxp, _dp]),
    "adscf_batch_get": (C.c_int, [_ctxp, C.c_int, _dp, _dp]),
    "adscf_batch_timing": (C.c_int, [_ctxp, _dp, _dp]),
    "adscf_batch_autotune": (C.c_int, [_ctxp, C.c_int, C.c_char_p, C.c_int]),
    "adscf_hipfft_step_time": (C.c_int, [_ctxp, C.c_int, _dp]),
    "adscf_hbm_bandwidth": (C.c_int, [_ctxp, C.c_long, C.c_int, _dp, _dp]),
    "adscf_adjoint": (C.c_int, [_ctxp, C.c_int, _ip, _ip, C.c_double, _dp, C.c_int, _dp, C.c_double, C.c_int,
                                _dp, _dp, _dp, _dp, _dp]),
}


class AdscfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"adscf error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libadscf_hip.so.  Raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m adscf_amd.build` "
                           "(the SCFT path runs only on the GPU)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "expected C-contiguous float64"
    return a.ctypes.data_as(_dp)


def _pc(a):
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.view(np.float64).ctypes.data_as(_dp)


def _pi(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


@dataclass
class ChainResult:
    """Outputs of Propagate_{Diblock,Triblock,Multiblock} (mod_multiblock.f90:19-33 components)."""
    distribution: np.ndarray
    density: np.ndarray
    Qsum: float
    mu: float
    rho: float
    energy: float


@dataclass
class ScfStatus:
    iters: int
    accuracy: float
    converged: bool


class Context:
    """A grid, box and contour step on one GPU (Initialize, sub_wrap.f90:8-40)."""

    def __init__(self, nx, ny, nz, lx, ly, lz, ds, ntypes=2, precision=FP64):
        self._lib = load_library()
        self.nx, self.ny, self.nz, self.ntypes = int(nx), int(ny), int(nz), int(ntypes)
        h = _ctxp()
        rc = self._lib.adscf_init(C.byref(h), self.nx, self.ny, self.nz, self.ntypes, float(lx), float(ly),
                                  float(lz), float(ds), int(precision))
        if rc != OK:
            raise AdscfError(rc, "adscf_init failed (is a gfx950 GPU visible?)")
        self._h = h
        self.ds = float(ds)
        self.box = (float(lx), float(ly), float(lz))
        self.precision = int(precision)

    # -- lifetime --------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.adscf_finalize(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc != OK:
            raise AdscfError(rc, self._lib.adscf_last_error(self._h).decode())

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def kshape(self):
        return (self.nz, self.ny, self.nx // 2 + 1)

    @property
    def ng(self):
        return self.nx * self.ny * self.nz

    def set_box(self, lx, ly, lz):
        self._chk(self._lib.adscf_set_box(self._h, float(lx), float(ly), float(lz)))
        self.box = (float(lx), float(ly), float(lz))

    def set_ds(self, ds):
        self._chk(self._lib.adscf_set_ds(self._h, float(ds)))
        self.ds = float(ds)

    def set_precision(self, precision):
        self._chk(self._lib.adscf_set_precision(self._h, int(precision)))
        self.precision = int(precision)

    def volume(self):
        v, dv = C.c_double(), C.c_double()
        self._chk(self._lib.adscf_volume(self._h, C.byref(v), C.byref(dv)))
        return v.value, dv.value

    # -- Fourier_fftw3 ----------------------------------------------------
    def real2fourier(self, prop):
        prop = _f64(prop, self.shape)
        out = np.empty(self.kshape, dtype=np.complex128)
        self._chk(self._lib.adscf_real2fourier(self._h, _p(prop), _pc(out)))
        return out

    def fourier2real(self, spec):
        spec = np.ascontiguousarray(spec, dtype=np.complex128).reshape(self.kshape)
        out = np.empty(self.shape)
        self._chk(self._lib.adscf_fourier2real(self._h, _pc(spec), _p(out)))
        return out

    def laplace(self):
        lap = np.empty(self.kshape, dtype=np.complex128)
        el = np.empty(self.kshape, dtype=np.complex128)
        nab = np.empty((3,) + self.kshape, dtype=np.complex128)
        self._chk(self._lib.adscf_laplace(self._h, _pc(lap), _pc(el), _pc(nab)))
        return lap, el, nab

    # -- Propagate ------------------------------------------------------
    def expfield(self, field):
        f = _f64(field, (self.ntypes,) + self.shape)
        out = np.empty_like(f)
        self._chk(self._lib.adscf_expfield(self._h, _p(f), _p(out)))
        return out

    def propagate_step(self, prop, expfield_type):
        q = np.array(prop, dtype=np.float64, order="C").reshape(self.shape)
        e = _f64(expfield_type, self.shape)
        self._chk(self._lib.adscf_propagate_step(self._h, _p(q), _p(e)))
        return q

    def propagate_chain_noalloc(self, block_types, block_segments, expfield_dev, distribution_dev, density_dev,
                                scalars_dev, ensemble=1, rho=1.0, mu=0.0):
        """adscf_propagate_chain_noalloc on device arrays given as integer addresses (allocated
        with the HIP runtime the library links): expfield (ntypes,nz,ny,nx) float64 in;
        distribution and density (nblocks,nz,ny,nx) float64 and scalars[4] = Qsum, mu, rho,
        energy out.  The caller's own pending work on these arrays must be complete first."""
        bt = np.ascontiguousarray(block_types, dtype=np.int32)
        bs = np.ascontiguousarray(block_segments, dtype=np.int32)
        self._chk(self._lib.adscf_propagate_chain_noalloc(self._h, len(bt), _pi(bt), _pi(bs), int(ensemble),
                                                          float(rho), float(mu), C.c_void_p(int(expfield_dev)),
                                                          C.c_void_p(int(distribution_dev)),
                                                          C.c_void_p(int(density_dev)), C.c_void_p(int(scalars_dev))))

    def propagate_chain(self, block_types, block_segments, expfield, ensemble=1, rho=1.0, mu=0.0,
                        want_distribution=True, want_density=True):
        bt = np.ascontiguousarray(block_types, dtype=np.int32)
        bs = np.ascontiguousarray(block_segments, dtype=np.int32)
        e = _f64(expfield, (self.ntypes,) + self.shape)
        nb = len(bt)
        dist = np.empty((nb,) + self.shape) if want_distribution else None
        den = np.empty((nb,) + self.shape) if want_density else None
        r, m, q, en = C.c_double(rho), C.c_double(mu), C.c_double(), C.c_double()
        self._chk(self._lib.adscf_propagate_chain(self._h, nb, _pi(bt), _pi(bs), int(ensemble), C.byref(r),
                                                  C.byref(m), _p(e), _p(dist), _p(den), C.byref(q),
                                                  C.byref(en)))
        return ChainResult(dist, den, q.value, m.value, r.value, en.value)

    def propagate_homopolymer(self, monomer_type, segments, expfield, ensemble=1, rho=1.0, mu=0.0):
        """Propagate_Homopolymer (mod_homopolymer.f90:90-141): distribution, density (nz,ny,nx)."""
        e = _f64(expfield, (self.ntypes,) + self.shape)
        dist = np.empty(self.shape)
        den = np.empty(self.shape)
        r, m, q, en = C.c_double(rho), C.c_double(mu), C.c_double(), C.c_double()
        self._chk(self._lib.adscf_propagate_homopolymer(self._h, int(monomer_type), int(segments), int(ensemble),
                                                        C.byref(r), C.byref(m), _p(e), _p(dist), _p(den),
                                                        C.byref(q), C.byref(en)))
        return ChainResult(dist, den, q.value, m.value, r.value, en.value)

    def update_solvent(self, field, ensemble=1, rho=1.0, mu=0.0):
        """Update_Solvent (mod_solvent.f90:85-108) for one field (nz,ny,nx)."""
        w = _f64(field, self.shape)
        dist = np.empty(self.shape)
        den = np.empty(self.shape)
        r, m, q, en = C.c_double(rho), C.c_double(mu), C.c_double(), C.c_double()
        self._chk(self._lib.adscf_update_solvent(self._h, int(ensemble), C.byref(r), C.byref(m), _p(w), _p(dist),
                                                 _p(den), C.byref(q), C.byref(en)))
        return ChainResult(dist, den, q.value, m.value, r.value, en.value)

    # -- Interactions_FloryHuggins ----------------------------------------
    def set_interactions(self, kappa, chi, monomer_volume):
        chi = np.asarray(chi, dtype=np.float64)
        chi_f = np.ascontiguousarray(chi.T)  # column-major
        vol = _f64(monomer_volume)
        self._chk(self._lib.adscf_set_interactions(self._h, float(kappa), _p(chi_f), _p(vol)))

    def mean_fields(self, density):
        d = _f64(density, (self.ntypes,) + self.shape)
        out = np.empty_like(d)
        self._chk(self._lib.adscf_mean_fields(self._h, _p(d), _p(out)))
        return out

    def interaction_energy(self, density):
        d = _f64(density, (self.ntypes,) + self.shape)
        e = C.c_double()
        self._chk(self._lib.adscf_interaction_energy(self._h, _p(d), C.byref(e)))
        return e.value

    def average_volume_fraction(self, density):
        d = _f64(density, (self.ntypes,) + self.shape)
        e = C.c_double()
        self._chk(self._lib.adscf_average_volume_fraction(self._h, _p(d), C.byref(e)))
        return e.value

    # -- Iterate (in place on (m+1, fieldsize) history arrays) ------------
    def mixing_simple(self, field, dfield, lam):
        self._chk(self._lib.adscf_mixing_simple(self._h, _p(field), _p(dfield), field.shape[1],
                                                field.shape[0] - 1, float(lam)))

    def mixing_lambda(self, field, dfield, lammax, it):
        self._chk(self._lib.adscf_mixing_lambda(self._h, _p(field), _p(dfield), field.shape[1],
                                                field.shape[0] - 1, float(lammax), int(it)))

    def mixing_anderson(self, field, dfield, lam, it):
        self._chk(self._lib.adscf_mixing_anderson(self._h, _p(field), _p(dfield), field.shape[1],
                                                  field.shape[0] - 1, float(lam), int(it)))

    def device_lu_solve(self, a, b):
        """The Anderson step's device LU (ludcmp + lubksb) on one (n,n) system, n <= 40;
        returns (x, lu, indx, singular) in the layout of ludcmp() and solve_lu()."""
        n = a.shape[0]
        af = np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)
        x = _f64(b).copy()
        lu = np.empty((n, n))
        indx = np.zeros(n, dtype=np.int32)
        sing = C.c_int()
        self._chk(self._lib.adscf_device_lu_solve(self._h, _p(af), n, _p(x), _p(lu), _pi(indx), C.byref(sing)))
        return x, lu.T.copy(), indx, bool(sing.value)

    # -- device-resident SCF ------------------------------------------------
    def scf_setup(self, block_types, block_segments, ensemble=1, rho=1.0, mu=0.0, mixing_dim=10, lam_mix=0.5):
        bt = np.ascontiguousarray(block_types, dtype=np.int32)
        bs = np.ascontiguousarray(block_segments, dtype=np.int32)
        self._chk(self._lib.adscf_scf_setup(self._h, len(bt), _pi(bt), _pi(bs), int(ensemble), float(rho),
                                            float(mu), int(mixing_dim), float(lam_mix)))

    def scf_set_field(self, field):
        f = _f64(field, (self.ntypes,) + self.shape)
        self._chk(self._lib.adscf_scf_set_field(self._h, _p(f)))

    def scf_iterate(self, niter, tol=0.0):
        it, acc, conv = C.c_int(), C.c_double(), C.c_int()
        self._chk(self._lib.adscf_scf_iterate(self._h, int(niter), float(tol), C.byref(it), C.byref(acc),
                                              C.byref(conv)))
        return ScfStatus(it.value, acc.value, bool(conv.value))

    def scf_evaluate(self):
        self._chk(self._lib.adscf_scf_evaluate(self._h))

    def scf_get(self):
        f = np.empty((self.ntypes,) + self.shape)
        d = np.empty((self.ntypes,) + self.shape)
        self._chk(self._lib.adscf_scf_get(self._h, _p(f), _p(d)))
        return f, d

    def scf_energy(self):
        F, Fc, q, mu, rho = (C.c_double() for _ in range(5))
        self._chk(self._lib.adscf_scf_energy(self._h, C.byref(F), C.byref(Fc), C.byref(q), C.byref(mu),
                                             C.byref(rho)))
        return dict(F=F.value, F_per_chain=Fc.value, Qsum=q.value, mu=mu.value, rho=rho.value)

    def scf_residuals(self, cap=200000):
        out = np.empty(cap)
        n = C.c_int()
        self._chk(self._lib.adscf_scf_residuals(self._h, _p(out), int(cap), C.byref(n)))
        return out[:min(n.value, cap)].copy()

    def scf_timing(self):
        a, b = C.c_double(), C.c_double()
        self._chk(self._lib.adscf_scf_timing(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def scf_set_residual(self, kind=0, measure=0, target=None):
        """Residual of scf_iterate: kind 0 mean fields - field, kind 1 densities - target;
        measure 0 max|dfield|, 1 sum|dfield| dvol, 2 sum of squared density changes."""
        t = None if target is None else _f64(target, (self.ntypes,) + self.shape)
        self._chk(self._lib.adscf_scf_set_residual(self._h, int(kind), int(measure), None if t is None else _p(t)))

    def scf_set_density(self, density):
        """Densities the next evaluation's measure 2 (sum of squared changes) compares with."""
        d = _f64(density, (self.ntypes,) + self.shape)
        self._chk(self._lib.adscf_scf_set_density(self._h, _p(d)))

    def scf_set_mixing(self, mixing_type=3):
        """Mixing of scf_iterate: 1 simple, 2 lambda (lammax = lam_mix), 3 Anderson."""
        self._chk(self._lib.adscf_scf_set_mixing(self._h, int(mixing_type)))

    def set_contour_rule(self, rule="trapezoid"):
        """Contour rule of every density evaluation: "trapezoid" (the reference's, default) or
        "simpson" (composite Simpson per block, a 3/8 panel closing odd blocks)."""
        code = {"trapezoid": 0, "simpson": 1}.get(rule, rule)
        self._chk(self._lib.adscf_set_contour_rule(self._h, int(code)))

    # -- conjugate fields and DDFT (main_multiblock.f90) -------------------------
    def conjugate_fields(self, density=None, thresh=5e-4, maxit=2000, measure=0):
        """Get_Conjugate_Fields (main_multiblock.f90:566-640) from the current field; density None
        uses the device-resident DDFT density.  Returns (iterations, accuracy)."""
        d = None if density is None else _f64(density, (self.ntypes,) + self.shape)
        it, acc = C.c_int(), C.c_double()
        self._chk(self._lib.adscf_conjugate_fields(self._h, None if d is None else _p(d), float(thresh), int(maxit),
                                                   int(measure), C.byref(it), C.byref(acc)))
        return it.value, acc.value

    def mobility_qq(self, table, dq, wcutoff):
        """Read_Mobilities' table part (main_multiblock.f90:486-504).  table: (nt, nt, nq) with
        table[t2, t1, iq-1] = mobility_table(iq, t1, t2).  Uploads the result for ddft_step and
        returns it as (nt, nt, nz, ny, nx/2+1)."""
        t = np.ascontiguousarray(table, dtype=np.float64)
        nt, nq = t.shape[0], t.shape[2]
        if t.shape != (self.ntypes, self.ntypes, nq):
            raise ValueError("table must be (ntypes, ntypes, nq)")
        out = np.empty((nt, nt) + self.kshape)
        self._chk(self._lib.adscf_mobility_qq(self._h, int(nq), float(dq), _p(t), float(wcutoff), _p(out)))
        return out

    def ddft_set_mobility(self, mob_qq):
        m = _f64(mob_qq, (self.ntypes, self.ntypes) + self.kshape)
        self._chk(self._lib.adscf_ddft_set_mobility(self._h, _p(m)))

    def ddft_set_density(self, density):
        d = _f64(density, (self.ntypes,) + self.shape)
        self._chk(self._lib.adscf_ddft_set_density(self._h, _p(d)))

    def ddft_get_density(self):
        d = np.empty((self.ntypes,) + self.shape)
        self._chk(self._lib.adscf_ddft_get_density(self._h, _p(d)))
        return d

    def ddft_step(self, timestep, qdmin=True, accuracy_scf=1e-3, itermax_scf=2000):
        """DDFT_Timestep_Euler (main_multiblock.f90:794-880) on the device-resident density and
        field.  Returns dict(gcfstat=[s1, s2], iterations=[i1, i2])."""
        st = np.zeros(2, dtype=np.int32)
        its = np.zeros(2, dtype=np.int32)
        self._chk(self._lib.adscf_ddft_step(self._h, float(timestep), int(bool(qdmin)), float(accuracy_scf),
                                            int(itermax_scf), _pi(st), _pi(its)))
        return dict(gcfstat=st.tolist(), iterations=its.tolist())

    # -- homopolymer blend and EPD (main_blend.f90) ------------------------------
    def blend_setup(self, monomer_types, segments, rho, mixing_dim=2, lam_mix=0.5):
        """The device-resident loop over a blend of homopolymers (canonical, one per entry)."""
        t = np.ascontiguousarray(monomer_types, dtype=np.int32)
        sg = np.ascontiguousarray(segments, dtype=np.int32)
        r = np.ascontiguousarray(rho, dtype=np.float64)
        self._npol = len(t)
        self._chk(self._lib.adscf_blend_setup(self._h, len(t), _pi(t), _pi(sg), _p(r), int(mixing_dim),
                                              float(lam_mix)))

    def blend_scalars(self):
        n = self._npol
        q, mu, e = np.empty(n), np.empty(n), np.empty(n)
        self._chk(self._lib.adscf_blend_scalars(self._h, _p(q), _p(mu), _p(e)))
        return dict(Qsum=q, mu=mu, energy=e)

    def epd_step(self, timestep, accuracy_scf=float(np.float32(0.01)), itermax=100000):
        """Chain_Dynamics_Timestep_Euler (main_blend.f90:605-667); returns the conjugate
        solve's loop index."""
        it = C.c_int()
        self._chk(self._lib.adscf_epd_step(self._h, float(timestep), float(accuracy_scf), int(itermax),
                                           C.byref(it)))
        return it.value

    def field_derivative(self):
        """Calculate_Field_Derivative (main_blend.f90:709-740) at the device-resident field:
        -laplace * (mean fields of the densities - field), (ntypes, nz, ny, nx)."""
        out = np.empty((self.ntypes,) + self.shape)
        self._chk(self._lib.adscf_field_derivative(self._h, _p(out)))
        return out

    def epd_field_step(self, timestep, method="euler"):
        """EPD_Timestep_Euler (main_blend.f90:580-596) or, method "rk4",
        EPD_Timestep_RungeKutta (:678-701) on the device-resident field."""
        code = {"euler": 0, "rk4": 1}.get(method, method)
        self._chk(self._lib.adscf_epd_field_step(self._h, float(timestep), int(code)))

    def ddft_energy(self):
        F, mu, rho = C.c_double(), C.c_double(), C.c_double()
        self._chk(self._lib.adscf_ddft_energy(self._h, C.byref(F), C.byref(mu), C.byref(rho)))
        return dict(F=F.value, mu=mu.value, rho=rho.value)


    def hipfft_step_us(self, reps=200):
        """Time of one single-chain step built from hipFFT r2c/c2r (comparator only)."""
        v = C.c_double()
        self._chk(self._lib.adscf_hipfft_step_time(self._h, int(reps), C.byref(v)))
        return v.value

    def hbm_bandwidth(self, nbytes=1 << 30, reps=7):
        """Measured (copy GB/s counting read + write, read GB/s) of the device."""
        cp, rd = C.c_double(), C.c_double()
        self._chk(self._lib.adscf_hbm_bandwidth(self._h, int(nbytes), int(reps), C.byref(cp), C.byref(rd)))
        return cp.value, rd.value

    # -- schedules ------------------------------------------------------------
    def autotune(self, reps=3):
        buf = C.create_string_buffer(1 << 16)
        self._chk(self._lib.adscf_autotune(self._h, int(reps), buf, len(buf)))
        return buf.value.decode()

    def schedule(self, axis):
        n, a, b, w = (C.c_int() for _ in range(4))
        self._chk(self._lib.adscf_get_schedule(self._h, int(axis), C.byref(n), C.byref(a), C.byref(b), C.byref(w)))
        return dict(n=n.value, N1=a.value, N2=b.value, waves=w.value)

    def set_middle_axis(self, axis=-1):
        """Axis with one line pass per contour step (-1: the default, y)."""
        self._chk(self._lib.adscf_set_middle_axis(self._h, int(axis)))

    def middle_axis(self):
        a = C.c_int()
        self._chk(self._lib.adscf_get_middle_axis(self._h, C.byref(a)))
        return a.value

    def set_schedule(self, axis, n1, n2, waves=1):
        self._chk(self._lib.adscf_set_schedule(self._h, int(axis), int(n1), int(n2), int(waves)))

    def set_pass_tiles(self, axis, tiles):
        """Tiles per workgroup (1, 2, 4) of the sweep passes along an axis."""
        self._chk(self._lib.adscf_set_pass_tiles(self._h, int(axis), int(tiles)))

    def pass_tiles(self, axis):
        t = np.zeros(1, dtype=np.int32)
        self._chk(self._lib.adscf_get_pass_tiles(self._h, int(axis), _pi(t)))
        return int(t[0])

    def set_plane_passes(self, mode=-1, n1=0, n2=0):
        """xy-plane passes of the 3-D sweep: -1 where the grid allows, 0 never, 1 required;
        (n1, n2) the engine shape of the plane operators (n1 = 0 keeps the current one)."""
        self._chk(self._lib.adscf_set_plane_passes(self._h, int(mode), int(n1), int(n2)))

    def plane_passes(self):
        """dict(active, N1, N2): whether the sweep of the context's precision uses plane passes."""
        t = np.zeros(3, dtype=np.int32)
        self._chk(self._lib.adscf_get_plane_passes(self._h, _pi(t[0:]), _pi(t[1:]), _pi(t[2:])))
        return dict(active=bool(t[0]), N1=int(t[1]), N2=int(t[2]))

    # -- adjoint ----------------------------------------------------------------
    def adjoint(self, block_types, block_segments, field, phi_target, lam=0.1, target_type=1, rho=1.0,
                checkpoint_every=0, gradients=True):
        """Loss L = sum (phi_target_type - phi_target)^2 + lam*F of one canonical evaluation and,
        with gradients=True, dL/dw (ntypes, nz, ny, nx) and dL/d(lx, ly, lz).
        Returns dict(loss, F, phi, grad_w, grad_box)."""
        bt = np.ascontiguousarray(block_types, dtype=np.int32)
        bs = np.ascontiguousarray(block_segments, dtype=np.int32)
        f = _f64(field, (self.ntypes,) + self.shape)
        pt = _f64(phi_target, self.shape)
        loss, F = C.c_double(), C.c_double()
        phi = np.empty((self.ntypes,) + self.shape)
        gw = np.empty_like(phi) if gradients else None
        gb = np.empty(3) if gradients else None
        self._chk(self._lib.adscf_adjoint(self._h, len(bt), _pi(bt), _pi(bs), float(rho), _p(f), int(target_type),
                                          _p(pt), float(lam), int(checkpoint_every), C.byref(loss), C.byref(F),
                                          _p(gw), _p(gb), _p(phi)))
        return dict(loss=loss.value, F=F.value, phi=phi, grad_w=gw, grad_box=gb)


class Batch:
    """Independent two-type SCF instances run together (adscf_batch_*).

    instances: list of dicts with keys grid (nx,ny,nz), box (lx,ly,lz), ds, blocks (types),
    segments, kappa, chi (2x2), vol (2)."""

    def __init__(self, instances, ensemble=1, rho=1.0, mixing_dim=20, lam_mix=0.5, precision=FP32):
        self._lib = load_library()
        self.instances = instances
        n = len(instances)
        grid = np.ascontiguousarray([v for d in instances for v in d["grid"]], dtype=np.int32)
        box = _f64([v for d in instances for v in d["box"]])
        ds = _f64([d["ds"] for d in instances])
        nb = np.ascontiguousarray([len(d["blocks"]) for d in instances], dtype=np.int32)
        bt = np.ascontiguousarray([v for d in instances for v in d["blocks"]], dtype=np.int32)
        bs = np.ascontiguousarray([v for d in instances for v in d["segments"]], dtype=np.int32)
        kappa = _f64([d["kappa"] for d in instances])
        chi = _f64([v for d in instances for v in np.asarray(d["chi"], dtype=np.float64).T.ravel()])
        vol = _f64([v for d in instances for v in d["vol"]])
        h = _ctxp()
        rc = self._lib.adscf_batch_create(C.byref(h), n, _pi(grid), _p(box), _p(ds), _pi(nb), _pi(bt), _pi(bs),
                                          _p(kappa), _p(chi), _p(vol), int(ensemble), float(rho), int(mixing_dim),
                                          float(lam_mix), int(precision))
        if rc != OK:
            raise AdscfError(rc, "adscf_batch_create failed")
        self._h = h
        self.n = n

    def close(self):
        if getattr(self, "_h", None):
            self._lib.adscf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc != OK:
            raise AdscfError(rc, self._lib.adscf_batch_last_error(self._h).decode())

    def shape(self, i):
        nx, ny, nz = self.instances[i]["grid"]
        return (2, nz, ny, nx)

    def set_fields(self, fields):
        flat = np.concatenate([_f64(f, self.shape(i)).ravel() for i, f in enumerate(fields)])
        self._chk(self._lib.adscf_batch_set_fields(self._h, _p(flat)))

    def set_precision(self, precision):
        self._chk(self._lib.adscf_batch_set_precision(self._h, int(precision)))

    def set_persistent(self, enable=True):
        """One persistent launch for the whole sweep (True) or one launch per pass (False);
        returns whether the persistent sweep is active for the current schedules."""
        act = np.zeros(1, dtype=np.int32)
        self._chk(self._lib.adscf_batch_set_persistent(self._h, int(bool(enable)), _pi(act)))
        return bool(act[0])

    def set_plane_passes(self, mode=-1):
        """xy-plane passes of the instances' sweeps: -1 where the grid allows, 0 never, 1 on
        every instance; returns the number of instances whose sweep uses them."""
        act = np.zeros(1, dtype=np.int32)
        self._chk(self._lib.adscf_batch_set_plane_passes(self._h, int(mode), _pi(act)))
        return int(act[0])

    def set_schedule(self, n, n1, n2):
        """Register shape n1 x n2 of every instance axis of length n (the batched line
        passes run four wavefronts of lines per workgroup)."""
        self._chk(self._lib.adscf_batch_set_schedule(self._h, int(n), int(n1), int(n2)))

    def iterate(self, niter, tol=0.0):
        it = np.zeros(self.n, dtype=np.int32)
        acc = np.zeros(self.n)
        conv = np.zeros(self.n, dtype=np.int32)
        self._chk(self._lib.adscf_batch_iterate(self._h, int(niter), float(tol), _pi(it), _p(acc), _pi(conv)))
        return it, acc, conv.astype(bool)

    def energy(self):
        f = np.zeros(self.n)
        self._chk(self._lib.adscf_batch_energy(self._h, _p(f)))
        return f

    def get(self, i):
        f = np.empty(self.shape(i))
        d = np.empty(self.shape(i))
        self._chk(self._lib.adscf_batch_get(self._h, int(i), _p(f), _p(d)))
        return f, d

    def timing(self):
        a, b = C.c_double(), C.c_double()
        self._chk(self._lib.adscf_batch_timing(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, reps=3):
        buf = C.create_string_buffer(1 << 16)
        self._chk(self._lib.adscf_batch_autotune(self._h, int(reps), buf, len(buf)))
        return buf.value.decode()


# ---- host routines of module Iterate (no GPU needed) -----------------------

def ludcmp(a):
    """ludcmp (mod_iterate.f90:232-298) on a numpy (n,n) matrix; returns (lu, indx, d)."""
    lib = load_library()
    n = a.shape[0]
    af = np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)
    indx = np.zeros(n, dtype=np.int32)
    d = C.c_double()
    rc = lib.adscf_ludcmp(_p(af), n, _pi(indx), C.byref(d))
    if rc != OK:
        raise AdscfError(rc, "ludcmp")
    return af.T.copy(), indx, d.value


def solve_lu(a, b):
    lib = load_library()
    n = a.shape[0]
    af = np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)
    x = np.empty(n)
    rc = lib.adscf_solve_lu(_p(af), _p(_f64(b)), _p(x), n)
    if rc != OK:
        raise AdscfError(rc, "solve_lu")
    return x


def solve_cholesky(a, b):
    lib = load_library()
    n = a.shape[0]
    af = np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)
    x = np.empty(n)
    rc = lib.adscf_solve_cholesky(_p(af), _p(_f64(b)), _p(x), n)
    if rc != OK:
        raise AdscfError(rc, "solve_cholesky")
    return x


# ---- reference-named aliases -------------------------------------------------

@dataclass
class Copolymer:
    """Mirror of diblock_/triblock_/multiblock_copolymer (mod_multiblock.f90:19-33)."""
    label: int
    monomer_type: list
    block_length: list
    ensemble: int
    rho: float
    mu: float
    block_segments: list = dc_field(default_factory=list)
    segments: int = 0
    energy: float = 0.0
    Qsum: float = 0.0
    distribution: np.ndarray = None
    density: np.ndarray = None


def _nint(x):
    # Fortran NINT: round half away from zero (mod_multiblock.f90:85)
    return int(np.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def Initialize_Multiblock(ctx: Context, label, monomer_type, block_length, ensemble, rho, mu):
    """Initialize_Multiblock (mod_multiblock.f90:48-99); also stands for _Diblock / _Triblock."""
    if ensemble not in (1, 2):
        print("Invalid Ensemble")  # mod_multiblock.f90:75
    segs = [_nint(b / ctx.ds) for b in block_length]
    return Copolymer(label, list(monomer_type), list(block_length), ensemble, rho, mu, segs, sum(segs))


Initialize_Diblock = Initialize_Multiblock
Initialize_Triblock = Initialize_Multiblock


def Calculate_Expfield(ctx: Context, field):
    """Calculate_Expfield (mod_propagate.f90:22-32)."""
    return ctx.expfield(field)


def Propagate_Step(ctx: Context, prop, expfield_type):
    """Propagate_Step (mod_propagate.f90:40-68); returns the new prop."""
    return ctx.propagate_step(prop, expfield_type)


def Propagate_Multiblock(ctx: Context, copolymer: Copolymer, expfield):
    """Propagate_Multiblock (mod_multiblock.f90:112-188): fills Qsum, mu/rho, energy,
    distribution and density of the copolymer.  An invalid ensemble raises AdscfError
    (the reference prints and stops, :181-183)."""
    r = ctx.propagate_chain(copolymer.monomer_type, copolymer.block_segments, expfield, copolymer.ensemble,
                            copolymer.rho, copolymer.mu)
    copolymer.Qsum, copolymer.mu, copolymer.rho, copolymer.energy = r.Qsum, r.mu, r.rho, r.energy
    copolymer.distribution, copolymer.density = r.distribution, r.density
    return copolymer


Propagate_Diblock = Propagate_Multiblock
Propagate_Triblock = Propagate_Multiblock


@dataclass
class Homopolymer:
    """Mirror of type homopolymer (mod_homopolymer.f90:17-29)."""
    label: int
    monomer_type: int
    length: float
    ensemble: int
    rho: float
    mu: float
    segments: int = 0
    energy: float = 0.0
    Qsum: float = 0.0
    distribution: np.ndarray = None
    density: np.ndarray = None


def Initialize_Homopolymer(ctx: Context, label, monomer_type, length, ensemble, rho, mu):
    """Initialize_Homopolymer (mod_homopolymer.f90:41-80)."""
    if ensemble not in (1, 2):
        print("Invalid Ensemble")  # mod_homopolymer.f90:58
    return Homopolymer(label, monomer_type, length, ensemble, rho, mu, _nint(length / ctx.ds))


def Propagate_Homopolymer(ctx: Context, polymer: Homopolymer, expfield):
    """Propagate_Homopolymer (mod_homopolymer.f90:90-141)."""
    r = ctx.propagate_homopolymer(polymer.monomer_type, polymer.segments, expfield, polymer.ensemble, polymer.rho,
                                  polymer.mu)
    polymer.Qsum, polymer.mu, polymer.rho, polymer.energy = r.Qsum, r.mu, r.rho, r.energy
    polymer.distribution, polymer.density = r.distribution, r.density
    return polymer


@dataclass
class MonomericSolvent:
    """Mirror of type monomeric_solvent (mod_solvent.f90:17-27)."""
    label: int
    monomer_type: int
    ensemble: int
    rho: float
    mu: float
    energy: float = 0.0
    Qsum: float = 0.0
    distribution: np.ndarray = None
    density: np.ndarray = None


def Update_Solvent(ctx: Context, solvent: MonomericSolvent, field):
    """Update_Solvent (mod_solvent.f90:85-108); field is the solvent's own field (nz,ny,nx)."""
    r = ctx.update_solvent(field, solvent.ensemble, solvent.rho, solvent.mu)
    solvent.Qsum, solvent.mu, solvent.rho, solvent.energy = r.Qsum, r.mu, r.rho, r.energy
    solvent.distribution, solvent.density = r.distribution, r.density
    return solvent


def Calculate_Mean_Fields(ctx: Context, density):
    """Calculate_Mean_Fields (mod_FloryHuggins.f90:104-129)."""
    return ctx.mean_fields(density)


def Interaction_Energy(ctx: Context, density):
    """Interaction_Energy (mod_FloryHuggins.f90:74-96)."""
    return ctx.interaction_energy(density)


def Mixing_Anderson(ctx: Context, field, dfield, lam_mix, it):
    """Mixing_Anderson (mod_iterate.f90:92-150), in place on (m+1, fieldsize) arrays."""
    ctx.mixing_anderson(field, dfield, lam_mix, it)
