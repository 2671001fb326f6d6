"""GPU: adscf_propagate_chain_noalloc (device arrays, every work array in the library's device
scratch, no allocation) against adscf_propagate_chain on the same inputs: bit for bit, on two
calls in a row, the second with a smaller chain, in FP64 and FP32-mixed and both ensembles;
a chain larger than the scratch is refused before any work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, L, DS = 32, 4.0, 0.01


class DevArray:
    """A float64 device array of the HIP runtime the library itself links (not torch's own
    runtime copy), allocated and copied by the test; the entry point under test allocates
    nothing."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import adscf_amd

            adscf_amd.load_library()  # the library's runtime, found through its RUNPATH
            cls._hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        return cls._hip

    def __init__(self, shape, fill=None):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.ptr = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        if fill is not None:
            self.upload(np.full(self.shape, fill, dtype=np.float64))

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes == self.nbytes
        assert self.hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1) == 0

    def download(self):
        a = np.empty(self.shape, dtype=np.float64)
        assert self.hip().hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) == 0
        return a

    @property
    def addr(self):
        return self.ptr.value

    def free(self):
        if self.ptr.value:
            self.hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_propagate_chain(prec):
    import adscf_amd as amd

    P = amd.FP64 if prec == "fp64" else amd.FP32
    rng = np.random.default_rng(11)
    w = 3.0 * rng.standard_normal((2, N, N, N))
    E = np.exp(-0.5 * DS * w)
    calls = [([1, 2], [50, 50], 1, 1.0, 0.0),      # config-1 chain, canonical
             ([2, 1, 2], [10, 15, 5], 2, 1.0, 0.4)]  # a smaller chain, grand canonical
    with amd.Context(N, N, N, L, L, L, DS, precision=P) as ctx:
        for bt, bs, ens, rho, mu in calls:
            print(f"noalloc {prec}: blocks {bt} segments {bs} ensemble {ens}", flush=True)
            ref = ctx.propagate_chain(bt, bs, E, ensemble=ens, rho=rho, mu=mu)
            e_d = DevArray(E.shape)
            e_d.upload(E)
            dist_d = DevArray((len(bt), N, N, N), fill=np.nan)
            dens_d = DevArray((len(bt), N, N, N), fill=np.nan)
            sc_d = DevArray((4,), fill=np.nan)
            ctx.propagate_chain_noalloc(bt, bs, e_d.addr, dist_d.addr, dens_d.addr, sc_d.addr, ensemble=ens,
                                        rho=rho, mu=mu)
            dist, dens, sc = dist_d.download(), dens_d.download(), sc_d.download()
            for a in (e_d, dist_d, dens_d, sc_d):
                a.free()
            assert np.array_equal(dist, ref.distribution)
            assert np.array_equal(dens, ref.density)
            assert sc[0] == ref.Qsum and sc[1] == ref.mu and sc[2] == ref.rho and sc[3] == ref.energy


def test_noalloc_refuses_a_chain_beyond_the_scratch():
    import adscf_amd as amd

    n = 64
    with amd.Context(n, n, n, 4.5, 4.5, 4.5, 0.005, precision=amd.FP64) as ctx:
        e_d = DevArray((2, n, n, n), fill=1.0)
        out = DevArray((2, n, n, n))
        sc = DevArray((4,))
        with pytest.raises(amd.AdscfError, match="scratch"):
            ctx.propagate_chain_noalloc([1, 2], [60, 140], e_d.addr, out.addr, out.addr, sc.addr)
        for a in (e_d, out, sc):
            a.free()
