"""GPU checks of Propagate_Homopolymer (mod_homopolymer.f90:90-141) and Update_Solvent
(mod_solvent.f90:85-108) through the C-ABI against the oracle, both ensembles, FP64 at
the plain 1e-10 bar and the FP32-mixed homopolymer at its stated 2e-5."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(shape, seed, amp=1.0):
    return amp * np.random.default_rng(seed).standard_normal((2,) + shape)


@pytest.mark.parametrize("dims,seg", [((16, 12, 10), 23), ((64, 16, 1), 10), ((86, 1, 75), 7), ((32, 32, 32), 40)])
@pytest.mark.parametrize("ens", [1, 2])
def test_homopolymer_fp64(orc, dims, seg, ens):
    import adscf_amd

    L = (3.0, 2.5, 2.0)
    g = orc.Grid(*dims, *L, 0.02)
    E = orc.expfield(g, _fields(g.shape, sum(dims) + seg))
    print(f"homopolymer dims={dims} segments={seg} ensemble={ens}", flush=True)
    ref = orc.propagate_homopolymer(g, 2, seg, E, ensemble=ens, rho=0.7, mu=0.3)
    with adscf_amd.Context(*dims, *L, 0.02) as ctx:
        got = ctx.propagate_homopolymer(2, seg, E, ensemble=ens, rho=0.7, mu=0.3)
    scale = np.max(np.abs(ref.density))
    assert np.max(np.abs(got.density - ref.density)) < 1e-10 * scale
    assert np.max(np.abs(got.distribution - ref.distribution)) < 1e-10 * np.max(ref.distribution)
    for k in ("Qsum", "mu", "rho", "energy"):
        assert abs(getattr(got, k) - getattr(ref, k)) < 1e-10 * max(1.0, abs(getattr(ref, k)))


def test_homopolymer_fp32_error_bound(orc):
    import adscf_amd

    dims, L, seg = (32, 32, 32), (4.0, 4.0, 4.0), 100
    g = orc.Grid(*dims, *L, 0.01)
    E = orc.expfield(g, _fields(g.shape, 3))
    ref = orc.propagate_homopolymer(g, 1, seg, E)
    with adscf_amd.Context(*dims, *L, 0.01, precision=adscf_amd.FP32) as ctx:
        got = ctx.propagate_homopolymer(1, seg, E)
    assert abs(got.Qsum - ref.Qsum) < 2e-5 * ref.Qsum
    assert np.max(np.abs(got.density - ref.density)) < 2e-5 * np.max(ref.density)


@pytest.mark.parametrize("ens", [1, 2])
def test_solvent(orc, ens):
    import adscf_amd

    dims, L = (24, 20, 18), (3.0, 2.5, 2.0)
    g = orc.Grid(*dims, *L, 0.05)
    w = _fields(g.shape, 9)[0]
    ref = orc.update_solvent(g, w, ensemble=ens, rho=0.4, mu=-0.2)
    with adscf_amd.Context(*dims, *L, 0.05) as ctx:
        got = ctx.update_solvent(w, ensemble=ens, rho=0.4, mu=-0.2)
    assert np.max(np.abs(got.distribution - ref.distribution)) < 1e-13 * np.max(ref.distribution)
    assert np.max(np.abs(got.density - ref.density)) < 1e-12 * np.max(ref.density)
    for k in ("Qsum", "mu", "rho", "energy"):
        assert abs(getattr(got, k) - getattr(ref, k)) < 1e-12 * max(1.0, abs(getattr(ref, k)))


def test_homopolymer_and_solvent_invalid_ensemble(orc):
    import adscf_amd

    with adscf_amd.Context(8, 8, 8, 2.0, 2.0, 2.0, 0.1) as ctx:
        E = np.ones((2, 8, 8, 8))
        with pytest.raises(adscf_amd.AdscfError) as e1:
            ctx.propagate_homopolymer(1, 5, E, ensemble=3)
        assert e1.value.code == adscf_amd.ERR_ENSEMBLE
        with pytest.raises(adscf_amd.AdscfError) as e2:
            ctx.update_solvent(np.zeros((8, 8, 8)), ensemble=0)
        assert e2.value.code == adscf_amd.ERR_ENSEMBLE
