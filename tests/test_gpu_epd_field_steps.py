"""GPU: the device-resident EPD field steps of Main_Blend (adscf_field_derivative,
adscf_epd_field_step) against the oracle's Calculate_Field_Derivative, EPD_Timestep_Euler and
EPD_Timestep_RungeKutta (main_blend.f90:580-596, 678-701, 709-740; the oracle's versions are
pinned in test_oracle_epd_steps.py).  The shipped driver calls Chain_Dynamics_Timestep_Euler
instead (adscf_epd_step, test_gpu_blend.py); these are its two alternatives.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS, BOX, DS = (32, 8, 1), (6.4, 1.6, 1.0), 0.1
TYPES, SEGS, RHO = [1, 2], [10, 10], [0.5, 0.5]
KAPPA, CHI = 20.0, [[0.0, 2.0], [2.0, 0.0]]


def _field():
    nx = DIMS[0]
    x = np.arange(nx) / nx
    rng = np.random.default_rng(1)
    shape = DIMS[::-1]
    w = np.zeros((2,) + shape)
    w[0] += 0.3 * np.cos(2 * np.pi * x)[None, None, :] + 0.05 * rng.standard_normal(shape)
    w[1] -= 0.3 * np.cos(2 * np.pi * x)[None, None, :]
    return w


def _ctx(amd, prec):
    ctx = amd.Context(*DIMS, *BOX, DS, precision=prec)
    ctx.set_interactions(KAPPA, CHI, [1.0, 1.0])
    ctx.blend_setup(TYPES, SEGS, RHO, mixing_dim=2, lam_mix=0.5)
    return ctx


def _oracle(orc):
    return orc.Grid(*DIMS, *BOX, DS), orc.Blend(TYPES, SEGS, RHO), orc.FH(KAPPA, CHI, [1.0, 1.0])


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_field_derivative_matches_oracle(orc, prec):
    import adscf_amd as amd

    g, bl, fh = _oracle(orc)
    w = _field()
    ref = orc.field_derivative(g, bl, fh, w)
    with _ctx(amd, amd.FP64 if prec == "fp64" else amd.FP32) as ctx:
        ctx.scf_set_field(w)
        got = ctx.field_derivative()
        f_after, _ = ctx.scf_get()
    err = _rel(got, ref)
    print(f"{prec}: field derivative {err:.2e}")
    assert err < (1e-10 if prec == "fp64" else 1e-4)
    assert np.array_equal(f_after, w)  # the device-resident field is left as it was


@pytest.mark.parametrize("method", ["euler", "rk4"])
def test_epd_field_steps_match_oracle(orc, method):
    import adscf_amd as amd

    g, bl, fh = _oracle(orc)
    w = _field()
    dt, nsteps = 1e-4, 3
    ref = w.copy()
    with _ctx(amd, amd.FP64) as ctx:
        ctx.scf_set_field(w)
        for k in range(nsteps):
            ref = orc.epd_timestep(g, bl, fh, ref, dt, method)
            ctx.epd_field_step(dt, method)
            got, _ = ctx.scf_get()
            err = _rel(got - w, ref - w)
            print(f"{method} step {k + 1}: field change {err:.2e}")
            assert err < 1e-10


def test_epd_step_restarts_the_scf_state(orc):
    """After a field step the context iterates as a fresh context started from the new field."""
    import adscf_amd as amd

    with _ctx(amd, amd.FP64) as a:
        a.scf_set_field(_field())
        a.scf_iterate(3, tol=-1.0)
        a.epd_field_step(2e-4, "rk4")
        w1, _ = a.scf_get()
        s1 = a.scf_iterate(4, tol=-1.0)
        fa, pa = a.scf_get()
    with _ctx(amd, amd.FP64) as b:
        b.scf_set_field(w1)
        s2 = b.scf_iterate(4, tol=-1.0)
        fb, pb = b.scf_get()
    assert s1.iters == s2.iters == 4
    assert np.array_equal(fa, fb) and np.array_equal(pa, pb)


def test_field_derivative_of_a_chain(orc):
    """On a single-chain context the derivative is -laplacian(mean fields - field) of the
    chain's species densities (the oracle's densities and mean fields, numpy's spectral
    Laplacian)."""
    import adscf_amd as amd

    dims, box, ds, segs = (16, 12, 10), (3.2, 2.8, 3.0), 0.02, [15, 20]
    w = 2.0 * np.random.default_rng(4).standard_normal((2,) + dims[::-1])
    g = orc.Grid(*dims, *box, ds)
    fh = orc.FH(25.0, [[0.0, 12.0], [12.0, 0.0]], [1.0, 1.0])
    den, _ = orc.species_densities(g, [1, 2], segs, w)
    mu = orc.mean_fields(g, fh, den) - w
    nz, ny, nx = mu.shape[1:]
    kx = 2 * np.pi * np.fft.rfftfreq(nx, d=box[0] / nx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=box[1] / ny)
    kz = 2 * np.pi * np.fft.fftfreq(nz, d=box[2] / nz)
    k2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
    ref = np.stack([np.fft.irfftn(k2 * np.fft.rfftn(mu[t]), s=mu[t].shape, axes=(0, 1, 2)) for t in range(2)])
    with amd.Context(*dims, *box, ds, precision=amd.FP64) as ctx:
        ctx.set_interactions(25.0, [[0.0, 12.0], [12.0, 0.0]], [1.0, 1.0])
        ctx.scf_setup([1, 2], segs, mixing_dim=4, lam_mix=0.5)
        ctx.scf_set_field(w)
        got = ctx.field_derivative()
    assert _rel(got, ref) < 1e-10


def test_epd_field_step_errors():
    import adscf_amd as amd

    with _ctx(amd, amd.FP64) as ctx:
        with pytest.raises(amd.AdscfError):
            ctx.epd_field_step(1e-4, "euler")  # no field yet
        ctx.scf_set_field(_field())
        with pytest.raises(amd.AdscfError):
            ctx.epd_field_step(1e-4, 2)
