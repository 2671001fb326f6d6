"""GPU: switching the loop of adscf_scf_iterate on one context.

A loop is its mixing type (1..3), residual kind (0..1) and accuracy measure (0..2).  Each
loop is captured as its own iteration graph; a switch starts a new loop on the current
field with the count at 1 (a new Find_SCF_Saddle_Point or Get_Conjugate_Fields loop of the
reference, main_multiblock.f90:566-700).  The switched context must therefore compute what
a fresh context started from the same field computes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX, NY, NZ, BOX, DS = 16, 12, 10, (3.2, 2.8, 3.0), 0.02
SEGS = [15, 35]


def _ctx(amd, m=6, lam=0.2):
    ctx = amd.Context(NX, NY, NZ, *BOX, DS, precision=amd.FP64)
    ctx.set_interactions(25.0, [[0.0, 18.0], [18.0, 0.0]], [1.0, 1.0])
    ctx.scf_setup([1, 2], SEGS, ensemble=1, rho=1.0, mixing_dim=m, lam_mix=lam)
    return ctx


def _field(seed=5):
    rng = np.random.default_rng(seed)
    w = 4.0 * rng.standard_normal((2, NZ, NY, NX))
    return w


@pytest.mark.parametrize("first,second", [
    # (mixing, kind, measure): pairs whose loop codes collided in an earlier graph key,
    # (m*2 + kind)*2 + measure
    ((2, 1, 2), (3, 0, 0)),
    ((3, 0, 2), (3, 1, 0)),
    ((1, 0, 0), (3, 0, 1)),
])
def test_loop_switch_matches_fresh_context(first, second):
    import adscf_amd as amd

    print(f"loop switch {first} -> {second}", flush=True)
    w0 = _field()
    with _ctx(amd) as a:
        a.scf_set_field(w0)
        target = None
        if first[1] == 1 or second[1] == 1:
            a.scf_evaluate()
            target = a.scf_get()[1] * 1.01  # densities near the field's own
        mix, kind, meas = first
        a.scf_set_mixing(mix)
        a.scf_set_residual(kind=kind, measure=meas, target=target)
        if meas == 2:
            a.scf_set_density(a.scf_get()[1])
        s1 = a.scf_iterate(4, tol=-1.0)
        assert s1.iters == 4
        w_switch, _ = a.scf_get()
        mix, kind, meas = second
        a.scf_set_mixing(mix)
        a.scf_set_residual(kind=kind, measure=meas, target=target)
        s2 = a.scf_iterate(9, tol=-1.0)
        fa, pa = a.scf_get()
        ra = a.scf_residuals()
    with _ctx(amd) as b:
        b.scf_set_field(w_switch)
        b.scf_set_mixing(mix)
        b.scf_set_residual(kind=kind, measure=meas, target=target)
        s3 = b.scf_iterate(9, tol=-1.0)
        fb, pb = b.scf_get()
        rb = b.scf_residuals()
    assert s2.iters == s3.iters == 9
    rel = lambda x, y: float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
    print(f"field {rel(fa, fb):.2e} density {rel(pa, pb):.2e} accuracy {s2.accuracy:.6e} {s3.accuracy:.6e}")
    assert rel(fa, fb) < 1e-12 and rel(pa, pb) < 1e-12
    assert abs(s2.accuracy - s3.accuracy) <= 1e-10 * abs(s3.accuracy)
    assert np.allclose(ra[-9:], rb[-9:], rtol=1e-10, atol=0.0)


def test_density_change_measure_rejected_for_blends():
    """measure 2 compares the densities with their previous contents, which a blend of
    several polymers accumulates in place: the loop refuses it."""
    import adscf_amd as amd

    with amd.Context(16, 8, 1, 4.0, 2.0, 1.0, 0.1, precision=amd.FP64) as ctx:
        ctx.set_interactions(10.0, [[0.0, 10.0], [10.0, 0.0]], [1.0, 1.0])
        ctx.blend_setup([1, 2], [10, 10], [1.0, 1.0], mixing_dim=2, lam_mix=0.5)
        ctx.scf_set_field(np.random.default_rng(1).standard_normal((2, 1, 8, 16)))
        ctx.scf_set_residual(kind=0, measure=2)
        with pytest.raises(amd.AdscfError, match="blend"):
            ctx.scf_iterate(1, tol=-1.0)
        ctx.scf_set_residual(kind=0, measure=1)
        st = ctx.scf_iterate(2, tol=-1.0)
        assert st.iters == 2
