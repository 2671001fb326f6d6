"""GPU checks of the device-resident homopolymer blend (Main_Blend/main_blend.f90): blend
densities, the saddle-point and conjugate-field loops over them, and the EPD step
Chain_Dynamics_Timestep_Euler (main_blend.f90:605-667), against the oracle.

As in tests/test_gpu_ddft.py: the conjugate-field solve is compared exactly through its
simple-mixing iterations (m = 2 here) and by outcome afterwards, because its first Anderson
step amplifies rounding; an EPD step's density update does not depend on that trajectory
and is compared at 1e-12 against the oracle's step from the GPU's state before it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# the shipped Main_Blend inputs on a smaller grid: chains of length 1 at ds = 0.1, kappa = chi = 10
KAPPA, CHI = 10.0, 10.0


def _grid(orc, nx=32, ny=8, nz=1, box=(10.0, 2.5, 1.0)):
    return orc.Grid(nx, ny, nz, *box, 0.1)


def _ctx(amd, g, types, segs, rho, m=2):
    ctx = amd.Context(g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, g.ds, precision=amd.FP64)
    ctx.set_interactions(KAPPA, [[0.0, CHI], [CHI, 0.0]], [1.0, 1.0])
    ctx.blend_setup(types, segs, rho, mixing_dim=m, lam_mix=0.5)
    return ctx


# (24, 6, 4): line passes; (32, 32, 4): square xy planes, so the sweeps run xy-plane passes
@pytest.mark.parametrize("dims", [(24, 6, 4, (6.0, 1.5, 1.0)), (32, 32, 4, (6.0, 6.0, 1.0))])
@pytest.mark.parametrize("types,segs", [([1, 2], [10, 10]), ([1, 2, 1], [10, 14, 6])])
def test_blend_densities_match_oracle(orc, types, segs, dims):
    import adscf_amd as amd

    g = _grid(orc, *dims)
    print(f"blend densities dims={dims[:3]} types={types}", flush=True)
    vol, dvol = g.volume()
    rho = [20.0 / vol * (i + 1) for i in range(len(types))]
    field = orc.MT(2007).grnd(2 * g.ng).reshape(2, g.nz, g.ny, g.nx).copy()
    bl = orc.Blend(types, segs, rho)
    ref = orc.blend_densities(g, bl, field)
    with _ctx(amd, g, types, segs, rho) as ctx:
        ctx.scf_set_field(field)
        ctx.scf_evaluate()
        _, phi = ctx.scf_get()
        sc = ctx.blend_scalars()
        en = ctx.scf_energy()
    assert np.max(np.abs(phi - ref)) <= 1e-12 * np.max(ref)
    assert np.allclose(sc["mu"], bl.mu, rtol=1e-12, atol=0) and np.allclose(sc["energy"], bl.energy, rtol=1e-12,
                                                                              atol=0)
    fh = orc.FH(KAPPA, [[0.0, CHI], [CHI, 0.0]], [1.0, 1.0])
    F = -dvol * np.sum(ref * field) + sum(bl.energy) + orc.interaction_energy(g, fh, ref)
    assert abs(en["F"] - F) <= 1e-12 * abs(F)


def test_blend_conjugate_fields_match_oracle(orc):
    import adscf_amd as amd

    g = _grid(orc)
    vol, dvol = g.volume()
    rho = [25.0 / vol, 25.0 / vol]
    bl = orc.Blend([1, 2], [10, 10], rho)
    rng = np.random.default_rng(4)
    w0 = rng.uniform(0, 1, (2, 1, g.ny, g.nx))
    target = orc.blend_densities(g, bl, w0 + 0.3 * rng.standard_normal(w0.shape))
    ref = orc.blend_conjugate_fields(g, bl, target, w0, m=2, thresh=-1.0, maxit=2)
    thresh = orc.BLEND_CONJ_ACCURACY
    with _ctx(amd, g, [1, 2], [10, 10], rho) as ctx:
        ctx.scf_set_field(w0)
        it, acc = ctx.conjugate_fields(target, thresh=-1.0, maxit=2, measure=1)
        f2, _ = ctx.scf_get()
        assert it == ref["iters"] == 3
        assert abs(acc - ref["accuracy"]) <= 1e-12 * ref["accuracy"]
        assert np.max(np.abs(f2 - ref["field"])) <= 1e-12 * np.max(np.abs(ref["field"]))
        ctx.scf_set_field(w0)
        it, acc = ctx.conjugate_fields(target, thresh=thresh, maxit=2000, measure=1)
        field, _ = ctx.scf_get()
    d = orc.blend_densities(g, bl, field)
    a = np.sum(np.abs(d - target)) * dvol
    print(f"GPU conjugate solve: {it} iterations, accuracy {acc:.4e} (oracle evaluation {a:.4e})")
    assert it <= 2000 and acc <= thresh and a <= thresh * (1 + 1e-9)


def test_blend_saddle_point_by_anderson(orc):
    """The device-resident saddle-point loop over the blend (Anderson, accuracy sum|dfield| dvol,
    main_blend.f90:567): the oracle finds the GPU's final field self-consistent."""
    import adscf_amd as amd

    g = _grid(orc)
    vol, dvol = g.volume()
    rho = [25.0 / vol, 25.0 / vol]
    w0 = orc.MT(2007).grnd(2 * g.ng).reshape(2, 1, g.ny, g.nx).copy()
    with _ctx(amd, g, [1, 2], [10, 10], rho, m=8) as ctx:
        ctx.scf_set_residual(kind=0, measure=1)
        ctx.scf_set_field(w0)
        st = ctx.scf_iterate(5000, tol=1e-9)
        field, phi = ctx.scf_get()
        en = ctx.scf_energy()
    assert st.converged, st
    bl = orc.Blend([1, 2], [10, 10], rho)
    fh = orc.FH(KAPPA, [[0.0, CHI], [CHI, 0.0]], [1.0, 1.0])
    d = orc.blend_densities(g, bl, field)
    resid = np.sum(np.abs(orc.mean_fields(g, fh, d) - field)) * dvol
    print(f"blend Anderson SCF: {st.iters} iterations, accuracy {st.accuracy:.3e}, oracle residual {resid:.3e}")
    assert resid <= 2e-9
    F = -dvol * np.sum(d * field) + sum(bl.energy) + orc.interaction_energy(g, fh, d)
    assert abs(en["F"] - F) <= 1e-10 * abs(F)


def test_epd_steps_match_oracle(orc):
    import adscf_amd as amd

    g = _grid(orc)
    vol, dvol = g.volume()
    rho = [25.0 / vol, 25.0 / vol]
    fh = orc.FH(KAPPA, [[0.0, CHI], [CHI, 0.0]], [1.0, 1.0])
    field = orc.MT(2007).grnd(2 * g.ng).reshape(2, 1, g.ny, g.nx).copy()
    bl = orc.Blend([1, 2], [10, 10], rho)
    den = orc.blend_densities(g, bl, field)
    with _ctx(amd, g, [1, 2], [10, 10], rho) as ctx:
        ctx.scf_set_field(field)
        ctx.ddft_set_density(den)
        for step in range(3):
            d_ref, f_ref = den.copy(), field.copy()
            it_ref = orc.epd_step(g, bl, fh, d_ref, f_ref, 0.001)
            it = ctx.epd_step(0.001)
            den = ctx.ddft_get_density()
            field, _ = ctx.scf_get()
            e = ctx.ddft_energy()
            chain = orc.blend_densities(g, bl, field)
            conj = np.sum(np.abs(chain - den)) * dvol
            e_ref = -dvol * np.sum(den * field) + sum(bl.energy) + orc.interaction_energy(g, fh, den)
            print(f"step {step}: GPU solve {it} iterations, oracle {it_ref}; "
                  f"max|dphi| {np.max(np.abs(den - d_ref)):.2e}, conjugacy {conj:.3e}, F {e['F']:.12f}")
            assert np.max(np.abs(den - d_ref)) <= 1e-12 * np.max(d_ref)
            assert conj <= orc.BLEND_CONJ_ACCURACY * (1 + 1e-9)
            assert abs(e["F"] - e_ref) <= 1e-10 * abs(e_ref)


def test_blend_saddle_point_lambda_mixing_matches_oracle(orc):
    """Main_Blend's own search (main_blend.f90:437-486): Mixing_Lambda with lammax 0.5 and the
    accuracy sum|dfield| dvol.  Lambda mixing has no Gram solve, but its step (a ratio of norms
    of differences of consecutive iterates) still amplifies rounding: on the oracle alone a
    1e-16 change of the start field moves the accuracy after 100 iterations by 7e-10 relative
    and after 150 by 1e-6.  The trajectory is compared after 100 iterations at 1e-7 (accuracy)
    and 1e-9 (field); the run to the driver's 1e-7 by its iteration count, and by the oracle."""
    import adscf_amd as amd

    g = _grid(orc)
    vol, dvol = g.volume()
    rho = [25.0 / vol, 25.0 / vol]
    bl = orc.Blend([1, 2], [10, 10], rho)
    fh = orc.FH(KAPPA, [[0.0, CHI], [CHI, 0.0]], [1.0, 1.0])
    w0 = orc.MT(2007).grnd(2 * g.ng).reshape(2, 1, g.ny, g.nx).copy()
    ref = orc.blend_scf(g, bl, fh, w0, m=2, mixing_type=2, lam=0.5, thresh=-1.0, maxit=100)
    conv = orc.blend_scf(g, bl, fh, w0, m=2, mixing_type=2, lam=0.5, maxit=20000)
    with _ctx(amd, g, [1, 2], [10, 10], rho, m=2) as ctx:
        ctx.scf_set_mixing(2)
        ctx.scf_set_residual(kind=0, measure=1)
        ctx.scf_set_field(w0)
        st = ctx.scf_iterate(100, tol=-1.0)
        field, phi = ctx.scf_get()
        print(f"100 lambda iterations: accuracy GPU {st.accuracy:.12e} oracle {ref['accuracy']:.12e}")
        assert st.iters == 100
        assert abs(st.accuracy - ref["accuracy"]) <= 1e-7 * ref["accuracy"]
        assert np.max(np.abs(field - ref["field"])) <= 1e-9 * np.max(np.abs(ref["field"]))
        st = ctx.scf_iterate(20000, tol=orc.BLEND_SCF_ACCURACY)
        field, _ = ctx.scf_get()
    assert st.converged
    print(f"converged: GPU {st.iters} iterations, oracle {conv['iters']}")
    assert abs(st.iters - conv["iters"]) <= 5  # the end game amplifies rounding further
    d = orc.blend_densities(g, bl, field)
    resid = np.sum(np.abs(orc.mean_fields(g, fh, d) - field)) * dvol
    print(f"lambda mixing converged after {st.iters} iterations; oracle residual {resid:.3e}")
    assert resid <= 1.01 * orc.BLEND_SCF_ACCURACY


def test_epd_step_fp32_mixed_error(orc):
    """FP32-mixed EPD step against FP64 from the same start.  The eps = 1e-4 finite difference
    divides the FP32 rounding of the density response by eps, so the stated bound is 1e-4 of
    max density at dt = 1e-3 (against 1e-5 for the DDFT step)."""
    import adscf_amd as amd

    g = _grid(orc)
    vol, dvol = g.volume()
    rho = [25.0 / vol, 25.0 / vol]
    field = orc.MT(2007).grnd(2 * g.ng).reshape(2, 1, g.ny, g.nx).copy()
    den = orc.blend_densities(g, orc.Blend([1, 2], [10, 10], rho), field)
    out = {}
    for prec in (amd.FP64, amd.FP32):
        ctx = amd.Context(g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, g.ds, precision=prec)
        with ctx:
            ctx.set_interactions(KAPPA, [[0.0, CHI], [CHI, 0.0]], [1.0, 1.0])
            ctx.blend_setup([1, 2], [10, 10], rho, mixing_dim=2, lam_mix=0.5)
            ctx.scf_set_field(field)
            ctx.ddft_set_density(den)
            ctx.epd_step(0.001)
            out[prec] = ctx.ddft_get_density()
    err = np.max(np.abs(out[amd.FP32] - out[amd.FP64])) / np.max(out[amd.FP64])
    print(f"FP32-mixed EPD step: density error {err:.2e}")
    assert err < 1e-4
