"""GPU: the Simpson contour rule (adscf_set_contour_rule) against the oracle in the same rule.

The trapezoid is the reference's rule and the default; Simpson is an extra mode (composite
Simpson per block, a 3/8 panel closing a block of odd length), whose oracle weights are
pinned by quadrature identities in test_oracle_contour_rules.py.  Checked here: chain and
homopolymer densities (FP64 to 1e-10, FP32-mixed to the stated 2e-5), one SCF iteration of
the captured loop, a rule switch on a context that has already captured its loop, and the
adjoint's gradients against central differences of its own Simpson forward.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b))))


@pytest.fixture
def simpson(orc):
    orc.set_contour_rule("simpson")
    yield orc
    orc.set_contour_rule("trapezoid")


@pytest.mark.parametrize("case", [
    ((16, 16, 16), (4.0, 4.0, 4.0), 0.02, [1, 2], [25, 25], 1),
    ((16, 12, 20), (3.0, 2.5, 4.5), 0.01, [1, 2, 1], [25, 50, 25], 1),
    ((12, 10, 8), (3.0, 3.0, 3.0), 0.02, [1, 2, 1, 2, 1, 2], [15, 3, 3, 2, 1, 4], 1),
    ((16, 16, 16), (4.0, 4.0, 4.0), 0.02, [1, 2], [15, 36], 2),
    ((12, 10, 8), (3.0, 3.0, 3.0), 0.02, [1, 2, 1], [10, 0, 13], 1),
    ((86, 1, 75), (17.2, 1.0, 1.0), 0.01, [1, 2], [11, 10], 1),
])
def test_propagate_chain_simpson_fp64(simpson, case):
    import adscf_amd as amd

    orc = simpson
    dims, L, ds, bt, bs, ens = case
    print(f"Simpson chain {case}", flush=True)
    g = orc.Grid(*dims, *L, ds)
    w = 2.0 * np.random.default_rng(3).standard_normal((2,) + g.shape)
    E = orc.expfield(g, w)
    ref = orc.propagate_chain(g, bt, bs, E, ensemble=ens, mu=0.3)
    with amd.Context(*dims, *L, ds) as ctx:
        ctx.set_contour_rule("simpson")
        got = ctx.propagate_chain(bt, bs, E, ensemble=ens, mu=0.3)
        ctx.set_contour_rule("trapezoid")
        trap = ctx.propagate_chain(bt, bs, E, ensemble=ens, mu=0.3)
    assert _rel(got.distribution, ref.distribution) < 1e-10
    assert _rel(got.density, ref.density) < 1e-10
    assert abs(got.Qsum - ref.Qsum) < 1e-10 * ref.Qsum
    assert abs(got.energy - ref.energy) < 1e-10 * abs(ref.energy)
    assert _rel(trap.distribution, ref.distribution) > 1e-8  # the rule took effect


def test_propagate_chain_simpson_fp32_bound(simpson):
    import adscf_amd as amd

    orc = simpson
    g = orc.Grid(32, 32, 32, 4.0, 4.0, 4.0, 0.01)
    E = orc.expfield(g, 2.0 * np.random.default_rng(4).standard_normal((2,) + g.shape))
    ref = orc.propagate_chain(g, [1, 2], [50, 51], E)
    with amd.Context(32, 32, 32, 4.0, 4.0, 4.0, 0.01, precision=amd.FP32) as ctx:
        ctx.set_contour_rule("simpson")
        got = ctx.propagate_chain([1, 2], [50, 51], E)
    assert abs(got.Qsum - ref.Qsum) < 2e-5 * ref.Qsum
    assert _rel(got.density, ref.density) < 2e-5


@pytest.mark.parametrize("segs", [1, 8, 13])
def test_propagate_homopolymer_simpson(simpson, segs):
    import adscf_amd as amd

    orc = simpson
    g = orc.Grid(16, 12, 10, 3.2, 2.8, 3.0, 0.05)
    E = orc.expfield(g, np.random.default_rng(5).standard_normal((2,) + g.shape))
    ref = orc.propagate_homopolymer(g, 2, segs, E)
    with amd.Context(16, 12, 10, 3.2, 2.8, 3.0, 0.05) as ctx:
        ctx.set_contour_rule("simpson")
        got = ctx.propagate_homopolymer(2, segs, E)
    assert _rel(got.density, ref.density) < 1e-10
    assert abs(got.Qsum - ref.Qsum) < 1e-10 * ref.Qsum


def _scf_ctx(amd, dims, L, ds, chin, segs, m):
    ctx = amd.Context(*dims, *L, ds, precision=amd.FP64)
    ctx.set_interactions(25.0, [[0.0, chin], [chin, 0.0]], [1.0, 1.0])
    ctx.scf_setup([1, 2], segs, mixing_dim=m, lam_mix=0.5)
    return ctx


def test_scf_iteration_simpson_matches_oracle(simpson):
    import adscf_amd as amd

    orc = simpson
    dims, L, ds, segs, chin, m = (16, 16, 16), (4.0, 4.0, 4.0), 0.02, [21, 28], 15.0, 5
    w = 3.0 * np.random.default_rng(6).standard_normal((2,) + dims[::-1])
    g = orc.Grid(*dims, *L, ds)
    fh = orc.FH(25.0, [[0.0, chin], [chin, 0.0]], [1.0, 1.0])
    ref = orc.scf_anderson(g, [1, 2], segs, fh, w, m, lam=0.5, tol=-1.0, maxit=1)
    with _scf_ctx(amd, dims, L, ds, chin, segs, m) as ctx:
        ctx.set_contour_rule("simpson")
        ctx.scf_set_field(w)
        st = ctx.scf_iterate(1, tol=-1.0)
        f, d = ctx.scf_get()
    assert st.iters == 1
    print(f"field {_rel(f, ref['field']):.2e} density {_rel(d, ref['density']):.2e}")
    assert _rel(d, ref["density"]) < 1e-10
    assert _rel(f, ref["field"]) < 1e-10
    assert abs(st.accuracy - ref["accuracy"]) < 1e-10 * ref["accuracy"]


def test_rule_switch_recaptures_the_loop(orc):
    """A context that has captured its iteration with the trapezoid evaluates with Simpson
    after the switch: its densities at the current field equal the oracle's Simpson ones."""
    import adscf_amd as amd

    dims, L, ds, segs, chin, m = (16, 12, 10), (3.2, 2.8, 3.0), 0.02, [15, 20], 12.0, 4
    w = 2.0 * np.random.default_rng(8).standard_normal((2,) + dims[::-1])
    g = orc.Grid(*dims, *L, ds)
    with _scf_ctx(amd, dims, L, ds, chin, segs, m) as ctx:
        ctx.scf_set_field(w)
        ctx.scf_iterate(3, tol=-1.0)
        ctx.set_contour_rule("simpson")
        ctx.scf_evaluate()
        f, d = ctx.scf_get()
        st = ctx.scf_iterate(2, tol=-1.0)
        assert np.isfinite(st.accuracy)
    orc.set_contour_rule("simpson")
    try:
        den, _ = orc.species_densities(g, [1, 2], segs, f)
    finally:
        orc.set_contour_rule("trapezoid")
    assert _rel(d, den) < 1e-10


def test_adjoint_simpson_central_differences():
    import adscf_amd as amd
    import adjoint_harness as H
    from adscf_amd import fields

    n, L, segs, lam, ck = 16, 3.5, [21, 30], 0.1, 7
    rng = np.random.default_rng(10)
    w = fields.lamellar_seed(n, L, 14.0) + 0.3 * rng.standard_normal((2, n, n, n))
    phit = fields.smooth_target(n, sigma=1.5, lo=0.2, hi=0.8, seed=11)
    box = np.array([L, L, L])
    with amd.Context(n, n, n, L, L, L, 1.0 / sum(segs), precision=amd.FP64) as ctx:
        ctx.set_interactions(25.0, [[0.0, 14.0], [14.0, 0.0]], [1.0, 1.0])
        ctx.set_contour_rule("simpson")
        out = ctx.adjoint([1, 2], segs, w, phit, lam=lam, checkpoint_every=ck)
        grad = np.concatenate([out["grad_w"].ravel(), out["grad_box"]])
        dirs = list(fields.unit_directions(12, [w.size, 3], seed=12) * np.sqrt(w.size + 3))
        for a in range(3):
            e = np.zeros(w.size + 3)
            e[w.size + a] = 1.0
            dirs.append(e)
        h = 1e-6
        worst = 0.0
        for d in dirs:
            dw, db = d[: w.size].reshape(w.shape), d[w.size:]
            lp = H.loss_only(ctx, ([1, 2], segs), w + h * dw, box + h * db, phit, lam, ck)
            lm = H.loss_only(ctx, ([1, 2], segs), w - h * dw, box - h * db, phit, lam, ck)
            fd = (lp - lm) / (2 * h)
            ad = float(grad @ d)
            worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-12))
        ctx.set_contour_rule("trapezoid")
        trap = ctx.adjoint([1, 2], segs, w, phit, lam=lam, checkpoint_every=ck)
    print(f"worst relative FD error over {len(dirs)} directions: {worst:.2e}")
    assert worst < 1e-6
    assert abs(trap["loss"] - out["loss"]) > 1e-9 * abs(out["loss"])


def test_rule_is_validated():
    import adscf_amd as amd

    with amd.Context(8, 8, 8, 2.0, 2.0, 2.0, 0.1) as ctx:
        with pytest.raises(amd.AdscfError):
            ctx.set_contour_rule(2)
