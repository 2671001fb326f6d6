"""GPU checks of the adjoint (adscf_adjoint): its forward part against the oracle,
its gradients against central finite differences of the FP64 forward, checkpoint
independence, the FP32 error against FP64, and the torch.autograd wrapper of the
verification harness.

Workload (the adjoint configuration): 32^3 box of side 4.0, Ns = 100 (50/50),
chiN = 15, kappa = 25, lamellar seed fields, loss
L = sum (phi_A - phi_target)^2 + 0.1 F with a seeded smooth target in [0.1, 0.9],
checkpoints every 14 contour steps, central differences with step 1e-6 along 64
seeded random directions in (w_A, w_B, Lx, Ly, Lz), scaled to unit RMS per
component (a unit 2-norm over 65,539 parameters would move each field value by
~4e-9 and leave the loss change only ~1e-11 relative, below the rounding noise of
the loss), plus the three box axes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, L, NS, CHIN, LAM, CKPT = 32, 4.0, 100, 15.0, 0.1, 14


@pytest.fixture(scope="module")
def setup():
    import adscf_amd
    from adscf_amd import fields

    w = fields.lamellar_seed(N, L, CHIN)
    phit = fields.smooth_target(N, sigma=2.0, lo=0.1, hi=0.9, seed=2008)
    return adscf_amd, fields, w, phit


def _ctx(amd, prec):
    ctx = amd.Context(N, N, N, L, L, L, 1.0 / NS, precision=prec)
    ctx.set_interactions(25.0, [[0.0, CHIN], [CHIN, 0.0]], [1.0, 1.0])
    return ctx


def test_adjoint_forward_matches_oracle(setup, orc):
    amd, fields, w, phit = setup
    with _ctx(amd, amd.FP64) as ctx:
        out = ctx.adjoint([1, 2], [50, 50], w, phit, lam=LAM, checkpoint_every=CKPT, gradients=False)
    g = orc.Grid(N, N, N, L, L, L, 1.0 / NS)
    fh = orc.FH(25.0, [[0.0, CHIN], [CHIN, 0.0]], [1.0, 1.0])
    den, info = orc.species_densities(g, [1, 2], [50, 50], w)
    F = orc.total_energy(g, info, fh, den, w)
    assert np.max(np.abs(out["phi"] - den)) < 1e-10 * np.max(den)
    assert abs(out["F"] - F) < 1e-10 * abs(F)
    loss = np.sum((den[0] - phit) ** 2) + LAM * F
    assert abs(out["loss"] - loss) < 1e-10 * abs(loss)


def test_adjoint_matches_central_differences(setup):
    amd, fields, w, phit = setup
    import adjoint_harness as H

    box = np.array([L, L, L])
    with _ctx(amd, amd.FP64) as ctx:
        out = ctx.adjoint([1, 2], [50, 50], w, phit, lam=LAM, checkpoint_every=CKPT)
        grad = np.concatenate([out["grad_w"].ravel(), out["grad_box"]])
        dirs = list(fields.unit_directions(64, [w.size, 3], seed=2009) * np.sqrt(w.size + 3))
        for a in range(3):
            e = np.zeros(w.size + 3)
            e[w.size + a] = 1.0
            dirs.append(e)
        h = 1e-6
        worst = 0.0
        rows = []
        for d in dirs:
            dw = d[: w.size].reshape(w.shape)
            db = d[w.size:]
            lp = H.loss_only(ctx, ([1, 2], [50, 50]), w + h * dw, box + h * db, phit, LAM, CKPT)
            lm = H.loss_only(ctx, ([1, 2], [50, 50]), w - h * dw, box - h * db, phit, LAM, CKPT)
            fd = (lp - lm) / (2 * h)
            ad = float(grad @ d)
            rel = abs(fd - ad) / max(abs(ad), 1e-12)
            rows.append((rel, fd, ad))
            worst = max(worst, rel)
        for rel, fd, ad in sorted(rows)[-5:]:
            print(f"  fd {fd:.12e} adjoint {ad:.12e} rel {rel:.2e}")
        print(f"adjoint vs central FD: worst relative error {worst:.3e} over {len(dirs)} directions; "
              f"box gradient {out['grad_box']}")
        assert worst < 1e-6


def test_adjoint_checkpoint_independent(setup):
    amd, fields, w, phit = setup
    with _ctx(amd, amd.FP64) as ctx:
        a = ctx.adjoint([1, 2], [50, 50], w, phit, lam=LAM, checkpoint_every=1)
        b = ctx.adjoint([1, 2], [50, 50], w, phit, lam=LAM, checkpoint_every=CKPT)
        c = ctx.adjoint([1, 2], [50, 50], w, phit, lam=LAM, checkpoint_every=0)
    for x in (b, c):
        assert np.max(np.abs(x["grad_w"] - a["grad_w"])) <= 1e-12 * np.max(np.abs(a["grad_w"]))
        assert np.max(np.abs(x["grad_box"] - a["grad_box"])) <= 1e-12 * np.max(np.abs(a["grad_box"]))


def test_adjoint_fp32_error_vs_fp64(setup):
    amd, fields, w, phit = setup
    with _ctx(amd, amd.FP64) as ctx:
        r64 = ctx.adjoint([1, 2], [50, 50], w, phit, lam=LAM, checkpoint_every=CKPT)
    with _ctx(amd, amd.FP32) as ctx:
        r32 = ctx.adjoint([1, 2], [50, 50], w, phit, lam=LAM, checkpoint_every=CKPT)
    ew = np.max(np.abs(r32["grad_w"] - r64["grad_w"])) / np.max(np.abs(r64["grad_w"]))
    eb = np.max(np.abs(r32["grad_box"] - r64["grad_box"])) / np.max(np.abs(r64["grad_box"]))
    el = abs(r32["loss"] - r64["loss"]) / abs(r64["loss"])
    print(f"FP32 adjoint vs FP64: loss {el:.2e}, grad_w {ew:.2e}, grad_box {eb:.2e}")
    # stated FP32 bounds
    assert el < 1e-5 and ew < 1e-3 and eb < 1e-3


def test_torch_autograd_wrapper(setup):
    amd, fields, w, phit = setup
    import adjoint_harness as H

    with _ctx(amd, amd.FP64) as ctx:
        direct = ctx.adjoint([1, 2], [50, 50], w, phit, lam=LAM, checkpoint_every=CKPT)
        loss, gw, gb = H.loss_and_grads(ctx, ([1, 2], [50, 50]), w, [L, L, L], phit, LAM, CKPT)
    assert loss == direct["loss"]
    assert np.array_equal(gw, direct["grad_w"]) and np.array_equal(gb, direct["grad_box"])


@pytest.mark.parametrize("dims,bt,bs", [((8, 6, 5), [1, 2, 1], [3, 4, 2]), ((16, 8, 12), [1, 2], [9, 6])])
def test_adjoint_matches_numpy_restatement(dims, bt, bs):
    """Exact comparison (no finite-difference noise) against oracle/np_adjoint.py."""
    import adscf_amd
    from oracle import np_adjoint

    nx, ny, nz = dims
    rng = np.random.default_rng(3)
    w = rng.standard_normal((2, nz, ny, nx))
    phit = rng.uniform(0.1, 0.9, (nz, ny, nx))
    Lb = (2.0, 1.7, 1.5)
    ds = 0.05
    chi = [[0.0, 12.0], [12.0, 0.0]]
    loss, F, gw, gb = np_adjoint.adjoint(w, np.array(Lb), bt, bs, ds, 25.0, chi, [1.0, 0.8], phit, 0.1)
    with adscf_amd.Context(nx, ny, nz, *Lb, ds) as ctx:
        ctx.set_interactions(25.0, chi, [1.0, 0.8])
        for ck in (1, 3, 0):
            out = ctx.adjoint(bt, bs, w, phit, lam=0.1, checkpoint_every=ck)
            assert abs(out["loss"] - loss) < 1e-11 * abs(loss)
            assert abs(out["F"] - F) < 1e-11 * abs(F)
            assert np.max(np.abs(out["grad_w"] - gw)) < 1e-10 * np.max(np.abs(gw))
            assert np.max(np.abs(out["grad_box"] - gb)) < 1e-10 * np.max(np.abs(gb))
