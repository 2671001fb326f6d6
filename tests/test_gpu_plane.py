"""GPU parity of the xy-plane passes: the x and y operators of a contour step applied in one
pass per chain and z-plane, alternating with z passes (one dependent pass per contour step).

Each compiled plane shape is checked against the CPU FP64 oracle at 1e-10 relative (plain
FP64) and against the line passes of the same context (the operators commute, so the two
sweeps differ only by rounding), over contour lengths whose sweep ends with a z pass (odd
Ns) or a plane pass (even Ns), a single-step chain, grids whose nz is not a multiple of the
8 planes a group of workgroups takes, and both ensembles.  FP32-mixed is checked at the
2e-5 bound stated for the line passes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-10
SHAPES = {32: [(4, 8), (8, 4)], 48: [(8, 6), (6, 8)], 64: [(4, 0), (8, 8)], 128: [(4, -8)]}
CASES = [
    # dims, box, ds, block types, segments, ensemble
    ((32, 32, 12), (4.0, 4.5, 3.0), 0.02, [1, 2], [7, 6], 1),          # odd Ns: last pass along z
    ((32, 32, 16), (4.0, 4.0, 4.0), 0.02, [1, 2, 1], [5, 10, 5], 1),   # even Ns: last pass a plane pass
    ((32, 32, 8), (4.0, 4.0, 4.0), 0.02, [2], [1], 1),                 # one contour step
    ((48, 48, 10), (4.5, 4.5, 3.0), 0.01, [1, 2], [15, 10], 2),        # grand canonical
    ((64, 64, 9), (4.5, 4.0, 4.5), 0.005, [1, 2], [12, 9], 1),
]


@pytest.fixture(scope="module")
def amd():
    import adscf_amd

    adscf_amd.load_library()
    return adscf_amd


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


def _chain(ctx, bt, bs, E, ens):
    return ctx.propagate_chain(bt, bs, E, ensemble=ens, rho=1.0, mu=0.3)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + "_Ns%d" % sum(c[4]))
@pytest.mark.parametrize("which", [0, 1])
def test_plane_sweep_matches_oracle_and_line_passes(amd, orc, case, which):
    dims, box, ds, bt, bs, ens = case
    shape = SHAPES[dims[0]][which]
    print(f"plane sweep dims={dims} shape={shape} Ns={sum(bs)} ens={ens}", flush=True)
    g = orc.Grid(*dims, *box, ds)
    rng = np.random.default_rng(21)
    w = 2.0 * rng.standard_normal((2,) + g.shape)
    E = orc.expfield(g, w)
    ref = orc.propagate_chain(g, bt, bs, E, ensemble=ens, rho=1.0, mu=0.3)
    with amd.Context(*dims, *box, ds) as ctx:
        ctx.set_plane_passes(1, *shape)
        pp = ctx.plane_passes()
        assert pp["active"] and (pp["N1"], pp["N2"]) == shape
        got = _chain(ctx, bt, bs, E, ens)
        ctx.set_plane_passes(0)
        assert not ctx.plane_passes()["active"]
        lines = _chain(ctx, bt, bs, E, ens)
    assert abs(got.Qsum - ref.Qsum) < FP64_RTOL * abs(ref.Qsum)
    assert _rel(got.distribution, ref.distribution) < FP64_RTOL
    assert _rel(got.density, ref.density) < FP64_RTOL
    assert abs(got.energy - ref.energy) < FP64_RTOL * abs(ref.energy)
    # the same propagator up to rounding
    assert _rel(got.density, lines.density) < 1e-12
    assert abs(got.Qsum - lines.Qsum) < 1e-12 * abs(lines.Qsum)


@pytest.mark.parametrize("n,shape", [(32, (4, 8)), (64, (4, 0)), (64, (8, 8)), (128, (4, -8))])
def test_plane_sweep_fp32_error_bound(amd, orc, n, shape):
    dims, L, ds = (n, n, 16), (4.0, 4.0, 3.0), 0.01
    print(f"plane sweep FP32 dims={dims} shape={shape}", flush=True)
    g = orc.Grid(*dims, *L, ds)
    w = 2.0 * np.random.default_rng(4).standard_normal((2,) + g.shape)
    E = orc.expfield(g, w)
    ref = orc.propagate_chain(g, [1, 2], [50, 50], E)
    with amd.Context(*dims, *L, ds, precision=amd.FP32) as ctx:
        ctx.set_plane_passes(1, *shape)
        got = ctx.propagate_chain([1, 2], [50, 50], E)
    assert abs(got.Qsum - ref.Qsum) < 2e-5 * abs(ref.Qsum)
    assert _rel(got.density, ref.density) < 2e-5


def test_plane_passes_refused_where_the_grid_does_not_allow_them(amd):
    with amd.Context(32, 16, 16, 3.0, 3.0, 3.0, 0.01) as ctx:  # nx != ny
        assert not ctx.plane_passes()["active"]
        with pytest.raises(amd.AdscfError):
            ctx.set_plane_passes(1)
        ctx.set_plane_passes(-1)
        assert not ctx.plane_passes()["active"]
    with amd.Context(32, 32, 16, 3.0, 3.0, 3.0, 0.01) as ctx:
        with pytest.raises(amd.AdscfError):
            ctx.set_plane_passes(1, 8, 8)  # no 32-point plane kernel with this shape


def test_plane_scf_iteration_matches_line_passes(amd):
    """The captured SCF iteration on plane passes follows the line-pass loop (FP64, 32^3 x 40,
    Anderson m = 4): the same residual up to rounding at each of the first m + 1 iterations.
    Later iterations go through Anderson solves whose first Gram system, built from the
    nearly collinear residuals of the lam00 = 0.001 steps, amplifies rounding (any two
    summation orders part there), so they are not compared."""
    n, L, ns = 32, 4.0, 40
    rng = np.random.default_rng(5)
    w = 3.0 * rng.standard_normal((2, n, n, n))
    acc = {}
    for mode in (0, 1):
        with amd.Context(n, n, n, L, L, L, 1.0 / ns) as ctx:
            ctx.set_plane_passes(mode)
            ctx.set_interactions(25.0, [[0.0, 15.0], [15.0, 0.0]], [1.0, 1.0])
            ctx.scf_setup([1, 2], [20, 20], mixing_dim=4, lam_mix=0.2)
            ctx.scf_set_field(w)
            hist = []
            for _ in range(5):
                hist.append(ctx.scf_iterate(1, tol=0.0).accuracy)
            acc[mode] = np.array(hist)
    print("line passes", acc[0], "plane passes", acc[1])
    assert np.all(np.abs(acc[1] - acc[0]) <= 1e-9 * np.abs(acc[0]))


@pytest.mark.parametrize("ns", [(7, 6), (5, 5)])
def test_plane_sweep_128_fp32_matches_line_passes(amd, orc, ns):
    """128-point planes (two lines per thread, FP32 only): against the line passes of the same
    context at the FP32 rounding level and against the oracle at the stated 2e-5, with the
    sweep ending on a z pass (odd Ns) and on a plane pass (even Ns)."""
    dims, L, ds = (128, 128, 10), (6.0, 6.0, 2.0), 0.01
    print(f"plane sweep 128 FP32 Ns={sum(ns)}", flush=True)
    g = orc.Grid(*dims, *L, ds)
    w = 2.0 * np.random.default_rng(8).standard_normal((2,) + g.shape)
    E = orc.expfield(g, w)
    ref = orc.propagate_chain(g, [1, 2], list(ns), E)
    with amd.Context(*dims, *L, ds, precision=amd.FP32) as ctx:
        ctx.set_plane_passes(1, 4, -8)
        assert ctx.plane_passes()["active"]
        got = ctx.propagate_chain([1, 2], list(ns), E)
        ctx.set_plane_passes(0)
        lines = ctx.propagate_chain([1, 2], list(ns), E)
    assert _rel(got.density, ref.density) < 2e-5 and abs(got.Qsum - ref.Qsum) < 2e-5 * abs(ref.Qsum)
    assert _rel(got.density, lines.density) < 2e-6


def test_plane_passes_not_offered_for_fp64_128(amd):
    with amd.Context(128, 128, 8, 6.0, 6.0, 2.0, 0.01) as ctx:  # FP64: the plane does not fit the LDS
        assert not ctx.plane_passes()["active"]
        with pytest.raises(amd.AdscfError):
            ctx.set_plane_passes(1, 4, -8)
