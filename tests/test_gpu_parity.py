"""GPU parity: the HIP path (through the C-ABI) against the CPU FP64 oracle.

Tolerances: the plain FP64 path must agree with the oracle to 1e-10 relative
(phi, Q, F); the FP32-mixed path is checked against stated error bounds.
Sizes are chosen so the oracle finishes in seconds; the metric-size runs are
checked through size-independent properties in test_gpu_properties.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-10


@pytest.fixture(scope="module")
def amd():
    import adscf_amd

    adscf_amd.load_library()
    return adscf_amd


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


def _fields(shape, seed, amp=3.0, nt=2):
    rng = np.random.default_rng(seed)
    return amp * rng.standard_normal((nt,) + shape)


@pytest.mark.parametrize("dims", [(8, 6, 10), (32, 32, 32), (86, 1, 75), (64, 1, 30), (9, 5, 7), (12, 1, 1)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_real2fourier_fourier2real(amd, orc, dims, prec):
    nx, ny, nz = dims
    print(f"r2c/c2r dims={dims} prec={prec}", flush=True)
    P = amd.FP64 if prec == "fp64" else amd.FP32
    tol = 1e-13 if prec == "fp64" else 2e-6
    with amd.Context(nx, ny, nz, 2.0, 3.0, 4.0, 0.01, precision=P) as ctx:
        g = orc.Grid(nx, ny, nz, 2.0, 3.0, 4.0, 0.01)
        a = np.random.default_rng(1).standard_normal(g.shape)
        c_gpu = ctx.real2fourier(a)
        c_ref = orc.real2fourier(g, a)
        assert _rel(c_gpu, c_ref) < tol
        b_gpu = ctx.fourier2real(c_ref)
        b_ref = orc.fourier2real(g, c_ref)
        assert _rel(b_gpu, b_ref) < tol


@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 32, 32), (48, 48, 48), (64, 1, 30), (86, 1, 75), (9, 5, 7),
                                  (1, 1, 16), (8, 1, 1)])
def test_propagate_step_fp64(amd, orc, dims):
    nx, ny, nz = dims
    print(f"propagate_step dims={dims}", flush=True)
    with amd.Context(nx, ny, nz, 3.0, 2.5, 3.5, 0.01) as ctx:
        g = orc.Grid(nx, ny, nz, 3.0, 2.5, 3.5, 0.01)
        rng = np.random.default_rng(2)
        q = 1.0 + 0.3 * rng.standard_normal(g.shape)
        E = np.exp(-0.5 * 0.01 * 4.0 * rng.standard_normal(g.shape))
        assert _rel(ctx.propagate_step(q, E), orc.propagate_step(g, q, E)) < 1e-13


@pytest.mark.parametrize("case", [
    # (dims, L, ds, block types, segments, ensemble)
    ((16, 16, 16), (4.0, 4.0, 4.0), 0.02, [1, 2], [25, 25], 1),
    ((32, 32, 32), (4.0, 4.0, 4.0), 0.01, [1, 2], [50, 50], 1),          # config-1 sizes
    ((16, 12, 20), (3.0, 2.5, 4.5), 0.01, [1, 2, 1], [25, 50, 25], 1),   # ABA triblock
    ((12, 10, 8), (3.0, 3.0, 3.0), 0.02, [1, 2, 1, 2, 1, 2], [15, 3, 3, 3, 3, 3], 1),  # Main_Multiblock blocks
    ((16, 16, 16), (4.0, 4.0, 4.0), 0.02, [1, 2], [15, 35], 2),          # grand canonical
    ((64, 1, 30), (3.6, 1.0, 3.6), 0.01, [1, 2], [70, 30], 2),           # degenerate y axis
    ((86, 1, 75), (17.2, 1.0, 1.0), 0.01, [1, 2], [10, 10], 1),          # radix 43 and 3*5*5
    ((9, 5, 7), (2.0, 2.0, 2.0), 0.05, [2, 1], [7, 13], 1),
    ((1, 1, 24), (1.0, 1.0, 5.0), 0.01, [1, 2], [30, 30], 1),            # single active axis
    ((12, 10, 8), (3.0, 3.0, 3.0), 0.02, [1, 2, 1], [10, 0, 12], 1),     # empty middle block
    ((12, 10, 8), (3.0, 3.0, 3.0), 0.02, [1, 2, 1], [0, 20, 3], 1),      # empty first block
    ((12, 10, 8), (3.0, 3.0, 3.0), 0.02, [1, 2, 1], [1, 19, 0], 2),      # one-segment and empty last block
])
def test_propagate_chain_fp64(amd, orc, case):
    dims, L, ds, bt, bs, ens = case
    print(f"propagate_chain {case}", flush=True)
    with amd.Context(*dims, *L, ds) as ctx:
        g = orc.Grid(*dims, *L, ds)
        w = _fields(g.shape, 3, amp=2.0)
        E = orc.expfield(g, w)
        ref = orc.propagate_chain(g, bt, bs, E, ensemble=ens, rho=1.0, mu=0.3)
        got = ctx.propagate_chain(bt, bs, E, ensemble=ens, rho=1.0, mu=0.3)
        assert abs(got.Qsum - ref.Qsum) < FP64_RTOL * abs(ref.Qsum)
        assert _rel(got.distribution, ref.distribution) < FP64_RTOL
        assert _rel(got.density, ref.density) < FP64_RTOL
        assert abs(got.energy - ref.energy) < FP64_RTOL * abs(ref.energy)
        if ens == 1:
            assert abs(got.mu - ref.mu) < FP64_RTOL * max(1.0, abs(ref.mu))
        else:
            assert abs(got.rho - ref.rho) < FP64_RTOL * abs(ref.rho)


def test_invalid_ensemble_is_an_error(amd):
    with amd.Context(8, 8, 8, 2.0, 2.0, 2.0, 0.1) as ctx:
        E = np.ones((2, 8, 8, 8))
        with pytest.raises(amd.AdscfError) as ei:
            ctx.propagate_chain([1, 2], [5, 5], E, ensemble=3)
        assert ei.value.code == amd.ERR_ENSEMBLE


def test_propagate_chain_fp32_error_bound(amd, orc):
    dims, L, ds = (32, 32, 32), (4.0, 4.0, 4.0), 0.01
    with amd.Context(*dims, *L, ds, precision=amd.FP32) as ctx:
        g = orc.Grid(*dims, *L, ds)
        w = _fields(g.shape, 4, amp=2.0)
        E = orc.expfield(g, w)
        ref = orc.propagate_chain(g, [1, 2], [50, 50], E)
        got = ctx.propagate_chain([1, 2], [50, 50], E)
        # FP32 propagators/FFTs, FP64 contour sums: stated bound 2e-5 relative
        assert abs(got.Qsum - ref.Qsum) < 2e-5 * abs(ref.Qsum)
        assert _rel(got.density, ref.density) < 2e-5


def test_flory_huggins_fp64(amd, orc):
    dims = (16, 8, 12)
    with amd.Context(*dims, 2.0, 1.0, 1.5, 0.01) as ctx:
        g = orc.Grid(*dims, 2.0, 1.0, 1.5, 0.01)
        chi = np.array([[0.0, 17.0], [17.0, 0.0]])
        ctx.set_interactions(25.0, chi, [1.2, 0.9])
        fh = orc.FH(25.0, chi, [1.2, 0.9])
        den = np.random.default_rng(5).uniform(0.1, 0.9, (2,) + g.shape)
        assert _rel(ctx.mean_fields(den), orc.mean_fields(g, fh, den)) < 1e-14
        e1, e2 = ctx.interaction_energy(den), orc.interaction_energy(g, fh, den)
        assert abs(e1 - e2) < 1e-12 * abs(e2)
        a1, a2 = ctx.average_volume_fraction(den), orc.average_volume_fraction(g, fh, den)
        assert abs(a1 - a2) < 1e-13


def test_expfield(amd, orc):
    with amd.Context(8, 4, 6, 1.0, 1.0, 1.0, 0.02) as ctx:
        g = orc.Grid(8, 4, 6, 1.0, 1.0, 1.0, 0.02)
        w = _fields(g.shape, 6)
        assert _rel(ctx.expfield(w), orc.expfield(g, w)) < 1e-15


@pytest.mark.parametrize("m,it", [(4, 3), (4, 5), (10, 11), (20, 25)])
def test_mixing_anderson(amd, orc, m, it):
    rng = np.random.default_rng(m * 100 + it)
    n = 3000
    F = rng.standard_normal((m + 1, n))
    D = rng.standard_normal((m + 1, n))
    with amd.Context(4, 4, 4, 1.0, 1.0, 1.0, 0.01) as ctx:
        f1, d1 = F.copy(), D.copy()
        ctx.mixing_anderson(f1, d1, 0.5, it)
    f2, d2 = F.copy(), D.copy()
    orc.mixing_anderson(f2, d2, 0.5, it)
    assert np.array_equal(d1, d2)
    assert _rel(f1, f2) < 1e-11  # random histories are well conditioned


@pytest.mark.parametrize("n", [1, 2, 3, 8, 20, 21, 33, 40])
def test_device_lu_matches_host_lu_bitwise(amd, orc, n):
    """The wavefront LU of the Anderson step gives the bits of the host ludcmp/lubksb (lu.hpp),
    which the oracle's ludcmp matches bit for bit: general, symmetric Gram, badly scaled, equal-row
    (zero pivot -> 1e-20) and zero-row (singular) matrices, pivoting included."""
    rng = np.random.default_rng(100 + n)
    a0 = rng.uniform(-0.5, 0.5, (n, n))
    v = rng.standard_normal((n, 3 * n + 1))
    cases = [a0, v @ v.T, a0 * 10.0 ** rng.integers(-6, 7, (n, n))]
    if n > 2:
        eq = a0.copy()
        eq[2] = eq[0]
        zr = a0.copy()
        zr[1] = 0.0
        cases += [eq, zr]
    pivots = 0
    with amd.Context(4, 4, 4, 1.0, 1.0, 1.0, 0.01) as ctx:
        for a in cases:
            b = rng.uniform(-1.0, 1.0, n)
            x, lu, indx, sing = ctx.device_lu_solve(a, b)
            lu_h, indx_h, _ = amd.ludcmp(a)
            x_h = amd.solve_lu(a, b)
            lu_o, indx_o, _, sing_o = orc.ludcmp(a)
            assert np.array_equal(indx, indx_h) and np.array_equal(indx, indx_o)
            assert np.array_equal(lu, lu_h, equal_nan=True) and np.array_equal(lu, lu_o, equal_nan=True)
            assert np.array_equal(x, x_h, equal_nan=True)
            assert sing == bool(sing_o)
            pivots += int(np.sum(indx != np.arange(1, n + 1)))
    if n > 2:
        assert pivots > 0


def test_mixing_lambda_and_simple(amd, orc):
    rng = np.random.default_rng(9)
    F = rng.standard_normal((3, 500))
    D = rng.standard_normal((3, 500))
    with amd.Context(4, 4, 4, 1.0, 1.0, 1.0, 0.01) as ctx:
        for it in (1, 2):
            f1, d1 = F.copy(), D.copy()
            ctx.mixing_lambda(f1, d1, 2.0, it)
            f2, d2 = F.copy(), D.copy()
            orc.mixing_lambda(f2, d2, 2.0, it)
            assert _rel(f1, f2) < 1e-13 and np.array_equal(d1, d2)
        f1, d1 = F.copy(), D.copy()
        ctx.mixing_simple(f1, d1, 0.3)
        f2, d2 = F.copy(), D.copy()
        orc.mixing_simple(f2, d2, 0.3)
        assert _rel(f1, f2) < 1e-15


def _scf_case(orc, n=8, chiN=12.0):
    g = orc.Grid(n, n, n, 3.5, 3.5, 3.5, 0.05)
    rng = np.random.default_rng(1)
    x = np.arange(n) * 3.5 / n
    w0 = np.stack([0.5 * chiN * (np.cos(2 * np.pi * x / 3.5)[None, None, :] + rng.normal(0, 0.1, g.shape))] * 2)
    w0[1] *= -1
    chi = np.array([[0.0, chiN], [chiN, 0.0]])
    return g, w0, chi


def test_scf_trajectory_matches_oracle(amd, orc):
    # Up to the first Anderson solve the trajectory is fixed to rounding; that solve is
    # ill-conditioned (see test_oracle.py), so the comparison stops at the residual it starts from.
    g, w0, chi = _scf_case(orc)
    m = 3
    fh = orc.FH(25.0, chi, (1.0, 1.0))
    ref = orc.scf_anderson(g, [1, 2], [10, 10], fh, w0, m, lam=0.5, tol=0.0, maxit=m)
    ref2 = orc.scf_anderson(g, [1, 2], [10, 10], fh, w0, m, lam=0.5, tol=0.0, maxit=m + 1)
    with amd.Context(8, 8, 8, 3.5, 3.5, 3.5, 0.05) as ctx:
        ctx.set_interactions(25.0, chi, [1.0, 1.0])
        ctx.scf_setup([1, 2], [10, 10], ensemble=1, rho=1.0, mixing_dim=m, lam_mix=0.5)
        ctx.scf_set_field(w0)
        st = ctx.scf_iterate(m, tol=0.0)
        assert st.iters == m and not st.converged
        f_gpu, den_gpu = ctx.scf_get()
        assert abs(st.accuracy - ref["residuals"][-1]) < 1e-10 * ref["residuals"][-1]
        assert _rel(f_gpu, ref["field"]) < 1e-10
        st = ctx.scf_iterate(1, tol=0.0)
        assert st.iters == m + 1
        assert abs(st.accuracy - ref2["residuals"][-1]) < 1e-10 * ref2["residuals"][-1]


def test_scf_converges_to_oracle_saddle_point(amd, orc):
    g, w0, chi = _scf_case(orc)
    fh = orc.FH(25.0, chi, (1.0, 1.0))
    ref = orc.scf_anderson(g, [1, 2], [10, 10], fh, w0, 10, lam=0.1, tol=1e-8, maxit=2000)
    den_ref, info = orc.species_densities(g, [1, 2], [10, 10], ref["field"])
    F_ref = orc.total_energy(g, info, fh, den_ref, ref["field"])
    with amd.Context(8, 8, 8, 3.5, 3.5, 3.5, 0.05) as ctx:
        ctx.set_interactions(25.0, chi, [1.0, 1.0])
        ctx.scf_setup([1, 2], [10, 10], ensemble=1, rho=1.0, mixing_dim=10, lam_mix=0.1)
        ctx.scf_set_field(w0)
        st = ctx.scf_iterate(2000, tol=1e-8)
        assert st.converged and st.accuracy <= 1e-8
        en = ctx.scf_energy()
        f_gpu, den = ctx.scf_get()
    # F is stationary at the saddle point, so two runs converged to 1e-8 agree to 1e-10
    assert abs(en["F"] - F_ref) < FP64_RTOL * abs(F_ref)
    # the GPU densities and F are those of its own final field to the FP64 bar
    den_o, info_o = orc.species_densities(g, [1, 2], [10, 10], f_gpu)
    assert _rel(den, den_o) < FP64_RTOL
    assert abs(en["F"] - orc.total_energy(g, info_o, fh, den_o, f_gpu)) < FP64_RTOL * abs(F_ref)
    # the two final states differ by the convergence tolerance times the density response
    assert np.max(np.abs(den - den_ref)) < 1e-6


def test_scf_continues_after_convergence(amd, orc):
    """A loop that stopped at a loose tolerance continues under a tighter one as if it had
    never stopped (same iterations, fields and residuals, bit for bit), and the FP32 phase
    followed by an FP64 refinement converges to the oracle's saddle point."""
    g, w0, chi = _scf_case(orc)

    def ctx_for(prec):
        ctx = amd.Context(8, 8, 8, 3.5, 3.5, 3.5, 0.05, precision=prec)
        ctx.set_interactions(25.0, chi, [1.0, 1.0])
        ctx.scf_setup([1, 2], [10, 10], ensemble=1, rho=1.0, mixing_dim=10, lam_mix=0.1)
        ctx.scf_set_field(w0)
        return ctx

    with ctx_for(amd.FP64) as a, ctx_for(amd.FP64) as b:
        s1 = a.scf_iterate(2000, tol=1e-4)
        assert s1.converged and s1.accuracy <= 1e-4
        again = a.scf_iterate(2000, tol=1e-3)  # already below: returns at once
        assert again.converged and again.iters == s1.iters
        s2 = a.scf_iterate(2000, tol=1e-9)
        ref = b.scf_iterate(2000, tol=1e-9)
        assert s2.converged and s2.iters == ref.iters > s1.iters and s2.accuracy == ref.accuracy
        fa, pa = a.scf_get()
        fb, pb = b.scf_get()
        assert np.array_equal(fa, fb) and np.array_equal(pa, pb)
        assert np.array_equal(a.scf_residuals(), b.scf_residuals())
        # another loop (here the accuracy sum|dfield| dvol) starts its count at 1 on the
        # current field, as a new saddle-point loop of the reference does
        a.scf_set_residual(kind=0, measure=1)
        s3 = a.scf_iterate(1, tol=-1.0)
        assert s3.iters == 1 and s3.accuracy != s2.accuracy and not s3.converged
    with ctx_for(amd.FP32) as c:
        s32 = c.scf_iterate(2000, tol=1e-4)
        assert s32.converged
        c.set_precision(amd.FP64)
        s64 = c.scf_iterate(2000, tol=1e-9)
        assert s64.converged and s64.iters > s32.iters and s64.accuracy <= 1e-9
        f, _ = c.scf_get()
    fh = orc.FH(25.0, chi, (1.0, 1.0))
    den, _ = orc.species_densities(g, [1, 2], [10, 10], f)
    assert np.max(np.abs(orc.mean_fields(g, fh, den) - f)) < 2e-9


@pytest.mark.parametrize("n,shape", [(32, (4, 8)), (32, (8, 4)), (32, (2, 16)), (48, (8, 6)), (48, (6, 8)),
                                     (48, (16, 3)), (64, (8, 8)), (64, (4, 16)), (64, (16, 4)), (24, (8, 3)),
                                     (24, (4, 6)), (128, (8, 16)), (128, (16, 8)), (16, (1, 16)), (64, (0, 0)),
                                     (64, (4, 0)), (128, (4, -8)), (256, (4, -16))])
@pytest.mark.parametrize("waves", [1, 2, 4])
def test_every_schedule_matches_oracle(amd, orc, n, shape, waves):
    if shape[0] == 0 and waves > 1:
        pytest.skip("the Stockham kernel has no wave setting")
    dims = (n, 6, 10) if n >= 64 else (n, n // 2, 12)
    print(f"schedule n={n} shape={shape} waves={waves}", flush=True)
    with amd.Context(*dims, 3.0, 2.5, 3.5, 0.01) as ctx:
        ctx.set_schedule(0, shape[0], shape[1], waves)
        g = orc.Grid(*dims, 3.0, 2.5, 3.5, 0.01)
        w = _fields(g.shape, 7, amp=2.0)
        E = orc.expfield(g, w)
        ref = orc.propagate_chain(g, [1, 2], [6, 9], E)
        got = ctx.propagate_chain([1, 2], [6, 9], E)
        assert _rel(got.density, ref.density) < FP64_RTOL
        assert abs(got.Qsum - ref.Qsum) < FP64_RTOL * ref.Qsum


@pytest.mark.parametrize("shape,tiles", [((4, 0), 2), ((4, 0), 4), ((8, 8), 2), ((16, 4), 4), ((4, 16), 2)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_pass_tiles_bitwise(amd, shape, tiles, prec):
    """Passes that take 2 or 4 tiles per workgroup (the loads of the next tile issued before
    the current one is transformed) give the one-tile results bit for bit, on every axis of a
    64 x 48 x 64 grid (x: contiguous lines, y/z: strided; 48 keeps its default shape)."""
    P = amd.FP64 if prec == "fp64" else amd.FP32
    dims = (64, 48, 64)
    print(f"pass tiles shape={shape} tiles={tiles} {prec}", flush=True)
    rng = np.random.default_rng(13)
    E = np.exp(-0.005 * 2.0 * rng.standard_normal((2, dims[2], dims[1], dims[0])))
    out = []
    for k in (1, tiles):
        with amd.Context(*dims, 3.0, 2.5, 3.5, 0.01, precision=P) as ctx:
            for axis in (0, 2):
                ctx.set_schedule(axis, shape[0], shape[1], 1)
                ctx.set_pass_tiles(axis, k)
            ctx.set_pass_tiles(1, k)
            r = ctx.propagate_chain([1, 2], [13, 8], E)
            out.append((r.density, r.Qsum))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]


@pytest.mark.parametrize("axis", [1, 2])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tri_engine_on_strided_axes(amd, orc, axis, prec):
    """The 8 x 4 x 4 engine of 128-point lines on the y and z axes (lines strided in memory,
    tiles of adjacent x), FP64 against the oracle at 1e-10 and FP32-mixed at its 2e-5 bound."""
    dims = [6, 10, 8]
    dims[axis] = 128
    P = amd.FP64 if prec == "fp64" else amd.FP32
    print(f"tri8x4x4 axis={axis} dims={dims} {prec}", flush=True)
    with amd.Context(*dims, 3.0, 2.5, 3.5, 0.01, precision=P) as ctx:
        ctx.set_schedule(axis, 4, -8, 1)
        g = orc.Grid(*dims, 3.0, 2.5, 3.5, 0.01)
        E = orc.expfield(g, _fields(g.shape, 9, amp=2.0))
        ref = orc.propagate_chain(g, [1, 2], [6, 9], E)
        got = ctx.propagate_chain([1, 2], [6, 9], E)
        tol = FP64_RTOL if prec == "fp64" else 2e-5
        assert _rel(got.density, ref.density) < tol
        assert abs(got.Qsum - ref.Qsum) < tol * ref.Qsum


def test_autotune_keeps_parity(amd, orc):
    n, L, ds = 32, 4.0, 0.02
    g = orc.Grid(n, n, n, L, L, L, ds)
    w = _fields(g.shape, 8, amp=2.0)
    with amd.Context(n, n, n, L, L, L, ds) as ctx:
        ctx.set_interactions(25.0, [[0.0, 15.0], [15.0, 0.0]], [1.0, 1.0])
        ctx.scf_setup([1, 2], [25, 25], mixing_dim=4)
        ctx.scf_set_field(w)
        rep = ctx.autotune(reps=2)
        print(rep)
        assert "chosen" in rep and "middle axis chosen" in rep
        E = orc.expfield(g, w)
        got = ctx.propagate_chain([1, 2], [25, 25], E)
        # every choice of the single-pass axis gives the same propagator (the axis operators commute)
        mids = []
        for mid in (0, 1, 2):
            ctx.set_middle_axis(mid)
            mids.append(ctx.propagate_chain([1, 2], [25, 25], E))
    ref = orc.propagate_chain(g, [1, 2], [25, 25], E)
    assert _rel(got.density, ref.density) < FP64_RTOL
    for r in mids:
        assert _rel(r.density, ref.density) < FP64_RTOL and abs(r.Qsum - ref.Qsum) < FP64_RTOL * ref.Qsum


def test_diblock_saddle_loop_matches_oracle(orc):
    """Main_Diblock_Melt's own loop (main_diblock.f90:398-451) on the device: lambda mixing
    (lammax 0.1), the accuracy sum((density - newdensity)**2) against the previous iteration's
    densities (measure 2; the first iteration against the driver's guessed density), grand
    canonical f_A = 0.7 from the zero field.  Lambda mixing amplifies rounding slowly (see
    tests/test_gpu_blend.py), so the trajectory is compared after 40 iterations; the run to
    convergence by the oracle's evaluation of its final state."""
    import adscf_amd as amd

    nx, ny, nz, box, ds = 32, 1, 40, (3.6, 1.0, 4.5), 0.01
    bs = [orc.block_segments(0.7, ds), orc.block_segments(0.3, ds)]
    g = orc.Grid(nx, ny, nz, *box, ds)
    fh = orc.FH(25.0, [[0.0, 20.0], [20.0, 0.0]], [1.0, 1.0])
    mu, rho = 2.5, 1.0
    x = np.arange(1, nx + 1)
    den0 = np.empty((2, nz, ny, nx))
    den0[0] = (rho * 0.7 * (1.0 + 0.5 * np.cos(x / nx * np.pi)))[None, None, :]  # main_diblock.f90:128-129
    den0[1] = (rho * 0.3 * (1.0 - 0.5 * np.cos(x / nx * np.pi)))[None, None, :]
    w0 = np.zeros_like(den0)
    ref = orc.diblock_saddle(g, [1, 2], bs, fh, w0, den0, m=2, mixing_type=2, lam=0.1, tol=-1.0, maxit=40,
                             ensemble=2, rho=rho, mu=mu)
    with amd.Context(nx, ny, nz, *box, ds, precision=amd.FP64) as ctx:
        ctx.set_interactions(25.0, [[0.0, 20.0], [20.0, 0.0]], [1.0, 1.0])
        ctx.scf_setup([1, 2], bs, ensemble=2, rho=rho, mu=mu, mixing_dim=2, lam_mix=0.1)
        ctx.scf_set_mixing(2)
        ctx.scf_set_residual(kind=0, measure=2)
        ctx.scf_set_field(w0)
        ctx.scf_set_density(den0)
        st = ctx.scf_iterate(40, tol=-1.0)
        field, phi = ctx.scf_get()
        print(f"40 iterations: accuracy GPU {st.accuracy:.12e} oracle {ref['accuracy']:.12e}")
        assert st.iters == 40
        assert abs(st.accuracy - ref["accuracy"]) <= 1e-8 * ref["accuracy"]
        assert np.max(np.abs(field - ref["mixed_field"])) <= 1e-9 * np.max(np.abs(ref["mixed_field"]))
        st = ctx.scf_iterate(20000, tol=1e-12)
        field, phi = ctx.scf_get()
    assert st.converged
    # at exit the field is the one whose densities converged (the loop exits before mixing)
    d, _ = orc.species_densities(g, [1, 2], bs, field, ensemble=2, rho=rho, mu=mu)
    print(f"converged after {st.iters} iterations, accuracy {st.accuracy:.3e}")
    assert st.accuracy <= 1e-12
    assert np.max(np.abs(d - phi)) <= 1e-12 * np.max(d)


def test_hbm_bandwidth_probe(amd):
    """The bandwidth diagnostic returns plausible rates and rejects sizes under 1 MiB (512 MiB
    exceeds the 256 MiB last-level cache, so the rates are HBM rates below the 8 TB/s peak)."""
    with amd.Context(8, 8, 8, 1.0, 1.0, 1.0, 0.1) as ctx:
        cp, rd = ctx.hbm_bandwidth(512 << 20, 3)
        print(f"copy {cp:.0f} GB/s, read {rd:.0f} GB/s")
        assert 500.0 < cp < 8000.0 and 500.0 < rd < 8000.0
        with pytest.raises(amd.AdscfError):
            ctx.hbm_bandwidth(1000, 3)
