"""Pins the oracle's DDFT restatement (orc_mobility_qq, orc_conjugate_fields, orc_gcfstat,
orc_ddft_step: Main_Multiblock/main_multiblock.f90:486-504, 566-640, 794-880) by what the
mathematics fixes, independently of the oracle's own code:

* a constant mobility table with every nonzero |q| inside the table's interior gives
  mobility_qq = L |q|^2, so one Euler step must change the density by
  timestep * sum_t2 L[t2][t1] laplacian(mulocal_t2), computed here with numpy's FFT;
  a non-symmetric L fixes the type order of the sum and the spectral factor fixes the
  transforms' normalisation;
* the q = 0 entry of mobility_qq is 0, so the step conserves the total density of every
  type when no floor is applied;
* a converged conjugate-field solve reproduces its target to the requested accuracy, in
  both accuracy measures, and a field that is already conjugate exits at iteration 1;
* gcfstat's branches at their thresholds.

The reference's own interpolation between table rows (its weights, main_multiblock.f90:497-499)
is pinned on the GPU side, where the unchanged Main_Multiblock driver computes the table with
the reference's Read_Mobilities and its DDFT result is compared with this restatement.
"""
import numpy as np
import pytest

NX, NY, NZ = 12, 1, 10
L = (3.3, 1.0, 2.9)
DS = 0.02
BT = [1, 2, 1]
BS = [10, 20, 20]


@pytest.fixture(scope="module")
def setup(orc):
    g = orc.Grid(NX, NY, NZ, *L, DS)
    fh = orc.FH(20.0, [[0.0, 12.0], [12.0, 0.0]], [1.0, 1.0])
    rng = np.random.default_rng(11)
    w = 0.8 * rng.standard_normal((2, NZ, NY, NX))
    den, info = orc.species_densities(g, BT, BS, w, ensemble=1, rho=1.0)
    return g, fh, w, den


def _k2(g):
    kx = 2 * np.pi * np.fft.rfftfreq(g.nx, d=1.0 / g.nx) / g.lx
    ky = 2 * np.pi * np.fft.fftfreq(g.ny, d=1.0 / g.ny) / g.ly
    kz = 2 * np.pi * np.fft.fftfreq(g.nz, d=1.0 / g.nz) / g.lz
    return kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2


def _constant_table(Lmat, nq, dq):
    """C view (t2, t1, iq) of a Fortran mobility_table(iq, t1, t2) = Lmat[t1][t2]."""
    t = np.empty((2, 2, nq))
    for a in range(2):
        for b in range(2):
            t[b, a, :] = Lmat[a][b]
    return t


def test_mobility_constant_table_is_L_q2(orc, setup):
    g = setup[0]
    Lmat = [[0.7, 0.2], [0.35, 0.5]]
    kmin = 2 * np.pi / max(g.lx, g.lz)
    dq = 0.5 * kmin
    nq = int(np.sqrt(_k2(g).max()) / dq) + 5
    mob = orc.mobility_qq(g, _constant_table(Lmat, nq, dq), dq, 1e300)
    k2 = _k2(g)
    for a in range(2):
        for b in range(2):
            # mobility_qq(:,:,:,a,b) is the C block [b, a]
            assert np.allclose(mob[b, a], Lmat[a][b] * k2, rtol=1e-13, atol=0.0)
    assert np.all(mob[:, :, 0, 0, 0] == 0.0)
    capped = orc.mobility_qq(g, _constant_table(Lmat, nq, dq), dq, 1.5)
    assert np.all(capped <= 1.5) and np.any(capped == 1.5)


def test_ddft_step_is_spectral_laplacian(orc, setup):
    """With a conjugate field (both solves exit at iteration 1) and no floor, the step is
    density += dt * sum_t2 L[t2][t1] laplacian(mulocal_t2), mobility_qq(..,t2,t1) = L[t2][t1] q^2."""
    g, fh, w, den0 = setup
    Lmat = [[0.7, 0.2], [0.35, 0.5]]
    dq = 0.5 * 2 * np.pi / max(g.lx, g.lz)
    nq = int(np.sqrt(_k2(g).max()) / dq) + 5
    mob = orc.mobility_qq(g, _constant_table(Lmat, nq, dq), dq, 1e300)
    den = den0.copy()
    field = w.copy()
    dt = 1e-3
    r = orc.ddft_step(g, BT, BS, fh, den, field, mob, dt, qdmin=False, accuracy_scf=1e3, itermax_scf=5, m=3)
    assert r["rc"] == 0 and r["iters"] == [1, 1] and r["gcfstat"] == [0, 0]
    assert np.array_equal(field, w)
    mu = orc.mean_fields(g, fh, den0) - w
    lap = np.stack([np.fft.irfftn(-_k2(g) * np.fft.rfftn(mu[t]), s=mu[t].shape, axes=(0, 1, 2)) for t in range(2)])
    expect = np.stack([sum(Lmat[t2][t1] * lap[t2] for t2 in range(2)) for t1 in range(2)])
    got = (den - den0) / dt
    assert np.max(np.abs(got - expect)) < 1e-9 * np.max(np.abs(expect))


def test_ddft_step_conserves_mass(orc, setup):
    g, fh, w, den0 = setup
    from adscf_amd import formats
    import os

    m = formats.read_mobilities(os.path.join(os.path.dirname(__file__), "golden", "multiblock_input_mobilities"))
    mob = orc.mobility_qq(g, m.table, m.dq, 10.0)
    assert np.all(mob[:, :, 0, 0, 0] == 0.0)
    den = den0.copy()
    field = w.copy()
    r = orc.ddft_step(g, BT, BS, fh, den, field, mob, 5e-3, qdmin=False, accuracy_scf=1e-3, itermax_scf=400, m=4)
    assert r["rc"] == 0 and r["gcfstat"] == [0, 0] and r["iters"][0] == 1
    assert np.max(np.abs(den - den0)) > 1e-3 * np.max(den0)  # the step moved the density
    for t in range(2):
        assert abs(den[t].sum() - den0[t].sum()) < 1e-12 * den0[t].sum()
    # the second solve made the field conjugate to the stepped density
    newden, _ = orc.species_densities(g, BT, BS, field, ensemble=1, rho=1.0)
    assert np.max(np.abs(newden - den)) <= 0.5e-3


@pytest.mark.parametrize("measure", [0, 1])
def test_conjugate_fields_reaches_target(orc, setup, measure):
    g, fh, w, den0 = setup
    vol, dvol = g.volume()
    rng = np.random.default_rng(5)
    # a reachable target: the densities of a perturbed field
    target, _ = orc.species_densities(g, BT, BS, w + 0.3 * rng.standard_normal(w.shape), ensemble=1, rho=1.0)
    thresh = 1e-6 if measure == 0 else 1e-6 * vol
    r = orc.conjugate_fields(g, BT, BS, target, w, m=5, thresh=thresh, maxit=3000, measure=measure)
    assert r["iters"] <= 3000
    newden, _ = orc.species_densities(g, BT, BS, r["field"], ensemble=1, rho=1.0)
    d = np.abs(newden - target)
    acc = d.max() if measure == 0 else d.sum() * dvol
    assert acc == pytest.approx(r["accuracy"], rel=1e-12)
    assert acc <= thresh
    # an already conjugate field exits at the first evaluation, unchanged
    r1 = orc.conjugate_fields(g, BT, BS, den0, w, m=5, thresh=1e-12, maxit=10, measure=measure)
    assert r1["iters"] == 1 and np.array_equal(r1["field"], w)


def test_gcfstat_branches(orc):
    acc_scf = 1e-3
    assert orc.gcfstat(0.49e-3, acc_scf) == 0
    assert orc.gcfstat(0.5e-3, acc_scf) == 1
    assert orc.gcfstat(0.99e-3, acc_scf) == 1
    assert orc.gcfstat(1e-3, acc_scf) == 2
    assert orc.gcfstat(0.0999, acc_scf) == 2
    assert orc.gcfstat(0.1, acc_scf) == 3
    assert orc.gcfstat(float("nan"), acc_scf) == 3


def test_conjugate_solve_first_anderson_step_amplifies_rounding(orc):
    """Get_Conjugate_Fields takes mixing_dim steps with lam00 = 0.001 before Anderson starts
    (mod_iterate.f90:94, 136-139), so the history the first Anderson solve sees holds nearly
    equal residuals and its Gram system is singular to rounding.  A 1e-16 relative change of
    the start field stays at rounding level through the simple-mixing steps and reaches the
    1e-3 level at iteration m+1: a property of the reference's algorithm that fixes how its
    results can be compared (exactly up to iteration m, by outcome afterwards)."""
    BT6, L6 = [1, 2, 1, 2, 1, 2], [0.5, 0.1, 0.1, 0.1, 0.1, 0.1]
    nx, ny, nz, box, ds = 16, 8, 12, (3.2, 1.6, 2.4), 0.02
    bs = [orc.block_segments(l, ds) for l in L6]
    g = orc.Grid(nx, ny, nz, *box, ds)
    rng = np.random.default_rng(21)
    w0 = 0.5 * rng.standard_normal((2, nz, ny, nx))
    target, _ = orc.species_densities(g, BT6, bs, w0 + 0.2 * rng.standard_normal(w0.shape), ensemble=1, rho=1.0)
    w1 = w0 * (1 + 1e-16 * np.random.default_rng(1).standard_normal(w0.shape))
    diff = []
    for n in (10, 11):
        a = orc.conjugate_fields(g, BT6, bs, target, w0, m=10, thresh=-1.0, maxit=n)
        b = orc.conjugate_fields(g, BT6, bs, target, w1, m=10, thresh=-1.0, maxit=n)
        diff.append(np.max(np.abs(a["field"] - b["field"])))
    assert diff[0] < 1e-14 and diff[1] > 1e-5


# ---- Main_Blend EPD (Chain_Dynamics_Timestep_Euler, main_blend.f90:605-667) ----------

def _blend_setup(orc, nx=32, ny=8, lx=10.0, ly=2.5):
    g = orc.Grid(nx, ny, 1, lx, ly, 1.0, 0.1)
    vol, dvol = g.volume()
    bl = orc.Blend([1, 2], [10, 10], [25.0 / vol, 25.0 / vol])
    fh = orc.FH(10.0, [[0.0, 10.0], [10.0, 0.0]], [1.0, 1.0])
    return g, vol, dvol, bl, fh


def _blend_energy(orc, g, dvol, bl, fh, den, field):
    orc.blend_densities(g, bl, field)  # chain energies of the field (Calculate_Energy, main_blend.f90:356-366)
    return -dvol * np.sum(den * field) + sum(bl.energy) + orc.interaction_energy(g, fh, den)


def test_epd_step_conserves_mass_and_lowers_free_energy(orc):
    """Chain dynamics is a gradient flow of F: with conjugate fields solved tightly, F falls
    at every small step and each type's mass stays; a sign slip in either nabla product would
    make F rise."""
    g, vol, dvol, bl, fh = _blend_setup(orc)
    field = orc.MT(2007).grnd(2 * g.ng).reshape(2, 1, g.ny, g.nx).copy()
    den = orc.blend_densities(g, bl, field)
    mass = den.sum(axis=(1, 2, 3))
    F = [_blend_energy(orc, g, dvol, bl, fh, den, field)]
    for _ in range(3):
        orc.epd_step(g, bl, fh, den, field, 0.001, accuracy_scf=1e-10)
        F.append(_blend_energy(orc, g, dvol, bl, fh, den, field))
        assert np.allclose(den.sum(axis=(1, 2, 3)), mass, rtol=1e-13, atol=0)
    assert all(b < a for a, b in zip(F, F[1:])), F
    # the returned field is conjugate to the new density
    d = orc.blend_densities(g, bl, field)
    assert np.sum(np.abs(d - den)) * dvol <= 1e-10


def test_epd_step_reduces_to_one_dimension(orc):
    """A state uniform along y evolves exactly as the same state on an (nx, 1, 1) grid: pins
    the per-axis wave numbers and spacings of the nabla products."""
    g, vol, dvol, bl, fh = _blend_setup(orc)
    g1 = orc.Grid(g.nx, 1, 1, g.lx, 1.0, 1.0, 0.1)
    bl1 = orc.Blend([1, 2], [10, 10], [25.0 / vol, 25.0 / vol])
    prof = orc.MT(2009).grnd(2 * g.nx).reshape(2, 1, 1, g.nx)
    f3, f1 = np.repeat(prof, g.ny, axis=2).copy(), prof.copy()
    d3, d1 = orc.blend_densities(g, bl, f3), orc.blend_densities(g1, bl1, f1)
    # two simple-mixing iterations of the solve (Anderson, m = 2, would start at the third)
    orc.epd_step(g, bl, fh, d3, f3, 0.001, accuracy_scf=-1.0, itermax=2)
    orc.epd_step(g1, bl1, fh, d1, f1, 0.001, accuracy_scf=-1.0, itermax=2)
    for y in range(g.ny):
        assert np.allclose(d3[:, :, y, :], d1[:, :, 0, :], rtol=0, atol=1e-12)
        assert np.allclose(f3[:, :, y, :], f1[:, :, 0, :], rtol=0, atol=1e-12)


def test_epd_step_uniform_state_is_stationary(orc):
    g, vol, dvol, bl, fh = _blend_setup(orc)
    field = np.empty((2, 1, g.ny, g.nx))
    field[0], field[1] = 0.3, -0.2
    den = orc.blend_densities(g, bl, field)
    d0, f0 = den.copy(), field.copy()
    it = orc.epd_step(g, bl, fh, den, field, 0.001)
    assert it == 1 and np.allclose(den, d0, rtol=1e-14, atol=0) and np.array_equal(field, f0)
