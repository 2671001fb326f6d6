"""Pins the CPU oracle (oracle/scft_oracle.c) before it is trusted as the parity reference.

The reference ships no tests or golden vectors for this path, so each oracle
function is pinned here by what the mathematics fixes: brute-force DFT sums,
closed forms, exact invariants and an independent numpy restatement
(oracle/np_ref.py).  Each check is chosen so that a plausible slip -- a wrong
sign or normalisation, a transposed axis, an off-by-one contour index, a
missing half weight -- fails it.
"""
import numpy as np
import pytest

from oracle import np_ref


def brute_dft(x, sign):
    n = len(x)
    j = np.arange(n)
    M = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
    return M @ x


def brute_dft3(a, sign):
    """O(Ng^2) 3-D DFT of a (nz,ny,nx) array, x fastest."""
    nz, ny, nx = a.shape
    out = np.zeros(a.shape, dtype=np.complex128)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx):
                ph = np.exp(sign * 2j * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz))
                out[kz, ky, kx] = np.sum(a * ph)
    return out


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 43, 64, 75, 86])
def test_dft_line_bruteforce(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        got = orc.dft_line(x, sign)
        ref = brute_dft(x, sign)
        assert np.max(np.abs(got - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("dims", [(4, 4, 4), (6, 5, 3), (3, 4, 5), (4, 1, 6)])
def test_real2fourier_bruteforce(orc, dims):
    nx, ny, nz = dims
    g = orc.Grid(nx, ny, nz, 1.0, 1.0, 1.0, 0.01)
    a = np.random.default_rng(7).standard_normal(g.shape)
    full = brute_dft3(a, -1) / g.ng  # forward normalised by Ng (mod_fourier.f90:102)
    got = orc.real2fourier(g, a)
    assert got.shape == (nz, ny, nx // 2 + 1)  # x is the halved axis (mod_fourier.f90:21)
    assert np.max(np.abs(got - full[:, :, : nx // 2 + 1])) < 1e-14


def test_fft_known_answers(orc):
    g = orc.Grid(8, 6, 4, 1.0, 1.0, 1.0, 0.01)
    # delta at the origin -> flat spectrum 1/Ng
    d = np.zeros(g.shape)
    d[0, 0, 0] = 1.0
    assert np.allclose(orc.real2fourier(g, d), 1.0 / g.ng, atol=1e-16)
    # single x mode cos(2*pi*m*x/nx) -> 0.5 in bin kx=m only
    x = np.arange(8)
    a = np.broadcast_to(np.cos(2 * np.pi * 3 * x / 8), g.shape).copy()
    c = orc.real2fourier(g, a)
    exp = np.zeros(g.kshape, dtype=complex)
    exp[0, 0, 3] = 0.5
    assert np.max(np.abs(c - exp)) < 1e-15
    # single y mode -> bins (0, m, 0) and (0, ny-m, 0)
    y = np.arange(6)
    a = np.broadcast_to(np.cos(2 * np.pi * 1 * y / 6)[:, None], g.shape).copy()
    c = orc.real2fourier(g, a)
    exp = np.zeros(g.kshape, dtype=complex)
    exp[0, 1, 0] = exp[0, 5, 0] = 0.5
    assert np.max(np.abs(c - exp)) < 1e-15
    # c2r is unnormalised: one bin at kx=1 -> 2*cos(2*pi*x/nx) (Hermitian partner implied)
    s = np.zeros(g.kshape, dtype=complex)
    s[0, 0, 1] = 1.0
    r = orc.fourier2real(g, s)
    assert np.max(np.abs(r - 2 * np.cos(2 * np.pi * x / 8)[None, None, :])) < 1e-14


@pytest.mark.parametrize("dims", [(8, 6, 10), (16, 16, 16), (86, 1, 75), (64, 1, 30), (9, 5, 7)])
def test_fft_matches_pocketfft_and_roundtrips(orc, dims):
    nx, ny, nz = dims
    g = orc.Grid(nx, ny, nz, 2.0, 3.0, 4.0, 0.01)
    a = np.random.default_rng(3).standard_normal(g.shape)
    c = orc.real2fourier(g, a)
    assert np.max(np.abs(c - np_ref.r2c(a))) < 1e-14
    b = orc.fourier2real(g, c)
    assert np.max(np.abs(b - a)) < 1e-12
    assert np.max(np.abs(b - np_ref.c2r(c, g.shape))) < 1e-12


def test_laplace_index_convention(orc):
    # nx=6 (x not folded), ny=4, nz=5; 1-based index y > ny/2+1 folds (mod_fourier.f90:181-186)
    lx, ly, lz, ds = 2.0, 3.0, 5.0, 0.02
    g = orc.Grid(6, 4, 5, lx, ly, lz, ds)
    lap, el, nab = orc.calculate_laplace(g)
    tp = 2 * np.pi
    # y indices 0..3 -> 0, 1, 2 (Nyquist stays positive), -1 ; z 0..4 -> 0, 1, 2, -2, -1
    qy = tp * np.array([0, 1, 2, -1]) / ly
    qz = tp * np.array([0, 1, 2, -2, -1]) / lz
    qx = tp * np.arange(4) / lx
    K2 = qx[None, None, :] ** 2 + qy[None, :, None] ** 2 + qz[:, None, None] ** 2
    assert np.allclose(lap.real, -K2, rtol=1e-15) and np.all(lap.imag == 0)
    assert np.allclose(el.real, np.exp(-ds * K2), rtol=1e-15) and np.all(el.imag == 0)
    dy = ly / 4
    assert np.allclose(nab[1].imag[0, :, 0], np.sin(qy * dy) / dy, atol=1e-15)
    assert nab[1].imag[0, 3, 0] < 0 and nab[1].imag[0, 1, 0] > 0
    dz = lz / 5
    assert np.allclose(nab[2].imag[:, 0, 0], np.sin(qz * dz) / dz, atol=1e-15)
    assert np.all(nab.real == 0)


def test_step_closed_forms(orc):
    g = orc.Grid(8, 8, 8, 3.0, 3.0, 3.0, 0.01)
    ones = np.ones(g.shape)
    # w = 0 -> q stays 1
    assert np.max(np.abs(orc.propagate_step(g, ones, ones) - 1.0)) < 1e-14
    # constant w = c -> q *= exp(-c ds) exactly (E and K commute)
    c = 1.7
    E = np.full(g.shape, np.exp(-0.5 * g.ds * c))
    q = orc.propagate_step(g, ones, E)
    assert np.allclose(q, np.exp(-c * g.ds), rtol=1e-14)
    # single Fourier mode with w=0 decays by exp(-k^2 ds)
    z = np.arange(8)
    mode = np.broadcast_to(np.cos(2 * np.pi * 2 * z / 8)[:, None, None], g.shape).copy()
    q = orc.propagate_step(g, mode, ones)
    k = 2 * np.pi * 2 / 3.0
    assert np.max(np.abs(q - np.exp(-k * k * g.ds) * mode)) < 1e-14


def _step1d(f, e, L, ds):
    n = len(f)
    j = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(j, j) / n) / n
    Fi = np.exp(2j * np.pi * np.outer(j, j) / n)
    kk = np.where(j + 1 > n // 2 + 1, j - n, j) * 2 * np.pi / L
    return (e * (Fi @ (np.exp(-ds * kk * kk) * (F @ (e * f))))).real


def test_step_separable_factorises(orc):
    # w = a(x)+b(y)+c(z), q0 = f(x)g(y)h(z): the 3-D step is the product of three 1-D steps
    nx, ny, nz, L, ds = 8, 6, 5, (2.5, 3.0, 3.5), 0.05
    g = orc.Grid(nx, ny, nz, *L, ds)
    rng = np.random.default_rng(11)
    a, b, c = rng.standard_normal(nx), rng.standard_normal(ny), rng.standard_normal(nz)
    f, gg, h = 1 + 0.3 * rng.standard_normal(nx), 1 + 0.3 * rng.standard_normal(ny), 1 + 0.3 * rng.standard_normal(nz)
    w = a[None, None, :] + b[None, :, None] + c[:, None, None]
    q0 = f[None, None, :] * gg[None, :, None] * h[:, None, None]
    q = orc.propagate_step(g, q0, np.exp(-0.5 * ds * w))
    ex = _step1d(f, np.exp(-0.5 * ds * a), L[0], ds)
    ey = _step1d(gg, np.exp(-0.5 * ds * b), L[1], ds)
    ez = _step1d(h, np.exp(-0.5 * ds * c), L[2], ds)
    ref = ex[None, None, :] * ey[None, :, None] * ez[:, None, None]
    assert np.max(np.abs(q - ref)) < 1e-13


def test_step_translation_equivariance(orc):
    g = orc.Grid(8, 6, 4, 2.0, 2.0, 2.0, 0.01)
    rng = np.random.default_rng(5)
    w = rng.standard_normal(g.shape)
    q0 = 1 + 0.2 * rng.standard_normal(g.shape)
    E = np.exp(-0.5 * g.ds * w)
    q = orc.propagate_step(g, q0, E)
    sh = (1, 2, 3)
    qs = orc.propagate_step(g, np.roll(q0, sh, axis=(0, 1, 2)), np.roll(E, sh, axis=(0, 1, 2)))
    assert np.max(np.abs(qs - np.roll(q, sh, axis=(0, 1, 2)))) < 1e-14


def _random_chain_setup(orc, dims=(8, 6, 10), L=(3.0, 2.5, 3.5), ds=0.02, seed=0, amp=2.0):
    g = orc.Grid(*dims, *L, ds)
    w = amp * np.random.default_rng(seed).standard_normal((2,) + g.shape)
    return g, w


def test_chain_contour_invariant_and_mass(orc):
    # dvol * sum_r q(r,s) q+(r,Ns-s) = Q for every s; sum of trapezoid weights = ds*Ns
    g, w = _random_chain_setup(orc)
    bt, bs = [1, 2, 1], [7, 11, 5]
    r = orc.propagate_chain(g, bt, bs, orc.expfield(g, w), ensemble=1, rho=0.8, keep_q=True)
    vol, dvol = g.volume()
    ns = sum(bs)
    for s in range(ns + 1):
        inv = dvol * np.sum(r.q[s] * r.qdag[ns - s])
        assert abs(inv - r.Qsum) < 1e-12 * r.Qsum
    assert abs(dvol * r.density.sum() - 0.8 * vol * g.ds * ns) < 1e-11 * vol
    assert abs(r.mu - np.log(0.8 * vol / r.Qsum)) < 1e-14
    assert abs(r.energy - 0.8 * vol * (r.mu - 1.0)) < 1e-12 * abs(r.energy)


def test_chain_reversal_symmetry(orc):
    # reversing the block sequence exchanges q and q+: same Q, block densities reversed
    g, w = _random_chain_setup(orc, seed=3)
    E = orc.expfield(g, w)
    a = orc.propagate_chain(g, [1, 2, 2], [6, 9, 4], E)
    b = orc.propagate_chain(g, [2, 2, 1], [4, 9, 6], E)
    assert abs(a.Qsum - b.Qsum) < 1e-13 * a.Qsum
    assert np.max(np.abs(a.density - b.density[::-1])) < 1e-12 * np.max(a.density)


def test_chain_constant_field(orc):
    g = orc.Grid(6, 4, 8, 2.0, 1.5, 3.0, 0.01)
    cA, cB = 0.7, -0.4
    w = np.stack([np.full(g.shape, cA), np.full(g.shape, cB)])
    r = orc.propagate_chain(g, [1, 2], [30, 70], orc.expfield(g, w))
    vol, _ = g.volume()
    assert abs(r.Qsum - vol * np.exp(-(cA * 30 + cB * 70) * g.ds)) < 1e-13 * r.Qsum


def test_chain_grand_canonical(orc):
    g, w = _random_chain_setup(orc, seed=9)
    E = orc.expfield(g, w)
    r = orc.propagate_chain(g, [1, 2], [10, 15], E, ensemble=2, mu=0.3)
    vol, dvol = g.volume()
    assert np.allclose(r.density, np.exp(0.3) * r.distribution, rtol=1e-15)
    assert abs(r.rho - r.density.sum() * dvol / vol / (g.ds * 25)) < 1e-15
    assert r.energy == -r.rho * vol
    with pytest.raises(ValueError):
        orc.propagate_chain(g, [1, 2], [10, 15], E, ensemble=3)


@pytest.mark.parametrize("ensemble", [1, 2])
def test_chain_matches_numpy_restatement(orc, ensemble):
    g, w = _random_chain_setup(orc, dims=(12, 10, 8), seed=21, amp=3.0)
    bt, bs = [1, 2, 1, 2], [5, 8, 3, 9]
    r = orc.propagate_chain(g, bt, bs, orc.expfield(g, w), ensemble=ensemble, rho=1.0, mu=0.2)
    n = np_ref.chain(w, bt, bs, (g.lx, g.ly, g.lz), g.ds, ensemble=ensemble, rho=1.0, mu=0.2)
    assert abs(r.Qsum - n["Qsum"]) < 1e-13 * n["Qsum"]
    assert np.max(np.abs(r.density - n["density"])) < 1e-12 * np.max(n["density"])
    assert abs(r.energy - n["energy"]) < 1e-12 * abs(n["energy"])


def test_block_segments_nint(orc):
    assert orc.block_segments(0.3, 0.005) == 60
    assert orc.block_segments(0.7, 0.001) == 700
    assert orc.block_segments(0.25, 0.1) == 3  # 2.5 -> 3: NINT rounds half away from zero
    assert orc.block_segments(0.05, 0.1) == 1  # 0.5 -> 1


def test_flory_huggins_closed_form(orc):
    g = orc.Grid(4, 3, 5, 1.0, 2.0, 3.0, 0.01)
    rng = np.random.default_rng(2)
    den = rng.uniform(0.1, 0.9, (2,) + g.shape)
    kappa, chi, v = 25.0, 17.0, (1.3, 0.8)
    fh = orc.FH(kappa, [[0.0, chi], [chi, 0.0]], v)
    f = orc.mean_fields(g, fh, den)
    pA, pB = den[0] * v[0], den[1] * v[1]
    tot = pA + pB - 1
    assert np.allclose(f[0], v[0] * (kappa * tot + chi * pB), rtol=1e-15)
    assert np.allclose(f[1], v[1] * (kappa * tot + chi * pA), rtol=1e-15)
    _, dvol = g.volume()
    e = orc.interaction_energy(g, fh, den)
    assert abs(e - (0.5 * kappa * dvol * np.sum(tot ** 2) + chi * dvol * np.sum(pA * pB))) < 1e-12 * abs(e)
    assert abs(orc.average_volume_fraction(g, fh, den) - np.mean(pA + pB)) < 1e-15
    assert np.allclose(f, np_ref.mean_fields(den, kappa, np.array([[0, chi], [chi, 0]]), np.array(v)), rtol=1e-14)


def test_lu_matches_lapack_and_pivots(orc):
    rng = np.random.default_rng(4)
    for n in (1, 2, 5, 10, 20):
        a = rng.standard_normal((n, n)) + n * np.eye(n) * 0.1
        b = rng.standard_normal(n)
        x = orc.solve_lu(a, b)
        assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)
        spd = a @ a.T + n * np.eye(n)
        assert np.allclose(orc.solve_cholesky(spd, b), np.linalg.solve(spd, b), rtol=1e-12)
    # implicit scaling: plain partial pivoting would keep row 1 (|2| > |1|), but scaled by its
    # largest entry (1000) row 1 loses to row 2, so the rows are swapped (mod_iterate.f90:241-283)
    lu, indx, d, sing = orc.ludcmp(np.array([[2.0, 1000.0], [1.0, 1.0]]))
    assert list(indx) == [2, 2] and d == -1.0 and not sing
    lu, indx, d, sing = orc.ludcmp(np.array([[1.0, 2.0], [3.0, 4.0]]))
    assert list(indx) == [2, 2] and d == -1.0
    assert np.allclose(lu, [[3.0, 4.0], [1.0 / 3.0, 2.0 - 4.0 / 3.0]])
    # a zero row is reported; duplicate rows hit the 1e-20 tiny-pivot substitution
    assert orc.ludcmp(np.array([[0.0, 0.0], [1.0, 2.0]]))[3]
    lu, indx, d, sing = orc.ludcmp(np.array([[1.0, 2.0], [1.0, 2.0]]))
    assert lu[1, 1] == 1e-20


def test_anderson_known_answer_and_numpy(orc):
    rng = np.random.default_rng(8)
    m, n = 4, 50
    F = rng.standard_normal((m + 1, n))
    D = rng.standard_normal((m + 1, n))
    # iter <= m: simple step with lam00 = 0.001, history shifted (mod_iterate.f90:136-148)
    f, d = F.copy(), D.copy()
    orc.mixing_anderson(f, d, 0.5, m)
    assert np.allclose(f[0], F[0] + 0.001 * D[0], rtol=1e-15)
    assert np.array_equal(f[1:], F[:-1]) and np.array_equal(d[1:], D[:-1])
    # iter > m: full Anderson against the numpy restatement with a LAPACK solve
    f, d = F.copy(), D.copy()
    orc.mixing_anderson(f, d, 0.5, m + 1)
    f2, d2 = F.copy(), D.copy()
    np_ref.anderson(f2, d2, 0.5, m + 1)
    assert np.allclose(f, f2, rtol=1e-10, atol=1e-12)
    assert np.array_equal(d, d2)
    # the mixed residual is the least-squares minimiser over the history's affine span
    dd = D[0][None, :] - D[1:]
    x = np.linalg.lstsq(dd.T, D[0], rcond=None)[0]
    dmix = D[0] - x @ dd
    assert np.allclose(f[0], F[0] - x @ (F[0][None, :] - F[1:]) + 0.5 * dmix, rtol=1e-9)


def test_lambda_and_simple_mixing(orc):
    rng = np.random.default_rng(12)
    F = rng.standard_normal((3, 20))
    D = rng.standard_normal((3, 20))
    f, d = F.copy(), D.copy()
    orc.mixing_simple(f, d, 0.25)
    assert np.allclose(f[0], F[0] + 0.25 * D[0]) and np.array_equal(f[1:], F[1:])
    f, d = F.copy(), D.copy()
    orc.mixing_lambda(f, d, 0.1, 1)
    assert np.allclose(f[0], F[0] + 1e-5 * D[0]) and np.array_equal(f[1], F[0])
    f, d = F.copy(), D.copy()
    orc.mixing_lambda(f, d, 10.0, 2)
    lam = np.sqrt(np.sum((F[0] - F[1]) ** 2) / np.sum((D[0] - D[1]) ** 2))
    assert np.allclose(f[0], F[0] + lam * D[0], rtol=1e-14)


def test_mt19937_against_numpy(orc):
    # sgrnd seeding (69069 LCG, mod_mt.f90:71-87) fed into numpy's independent MT19937
    seed = 2007
    key = np.empty(624, dtype=np.uint64)
    key[0] = seed
    for i in range(1, 624):
        key[i] = (69069 * int(key[i - 1])) & 0xFFFFFFFF
    bg = np.random.MT19937()
    bg.state = {"bit_generator": "MT19937", "state": {"key": key.astype(np.uint32), "pos": 624}}
    ref = bg.random_raw(2000).astype(np.uint32)
    mt = orc.MT(seed)
    got = mt.next32(2000)
    assert np.array_equal(got, ref)
    mt2 = orc.MT(seed)
    assert mt2.grnd() == float(ref[0]) / (2.0 ** 32 - 1.0)


def _np_scf(field, bt, bs, L, ds, chi, m, lam, its, tol=0.0):
    """Find_SCF_Saddle_Point (main_multiblock.f90:648-700) restated with np_ref."""
    var = np.zeros((m + 1, field.size))
    dvar = np.zeros((m + 1, field.size))
    res = []
    for it in range(1, its + 1):
        r = np_ref.chain(field, bt, bs, L, ds)
        den = np_ref.species(r["density"], bt, 2)
        dfield = np_ref.mean_fields(den, 25.0, chi, np.array([1.0, 1.0])) - field
        res.append(np.max(np.abs(dfield)))
        if not res[-1] > tol:
            break
        var[0] = field.ravel()
        dvar[0] = dfield.ravel()
        np_ref.anderson(var, dvar, lam, min(it, 1000))
        field = var[0].reshape(field.shape)
    return field, den, r, np.array(res)


def _scf_setup(orc, chiN):
    g = orc.Grid(8, 8, 8, 3.5, 3.5, 3.5, 0.05)
    rng = np.random.default_rng(1)
    x = np.arange(8) * 3.5 / 8
    w0 = np.stack([0.5 * chiN * (np.cos(2 * np.pi * x / 3.5)[None, None, :] + rng.normal(0, 0.1, g.shape))] * 2)
    w0[1] *= -1
    chi = np.array([[0.0, chiN], [chiN, 0.0]])
    return g, w0, chi


def test_scf_anderson_trajectory_matches_numpy_loop(orc):
    # Up to the first Anderson solve the trajectory is fixed to rounding; afterwards the m x m
    # system built from lam00=0.001 steps is ill-conditioned (cond ~1e12 at m=3), so the two
    # restatements are compared there through their converged fixed point instead.
    g, w0, chi = _scf_setup(orc, 12.0)
    bt, bs = [1, 2], [10, 10]
    fh = orc.FH(25.0, chi, (1.0, 1.0))
    m = 3
    out = orc.scf_anderson(g, bt, bs, fh, w0, m, lam=0.5, tol=0.0, maxit=m + 1)
    assert out["iters"] == m + 2  # DO-loop index after running to completion
    _, _, _, res = _np_scf(w0.copy(), bt, bs, (3.5,) * 3, 0.05, chi, m, 0.5, m + 1)
    assert np.allclose(out["residuals"], res, rtol=1e-12)


def test_scf_anderson_converged_state_matches_numpy_loop(orc):
    g, w0, chi = _scf_setup(orc, 12.0)
    bt, bs = [1, 2], [10, 10]
    fh = orc.FH(25.0, chi, (1.0, 1.0))
    out = orc.scf_anderson(g, bt, bs, fh, w0, 10, lam=0.1, tol=1e-8, maxit=2000)
    assert out["accuracy"] <= 1e-8 and out["iters"] < 2000
    f2, den2, r2, res2 = _np_scf(w0.copy(), bt, bs, (3.5,) * 3, 0.05, chi, 10, 0.1, 2000, tol=1e-8)
    assert res2[-1] <= 1e-8
    den, info = orc.species_densities(g, bt, bs, out["field"])
    F1 = orc.total_energy(g, info, fh, den, out["field"])
    vol, dvol = g.volume()
    F2 = r2["energy"] - dvol * np.sum(den2 * f2) + np_ref.interaction_energy(den2, 25.0, chi, [1.0, 1.0], dvol)
    assert abs(F1 - F2) < 1e-10 * abs(F1)
    assert np.max(np.abs(den - den2)) < 1e-7  # both within the 1e-8 field residual of the saddle point


# ---- Propagate_Homopolymer and Update_Solvent (mod_homopolymer.f90, mod_solvent.f90) ----

def test_homopolymer_equals_one_block_chain(orc):
    """The single-propagator restatement (q+ = q) and the one-block multiblock restatement
    (separate forward and complementary sweeps) are two independent codings of the same
    sum; they must agree to rounding in both ensembles."""
    g = orc.Grid(8, 6, 5, 2.0, 1.7, 1.5, 0.05)
    rng = np.random.default_rng(11)
    E = orc.expfield(g, rng.standard_normal((2,) + g.shape))
    for ens, rho, mu in ((1, 0.8, 0.0), (2, 1.0, 0.4)):
        h = orc.propagate_homopolymer(g, 2, 13, E, ensemble=ens, rho=rho, mu=mu)
        c = orc.propagate_chain(g, [2], [13], E, ensemble=ens, rho=rho, mu=mu)
        assert np.max(np.abs(h.distribution - c.distribution[0])) < 1e-14 * np.max(c.distribution)
        assert np.max(np.abs(h.density - c.density[0])) < 1e-14 * np.max(c.density)
        for k in ("Qsum", "mu", "rho", "energy"):
            assert abs(getattr(h, k) - getattr(c, k)) <= 1e-14 * max(1.0, abs(getattr(c, k)))


def test_homopolymer_constant_field_closed_form(orc):
    """w = c: q(s) = exp(-c s ds), Q = V exp(-c N ds), distribution = N ds exp(-c N ds),
    canonical density rho N ds and mu = ln(rho V / Q)."""
    g = orc.Grid(6, 4, 5, 2.0, 1.5, 1.7, 0.1)
    c, n = 0.7, 9
    E = orc.expfield(g, np.full((1,) + g.shape, c))
    h = orc.propagate_homopolymer(g, 1, n, E, ensemble=1, rho=1.3)
    vol = 2.0 * 1.5 * 1.7
    assert abs(h.Qsum - vol * np.exp(-c * n * 0.1)) < 1e-13 * h.Qsum
    assert np.allclose(h.distribution, n * 0.1 * np.exp(-c * n * 0.1), rtol=1e-13, atol=0)
    assert np.allclose(h.density, 1.3 * n * 0.1, rtol=1e-13, atol=0)
    assert abs(h.mu - np.log(1.3 * vol / h.Qsum)) < 1e-14


def test_solvent_closed_forms(orc):
    """Update_Solvent: distribution = exp(-w); canonical density rho V exp(-w)/Q with
    Q = dvol sum exp(-w); grand canonical rho = dvol sum exp(mu - w) / V."""
    g = orc.Grid(6, 5, 4, 2.0, 1.5, 1.2, 0.1)
    rng = np.random.default_rng(5)
    w = rng.standard_normal(g.shape)
    vol, dvol = g.volume()
    s1 = orc.update_solvent(g, w, ensemble=1, rho=0.6)
    q = dvol * np.exp(-w).sum()
    assert np.allclose(s1.distribution, np.exp(-w), rtol=2e-16, atol=0)
    assert abs(s1.Qsum - q) < 1e-14 * q
    assert np.allclose(s1.density, 0.6 * vol * np.exp(-w) / q, rtol=1e-14, atol=0)
    assert abs(s1.energy - 0.6 * vol * (np.log(0.6 * vol / q) - 1.0)) < 1e-13
    s2 = orc.update_solvent(g, w, ensemble=2, mu=0.3)
    assert abs(s2.rho - dvol * np.exp(0.3 - w).sum() / vol) < 1e-14
    assert abs(s2.energy + s2.rho * vol) < 1e-13
    with pytest.raises(ValueError):
        orc.update_solvent(g, w, ensemble=3)
