"""The oracle's contour rules, pinned by what the mathematics fixes.

Rule "trapezoid" is the reference's (Polymer_Modules/mod_multiblock.f90:157-168, pinned by
the chain tests of test_oracle.py).  Rule "simpson" (composite Simpson per block, a 3/8 panel
closing a block with an odd number of segments) has no counterpart in the reference, so its
weights are pinned by quadrature facts computed here without the oracle's formula:

* exactness: Simpson integrates every cubic in s exactly over a block and the trapezoid every
  line; s^4 (Simpson) and s^2 (trapezoid) are not integrated exactly;
* uniqueness: for 2 and 3 segments the weights exact on 1, s, s^2 (and s^3) are unique, so the
  Vandermonde solve gives them independently;
* Richardson: for an even block, composite Simpson = (4 T(ds) - T(2 ds)) / 3 with T the
  trapezoid; an odd block follows that on its first nb - 3 intervals and the 3-segment rule
  on its last three.

At chain level: both rules give the same canonical mass (the contour invariant makes every
slice pair sum to Q), the same densities in a homogeneous field (q(s) q+(Ns - s) does not
depend on s there), and the Simpson distribution equals the weighted sum of the oracle's own
propagator slices with the weights built here.
"""
import numpy as np
import pytest

DS = 0.1


def _moments_exact(w, nb, ds, k):
    s = np.arange(nb + 1) * ds
    return float(np.dot(w, s ** k)), (nb * ds) ** (k + 1) / (k + 1)


def _weights(orc, rule, nb, ds=DS):
    return np.array([orc.contour_weight(rule, j, nb, ds) for j in range(nb + 1)])


def _trapezoid(nb, h):
    w = np.full(nb + 1, h)
    w[0] = w[-1] = 0.5 * h
    return w


def _vandermonde(nb, ds):
    s = np.arange(nb + 1) * ds
    V = np.vstack([s ** k for k in range(nb + 1)])
    m = np.array([(nb * ds) ** (k + 1) / (k + 1) for k in range(nb + 1)])
    return np.linalg.solve(V, m)


def _richardson(nb, ds):
    """(4 T(ds) - T(2 ds)) / 3 on an even block."""
    t1 = _trapezoid(nb, ds)
    t2 = np.zeros(nb + 1)
    t2[::2] = _trapezoid(nb // 2, 2 * ds)
    return (4.0 * t1 - t2) / 3.0


def _simpson_independent(nb, ds):
    if nb == 0:
        return np.array([ds])
    if nb == 1:
        return _trapezoid(1, ds)
    if nb % 2 == 0:
        return _richardson(nb, ds)
    w = np.zeros(nb + 1)
    if nb > 3:
        w[: nb - 2] += _richardson(nb - 3, ds)
    w[nb - 3:] += _vandermonde(3, ds)
    return w


@pytest.mark.parametrize("nb", list(range(0, 16)))
def test_weight_sums(orc, nb):
    for rule in ("trapezoid", "simpson"):
        w = _weights(orc, rule, nb)
        assert abs(w.sum() - max(nb, 1) * DS) < 1e-14, (rule, nb)


@pytest.mark.parametrize("nb", list(range(2, 16)))
def test_simpson_exact_for_cubics_only(orc, nb):
    w = _weights(orc, "simpson", nb)
    for k in range(4):
        got, ref = _moments_exact(w, nb, DS, k)
        assert abs(got - ref) <= 1e-12 * ref, (nb, k)
    got, ref = _moments_exact(w, nb, DS, 4)
    assert abs(got - ref) > 1e-6 * ref


@pytest.mark.parametrize("nb", list(range(1, 16)))
def test_trapezoid_exact_for_lines_only(orc, nb):
    w = _weights(orc, "trapezoid", nb)
    for k in range(2):
        got, ref = _moments_exact(w, nb, DS, k)
        assert abs(got - ref) <= 1e-12 * ref
    got, ref = _moments_exact(w, nb, DS, 2)
    assert abs(got - ref) > 1e-6 * ref


@pytest.mark.parametrize("nb", [2, 3])
def test_simpson_unique_small_blocks(orc, nb):
    assert np.allclose(_weights(orc, "simpson", nb), _vandermonde(nb, DS), rtol=1e-12, atol=0.0)


@pytest.mark.parametrize("nb", list(range(0, 22)))
def test_simpson_matches_richardson_construction(orc, nb):
    np.testing.assert_allclose(_weights(orc, "simpson", nb), _simpson_independent(nb, DS), rtol=1e-12, atol=1e-16)


CHAINS = [
    ([1, 2], [6, 9], 1),
    ([1, 2, 1], [5, 8, 3], 1),
    ([1, 2, 1], [4, 0, 7], 1),
    ([2, 1], [1, 10], 2),
    ([1, 2, 1, 2], [2, 3, 2, 5], 1),
]


@pytest.mark.parametrize("chain", CHAINS)
def test_chain_simpson_assembles_its_own_slices(orc, chain):
    bt, bs, ens = chain
    g = orc.Grid(6, 5, 4, 2.0, 1.8, 1.6, 0.05)
    w = 1.5 * np.random.default_rng(7).standard_normal((2,) + g.shape)
    E = orc.expfield(g, w)
    trap = orc.propagate_chain(g, bt, bs, E, ensemble=ens, mu=0.2)
    orc.set_contour_rule("simpson")
    try:
        simp = orc.propagate_chain(g, bt, bs, E, ensemble=ens, mu=0.2, keep_q=True)
    finally:
        orc.set_contour_rule("trapezoid")
    ns = sum(bs)
    s0 = 0
    for b, n in enumerate(bs):
        wts = _simpson_independent(n, g.ds)
        ref = sum(wts[j] * simp.q[s0 + j] * simp.qdag[ns - s0 - j] for j in range(n + 1))
        np.testing.assert_allclose(simp.distribution[b], ref, rtol=1e-12, atol=0.0)
        s0 += n
    assert simp.Qsum == trap.Qsum
    # the contour invariant: dvol sum_r q(s) q+(Ns - s) = Q for every s, so each block's
    # dvol sum of the distribution is Q times its weight sum under either rule
    dvol = g.volume()[1]
    for b in range(len(bs)):
        assert abs(dvol * simp.distribution[b].sum() - dvol * trap.distribution[b].sum()) < 1e-11 * trap.Qsum
    assert not np.allclose(simp.distribution, trap.distribution, rtol=1e-9)


def test_homogeneous_field_rules_agree(orc):
    g = orc.Grid(4, 4, 4, 2.0, 2.0, 2.0, 0.05)
    w = np.stack([np.full(g.shape, 0.7), np.full(g.shape, -0.4)])
    E = orc.expfield(g, w)
    for bt, bs, ens in CHAINS:
        trap = orc.propagate_chain(g, bt, bs, E, ensemble=ens)
        orc.set_contour_rule("simpson")
        try:
            simp = orc.propagate_chain(g, bt, bs, E, ensemble=ens)
        finally:
            orc.set_contour_rule("trapezoid")
        np.testing.assert_allclose(simp.density, trap.density, rtol=1e-13, atol=0.0)


@pytest.mark.parametrize("segs", [1, 2, 7, 12])
def test_homopolymer_simpson_is_the_one_block_chain(orc, segs):
    g = orc.Grid(5, 4, 3, 2.0, 1.5, 1.2, 0.05)
    w = np.random.default_rng(3).standard_normal((2,) + g.shape)
    E = orc.expfield(g, w)
    orc.set_contour_rule("simpson")
    try:
        h = orc.propagate_homopolymer(g, 2, segs, E)
        c = orc.propagate_chain(g, [2], [segs], E)
    finally:
        orc.set_contour_rule("trapezoid")
    np.testing.assert_allclose(h.distribution, c.distribution[0], rtol=1e-12, atol=0.0)
    assert abs(h.Qsum - c.Qsum) < 1e-13 * c.Qsum
