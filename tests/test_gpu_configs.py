"""The plain FP64 configuration at the oracle's size and the large mixed-precision
configuration, at full size.

* 32^3, box 4.0 Rg, Ns = 100 (50/50), f_A = 0.5, chiN = 15, kappa = 25, lamellar seed
  (eta ~ N(0, 0.1^2), MT19937 seed 2007), Anderson m = 10, lam_mix = 0.5, tolerance 1e-8
  on max|dfield|: the GPU's converged state is checked against the oracle evaluated at
  the GPU's own final field (densities and F to 1e-10, self-consistency to the tolerance),
  which is what the oracle can afford at this size in a test.
* 128^3, box 6.0 Rg, Ns = 400, ABA triblock 0.25/0.5/0.25 (segments 100/200/100),
  chiN = 25, BCC seed (eta ~ N(0, 0.1^2)), m = 20: the errors of phi, F and the box
  stress of the FP32-mixed path against FP64 at a field on the SCF path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_config_32cubed_fp64_saddle_point_matches_oracle(orc):
    import adscf_amd
    from adscf_amd import fields

    n, L, ns, chin = 32, 4.0, 100, 15.0
    segs = fields.block_segments([0.5, 0.5], ns)
    w0 = fields.lamellar_seed(n, L, chin)
    chi = [[0.0, chin], [chin, 0.0]]
    with adscf_amd.Context(n, n, n, L, L, L, 1.0 / ns) as ctx:
        ctx.set_interactions(25.0, chi, [1.0, 1.0])
        ctx.scf_setup([1, 2], segs, ensemble=1, rho=1.0, mixing_dim=10, lam_mix=0.5)
        ctx.scf_set_field(w0)
        st = ctx.scf_iterate(3000, tol=1e-8)
        en = ctx.scf_energy()
        f_gpu, den = ctx.scf_get()
    print(f"32^3 x 100 FP64: {st.iters} iterations, residual {st.accuracy:.2e}, F per chain {en['F_per_chain']:.15f}")
    assert st.converged and st.accuracy <= 1e-8
    g = orc.Grid(n, n, n, L, L, L, 1.0 / ns)
    fh = orc.FH(25.0, chi, (1.0, 1.0))
    den_o, info = orc.species_densities(g, [1, 2], segs, f_gpu)
    F_o = orc.total_energy(g, info, fh, den_o, f_gpu)
    assert np.max(np.abs(den - den_o)) < 1e-10 * np.max(den_o)
    assert abs(en["F"] - F_o) < 1e-10 * abs(F_o)
    resid = np.max(np.abs(orc.mean_fields(g, fh, den_o) - f_gpu))
    print(f"oracle residual at the GPU saddle point {resid:.2e}")
    assert resid < 1.01e-8


def test_config_128cubed_fp32_mixed_errors_against_fp64():
    """From the BCC seed, 300 FP64 SCF iterations (Anderson m = 20) move the field into the
    basin of the ordered state; at that field one FP64 and one FP32-mixed evaluation give
    phi, F per chain and the box stress dF/dL (adjoint with lam = 1 and a zero misfit), and
    the FP32-mixed errors are the stated ones.  (The reference's Anderson form converges
    slowly from this seed: max|dfield| is still ~1e-2 after 1200 iterations.)"""
    import adscf_amd
    from adscf_amd import fields

    n, L, ns, chin, m = 128, 6.0, 400, 25.0, 20
    segs = fields.block_segments([0.25, 0.5, 0.25], ns)
    assert segs == [100, 200, 100]
    w0 = fields.bcc_seed(n, L, chin)
    chi = [[0.0, chin], [chin, 0.0]]
    with adscf_amd.Context(n, n, n, L, L, L, 1.0 / ns) as ctx:
        ctx.set_interactions(25.0, chi, [1.0, 1.0])
        ctx.scf_setup([1, 2, 1], segs, ensemble=1, rho=1.0, mixing_dim=m, lam_mix=0.5)
        ctx.scf_set_field(w0)
        st = ctx.scf_iterate(300, tol=-1.0)
        w, _ = ctx.scf_get()
    print(f"128^3 x 400 ABA: field after {st.iters} FP64 iterations, residual {st.accuracy:.2e}", flush=True)
    out = {}
    for name, prec in (("fp64", adscf_amd.FP64), ("fp32mixed", adscf_amd.FP32)):
        with adscf_amd.Context(n, n, n, L, L, L, 1.0 / ns, precision=prec) as ctx:
            ctx.set_interactions(25.0, chi, [1.0, 1.0])
            fwd = ctx.adjoint([1, 2, 1], segs, w, np.zeros((n, n, n)), lam=1.0, gradients=False)
            adj = ctx.adjoint([1, 2, 1], segs, w, fwd["phi"][0], lam=1.0)
            out[name] = dict(phi=fwd["phi"], F=fwd["F"], stress=adj["grad_box"])
    a, b = out["fp64"], out["fp32mixed"]
    eF = abs(a["F"] - b["F"]) / abs(a["F"])
    ephi = float(np.max(np.abs(a["phi"] - b["phi"])) / np.max(np.abs(a["phi"])))
    est = float(np.max(np.abs(a["stress"] - b["stress"])) / np.max(np.abs(a["stress"])))
    print(f"128^3 x 400 FP32-mixed vs FP64 at that field: F {eF:.2e}, phi {ephi:.2e}, stress {est:.2e}; "
          f"stress FP64 {a['stress']}")
    assert eF < 1e-6 and ephi < 1e-4 and est < 1e-3


def test_config_128cubed_refinement_switch():
    """The FP64 refinement phase of the large configuration: the FP32-mixed loop (lam_mix 0.3)
    runs until max|dfield| < 3e-2, then adscf_set_precision(FP64) continues the same loop and
    history.  After the switch the loop computes in FP64: an evaluation of its field agrees
    with an independent FP64 evaluation of the same field to FP64 rounding (1e-12; the two
    contexts keep the line-FFT schedules of their starting precisions), where before the switch
    the two differ by FP32 rounding, and the iteration count carries on.  (The reference's Anderson form
    approaches the homogeneous state F = chiN/4 - 1 per chain from this seed too slowly to
    reach 1e-5 in a test: 2.0e-3 after 3000 iterations at lam_mix 0.3.)"""
    import adscf_amd
    from adscf_amd import fields

    n, L, ns, chin, m = 128, 6.0, 400, 25.0, 20
    segs = fields.block_segments([0.25, 0.5, 0.25], ns)
    chi = [[0.0, chin], [chin, 0.0]]

    def ctx_for(prec):
        ctx = adscf_amd.Context(n, n, n, L, L, L, 1.0 / ns, precision=prec)
        ctx.set_interactions(25.0, chi, [1.0, 1.0])
        ctx.scf_setup([1, 2, 1], segs, ensemble=1, rho=1.0, mixing_dim=m, lam_mix=0.3)
        return ctx

    with ctx_for(adscf_amd.FP32) as mixed:
        mixed.scf_set_field(fields.bcc_seed(n, L, chin))
        st = mixed.scf_iterate(2000, tol=3e-2)
        print(f"FP32-mixed phase: {st.iters} iterations to residual {st.accuracy:.3e}", flush=True)
        assert st.converged
        w, _ = mixed.scf_get()
        mixed.scf_evaluate()
        _, d32 = mixed.scf_get()
        mixed.set_precision(adscf_amd.FP64)
        mixed.scf_evaluate()
        _, d_after = mixed.scf_get()
        st2 = mixed.scf_iterate(20, tol=-1.0)
        assert st2.iters == st.iters + 20
    with ctx_for(adscf_amd.FP64) as ref:
        ref.scf_set_field(w)
        ref.scf_evaluate()
        _, d64 = ref.scf_get()
    e32 = float(np.max(np.abs(d32 - d64)) / np.max(d64))
    e64 = float(np.max(np.abs(d_after - d64)) / np.max(d64))
    print(f"densities at the switch field: FP32-mixed vs FP64 {e32:.2e}; after the switch {e64:.1e}")
    assert 0.0 < e32 < 1e-4
    assert e64 < 1e-12 and e64 < 1e-4 * e32
