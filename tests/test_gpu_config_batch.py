"""The 64-instance parameter sweep at full size (adscf_amd.fields.batched_instances, seed 2007):
16 instances on 32^3, 32 on 48^3 and 16 on 64^3 grids, Ns in {100, 150, 200}, chiN ~ U[12, 30],
f_A ~ U[0.2, 0.5], box ~ U[3.5, 5.0], Anderson m = 20, lam_mix = 0.5, each instance the
reference's Propagate_Diblock + Mixing_Anderson (Polymer_Modules/mod_diblock.f90:100-174,
General_Routines/mod_iterate.f90:92-150), all run by one batched graph replay per iteration.

* One batched iteration in FP64 and in FP32-mixed: every instance's canonical mass
  (dvol * sum(phi) = rho V, the trapezoid weights summing to ds Ns), and all sixteen 32^3
  instances plus the first 48^3 and the first 64^3 instance against the oracle's
  Find_SCF_Saddle_Point iteration from the same field (FP64 to 1e-10 relative; FP32-mixed
  to the stated bounds: densities 5e-5, mixed field 1e-7).
* 23 FP64 iterations (three Anderson solves past m = 20) of one instance per grid size and
  contour length, each bit for bit equal to a single-instance context on the same schedules:
  the batched passes, Gram sums, wavefront LU and mix are the single-instance ones; with the
  sweep as per-pass launches of xy-plane and z passes (the default), as per-pass line passes
  and as one persistent launch (line passes).
* The persistent sweep with the 48-point axes pinned (adscf_batch_set_schedule) to each shape
  it carries for that length, against the per-pass launches bit for bit; the refusal of
  shapes the batched kernels lack.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M, LAM = 20, 0.5


@pytest.fixture(scope="module")
def sweep():
    from adscf_amd import fields

    return fields.batched_instances(seed=2007)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _oracle_subset(inst):
    ids = list(range(16)) + [16, 48]
    assert [inst[i]["grid"][0] for i in ids] == [32] * 16 + [48, 64]
    return ids


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_sweep_one_iteration_against_oracle(orc, sweep, prec):
    import adscf_amd

    inst, w = sweep
    P = adscf_amd.FP64 if prec == "fp64" else adscf_amd.FP32
    print(f"64-instance sweep, one batched iteration, {prec}", flush=True)
    with adscf_amd.Batch(inst, mixing_dim=M, lam_mix=LAM, precision=P) as b:
        b.set_fields(w)
        it, acc, conv = b.iterate(1, tol=-1.0)
        got = [b.get(i) for i in range(len(inst))]
        F = b.energy()
    assert np.all(it == 1) and not conv.any()
    mass_tol = 1e-10 if prec == "fp64" else 1e-5
    for i, d in enumerate(inst):
        n = d["grid"][0]
        V = d["box"][0] ** 3
        mass = V / n ** 3 * got[i][1].sum()
        assert abs(mass - V) < mass_tol * V, (i, mass, V)
    phi_tol, field_tol = (1e-10, 1e-10) if prec == "fp64" else (5e-5, 1e-7)
    worst = [0.0, 0.0, 0.0]
    for i in _oracle_subset(inst):
        d = inst[i]
        n, L = d["grid"][0], d["box"][0]
        g = orc.Grid(n, n, n, L, L, L, d["ds"])
        fh = orc.FH(d["kappa"], d["chi"], d["vol"])
        ref = orc.scf_anderson(g, d["blocks"], d["segments"], fh, w[i], M, lam=LAM, tol=-1.0, maxit=1)
        ephi = _rel(got[i][1], ref["density"])
        efield = _rel(got[i][0], ref["field"])
        eacc = abs(acc[i] - ref["accuracy"]) / ref["accuracy"]
        print(f"  instance {i:2d} ({n}^3, Ns {sum(d['segments'])}): phi {ephi:.1e} field {efield:.1e} "
              f"accuracy {eacc:.1e}", flush=True)
        worst = [max(worst[0], ephi), max(worst[1], efield), max(worst[2], eacc)]
        assert ephi < phi_tol and efield < field_tol and eacc < max(phi_tol, 1e-10) * 10
    print(f"worst over the oracle subset: phi {worst[0]:.2e} field {worst[1]:.2e} accuracy {worst[2]:.2e}")
    assert np.all(np.isfinite(F))


@pytest.mark.parametrize("persistent,plane", [(False, -1), (False, 0), (True, 0)])
def test_sweep_instances_match_single_contexts_past_m(sweep, persistent, plane):
    import adscf_amd

    inst, w = sweep
    # one instance of every (grid, Ns) pair present
    pick, seen = [], set()
    for i, d in enumerate(inst):
        key = (d["grid"][0], sum(d["segments"]))
        if key not in seen:
            seen.add(key)
            pick.append(i)
    sub = [inst[i] for i in pick]
    ws = [w[i] for i in pick]
    iters = M + 3
    print(f"instances {pick} ({sorted(seen)}), {iters} FP64 iterations, persistent {persistent}, plane {plane}",
          flush=True)
    with adscf_amd.Batch(sub, mixing_dim=M, lam_mix=LAM, precision=adscf_amd.FP64) as b:
        assert b.set_plane_passes(plane) == (0 if plane == 0 else len(sub))
        assert b.set_persistent(persistent) == persistent
        b.set_fields(ws)
        it, acc, conv = b.iterate(iters, tol=-1.0)
        got = [b.get(k) for k in range(len(sub))]
    assert np.all(it == iters) and not conv.any()
    for k, d in enumerate(sub):
        # a context's default line schedules (FP64: cube4 at 64, 4x8 at 32, 8x6 at 48) are the
        # one-wave register schedules a batch instance keeps, and its default plane shapes the
        # batch's
        with adscf_amd.Context(*d["grid"], *d["box"], d["ds"], precision=adscf_amd.FP64) as ctx:
            ctx.set_plane_passes(plane)
            ctx.set_interactions(d["kappa"], d["chi"], d["vol"])
            ctx.scf_setup(d["blocks"], d["segments"], mixing_dim=M, lam_mix=LAM)
            ctx.scf_set_field(ws[k])
            st = ctx.scf_iterate(iters, tol=-1.0)
            f1, d1 = ctx.scf_get()
        print(f"  {d['grid'][0]}^3 Ns {sum(d['segments'])}: batch accuracy {acc[k]:.6e} single {st.accuracy:.6e}")
        assert st.iters == it[k]
        assert st.accuracy == acc[k]
        assert np.array_equal(f1, got[k][0]) and np.array_equal(d1, got[k][1])


def test_sweep_persistent_fp32_matches_per_pass(sweep):
    """FP32-mixed, all 64 instances, 5 iterations: the persistent sweep and the per-pass
    launches give the same fields, densities and residuals bit for bit."""
    import adscf_amd

    inst, w = sweep
    out = {}
    for persistent in (False, True):
        with adscf_amd.Batch(inst, mixing_dim=M, lam_mix=LAM, precision=adscf_amd.FP32) as b:
            b.set_plane_passes(0)  # the persistent sweep carries line passes
            assert b.set_persistent(persistent) == persistent
            b.set_fields(w)
            it, acc, conv = b.iterate(5, tol=-1.0)
            out[persistent] = (acc, [b.get(i) for i in range(len(inst))])
    assert np.array_equal(out[False][0], out[True][0])
    for a, c in zip(out[False][1], out[True][1]):
        assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])


def test_sweep_persistent_pinned_shapes_match_per_pass(sweep):
    """The persistent sweep with the 48-point axes pinned to each shape the persistent kernel
    carries for that length (8x6 and 6x8): the same fields, densities and residuals as the
    per-pass launches, bit for bit (FP32-mixed, all 64 instances, 3 iterations)."""
    import adscf_amd

    inst, w = sweep
    for shape in ((8, 6), (6, 8)):
        print(f"48-point axes pinned to {shape}", flush=True)
        out = {}
        for persistent in (False, True):
            with adscf_amd.Batch(inst, mixing_dim=M, lam_mix=LAM, precision=adscf_amd.FP32) as b:
                b.set_plane_passes(0)
                b.set_schedule(48, *shape)
                assert b.set_persistent(persistent) == persistent
                b.set_fields(w)
                it, acc, conv = b.iterate(3, tol=-1.0)
                out[persistent] = (acc, [b.get(i) for i in range(len(inst))])
        assert np.array_equal(out[False][0], out[True][0])
        for a, c in zip(out[False][1], out[True][1]):
            assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])


def test_batch_schedule_refuses_unknown_shapes(sweep):
    import adscf_amd

    inst, _ = sweep
    with adscf_amd.Batch(inst[:2], mixing_dim=M, lam_mix=LAM, precision=adscf_amd.FP32) as b:
        with pytest.raises(adscf_amd.AdscfError):
            b.set_schedule(48, 4, 8)  # 4 x 8 is 32 points
        with pytest.raises(adscf_amd.AdscfError):
            b.set_schedule(40, 8, 5)  # no batched 8 x 5 kernel
