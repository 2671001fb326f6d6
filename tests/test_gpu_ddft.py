"""GPU checks of the DDFT path of Main_Multiblock (main_multiblock.f90:441-506, 566-640,
794-880): the device-resident conjugate-field solve and DDFT Euler step against the oracle,
and the unchanged Main_Multiblock driver, built against the Fortran replacement modules
(the reference's own build of it fails on the mixing_dim clash), running its DDFT on the
GPU path, against the oracle's restatement of the same run.

What is compared exactly and what by outcome.  Get_Conjugate_Fields mixes with lam00 =
0.001 for its first mixing_dim iterations, so the residual history Anderson first solves
with holds nearly equal vectors and its Gram system is singular to rounding: a 1e-16
relative change of the start field moves the field of iteration m+1 by ~1e-3 (shown on
the oracle alone in tests/test_oracle_ddft.py).  Two correct implementations, like the
reference built by two compilers, part there.  Hence:
* the first m iterations of a solve (simple mixing) are compared at 1e-12;
* a full solve is checked by its outcome: the oracle evaluates the densities of the GPU's
  field and finds them within the exit threshold of the target;
* a DDFT step's density update is deterministic given the density and field it starts from
  (its first solve exits at iteration 1 when the field is already conjugate), so every step
  is compared at 1e-12 against the oracle's step from the same start; the field the second
  solve returns is checked by its outcome, the energy by recomputation from the state."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "adscf_amd", "fortran", "_build", "Main_Multiblock", "main_multiblock.exe")

# the hard-coded polymer of main_multiblock.f90:76-92: six blocks A B A B A B with lengths
# 0.5 0.1 0.1 0.1 0.1 0.1, canonical
BT = [1, 2, 1, 2, 1, 2]
LENGTHS = [0.5, 0.1, 0.1, 0.1, 0.1, 0.1]
KAPPA, CHI = 200.0, 70.0  # the shipped input_interactions


def _mobilities():
    from adscf_amd import formats

    return formats.read_mobilities(os.path.join(GOLDEN, "multiblock_input_mobilities"))


def _segments(orc, ds):
    return [orc.block_segments(l, ds) for l in LENGTHS]


def _ctx(amd, nx, ny, nz, box, ds, bs, m=10):
    ctx = amd.Context(nx, ny, nz, *box, ds, precision=amd.FP64)
    ctx.set_interactions(KAPPA, [[0.0, CHI], [CHI, 0.0]], [1.0, 1.0])
    ctx.scf_setup(BT, bs, ensemble=1, rho=1.0, mixing_dim=m, lam_mix=0.5)
    return ctx


def test_mobility_qq_matches_oracle(orc):
    import adscf_amd as amd

    mob = _mobilities()
    g = orc.Grid(24, 6, 20, 4.8, 1.7, 4.0, 0.01)
    ref = orc.mobility_qq(g, mob.table, mob.dq, 10.0)
    with amd.Context(24, 6, 20, 4.8, 1.7, 4.0, 0.01) as ctx:
        got = ctx.mobility_qq(mob.table, mob.dq, 10.0)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("measure", [0, 1])
def test_conjugate_fields_matches_oracle(orc, measure):
    import adscf_amd as amd

    nx, ny, nz, box, ds = 16, 8, 12, (3.2, 1.6, 2.4), 0.02
    bs = _segments(orc, ds)
    g = orc.Grid(nx, ny, nz, *box, ds)
    vol, dvol = g.volume()
    rng = np.random.default_rng(21 + measure)
    w0 = 0.5 * rng.standard_normal((2, nz, ny, nx))
    target, _ = orc.species_densities(g, BT, bs, w0 + 0.2 * rng.standard_normal(w0.shape), ensemble=1, rho=1.0)
    # the simple-mixing iterations (thresh -1: no exit), exactly
    ref = orc.conjugate_fields(g, BT, bs, target, w0, m=10, thresh=-1.0, maxit=10, measure=measure)
    with _ctx(amd, nx, ny, nz, box, ds, bs) as ctx:
        ctx.scf_set_field(w0)
        it, acc = ctx.conjugate_fields(target, thresh=-1.0, maxit=10, measure=measure)
        field, _ = ctx.scf_get()
        assert it == ref["iters"] == 11
        assert abs(acc - ref["accuracy"]) <= 1e-12 * ref["accuracy"]
        assert np.max(np.abs(field - ref["field"])) <= 1e-12 * np.max(np.abs(ref["field"]))
        # a full solve, by its outcome
        thresh = 1e-7 if measure == 0 else 1e-7 * vol
        ctx.scf_set_field(w0)
        it, acc = ctx.conjugate_fields(target, thresh=thresh, maxit=600, measure=measure)
        field, _ = ctx.scf_get()
    ref = orc.conjugate_fields(g, BT, bs, target, w0, m=10, thresh=thresh, maxit=600, measure=measure)
    den, _ = orc.species_densities(g, BT, bs, field, ensemble=1, rho=1.0)
    d = np.abs(den - target)
    a = d.max() if measure == 0 else d.sum() * dvol
    print(f"measure {measure}: GPU {it} iterations, accuracy {acc:.3e} (oracle evaluation {a:.3e}); "
          f"oracle {ref['iters']} iterations")
    assert it <= 600 and acc <= thresh
    # the two evaluations differ by rounding of the densities (~1e-16 of their size per point)
    tol = 1e-13 * np.max(target) * (vol if measure else 1.0)
    assert abs(a - acc) <= tol and a <= thresh + tol


def _state_step(orc, g, bs, fh, mq, den, field, dt, **kw):
    d, f = den.copy(), field.copy()
    r = orc.ddft_step(g, BT, bs, fh, d, f, mq, dt, m=10, **kw)
    assert r["rc"] == 0
    return d, f, r


def test_ddft_steps_match_oracle(orc):
    """Three DDFT_Timestep_Euler steps (the driver's 1e-3 / 2000 solves, qdmin) of the
    shipped polymer, interactions and mobility table on a 24 x 1 x 20 grid, each compared with
    the oracle's step from the GPU's state before it."""
    import adscf_amd as amd

    nx, ny, nz, box, ds = 24, 1, 20, (4.8, 1.0, 4.0), 0.01
    bs = _segments(orc, ds)
    mob = _mobilities()
    g = orc.Grid(nx, ny, nz, *box, ds)
    vol, dvol = g.volume()
    fh = orc.FH(KAPPA, [[0.0, CHI], [CHI, 0.0]], [1.0, 1.0])
    mq = orc.mobility_qq(g, mob.table, mob.dq, 10.0)
    field = orc.MT(2007).grnd(2 * nx * ny * nz).reshape(2, nz, ny, nx).copy()
    den, _ = orc.species_densities(g, BT, bs, field, ensemble=1, rho=1.0)
    with _ctx(amd, nx, ny, nz, box, ds, bs) as ctx:
        ctx.mobility_qq(mob.table, mob.dq, 10.0)
        ctx.scf_set_field(field)
        ctx.ddft_set_density(den)
        for step in range(3):
            d_ref, f_ref, r = _state_step(orc, g, bs, fh, mq, den, field, 5e-5, qdmin=True)
            s = ctx.ddft_step(5e-5, qdmin=True)
            den = ctx.ddft_get_density()
            field, _ = ctx.scf_get()
            e = ctx.ddft_energy()
            chain, info = orc.species_densities(g, BT, bs, field, ensemble=1, rho=1.0)
            e_ref = info.energy - dvol * np.sum(den * field) + orc.interaction_energy(g, fh, den)
            print(f"step {step}: GPU solves {s['iterations']} {s['gcfstat']}, oracle {r['iters']} {r['gcfstat']}; "
                  f"max|dphi| {np.max(np.abs(den - d_ref)):.2e}, conjugacy {np.max(np.abs(chain - den)):.2e}, "
                  f"F {e['F']:.12f} ({e_ref:.12f})")
            assert s["iterations"][0] == r["iters"][0] == 1 and s["gcfstat"] == [0, 0]
            assert np.max(np.abs(den - d_ref)) <= 1e-12 * np.max(d_ref)
            assert np.max(np.abs(chain - den)) <= 0.5e-3
            assert abs(e["F"] - e_ref) <= 1e-10 * abs(e_ref)


def test_ddft_step_slow_solve_takes_densities_of_field(orc):
    """qdmin with a second solve that does not reach accuracy_scf (gcfstat 2): the density
    becomes the densities of the field (main_multiblock.f90:874-876)."""
    import adscf_amd as amd

    nx, ny, nz, box, ds = 24, 1, 20, (4.8, 1.0, 4.0), 0.01
    bs = _segments(orc, ds)
    mob = _mobilities()
    g = orc.Grid(nx, ny, nz, *box, ds)
    fh = orc.FH(KAPPA, [[0.0, CHI], [CHI, 0.0]], [1.0, 1.0])
    mq = orc.mobility_qq(g, mob.table, mob.dq, 10.0)
    field = orc.MT(2008).grnd(2 * nx * ny * nz).reshape(2, nz, ny, nx).copy()
    den, _ = orc.species_densities(g, BT, bs, field, ensemble=1, rho=1.0)
    d_ref, f_ref, r = _state_step(orc, g, bs, fh, mq, den, field, 5e-5, qdmin=True, accuracy_scf=1e-9,
                                  itermax_scf=4)
    with _ctx(amd, nx, ny, nz, box, ds, bs) as ctx:
        ctx.mobility_qq(mob.table, mob.dq, 10.0)
        ctx.scf_set_field(field)
        ctx.ddft_set_density(den)
        s = ctx.ddft_step(5e-5, qdmin=True, accuracy_scf=1e-9, itermax_scf=4)
        gd = ctx.ddft_get_density()
        gf, _ = ctx.scf_get()
    assert r["gcfstat"] == [0, 2] and r["iters"] == [1, 5]
    assert s["gcfstat"] == [0, 2] and s["iterations"] == [1, 5]
    # four simple-mixing iterations: no Anderson yet, so the field is comparable exactly
    assert np.max(np.abs(gf - f_ref)) <= 1e-12 * np.max(np.abs(f_ref))
    chain, _ = orc.species_densities(g, BT, bs, gf, ensemble=1, rho=1.0)
    assert np.max(np.abs(gd - chain)) <= 1e-12 * np.max(chain)
    assert np.max(np.abs(gd - d_ref)) <= 1e-12 * np.max(d_ref)


MB_GENERAL = """17.2d0\t\tsizex
1.0d0\t\tsizey
1.0d0\t\tsizez

1.0d0\t\tC
5.65d0 \t\tmu_scf

0.01d0\t\tds

0.          start_time
1\t        number_of_timesteps
0.00005\t\ttimestep

.false.\t\tqread
3           qinit
.true.      qddft

10.          upper omega cutoff for DDFT calculations
"""
MB_INTERACTIONS = """200.0d0        ! kappa

70.0d0         ! chi

1.0d0          ! volume of monomer_type 1
1.0d0          ! volume of monomer_type 2
"""


@pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built (adscf_amd/fortran/Makefile needs the reference tree)")
def test_unchanged_multiblock_driver_ddft(tmp_path, orc):
    """The unchanged Main_Multiblock driver (its 86 x 1 x 75 mod_global.f90, mixing_dim 10) with
    the shipped inputs, qinit = 3 (random fields from grnd, seed 2007) and one DDFT step:
    Read_Mobilities (the reference's own code, on the replacement Fourier module's laplace),
    Calculate_Densities, DDFT_Timestep_Euler and Calculate_Energy run on the GPU path and
    write final_dens.dat, final_field.dat and the energy; the oracle restates the same run,
    and the device-resident DDFT step repeats it from the same start."""
    import adscf_amd as amd
    from adscf_amd import formats

    (tmp_path / "input_general").write_text(MB_GENERAL)
    (tmp_path / "input_interactions").write_text(MB_INTERACTIONS)
    with open(os.path.join(GOLDEN, "multiblock_input_mobilities")) as f:
        (tmp_path / "input_mobilities").write_text(f.read())
    res = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(res.stdout[-1500:], res.stderr[-1500:])
    assert res.returncode == 0 and "BEGIN DDFT SIMULATION" in res.stdout
    nx, ny, nz, box, ds = 86, 1, 75, (17.2, 1.0, 1.0), 0.01
    shape = (2, nz, ny, nx)
    den_drv = formats.read_profile(str(tmp_path / "final_dens.dat"), shape, "xyz")
    field_drv = formats.read_profile(str(tmp_path / "final_field.dat"), shape, "xyz")
    cnf = formats.read_cnf(str(tmp_path / "cnf.out"), shape)
    energy_drv = formats.parse_real(res.stdout.split("Energy:")[-1].split()[0])

    bs = _segments(orc, ds)
    mob = _mobilities()
    g = orc.Grid(nx, ny, nz, *box, ds)
    vol, dvol = g.volume()
    fh = orc.FH(KAPPA, [[0.0, CHI], [CHI, 0.0]], [1.0, 1.0])
    mq = orc.mobility_qq(g, mob.table, mob.dq, 10.0)
    field0 = orc.MT(2007).grnd(2 * nx * ny * nz).reshape(shape).copy()
    den0, _ = orc.species_densities(g, BT, bs, field0, ensemble=1, rho=1.0)
    den, field, r = _state_step(orc, g, bs, fh, mq, den0, field0, 5e-5, qdmin=True)
    # the density after the step is deterministic (the first solve exits at iteration 1)
    print(f"oracle solves {r['iters']} {r['gcfstat']}; max|dphi| driver-oracle {np.max(np.abs(den_drv - den)):.2e}")
    assert r["iters"][0] == 1
    assert np.max(np.abs(den_drv - den)) <= 1e-12 * np.max(den)
    # the field of the second solve, by its outcome, and the printed energy from the written state
    chain, info = orc.species_densities(g, BT, bs, field_drv, ensemble=1, rho=1.0)
    assert np.max(np.abs(chain - den_drv)) <= 0.5e-3
    e_ref = info.energy - dvol * np.sum(den_drv * field_drv) + orc.interaction_energy(g, fh, den_drv)
    print(f"driver energy {energy_drv!r}, from its written state {e_ref!r}")
    assert abs(energy_drv - e_ref) <= 1e-10 * abs(e_ref)
    assert np.max(np.abs(cnf - field_drv)) <= 1e-15 * np.max(np.abs(cnf))

    # the device-resident step from the same start
    with _ctx(amd, nx, ny, nz, box, ds, bs) as ctx:
        ctx.mobility_qq(mob.table, mob.dq, 10.0)
        ctx.scf_set_field(field0)
        ctx.ddft_set_density(den0)
        s = ctx.ddft_step(5e-5, qdmin=True)
        gd = ctx.ddft_get_density()
        gf, _ = ctx.scf_get()
    assert s["iterations"][0] == 1 and s["gcfstat"] == [0, 0]
    assert np.max(np.abs(gd - den)) <= 1e-12 * np.max(den)
    chain, _ = orc.species_densities(g, BT, bs, gf, ensemble=1, rho=1.0)
    assert np.max(np.abs(chain - gd)) <= 0.5e-3


def test_ddft_step_fp32_mixed_error(orc):
    """FP32-mixed DDFT step (float propagators and transforms; densities, fields, the mobility
    operator's inputs and outputs and Anderson in double) against the FP64 step from the same
    start: stated bound 1e-5 of max density for the density after one step."""
    import adscf_amd as amd

    nx, ny, nz, box, ds = 24, 1, 20, (4.8, 1.0, 4.0), 0.01
    bs = _segments(orc, ds)
    mob = _mobilities()
    g = orc.Grid(nx, ny, nz, *box, ds)
    field = orc.MT(2007).grnd(2 * nx * ny * nz).reshape(2, nz, ny, nx).copy()
    den, _ = orc.species_densities(g, BT, bs, field, ensemble=1, rho=1.0)
    out = {}
    for prec in (amd.FP64, amd.FP32):
        with amd.Context(nx, ny, nz, *box, ds, precision=prec) as ctx:
            ctx.set_interactions(KAPPA, [[0.0, CHI], [CHI, 0.0]], [1.0, 1.0])
            ctx.scf_setup(BT, bs, ensemble=1, rho=1.0, mixing_dim=10, lam_mix=0.5)
            ctx.mobility_qq(mob.table, mob.dq, 10.0)
            ctx.scf_set_field(field)
            ctx.ddft_set_density(den)
            s = ctx.ddft_step(5e-5, qdmin=True)
            out[prec] = (ctx.ddft_get_density(), s)
    err = np.max(np.abs(out[amd.FP32][0] - out[amd.FP64][0])) / np.max(out[amd.FP64][0])
    print(f"FP32-mixed DDFT step: density error {err:.2e}; solves {out[amd.FP32][1]} vs {out[amd.FP64][1]}")
    assert err < 1e-5
