#!/usr/bin/env python3
"""SCF iterations/s of the device-resident SCFT loop at 64^3 x 200 contour steps.

Workload (AB diblock melt, the plain-method GPU configuration): 64^3 grid, box
4.5 Rg, Ns = 200 (ds = 0.005, block segments nint(0.3*200) = 60 / 140), f_A = 0.3,
chiN = 20, Helfand kappa = 25, monomer volumes 1, canonical ensemble rho = 1,
Anderson mixing with history 20 and lam_mix = 0.5.  Initial fields: seeded
N(0, 0.2^2) noise plus the 6-fold cylinder seed (adscf_amd/fields.py).

A "step" is one SCF iteration: the forward and complementary propagator
sweeps (2 x 200 contour steps), Q, the contour integral, the mean fields and
residual, and the Anderson update, all resident in HBM (one replay of the
captured HIP graph).  value = SCF iterations/s of the FP32-mixed path; the
plain FP64 path is timed beside it.  The roofline entry covers the dominant
kernel, the propagator line pass: its algorithmic bytes are 3*Ng*P per contour
step per chain (read q_s, read E, write q_{s+1}), so a sweep of both chains
moves 2*Ns*3*Ng*P bytes over 2*Ns+1 passes; the pass duration comes from HIP
events around replays of the captured sweep on the library's stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SCF iters/s at 64³×200: FP32-mixed vs FP64 GPU at |ΔF|<1e-6; % HBM peak"
HBM_PEAK_GBS = 8000.0

N, L, NS, FA, CHIN, KAPPA, M, LAM = 64, 4.5, 200, 0.3, 20.0, 25.0, 20, 0.5


def algorithmic_bytes_iter(ng, ns, m, p):
    """B_iter: sweeps, contour integral (read every q/q+ slice, write phi), Q, E from w,
    and the FP64 field update (read phi, w, write D; Gram row over m D's; mix over m W's and
    m D's; write w; push W and D)."""
    return (2 * ns * 3 * ng * p + (2 * (ns + 1) * ng * p + 2 * ng * 8) + ng * p + 2 * ng * (8 + p)
            + (3 * m + 6) * 2 * ng * 8)


def shape_name(sch):
    if sch["N1"] == 0:
        return "stockham"
    return "cube%d" % sch["N1"] if sch["N2"] == 0 else "%dx%d" % (sch["N1"], sch["N2"])


def run_gpu(adscf_amd, fields, precision, steps, warmup, w0, hipfft=True, tune=False):
    segs = fields.block_segments([FA, 1 - FA], NS)
    ctx = adscf_amd.Context(N, N, N, L, L, L, 1.0 / NS, precision=precision)
    chi = [[0.0, CHIN], [CHIN, 0.0]]
    ctx.set_interactions(KAPPA, chi, [1.0, 1.0])
    ctx.scf_setup([1, 2], segs, ensemble=1, rho=1.0, mixing_dim=M, lam_mix=LAM)
    ctx.scf_set_field(w0)
    # per-axis line-pass schedule (shape, waves per workgroup)
    report = ctx.autotune(reps=3) if tune else "".join(
        "axis %d chosen: %s waves=%d (default)\n" % (a, shape_name(ctx.schedule(a)), ctx.schedule(a)["waves"])
        for a in range(3))
    ctx.scf_iterate(warmup, tol=-1.0)
    t0 = time.perf_counter()
    st = ctx.scf_iterate(steps, tol=-1.0)  # synchronises the library stream on return
    t1 = time.perf_counter()
    ms_iter_ev, ms_sweep = ctx.scf_timing()
    hipfft_us = ctx.hipfft_step_us(200) if hipfft else None
    wall = t1 - t0
    plane = ctx.plane_passes()
    out = dict(iters_per_s=steps / wall, ms_per_iter=1000.0 * wall / steps, ms_per_iter_events=ms_iter_ev,
               ms_sweep=ms_sweep, iterations_run=st.iters, stopped_early=st.converged, hipfft_step_us=hipfft_us,
               schedule=[l for l in report.splitlines() if "chosen" in l], plane=plane,
               npass=sweep_passes(NS, plane["active"]))
    if plane["active"]:
        out["schedule"].append("xy-plane passes: %s" % shape_name(plane))
    return ctx, out


def sweep_passes(ns, plane):
    """Dependent passes of one propagator sweep of both chains over a 3-D grid: with xy-plane
    passes one per contour step plus one (plane and z passes alternate, the operators of
    consecutive steps meeting in one pass), with line passes two per step plus one."""
    return ns + 1 if plane else 2 * ns + 1


def sweep_kernels(prec, plane):
    t = "float" if prec == "fp32" else "double"
    if plane["active"]:
        return ("k_plane_pass<%s, %d, %d, *> (x and y operators of a step, one workgroup per chain and z-plane) "
                "alternating with k_line_pass_reg<%s, *> (z)" % (t, plane["N1"], plane["N2"], t))
    return "k_line_pass_reg<%s, *> (propagator line passes, all variants of the sweep)" % t


def _metric_ctx(adscf_amd, fields, prec, w0):
    ctx = adscf_amd.Context(N, N, N, L, L, L, 1.0 / NS, precision=prec)
    ctx.set_interactions(KAPPA, [[0.0, CHIN], [CHIN, 0.0]], [1.0, 1.0])
    ctx.scf_setup([1, 2], fields.block_segments([FA, 1 - FA], NS), ensemble=1, rho=1.0, mixing_dim=M, lam_mix=LAM)
    ctx.scf_set_field(w0)
    return ctx


def iterations_to_dF(adscf_amd, fields, prec, w0, F_ref, max_iter, thresh=1e-6, chunk=5):
    """First iteration count k after which |F per chain - F_ref| < thresh holds at every probe
    (every `chunk` iterations, for 3 probes running); then a fresh timed run of exactly k
    iterations (graph captured beforehand).  Returns dict(iters, seconds, dF) or None."""
    k = None
    below = 0
    with _metric_ctx(adscf_amd, fields, prec, w0) as ctx:
        it = 0
        while it < max_iter:
            st = ctx.scf_iterate(chunk, tol=-1.0)
            it = st.iters
            dF = abs(ctx.scf_energy()["F_per_chain"] - F_ref)
            if dF < thresh:
                below += 1
                if below == 1:
                    k, dF_k = it, dF
                if below == 3:
                    break
            else:
                below, k = 0, None
    if k is None or below < 3:
        return None
    with _metric_ctx(adscf_amd, fields, prec, w0) as ctx:
        ctx.scf_iterate(0, tol=-1.0)  # captures the iteration graph outside the timed region
        t0 = time.perf_counter()
        st = ctx.scf_iterate(k, tol=-1.0)
        t = time.perf_counter() - t0
        dF = abs(ctx.scf_energy()["F_per_chain"] - F_ref)
    return dict(iters=k, seconds=t, iters_per_s=k / t, dF_per_chain=dF)


def converge(adscf_amd, fields, w0, max_iter):
    """The FP64 saddle point (max|dfield| < 1e-8) gives the reference F per chain.  Then, per
    path, the iterations and seconds until |dF| < 1e-6 per chain against it: plain FP64, and
    FP32-mixed (FP32 propagators and FFTs, FP64 Q/F/phi sums and Anderson) without refinement.
    Last, FP32-mixed with FP64 refinement (switch once max|dfield| < 1e-4 or the FP32 residual
    stops improving) to 1e-8, as the full solve."""
    res = {}
    with _metric_ctx(adscf_amd, fields, adscf_amd.FP64, w0) as ctx:
        ctx.scf_iterate(0, tol=1e-8)
        t0 = time.perf_counter()
        st = ctx.scf_iterate(max_iter, tol=1e-8)
        t = time.perf_counter() - t0
        F_ref = ctx.scf_energy()["F_per_chain"]
        res["fp64"] = dict(iters=st.iters, accuracy=st.accuracy, converged=st.converged, F_per_chain=F_ref,
                           seconds=t)
    dF = {"threshold": 1e-6, "F_ref_per_chain": F_ref}
    for name, prec in (("fp64", adscf_amd.FP64), ("fp32mixed", adscf_amd.FP32)):
        dF[name] = iterations_to_dF(adscf_amd, fields, prec, w0, F_ref, max_iter)
    if dF["fp64"] and dF["fp32mixed"]:
        dF["speedup_fp32mixed_vs_fp64_time"] = dF["fp64"]["seconds"] / dF["fp32mixed"]["seconds"]
        dF["speedup_fp32mixed_vs_fp64_iters_per_s"] = dF["fp32mixed"]["iters_per_s"] / dF["fp64"]["iters_per_s"]
    res["to_dF_1e-6"] = dF
    with _metric_ctx(adscf_amd, fields, adscf_amd.FP32, w0) as ctx:
        ctx.scf_iterate(0, tol=1e-4)
        t0 = time.perf_counter()
        # FP32 phase until max|dfield| < 1e-4 or it stops improving (FP32 rounding of phi,
        # amplified by kappa, sets a residual floor near 3e-5 here), then the FP64 refinement
        # continues the same loop and history
        best, it32 = float("inf"), 0
        while it32 < max_iter:
            st = ctx.scf_iterate(50, tol=1e-4)
            it32 = st.iters
            r = ctx.scf_residuals()[-50:]
            if st.converged or r.min() > 0.7 * best:
                break
            best = min(best, r.min())
        res["fp32_residual_at_switch"] = float(st.accuracy)
        ctx.set_precision(adscf_amd.FP64)
        st = ctx.scf_iterate(max_iter, tol=1e-8)
        t = time.perf_counter() - t0
        en = ctx.scf_energy()
        res["fp32mixed"] = dict(iters=st.iters, accuracy=st.accuracy, converged=st.converged,
                                F_per_chain=en["F_per_chain"], seconds=t, fp32_iters=it32)
    res["dF_per_chain"] = abs(res["fp32mixed"]["F_per_chain"] - res["fp64"]["F_per_chain"])
    return res


def run_batched(adscf_amd, fields, steps, warmup):
    """The 64-instance parameter sweep (32^3/48^3/64^3 grids, Ns in {100,150,200}), FP32-mixed,
    one graph replay per SCF iteration of all instances, schedules from the batch autotuner.
    The default sweep (xy-plane passes alternating with z passes) is the reported one; the
    same iteration with line passes only, as per-pass launches and as one persistent launch,
    is timed beside it."""
    inst, w = fields.batched_instances(seed=2007)

    def timed(b):
        b.set_fields(w)
        b.iterate(warmup, tol=-1.0)
        t0 = time.perf_counter()
        it, acc, conv = b.iterate(steps, tol=-1.0)
        wall = time.perf_counter() - t0
        return steps / wall, b.timing()[1], bool(conv.any())

    with adscf_amd.Batch(inst, mixing_dim=M, lam_mix=LAM, precision=adscf_amd.FP32) as b:
        b.set_fields(w)
        report = b.autotune(reps=2)
        nplane = b.set_plane_passes(-1)
        ips, ms_sweep, conv = timed(b)
        b.set_plane_passes(0)
        l_ips, l_sweep, l_conv = timed(b)
        lines = dict(batch_iters_per_s=l_ips, ms_per_batch_iter=1000.0 / l_ips, ms_sweep=l_sweep)
        # the line-pass iteration with the sweep as one persistent launch
        persistent = {"active": b.set_persistent(True)}
        if persistent["active"]:
            p_ips, p_sweep, _ = timed(b)
            persistent.update(batch_iters_per_s=p_ips, ms_per_batch_iter=1000.0 / p_ips, ms_sweep=p_sweep)
        b.set_persistent(False)
        b.set_plane_passes(-1)
    bytes_iter = sum(algorithmic_bytes_iter(d["grid"][0] ** 3, sum(d["segments"]), M, 4) for d in inst)
    sweep_bytes = sum(2 * sum(d["segments"]) * 3 * d["grid"][0] ** 3 * 4 for d in inst)
    lines["iteration_bytes_frac"] = bytes_iter * l_ips / (HBM_PEAK_GBS * 1e9)
    lines["sweep_roofline_frac"] = sweep_bytes / (l_sweep * 1e-3) / (HBM_PEAK_GBS * 1e9)
    return dict(instances=len(inst), batch_iters_per_s=ips, instance_iters_per_s=ips * len(inst),
                ms_per_batch_iter=1000.0 / ips,
                contour_steps_per_s=ips * sum(2 * sum(d["segments"]) for d in inst),
                iteration_bytes_frac=bytes_iter * ips / (HBM_PEAK_GBS * 1e9),
                sweep_roofline_frac=sweep_bytes / (ms_sweep * 1e-3) / (HBM_PEAK_GBS * 1e9),
                timing_valid=not (conv or l_conv), ms_sweep=ms_sweep,
                sweep_mode="xy-plane passes alternating with z passes on %d of %d instances, one launch per "
                           "pass, kernel and lane" % (nplane, len(inst)),
                line_passes=lines, persistent_sweep=persistent,
                schedules=[l for l in report.splitlines() if "chosen" in l])


def run_ddft(adscf_amd, fields, nsteps=6):
    """Main_Multiblock's DDFT as shipped (86 x 1 x 75 grid, box 17.2 x 1 x 1, the ABABAB
    polymer of lengths 0.5/0.1 x 5, ds = 0.01, kappa 200, chi 70, the shipped mobility table,
    wcutoff 10, dt = 5e-5, qdmin) from grnd's random fields (seed 2007): the device-resident
    DDFT_Timestep_Euler, and the unchanged driver on the compat path when it is built
    (wall time of nsteps steps minus a zero-step run, as one times the reference)."""
    import subprocess
    import tempfile
    from adscf_amd import formats

    nx, ny, nz, box, ds = 86, 1, 75, (17.2, 1.0, 1.0), 0.01
    bt = [1, 2, 1, 2, 1, 2]
    bs = fields.block_segments([0.5, 0.1, 0.1, 0.1, 0.1, 0.1], 100)
    mob_path = os.path.join(ROOT, "tests", "golden", "multiblock_input_mobilities")
    mob = formats.read_mobilities(mob_path)
    field = fields.MTStream(2007).grnd(2 * nx * ny * nz).reshape(2, nz, ny, nx)
    with adscf_amd.Context(nx, ny, nz, *box, ds, precision=adscf_amd.FP64) as ctx:
        ctx.set_interactions(200.0, [[0.0, 70.0], [70.0, 0.0]], [1.0, 1.0])
        ctx.scf_setup(bt, bs, ensemble=1, rho=1.0, mixing_dim=10, lam_mix=0.5)
        ctx.mobility_qq(mob.table, mob.dq, 10.0)
        ctx.scf_set_field(field)
        ctx.scf_evaluate()
        ctx.ddft_set_density(ctx.scf_get()[1])
        ctx.ddft_step(5e-5)  # warm-up: graph capture and work buffers
        iters = []
        t0 = time.perf_counter()
        for _ in range(nsteps):
            iters.append(ctx.ddft_step(5e-5)["iterations"])
        wall = time.perf_counter() - t0
        ms_it = ctx.scf_timing()[0]
    out = dict(workload="Main_Multiblock DDFT, 86x1x75, Ns=100 (6 blocks), shipped inputs, dt=5e-5",
               steps=nsteps, ms_per_step=1000.0 * wall / nsteps, solve_iterations=iters,
               ms_per_solve_iteration=ms_it)
    exe = os.path.join(ROOT, "adscf_amd", "fortran", "_build", "Main_Multiblock", "main_multiblock.exe")
    if os.path.exists(exe):
        gen = ("17.2d0\n1.0d0\n1.0d0\n\n1.0d0\n5.65d0\n\n0.01d0\n\n0.\n{n}\n0.00005\n\n"
               ".false.\n3\n.true.\n\n10.\n")
        inter = "200.0d0\n\n70.0d0\n\n1.0d0\n1.0d0\n"
        walls = {}
        for n in (0, nsteps):
            with tempfile.TemporaryDirectory() as d:
                with open(os.path.join(d, "input_general"), "w") as f:
                    f.write(gen.format(n=n))
                with open(os.path.join(d, "input_interactions"), "w") as f:
                    f.write(inter)
                with open(mob_path) as src, open(os.path.join(d, "input_mobilities"), "w") as dst:
                    dst.write(src.read())
                t0 = time.perf_counter()
                r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=600)
                walls[n] = time.perf_counter() - t0
                if r.returncode != 0:
                    walls = None
                    break
        if walls:
            out["unchanged_driver_ms_per_step"] = 1000.0 * (walls[nsteps] - walls[0]) / nsteps
    return out


def run_epd(adscf_amd, fields, nsteps=10):
    """Main_Blend's EPD as shipped (64 x 16 x 1 grid, box 20 x 5 x 1, two homopolymers of 50
    chains each, length 1 at ds = 0.1, kappa = chi = 10, dt = 1e-3, Anderson history 2) from
    grnd's random fields (seed 2007): the device-resident Chain_Dynamics_Timestep_Euler, and
    the unchanged driver on the compat path when it is built (nsteps minus 0 steps)."""
    import subprocess
    import tempfile

    nx, ny, nz, box, ds = 64, 16, 1, (20.0, 5.0, 1.0), 0.1
    vol = box[0] * box[1] * box[2]
    rho = [50.0 / vol, 50.0 / vol]
    field = fields.MTStream(2007).grnd(2 * nx * ny * nz).reshape(2, nz, ny, nx)
    with adscf_amd.Context(nx, ny, nz, *box, ds, precision=adscf_amd.FP64) as ctx:
        ctx.set_interactions(10.0, [[0.0, 10.0], [10.0, 0.0]], [1.0, 1.0])
        ctx.blend_setup([1, 2], [10, 10], rho, mixing_dim=2, lam_mix=0.5)
        ctx.scf_set_field(field)
        ctx.scf_evaluate()
        ctx.ddft_set_density(ctx.scf_get()[1])
        ctx.epd_step(0.001)  # warm-up
        iters = []
        t0 = time.perf_counter()
        for _ in range(nsteps):
            iters.append(ctx.epd_step(0.001))
        wall = time.perf_counter() - t0
    out = dict(workload="Main_Blend EPD, 64x16x1, two homopolymers (10 segments), shipped inputs, dt=1e-3",
               steps=nsteps, ms_per_step=1000.0 * wall / nsteps, solve_iterations=iters)
    exe = os.path.join(ROOT, "adscf_amd", "fortran", "_build", "Main_Blend", "main_blend.exe")
    if os.path.exists(exe):
        gen = ("20.0d0\n5.0d0\n1.0d0\n\n50.0d0\n27.0d0\n1.0d0\n\n50.0d0\n27.0d0\n1.0d0\n\n0.1d0\n\n"
               "{n}\n0.001\n\n.false.\n3\n.true.\n")
        walls = {}
        for n in (0, nsteps):
            with tempfile.TemporaryDirectory() as d:
                with open(os.path.join(d, "input_general"), "w") as f:
                    f.write(gen.format(n=n))
                with open(os.path.join(d, "input_interactions"), "w") as f:
                    f.write("10.0d0\n\n10.0d0\n\n1.0d0\n1.0d0\n")
                t0 = time.perf_counter()
                r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=600)
                walls[n] = time.perf_counter() - t0
                if r.returncode != 0:
                    walls = None
                    break
        if walls:
            out["unchanged_driver_ms_per_step"] = 1000.0 * (walls[nsteps] - walls[0]) / nsteps
    return out


TRAFFIC_SOURCES = {False: "profiles/line_pass_64cubed_pmc.md", True: "profiles/sweep_passes_64cubed_pmc.md"}


def profiled_traffic(plane):
    """HBM traffic per pass launch (bytes, mean over a sweep) from the committed rocprofv3 PMC
    summary of the sweep mode in use (FETCH_SIZE and WRITE_SIZE runs, copy-calibrated):
    line passes, or xy-plane passes alternating with z passes.  None if it is absent."""
    path = os.path.join(ROOT, TRAFFIC_SOURCES[bool(plane)])
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("bench traffic_bytes_per_launch"):
                    kv = dict(tok.split("=") for tok in line.split()[2:])
                    return {k: float(v) for k, v in kv.items()}
    except OSError:
        pass
    return None


def copy_probe(mib=256, reps=3):
    """Plain device copies of known size for calibrating FETCH_SIZE / WRITE_SIZE against bytes:
    a contiguous copy of mib MiB (hipMemcpy's kernel, wide requests) and a copy that reads
    mib/2 MiB in 64-byte pieces (every other 64 B; the segment size of the line passes)."""
    import torch

    x = torch.ones(mib * 2 ** 20 // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    for _ in range(reps):
        z = x.view(-1, 32)[:, :16].contiguous()
    torch.cuda.synchronize()
    del x, y, z


def measured_hbm(adscf_amd, nbytes=1 << 30, reps=7):
    """Device bandwidth beside the 8 TB/s peak: a 1 GiB device-to-device copy (read + write)
    and a 1 GiB streaming read, median of reps (adscf_hbm_bandwidth, HIP events)."""
    with adscf_amd.Context(8, 8, 8, 1.0, 1.0, 1.0, 0.1) as ctx:
        cp, rd = ctx.hbm_bandwidth(nbytes, reps)
    return {"copy_GBs": cp, "read_GBs": rd, "bytes": nbytes,
            "method": "hipMemcpyAsync device-to-device and a streaming-read kernel over 1 GiB, median of %d" % reps}


def run_large(adscf_amd, fields, steps, warmup):
    """The large mixed-precision configuration, 128^3 x 400 ABA triblock (0.25/0.5/0.25,
    chiN = 25, box 6.0, BCC seed, m = 20): iterations/s and the line-pass roofline for
    FP32-mixed and FP64 (default schedules)."""
    n, L, ns, chin = 128, 6.0, 400, 25.0
    segs = fields.block_segments([0.25, 0.5, 0.25], ns)
    w0 = fields.bcc_seed(n, L, chin)
    out = {}
    for name, prec, p in (("fp32mixed", adscf_amd.FP32, 4), ("fp64", adscf_amd.FP64, 8)):
        with adscf_amd.Context(n, n, n, L, L, L, 1.0 / ns, precision=prec) as ctx:
            ctx.set_interactions(KAPPA, [[0.0, chin], [chin, 0.0]], [1.0, 1.0])
            ctx.scf_setup([1, 2, 1], segs, ensemble=1, rho=1.0, mixing_dim=M, lam_mix=LAM)
            ctx.scf_set_field(w0)
            ctx.scf_iterate(warmup, tol=-1.0)
            t0 = time.perf_counter()
            st = ctx.scf_iterate(steps, tol=-1.0)
            wall = time.perf_counter() - t0
            _, ms_sweep = ctx.scf_timing()
            npass = sweep_passes(ns, ctx.plane_passes()["active"])
        bytes_launch = 2 * ns * 3 * n ** 3 * p / npass
        ach = bytes_launch / (ms_sweep * 1e-3 / npass) / 1e9
        out[name] = dict(iters_per_s=steps / wall, ms_per_iter=1e3 * wall / steps, us_per_pass=1e3 * ms_sweep / npass,
                         roofline_achieved_GBs=ach, roofline_frac=ach / HBM_PEAK_GBS,
                         contour_steps_per_s=2 * ns * steps / wall, timing_valid=not st.converged)
    out["speedup_fp32mixed_vs_fp64"] = out["fp32mixed"]["iters_per_s"] / out["fp64"]["iters_per_s"]
    out["config"] = "ABA triblock melt SCF, 128^3 grid, box 6.0 Rg, Ns=400 (100/200/100), chiN=25, kappa=25, m=20"
    return out


def host_cpu():
    """CPU model and thread count of this host, and whether the reference's own CPU path
    could be built here: a Fortran compiler (gfortran, or flang) and FFTW3's libraries."""
    import ctypes.util
    import shutil

    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(),
            "gfortran": shutil.which("gfortran"), "flang": shutil.which("flang") or
            (shutil.which("/opt/rocm/llvm/bin/flang") if os.path.exists("/opt/rocm/llvm/bin/flang") else None),
            "libfftw3": ctypes.util.find_library("fftw3"), "libfftw3f": ctypes.util.find_library("fftw3f")}


_ORACLE_CHILD = """
import sys, time
import numpy as np
sys.path.insert(0, %(root)r)
import oracle
w0 = np.load(%(path)r)
g = oracle.Grid(%(n)d, %(n)d, %(n)d, %(L)r, %(L)r, %(L)r, %(ds)r)
fh = oracle.FH(%(kappa)r, [[0.0, %(chin)r], [%(chin)r, 0.0]], [1.0, 1.0])
t0 = time.perf_counter()
oracle.scf_anderson(g, [1, 2], %(segs)r, fh, w0, %(m)d, lam=%(lam)r, tol=-1.0, maxit=1)
print(time.perf_counter() - t0)
"""


def cpu_baseline(adscf_amd, fields, w0):
    """The CPU FP64 oracle on one SCF iteration of the same workload (bounded sample): on one
    thread, and as nproc independent single-threaded processes at once (the all-core
    throughput of independent SCF problems, the reference's single-threaded path run once
    per core).  The one-thread result also checks the GPU's FP64 and FP32-mixed iteration from
    the same fields.  The oracle is a plain C restatement (its FFT a recursive radix-2/3/5 /
    generic DFT, not FFTW): the reference itself needs FFTW3, which this host lacks."""
    import subprocess
    import tempfile

    import numpy as np
    import oracle

    segs = fields.block_segments([FA, 1 - FA], NS)
    g = oracle.Grid(N, N, N, L, L, L, 1.0 / NS)
    fh = oracle.FH(KAPPA, [[0.0, CHIN], [CHIN, 0.0]], [1.0, 1.0])
    t0 = time.perf_counter()
    ref = oracle.scf_anderson(g, [1, 2], segs, fh, w0, M, lam=LAM, tol=-1.0, maxit=1)
    t = time.perf_counter() - t0
    host = host_cpu()
    nproc = int(host["nproc"] or 1)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "w0.npy")
        np.save(path, np.ascontiguousarray(w0))
        code = _ORACLE_CHILD % dict(root=ROOT, path=path, n=N, L=L, ds=1.0 / NS, kappa=KAPPA, chin=CHIN,
                                    segs=list(segs), m=M, lam=LAM)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        t0 = time.perf_counter()
        procs = [subprocess.Popen([sys.executable, "-c", code], stdout=subprocess.PIPE, text=True, env=env)
                 for _ in range(nproc)]
        outs = [p.communicate()[0] for p in procs]
        wall = time.perf_counter() - t0
        ok = all(p.returncode == 0 for p in procs)
    parity = {}
    for name, prec in (("fp64", adscf_amd.FP64), ("fp32mixed", adscf_amd.FP32)):
        with _metric_ctx(adscf_amd, fields, prec, w0) as ctx:
            st = ctx.scf_iterate(1, tol=-1.0)
            field, den = ctx.scf_get()
        scale = lambda a: float(np.max(np.abs(a)))
        parity[name] = {"phi_max_rel": float(np.max(np.abs(den - ref["density"]))) / scale(ref["density"]),
                        "field_max_rel": float(np.max(np.abs(field - ref["field"]))) / scale(ref["field"]),
                        "residual_rel": abs(st.accuracy - ref["accuracy"]) / ref["accuracy"]}
    all_core = None
    if ok:
        per = [float(o.strip().splitlines()[-1]) for o in outs]
        all_core = {"value": nproc / wall, "unit": "SCF iters/s", "cores": nproc, "processes": nproc,
                    "wall_seconds": wall, "seconds_per_process_median": float(np.median(per))}
    return dict(value=1.0 / t, unit="SCF iters/s", cores=1, kind="port",
                sample="1 SCF iteration (2x200 propagator steps, contour integral, mean fields, mixing) of the "
                       "64^3 workload by the single-threaded C FP64 oracle (oracle/scft_oracle.c; recursive "
                       "mixed-radix DFT, not FFTW)",
                seconds=t, host=host, all_cores=all_core,
                reference_buildable=bool((host["gfortran"] or host["flang"]) and host["libfftw3"]),
                gpu_vs_oracle_same_iteration=parity)


def profile_run(adscf_amd, fields, what, iters, tune=False):
    """A workload without timing legs, for rocprofv3 kernel traces and counter runs:
    metric32 / metric64: `iters` SCF iterations of the 64^3 x 200 diblock workload above
    (the first m + 1 = 21 iterations mix without the Anderson solve; later ones run every
    kernel of the steady-state iteration); large32 / large64: `iters` iterations of the
    128^3 x 400 ABA triblock (lamellar seed); batch: `iters` batched iterations of the
    64-instance sweep (FP32-mixed).  With ADSCF_GRAPH=0 in the environment the iterations
    are stream launches instead of graph replays; with tune the metric workload first runs
    the autotuner (many captured sweeps, each replayed several times)."""
    if what in ("metric32", "metric64"):
        prec = adscf_amd.FP32 if what == "metric32" else adscf_amd.FP64
        w0 = fields.cylinder_seed(N, L, CHIN, sigma=0.2, seed=2007)
        with adscf_amd.Context(N, N, N, L, L, L, 1.0 / NS, precision=prec) as ctx:
            ctx.set_interactions(KAPPA, [[0.0, CHIN], [CHIN, 0.0]], [1.0, 1.0])
            ctx.scf_setup([1, 2], fields.block_segments([FA, 1 - FA], NS), ensemble=1, rho=1.0, mixing_dim=M,
                          lam_mix=LAM)
            ctx.scf_set_field(w0)
            if tune:
                print(ctx.autotune(reps=3), flush=True)
            st = ctx.scf_iterate(iters, tol=-1.0)
    elif what in ("large32", "large64"):
        n, Lb, ns = 128, 6.0, 400
        prec = adscf_amd.FP32 if what == "large32" else adscf_amd.FP64
        with adscf_amd.Context(n, n, n, Lb, Lb, Lb, 1.0 / ns, precision=prec) as ctx:
            ctx.set_interactions(KAPPA, [[0.0, 25.0], [25.0, 0.0]], [1.0, 1.0])
            ctx.scf_setup([1, 2, 1], fields.block_segments([0.25, 0.5, 0.25], ns), ensemble=1, rho=1.0,
                          mixing_dim=M, lam_mix=LAM)
            ctx.scf_set_field(fields.lamellar_seed(n, Lb, 25.0, sigma=0.1, seed=2007))
            st = ctx.scf_iterate(iters, tol=-1.0)
    elif what == "batch":
        inst, w = fields.batched_instances(seed=2007)
        with adscf_amd.Batch(inst, mixing_dim=M, lam_mix=LAM, precision=adscf_amd.FP32) as b:
            b.set_fields(w)
            st = b.iterate(iters, tol=-1.0)
    else:
        raise SystemExit("unknown --profile-run workload %r" % what)
    print("done", what, iters, st, flush=True)


def pmc_summary(db, pattern="."):
    """Per-kernel summary of the hardware counters in a rocprofv3 database (rocpd, SQLite):
    for every kernel whose name matches `pattern`, the dispatch count and the median and
    maximum of each counter over its dispatches (summed over hardware instances per dispatch,
    as rocprofv3 reports *_sum counters).  Turns `rocprofv3 --pmc FETCH_SIZE ...` and
    `--pmc WRITE_SIZE ...` runs into per-kernel traffic figures."""
    import collections
    import re
    import sqlite3
    import statistics

    con = sqlite3.connect("file:%s?mode=ro" % db, uri=True)
    rows = con.execute("select name, dispatch_id, counter_name, counter_value from pmc_events").fetchall()
    per = collections.defaultdict(lambda: collections.defaultdict(lambda: collections.defaultdict(float)))
    rx = re.compile(pattern)
    for name, did, cname, val in rows:
        if rx.search(name):
            per[name][cname][did] += val
    for name in sorted(per):
        for cname, vals in sorted(per[name].items()):
            v = list(vals.values())
            print("%s\tdispatches=%d\tmedian=%.1f\tmax=%.1f\t%s" % (cname, len(v), statistics.median(v), max(v),
                                                                      name[:100]))


def trace_summary(db, skip=0.0):
    """Per-kernel durations of a rocprofv3 kernel-trace database (rocpd, SQLite), as a
    markdown table: calls, mean and median duration, total, share of the summed kernel
    time, grid, VGPRs and LDS; the first `skip` fraction of the dispatches (warm-up) is left
    out."""
    import collections
    import re
    import sqlite3
    import statistics

    con = sqlite3.connect("file:%s?mode=ro" % db, uri=True)
    rows = con.execute("select name, start, end, grid_x, workgroup_x, vgpr_count, lds_size from kernels "
                       "order by start").fetchall()
    rows = rows[int(len(rows) * skip):]
    agg, meta = collections.defaultdict(list), {}
    for name, t0, t1, gx, wx, vg, lds in rows:
        agg[name].append(t1 - t0)
        meta[name] = (gx // max(wx, 1), wx, vg, lds)
    tot = sum(sum(v) for v in agg.values())
    print("| kernel | calls | mean us | median us | total ms | share | workgroups x threads | VGPR | LDS B |")
    print("|:--|--:|--:|--:|--:|--:|--:|--:|--:|")
    for name, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        wg, wx, vg, lds = meta[name]
        print("| `%s` | %d | %.2f | %.2f | %.3f | %.1f%% | %d x %d | %d | %d |" % (
            re.sub(r"\(.*", "", name)[:70], len(v), statistics.mean(v) / 1e3, statistics.median(v) / 1e3,
            sum(v) / 1e6, 100.0 * sum(v) / tot, wg, wx, vg, lds))
    if rows:
        print("\nspan of the listed dispatches %.3f ms, kernel time %.3f ms" % ((rows[-1][2] - rows[0][1]) / 1e6,
                                                                              tot / 1e6))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--converge-max-iter", type=int, default=4000)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the 128^3 x 400 configuration")
    ap.add_argument("--no-hipfft", action="store_true", help="skip the hipFFT comparator")
    ap.add_argument("--no-ddft", action="store_true",
                    help="skip the Main_Multiblock DDFT and Main_Blend EPD timings")
    ap.add_argument("--autotune", action="store_true",
                    help="re-tune the line-pass schedules first (the defaults are the tuner's MI355X choice)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 256 MiB device-to-device copy (counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--profile-run", metavar="WORKLOAD",
                    help="only run WORKLOAD (metric32, metric64, large32, large64, batch) for --iters iterations, "
                         "for rocprofv3 traces and counters")
    ap.add_argument("--iters", type=int, default=30, help="iterations of --profile-run")
    ap.add_argument("--pmc-summary", metavar="DB", help="summarise the counters of a rocprofv3 database and exit")
    ap.add_argument("--pmc-regex", default=".", help="kernel-name filter of --pmc-summary")
    ap.add_argument("--trace-summary", metavar="DB", help="summarise a rocprofv3 kernel-trace database and exit")
    ap.add_argument("--trace-skip", type=float, default=0.0,
                    help="fraction of the first dispatches --trace-summary leaves out (warm-up)")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(args.pmc_summary, args.pmc_regex)
        return
    if args.trace_summary:
        trace_summary(args.trace_summary, args.trace_skip)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the SCFT path runs on a single MI355X; "
              "runs on more than one GPU are not supported", file=sys.stderr)
        sys.exit(2)
    import adscf_amd
    from adscf_amd import fields

    if args.profile_run:
        if args.copy_probe:
            copy_probe()
        profile_run(adscf_amd, fields, args.profile_run, args.iters, args.autotune)
        return
    w0 = fields.cylinder_seed(N, L, CHIN, sigma=0.2, seed=2007)
    ng = N ** 3
    if args.copy_probe:
        copy_probe()
    ctx64, r64 = run_gpu(adscf_amd, fields, adscf_amd.FP64, args.steps, args.warmup, w0, not args.no_hipfft,
                         args.autotune)
    ctx64.close()
    ctx32, r32 = run_gpu(adscf_amd, fields, adscf_amd.FP32, args.steps, args.warmup, w0, not args.no_hipfft,
                         args.autotune)
    ctx32.close()
    npass, npass64 = r32["npass"], r64["npass"]
    sweep_bytes32 = 2 * NS * 3 * ng * 4
    sweep_bytes64 = 2 * NS * 3 * ng * 8
    ach32 = sweep_bytes32 / (r32["ms_sweep"] * 1e-3) / 1e9
    ach64 = sweep_bytes64 / (r64["ms_sweep"] * 1e-3) / 1e9
    traffic = profiled_traffic(r32["plane"]["active"])
    traffic64 = profiled_traffic(r64["plane"]["active"])
    biter32 = algorithmic_bytes_iter(ng, NS, M, 4)
    biter64 = algorithmic_bytes_iter(ng, NS, M, 8)
    line = {
        "metric": METRIC,
        "value": r32["iters_per_s"],
        "unit": "SCF iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_iter"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed (f32 propagators/FFTs; f64 Q, F, phi sums, Anderson)",
        "data": "synthetic (seeded cylinder fields, MT19937 seed 2007)",
        "config": {"workload": "AB diblock melt SCF, 64^3 grid, box 4.5 Rg, Ns=200 (60/140), f_A=0.3, chiN=20, "
                               "kappa=25, Anderson m=20, lam_mix=0.5",
                   "grid": [N, N, N], "contour_steps": NS, "mixing_dim": M, "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": ach32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach32 / HBM_PEAK_GBS, "traffic": traffic.get("fp32") if traffic else None,
                     "traffic_source": TRAFFIC_SOURCES[r32["plane"]["active"]] if traffic else None,
                     "kernel": sweep_kernels("fp32", r32["plane"]),
                     "launches_per_sweep": npass,
                     "bytes_per_launch": sweep_bytes32 / npass,
                     "us_per_launch": r32["ms_sweep"] * 1e3 / npass},
        "schedule": r32["schedule"],
        "fp64": {"iters_per_s": r64["iters_per_s"], "ms_per_iter": r64["ms_per_iter"], "schedule": r64["schedule"],
                 "roofline_achieved_GBs": ach64, "roofline_frac": ach64 / HBM_PEAK_GBS,
                 "kernel": sweep_kernels("fp64", r64["plane"]), "launches_per_sweep": npass64,
                 "us_per_pass": r64["ms_sweep"] * 1e3 / npass64,
                 "traffic_per_pass": traffic64.get("fp64") if traffic64 else None,
                 "iteration_bytes_frac": biter64 * r64["iters_per_s"] / (HBM_PEAK_GBS * 1e9)},
        "fp32mixed_iteration_bytes_frac": biter32 * r32["iters_per_s"] / (HBM_PEAK_GBS * 1e9),
        "speedup_fp32mixed_vs_fp64": r32["iters_per_s"] / r64["iters_per_s"],
        "contour_steps_per_s": r32["iters_per_s"] * 2 * NS,
        "timing_valid": not (r32["stopped_early"] or r64["stopped_early"]),
    }
    if not args.no_hipfft:
        # one propagator step of one chain: ours = sweep time / (2 chains x Ns steps);
        # hipFFT = E * c2r(K * r2c(E * q)) with hipFFT plans and two pointwise kernels
        line["hipfft_comparator"] = {
            "us_per_chain_step": {"fp32": {"adscf": r32["ms_sweep"] * 1e3 / (2 * NS), "hipfft": r32["hipfft_step_us"]},
                                  "fp64": {"adscf": r64["ms_sweep"] * 1e3 / (2 * NS), "hipfft": r64["hipfft_step_us"]}}}
    if not args.no_batched:
        line["batched"] = run_batched(adscf_amd, fields, max(5, args.steps // 5), 2)
    if not args.no_large:
        line["large"] = run_large(adscf_amd, fields, max(5, args.steps // 10), 3)
    if not args.no_ddft:
        line["ddft"] = run_ddft(adscf_amd, fields)
        line["epd"] = run_epd(adscf_amd, fields)
    if not args.no_converge:
        line["convergence"] = converge(adscf_amd, fields, w0, args.converge_max_iter)
    hbm = measured_hbm(adscf_amd)
    hbm["line_pass_frac_of_copy"] = ach32 / hbm["copy_GBs"]
    line["hbm_measured"] = hbm
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(adscf_amd, fields, w0)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
