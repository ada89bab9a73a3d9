"""Benchmark: time to relative residual 1e-10 of the 256^3 vec3 projection solve.

Workload (one "step" = one full solve from x = 0): the transformed-coordinate
projection operator of trans_sim_3d for simple shear T = [[1,0,0.5],[0,1,0],
[0,0,1]], 256^3 vertices, fully periodic, mu = 1, lamb = 1.5, h = 1/256,
dt = 1; RHS = Gaussian random field (seed 2, correlation length 8 cells, unit
variance per component, mean removed).  V(2,2) cycles with the reference's
level schedule and Galerkin coarse operators (6 levels, 256 -> 8).

The timed solve (`value`) is GMRES-IR: iterative refinement in FP64 whose
every correction is an FP32 GMRES(10) solve, right-preconditioned by one
FP32 V-cycle, run until its residual estimate fell by 1e-4; the true FP64
residual decides after each correction.  Timed beside it on the same
system: the plain FP64 multigrid (V-cycles to 1e-10), the stationary FP32+IR
solve (three FP32 V-cycles per FP64 refinement step), flexible GMRES(10) in
FP64 preconditioned by one FP32 V-cycle (FGMRES-IR), and flexible GMRES(10)
entirely in FP64 preconditioned by one FP64 V-cycle, which isolates what the
FP32 precision gains from what the Krylov acceleration gains
(`precision_speedup` = FP64 GMRES / GMRES-IR; `stationary_speedup` = FP64
multigrid / stationary FP32+IR).  The error of the timed solution against
the FP64 one is reported.  Inputs are resident in HBM when the timed region
starts.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the dominant kernel (fine-level FP32 fused
# smoother, one z parity of the 256^3 workload), from two rocprofv3 --pmc
# runs of `python3 bench.py --steps 1 --warmup 0 --no-cpu --no-graph`:
# FETCH_SIZE 176.49 MiB x 2.0 + WRITE_SIZE 96.05 MiB x 1.0 per dispatch
# (average of the 722 dispatches that stage every plane; the 134 first
# sweeps of V-cycles that start from x = 0 skip the zero planes and read
# 86.9 MiB on average).  The factors come from the plain copy kernel
# k_convert<double,double> of the same run (3 x 256^3 doubles = 384 MiB each
# way), which reads FETCH_SIZE 192.0 MiB and WRITE_SIZE 384.0 MiB.  Summary:
# profiles/mg_solve_256_trans_kernels.txt.
TRAFFIC_256_BYTES_PER_LAUNCH = (176.49 * 2.0 + 96.05) * 2 ** 20


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_problem(n):
    from stzpp_amd import fields
    from stzpp_amd import multigrid as MG

    T = np.array([[1.0, 0.0, 0.5], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    h = 1.0 / n
    st = MG.trans_stencil(T, 1.0, 1.5, h, h, h, 1.0)
    mg = MG.Multigrid(n, n, n, (1, 1, 1))
    mg.setup_matrices_const(st, None)
    b = fields.grf_vec3((n, n, n), 8.0, 2)
    b -= b.mean(axis=0)
    return mg, b, st


def host_cpu():
    """(CPU model, nproc, threads for the CPU baseline): the OpenMP thread
    count this process may use (OMP_NUM_THREADS within the affinity mask),
    at most 16."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = nproc
    try:
        want = int(os.environ.get("OMP_NUM_THREADS", "") or aff)
    except ValueError:
        want = aff
    return model, nproc, max(1, min(want, aff, 16))


def rank_grid(threads, dims, prd):
    """The reference's subdomain grid for the largest rank count <= threads
    that calc_processor_division accepts (src/common.cpp:7-60)."""
    import oracle as O

    for P in range(threads, 0, -1):
        r = O.processor_division(P, *dims, prd)
        if r:
            return P, r
    return 1, (1, 1, 1)


def calibration():
    """Timings of the compiled reference beside the oracle on one machine
    (written by oracle/ref/time_reference.py), or None."""
    path = os.path.join(ROOT, "oracle", "ref", "reference_timings.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_baseline(n, st, b, gpu_fp64_cycles):
    """The reference's algorithm (lexicographic block GS on the subdomain grid
    of P ranks, reference schedule) in the C oracle on P OpenMP threads of
    this host: one timed V-cycle x the oracle's V-cycle count to 1e-10 for
    that grid (measured once by oracle/ref/time_reference.py; the GPU's FP64
    count where none is recorded).  The reference-equivalent time scales it by
    the reference/oracle V-cycle ratio measured on one machine at equal
    ranks."""
    import oracle as O

    model, nproc, threads = host_cpu()
    P, ranks = rank_grid(threads, (n, n, n), (1, 1, 1))
    O.set_threads(P)
    orc = O.OracleMG(n, n, n, (1, 1, 1), mode=O.PLANE)
    orc.set_ops(0, O.plane_operator(n, 1, st, None))
    orc.rat()
    orc.params(smoother=0)
    orc.set_ranks(ranks)
    orc.set_fields(0, np.zeros((n ** 3, 3)), b)
    t0 = time.perf_counter()
    orc.vcycle(1)
    cyc_s = time.perf_counter() - t0
    orc.close()
    cal = calibration()
    cycles, src = gpu_fp64_cycles, "the GPU FP64 multigrid's count"
    ratio = None
    if cal and n == 256:
        rec = cal.get("trans_256_cycles_to_1e-10", {}).get(str(P))
        if rec and rec.get("cycles"):
            cycles, src = rec["cycles"], f"the oracle's count for ranks {tuple(rec['ranks'])}"
        t = cal.get("trans_256", {})
        ratio = t.get("calibration_8" if P > 1 else "calibration_1")
    out = {
        "value": cyc_s * cycles,
        "unit": "s (time to 1e-10: one measured V-cycle x cycle count)",
        "cores": P,
        "kind": "port",
        "sample": f"one lexicographic V-cycle of the C oracle at {n}^3 on {P} threads as the reference's "
                  f"{ranks[0]}x{ranks[1]}x{ranks[2]} rank grid took {cyc_s:.2f} s; x {cycles} cycles ({src})",
        "cpu_model": model,
        "nproc": nproc,
        "ranks": list(ranks),
        "vcycle_s": cyc_s,
        "cycles": cycles,
    }
    if ratio:
        out["reference_equivalent_s"] = cyc_s * cycles * ratio
        out["calibration"] = {
            "reference_over_oracle": ratio,
            "measured_on": cal.get("cpu_model"),
            "ranks": 8 if P > 1 else 1,
            "source": "oracle/ref/reference_timings.json (reference trans_sim_3d V-cycle at 256^3 vs the oracle's)",
        }
    return out


def step_roofline(sim, n_points):
    """Roofline of the step's dominant kernel, the fine-level FP64 fused
    Gauss-Seidel sweep of its projection solver (two launches, one per z
    parity, 72 algorithmic bytes per updated point): after five unrecorded
    sweeps, 20 back-to-back sweeps under one HIP event pair on the solver's
    stream, and an event pair around every launch beside it."""
    import ctypes

    from stzpp_amd import _lib

    L = _lib.load()
    h = L.stz_multigrid(sim.h)
    v = ctypes.c_double()
    if L.stzmg_time_kernel(h, 0, 0, 5, ctypes.byref(v)):
        return None
    if L.stzmg_time_kernel(h, 0, 0, 20, ctypes.byref(v)):
        return None
    launch_s = v.value * 1e-3 / 2
    if L.stzmg_time_kernel(h, 4, 0, 20, ctypes.byref(v)):
        return None
    launch_ev_s = v.value * 1e-3
    alg = 36.0 * n_points
    achieved = alg / launch_s / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": "k_gs_march<double, ...> (fine-level fused GS of the projection, one z parity)",
            "alg_bytes_per_launch": alg, "launch_us": launch_s * 1e6, "launch_us_event_pairs": launch_ev_s * 1e6}


def oracle_step_baseline(N, prd, params, state0, dt, family=None, Ut=0.0):
    """One step of the oracle (the reference's step and projection loop with
    the lexicographic sweep on the subdomain grid of P ranks, P OpenMP
    threads) from the workload's initial state.  The advection pass of the
    oracle is sequential (it keeps the reference's in-place ad_Dpl order)."""
    import oracle as O

    model, nproc, threads = host_cpu()
    nzv = N[2] + (0 if prd[2] else 1)
    P, ranks = rank_grid(threads, (N[0], N[1], nzv), prd)
    O.set_threads(P)
    box = (params["ax"], params["bx"], params["ay"], params["by"], params["az"], params["bz"])
    orc = O.OracleShear(N, params, box=box, prd=prd)
    if family:
        orc.set_transform(family, Ut)
    orc.set_ranks(ranks)
    orc.set_state(state0["t"], state0["u"], state0["sig"], state0["chi"])
    t0 = time.perf_counter()
    orc.step(dt, 1)
    el = time.perf_counter() - t0
    cyc = orc.vcycles
    orc.close()
    return {"value": 1.0 / el, "unit": "steps/s", "cores": P, "kind": "port",
            "sample": f"the first step of the workload by the C oracle on {P} threads (rank grid "
                      f"{ranks[0]}x{ranks[1]}x{ranks[2]}): {el:.2f} s, {cyc} V-cycles",
            "cpu_model": model, "nproc": nproc, "ranks": list(ranks), "step_s": el, "vcycles": cyc}


def step_workload(args):
    """QS steps of shear_sim_3d at 192 x 192 x 96 cells (MG 192 x 192 x 97),
    config-0 material, chi0 = 0.09 + 0.01 GRF(seed 3, l = 4 cells), dt = 0.005,
    the reference's projection loop; value = steps per second."""
    from stzpp_amd import shear_sim as SS

    Nx, Ny, Nz = 192, 192, 96
    p = SS.config0_params(Nx)
    sim = SS.ShearSim(Nx, Ny, Nz, params=p)
    # chi0 from the reference's random initial condition (cube-cutoff
    # Gaussian, l = 4 cells, cut 20) on the GPU, noise seed 3
    sim.set_random_initial_conditions(gauss_fac=5, conv_l=4.0, chi_avg=0.09, sigma=0.01, seed=3)
    state0 = sim.state()
    dt = 0.005
    log(f"[bench] step workload {Nx}x{Ny}x{Nz}, warmup {args.warmup}")
    graph = not args.no_graph
    solve = {"ref": None, "fp64": SS.SOLVE_RELRES, "fp32ir": SS.SOLVE_FP32_IR, "fgmres": SS.SOLVE_FP32_GMRES,
             "gmresir": SS.SOLVE_FP32_GMRES_IR}[
        args.qs_solve]
    for _ in range(args.warmup):
        sim.step_forward_qs(dt, 1, use_graph=graph, solve=solve, rtol=1e-12)
    t0 = time.perf_counter()
    cyc, solve_ms = 0, 0.0
    for _ in range(args.steps):
        st = sim.step_forward_qs(dt, 1, use_graph=graph, solve=solve, rtol=1e-12)
        cyc += st["vcycles"]
        solve_ms += st["solve_ms"]
    el = time.perf_counter() - t0
    how = {"ref": "reference projection loop", "fp64": "FP64 multigrid to 1e-12",
           "fp32ir": "FP32+IR to 1e-12", "fgmres": "FGMRES-IR to 1e-12",
           "gmresir": "GMRES-IR to 1e-12"}[args.qs_solve]
    out = {
        "metric": f"quasi-static steps per second, shear_sim_3d 192x192x96 cells, {how}",
        "value": args.steps / el, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "shear_sim_3d QS step 192x192x96 cells (MG 192x192x97, z walled), config-0 "
                               "material, chi0 GRF seed 3 l=4, dt 0.005",
                   "projection": how, "vcycles_per_step": cyc / args.steps,
                   "solve_ms_per_step": solve_ms / args.steps, "v_cycle_graph": graph},
        "roofline": step_roofline(sim, Nx * Ny * (Nz + 1)), "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle step, all cores)")
        out["cpu_baseline"] = oracle_step_baseline((Nx, Ny, Nz), (1, 1, 0), p, state0, dt)
    print(json.dumps(out), flush=True)


def grf_workload(args):
    """The random initial chi at 192 x 192 x 96 cells, z walled: the
    reference's cube-cutoff Gaussian (l = 4 cells, gauss_fac 5, cut 20),
    chi = 0.09 + 0.01 GRF, noise drawn on the GPU (seed 3).  value = cells per
    second of its four kernels (noise draw, x, y and z passes), each timed
    with HIP events on the library's stream; the API call with its
    allocations and the copy-out is timed beside it.  CPU baseline: the
    reference's brute-force convolution (the oracle) on one thread at
    sampled cells of the same grid."""
    from stzpp_amd import grf as G

    N, prd = (192, 192, 96), (1, 1, 0)
    p = dict(gauss_fac=5, conv_l=4.0, chi_avg=0.09, sigma=0.01, seed=3)
    cut, nn, chi_1 = G.geometry(N, prd, **p)
    nx, ny, nz = N
    nxt, nyt, nzt = G.noise_shape(N, prd, **p)
    nc = nx * ny * nz
    log(f"[bench] grf workload {N}, cut {cut}, noise grid {nxt}x{nyt}x{nzt}")
    for _ in range(args.warmup):
        G.grf_field(N, prd, **p)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        chi = G.grf_field(N, prd, **p)
    call_s = (time.perf_counter() - t0) / args.steps
    ms = G.time_kernels(N, prd, reps=20, **p)
    total_s = float(ms.sum()) * 1e-3
    names = ["k_grf_noise (Philox4x64 + polar Box-Muller)", "k_grf_x", "k_grf_axis_lds (y pass)",
             "k_grf_axis_lds (z pass)"]
    # algorithmic bytes per launch: each input element read once and each
    # output written once (windows are re-read from LDS)
    alg = [8.0 * nn, 8.0 * (nn + nx * nyt * nzt), 8.0 * (nx * nyt * nzt + nx * ny * nzt),
           8.0 * (nx * ny * nzt + nc)]
    dom = int(np.argmax(ms))
    achieved = alg[dom] / (ms[dom] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        import ctypes

        import oracle as O

        try:
            ctypes.CDLL("libgomp.so.1").omp_set_num_threads(1)
        except OSError:
            pass
        noise = O.grf_noise(3, nn)
        rng = np.random.default_rng(1)
        idx = rng.integers(0, nc, 16384)
        t0 = time.perf_counter()
        O.grf_convolve_points(N, prd, cut, 1 / p["conv_l"], p["chi_avg"], chi_1, noise, idx)
        el = time.perf_counter() - t0
        cpu = {"value": idx.size / el, "unit": "cells/s", "cores": 1, "kind": "port",
               "sample": f"brute-force cube convolution (the reference's convolve, {(2 * cut + 1) ** 3} taps per "
                         f"cell) at {idx.size} random cells of the same grid: {el:.1f} s; the whole field "
                         f"would take {nc / (idx.size / el) / 3600:.1f} h"}
    out = {
        "metric": "random initial chi: cells per second, 192x192x96 cells, reference cube-cutoff Gaussian l=4",
        "value": nc / total_s, "unit": "cells/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": total_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "set_random_initial_conditions 192x192x96 cells, z walled (noise padded by the "
                               "cutoff), gauss_fac 5, conv_l 4 (cut 20), chi_avg 0.09, sigma 0.01, seed 3",
                   "kernel_ms": dict(zip(["noise", "x", "y", "z"], [float(v) for v in ms])),
                   "api_call_ms": call_s * 1e3, "chi_mean": float(chi.mean()), "chi_std": float(chi.std())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": names[dom],
                     "alg_bytes_per_launch": alg[dom], "launch_us": float(ms[dom]) * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def adjoint_workload(args):
    """Gradient of J = mean (ad_Dpl/dt)^2 after 20 QS steps at 128^3 cells
    (MG 128 x 128 x 129) with respect to chi0 and (s_y, eps0, chi_inf);
    checkpoints every 5 steps, projections to relative residual 1e-12;
    value = seconds per gradient (forward + reverse)."""
    from stzpp_amd import fields
    from stzpp_amd import shear_sim as SS

    N = 128
    p = SS.config0_params(N)
    p["eps0"] = 0.3
    sim = SS.ShearSim(N, N, N, params=p)
    chi0 = 0.09 + 0.01 * fields.grf((N, N, N), 3.0, 4).ravel()
    nc = N ** 3
    sig0 = np.zeros((nc, 6))
    sig0[:, 1] = 1.05
    ts = []
    for q in range(args.warmup + args.steps):
        sim.set_initial_conditions(chi0, sig=sig0)
        t0 = time.perf_counter()
        J, g, gp = sim.adjoint(0.01, 20, checkpoint_every=5, rtol=1e-12)
        el = time.perf_counter() - t0
        log(f"[bench] adjoint {q}: {el:.2f} s, J {J:.4e}")
        if q >= args.warmup:
            ts.append(el)
    el = float(np.mean(ts))
    out = {
        "metric": "seconds per adjoint gradient, 20 QS steps at 128^3 cells, checkpoints every 5",
        "value": el, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "adjoint of 20 shear_sim_3d QS steps, 128^3 cells (MG 128x128x129), "
                               "eps0 0.3, s12 pre-stress 1.05, chi0 GRF seed 4 l=3, dt 0.01, RELRES 1e-12",
                   "loss": J, "grad_params": gp.tolist()},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def trans_workload(args):
    """QS steps of trans_sim_3d at the trans_shear_sim driver's size,
    256 x 256 x 128 cells, fully periodic (src/trans_shear_sim.cpp:19-21,
    :96): pure shear T = diag(e^{Ut t}, 1, e^{-Ut t}) at Ut = 0.5, config-0
    material with eps0 = 0.3, chi0 = 0.09 + 0.01 GRF (seed 3, l = 4), dt =
    0.005, the reference's projection loop (norm on the std_iters schedule).
    Every step rebuilds the transformed operator from T(t + dt), its Galerkin
    hierarchy (RAT) and inverse tables on the GPU before the solve.  value =
    steps per second."""
    from stzpp_amd import shear_sim as SS

    N = (256, 256, 128)
    p = SS.config0_params(N[0])
    p["eps0"] = 0.3
    sim = SS.TransSim(*N, 0.5, family="pure_shear", params=p, prd=(1, 1, 1))
    sim.set_random_initial_conditions(gauss_fac=5, conv_l=4.0, chi_avg=0.09, sigma=0.01, seed=3)
    state0 = sim.state()
    dt = 0.005
    graph = not args.no_graph
    log(f"[bench] trans workload {N}, warmup {args.warmup}")
    for _ in range(args.warmup):
        sim.step_forward_qs(dt, 1, use_graph=graph)
    t0 = time.perf_counter()
    cyc, solve_ms = 0, 0.0
    for _ in range(args.steps):
        st = sim.step_forward_qs(dt, 1, use_graph=graph)
        cyc += st["vcycles"]
        solve_ms += st["solve_ms"]
    el = time.perf_counter() - t0
    out = {
        "metric": "transformed QS steps per second, trans_sim_3d 256x256x128 cells fully periodic, per-step "
                  "operator and Galerkin rebuild, reference projection loop",
        "value": args.steps / el, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "trans_sim_3d QS step 256x256x128 cells (MG 256x256x128, periodic), pure shear "
                               "Ut 0.5, config-0 material eps0 0.3, chi0 GRF seed 3 l=4, dt 0.005",
                   "vcycles_per_step": cyc / args.steps, "solve_ms_per_step": solve_ms / args.steps,
                   "v_cycle_graph": graph},
        "roofline": step_roofline(sim, N[0] * N[1] * N[2]), "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle step, all cores)")
        out["cpu_baseline"] = oracle_step_baseline(N, (1, 1, 1), p, state0, dt, family="pure_shear", Ut=0.5)
    print(json.dumps(out), flush=True)


def output_workload(args):
    """The reference's output cubes (write_files modes 0-12: u v w, the six
    stresses, p, dev, chi, tem) of a 192 x 192 x 96 state (vertices
    192 x 192 x 97, z walled): float32, x slowest and z fastest, through
    stz_output_cube (device kernel, copy to host).  value = cubes per second
    of the API calls; each mode's kernel is timed on the device beside it.
    CPU baseline: the oracle's numpy restatement of the same cubes on one
    core."""
    from stzpp_amd import shear_sim as SS

    Nx, Ny, Nz = 192, 192, 96
    nc, nv = Nx * Ny * Nz, Nx * Ny * (Nz + 1)
    sim = SS.ShearSim(Nx, Ny, Nz, params=SS.config0_params(Nx))
    rng = np.random.default_rng(4)
    u0, sig0, chi0 = 1e-3 * rng.standard_normal((nv, 3)), rng.standard_normal((nc, 6)), 0.09 + 0.01 * rng.random(nc)
    sim.set_initial_conditions(chi0, u=u0, sig=sig0, t=0.25)
    modes = list(range(13))
    log(f"[bench] output workload {Nx}x{Ny}x{Nz}, {len(modes)} cubes")
    for _ in range(args.warmup):
        for m in modes:
            sim.output_cube(m)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for m in modes:
            sim.output_cube(m)
    el = time.perf_counter() - t0
    ms = np.array([sim.time_cube_kernel(m, reps=20) for m in modes])
    # algorithmic bytes per launch: the FP64 values read once, the FP32 cube
    # written once (p reads three stresses, dev six)
    alg = np.array([nv * 12.0 if m <= 2 else nc * (28.0 if m == 9 else (52.0 if m == 10 else 12.0))
                    for m in modes])
    dom = int(np.argmax(ms))
    achieved = alg[dom] / (ms[dom] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        import oracle as O

        s = sim.state()
        t0 = time.perf_counter()
        n_cpu = 0
        while time.perf_counter() - t0 < 10.0 and n_cpu < len(modes):
            O.cube_values(modes[n_cpu], (Nx, Ny, Nz), (1, 1, 0), s["u"], s["sig"], s["chi"])
            n_cpu += 1
        el_cpu = time.perf_counter() - t0
        cpu = {"value": n_cpu / el_cpu, "unit": "cubes/s", "cores": 1, "kind": "port",
               "sample": f"{n_cpu} cubes of the same state by the oracle's numpy restatement: {el_cpu:.1f} s"}
    out = {
        "metric": "output cubes per second (write_files modes 0-12, float32 to host), 192x192x96 cells",
        "value": args.steps * len(modes) / el, "unit": "cubes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64->f32", "data": "synthetic",
        "config": {"workload": "output_cube modes 0-12 of a 192x192x96 state (vertices 192x192x97, z walled)",
                   "kernel_ms": {str(m): float(v) for m, v in zip(modes, ms)}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": f"k_cube (mode {modes[dom]})",
                     "alg_bytes_per_launch": float(alg[dom]), "launch_us": float(ms[dom]) * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


DIRECT_N = (192, 192, 96)
# HBM bytes per launch of the direct step's cell and vertex kernels at
# 192 x 192 x 96, from two rocprofv3 --pmc runs of
# `python3 bench.py --workload direct --steps 20 --warmup 2 --no-cpu`
# (FETCH_SIZE x 2.0 + WRITE_SIZE, the copy-kernel factors of the solve
# profile): cell 445.7 MiB x 2 + 189.0 MiB, vertex 372.3 MiB x 2 + 163.7 MiB.
# Summary: profiles/direct_step_192x192x96_kernels.txt.
TRAFFIC_DIRECT_BYTES_PER_LAUNCH = {"k_direct_cell": (445.7 * 2 + 189.0) * 2 ** 20,
                                   "k_direct_vertex": (372.3 * 2 + 163.7) * 2 ** 20}


def direct_workload(args):
    """Direct (explicit dynamics) steps of shear_sim_3d at 192 x 192 x 96
    cells (vertices 192 x 192 x 97, z walled): config-0 material, density 1,
    viscous damping 0.35, the reference's dt = min(tmult dx^2 / (6 kappa),
    dx / 4) with tmult 0.5; chi0 = 0.09 + 0.01 GRF (seed 3, l = 4 cells),
    s12 pre-stress 1.05, at rest: every cell yields in the first step and
    relaxes onto the yield surface within it, so the timed steps run near
    yield (the last step's yielding count is reported).  value = steps per
    second over K steps in one call.
    The four kernels are timed with HIP events on the handle's stream.  CPU
    baseline: the oracle's step (bit-exact to the reference's step_forward
    where no exp enters) on one thread, on the same grid."""
    from stzpp_amd import shear_sim as SS

    Nx, Ny, Nz = DIRECT_N
    p = SS.config0_params(Nx)
    p.update(rho=1.0, kappa=0.35, tmult=0.5)
    nc, nv = Nx * Ny * Nz, Nx * Ny * (Nz + 1)
    sim = SS.ShearSim(Nx, Ny, Nz, params=p)
    sig0 = np.zeros((nc, 6))
    sig0[:, 1] = 1.05
    sim.set_initial_conditions(np.zeros(nc), sig=sig0)
    sim.set_random_initial_conditions(gauss_fac=5, conv_l=4.0, chi_avg=0.09, sigma=0.01, seed=3)
    sim.init_ref()
    dt = sim.direct_dt()
    log(f"[bench] direct workload {Nx}x{Ny}x{Nz}, dt {dt:.4g}, warmup {args.warmup}")
    if args.warmup:
        sim.step_forward(dt, args.warmup)
    ms = sim.time_direct_kernels(dt, reps=20)
    t0 = time.perf_counter()
    st = sim.step_forward(dt, args.steps)
    el = time.perf_counter() - t0
    log(f"[bench] {args.steps} steps in {el * 1e3:.2f} ms; kernels (ms) {ms.round(4).tolist()}")
    names = ["k_direct_plastic", "k_direct_cell", "k_direct_vertex", "k_direct_plate"]
    # algorithmic bytes per launch: every input read once, every output
    # written once (neighbour re-reads come from cache)
    alg = [80.0 * nc, 144.0 * nc + 24.0 * nv, 96.0 * nv + 48.0 * nc, 48.0 * Nx * Ny]
    dom = int(np.argmax(ms))
    achieved = alg[dom] / (ms[dom] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        import oracle as O

        P = {k: p[k] for k in O.STEP_PARAM_ORDER}
        orc = O.OracleShear((Nx, Ny, Nz), P, prd=(1, 1, 0))
        orc.set_direct(p["rho"], p["kappa"])
        s = sim.state()
        orc.set_state(s["t"], s["u"], s["sig"], s["chi"])
        orc.init_ref()
        n_cpu = 6
        t0 = time.perf_counter()
        orc.step_direct(dt, n_cpu)
        el_cpu = time.perf_counter() - t0
        cpu = {"value": n_cpu / el_cpu, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"{n_cpu} direct steps of the oracle (the reference's step_forward restated in C, one "
                         f"thread) on the same 192x192x96 grid from the GPU's state after warmup: {el_cpu:.1f} s"}
    out = {
        "metric": "direct (explicit dynamics) steps per second, shear_sim_3d 192x192x96 cells",
        "value": args.steps / el, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "shear_sim_3d direct step 192x192x96 cells (vertices 192x192x97, z walled), "
                               "config-0 material, rho 1, kappa 0.35, tmult 0.5, chi0 GRF seed 3 l=4, "
                               "s12 pre-stress 1.05",
                   "dt": dt, "yielded_cells": st["yielded_cells"],
                   "kernel_ms": dict(zip(["plastic", "cell", "vertex", "plate"], [float(v) for v in ms]))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": TRAFFIC_DIRECT_BYTES_PER_LAUNCH.get(names[dom]) if (Nx, Ny, Nz) == DIRECT_N else None,
                     "kernel": names[dom], "alg_bytes_per_launch": alg[dom], "launch_us": float(ms[dom]) * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--inner", type=int, default=3, help="FP32 V-cycles per stationary refinement step")
    ap.add_argument("--restart", type=int, default=10, help="GMRES restart length")
    ap.add_argument("--inner-rtol", type=float, default=0.0,
                    help="residual reduction of each FP32 GMRES solve of GMRES-IR (0: library default)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch V-cycles kernel by kernel instead of replaying a hipGraph")
    ap.add_argument("--qs-solve", choices=("ref", "fp64", "fp32ir", "fgmres", "gmresir"), default="ref",
                    help="projection solve of the step workload: the reference's loop (default) or a "
                         "solve to 1e-12 relative residual")
    ap.add_argument("--workload", choices=("solve", "step", "adjoint", "grf", "direct", "trans", "output"), default="solve",
                    help="solve (default): the 256^3 projection solve; step: QS steps at 192x192x96; "
                         "adjoint: gradient of 20 QS steps at 128^3; grf: random initial chi at 192x192x96; "
                         "direct: explicit dynamics steps at 192x192x96; trans: transformed QS steps at 256x256x128; "
                         "output: the reference's output cubes at 192x192x96")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(one process drives one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "step":
        return step_workload(args)
    if args.workload == "adjoint":
        return adjoint_workload(args)
    if args.workload == "grf":
        return grf_workload(args)
    if args.workload == "direct":
        return direct_workload(args)
    if args.workload == "trans":
        return trans_workload(args)
    if args.workload == "output":
        return output_workload(args)

    from stzpp_amd import multigrid as MG

    n = args.size
    log(f"[bench] building {n}^3 problem")
    mg, b, st = make_problem(n)
    N = n ** 3
    zeros = np.zeros((N, 3))
    log(f"[bench] levels {[mg.dims(l) for l in range(mg.nlevels)]}")

    # Host->device upload once; each solve restarts from x = 0 with a
    # device-to-device copy of the resident RHS/zero fields.
    import ctypes

    mg.setup_fields(zeros, b)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipDeviceSynchronize.argtypes = []
    bytes3 = 8 * 3 * N
    dx0 = ctypes.c_void_p()
    db = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dx0), bytes3) == 0
    assert hip.hipMalloc(ctypes.byref(db), bytes3) == 0
    assert hip.hipMemset(dx0, 0, bytes3) == 0
    # resident SoA copy of the rhs (component-major), uploaded once
    bsoa = np.ascontiguousarray(b.T)
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert hip.hipMemcpy(db, bsoa.ctypes.data, bytes3, 1) == 0  # hipMemcpyHostToDevice

    def solve(mode, inner=None):
        mg.set_fields_device(dx0, db)
        hip.hipDeviceSynchronize()
        t0 = time.perf_counter()
        s = mg.solve(mode=mode, rtol=1e-10, inner_cycles=inner, max_cycles=400,
                     use_graph=not args.no_graph, restart=args.restart, inner_rtol=args.inner_rtol)
        hip.hipDeviceSynchronize()
        return time.perf_counter() - t0, s

    for w in range(args.warmup):
        log(f"[bench] warmup {w}")
        solve(MG.SOLVE_FP32_GMRES)
        solve(MG.SOLVE_FP32_GMRES_IR)
        solve(MG.SOLVE_FP32_IR, args.inner)
        solve(MG.SOLVE_RELRES)
        solve(MG.SOLVE_GMRES64)

    t64s, s64 = [], None
    for k in range(args.steps):
        t, s64 = solve(MG.SOLVE_RELRES)
        t64s.append(t)
    x64 = mg.x(0)
    log(f"[bench] fp64: {np.mean(t64s)*1e3:.1f} ms, {s64}")

    tirs, sir = [], None
    for k in range(args.steps):
        t, sir = solve(MG.SOLVE_FP32_IR, args.inner)
        tirs.append(t)
    log(f"[bench] fp32+ir (stationary): {np.mean(tirs)*1e3:.1f} ms, {sir}")

    t32s, s32 = [], None
    for k in range(args.steps):
        t, s32 = solve(MG.SOLVE_FP32_GMRES)
        t32s.append(t)
    x32 = mg.x(0)
    log(f"[bench] fp32+ir (fgmres): {np.mean(t32s)*1e3:.1f} ms, {s32}")

    tk64s, sk64 = [], None
    for k in range(args.steps):
        t, sk64 = solve(MG.SOLVE_GMRES64)
        tk64s.append(t)
    xk64 = mg.x(0)
    log(f"[bench] fp64 gmres (fp64 v-cycle preconditioner): {np.mean(tk64s)*1e3:.1f} ms, {sk64}")

    tgis, sgi = [], None
    for k in range(args.steps):
        t, sgi = solve(MG.SOLVE_FP32_GMRES_IR)
        tgis.append(t)
    xgi = mg.x(0)

    # Roofline of the dominant kernel, the fine-level FP32 fused smoother:
    # one sweep is two launches (one per z parity), each updating N/2 points
    # at 36 algorithmic bytes per point (read x, read r, write x, FP32 vec3),
    # so 18 N bytes per launch.  Timed right after the timed solves, on the
    # level's fields (which it overwrites; the solution was copied out above),
    # after five unrecorded sweeps that bring the device back to the busy
    # state of the solves.  The launch duration is that of 20 back-to-back
    # sweeps (40 launches, as the V-cycles issue them) under one HIP event
    # pair on the solver's stream, gaps included.  An event pair around
    # every launch (launch_us_event_pairs) also holds each launch's dispatch
    # and isolates it from the next, and reads a few microseconds longer.
    mg.time_kernel(2, 0, 5)
    sweep_s = mg.time_kernel(2, 0, 20) * 1e-3
    launch_s = sweep_s / 2
    launch_ev_s = mg.time_kernel(5, 0, 20) * 1e-3
    sweep64_s = mg.time_kernel(0, 0, 10) * 1e-3
    log(f"[bench] fp32+ir (gmres-ir): {np.mean(tgis)*1e3:.1f} ms, {sgi}")

    # What the timed solve leaves out when the operator changes (the
    # transformed step rebuilds it every step): the operator upload with its
    # FP64 Galerkin hierarchy, centre inverses and packed tables, then the
    # first GMRES-IR solve after it, which also tabulates the coarsest level
    # (x = Q r) and captures the V-cycle graphs again.
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    mg.setup_matrices_const(st, None)
    hip.hipDeviceSynchronize()
    setup_s = time.perf_counter() - t0
    first_s, _ = solve(MG.SOLVE_FP32_GMRES_IR)
    log(f"[bench] operator setup {setup_s*1e3:.1f} ms, first GMRES-IR solve after it {first_s*1e3:.1f} ms")

    # error of the timed solution against the FP64 one (periodic system:
    # compare after removing the constant null-space component)
    d32 = x32 - x32.mean(axis=0)
    d64 = x64 - x64.mean(axis=0)
    err = float(np.linalg.norm(d32 - d64) / np.linalg.norm(d64))
    dgi = xgi - xgi.mean(axis=0)
    err_fg = err
    err = float(np.linalg.norm(dgi - d64) / np.linalg.norm(d64))
    dk = xk64 - xk64.mean(axis=0)
    err_k64 = float(np.linalg.norm(dk - d64) / np.linalg.norm(d64))

    # point updates of one V-cycle (3 unknowns per point): the reference
    # schedule sweeps level l (v_down + l) + (v_up + l) times, the coarsest
    # level's 24 bottom sweeps are applied as one tabulated map (counted as
    # its 24 sweeps)
    dof_per_cycle = 0
    L = mg.nlevels
    for l in range(L):
        m_, n_, o_ = mg.dims(l)
        sweeps = 24 if l == L - 1 else (2 + l) + (2 + l)
        dof_per_cycle += 3 * sweeps * m_ * n_ * o_

    alg_launch = 18.0 * N
    achieved = alg_launch / launch_s / 1e9
    achieved64 = 72.0 * N / sweep64_s / 1e9

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle, all cores, one V-cycle)")
        cpu = cpu_baseline(n, st, b, s64["cycles"])

    t32 = float(np.mean(tgis))
    tfg = float(np.mean(t32s))
    t64 = float(np.mean(t64s))
    tir = float(np.mean(tirs))
    tk64 = float(np.mean(tk64s))
    out = {
        "metric": "time-to-1e-10 residual, 256^3 vec3 MG solve: FP32+IR vs FP64; achieved HBM GB/s",
        "value": t32,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"trans-operator projection solve {n}^3 vec3, simple shear T(gamma=0.5), "
                        f"fully periodic, GRF rhs seed 2 l=8, V(2,2)+reference growth, rtol 1e-10",
            "solver": f"GMRES-IR: FP64 refinement, each correction FP32 GMRES({args.restart}) to "
                      f"{args.inner_rtol or 1e-4:g} preconditioned by one FP32 V-cycle",
            "fp64_s": t64,
            "fp64_cycles": s64["cycles"],
            "fp32_ir_stationary_s": tir,
            "fp32_ir_stationary_cycles": sir["cycles"],
            "fp32_cycles": sgi["cycles"],
            "refinements": sgi["outer"],
            "fgmres_s": tfg,
            "fgmres_fp32_cycles": s32["cycles"],
            "fgmres_restarts": s32["outer"],
            "fgmres_rel_err_vs_fp64": err_fg,
            "fp64_gmres_s": tk64,
            "fp64_gmres_cycles": sk64["cycles"],
            "fp64_gmres_rel_err_vs_fp64": err_k64,
            "speedup_vs_fp64": t64 / t32,
            "precision_speedup": tk64 / t32,
            "stationary_speedup": t64 / tir,
            "vcycles_per_s": sgi["cycles"] / t32,
            "fp64_vcycles_per_s": s64["cycles"] / t64,
            "dof_updates_per_s": dof_per_cycle * sgi["cycles"] / t32,
            "fp64_dof_updates_per_s": dof_per_cycle * s64["cycles"] / t64,
            "dof_updates_per_vcycle": dof_per_cycle,
            "rel_err_vs_fp64": err,
            "levels": mg.nlevels,
            "fp64_gs_sweep_GBs": achieved64,
            "fp32_gs_sweep_ms": sweep_s * 1e3,
            "fp64_gs_sweep_ms": sweep64_s * 1e3,
            "operator_setup_ms": setup_s * 1e3,
            "first_gmresir_after_setup_ms": first_s * 1e3,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": TRAFFIC_256_BYTES_PER_LAUNCH if n == 256 else None,
            "kernel": "k_gs_march<float, PAT_TRANS_XZ | centrosymmetric, 64, 16, 512> (fine-level fused GS, "
                      "one z parity)",
            "alg_bytes_per_launch": alg_launch,
            "launch_us": launch_s * 1e6,
            "launch_us_event_pairs": launch_ev_s * 1e6,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
