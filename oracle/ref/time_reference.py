"""Times the compiled reference beside the oracle on this machine's cores
(test infrastructure: the CPU baseline's calibration).

Runs, with OMP_NUM_THREADS=1 and MPICH's mpiexec from /opt/conda:
* mg_parallel (the reference's own multigrid benchmark,
  mg_3d_template/mg_parallel.cc: 129^3 walled manufactured vec3 problem,
  5 V-cycles x 4 tests) on 1 and 8 ranks;
* ref_time (oracle/ref/ref_time.cc) on 1 and 8 ranks: the reference's
  V-cycle for the shear_sim_3d operator at 128^3 cells (multigrid
  128 x 128 x 129) and for the trans_sim_3d operator at 256^3 (multigrid
  256^3, periodic);
and the oracle's V-cycle (lexicographic sweep, reference schedule) on the
same operators, on 1 thread as one rank and on 8 threads as the reference's
8-rank subdomain grid.  The ratio reference / oracle at equal ranks is the
calibration that turns an oracle time measured elsewhere into a
reference-equivalent time.  Also records the oracle's V-cycle count to
relative residual 1e-10 on the 256^3 simple-shear solve (GRF right-hand side,
seed 2, l = 8) for the subdomain grids of 1, 8 and 16 ranks.

Writes oracle/ref/reference_timings.json.  Needs oracle/_ref built
(make -C oracle/ref) and about 15 minutes and 20 GB of memory.

    python oracle/ref/time_reference.py [--quick]
"""
import argparse
import json
import os
import platform
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

REFBIN = os.path.join(ROOT, "oracle", "_ref")
MPIEXEC = "/opt/conda/bin/mpiexec"
T_SHEAR = np.array([[1.0, 0.0, 0.5], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def mpirun(P, args, cwd):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    p = subprocess.run([MPIEXEC, "-n", str(P)] + args, cwd=cwd, env=env, capture_output=True, text=True,
                       check=True, timeout=3600)
    return p.stdout


def mg_parallel(P, cwd):
    out = mpirun(P, [os.path.join(REFBIN, "mg_parallel")], cwd)
    tests = [float(v) for v in re.findall(r"Test \d+ : ([0-9.eE+-]+) s", out)]
    return {"ranks": P, "tests_s": tests, "vcycle_s": float(np.mean(tests)) / 5}


def ref_time(kind, N, P, reps, p, cwd):
    a = [1.0, p["mu"], p["lamb"], p["sy"], p["tau0"], p["c0"], p["eps0"], p["delta"], p["omega"],
         p["bath_temp"], p["chi_inf"], p["ez"], 21000.0, p["diff_l"], p["U"]]
    out = mpirun(P, [os.path.join(REFBIN, "ref_time"), kind, str(N), str(reps)] + [repr(float(v)) for v in a], cwd)
    line = [ln for ln in out.splitlines() if ln.startswith("{")][-1]
    return json.loads(line)


def oracle_problem(kind, N):
    h = 1.0 / N
    if kind == "shear":
        st, fix = O.shear_operator(1.0, 1.5, h, h, h, 1.0)
        dims, prd = (N, N, N + 1), (1, 1, 0)
        ops = O.plane_operator(N + 1, 0, st, fix)
    else:
        st = -O.trans_stencil_blocks(T_SHEAR, 1.0, 1.5, h, h, h)
        dims, prd = (N, N, N), (1, 1, 1)
        ops = O.plane_operator(N, 1, st, None)
    return dims, prd, ops


def oracle_time(kind, N, threads, ranks, reps):
    dims, prd, ops = oracle_problem(kind, N)
    O.set_threads(threads)
    orc = O.OracleMG(*dims, prd, mode=O.PLANE)
    orc.set_ops(0, ops)
    orc.rat()
    orc.set_ranks(ranks)
    Np = dims[0] * dims[1] * dims[2]
    orc.set_fields(0, np.zeros((Np, 3)), O.det_field(Np, 11))
    orc.vcycle(1)
    t0 = time.perf_counter()
    orc.vcycle(reps)
    dt = (time.perf_counter() - t0) / reps
    orc.close()
    return {"threads": threads, "ranks": list(ranks), "vcycle_s": dt}


def oracle_cycles(N, threads, ranks, rtol=1e-10, cap=200):
    from stzpp_amd import fields

    dims, prd, ops = oracle_problem("trans", N)
    O.set_threads(threads)
    orc = O.OracleMG(*dims, prd, mode=O.PLANE)
    orc.set_ops(0, ops)
    orc.rat()
    orc.set_ranks(ranks)
    b = fields.grf_vec3(dims, 8.0, 2)
    b -= b.mean(axis=0)
    orc.set_fields(0, np.zeros_like(b), b)
    r0 = np.sqrt(orc.norm2(0))
    for c in range(1, cap + 1):
        orc.vcycle(1)
        if np.sqrt(orc.norm2(0)) <= rtol * r0:
            orc.close()
            return c
    orc.close()
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="small sizes, to check the recipe")
    ap.add_argument("--keep-cycles", action="store_true",
                    help="reuse the cycle counts of the existing file (they do not depend on the machine)")
    args = ap.parse_args()
    from stzpp_amd import shear_sim as SS

    n_shear, n_trans = (32, 32) if args.quick else (128, 256)
    out = {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "mpi": "MPICH (/opt/conda)",
           "flags": "g++ -O3 -std=c++11 -fopenmp (the reference's src/Makefile), OMP_NUM_THREADS=1"}
    with tempfile.TemporaryDirectory() as wd:
        out["mg_parallel_129"] = [mg_parallel(P, wd) for P in (1, 8)] if not args.quick else []
        log(f"mg_parallel: {out['mg_parallel_129']}")
        for kind, N, reps in (("shear", n_shear, 3), ("trans", n_trans, 2)):
            p = SS.config0_params(N)
            ref = [ref_time(kind, N, P, reps, p, wd) for P in (1, 8)]
            log(f"reference {kind} {N}: {ref}")
            orc = [oracle_time(kind, N, 1, (1, 1, 1), reps),
                   oracle_time(kind, N, 8, ref[1]["dims"], reps)]
            log(f"oracle {kind} {N}: {orc}")
            out[f"{kind}_{N}"] = {
                "reference": ref, "oracle": orc,
                "calibration_1": ref[0]["vcycle_s"] / orc[0]["vcycle_s"],
                "calibration_8": ref[1]["vcycle_s"] / orc[1]["vcycle_s"],
            }
    path = os.path.join(HERE, "reference_timings_quick.json" if args.quick else "reference_timings.json")
    key = f"trans_{n_trans}_cycles_to_1e-10"
    if args.keep_cycles and os.path.exists(path):
        with open(path) as f:
            out[key] = json.load(f)[key]
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        log(f"wrote {path}")
        return
    cyc = {}
    for P in (1, 8, 16):
        ranks = O.processor_division(P, n_trans, n_trans, n_trans, (1, 1, 1))
        cyc[str(P)] = {"ranks": list(ranks), "cycles": oracle_cycles(n_trans, 8, ranks)}
        log(f"oracle cycles to 1e-10, {P} ranks {ranks}: {cyc[str(P)]['cycles']}")
    out[key] = cyc
    O.set_threads(os.cpu_count())
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    log(f"wrote {path}")


if __name__ == "__main__":
    main()
