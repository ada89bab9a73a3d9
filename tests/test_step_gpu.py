"""Parity of the GPU QS step (C ABI include/stz.h) with the reference's
golden fields and with the CPU oracle."""
import numpy as np
import pytest

import oracle as O
from stzpp_amd import shear_sim as SS

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def params_of(g):
    return {k[2:]: float(v) for k, v in g.items() if k.startswith("p_")}


def gpu_run(g, solve=None, rtol=1e-12):
    N = tuple(int(v) for v in g["N"])
    sim = SS.ShearSim(*N, params=params_of(g))
    sim.set_initial_conditions(g["chi0"], sig=g["sig0"])
    cyc = []
    for _ in range(int(g["nsteps"])):
        st = sim.step_forward_qs(float(g["dt"]), 1, solve=solve, rtol=rtol)
        cyc.append(st["vcycles"])
    return sim.state(), cyc


def test_gpu_step_plastic_matches_reference(golden):
    g = golden("step_plastic.npz")
    s, cyc = gpu_run(g)
    assert abs(s["t"] - float(g["t"])) < 1e-15
    # Projection solved to the reference's own tolerance with 8-colour
    # smoothing: fields agree to the solve tolerance (1e-10 relative).
    assert relerr(s["u"], g["u"]) < 1e-10
    assert relerr(s["sig"], g["sig"]) < 1e-10
    assert relerr(s["chi"], g["chi"]) < 1e-12
    assert relerr(s["dpl"], g["dpl"]) < 1e-10
    assert all(c >= 3 for c in cyc)


def test_gpu_step_config0_matches_reference(golden):
    g = golden("step_c0.npz")
    s, cyc = gpu_run(g)
    Nx, Ny, Nz = (int(v) for v in g["N"])
    for key, nz in (("u", Nz + 1), ("sig", Nz), ("chi", Nz)):
        a = s[key].reshape(nz, Nx * Ny, -1)
        assert relerr(a[8], g[f"{key}_k8"]) < 1e-9, key
        assert relerr(a[20], g[f"{key}_k20"]) < 1e-9, key
        assert relerr((a * a).sum(axis=1), g[f"{key}_plane_sq"]) < 1e-9, key


def test_gpu_step_fp32_ir_projection_within_tolerance(golden):
    # FP32 smoothing inside FP64 refinement, each projection to 1e-12
    # relative: fields within 1e-9 of the reference after 10 steps.
    g = golden("step_c0.npz")
    s, _ = gpu_run(g, solve=SS.SOLVE_FP32_IR, rtol=1e-12)
    Nx, Ny, Nz = (int(v) for v in g["N"])
    for key, nz in (("u", Nz + 1), ("sig", Nz), ("chi", Nz)):
        a = s[key].reshape(nz, Nx * Ny, -1)
        assert relerr(a[8], g[f"{key}_k8"]) < 1e-9, key


def test_gpu_step_matches_oracle_with_plasticity_and_motion():
    # Nonzero initial velocity, yielding band, ramped plates.
    N = (16, 16, 8)
    p = SS.config0_params(16)
    p["eps0"] = 0.3
    nc, nv = 16 * 16 * 8, 16 * 16 * 9
    rng = np.random.default_rng(4)
    u = 1e-3 * rng.standard_normal((nv, 3))
    sig = np.zeros((nc, 6))
    k = np.arange(nc) // 256
    sig[:, 1] = np.where((k > 2) & (k < 6), 1.15, 0.4)
    sig[:, 0] = 0.05 * rng.standard_normal(nc)
    chi = 0.09 + 0.005 * rng.standard_normal(nc)
    orc = O.OracleShear(N, {k_: p[k_] for k_ in O.STEP_PARAM_ORDER})
    orc.set_state(0.3, u, sig, chi)
    sim = SS.ShearSim(*N, params=p)
    sim.set_initial_conditions(chi, u=u, sig=sig, t=0.3)
    for _ in range(3):
        orc.step(0.01, 1)
        sim.step_forward_qs(0.01, 1)
    a, b = sim.state(), orc.state()
    assert np.abs(b["dpl"]).max() > 1e-4
    for key in ("u", "sig", "chi", "dpl"):
        assert relerr(a[key], b[key]) < 1e-9, key
