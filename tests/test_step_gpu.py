"""Parity of the GPU QS step (C ABI include/stz.h) with the CPU oracle and
the reference's golden fields.

The reference stops each projection solve when sum |r|^4 <= (eps a11 1e3)^2
(src/shear_sim_3d.cpp:551-582 with mg3d_mod_sq(vec3) = |r|^4): a per-point
residual near (eps a11 1e3)^(1/2), which leaves the reference's own fields
about 1e-5 (relative) from the converged projection in these cases.  Hence:
  * GPU vs oracle with the same 8-colour smoother and the same stopping rule:
    the same iterates, agreement to 1e-10;
  * GPU vs oracle, both solved to ||r|| <= 1e-13 ||r0||: the converged
    projection, agreement to 1e-10 (lexicographic oracle, 8-colour GPU);
  * GPU vs the reference's fields: within the reference's solve error,
    stated as 1e-4 relative (measured 9e-6 for u and 1.9e-5 for sigma).
"""
import numpy as np
import pytest

import oracle as O
from stzpp_amd import shear_sim as SS

pytestmark = pytest.mark.gpu

KEYS = ("u", "sig", "chi", "dpl")


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


def oracle_run(g, smoother, rtol):
    N = tuple(int(v) for v in g["N"])
    orc = O.OracleShear(N, params_of(g))
    orc.set_solver(smoother, rtol)
    nv = N[0] * N[1] * (N[2] + 1)
    orc.set_state(0.0, np.zeros((nv, 3)), g["sig0"], g["chi0"])
    cyc = []
    for _ in range(int(g["nsteps"])):
        orc.step(float(g["dt"]), 1)
        cyc.append(orc.vcycles)
    return orc.state(), cyc


@pytest.mark.parametrize("case", ["step_plastic.npz", "step_c0.npz"])
def test_gpu_step_matches_oracle_same_smoother(golden, case):
    g = golden(case)
    a, ca = gpu_run(g)
    b, cb = oracle_run(g, smoother=1, rtol=0.0)
    assert ca == cb
    assert abs(a["t"] - b["t"]) < 1e-15
    for key in KEYS:
        assert relerr(a[key], b[key]) < 1e-10, key


@pytest.mark.parametrize("case", ["step_plastic.npz", "step_c0.npz"])
@pytest.mark.parametrize("solve", [SS.SOLVE_RELRES, SS.SOLVE_FP32_IR, SS.SOLVE_FP32_GMRES_IR])
def test_gpu_step_converged_matches_oracle(golden, case, solve):
    g = golden(case)
    a, _ = gpu_run(g, solve=solve, rtol=1e-13)
    b, _ = oracle_run(g, smoother=0, rtol=1e-13)
    for key in KEYS:
        assert relerr(a[key], b[key]) < 1e-10, key


@pytest.mark.parametrize("case", ["step_plastic.npz", "step_c0.npz"])
def test_gpu_step_within_reference_solve_error(golden, case):
    g = golden(case)
    a, _ = gpu_run(g)
    if case == "step_plastic.npz":
        assert np.abs(g["dpl"]).max() > 1e-3
        for key in KEYS:
            assert relerr(a[key], g[key]) < 1e-4, key
    else:
        Nx, Ny, Nz = (int(v) for v in g["N"])
        for key, nz in (("u", Nz + 1), ("sig", Nz), ("chi", Nz)):
            p = a[key].reshape(nz, Nx * Ny, -1)
            assert relerr(p[8], g[f"{key}_k8"]) < 1e-4, key
            assert relerr(p[20], g[f"{key}_k20"]) < 1e-4, key
            assert relerr((p * p).sum(axis=1), g[f"{key}_plane_sq"]) < 1e-4, key


def test_gpu_step_matches_oracle_with_plasticity_and_motion():
    # Nonzero initial velocity, a yielding band, plates on the ramp at t=0.3.
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
    orc.set_solver(1, 0.0)
    orc.set_state(0.3, u, sig, chi)
    sim = SS.ShearSim(*N, params=p)
    sim.set_initial_conditions(chi, u=u, sig=sig, t=0.3)
    for _ in range(3):
        orc.step(0.01, 1)
        st = sim.step_forward_qs(0.01, 1)
    a, b = sim.state(), orc.state()
    assert st["yielded_cells"] > 0
    assert np.abs(b["dpl"]).max() > 1e-4
    for key in KEYS:
        assert relerr(a[key], b[key]) < 1e-10, key


# ---- transformed-coordinate step (trans_sim_3d) ---------------------------

def gpu_trans_run(g, solve=None, rtol=1e-12, family="pure_shear"):
    N = tuple(int(v) for v in g["N"])
    prd = tuple(int(v) for v in g["prd"])
    sim = SS.TransSim(*N, float(g["Ut"]), family=family, params=params_of(g), prd=prd)
    sim.set_initial_conditions(g["chi0"], u=g["u0"], sig=g["sig0"])
    cyc = []
    for _ in range(int(g["nsteps"])):
        st = sim.step_forward_qs(float(g["dt"]), 1, solve=solve, rtol=rtol)
        cyc.append(st["vcycles"])
    return sim.state(), cyc


def oracle_trans_run(g, smoother, rtol, family="pure_shear"):
    N = tuple(int(v) for v in g["N"])
    prd = tuple(int(v) for v in g["prd"])
    orc = O.OracleShear(N, params_of(g), prd=prd)
    orc.set_transform(family, float(g["Ut"]))
    orc.set_solver(smoother, rtol)
    orc.set_state(0.0, g["u0"], g["sig0"], g["chi0"])
    cyc = []
    for _ in range(int(g["nsteps"])):
        orc.step(float(g["dt"]), 1)
        cyc.append(orc.vcycles)
    return orc.state(), cyc


TRANS_CASES = ["trans_step_walled.npz", "trans_step_periodic.npz"]


@pytest.mark.parametrize("case", TRANS_CASES)
def test_gpu_trans_step_matches_oracle_same_smoother(golden, case):
    # same 8-colour smoother and the reference's stopping rule (norm on the
    # std_iters schedule): same V-cycle counts, fields to 1e-10
    g = golden(case)
    a, ca = gpu_trans_run(g)
    b, cb = oracle_trans_run(g, smoother=1, rtol=0.0)
    assert ca == cb
    assert abs(a["t"] - b["t"]) < 1e-15
    assert np.abs(b["dpl"]).max() > 1e-3
    for key in KEYS:
        assert relerr(a[key], b[key]) < 1e-10, key


@pytest.mark.parametrize("case,solve,smoother", [
    ("trans_step_walled.npz", SS.SOLVE_RELRES, 0),
    ("trans_step_walled.npz", SS.SOLVE_FP32_IR, 0),
    ("trans_step_walled.npz", SS.SOLVE_FP32_GMRES_IR, 0),
    # fully periodic: the projection operator is singular (constant vectors),
    # and the constant the solve leaves in U depends on the iterates and is
    # advected in the next step; the oracle takes the GPU's 8-colour iterates
    ("trans_step_periodic.npz", SS.SOLVE_RELRES, 1),
])
def test_gpu_trans_step_converged_matches_oracle(golden, case, solve, smoother):
    g = golden(case)
    a, _ = gpu_trans_run(g, solve=solve, rtol=1e-13)
    b, _ = oracle_trans_run(g, smoother=smoother, rtol=1e-13)
    for key in KEYS:
        assert relerr(a[key], b[key]) < 1e-10, key


def test_gpu_trans_step_within_reference_solve_error(golden):
    # Walled z only.  With z periodic the projection is singular and the
    # reference's loose stopping rule leaves U dependent on the smoother's
    # iterates: the lexicographic and the 8-colour oracle differ by 3e-3 in
    # U (3e-4 in sigma) after 3 steps while the lexicographic oracle matches
    # the reference to 2e-14.  There parity is the chain GPU == 8-colour
    # oracle (test above, 1e-10) and lexicographic oracle == reference
    # (tests/test_oracle_step.py, 1e-11).
    g = golden("trans_step_walled.npz")
    a, _ = gpu_trans_run(g)
    for key in KEYS:
        assert relerr(a[key], g[key]) < 1e-4, key


def test_gpu_trans_step_graph_and_launches_agree(golden):
    # the reference loop launched kernel by kernel (explicit options, mode
    # REF with tol 0) and replayed from the captured graph: the stopping
    # tolerance follows each step's operator in both, from the first step of
    # a fresh simulation on, so the V-cycle counts and the fields agree
    g = golden("trans_step_walled.npz")
    N = tuple(int(v) for v in g["N"])
    prd = tuple(int(v) for v in g["prd"])
    out = []
    for use_graph in (True, False):
        sim = SS.TransSim(*N, float(g["Ut"]), family="pure_shear", params=params_of(g), prd=prd)
        sim.set_initial_conditions(g["chi0"], u=g["u0"], sig=g["sig0"])
        cyc = [sim.step_forward_qs(float(g["dt"]), 1, use_graph=use_graph)["vcycles"]
               for _ in range(int(g["nsteps"]))]
        out.append((sim.state(), cyc))
    (a, ca), (b, cb) = out
    assert ca == cb
    for key in KEYS:
        assert relerr(a[key], b[key]) < 1e-12, key


@pytest.mark.parametrize("family", ["simple_shear", "homogeneous", "identity"])
def test_gpu_trans_families_match_oracle(golden, family):
    # the other transform families: the simple shear couples x and z in the
    # operator, the metric and the L term
    g = dict(golden("trans_step_walled.npz"))
    a, ca = gpu_trans_run(g, family=family)
    b, cb = oracle_trans_run(g, smoother=1, rtol=0.0, family=family)
    assert ca == cb
    for key in KEYS:
        assert relerr(a[key], b[key]) < 1e-10, key


def test_gpu_transform_matches_oracle():
    for fam in ("pure_shear", "simple_shear", "homogeneous", "identity"):
        for t in (0.0, 0.37):
            A = SS.transform(fam, 0.8, t)
            B = O.transform_at(fam, 0.8, 0.5, t)
            for x, y in zip(A, B):
                assert np.array_equal(x, y), fam


def test_gpu_user_transform_reproduces_simple_shear(golden):
    # stz_set_transform with the simple shear's T(t0) and dT/dt: the family
    # is affine in t, so the user transform gives the same steps
    g = golden("trans_step_walled.npz")
    N = tuple(int(v) for v in g["N"])
    prd = tuple(int(v) for v in g["prd"])
    Ut, dt = float(g["Ut"]), float(g["dt"])
    out = []
    for user in (False, True):
        sim = SS.TransSim(*N, Ut, family="identity" if user else "simple_shear", params=params_of(g), prd=prd)
        sim.set_initial_conditions(g["chi0"], u=g["u0"], sig=g["sig0"])
        if user:
            T, _, dT = SS.transform("simple_shear", Ut, 0.0)
            sim.set_transform(T, dT)
        for _ in range(int(g["nsteps"])):
            sim.step_forward_qs(dt, 1, solve=SS.SOLVE_RELRES, rtol=1e-13)
        out.append(sim.state())
    a, b = out
    for key in KEYS:
        assert relerr(a[key], b[key]) < 1e-10, key


def test_gpu_user_transform_operator_and_errors(golden):
    # a general (non-family) constant-rate transform: the step's projection
    # operator is the transformed stencil at T(t + dt); an untransformed
    # handle and a singular T are refused
    import ctypes

    from stzpp_amd import _lib
    from stzpp_amd import multigrid as MG

    g = golden("trans_step_walled.npz")
    N = tuple(int(v) for v in g["N"])
    prd = tuple(int(v) for v in g["prd"])
    p = params_of(g)
    dt = float(g["dt"])
    sim = SS.TransSim(*N, 0.1, family="identity", params=p, prd=prd)
    sim.set_initial_conditions(g["chi0"], u=g["u0"], sig=g["sig0"])
    T0 = np.array([[1.05, 0.1, 0.2], [0.0, 0.97, 0.05], [0.0, 0.0, 1.02]])
    dT = np.array([[0.1, 0.0, 0.3], [0.0, -0.05, 0.0], [0.0, 0.0, 0.02]])
    sim.set_transform(T0, dT)
    st = sim.step_forward_qs(dt, 1, solve=SS.SOLVE_RELRES, rtol=1e-12)
    assert st["steps"] == 1
    s = sim.state()
    assert all(np.isfinite(s[k]).all() for k in KEYS)
    L = _lib.load()
    mg = L.stz_multigrid(sim.h)
    n = ctypes.c_long()
    assert L.stzmg_get_level_operator(mg, 0, None, ctypes.byref(n)) == 0
    ops = np.zeros((n.value, 27, 3, 3))
    assert L.stzmg_get_level_operator(mg, 0, _lib.dptr(ops), None) == 0
    q = sim.params
    h = [(q["bx"] - q["ax"]) / N[0], (q["by"] - q["ay"]) / N[1], (q["bz"] - q["az"]) / N[2]]
    ref = MG.trans_stencil(T0 + dt * dT, q["mu"], q["lamb"], *h, dt)
    assert relerr(ops[n.value // 2], ref) < 1e-14
    # refusals
    plain = SS.ShearSim(*N, params=p)
    Ta = np.eye(3).ravel().copy()
    assert L.stz_set_transform(plain.h, _lib.dptr(Ta), _lib.dptr(np.zeros(9))) == -5
    sing = np.zeros(9)
    assert L.stz_set_transform(sim.h, _lib.dptr(sing), _lib.dptr(np.zeros(9))) == -8
