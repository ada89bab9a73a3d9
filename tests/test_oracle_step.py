"""Pins the CPU oracle of the QS step against the compiled reference's
golden fields (tests/golden/step_*.npz) and the operator builders against
the reference's own stencil."""
import numpy as np
import pytest

import oracle as O


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def params_of(g):
    return {k[2:]: float(v) for k, v in g.items() if k.startswith("p_")}


def run_oracle(g):
    N = tuple(int(v) for v in g["N"])
    p = params_of(g)
    orc = O.OracleShear(N, p)
    nv = N[0] * N[1] * (N[2] + 1)
    orc.set_state(0.0, np.zeros((nv, 3)), g["sig0"], g["chi0"])
    cyc = []
    for _ in range(int(g["nsteps"])):
        orc.step(float(g["dt"]), 1)
        cyc.append(orc.vcycles)
    return orc.state(), cyc


def test_oracle_step_plastic_matches_reference(golden):
    g = golden("step_plastic.npz")
    s, cyc = run_oracle(g)
    assert cyc == list(g["vcycles"])
    assert abs(s["t"] - float(g["t"])) < 1e-15
    assert np.abs(g["dpl"]).max() > 1e-3  # plasticity is active in this case
    assert relerr(s["u"], g["u"]) < 1e-12
    assert relerr(s["sig"], g["sig"]) < 1e-12
    assert relerr(s["chi"], g["chi"]) < 1e-13
    assert relerr(s["dpl"], g["dpl"]) < 1e-12


def test_oracle_step_config0_matches_reference(golden):
    g = golden("step_c0.npz")
    s, cyc = run_oracle(g)
    assert cyc == list(g["vcycles"])
    N = tuple(int(v) for v in g["N"])
    Nx, Ny, Nz = N
    for key, nz in (("u", Nz + 1), ("sig", Nz), ("chi", Nz)):
        a = s[key].reshape(nz, Nx * Ny, -1)
        assert relerr(a.sum(axis=1), g[f"{key}_plane_sum"]) < 1e-10, key
        assert relerr((a * a).sum(axis=1), g[f"{key}_plane_sq"]) < 1e-10, key
        assert relerr(a[8], g[f"{key}_k8"]) < 1e-11, key
        assert relerr(a[20], g[f"{key}_k20"]) < 1e-11, key


@pytest.mark.parametrize("case", ["step_plastic.npz", "step_c0.npz"])
def test_reference_stencil_matches_builder(golden, case):
    g = golden(case)
    N = tuple(int(v) for v in g["N"])
    p = params_of(g)
    dx, dy, dz = 2.0 / N[0], 2.0 / N[1], 1.0 / N[2]
    st, fix = O.shear_operator(p["mu"], p["lamb"], dx, dy, dz, float(g["dt"]))
    sym = np.array([[[b[0], b[1], b[2]], [b[1], b[3], b[4]], [b[2], b[4], b[5]]] for b in g["stencil_sym"]])
    assert relerr(st, sym) < 1e-14
    f = g["fix_sym"]
    assert relerr(fix, np.array([[f[0], f[1], f[2]], [f[1], f[3], f[4]], [f[2], f[4], f[5]]])) < 1e-15


def test_step_invariants_uniform_state():
    # A uniform stress/chi state at rest with plates at rest stays uniform:
    # advection, ENO and diffusion of a constant vanish, divergence is zero.
    N = (8, 8, 8)
    p = dict(mu=1.0, lamb=1.5, sy=1.0, tau0=1.0, c0=0.4, eps0=0.3, delta=0.38, omega=1.35,
             bath_temp=0.019, chi_inf=0.15, ez=1.0, diff_l=0.1, U=0.0)
    orc = O.OracleShear(N, p)
    nc, nv = 512, 8 * 8 * 9
    sig = np.tile([0.3, 0.1, 0.0, -0.2, 0.05, -0.1], (nc, 1))
    chi = np.full(nc, 0.1)
    orc.set_state(0.0, np.zeros((nv, 3)), sig, chi)
    orc.step(0.01, 2)
    s = orc.state()
    assert np.abs(s["u"]).max() < 1e-12
    assert relerr(s["sig"], sig) < 1e-12
    assert relerr(s["chi"], chi) < 1e-14


# ---- transformed-coordinate step (trans_sim_3d) ---------------------------

def run_oracle_trans(g, smoother=0, rtol=0.0):
    N = tuple(int(v) for v in g["N"])
    prd = tuple(int(v) for v in g["prd"])
    orc = O.OracleShear(N, params_of(g), prd=prd)
    orc.set_transform("pure_shear", float(g["Ut"]))
    orc.set_solver(smoother, rtol)
    orc.set_state(0.0, g["u0"], g["sig0"], g["chi0"])
    cyc = []
    for _ in range(int(g["nsteps"])):
        orc.step(float(g["dt"]), 1)
        cyc.append(orc.vcycles)
    return orc.state(), cyc


@pytest.mark.parametrize("case", ["trans_step_walled.npz", "trans_step_periodic.npz"])
def test_oracle_trans_step_matches_reference(golden, case):
    # the oracle's trans step (L and T terms, physical-stress plasticity,
    # metric chi diffusion with its quirks, per-step operator, T div sigma'
    # right-hand side, trans norm schedule) against trans_sim_3d
    g = golden(case)
    s, cyc = run_oracle_trans(g)
    assert cyc == list(g["vcycles"])
    assert abs(s["t"] - float(g["t"])) < 1e-15
    assert np.abs(g["dpl"]).max() > 1e-3  # plasticity is active
    assert relerr(s["u"], g["u"]) < 1e-11
    assert relerr(s["sig"], g["sig"]) < 1e-12
    assert relerr(s["chi"], g["chi"]) < 1e-12
    assert relerr(s["dpl"], g["dpl"]) < 1e-11


def test_oracle_trans_fixture_pins_the_chi_eno_quirk(golden):
    # planted in the oracle: the textbook x-ENO argument me[2] instead of
    # the reference's me[1] (src/trans_sim_3d.cpp:985) must move chi away
    # from the reference fixture, so the fixture pins that form
    g = golden("trans_step_periodic.npz")
    N = tuple(int(v) for v in g["N"])
    orc = O.OracleShear(N, params_of(g), prd=(1, 1, 1))
    orc.set_transform("pure_shear", float(g["Ut"]))
    orc.L.sto_set_textbook(orc.h, 1)
    orc.set_state(0.0, g["u0"], g["sig0"], g["chi0"])
    orc.step(float(g["dt"]), int(g["nsteps"]))
    assert relerr(orc.state()["chi"], g["chi"]) > 1e-9


def test_oracle_transform_matches_reference(golden):
    g = golden("trans_step_walled.npz")
    T, Ti, dT = O.transform_at("pure_shear", float(g["Ut"]), 0.5, float(g["t"]))
    assert relerr(T, g["T"]) < 1e-15
    assert relerr(Ti, g["Tinv"]) < 1e-15
    assert relerr(dT, g["dTdt"]) < 1e-15


@pytest.mark.parametrize("family", ["pure_shear", "simple_shear", "homogeneous", "identity"])
def test_transform_families_closed_forms(family):
    # the families of src/trans_sim_3d.hh:57-125: T T^-1 = I and dT/dt equals
    # the derivative of T (central difference); identity and the shape of
    # each family fixed by its definition
    Ut, bz, t, h = 0.7, 0.5, 0.3, 1e-6
    T, Ti, dT = O.transform_at(family, Ut, bz, t)
    assert np.abs(T @ Ti - np.eye(3)).max() < 1e-14
    Tp, _, _ = O.transform_at(family, Ut, bz, t + h)
    Tm, _, _ = O.transform_at(family, Ut, bz, t - h)
    assert np.abs((Tp - Tm) / (2 * h) - dT).max() < 1e-8
    if family == "identity":
        assert np.array_equal(T, np.eye(3)) and not dT.any()
    if family == "simple_shear":
        assert T[0, 2] == Ut * t / bz and np.count_nonzero(T - np.eye(3)) == 1
    if family == "homogeneous":
        assert np.allclose(T, np.exp(Ut * t) * np.eye(3), rtol=1e-15)
    if family == "pure_shear":
        assert abs(np.linalg.det(T) - 1) < 1e-14  # volume preserving


@pytest.mark.parametrize("case", ["trans_step_walled.npz", "trans_step_periodic.npz"])
def test_c_trans_stencil_matches_reference(golden, case):
    # the C oracle's closed form of stenc_funcs at the final T, scaled by
    # -dt, against the reference's stencil; fix = -dt 2((l+2m)/dx^2+m/dy^2+m/dz^2) I
    g = golden(case)
    N = tuple(int(v) for v in g["N"])
    p = params_of(g)
    dx, dy, dz = 2.0 / N[0], 2.0 / N[1], 1.0 / N[2]
    dt = float(g["dt"])
    st = -dt * O.trans_stencil_c(g["T"], p["mu"], p["lamb"], dx, dy, dz)
    assert relerr(st, g["stencil"]) < 1e-13
    fx = -dt * 2 * ((p["lamb"] + 2 * p["mu"]) / dx ** 2 + p["mu"] / dy ** 2 + p["mu"] / dz ** 2)
    assert relerr(g["fix"], fx * np.eye(3)) < 1e-15


def test_c_trans_stencil_matches_generated_functions(golden):
    g = golden("trans_stencil.npz")
    for q, T in enumerate(g["T"]):
        st = O.trans_stencil_c(T, float(g["mu"]), float(g["lamb"]), *(float(v) for v in g["h"]))
        assert relerr(st, g["blocks"][q].reshape(27, 3, 3)) < 1e-13, q
