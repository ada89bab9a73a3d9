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
