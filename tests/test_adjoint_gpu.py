"""The adjoint of the QS step (include/stz.h, stz_adjoint) against central
finite differences of the converged forward model: the GPU step's and the
CPU oracle's (every projection solved to ||r|| <= 1e-13 ||r0||).

Loss J = mean over cells of (ad_Dpl/dt)^2 after 3 steps of a 16x16x8 walled
case whose pre-stress yields (plasticity, chi diffusion and the in-place
ad_Dpl ordering all active).  Tolerance of the directional derivatives:
1e-5 relative (central differences with step h: truncation O(h^2), solve
noise ~1e-13 J / h)."""
import numpy as np
import pytest

import oracle as O
from stzpp_amd import shear_sim as SS

pytestmark = pytest.mark.gpu

RTOL = 1e-13
N = (16, 16, 8)
DT, NSTEPS = 0.01, 3


def case():
    p = SS.config0_params(16)
    p["eps0"] = 0.3
    nc, nv = 16 * 16 * 8, 16 * 16 * 9
    rng = np.random.default_rng(7)
    sig = np.zeros((nc, 6))
    k = np.arange(nc) // 256
    sig[:, 1] = np.where((k > 1) & (k < 6), 1.12, 0.5) + 0.02 * rng.standard_normal(nc)
    sig[:, 0] = 0.05 * rng.standard_normal(nc)
    chi = 0.09 + 0.005 * rng.standard_normal(nc)
    u = 1e-3 * rng.standard_normal((nv, 3))
    return p, u, sig, chi


def gpu_loss(p, u, sig, chi):
    sim = SS.ShearSim(*N, params=p)
    sim.set_initial_conditions(chi, u=u, sig=sig, t=0.3)
    sim.step_forward_qs(DT, NSTEPS, solve=SS.SOLVE_RELRES, rtol=RTOL)
    d = sim.state()["dpl"]
    return float(np.mean((d / DT) ** 2))


def oracle_loss(p, u, sig, chi):
    orc = O.OracleShear(N, {k: p[k] for k in O.STEP_PARAM_ORDER})
    orc.set_solver(0, RTOL)
    orc.set_state(0.3, u, sig, chi)
    orc.step(DT, NSTEPS)
    d = orc.state()["dpl"]
    return float(np.mean((d / DT) ** 2))


def adjoint(p, u, sig, chi, k=2):
    sim = SS.ShearSim(*N, params=p)
    sim.set_initial_conditions(chi, u=u, sig=sig, t=0.3)
    return sim.adjoint(DT, NSTEPS, checkpoint_every=k, rtol=RTOL), sim


@pytest.fixture(scope="module")
def grad():
    p, u, sig, chi = case()
    (J, g, gp), sim = adjoint(p, u, sig, chi)
    return p, u, sig, chi, J, g, gp, sim


def test_adjoint_loss_and_state(grad):
    p, u, sig, chi, J, g, gp, sim = grad
    s = sim.state()
    assert abs(J - np.mean((s["dpl"] / DT) ** 2)) <= 1e-14 * J
    assert J > 0 and np.abs(g).max() > 0
    # the forward sweep of the adjoint is the RELRES forward step
    assert abs(J / gpu_loss(p, u, sig, chi) - 1) < 1e-11


@pytest.mark.parametrize("forward", ["gpu", "oracle"])
def test_adjoint_chi0_matches_finite_differences(grad, forward):
    p, u, sig, chi, J, g, gp, _ = grad
    f = gpu_loss if forward == "gpu" else oracle_loss
    rng = np.random.default_rng(11)
    for trial in range(2):
        d = rng.standard_normal(chi.size)
        h = 1e-5
        fd = (f(p, u, sig, chi + h * d) - f(p, u, sig, chi - h * d)) / (2 * h)
        ad = float(g @ d)
        assert abs(ad - fd) <= 1e-5 * abs(fd), (trial, ad, fd)


@pytest.mark.parametrize("which,name", [(0, "sy"), (1, "eps0"), (2, "chi_inf")])
def test_adjoint_params_match_finite_differences(grad, which, name):
    p, u, sig, chi, J, g, gp, _ = grad
    if name == "sy":
        # With eps0 = 0.3 the yielding cells relax onto the yield surface
        # inside the adaptive sub-steps, where the reference switches the
        # plastic rate off (sbar >= s_y, src/shear_sim_3d.cpp:1186-1188): J is
        # not smooth in s_y there (central differences at h = 1e-6 and 1e-5
        # differ by 50 %) and the adjoint gives the derivative almost
        # everywhere.  With eps0 = 0.1 the flow stops before the surface and
        # J is smooth in s_y.
        p = dict(p)
        p["eps0"] = 0.1
        (J, g, gp), _ = adjoint(p, u, sig, chi)
    h = 1e-6 * abs(p[name])
    pp, pm = dict(p), dict(p)
    pp[name] += h
    pm[name] -= h
    fd = (gpu_loss(pp, u, sig, chi) - gpu_loss(pm, u, sig, chi)) / (2 * h)
    assert abs(gp[which] - fd) <= 1e-5 * abs(fd) + 1e-12 * abs(J), (name, gp[which], fd)


def test_adjoint_independent_of_checkpoint_spacing(grad):
    p, u, sig, chi, J, g, gp, _ = grad
    for k in (1, 3):
        (J2, g2, gp2), _ = adjoint(p, u, sig, chi, k=k)
        assert abs(J2 - J) <= 1e-14 * J
        assert np.abs(g2 - g).max() <= 1e-10 * np.abs(g).max()
        assert np.abs(gp2 - gp).max() <= 1e-10 * np.abs(gp).max()


def test_adjoint_rejects_unsupported():
    p, u, sig, chi = case()
    sim = SS.ShearSim(16, 16, 8, params=p, prd=(1, 1, 1))
    sim.set_initial_conditions(chi)
    with pytest.raises(SS.StepError):
        sim.adjoint(DT, 1)
    tsim = SS.TransSim(16, 16, 8, 0.5, params=p)
    tsim.set_initial_conditions(chi)
    with pytest.raises(SS.StepError):
        tsim.adjoint(DT, 1)
