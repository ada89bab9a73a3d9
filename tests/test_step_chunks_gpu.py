"""Parity of the plane-marching step kernels at sizes where they march.

k_advect_march (stzpp_amd/csrc/stz_advect.hpp) cuts z into as many chunks as
the device keeps workgroups of 64 x 4 cell tiles resident; the golden step
cases are small enough that every chunk is one plane.  Here the grids are
wide enough that chunks hold several planes (256 x 192 x 16 walled: 192
tiles, chunks of about 8 planes; 128 x 128 x 32 periodic: 64 tiles, chunks
of about 4), with ragged last chunks, and one grid has a row count that is
not a multiple of the tile (16 x 18 x 4: the last tile's rows 18 and 19 are
outside the grid).  The GPU step (8-colour smoother, the reference's
stopping rule) is compared with the oracle's step with the same smoother and
rule (reference loops: src/shear_sim_3d.cpp:409-525, src/trans_sim_3d.cpp:
880-992): the same iterates, equal V-cycle counts, velocity, stress and chi
to 1e-10.  ad_Dpl is the STZ rate exp(-e_z/chi) C(sbar) (1 - s_y/sbar)/tau0,
whose relative change is about omega eps0 sbar / T_bath = 23 times that of
sbar here, so it is held to 23 x 1e-10 (measured 2.2e-10 on the walled
transformed case).  The fully periodic case has a singular projection: over
its 22 V-cycles of two steps the rounding of the singular iterates leaves
the GPU and the oracle 1.2e-10 apart in U, 4e-11 in sigma and chi and
1.8e-9 in ad_Dpl -- the same, to the digits printed, with the per-point
advection kernels the marching ones replaced -- and its bounds are 3e-10
and 5e-9.  The
states carry a yielding band, advecting velocities and chi variations, so
every term of the advection (ENO branches, Truesdell or L and T terms,
plastic update, chi diffusion with the in-place ad_Dpl order, wall ghosts
and periodic images) takes part.
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


def initial_state(N, prd, seed):
    Nx, Ny, Nz = N
    nc = Nx * Ny * Nz
    nv = Nx * Ny * (Nz + (0 if prd[2] else 1))
    rng = np.random.default_rng(seed)
    k = np.arange(nc) // (Nx * Ny)
    sig = np.zeros((nc, 6))
    # a yielding band in the middle planes, elastic elsewhere
    sig[:, 1] = np.where((k >= Nz // 4) & (k < 3 * Nz // 4), 1.1, 0.5) + 0.03 * rng.standard_normal(nc)
    sig[:, 0] = 0.05 * rng.standard_normal(nc)
    sig[:, 4] = 0.05 * rng.standard_normal(nc)
    chi = 0.09 + 0.005 * rng.standard_normal(nc)
    u = 2e-3 * rng.standard_normal((nv, 3))
    return u, sig, chi


def params(N):
    p = SS.config0_params(N[0])
    p["eps0"] = 0.3
    return p


def run_pair(N, prd, family=None, Ut=0.0, nsteps=2, dt=0.01, seed=5):
    p = params(N)
    u, sig, chi = initial_state(N, prd, seed)
    if family:
        sim = SS.TransSim(*N, Ut, family=family, params=p, prd=prd)
    else:
        sim = SS.ShearSim(*N, params=p, prd=prd)
    sim.set_initial_conditions(chi, u=u, sig=sig, t=0.3)
    box = (p["ax"], p["bx"], p["ay"], p["by"], p["az"], p["bz"])
    orc = O.OracleShear(N, p, box=box, prd=prd)
    if family:
        orc.set_transform(family, Ut)
    orc.set_solver(1, 0.0)
    orc.set_state(0.3, u, sig, chi)
    ca, cb = [], []
    for _ in range(nsteps):
        ca.append(sim.step_forward_qs(dt, 1)["vcycles"])
        orc.step(dt, 1)
        cb.append(orc.vcycles)
    a, b = sim.state(), orc.state()
    orc.close()
    return a, b, ca, cb


@pytest.mark.parametrize("N,prd,family", [
    ((256, 192, 16), (1, 1, 0), None),
    ((256, 192, 16), (1, 1, 0), "pure_shear"),
    ((128, 128, 32), (1, 1, 1), "pure_shear"),
    ((16, 18, 4), (1, 1, 0), None),
    ((16, 18, 4), (1, 1, 0), "simple_shear"),
])
def test_gpu_marching_step_matches_oracle(N, prd, family):
    print(f"step parity N={N} prd={prd} family={family}", flush=True)
    a, b, ca, cb = run_pair(N, prd, family=family, Ut=0.5 if family else 0.0)
    assert ca == cb
    assert np.abs(b["dpl"]).max() > 1e-4  # the band yields
    errs = {key: relerr(a[key], b[key]) for key in KEYS}
    if prd[2]:
        # fully periodic: the projection is singular (constant velocities);
        # the constant each solve leaves in U follows the iterates to the
        # rounding of the singular system, so U is compared without it
        errs["u"] = relerr(a["u"] - a["u"].mean(axis=0), b["u"] - b["u"].mean(axis=0))
        assert relerr(a["u"], b["u"]) < 1e-9
    print("relative errors", errs, flush=True)
    for key in KEYS:
        if key == "dpl":
            tol = 5e-9 if prd[2] else 2.3e-9
        else:
            tol = 3e-10 if prd[2] else 1e-10
        assert errs[key] < tol, (key, errs[key])
