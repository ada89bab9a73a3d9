"""The direct (explicit dynamics) step of shear_sim_3d restated in the
oracle (step_forward, src/shear_sim_3d.cpp:206-213) against the reference's
own runs (tests/golden/direct_*.npz): bit for bit after 1 and 5 steps,
velocity, stress, chi, ad_Dpl and the reference map.

The single-rank quirk is pinned as well: the vertex update skips the plane
zz = 0 whenever the rank's x coordinate is 0 (src/shear_sim_3d.cpp:101), and
with one rank in z the boundary conditions never set the bottom plate
(:2393, :2468), so the bottom vertex plane -- and, with z periodic, the
plane z = 0 -- keeps its initial velocity and reference map."""
import numpy as np
import pytest

import oracle as O

CASES = ["direct_walled", "direct_periodic"]


def run_oracle(g, steps):
    N, prd = tuple(g["N"]), tuple(g["prd"])
    orc = O.OracleShear(N, {k: float(g[f"p_{k}"]) for k in O.STEP_PARAM_ORDER}, prd=prd)
    orc.set_direct(float(g["rho"]), float(g["kappa"]))
    orc.set_state(float(g["t0"]), g["u0"], g["sig0"], g["chi0"])
    orc.init_ref()
    out = {}
    for s in range(steps):
        orc.step_direct(float(g["dt"]), 1)
        if s + 1 in (1, 5):
            st = orc.state()
            st["xi"] = orc.ref_map()
            out[s + 1] = st
    return orc, out


@pytest.mark.parametrize("case", CASES)
def test_oracle_direct_step_is_the_reference(golden, case):
    g = golden(f"{case}.npz")
    orc, out = run_oracle(g, 5)
    for n in (1, 5):
        st = out[n]
        assert st["t"] == float(g[f"t{n}"])
        for key in ("u", "sig", "chi", "dpl", "xi"):
            assert np.array_equal(st[key], g[f"{key}{n}"]), (n, key)


@pytest.mark.parametrize("case", CASES)
def test_direct_bottom_plane_never_moves(golden, case):
    g = golden(f"{case}.npz")
    N = tuple(g["N"])
    nxy = N[0] * N[1]
    assert np.array_equal(g["u5"][:nxy], g["u0"][:nxy])
    assert np.array_equal(g["xi5"][:nxy], g["xi0"][:nxy])
    assert np.abs(g["u5"][nxy:2 * nxy] - g["u0"][nxy:2 * nxy]).max() > 0


def test_reference_map_starts_at_the_positions(golden):
    g = golden("direct_walled.npz")
    Nx, Ny, Nz = (int(v) for v in g["N"])
    xi = g["xi0"].reshape(Nz + 1, Ny, Nx, 3)
    i, j, k = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="xy")
    dx, dy, dz = 2.0 / Nx, 2.0 / Ny, 1.0 / Nz
    assert np.array_equal(xi[:Nz, ..., 0], np.transpose(-1 + dx * i, (2, 0, 1)))
    assert np.array_equal(xi[:Nz, ..., 2], np.transpose(-0.5 + dz * k, (2, 0, 1)))
    # top plate: x and y from the plate motion, z from the zero of the
    # extrapolated Field (src/field.hh:49-112)
    t, U = float(g["t0"]), float(g["p_U"])
    assert np.allclose(xi[Nz, :, :, 0], -1 + dx * np.arange(Nx)[None, :] + U * t * t * .5, rtol=0, atol=1e-15)
    assert np.all(xi[Nz, :, :, 2] == 0)
