"""l2_comparison (src/shear_sim_3d.cpp:3122-3158) restated in the oracle.

* Against the reference's own l2_comparison of a direct run against a
  quasi-static one from the same state (tests/golden/l2_*.npz, as
  shear_compare does it): to 1e-13 relative.
* Closed forms the normalisation fixes: a constant velocity difference c
  gives (c/U)^2 (the trapezoid weights in z sum to Nz on Nz + 1 vertex
  planes, walled or periodic), a constant stress-component difference d
  gives d^2, a constant chi difference e gives (e/chi_inf)^2.
* The trapezoid weights themselves: a difference on one vertex of plane 0
  or of the top plane counts 1/2, one inside counts 1.
* The reference's accumulators start from an indeterminate value: its l2 of
  identical states is not exactly zero (tiny in the fixture's run); the
  oracle's is zero."""
import numpy as np
import pytest

import oracle as O

BOX = (-1, 1, -1, 1, -0.5, 0.5)


def states(g):
    a = {k: g[f"d_{k}"] for k in ("u", "sig", "chi")}
    b = {k: g[f"q_{k}"] for k in ("u", "sig", "chi")}
    return a, b


@pytest.mark.parametrize("case", ["l2_walled", "l2_periodic"])
def test_oracle_l2_is_the_reference(golden, case):
    g = golden(f"{case}.npz")
    a, b = states(g)
    got = O.l2_comparison(tuple(g["N"]), tuple(g["prd"]), BOX, float(g["p_U"]), float(g["p_chi_inf"]), a, b)
    assert np.allclose(got, g["l2"], rtol=1e-13, atol=0)
    assert min(got) > 0


@pytest.mark.parametrize("prd", [(1, 1, 0), (1, 1, 1)])
def test_l2_of_constant_differences(prd):
    N = (6, 5, 4)
    nc = 6 * 5 * 4
    nv = 6 * 5 * (4 + (0 if prd[2] else 1))
    rng = np.random.default_rng(5)
    a = {"u": rng.standard_normal((nv, 3)), "sig": rng.standard_normal((nc, 6)), "chi": rng.standard_normal(nc)}
    U, chi_inf, c, d, e = 0.01, 0.15, 3e-3, 0.25, 0.02
    b = {"u": a["u"] + np.array([c, 0, 0]), "sig": a["sig"].copy(), "chi": a["chi"] + e}
    b["sig"][:, 4] += d
    got = O.l2_comparison(N, prd, BOX, U, chi_inf, a, b)
    assert np.allclose(got, [(c / U) ** 2, d * d, (e / chi_inf) ** 2], rtol=1e-12)


@pytest.mark.parametrize("k,w", [(0, 0.5), (2, 1.0), (4, 0.5)])
def test_l2_trapezoid_weights(k, w):
    N, prd = (6, 5, 4), (1, 1, 0)
    nc, nv = 120, 150
    a = {"u": np.zeros((nv, 3)), "sig": np.zeros((nc, 6)), "chi": np.zeros(nc)}
    b = {"u": np.zeros((nv, 3)), "sig": np.zeros((nc, 6)), "chi": np.zeros(nc)}
    b["u"][3 + 6 * (2 + 5 * k), 1] = 2.0
    got = O.l2_comparison(N, prd, BOX, 1.0, 1.0, a, b)
    vol_per_point = 1.0 / nc
    assert got[0] == pytest.approx(w * 4.0 * vol_per_point, rel=1e-14)
    assert got[1] == 0 and got[2] == 0


def test_reference_l2_of_identical_states_is_not_exactly_zero(golden):
    # the fixture's l2 at the start: both simulations in the same state
    g = golden("l2_walled.npz")
    assert np.all(np.abs(g["l2_0"]) < 1e-300) and np.any(g["l2_0"] != 0)
    a, _ = states(g)
    assert O.l2_comparison(tuple(g["N"]), tuple(g["prd"]), BOX, 0.01, 0.15, a, a) == [0.0, 0.0, 0.0]
