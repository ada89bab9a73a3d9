"""The reference's output files, restated in the oracle and pinned to the
files the reference itself wrote (tests/golden/output_*.npz: output_cube,
compute_net_force and proj_mag.dat through oracle/_ref).

* Cubes of write_files' modes 0-12 (float32, x slowest, z fastest; the
  velocity carries the top plate plane, or the periodic image of plane 0):
  byte for byte.
* traction.dat (the set state) and the binary trac_grid.bin record: text
  and bytes exactly; the transformed step's traction line (physical stress,
  %g) exactly.
* proj_mag.dat and the traction lines after two reference steps, from the
  oracle's step with the reference's smoother and stopping rule: the %g
  text of proj_mag exactly, the %.16f tractions to 1e-9 (the oracle's step
  agrees with the reference's to ~1e-11)."""
import numpy as np
import pytest

import oracle as O

BOX = (-1, 1, -1, 1, -0.5, 0.5)


def params_of(g):
    return {k: float(g[f"p_{k}"]) for k in O.STEP_PARAM_ORDER}


def with_top_plate(g):
    """The state after the reference's boundary conditions: the top vertex
    plane moves with the plate, (U t, 0, 0) for t < 1
    (set_boundary_conditions, src/shear_sim_3d.cpp:2362-2658)."""
    u = np.array(g["u0"])
    N, prd = tuple(g["N"]), tuple(g["prd"])
    if not prd[2]:
        top = N[0] * N[1] * N[2]
        U, t = float(g["p_U"]), float(g["t0"])
        u[top:] = [U * t if t < 1 else U, 0.0, 0.0]
    return u


@pytest.mark.parametrize("case", ["output_walled", "output_periodic"])
def test_oracle_cubes_are_the_reference_files(golden, case):
    g = golden(f"{case}.npz")
    N, prd = tuple(g["N"]), tuple(g["prd"])
    u = with_top_plate(g)
    for mode, nm in enumerate(O.CUBE_NAMES):
        c = O.cube_values(mode, N, prd, u, g["sig0"], g["chi0"])
        assert c.tobytes() == g[f"cube_{nm}"].tobytes(), nm


@pytest.mark.parametrize("case", ["output_walled", "output_periodic"])
def test_oracle_traction_is_the_reference_record(golden, case):
    g = golden(f"{case}.npz")
    nf = O.net_force(tuple(g["N"]), BOX, g["sig0"])
    assert O.traction_line(float(g["t0"]), nf) == str(g["traction_txt"]).splitlines(True)[0]
    assert np.array(nf).tobytes() == g["trac_bin"].tobytes()[:48]


def test_bottom_traction_pairs_x_neighbours(golden):
    # the reference's bottom "extrapolation" uses the +x neighbour (grido + 1,
    # src/shear_sim_3d.cpp:1068); over a periodic row it sums to the plain
    # sum of the bottom cells, unlike a z extrapolation
    g = golden("output_walled.npz")
    N = tuple(g["N"])
    s = np.asarray(g["sig0"]).reshape(N[2], N[1], N[0], 6)
    dxdy = (2.0 / N[0]) * (2.0 / N[1])
    nf = O.net_force(N, BOX, g["sig0"])
    assert abs(nf[3] - (-dxdy * s[0, :, :, 2].sum())) < 1e-13
    zext = -dxdy * (1.5 * s[0, :, :, 2] - 0.5 * s[1, :, :, 2]).sum()
    assert abs(nf[3] - zext) > 1e-3


def test_oracle_trans_traction_is_the_reference_line(golden):
    g = golden("output_trans.npz")
    T, _, _ = O.transform_at("pure_shear", float(g["Ut"]), 0.5, float(g["t0"]))
    nf = O.net_force(tuple(g["N"]), BOX, g["sig0"], T=T)
    assert O.traction_line(float(g["t0"]), nf, trans=True) == str(g["traction_txt"])


def test_oracle_steps_reproduce_proj_mag_and_tractions(golden):
    g = golden("output_walled.npz")
    N = tuple(g["N"])
    orc = O.OracleShear(N, params_of(g))
    orc.set_solver(0, 0.0)
    orc.set_state(float(g["t0"]), g["u0"], g["sig0"], g["chi0"])
    lines = str(g["traction_txt"]).splitlines()
    pm_lines = str(g["proj_mag_txt"]).splitlines()
    for q in range(2):
        orc.step(0.01, 1)
        s = orc.state()
        assert "%g %g" % (s["t"], orc.proj_mag) == pm_lines[q]
        want = [float(v) for v in lines[q + 1].split("\t")]
        got = [s["t"]] + O.net_force(N, BOX, s["sig"])
        assert np.allclose(got, want, rtol=1e-9, atol=1e-12), (q, got, want)
