"""Output in the reference's file formats, from the GPU state.

* Cubes of write_files' modes 0-12 for a state set through the C ABI equal
  the files the reference's output_cube wrote for the same state, byte for
  byte (tests/golden/output_walled.npz, output_periodic.npz).
* The wall tractions of the set state equal the reference's traction.dat
  record to 1e-13 (the GPU sums in another order); the transformed step's
  line equals the reference's "%g" line.
* During steps (per-step output on) proj_mag.dat and traction.dat follow the
  8-colour oracle's step (the GPU's smoother) to 1e-9; proj_mag of an
  elastic step from rest is sum of modsq(sigma_new - sigma_0) to 1e-12.
* write_files writes the cubes the reference's selection rules pick."""
import os

import numpy as np
import pytest

import oracle as O
from stzpp_amd import shear_sim as SS

pytestmark = pytest.mark.gpu


def params_of(g):
    return {k: float(g[f"p_{k}"]) for k in O.STEP_PARAM_ORDER}


def sim_from(g, family=None):
    N, prd = tuple(int(v) for v in g["N"]), tuple(int(v) for v in g["prd"])
    if family is None:
        sim = SS.ShearSim(*N, params=params_of(g), prd=prd)
    else:
        sim = SS.TransSim(*N, float(g["Ut"]), family=family, params=params_of(g), prd=prd)
    sim.set_initial_conditions(g["chi0"], u=g["u0"], sig=g["sig0"], t=float(g["t0"]))
    return sim


@pytest.mark.parametrize("case", ["output_walled", "output_periodic"])
def test_gpu_cubes_are_the_reference_files(golden, case):
    g = golden(f"{case}.npz")
    sim = sim_from(g)
    for mode, nm in enumerate(O.CUBE_NAMES):
        c = sim.output_cube(mode)
        assert c.tobytes() == g[f"cube_{nm}"].tobytes(), nm


@pytest.mark.parametrize("case", ["output_walled", "output_periodic"])
def test_gpu_traction_matches_reference_record(golden, case, tmp_path):
    g = golden(f"{case}.npz")
    sim = sim_from(g)
    want = [float(v) for v in str(g["traction_txt"]).splitlines()[0].split("\t")]
    nf = sim.compute_net_force()
    assert np.allclose(nf, want[1:], rtol=1e-13, atol=1e-16)
    sim.write_traction(tmp_path)
    line = (tmp_path / "traction.dat").read_text()
    assert line.count("\n") == 1 and line.split("\t")[0] == "%.16f" % float(g["t0"])
    assert np.allclose([float(v) for v in line.split("\t")], want, rtol=1e-13, atol=1e-16)
    rec = np.fromfile(tmp_path / "trac_grid.bin", dtype="<f8")
    assert np.allclose(rec, np.frombuffer(g["trac_bin"].tobytes()[:48], dtype="<f8"), rtol=1e-13, atol=1e-16)


def test_gpu_trans_traction_is_the_reference_line(golden, tmp_path):
    g = golden("output_trans.npz")
    sim = sim_from(g, family="pure_shear")
    sim.write_traction(tmp_path)
    assert (tmp_path / "traction.dat").read_text() == str(g["traction_txt"])
    T, _, _ = O.transform_at("pure_shear", float(g["Ut"]), 0.5, float(g["t0"]))
    want = O.net_force(tuple(g["N"]), (-1, 1, -1, 1, -0.5, 0.5), g["sig0"], T=T)
    assert np.allclose(sim.compute_net_force(), want, rtol=1e-13, atol=1e-16)


def test_gpu_step_output_follows_the_oracle(golden, tmp_path):
    g = golden("output_walled.npz")
    N = tuple(int(v) for v in g["N"])
    sim = sim_from(g)
    sim.set_output(tmp_path, traction=True, proj_mag=True)
    orc = O.OracleShear(N, params_of(g))
    orc.set_solver(1, 0.0)
    orc.set_state(float(g["t0"]), g["u0"], g["sig0"], g["chi0"])
    pm, tr = [], []
    for _ in range(2):
        sim.step_forward_qs(0.01, 1)
        orc.step(0.01, 1)
        s = orc.state()
        pm.append((s["t"], orc.proj_mag))
        tr.append([s["t"]] + O.net_force(N, (-1, 1, -1, 1, -0.5, 0.5), s["sig"]))
    lines = (tmp_path / "proj_mag.dat").read_text().splitlines()
    assert len(lines) == 2
    for q, ln in enumerate(lines):
        t, v = (float(x) for x in ln.split())
        assert "%g" % t == "%g" % pm[q][0]
        assert abs(v / pm[q][1] - 1) < 1e-5  # "%g" keeps 6 digits
    assert abs(sim.proj_mag / pm[-1][1] - 1) < 1e-9
    got = [[float(v) for v in ln.split("\t")] for ln in (tmp_path / "traction.dat").read_text().splitlines()]
    assert np.allclose(got, tr, rtol=1e-9, atol=1e-13)
    assert os.path.getsize(tmp_path / "trac_grid.bin") == 2 * 48


def test_gpu_proj_mag_of_an_elastic_step_from_rest():
    # at rest and unstressed the advection leaves sigma* = 0 (its Truesdell
    # terms are linear in sigma, and the moving plate already gives the top
    # cells a velocity gradient), so the projection's increments are the new
    # stress itself
    p = SS.config0_params(16)
    sim = SS.ShearSim(16, 12, 8, params=p)
    nc = 16 * 12 * 8
    sig0 = np.zeros((nc, 6))
    sim.set_initial_conditions(np.full(nc, 0.09), sig=sig0, t=0.5)
    sim.step_forward_qs(0.01, 1, solve=SS.SOLVE_RELRES, rtol=1e-13)
    d = sim.state()["sig"] - sig0
    want = 0.5 * (d[:, 0] ** 2 + 2 * d[:, 1] ** 2 + 2 * d[:, 2] ** 2 + d[:, 3] ** 2 + 2 * d[:, 4] ** 2 + d[:, 5] ** 2)
    assert want.sum() > 0
    assert abs(sim.proj_mag / want.sum() - 1) < 1e-12


def test_gpu_write_files_selection(golden, tmp_path):
    g = golden("output_walled.npz")
    sim = sim_from(g)
    d = tmp_path / "frames"
    sim.write_files(2, d, n_tp=20)            # every 2 frames: u..s33; chi since 2 % 2 == 0; tem
    sim.write_files(3, d, n_tp=20)            # tem only
    sim.write_files(3, d, n_tp=20, fflags=1)  # u appended to u_grid.3
    names = sorted(os.listdir(d))
    want = sorted([f"{n}_grid.2" for n in O.CUBE_NAMES if n not in ("p", "dev")] + ["tem_grid.3", "u_grid.3"])
    assert names == want
    assert (d / "u_grid.3").stat().st_size == sim.cube_size(0) * 4
    assert (d / "tem_grid.3").stat().st_size == sim.cube_size(12) * 4
    assert (d / "s11_grid.2").read_bytes() == g["cube_s11"].tobytes()
    with pytest.raises(SS.StepError):
        sim.write_files(1, d, n_tp=5)
    with pytest.raises(SS.StepError):
        sim.output_cube(13)
