"""The direct (explicit dynamics) step on the GPU (stz_step_direct) against
the reference's own runs (tests/golden/direct_*.npz, 16 x 12 x 8 cells,
walled and periodic z) and the oracle.

* The reference map from init_ref equals the reference's bit for bit.
* After 1 and 5 steps velocity, stress, chi, ad_Dpl and the reference map
  equal the reference's to 1e-12 of each field's largest magnitude (the
  STZ rate goes through exp, whose last bit the GPU's library may round
  otherwise than glibc; everything else rounds as the reference does).
* Without yielding (s_y above every sbar) no transcendental function
  enters, and 4 steps equal the oracle's bit for bit.
* The vertex plane k = 0 never moves, the time advances by dt per step,
  and the reference map cubes (modes 13-15) are the stored map in the
  reference's cube layout."""
import numpy as np
import pytest

import oracle as O
from stzpp_amd import shear_sim as SS

pytestmark = pytest.mark.gpu

CASES = ["direct_walled", "direct_periodic"]
FIELDS = ("u", "sig", "chi", "dpl", "xi")


def params_of(g, **over):
    p = {k: float(g[f"p_{k}"]) for k in O.STEP_PARAM_ORDER}
    p.update(rho=float(g["rho"]), kappa=float(g["kappa"]), tmult=float(g["tmult"]))
    p.update(over)
    return p


def sim_from(g, **over):
    N, prd = tuple(int(v) for v in g["N"]), tuple(int(v) for v in g["prd"])
    sim = SS.ShearSim(*N, params=params_of(g, **over), prd=prd)
    sim.set_initial_conditions(g["chi0"], u=g["u0"], sig=g["sig0"], t=float(g["t0"]))
    sim.init_ref()
    return sim


def snapshot(sim):
    st = sim.state()
    st["xi"] = sim.ref_map()
    return st


def close(a, b, rel=1e-12):
    scale = max(1.0, float(np.abs(b).max()))
    return float(np.abs(a - b).max()) <= rel * scale


@pytest.mark.parametrize("case", CASES)
def test_gpu_direct_dt_and_init_ref(golden, case):
    g = golden(f"{case}.npz")
    sim = sim_from(g)
    assert sim.direct_dt() == float(g["dt"])
    assert np.array_equal(sim.ref_map(), g["xi0"])


@pytest.mark.parametrize("case", CASES)
def test_gpu_direct_steps_follow_the_reference(golden, case):
    g = golden(f"{case}.npz")
    sim = sim_from(g)
    dt = float(g["dt"])
    for n in range(1, 6):
        stats = sim.step_forward(dt)
        assert stats["steps"] == 1
        if n in (1, 5):
            st = snapshot(sim)
            assert st["t"] == float(g[f"t{n}"])
            for key in FIELDS:
                assert close(st[key], g[f"{key}{n}"]), (n, key, np.abs(st[key] - g[f"{key}{n}"]).max())
    assert sim.last_stats["yielded_cells"] > 0
    nxy = int(g["N"][0]) * int(g["N"][1])
    st = snapshot(sim)
    assert np.array_equal(st["u"][:nxy], g["u0"][:nxy])
    assert np.array_equal(st["xi"][:nxy], g["xi0"][:nxy])


@pytest.mark.parametrize("case", CASES)
def test_gpu_direct_elastic_steps_are_the_oracle_bit_for_bit(golden, case):
    g = golden(f"{case}.npz")
    over = {"sy": 1e6}
    sim = sim_from(g, **over)
    N, prd = tuple(int(v) for v in g["N"]), tuple(int(v) for v in g["prd"])
    P = {k: v for k, v in params_of(g, **over).items() if k in O.STEP_PARAM_ORDER}
    orc = O.OracleShear(N, P, prd=prd)
    orc.set_direct(float(g["rho"]), float(g["kappa"]))
    orc.set_state(float(g["t0"]), g["u0"], g["sig0"], g["chi0"])
    orc.init_ref()
    dt = float(g["dt"])
    for n in range(4):
        sim.step_forward(dt)
        orc.step_direct(dt, 1)
    assert sim.last_stats["yielded_cells"] == 0
    st, want = snapshot(sim), orc.state()
    want["xi"] = orc.ref_map()
    assert st["t"] == want["t"]
    for key in FIELDS:
        assert np.array_equal(st[key], want[key]), (key, np.abs(st[key] - want[key]).max())


def test_gpu_direct_several_steps_in_one_call_equal_single_steps(golden):
    g = golden("direct_walled.npz")
    a, b = sim_from(g), sim_from(g)
    dt = float(g["dt"])
    a.step_forward(dt, 3)
    for _ in range(3):
        b.step_forward(dt)
    sa, sb = snapshot(a), snapshot(b)
    for key in FIELDS:
        assert np.array_equal(sa[key], sb[key]), key


@pytest.mark.parametrize("case", CASES)
def test_gpu_reference_map_cubes(golden, case):
    g = golden(f"{case}.npz")
    sim = sim_from(g)
    sim.step_forward(float(g["dt"]), 2)
    Nx, Ny, Nz = (int(v) for v in g["N"])
    pz = int(g["prd"][2])
    xi = sim.ref_map().reshape(Nz + (0 if pz else 1), Ny, Nx, 3)
    if pz:  # plane Nz is the periodic image of plane 0, z shifted by the domain length
        top = xi[:1].copy()
        top[..., 2] += 1.0
        xi = np.concatenate([xi, top], axis=0)
    for q, mode in enumerate((13, 14, 15)):
        want = np.transpose(xi[..., q], (2, 1, 0)).astype(np.float32).ravel()
        assert sim.cube_size(mode) == want.size
        assert np.array_equal(sim.output_cube(mode), want), mode


def test_gpu_direct_argument_errors(golden):
    g = golden("direct_walled.npz")
    N = tuple(int(v) for v in g["N"])
    sim = SS.ShearSim(*N, params=params_of(g))
    with pytest.raises(SS.StepError):
        sim.output_cube(13)  # no reference map yet
    with pytest.raises(SS.StepError):
        sim.step_forward(-1.0)
    bad = SS.ShearSim(*N, params=params_of(g, rho=0.0))
    with pytest.raises(SS.StepError):
        bad.step_forward(1e-3)
    tr = SS.TransSim(16, 12, 8, 0.2, params=params_of(g), prd=(1, 1, 0))
    with pytest.raises(SS.StepError):
        tr.step_forward(1e-3)


# ---- l2_comparison and the shear_compare loop --------------------------------

def l2_sims(g):
    N, prd = tuple(int(v) for v in g["N"]), tuple(int(v) for v in g["prd"])
    p = params_of(g)
    d = SS.ShearSim(*N, params=p, prd=prd)
    q = SS.ShearSim(*N, params=p, prd=prd)
    return d, q


@pytest.mark.parametrize("case", ["l2_walled", "l2_periodic"])
def test_gpu_l2_comparison_is_the_reference(golden, case):
    g = golden(f"{case}.npz")
    d, q = l2_sims(g)
    d.set_initial_conditions(g["d_chi"], u=g["d_u"], sig=g["d_sig"], t=float(g["d_t"]))
    q.set_initial_conditions(g["q_chi"], u=g["q_u"], sig=g["q_sig"], t=float(g["q_t"]))
    got = d.l2_comparison(q)
    assert np.allclose(got, g["l2"], rtol=1e-13, atol=0), (got, g["l2"])
    assert np.array_equal(d.l2_comparison(d), np.zeros(3))


def test_gpu_shear_compare_loop_follows_the_reference(golden, tmp_path):
    # direct: the reference's steps to t0 + 0.02 (5 full steps and a
    # shortened one), bit-for-bit up to exp's last bit; QS: the GPU's
    # 8-colour projection loop against the reference's lexicographic one
    g = golden("l2_walled.npz")
    d, q = l2_sims(g)
    for s in (d, q):
        s.set_initial_conditions(g["chi0"], u=g["u0"], sig=g["sig0"], t=float(g["t0"]))
    d.init_ref()
    rows = SS.shear_compare(d, q, 1, 2 * float(g["qdt"]), float(g["qdt"]), out_path=tmp_path / "sc0.dat")
    st = d.state()
    assert st["t"] == float(g["d_t"])
    for key in ("u", "sig", "chi"):
        assert close(st[key], g[f"d_{key}"]), key
    sq = q.state()
    assert abs(sq["t"] - float(g["q_t"])) < 1e-15
    l2 = d.l2_comparison(q)
    assert np.allclose(l2, g["l2"], rtol=1e-6), (l2, g["l2"])
    lines = (tmp_path / "sc0.dat").read_text().splitlines()
    assert len(lines) == 2 and len(rows) == 2
    assert lines[0].split()[1:] == ["0", "0", "0"]
    assert np.allclose([float(v) for v in lines[1].split()[1:]], np.sqrt(g["l2"]), rtol=1e-6)
