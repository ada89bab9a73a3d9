"""The random initial chi field on the GPU against the reference and the
oracle.

* On the reference's own noise grids (tests/golden/grf_*.npz) the GPU field
  equals the reference's chi to 1e-12 sigma (the three 1-D passes sum in a
  different order from the reference's cube loop; the error is rounding).
* The noise drawn on the GPU equals the oracle's Philox + polar Box-Muller
  stream to 1e-15 relative (device and host log differ in the last bit).
* At full sizes (192 x 192 x 96 cells with a 20-cell cutoff, the LDS
  passes; 128^3 with the driver's 25-cell cutoff and 256^3 periodic with a
  40-cell cutoff, the direct passes) the GPU field equals the oracle's
  brute-force sum at 128 sampled cells, is linear in the noise, and has the
  requested mean and standard deviation.
* set_random_initial_conditions writes the simulation's chi and the
  reference's cache file, and loads an existing one (the reference's own
  bytes give the reference's chi exactly)."""
import os

import numpy as np
import pytest

import oracle as O
from stzpp_amd import grf as G
from stzpp_amd import shear_sim as SS

pytestmark = pytest.mark.gpu

CASES = ["grf_walled", "grf_periodic", "grf_open", "grf_wrap", "grf_c0"]


def params_of(g):
    return dict(gauss_fac=float(g["gauss_fac"]), conv_l=float(g["conv_l"]), chi_avg=float(g["chi_avg"]),
                sigma=float(g["sigma"]), TZ=float(g["TZ"]), sim_case=int(g["sim_case"]))


@pytest.mark.parametrize("case", CASES)
def test_gpu_grf_on_reference_noise_matches_reference(golden, case):
    g = golden(f"{case}.npz")
    N, prd = tuple(int(v) for v in g["N"]), tuple(int(v) for v in g["prd"])
    print(f"case {case} N={N} prd={prd} cut={int(g['cut'])}", flush=True)
    chi = G.grf_field(N, prd, noise=g["noise"], **params_of(g))
    assert np.abs(chi - g["chi"]).max() <= 1e-12 * float(g["sigma"])


def test_gpu_noise_matches_oracle_stream():
    # cut 0, mean 0, sigma 1: chi is the noise itself (odd count: the last
    # pair keeps its first value)
    N, prd, seed = (37, 21, 13), (1, 1, 1), 2024
    noise = G.grf_field(N, prd, gauss_fac=0, conv_l=1.0, chi_avg=0.0, sigma=1.0, seed=seed)
    want = O.grf_noise(seed, noise.size)
    rel = np.abs(noise - want) / np.maximum(np.abs(want), 1e-300)
    w = int(np.argmax(rel))
    print(f"differing {int((rel > 0).sum())} of {rel.size}; worst {w}: gpu {noise[w]!r} oracle {want[w]!r} "
          f"rel {rel[w]:.3e}", flush=True)
    assert rel.max() <= 1e-15, (w, noise[w], want[w], rel[w], int((rel > 0).sum()))
    assert abs(noise.mean()) < 0.05 and abs(noise.std() - 1) < 0.05


def test_gpu_seeded_field_matches_oracle():
    N, prd, seed = (24, 20, 12), (1, 1, 0), 5
    p = dict(gauss_fac=2, conv_l=3.0, chi_avg=0.09, sigma=0.01)
    chi = G.grf_field(N, prd, seed=seed, **p)
    want = O.grf_field(N, prd, p["gauss_fac"], p["conv_l"], p["chi_avg"], p["sigma"], seed=seed)
    assert np.abs(chi - want).max() <= 1e-12 * p["sigma"]


FULL = [
    # name, N, prd, gauss_fac, conv_l  (cut)
    ("cells192x192x96_l4", (192, 192, 96), (1, 1, 0), 5, 4.0),      # 20: LDS passes
    ("cells128_driver", (128, 128, 128), (1, 1, 0), 5, 5.0),       # 25: direct passes
    ("cells256_periodic_l8", (256, 256, 256), (1, 1, 1), 5, 8.0),  # 40: direct passes
]


@pytest.mark.parametrize("name,N,prd,gf,cl", FULL, ids=[f[0] for f in FULL])
def test_gpu_grf_full_size(name, N, prd, gf, cl):
    print(f"case {name}", flush=True)
    p = dict(gauss_fac=gf, conv_l=cl, chi_avg=0.09, sigma=0.01)
    cut, nn, chi_1 = G.geometry(N, prd, **p)
    a = O.grf_noise(3, nn)
    b = O.grf_noise(4, nn)
    chi_a = G.grf_field(N, prd, noise=a, **p)
    # the oracle's brute force at sampled cells (corners and faces included)
    rng = np.random.default_rng(0)
    nc = int(np.prod(N))
    idx = np.unique(np.concatenate([rng.integers(0, nc, 120), [0, N[0] - 1, nc - 1, nc - N[0] * N[1]]]))
    want = O.grf_convolve_points(N, prd, cut, 1 / cl, 0.09, chi_1, a, idx)
    assert np.abs(chi_a[idx] - want).max() <= 1e-12 * p["sigma"]
    # linear in the noise
    chi_b = G.grf_field(N, prd, noise=b, **p)
    chi_ab = G.grf_field(N, prd, noise=a + b, **p)
    assert np.abs((chi_ab - 0.09) - ((chi_a - 0.09) + (chi_b - 0.09))).max() <= 1e-12 * p["sigma"]
    # the noise drawn on the GPU from seed 3 is the oracle's stream
    chi_s = G.grf_field(N, prd, seed=3, **p)
    assert np.abs(chi_s - chi_a).max() <= 1e-12 * p["sigma"]
    # mean and standard deviation (correlated samples: loose bounds)
    assert abs(chi_a.mean() - 0.09) < 0.2 * p["sigma"]
    assert abs(chi_a.std() / p["sigma"] - 1) < 0.2


def test_sim_random_initial_conditions_and_cache(golden, tmp_path):
    g = golden("grf_walled.npz")
    N = tuple(int(v) for v in g["N"])
    p = params_of(g)
    d1 = tmp_path / "computed"
    d1.mkdir()
    sim = SS.ShearSim(*N)
    sim.set_random_initial_conditions(noise=g["noise"], cache_dir=d1, **p)
    chi = sim.state()["chi"]
    assert np.abs(chi - g["chi"]).max() <= 1e-12 * p["sigma"]
    files = os.listdir(d1)
    assert files == [str(g["cache_name"])]
    assert np.array_equal(np.fromfile(d1 / files[0], dtype="<f8"), chi)
    # an existing cache is loaded (the seed is not used)
    sim2 = SS.ShearSim(*N)
    sim2.set_random_initial_conditions(cache_dir=d1, seed=99, **p)
    assert np.array_equal(sim2.state()["chi"], chi)
    # the reference's own cache file gives the reference's chi exactly
    d2 = tmp_path / "reference"
    d2.mkdir()
    (d2 / str(g["cache_name"])).write_bytes(g["cache_bytes"].tobytes())
    sim3 = SS.ShearSim(*N)
    sim3.set_random_initial_conditions(cache_dir=d2, **p)
    assert np.array_equal(sim3.state()["chi"], g["chi"])
    # a truncated cache is an error, not a silent field
    d3 = tmp_path / "short"
    d3.mkdir()
    (d3 / str(g["cache_name"])).write_bytes(g["cache_bytes"].tobytes()[:800])
    with pytest.raises(SS.StepError):
        SS.ShearSim(*N).set_random_initial_conditions(cache_dir=d3, **p)
