"""The random initial chi field: the oracle pinned to the reference, and the
library's host-side pieces (cutoff, normalisation, cache file name).

Fixtures tests/golden/grf_*.npz come from the reference itself
(set_random_initial_conditions, src/shear_sim_3d.cpp:2793-2934, through
oracle/_ref): its own noise grid (glibc rand() from the unseeded state), the
chi it computes and the name of the cache file it writes.  The noise
generator of this library (Philox4x64-10 + polar Box-Muller) has no
reference counterpart; it is pinned to numpy's independent Philox and to
Random123's published known-answer block."""
import numpy as np
import pytest

import oracle as O
from stzpp_amd import grf as G

CASES = ["grf_walled", "grf_periodic", "grf_open", "grf_wrap", "grf_c0"]


def params_of(g):
    return dict(gauss_fac=float(g["gauss_fac"]), conv_l=float(g["conv_l"]), chi_avg=float(g["chi_avg"]),
                sigma=float(g["sigma"]), TZ=float(g["TZ"]), sim_case=int(g["sim_case"]))


@pytest.mark.parametrize("case", CASES)
def test_oracle_convolution_is_the_reference(golden, case):
    g = golden(f"{case}.npz")
    N, prd = tuple(g["N"]), tuple(g["prd"])
    cut = O.grf_cut(float(g["gauss_fac"]), float(g["conv_l"]))
    assert cut == int(g["cut"])
    llinv = 1.0 / float(g["conv_l"])
    assert O.grf_normal_factor(cut, llinv) == float(g["gnf"])
    assert float(g["sigma"]) / float(g["gnf"]) == float(g["chi_1"])
    chi = O.grf_convolve(N, prd, cut, llinv, float(g["chi_avg"]), float(g["chi_1"]), g["noise"])
    assert np.array_equal(chi, g["chi"])
    idx = np.arange(0, chi.size, 7)
    pts = O.grf_convolve_points(N, prd, cut, llinv, float(g["chi_avg"]), float(g["chi_1"]), g["noise"], idx)
    assert np.array_equal(pts, g["chi"][idx])
    assert O.grf_cache_name(N, prd, int(g["sim_case"]), float(g["TZ"]), float(g["chi_avg"]),
                            float(g["chi_1"]), float(g["gnf"])) == str(g["cache_name"])


def test_fractional_gauss_fac_is_truncated(golden):
    # the reference stores gauss_fac in an int (src/shear_sim_3d.hh:140):
    # 1.5 x 3 cells gives a cutoff of 3, not 4
    g = golden("grf_periodic.npz")
    assert float(g["gauss_fac"]) == 1.5 and float(g["conv_l"]) == 3.0 and int(g["cut"]) == 3


def test_reference_cache_file_is_raw_chi(golden):
    g = golden("grf_walled.npz")
    assert np.array_equal(np.frombuffer(g["cache_bytes"].tobytes(), dtype="<f8"), g["chi"])


def test_oracle_impulse_response_is_the_kernel():
    # a single noise value at one cell reproduces the cube-cutoff Gaussian,
    # placed through the padding of the walled axis
    N, prd, cut, ell = (12, 10, 9), (1, 1, 0), 3, 2.0
    nt = O.grf_noise_shape(N, prd, cut)
    noise = np.zeros(nt[::-1])
    x0, y0, z0 = 2, 9, 1
    noise[z0 + cut, y0, x0] = 1.0
    chi = O.grf_convolve(N, prd, cut, 1 / ell, 0.0, 1.0, noise).reshape(N[::-1])
    z, y, x = np.meshgrid(np.arange(N[2]), np.arange(N[1]), np.arange(N[0]), indexing="ij")
    dx = (x - x0 + N[0] // 2) % N[0] - N[0] // 2
    dy = (y - y0 + N[1] // 2) % N[1] - N[1] // 2
    dz = z - z0
    inside = (np.abs(dx) <= cut) & (np.abs(dy) <= cut) & (np.abs(dz) <= cut)
    want = np.where(inside, np.exp(-(dx * dx + dy * dy + dz * dz) / ell), 0.0)
    assert np.abs(chi - want).max() <= 1e-15


def test_normal_factor_closed_form():
    for cut, ell in [(0, 1.0), (3, 2.0), (15, 3.0), (25, 5.0)]:
        s1 = sum(np.exp(-2 * i * i / ell) for i in range(-cut, cut + 1))
        assert abs(O.grf_normal_factor(cut, 1 / ell) - s1 ** 1.5) <= 1e-13 * s1 ** 1.5


def test_philox_matches_numpy_and_known_answer():
    # numpy's Philox increments its 256-bit counter before each block
    rng = np.random.default_rng(3)
    for _ in range(6):
        ctr = [int(v) for v in rng.integers(0, 2**63, 4)]
        ctr[0] |= 1
        key = [int(v) for v in rng.integers(0, 2**63, 2)]
        prev = list(ctr)
        prev[0] -= 1
        ref = np.random.Philox(counter=np.array(prev, dtype=np.uint64),
                               key=np.array(key, dtype=np.uint64)).random_raw(4)
        assert np.array_equal(O.philox4x64(ctr, key), ref)
    # Random123 known answer for Philox4x64-10 with zero counter and key
    kat = [0x16554D9ECA36314C, 0xDB20FE9D672D0FDC, 0xD7E772CEE186176B, 0x7E68B68AEC7BA23B]
    assert [int(v) for v in O.philox4x64([0, 0, 0, 0], [0, 0])] == kat


def test_noise_is_the_polar_method_on_philox():
    seed, n = 12345, 4001
    noise = O.grf_noise(seed, n)
    key1 = 0x53545A2B2B475246
    for p in (0, 1, 17, 999, 2000):
        for a in range(64):
            o = O.philox4x64([p, a, 0, 0], [seed, key1])
            x = 2.0 * (float(int(o[0]) >> 11) * 2.0**-53) - 1.0
            y = 2.0 * (float(int(o[1]) >> 11) * 2.0**-53) - 1.0
            r = x * x + y * y
            if 0.0 < r <= 1.0:
                break
        d = np.sqrt(-2.0 * (np.log1p(r - 1.0) if r >= 0.5 else np.log(r)) / r)
        assert abs(noise[2 * p] - x * d) <= 4e-16 * abs(x * d)
        if 2 * p + 1 < n:
            assert abs(noise[2 * p + 1] - y * d) <= 4e-16 * abs(y * d)
    # prefix-stable and standard normal
    assert np.array_equal(O.grf_noise(seed, n - 1), noise[:-1])
    big = O.grf_noise(7, 400000)
    assert abs(big.mean()) < 5 / np.sqrt(big.size) and abs(big.std() - 1) < 5 / np.sqrt(big.size)


# ---- the library's host-side pieces (no GPU calls) ------------------------

@pytest.mark.parametrize("case", CASES)
def test_library_geometry_and_cache_name(golden, case):
    g = golden(f"{case}.npz")
    N, prd = tuple(int(v) for v in g["N"]), tuple(int(v) for v in g["prd"])
    cut, nn, chi_1 = G.geometry(N, prd, **params_of(g))
    assert cut == int(g["cut"]) and nn == g["noise"].size
    assert chi_1 == float(g["chi_1"])
    assert G.noise_shape(N, prd, **params_of(g)) == O.grf_noise_shape(N, prd, cut)
    assert G.cache_name(N, prd, **params_of(g)) == str(g["cache_name"])


def test_library_defaults_are_the_driver_values():
    d = G.default_params()
    assert d["gauss_fac"] == 5 and d["conv_l"] == 5 and d["TZ"] == 21000
    assert d["chi_avg"] == 550.0 / 21000 and d["sigma"] == 15.0 / 21000


def test_library_rejects_bad_configurations():
    with pytest.raises(G.GrfError):
        G.geometry((8, 8, 8), (1, 1, 1), gauss_fac=3, conv_l=3)  # cut 9 > periodic 8
    with pytest.raises(G.GrfError):
        G.geometry((8, 8, 8), (1, 1, 0), conv_l=0.0)
    # the largest cutoff the single wrap reaches is accepted
    assert G.geometry((8, 8, 8), (1, 1, 0), gauss_fac=2, conv_l=4)[0] == 8
