"""Parity of the z-chunked paths of the hot multigrid kernels with the oracle.

The fused smoother (every tile class: 64-, 32-, 16- and 8-wide tiles) and
the plane-marching residual cut z into chunks sized to the workgroups the
device keeps resident; the restriction marches a run
of coarse planes per thread.  At the sizes an oracle finishes in seconds the
device's resident count gives one plane per chunk, so these tests cap it
(STZMG_OPT_CHUNK_CAP) and fix the restriction's run (STZMG_OPT_RESTRICT_PLANES)
to make every workgroup march several planes, ragged last chunks included,
and compare with the oracle (reference semantics: region.cc:1009-1055 for
the residual, :1558-1762 for the sweep, :1767-1812 for the restriction).

The FP32 smoother and residual of the refinement's inner solve are compared
with the oracle's FP64 sweep and residual of the same FP32-rounded inputs:
the only difference is FP32 arithmetic, bounded here by 2e-6 of the largest
value (measured below 5e-7).
"""
import numpy as np
import pytest

import oracle as O
from stzpp_amd import multigrid as MG

pytestmark = pytest.mark.gpu

T_SHEAR = np.array([[1.0, 0.0, 0.5], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
# caps in workgroups: 1 marches all planes of a parity in one workgroup per
# tile; the others give chunks of several planes with a ragged last one
CAPS = [0, 1, 24, 88]


def relerr(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def operator(kind, n, prd):
    h = 1.0 / n
    o = n + (0 if prd[2] else 1)
    st, fix = MG.shear_stencil(1.0, 1.5, h, h, h, 1.0)
    if kind == "trans":
        st = MG.trans_stencil(T_SHEAR, 1.0, 1.5, h, h, h, 1.0)
    elif kind == "rat":
        big = MG.Multigrid(2 * n, 2 * n, 2 * n, (1, 1, 1))
        big.setup_matrices_const(MG.trans_stencil(T_SHEAR, 1.0, 1.5, h / 2, h / 2, h / 2, 1.0), None)
        st = big.level_operator(1)[0]
    return o, O.plane_operator(o, prd[2], st, fix if not prd[2] else None)


def oracle_for(n, o, prd, planes, x, r):
    orc = O.OracleMG(n, n, o, prd, mode=O.PLANE)
    orc.set_ops(0, planes)
    orc.params(smoother=1)
    orc.set_fields(0, x, r)
    return orc


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


@pytest.mark.parametrize("cap", CAPS)
@pytest.mark.parametrize("kind,prd,n", [("trans", (1, 1, 1), 64), ("shear", (1, 1, 0), 64),
                                        ("rat", (1, 1, 0), 64), ("trans", (1, 1, 0), 32),
                                        # 24 x 24 planes: the 8 x 8 tile class
                                        ("trans", (1, 1, 1), 24), ("shear", (1, 1, 0), 24)])
def test_gpu_fused_sweep_chunked_matches_oracle(kind, prd, n, cap):
    o, planes = operator(kind, n, prd)
    N = n * n * o
    x, r = O.det_field(N, 71), O.det_field(N, 72)
    mg = MG.Multigrid(n, n, o, prd)
    mg.set_option(MG.OPT_CHUNK_CAP, cap)
    mg.setup_matrices_planes(planes)
    # FP64: three sweeps against the oracle's 8-colour sweeps
    orc = oracle_for(n, o, prd, planes, x, r)
    orc.gs(0, 3)
    mg.setup_fields(x, r)
    mg.gauss_seidel(0, 3)
    assert relerr(mg.x(0), orc.fields(0)[0]) < 1e-12
    # FP32: one sweep of the FP32-rounded fields
    orc = oracle_for(n, o, prd, planes, f32(x), f32(r))
    orc.gs(0, 1)
    mg.setup_fields(x, r)
    mg.gauss_seidel_fp32(0, 1)
    assert relerr(mg.x(0), orc.fields(0)[0]) < 2e-6


@pytest.mark.parametrize("cap", CAPS)
@pytest.mark.parametrize("kind,prd,n", [("trans", (1, 1, 1), 64), ("shear", (1, 1, 0), 64),
                                        ("rat", (1, 1, 1), 64)])
def test_gpu_tiled_residual_chunked_matches_oracle(kind, prd, n, cap):
    o, planes = operator(kind, n, prd)
    N = n * n * o
    x, r = O.det_field(N, 73), O.det_field(N, 74)
    mg = MG.Multigrid(n, n, o, prd)
    mg.set_option(MG.OPT_CHUNK_CAP, cap)
    mg.setup_matrices_planes(planes)
    mg.setup_fields(x, r)
    orc = oracle_for(n, o, prd, planes, x, r)
    assert relerr(mg.residual(0), orc.residual(0)) < 1e-13
    assert abs(mg.norm(0, MG.NORM_QUARTIC) / orc.norm_quartic(0) - 1) < 1e-12
    assert abs(mg.norm(0, MG.NORM_L2SQ) / orc.norm2(0) - 1) < 1e-12
    orc32 = oracle_for(n, o, prd, planes, f32(x), f32(r))
    assert relerr(mg.residual_fp32(0), orc32.residual(0)) < 2e-6


@pytest.mark.parametrize("planes_per_thread", [1, 2, 3, 5, 64])
@pytest.mark.parametrize("prd", [(1, 1, 1), (1, 1, 0)])
def test_gpu_restriction_marching_matches_oracle(prd, planes_per_thread):
    n = 64
    o, planes = operator("trans", n, prd)
    N = n * n * o
    x, r = O.det_field(N, 75), O.det_field(N, 76)
    mg = MG.Multigrid(n, n, o, prd)
    mg.set_option(MG.OPT_RESTRICT_PLANES, planes_per_thread)
    mg.setup_matrices_planes(planes)
    mg.setup_fields(x, r)
    mg.restriction(1)
    orc = O.OracleMG(n, n, o, prd, mode=O.PLANE)
    orc.set_ops(0, planes)
    orc.rat()
    orc.set_fields(0, x, r)
    orc.restrict(1)
    xc, rc = mg.fields(1)
    assert relerr(rc, orc.fields(1)[1]) < 1e-13
    assert not xc.any()


@pytest.mark.parametrize("cap", [1, 88])
def test_gpu_vcycle_chunked_matches_oracle(cap):
    # a whole V-cycle with every chunked kernel marching several planes
    n = 64
    prd = (1, 1, 0)
    o, planes = operator("shear", n, prd)
    N = n * n * o
    r = O.det_field(N, 77)
    mg = MG.Multigrid(n, n, o, prd)
    mg.set_option(MG.OPT_CHUNK_CAP, cap)
    mg.set_option(MG.OPT_RESTRICT_PLANES, 3)
    mg.setup_matrices_planes(planes)
    mg.setup_fields(np.zeros((N, 3)), r)
    orc = O.OracleMG(n, n, o, prd, mode=O.PLANE)
    orc.set_ops(0, planes)
    orc.rat()
    orc.params(smoother=1)
    orc.set_fields(0, np.zeros((N, 3)), r)
    mg.v_cycle(2)
    orc.vcycle(2)
    assert relerr(mg.x(0), orc.fields(0)[0]) < 1e-11


def test_gpu_c1_solution_checked_by_per_point_residual():
    # config 1 at full size (128 x 128 x 129, uniform RHS seed 1) solved with
    # the tiled, z-chunked kernels; the residual of the result is then taken
    # with the one-thread-per-point kernel, which shares no code with the
    # tiled residual that stopped the solve
    n = 128
    h = 1.0 / n
    st, fix = MG.shear_stencil(1.0, 1.5, h, h, h, 1.0)
    mg = MG.Multigrid(n, n, n + 1, (1, 1, 0))
    mg.setup_matrices_const(st, fix)
    N = n * n * (n + 1)
    rng = np.random.default_rng(1)
    b = rng.uniform(-1.0, 1.0, size=(N, 3))
    b -= b.mean(axis=0)
    b[: n * n] = 0
    b[-n * n:] = 0
    for mode in (MG.SOLVE_RELRES, MG.SOLVE_FP32_GMRES_IR):
        mg.set_option(MG.OPT_TILED_RESIDUAL, 1)
        mg.setup_fields(np.zeros((N, 3)), b)
        s = mg.solve(mode=mode, rtol=1e-10)
        assert s["res"] <= 1e-10 * s["res0"], mode
        mg.set_option(MG.OPT_TILED_RESIDUAL, 0)
        res = mg.residual(0)
        assert np.linalg.norm(res) <= 1.01e-10 * np.linalg.norm(b), mode


def test_gpu_time_kernel_leaves_level_and_reports_launches():
    # stzmg_time_kernel (no reference counterpart): the timings are positive,
    # the per-launch duration of the fused sweep (kinds 4, 5) is at most
    # about half a sweep, the level's buffers stay consistent (a sweep after
    # it still equals the oracle's), and a level without the fused sweep is
    # refused for kinds 4, 5
    n = 64
    o, planes = operator("trans", n, (1, 1, 1))
    N = n * n * o
    x, r = O.det_field(N, 81), O.det_field(N, 82)
    mg = MG.Multigrid(n, n, o, (1, 1, 1))
    mg.setup_matrices_planes(planes)
    mg.setup_fields(x, r)
    t = {k: mg.time_kernel(k, 0, 5) for k in range(6)}
    assert all(np.isfinite(v) and v > 0 for v in t.values()), t
    # (ms; the event pair around a launch also holds its dispatch, a few
    # microseconds that weigh on launches of this size)
    assert t[5] <= 0.75 * t[2] + 0.006 and t[4] <= 0.75 * t[0] + 0.006, t
    mg.setup_fields(x, r)
    mg.gauss_seidel(0, 1)
    orc = oracle_for(n, o, (1, 1, 1), planes, x, r)
    orc.gs(0, 1)
    assert relerr(mg.x(0), orc.fields(0)[0]) < 1e-12
    # the coarsest levels have no fused sweep (at most 1024 points)
    last = mg.nlevels - 1
    with pytest.raises(MG.MultigridError):
        mg.time_kernel(5, last, 2)
