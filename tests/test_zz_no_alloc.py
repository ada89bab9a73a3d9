"""stzmg_norm_noalloc (the residual norm without device allocation) against
stzmg_norm through the C ABI: the same kernels and summation order write to
the library's fixed device array instead of the handle's partials buffer,
so the values must be identical, call after call, also when a later call
covers fewer points (its partials are a prefix of the array the earlier
call filled) and from another handle on the same GPU."""
import ctypes

import numpy as np
import pytest

import oracle as O
from stzpp_amd import _lib
from stzpp_amd import multigrid as MG

pytestmark = pytest.mark.gpu


def norms(mg, level, fn):
    out = []
    for kind in (MG.NORM_QUARTIC, MG.NORM_L2SQ):
        v = ctypes.c_double()
        rc = fn(mg.h, level, kind, ctypes.byref(v))
        assert rc == 0, mg.L.stzmg_last_error(mg.h).decode()
        out.append(v.value)
    return out


def handle(n, prd, seed, tiled=1):
    h = 1.0 / n
    o = n + (0 if prd[2] else 1)
    st, fix = MG.shear_stencil(1.0, 1.5, h, h, h, 1.0)
    mg = MG.Multigrid(n, n, o, prd)
    mg.set_option(MG.OPT_TILED_RESIDUAL, tiled)
    mg.setup_matrices_const(st, fix if not prd[2] else None)
    N = n * n * o
    mg.setup_fields(O.det_field(N, seed), O.det_field(N, seed + 1))
    for l in range(1, mg.nlevels):
        M = mg.npts(l)
        mg.setup_fields(O.det_field(M, seed + 10 * l), O.det_field(M, seed + 10 * l + 1), level=l)
    return mg


@pytest.mark.parametrize("tiled", [1, 0])
def test_gpu_norm_noalloc_matches_norm(tiled):
    L = _lib.load()
    big = handle(128, (1, 1, 0), 81, tiled)
    small = handle(32, (1, 1, 1), 91, tiled)
    # two calls in a row, the second on fewer points (a coarser level, then
    # a smaller handle), each against the allocating entry point
    for mg, level in ((big, 0), (big, 1), (small, 0), (big, 0), (small, 1)):
        a = norms(mg, level, L.stzmg_norm_noalloc)
        b = norms(mg, level, L.stzmg_norm)
        assert a == b, (mg.dims(level), level, a, b)
        assert a[0] > 0 and a[1] > 0


def test_gpu_norm_noalloc_rejects_oversize():
    # the per-point residual of 256 x 256 x 257 points needs 256 x 257
    # partials, more than the 65536 the fixed array holds: rejected, while
    # the allocating entry point still serves it; the same level with the
    # tiled residual (a few hundred partials) is accepted
    L = _lib.load()
    n = 256
    mg = MG.Multigrid(n, n, n + 1, (1, 1, 0))
    st, fix = MG.shear_stencil(1.0, 1.5, 1.0 / n, 1.0 / n, 1.0 / n, 1.0)
    mg.setup_matrices_const(st, fix)
    N = n * n * (n + 1)
    rng = np.random.default_rng(5)
    mg.setup_fields(rng.standard_normal((N, 3)), rng.standard_normal((N, 3)))
    mg.set_option(MG.OPT_TILED_RESIDUAL, 0)
    v = ctypes.c_double()
    assert L.stzmg_norm_noalloc(mg.h, 0, MG.NORM_L2SQ, ctypes.byref(v)) == -2
    assert b"65536" in L.stzmg_last_error(mg.h)
    ref = mg.norm(0, MG.NORM_L2SQ)
    mg.set_option(MG.OPT_TILED_RESIDUAL, 1)
    a = norms(mg, 0, L.stzmg_norm_noalloc)
    b = norms(mg, 0, L.stzmg_norm)
    assert a == b
    assert abs(a[1] / ref - 1) < 1e-12
