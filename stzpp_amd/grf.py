"""Random initial chi field on the GPU (include/stz.h, stz_grf_*).

Host-side mirror of the reference's random initial condition
(shear_sim_3d::set_random_initial_conditions, src/shear_sim_3d.cpp:
2793-3078): chi = chi_avg + chi_1 (noise * G) with the cube-cutoff kernel
G = exp(-d^2 / conv_l), |i|, |j|, |k| <= cut, cut = int(int(gauss_fac) *
conv_l), chi_1 = sigma / sqrt(sum G^2).  The convolution runs on the GPU as
three 1-D passes; the noise is the caller's (e.g. the reference's own grid)
or drawn on the GPU from a seed.  Arrays are flat, x fastest, like the
simulation state.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class GrfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


def default_params():
    """The shear_sim driver's values (src/shear_sim.cpp:8, :83-86) as a dict."""
    L = _lib.load()
    g = _lib.StzGrfParams()
    L.stz_grf_params_default(ctypes.byref(g))
    return {k: getattr(g, k) for k, _ in _lib.StzGrfParams._fields_}


def _params(kw):
    L = _lib.load()
    g = _lib.StzGrfParams()
    L.stz_grf_params_default(ctypes.byref(g))
    for k, v in kw.items():
        if v is None:
            continue
        if k in ("sim_case", "seed"):
            setattr(g, k, int(v))
        else:
            setattr(g, k, float(v))
    return g


def _geom(N, prd):
    return np.array(N, dtype=np.int32), np.array([1 if p else 0 for p in prd], dtype=np.int32)


def _chk(rc):
    if rc:
        raise GrfError(rc, _lib.load().stz_grf_last_error().decode())


def geometry(N, prd=(1, 1, 0), **params):
    """(cut, noise-grid size, chi_1) of a configuration."""
    L = _lib.load()
    n3, p3 = _geom(N, prd)
    cut, nn, c1 = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_double()
    _chk(L.stz_grf_geometry(_lib.iptr(n3), _lib.iptr(p3), ctypes.byref(_params(params)), ctypes.byref(cut),
                            ctypes.byref(nn), ctypes.byref(c1)))
    return cut.value, nn.value, c1.value


def noise_shape(N, prd=(1, 1, 0), **params):
    """(nx, ny, nz) of the noise grid (padded by cut on non-periodic axes)."""
    cut = geometry(N, prd, **params)[0]
    return tuple(int(n) if p else int(n) + 2 * cut for n, p in zip(N, prd))


def cache_name(N, prd=(1, 1, 0), **params):
    """File name of the reference's convolution cache."""
    L = _lib.load()
    n3, p3 = _geom(N, prd)
    buf = ctypes.create_string_buffer(256)
    _chk(L.stz_grf_cache_name(_lib.iptr(n3), _lib.iptr(p3), ctypes.byref(_params(params)), buf, 256))
    return buf.value.decode()


def grf_field(N, prd=(1, 1, 0), noise=None, **params):
    """chi on the N cells (flat, x fastest).  params: gauss_fac, conv_l,
    chi_avg, sigma, TZ, sim_case, seed (see include/stz.h)."""
    L = _lib.load()
    n3, p3 = _geom(N, prd)
    g = _params(params)
    nz = None
    if noise is not None:
        nz = np.ascontiguousarray(noise, dtype=np.float64).ravel()
        if nz.size != geometry(N, prd, **params)[1]:
            raise ValueError("noise grid has the wrong size")
    chi = np.zeros(int(np.prod(N)))
    _chk(L.stz_grf_field(_lib.iptr(n3), _lib.iptr(p3), ctypes.byref(g), _lib.dptr(nz), _lib.dptr(chi)))
    return chi


def time_kernels(N, prd=(1, 1, 0), reps=20, **params):
    """Average device ms per launch of (noise, x pass, y pass, z pass)."""
    L = _lib.load()
    n3, p3 = _geom(N, prd)
    ms = np.zeros(4)
    _chk(L.stz_grf_time_kernels(_lib.iptr(n3), _lib.iptr(p3), ctypes.byref(_params(params)), int(reps),
                                _lib.dptr(ms)))
    return ms


__all__ = ["GrfError", "time_kernels", "default_params", "geometry", "noise_shape", "cache_name", "grf_field"]
