"""Host-side mirror of the reference ``multigrid<vec3,M>`` interface.

``Multigrid`` keeps the names and argument meaning of
mg_3d_template/multigrid.hh:19-113 (v_down / v_up, nlevels, setup_matrices,
setup_fields, v_cycle, gauss_seidel, restriction, interpolation, l2_error)
and runs every operation in the HIP library through the C ABI of
include/stzmg.h.  Errors raise ``MultigridError`` carrying the library's
message instead of aborting the process (the reference calls MPI_Abort,
mg_3d_template/common.cc:52-55).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

PLANE, POINT = 0, 1
NORM_QUARTIC, NORM_L2SQ = 0, 1
SOLVE_REF, SOLVE_RELRES, SOLVE_FP32_IR, SOLVE_FP32_GMRES, SOLVE_FP32_GMRES_IR, SOLVE_GMRES64 = 0, 1, 2, 3, 4, 5
OPT_COARSE_KERNEL, OPT_FUSED_SMOOTHER, OPT_TILED_RESIDUAL, OPT_BOTTOM_MAP = 0, 1, 2, 3
OPT_CHUNK_CAP, OPT_RESTRICT_PLANES, OPT_CENTRO_GALERKIN = 4, 5, 6


class MultigridError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"stzmg error {code}: {msg}")
        self.code = code


class Multigrid:
    """Geometric multigrid for vec3 unknowns with 27-point 3x3-block stencils.

    Parameters mirror the reference's top-level region: (m, n, o) vertices and
    periodicity flags; coarsening stops once m*n*o <= coarsest_total (1024 in
    the reference, multigrid.cc:157)."""

    def __init__(self, m, n, o, prd=(1, 1, 0), op_mode=None, coarsest_total=1024,
                 v_down=2, v_up=2, bottom_sweeps=24, sweep_growth=1):
        self.L = _lib.load()
        d = _lib.StzmgDesc()
        self.L.stzmg_desc_default(ctypes.byref(d), m, n, o, int(prd[0]), int(prd[1]), int(prd[2]))
        d.coarsest_total = coarsest_total
        d.v_down, d.v_up = v_down, v_up
        d.bottom_sweeps, d.sweep_growth = bottom_sweeps, sweep_growth
        if op_mode is not None:
            d.op_mode = op_mode
        self.desc = d
        h = ctypes.c_void_p()
        rc = self.L.stzmg_create(ctypes.byref(d), ctypes.byref(h))
        if rc != 0:
            raise MultigridError(rc, self.L.stzmg_last_error(None).decode())
        self.h = h
        self.op_mode = d.op_mode

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.L.stzmg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            raise MultigridError(rc, self.L.stzmg_last_error(self.h).decode())

    # -- geometry -----------------------------------------------------------
    @property
    def nlevels(self):
        return self.L.stzmg_nlevels(self.h)

    mlevels = nlevels

    def dims(self, level=0):
        d = np.zeros(3, np.int32)
        self._chk(self.L.stzmg_level_dims(self.h, level, _lib.iptr(d)))
        return tuple(int(v) for v in d)

    def npts(self, level=0):
        m, n, o = self.dims(level)
        return m * n * o

    # -- operator (setup_matrices, multigrid.hh:195-200) ----------------------
    def setup_matrices_const(self, stencil, fix=None):
        s = np.ascontiguousarray(stencil, dtype=np.float64).reshape(243)
        f = None if fix is None else np.ascontiguousarray(fix, dtype=np.float64).reshape(9)
        self._chk(self.L.stzmg_set_operator_const(self.h, _lib.dptr(s), _lib.dptr(f)))

    def setup_matrices_planes(self, planes):
        p = np.ascontiguousarray(planes, dtype=np.float64).reshape(-1)
        if p.size != self.dims(0)[2] * 243:
            raise ValueError("planes must hold o x 27 x 3 x 3 doubles")
        self._chk(self.L.stzmg_set_operator_planes(self.h, _lib.dptr(p)))

    def setup_matrices_points(self, points):
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1)
        if p.size != self.npts(0) * 243:
            raise ValueError("points must hold m*n*o x 27 x 3 x 3 doubles")
        self._chk(self.L.stzmg_set_operator_points(self.h, _lib.dptr(p)))

    def level_operator(self, level):
        n = ctypes.c_long()
        self._chk(self.L.stzmg_get_level_operator(self.h, level, None, ctypes.byref(n)))
        out = np.zeros((n.value, 27, 3, 3))
        self._chk(self.L.stzmg_get_level_operator(self.h, level, _lib.dptr(out), None))
        return out

    # -- fields (setup_fields, region.hh:209-241) -------------------------------
    def setup_fields(self, x=None, r=None, level=0):
        N = self.npts(level)
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float64).reshape(N, 3)
        ra = None if r is None else np.ascontiguousarray(r, dtype=np.float64).reshape(N, 3)
        self._chk(self.L.stzmg_set_level_fields(self.h, level, _lib.dptr(xa), _lib.dptr(ra)))

    def fields(self, level=0):
        N = self.npts(level)
        x = np.zeros((N, 3))
        r = np.zeros((N, 3))
        self._chk(self.L.stzmg_get_level_fields(self.h, level, _lib.dptr(x), _lib.dptr(r)))
        return x, r

    def x(self, level=0):
        return self.fields(level)[0]

    def set_fields_device(self, x_ptr=None, r_ptr=None):
        """Level-0 fields from device memory (SoA, 3 planes of m*n*o doubles)."""
        self._chk(self.L.stzmg_set_fields_device(self.h, x_ptr, r_ptr))

    def get_x_device(self, x_ptr):
        self._chk(self.L.stzmg_get_x_device(self.h, x_ptr))

    # -- level operations (multigrid.cc:412-458) -------------------------------
    def gauss_seidel(self, level=0, nsweeps=1):
        self._chk(self.L.stzmg_gauss_seidel(self.h, level, nsweeps))

    def gauss_seidel_fp32(self, level=0, nsweeps=1):
        """nsweeps sweeps of the FP32 smoother of the refinement's inner solve
        on the level's fields rounded to FP32; x is widened back to FP64."""
        self._chk(self.L.stzmg_gauss_seidel_fp32(self.h, level, nsweeps))

    def residual_fp32(self, level=0):
        """r - A x in FP32 on the level's fields rounded to FP32."""
        out = np.zeros((self.npts(level), 3))
        self._chk(self.L.stzmg_residual_fp32(self.h, level, _lib.dptr(out)))
        return out

    def restriction(self, level):
        self._chk(self.L.stzmg_restriction(self.h, level))

    def interpolation(self, level):
        self._chk(self.L.stzmg_interpolation(self.h, level))

    def residual(self, level=0):
        out = np.zeros((self.npts(level), 3))
        self._chk(self.L.stzmg_residual(self.h, level, _lib.dptr(out)))
        return out

    def norm(self, level=0, kind=NORM_L2SQ):
        v = ctypes.c_double()
        self._chk(self.L.stzmg_norm(self.h, level, kind, ctypes.byref(v)))
        return v.value

    def l2_error_all(self, level=0):
        """sum over points of |r|^4: the reference's stopping quantity
        (region.cc:3031-3057 with mg3d_mod_sq(vec3) = |r|^4)."""
        return self.norm(level, NORM_QUARTIC)

    l2_error = l2_error_all

    def v_cycle(self, n=1):
        self._chk(self.L.stzmg_vcycle(self.h, n))

    def v_cycle_fp32(self, n=1):
        self._chk(self.L.stzmg_vcycle_fp32(self.h, n))

    # -- solve -------------------------------------------------------------------
    def solve(self, mode=SOLVE_RELRES, rtol=1e-10, tol=0.0, max_cycles=1000, inner_cycles=None,
              extra_cycles=2, ref_norm_schedule=0, use_graph=True, restart=10, inner_rtol=0.0):
        """inner_cycles: FP32 V-cycles per refinement step (FP32_IR, default 3)
        or per preconditioner application (FP32_GMRES and FP32_GMRES_IR,
        default 1).  inner_rtol: reduction of each FP32 GMRES solve of
        FP32_GMRES_IR (0: the library's 1e-4)."""
        if inner_cycles is None:
            inner_cycles = 1 if mode in (SOLVE_FP32_GMRES, SOLVE_FP32_GMRES_IR, SOLVE_GMRES64) else 3
        o = _lib.StzmgSolveOpts()
        self.L.stzmg_solve_opts_default(ctypes.byref(o))
        o.restart, o.inner_rtol = restart, inner_rtol
        o.mode, o.rtol, o.tol = mode, rtol, tol
        o.max_cycles, o.inner_cycles, o.extra_cycles = max_cycles, inner_cycles, extra_cycles
        o.ref_norm_schedule, o.use_graph = ref_norm_schedule, int(bool(use_graph))
        st = _lib.StzmgSolveStats()
        self._chk(self.L.stzmg_solve(self.h, ctypes.byref(o), ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _lib.StzmgSolveStats._fields_}

    def time_kernel(self, kind, level=0, reps=10):
        v = ctypes.c_double()
        self._chk(self.L.stzmg_time_kernel(self.h, kind, level, reps, ctypes.byref(v)))
        return v.value

    def set_option(self, option, value):
        self._chk(self.L.stzmg_set_option(self.h, option, int(value)))

    @property
    def stream(self):
        return self.L.stzmg_stream(self.h)


def shear_stencil(mu, lamb, dx, dy, dz, dt, kappa=0.0):
    """(stencil[27,3,3], fix[3,3]) of shear_sim_3d (src/shear_sim_3d.cpp:216-331,
    src/shear_sim_3d.hh:45-55), built by the library."""
    L = _lib.load()
    B = np.zeros((27, 3, 3))
    F = np.zeros((3, 3))
    rc = L.stzmg_shear_stencil(mu, lamb, dx, dy, dz, dt, kappa, _lib.dptr(B), _lib.dptr(F))
    if rc:
        raise MultigridError(rc, "shear stencil")
    return B, F


def trans_stencil(T, mu, lamb, dx, dy, dz, dt):
    """Transformed-coordinate stencil of trans_sim_3d (src/trans_sim_3d.cpp:1187-1489)
    scaled by -dt, built by the library."""
    L = _lib.load()
    Ta = np.ascontiguousarray(T, dtype=np.float64).reshape(9)
    B = np.zeros((27, 3, 3))
    rc = L.stzmg_trans_stencil(_lib.dptr(Ta), mu, lamb, dx, dy, dz, dt, _lib.dptr(B))
    if rc:
        raise MultigridError(rc, "trans stencil")
    return B
