"""CPU oracle for the STZ++ hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  It holds:

* ``liboracle.so`` (mg_oracle.c): a plain-C FP64 restatement of the reference
  multigrid (mg_3d_template/region.cc, multigrid.cc) -- residual, lexicographic
  and 8-colour block Gauss-Seidel, restriction, interpolation, Galerkin RAT,
  V-cycle, the sum |r|^4 stopping quantity;
* numpy restatements of the operator builders: the shear stencil
  (src/shear_sim_3d.cpp:216-331, src/shear_sim_3d.hh:45-55) and the
  transformed-coordinate stencil (src/trans_sim_3d.cpp:1187-1489 with
  src/stenc_funcs.hh);
* ``RefMG`` / ``RefShear``: ctypes views of oracle/_ref/libstzref.so, the
  reference sources compiled in place (built only where the reference tree
  exists).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        L.mgo_new.restype = ctypes.c_void_p
        L.mgo_new.argtypes = [ctypes.c_int] * 3 + [_ip, ctypes.c_int, ctypes.c_long]
        vp = ctypes.c_void_p
        for name, args, res in [
            ("mgo_delete", [vp], None),
            ("mgo_nlev", [vp], ctypes.c_int),
            ("mgo_dims", [vp, ctypes.c_int, _ip], None),
            ("mgo_set_level_ops", [vp, ctypes.c_int, _dp], None),
            ("mgo_get_level_ops", [vp, ctypes.c_int, _dp], None),
            ("mgo_set_fields", [vp, ctypes.c_int, _dp, _dp], None),
            ("mgo_get_fields", [vp, ctypes.c_int, _dp, _dp], None),
            ("mgo_params", [vp] + [ctypes.c_int] * 5, None),
            ("mgo_set_ranks", [vp] + [ctypes.c_int] * 3, None),
            ("mgo_set_threads", [ctypes.c_int], None),
            ("mgo_get_threads", [], ctypes.c_int),
            ("mgo_do_rat", [vp], None),
            ("mgo_do_residual", [vp, ctypes.c_int, _dp], None),
            ("mgo_do_apply", [vp, ctypes.c_int, _dp, _dp], None),
            ("mgo_do_gs", [vp, ctypes.c_int, ctypes.c_int], None),
            ("mgo_do_restrict", [vp, ctypes.c_int], None),
            ("mgo_do_interp", [vp, ctypes.c_int], None),
            ("mgo_do_vcycle", [vp, ctypes.c_int], None),
            ("mgo_do_norm_quartic", [vp, ctypes.c_int], ctypes.c_double),
            ("mgo_do_norm2", [vp, ctypes.c_int], ctypes.c_double),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


PLANE, POINT = 0, 1


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops (0 leaves the default)."""
    lib().mgo_set_threads(int(n))


def get_threads():
    return lib().mgo_get_threads()


def processor_division(procs, m, n, o, prd):
    """Subdomain grid of the reference's calc_processor_division
    (src/common.cpp:7-60): among i*j*k = procs splits that divide m, n, o
    evenly, the first (k outermost, then j) with the smallest cut surface
    m*n*(k or k-1) + m*o*(j or j-1) + n*o*(i or i-1), the -1 where that
    dimension is not periodic.  None when no split divides evenly."""
    best, dims = m * n * o, None
    for k in range(1, procs + 1):
        if procs % k or o % k:
            continue
        p2 = procs // k
        for j in range(1, p2 + 1):
            if p2 % j or n % j:
                continue
            i = p2 // j
            if m % i:
                continue
            t = m * n * (k if prd[2] else k - 1) + m * o * (j if prd[1] else j - 1) + n * o * (i if prd[0] else i - 1)
            if t < best:
                best, dims = t, (i, j, k)
    return dims


class OracleMG:
    """Multigrid hierarchy restated from the reference (FP64, CPU)."""

    def __init__(self, m, n, o, prd, mode=PLANE, coarsest_total=1024):
        self.L = lib()
        p = np.ascontiguousarray(prd, dtype=np.int32)
        self.h = self.L.mgo_new(m, n, o, _i(p), mode, coarsest_total)
        if not self.h:
            raise ValueError("oracle hierarchy construction failed")
        self.mode = mode

    def close(self):
        if self.h:
            self.L.mgo_delete(self.h)
            self.h = None

    __del__ = close

    @property
    def nlevels(self):
        return self.L.mgo_nlev(self.h)

    def dims(self, l):
        d = np.zeros(3, np.int32)
        self.L.mgo_dims(self.h, l, _i(d))
        return tuple(int(v) for v in d)

    def npts(self, l):
        m, n, o = self.dims(l)
        return m * n * o

    def ncls(self, l):
        m, n, o = self.dims(l)
        return o if self.mode == PLANE else m * n * o

    def set_ops(self, l, ops):
        a = np.ascontiguousarray(ops, dtype=np.float64).reshape(self.ncls(l), 243)
        self.L.mgo_set_level_ops(self.h, l, _d(a))

    def ops(self, l):
        a = np.zeros((self.ncls(l), 27, 3, 3))
        self.L.mgo_get_level_ops(self.h, l, _d(a))
        return a

    def set_fields(self, l, x=None, r=None):
        N = self.npts(l)
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float64).reshape(N, 3)
        ra = None if r is None else np.ascontiguousarray(r, dtype=np.float64).reshape(N, 3)
        self.L.mgo_set_fields(self.h, l, _d(xa), _d(ra))

    def fields(self, l):
        N = self.npts(l)
        x = np.zeros((N, 3))
        r = np.zeros((N, 3))
        self.L.mgo_get_fields(self.h, l, _d(x), _d(r))
        return x, r

    def params(self, v_down=2, v_up=2, bottom=24, growth=1, smoother=0):
        self.L.mgo_params(self.h, v_down, v_up, bottom, growth, smoother)

    def set_ranks(self, ranks):
        """Subdomain grid of the lexicographic sweep: the reference's sweep on
        ranks[0] x ranks[1] x ranks[2] MPI ranks (subdomains swept in parallel
        on OpenMP threads)."""
        self.L.mgo_set_ranks(self.h, *(int(r) for r in ranks))

    def rat(self):
        self.L.mgo_do_rat(self.h)

    def residual(self, l=0):
        out = np.zeros((self.npts(l), 3))
        self.L.mgo_do_residual(self.h, l, _d(out))
        return out

    def apply(self, l, x):
        xa = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(xa)
        self.L.mgo_do_apply(self.h, l, _d(xa), _d(out))
        return out

    def gs(self, l=0, n=1):
        self.L.mgo_do_gs(self.h, l, n)

    def restrict(self, l):
        self.L.mgo_do_restrict(self.h, l)

    def interp(self, l):
        self.L.mgo_do_interp(self.h, l)

    def vcycle(self, n=1):
        self.L.mgo_do_vcycle(self.h, n)

    def norm_quartic(self, l=0):
        return self.L.mgo_do_norm_quartic(self.h, l)

    def norm2(self, l=0):
        return self.L.mgo_do_norm2(self.h, l)


# ---------------------------------------------------------------------------
# Operator builders (numpy restatements).

def block_index(di, dj, dk):
    """Stencil slot of offset (di, dj, dk): 13 + di + 3 dj + 9 dk
    (src/shear_sim_3d.cpp:222)."""
    return 13 + di + 3 * dj + 9 * dk


def shear_stencil_base(mu, lamb, dx, dy, dz, kappa=0.0, dt=1.0):
    """stencil_base of shear_sim_3d (src/shear_sim_3d.cpp:216-331): the 27
    3x3 blocks of the staggered elastic operator, plus the kappa/dt viscous
    terms (zero in QS runs).  The centre block keeps the reference's a33
    parenthesisation, in which kappa/dt*(...) is added rather than subtracted
    (src/shear_sim_3d.cpp:276-278); it is inert for kappa = 0."""
    S = np.zeros((27, 3, 3))
    L2 = lamb + 2 * mu
    kd = kappa / dt
    ax = np.array([L2, mu, mu]) / dx**2 + kd / dx**2
    ay = np.array([mu, L2, mu]) / dy**2 + kd / dy**2
    az = np.array([mu, mu, L2]) / dz**2 + kd / dz**2
    for s in (-1, 1):
        S[block_index(s, 0, 0)] = np.diag(ax)
        S[block_index(0, s, 0)] = np.diag(ay)
        S[block_index(0, 0, s)] = np.diag(az)
    c = (lamb + mu) / 4
    # mixed x-y blocks: +c/(dx dy) at (+,+) and (-,-), -c at (+,-), (-,+)
    for si in (-1, 1):
        for sj in (-1, 1):
            v = si * sj * c / (dx * dy)
            B = np.zeros((3, 3)); B[0, 1] = B[1, 0] = v
            S[block_index(si, sj, 0)] = B
            v = si * sj * c / (dx * dz)
            B = np.zeros((3, 3)); B[0, 2] = B[2, 0] = v
            S[block_index(si, 0, sj)] = B
            v = si * sj * c / (dy * dz)
            B = np.zeros((3, 3)); B[1, 2] = B[2, 1] = v
            S[block_index(0, si, sj)] = B
    ksum = kd * (1 / dx**2 + 1 / dy**2 + 1 / dz**2)
    S[13] = np.diag([
        -2 * (L2 / dx**2 + mu / dy**2 + mu / dz**2 + ksum),
        -2 * (mu / dx**2 + L2 / dy**2 + mu / dz**2 + ksum),
        -2 * (mu / dx**2 + mu / dy**2 + L2 / dz**2) + ksum,
    ])
    return S


def shear_operator(mu, lamb, dx, dy, dz, dt, kappa=0.0):
    """(stencil, fix) of set_up_stencil_and_matrices (src/shear_sim_3d.hh:45-55):
    stencil = -dt * base; fix = -dt*2*((lamb+2mu)/dx^2 + mu/dy^2 + mu/dz^2) I."""
    st = -dt * shear_stencil_base(mu, lamb, dx, dy, dz, kappa, dt)
    fix = -dt * 2 * ((lamb + 2 * mu) / dx**2 + mu / dy**2 + mu / dz**2) * np.eye(3)
    return st, fix


def trans_stencil_blocks(T, mu, lamb, dx, dy, dz):
    """The 19 non-zero blocks of src/stenc_funcs.hh as one closed form.

    With S the symmetric matrix built from the upper triangle of T (the
    generated functions never read t21, t31, t32) and W = adj(S)/det(T), the
    coefficient of d_a d_b U_m in row i of the transformed elastic operator is
        K[i,m,a,b] = mu S_im sum_j W_aj W_bj + (lamb+mu) W_ai delta_bm
    (symmetrised over a<->b).  Axis blocks take K/h_a^2, edge blocks
    +-K/(4 h_a h_b), the centre block -2 sum_a K[.,.,a,a]/h_a^2."""
    T = np.asarray(T, dtype=np.float64).reshape(3, 3)
    S = np.triu(T) + np.triu(T, 1).T
    D = np.linalg.det(T)
    adj = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            minor = np.delete(np.delete(S, b, 0), a, 1)
            adj[a, b] = (-1) ** (a + b) * np.linalg.det(minor)
    W = adj / D
    h = np.array([dx, dy, dz])
    WW = W @ W.T
    K = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for m in range(3):
            for a in range(3):
                for b in range(3):
                    K[i, m, a, b] = mu * S[i, m] * WW[a, b] + (lamb + mu) * W[a, i] * (b == m)
    out = np.zeros((27, 3, 3))
    for a in range(3):
        off = [0, 0, 0]
        off[a] = 1
        blk = K[:, :, a, a] / h[a] ** 2
        out[block_index(*off)] += blk
        off[a] = -1
        out[block_index(*off)] += blk
        out[13] -= 2 * blk
    for a in range(3):
        for b in range(a + 1, 3):
            Ks = (K[:, :, a, b] + K[:, :, b, a]) / (4 * h[a] * h[b])
            for sa in (-1, 1):
                for sb in (-1, 1):
                    off = [0, 0, 0]
                    off[a] = sa
                    off[b] = sb
                    out[block_index(*off)] += sa * sb * Ks
    return out


def point_operator(dims, prd, stencil, fix=None):
    """Per-point boxes S[6] and 27-block stencils for a problem whose interior
    points carry ``stencil`` and whose points on a non-periodic face carry the
    one-block ``fix`` with box [0,1)^3 (src/shear_sim_3d.hh:388-402)."""
    m, n, o = dims
    N = m * n * o
    S = np.zeros((N, 6), np.int32)
    B = np.zeros((N, 27, 3, 3))
    k, j, i = np.meshgrid(np.arange(o), np.arange(n), np.arange(m), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    bnd = np.zeros(N, bool)
    if not prd[0]:
        bnd |= (i == 0) | (i == m - 1)
    if not prd[1]:
        bnd |= (j == 0) | (j == n - 1)
    if not prd[2]:
        bnd |= (k == 0) | (k == o - 1)
    S[:] = [-1, 2, -1, 2, -1, 2]
    S[bnd] = [0, 1, 0, 1, 0, 1]
    B[~bnd] = stencil
    if fix is not None:
        B[bnd, 13] = fix
    return S, B


def plane_operator(o, prd_z, stencil, fix):
    """Per-plane table for an operator with Dirichlet planes k=0, o-1 (z walled)."""
    ops = np.repeat(np.asarray(stencil)[None], o, axis=0).copy()
    if not prd_z:
        for k in (0, o - 1):
            ops[k] = 0
            ops[k, 13] = fix
    return ops


# ---------------------------------------------------------------------------
# Reference library (compiled from the reference tree where it exists).

def ref_lib():
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libstzref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        vp = ctypes.c_void_p
        L.ref_mg_create.restype = vp
        L.ref_mg_create.argtypes = [ctypes.c_int] * 3 + [_ip, _ip, _dp]
        for name, args, res in [
            ("ref_mg_destroy", [vp], None),
            ("ref_mg_nlevels", [vp], ctypes.c_int),
            ("ref_mg_dims", [vp, ctypes.c_int, _ip], None),
            ("ref_mg_set_sweeps", [vp, ctypes.c_int, ctypes.c_int], None),
            ("ref_mg_get_operator", [vp, ctypes.c_int, _dp], ctypes.c_int),
            ("ref_mg_get_boxes", [vp, ctypes.c_int, _ip], None),
            ("ref_mg_set_fields", [vp, ctypes.c_int, _dp, _dp], None),
            ("ref_mg_get_fields", [vp, ctypes.c_int, _dp, _dp], None),
            ("ref_mg_residual", [vp, ctypes.c_int, _dp], None),
            ("ref_mg_gs", [vp, ctypes.c_int, ctypes.c_int], None),
            ("ref_mg_restrict", [vp, ctypes.c_int], None),
            ("ref_mg_interp", [vp, ctypes.c_int], None),
            ("ref_mg_vcycle", [vp, ctypes.c_int], None),
            ("ref_mg_l2_error_all", [vp, ctypes.c_int], ctypes.c_double),
            ("ref_trans_stencil", [_dp] + [ctypes.c_double] * 5 + [_dp], None),
            ("ref_shear_create", [ctypes.c_int] * 3 + [_dp, _dp, _ip, ctypes.c_char_p], vp),
            ("ref_shear_setup", [vp, ctypes.c_double], None),
            ("ref_shear_get_stencil", [vp, _dp, _dp], None),
            ("ref_shear_set_state", [vp, ctypes.c_double, _dp, _dp, _dp], None),
            ("ref_shear_step", [vp, ctypes.c_double, ctypes.c_int], None),
            ("ref_shear_get_state", [vp, _dp, _dp, _dp, _dp, _dp], None),
            ("ref_shear_vcycles", [vp], ctypes.c_int),
            ("ref_shear_destroy", [vp], None),
            ("ref_trans_create", [ctypes.c_int] * 3 + [_dp, _dp, ctypes.c_double, _ip, ctypes.c_char_p], vp),
            ("ref_trans_setup", [vp, ctypes.c_double], None),
            ("ref_trans_get_stencil", [vp, _dp, _dp], None),
            ("ref_trans_matrices", [vp, _dp, _dp, _dp], None),
            ("ref_trans_set_state", [vp, ctypes.c_double, _dp, _dp, _dp], None),
            ("ref_trans_step", [vp, ctypes.c_double, ctypes.c_int], None),
            ("ref_trans_get_state", [vp, _dp, _dp, _dp, _dp, _dp], None),
            ("ref_trans_vcycles", [vp], ctypes.c_int),
            ("ref_trans_destroy", [vp], None),
            ("ref_grf", [ctypes.c_int] * 3 + [_ip, _dp, ctypes.c_int, ctypes.c_char_p, _dp, ctypes.c_long, _dp,
                         _dp], ctypes.c_int),
            ("ref_shear_output_cube", [vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int], None),
            ("ref_shear_create_direct", [ctypes.c_int] * 3 + [_dp, _dp, ctypes.c_double, ctypes.c_double, _ip,
                                                              ctypes.c_char_p], vp),
            ("ref_shear_init_ref", [vp], None),
            ("ref_shear_step_direct", [vp, ctypes.c_double, ctypes.c_int], None),
            ("ref_shear_get_ref", [vp, _dp], None),
            ("ref_shear_net_force", [vp], None),
            ("ref_shear_l2_comparison", [vp, vp, _dp], None),
            ("ref_trans_output_cube", [vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int], None),
            ("ref_trans_net_force", [vp], None),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _REF = L
    return _REF


class RefMG:
    """The reference multigrid<vec3,mat3> on one rank, fed per-point stencils."""

    def __init__(self, dims, prd, S, B):
        self.L = ref_lib()
        if self.L is None:
            raise RuntimeError("reference library not built")
        self._S = np.ascontiguousarray(S, dtype=np.int32)
        self._B = np.ascontiguousarray(B, dtype=np.float64)
        p = np.ascontiguousarray(prd, dtype=np.int32)
        self.h = self.L.ref_mg_create(dims[0], dims[1], dims[2], _i(p), _i(self._S), _d(self._B))

    def close(self):
        if self.h:
            self.L.ref_mg_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def nlevels(self):
        return self.L.ref_mg_nlevels(self.h)

    def dims(self, l):
        d = np.zeros(3, np.int32)
        self.L.ref_mg_dims(self.h, l, _i(d))
        return tuple(int(v) for v in d)

    def npts(self, l):
        m, n, o = self.dims(l)
        return m * n * o

    def operator(self, l):
        out = np.zeros((self.npts(l), 27, 3, 3))
        if self.L.ref_mg_get_operator(self.h, l, _d(out)) != 0:
            raise ValueError("reference box wider than 3x3x3")
        return out

    def set_fields(self, l, x=None, r=None):
        N = self.npts(l)
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float64).reshape(N, 3)
        ra = None if r is None else np.ascontiguousarray(r, dtype=np.float64).reshape(N, 3)
        self.L.ref_mg_set_fields(self.h, l, _d(xa), _d(ra))

    def fields(self, l):
        N = self.npts(l)
        x = np.zeros((N, 3))
        r = np.zeros((N, 3))
        self.L.ref_mg_get_fields(self.h, l, _d(x), _d(r))
        return x, r

    def residual(self, l=0):
        out = np.zeros((self.npts(l), 3))
        self.L.ref_mg_residual(self.h, l, _d(out))
        return out

    def gs(self, l=0, n=1):
        self.L.ref_mg_gs(self.h, l, n)

    def restrict(self, l):
        self.L.ref_mg_restrict(self.h, l)

    def interp(self, l):
        self.L.ref_mg_interp(self.h, l)

    def vcycle(self, n=1):
        self.L.ref_mg_vcycle(self.h, n)

    def l2_error_all(self, l=0):
        return self.L.ref_mg_l2_error_all(self.h, l)


def ref_trans_stencil(T, mu, lamb, dx, dy, dz):
    L = ref_lib()
    Ta = np.ascontiguousarray(T, dtype=np.float64).reshape(9)
    out = np.zeros((27, 3, 3))
    L.ref_trans_stencil(_d(Ta), mu, lamb, dx, dy, dz, _d(out))
    return out


def det_field(N, salt, comps=3):
    """Deterministic pseudo-random field in [-1, 1) built from integer
    arithmetic only, so a fixture can be regenerated bit for bit anywhere."""
    p = np.arange(N, dtype=np.int64)[:, None]
    c = np.arange(comps, dtype=np.int64)[None, :]
    v = (p * 7919 + c * 104729 + salt * 15485863) % 2000003
    v = (v * 48271) % 2000003
    return v.astype(np.float64) / 1000001.5 - 1.0


# ---------------------------------------------------------------------------
# Quasi-static step oracle (stz_oracle.c) and the reference step harness.

def _step_lib():
    L = lib()
    if not hasattr(L, "_sto_bound"):
        vp = ctypes.c_void_p
        L.sto_new.restype = vp
        L.sto_new.argtypes = [ctypes.c_int] * 3 + [_ip, _dp, _dp]
        L.sto_delete.argtypes = [vp]
        L.sto_set_state.argtypes = [vp, ctypes.c_double, _dp, _dp, _dp]
        L.sto_get_state.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
        L.sto_step.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.sto_vcycles.argtypes = [vp]
        L.sto_vcycles.restype = ctypes.c_int
        L.sto_set_solver.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.sto_set_ranks.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.sto_set_transform.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.sto_set_eta.argtypes = [vp, ctypes.c_double]
        L.sto_proj_mag.argtypes = [vp]
        L.sto_proj_mag.restype = ctypes.c_double
        L.sto_set_direct.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.sto_init_ref.argtypes = [vp]
        L.sto_step_direct.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.sto_get_ref.argtypes = [vp, _dp]
        L.sto_set_textbook.argtypes = [vp, ctypes.c_int]
        L.sto_trans_stencil.argtypes = [_dp] + [ctypes.c_double] * 5 + [_dp]
        L.sto_transform_at.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       _dp, _dp, _dp]
        L._sto_bound = True
    return L


TRANSFORMS = {"none": 0, "pure_shear": 1, "simple_shear": 2, "homogeneous": 3, "identity": 4}


def transform_at(family, Ut, bz, t):
    """(T, T^-1, dT/dt) of a transform family (src/trans_sim_3d.hh:57-125)."""
    L = _step_lib()
    T, Ti, dT = np.zeros(9), np.zeros(9), np.zeros(9)
    L.sto_transform_at(TRANSFORMS[family], Ut, bz, t, _d(T), _d(Ti), _d(dT))
    return T.reshape(3, 3), Ti.reshape(3, 3), dT.reshape(3, 3)


def trans_stencil_c(T, mu, lamb, dx, dy, dz):
    """The C oracle's closed form of src/stenc_funcs.hh (27 x 3 x 3)."""
    L = _step_lib()
    Ta = np.ascontiguousarray(T, dtype=np.float64).reshape(9)
    out = np.zeros(243)
    L.sto_trans_stencil(_d(Ta), mu, lamb, dx, dy, dz, _d(out))
    return out.reshape(27, 3, 3)


STEP_PARAM_ORDER = ("mu", "lamb", "sy", "tau0", "c0", "eps0", "delta", "omega", "bath_temp",
                    "chi_inf", "ez", "diff_l", "U")


def _nzv(N, prd):
    return N[2] + (0 if prd[2] else 1)


class OracleShear:
    """shear_sim_3d QS step restated in C (padded ghost grid, in-place order)."""

    def __init__(self, N, params, box=(-1, 1, -1, 1, -0.5, 0.5), prd=(1, 1, 0)):
        self.L = _step_lib()
        self.N, self.prd = tuple(N), tuple(prd)
        p = np.array([params[k] for k in STEP_PARAM_ORDER], dtype=np.float64)
        b = np.array(box, dtype=np.float64)
        pr = np.array(prd, dtype=np.int32)
        self.h = self.L.sto_new(N[0], N[1], N[2], _i(pr), _d(b), _d(p))

    def close(self):
        if self.h:
            self.L.sto_delete(self.h)
            self.h = None

    __del__ = close

    def set_state(self, t, u, sig, chi):
        u, sig, chi = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, sig, chi))
        self.L.sto_set_state(self.h, t, _d(u), _d(sig), _d(chi))

    def state(self):
        Nx, Ny, Nz = self.N
        nv = Nx * Ny * _nzv(self.N, self.prd)
        nc = Nx * Ny * Nz
        t = np.zeros(1)
        u, sig, chi, dpl = np.zeros((nv, 3)), np.zeros((nc, 6)), np.zeros(nc), np.zeros(nc)
        self.L.sto_get_state(self.h, _d(t), _d(u), _d(sig), _d(chi), _d(dpl))
        return dict(t=float(t[0]), u=u, sig=sig, chi=chi, dpl=dpl)

    def step(self, dt, n=1):
        self.L.sto_step(self.h, dt, n)

    def set_solver(self, smoother=0, rtol=0.0):
        """smoother 0 = lexicographic (reference), 1 = 8-colour; rtol 0 = the
        reference stopping rule, else a relative-residual solve."""
        self.L.sto_set_solver(self.h, smoother, rtol)

    def set_ranks(self, ranks):
        """The projection's lexicographic sweep as the reference's on
        ranks[0] x ranks[1] x ranks[2] MPI ranks (subdomains on OpenMP
        threads)."""
        self.L.sto_set_ranks(self.h, *(int(r) for r in ranks))

    def set_transform(self, family, Ut):
        """Run the transformed-coordinate step of trans_sim_3d with transform
        family ``family`` (TRANSFORMS) at rate Ut."""
        self.L.sto_set_transform(self.h, TRANSFORMS[family] if isinstance(family, str) else family, Ut)

    @property
    def proj_mag(self):
        """Sum of sym_mat3::modsq of the last projection's stress increments."""
        return self.L.sto_proj_mag(self.h)

    # direct (explicit dynamics) step, src/shear_sim_3d.cpp:166-213, :774-1030
    def set_direct(self, rho, kappa):
        self.L.sto_set_direct(self.h, float(rho), float(kappa))

    def init_ref(self):
        """init_ref: the reference map at the grid positions."""
        self.L.sto_init_ref(self.h)

    def step_direct(self, dt, n=1):
        self.L.sto_step_direct(self.h, float(dt), int(n))

    def ref_map(self):
        Nx, Ny, Nz = self.N
        xi = np.zeros((Nx * Ny * _nzv(self.N, self.prd), 3))
        self.L.sto_get_ref(self.h, _d(xi))
        return xi

    def set_eta(self, eta):
        """Override the adaptive-step tolerance eta of the plastic update."""
        self.L.sto_set_eta(self.h, eta)

    @property
    def vcycles(self):
        return self.L.sto_vcycles(self.h)


class RefShear:
    """The reference shear_sim_3d<sym_mat3> driven through oracle/_ref."""

    def __init__(self, N, params, outdir, box=(-1, 1, -1, 1, -0.5, 0.5), prd=(1, 1, 0)):
        self.L = ref_lib()
        if self.L is None:
            raise RuntimeError("reference library not built")
        self.N, self.prd = tuple(N), tuple(prd)
        p = np.array([params[k] for k in ("mu", "lamb", "sy", "tau0", "c0", "eps0", "delta", "omega",
                                          "bath_temp", "chi_inf", "ez")] + [21000.0, params["diff_l"],
                                                                               params["U"]], dtype=np.float64)
        self._p = p
        b = np.array([box[0], box[1], box[2], box[3], box[4], box[5]], dtype=np.float64)
        pr = np.array(prd, dtype=np.int32)
        self.h = self.L.ref_shear_create(N[0], N[1], N[2], _d(b), _d(p), _i(pr), outdir.encode())

    def close(self):
        if self.h:
            self.L.ref_shear_destroy(self.h)
            self.h = None

    __del__ = close

    @classmethod
    def direct(cls, N, params, outdir, rho, kappa, box=(-1, 1, -1, 1, -0.5, 0.5), prd=(1, 1, 0)):
        """A reference shear_sim_3d for the direct step (density rho,
        viscous damping kappa)."""
        self = cls.__new__(cls)
        self.L = ref_lib()
        if self.L is None:
            raise RuntimeError("reference library not built")
        self.N, self.prd = tuple(N), tuple(prd)
        p = np.array([params[k] for k in ("mu", "lamb", "sy", "tau0", "c0", "eps0", "delta", "omega",
                                          "bath_temp", "chi_inf", "ez")] + [21000.0, params["diff_l"],
                                                                               params["U"]], dtype=np.float64)
        b = np.array(box, dtype=np.float64)
        pr = np.array(prd, dtype=np.int32)
        self.h = self.L.ref_shear_create_direct(N[0], N[1], N[2], _d(b), _d(p), float(rho), float(kappa), _i(pr),
                                                outdir.encode())
        return self

    def init_ref(self):
        self.L.ref_shear_init_ref(self.h)

    def step_direct(self, dt, n=1):
        """n calls of the reference's step_forward(dt)."""
        self.L.ref_shear_step_direct(self.h, float(dt), int(n))

    def ref_map(self):
        Nx, Ny, Nz = self.N
        xi = np.zeros((Nx * Ny * _nzv(self.N, self.prd), 3))
        self.L.ref_shear_get_ref(self.h, _d(xi))
        return xi

    def output_cube(self, prefix, mode, sn):
        """The reference's output_cube into the simulation's directory."""
        self.L.ref_shear_output_cube(self.h, prefix.encode(), int(mode), int(sn))

    def net_force(self):
        """The reference's compute_net_force (appends to traction.dat)."""
        self.L.ref_shear_net_force(self.h)

    def l2_comparison(self, other):
        """The reference's l2_comparison of this simulation against other
        (three squared, normalised differences)."""
        out = np.zeros(3)
        self.L.ref_shear_l2_comparison(self.h, other.h, _d(out))
        return out

    def setup(self, dt):
        self.L.ref_shear_setup(self.h, dt)

    def stencil(self):
        b, f = np.zeros((27, 6)), np.zeros(6)
        self.L.ref_shear_get_stencil(self.h, _d(b), _d(f))
        return b, f

    def set_state(self, t, u, sig, chi):
        u, sig, chi = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, sig, chi))
        self.L.ref_shear_set_state(self.h, t, _d(u), _d(sig), _d(chi))

    def step(self, dt, n=1):
        self.L.ref_shear_step(self.h, dt, n)

    def state(self):
        Nx, Ny, Nz = self.N
        nv = Nx * Ny * _nzv(self.N, self.prd)
        nc = Nx * Ny * Nz
        t = np.zeros(1)
        u, sig, chi, dpl = np.zeros((nv, 3)), np.zeros((nc, 6)), np.zeros(nc), np.zeros(nc)
        self.L.ref_shear_get_state(self.h, _d(t), _d(u), _d(sig), _d(chi), _d(dpl))
        return dict(t=float(t[0]), u=u, sig=sig, chi=chi, dpl=dpl)

    @property
    def vcycles(self):
        return self.L.ref_shear_vcycles(self.h)


class RefTrans:
    """The reference trans_sim_3d (pure-shear transform at rate Ut) driven
    through oracle/_ref."""

    def __init__(self, N, params, Ut, outdir, box=(-1, 1, -1, 1, -0.5, 0.5), prd=(1, 1, 0)):
        self.L = ref_lib()
        if self.L is None:
            raise RuntimeError("reference library not built")
        self.N, self.prd = tuple(N), tuple(prd)
        p = np.array([params[k] for k in ("mu", "lamb", "sy", "tau0", "c0", "eps0", "delta", "omega",
                                          "bath_temp", "chi_inf", "ez")] + [21000.0, params["diff_l"],
                                                                               params["U"]], dtype=np.float64)
        b = np.array(box, dtype=np.float64)
        pr = np.array(prd, dtype=np.int32)
        self.h = self.L.ref_trans_create(N[0], N[1], N[2], _d(b), _d(p), Ut, _i(pr), outdir.encode())

    def close(self):
        if self.h:
            self.L.ref_trans_destroy(self.h)
            self.h = None

    __del__ = close

    def output_cube(self, prefix, mode, sn):
        self.L.ref_trans_output_cube(self.h, prefix.encode(), int(mode), int(sn))

    def net_force(self):
        """compute_net_force at the current time (appends to traction.dat)."""
        self.L.ref_trans_net_force(self.h)

    def setup(self, dt):
        self.L.ref_trans_setup(self.h, dt)

    def stencil(self):
        b, f = np.zeros((27, 9)), np.zeros(9)
        self.L.ref_trans_get_stencil(self.h, _d(b), _d(f))
        return b.reshape(27, 3, 3), f.reshape(3, 3)

    def matrices(self):
        T, Ti, dT = np.zeros(9), np.zeros(9), np.zeros(9)
        self.L.ref_trans_matrices(self.h, _d(T), _d(Ti), _d(dT))
        return T.reshape(3, 3), Ti.reshape(3, 3), dT.reshape(3, 3)

    def set_state(self, t, u, sig, chi):
        u, sig, chi = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, sig, chi))
        self.L.ref_trans_set_state(self.h, t, _d(u), _d(sig), _d(chi))

    def step(self, dt, n=1):
        self.L.ref_trans_step(self.h, dt, n)

    def state(self):
        Nx, Ny, Nz = self.N
        nv = Nx * Ny * _nzv(self.N, self.prd)
        nc = Nx * Ny * Nz
        t = np.zeros(1)
        u, sig, chi, dpl = np.zeros((nv, 3)), np.zeros((nc, 6)), np.zeros(nc), np.zeros(nc)
        self.L.ref_trans_get_state(self.h, _d(t), _d(u), _d(sig), _d(chi), _d(dpl))
        return dict(t=float(t[0]), u=u, sig=sig, chi=chi, dpl=dpl)

    @property
    def vcycles(self):
        return self.L.ref_trans_vcycles(self.h)


# ---------------------------------------------------------------------------
# Random initial chi (grf_oracle.c): the reference's cube-cutoff Gaussian
# convolution (src/shear_sim_3d.cpp:2793-3078) and this library's own
# Philox4x64-10 noise.

_u64p = ctypes.POINTER(ctypes.c_uint64)


def _grf_lib():
    L = lib()
    if not hasattr(L, "_grf_bound"):
        L.grf_philox4x64.argtypes = [_u64p, _u64p, _u64p]
        L.grf_noise.argtypes = [ctypes.c_uint64, ctypes.c_long, _dp]
        L.grf_normal_factor.argtypes = [ctypes.c_int, ctypes.c_double]
        L.grf_normal_factor.restype = ctypes.c_double
        L.grf_convolve.argtypes = [_ip, _ip, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   _dp, _dp]
        L.grf_convolve_points.argtypes = [_ip, _ip, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, _dp, ctypes.c_long,
                                          ctypes.POINTER(ctypes.c_long), _dp]
        L._grf_bound = True
    return L


def philox4x64(ctr, key):
    """One Philox4x64-10 block (4 uint64) for counter ctr (4) and key (2)."""
    c = np.ascontiguousarray(ctr, dtype=np.uint64)
    k = np.ascontiguousarray(key, dtype=np.uint64)
    o = np.zeros(4, dtype=np.uint64)
    _grf_lib().grf_philox4x64(c.ctypes.data_as(_u64p), k.ctypes.data_as(_u64p), o.ctypes.data_as(_u64p))
    return o


def grf_cut(gauss_fac, conv_l):
    """Cutoff in grid points: the reference's int member initialised from
    gauss_fac * ll (src/shear_sim_3d.cpp:57), where gauss_fac itself is an int
    member (src/shear_sim_3d.hh:140): a fractional factor is truncated first."""
    return int(int(gauss_fac) * conv_l)


def grf_noise_shape(N, prd, cut):
    """(nx, ny, nz) of the noise grid: a non-periodic axis is padded by cut on
    each side (src/shear_sim_3d.cpp:2802-2805)."""
    return tuple(int(n) if p else int(n) + 2 * cut for n, p in zip(N, prd))


def grf_noise(seed, n):
    out = np.zeros(int(n))
    _grf_lib().grf_noise(int(seed), int(n), _d(out))
    return out


def grf_normal_factor(cut, llinv):
    return _grf_lib().grf_normal_factor(int(cut), float(llinv))


def grf_convolve(N, prd, cut, llinv, chi_avg, chi_1, noise):
    """chi on N (x fastest) by the reference's brute-force convolution."""
    N3 = np.array(N, dtype=np.int32)
    P3 = np.array([1 if p else 0 for p in prd], dtype=np.int32)
    noise = np.ascontiguousarray(noise, dtype=np.float64).ravel()
    assert noise.size == int(np.prod(grf_noise_shape(N, prd, cut)))
    out = np.zeros(int(np.prod(N)))
    _grf_lib().grf_convolve(_i(N3), _i(P3), int(cut), float(llinv), float(chi_avg), float(chi_1), _d(noise),
                            _d(out))
    return out


def grf_convolve_points(N, prd, cut, llinv, chi_avg, chi_1, noise, idx):
    """grf_convolve at the cells idx (flat, x fastest) only."""
    N3 = np.array(N, dtype=np.int32)
    P3 = np.array([1 if p else 0 for p in prd], dtype=np.int32)
    noise = np.ascontiguousarray(noise, dtype=np.float64).ravel()
    assert noise.size == int(np.prod(grf_noise_shape(N, prd, cut)))
    ix = np.ascontiguousarray(idx, dtype=np.int64)
    assert ix.min() >= 0 and ix.max() < int(np.prod(N))
    out = np.zeros(ix.size)
    _grf_lib().grf_convolve_points(_i(N3), _i(P3), int(cut), float(llinv), float(chi_avg), float(chi_1),
                                   _d(noise), ix.size, ix.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), _d(out))
    return out


def grf_field(N, prd, gauss_fac, conv_l, chi_avg, sigma, seed=0, noise=None):
    """The whole initialiser: noise (drawn from seed unless given), chi_1 =
    sigma / gauss_normal_factor_long, brute-force convolution."""
    cut = grf_cut(gauss_fac, conv_l)
    llinv = 1.0 / conv_l
    if noise is None:
        noise = grf_noise(seed, int(np.prod(grf_noise_shape(N, prd, cut))))
    chi_1 = sigma / grf_normal_factor(cut, llinv)
    return grf_convolve(N, prd, cut, llinv, chi_avg, chi_1, noise)


def grf_cache_name(N, prd, sim_case, TZ, chi_avg, chi_1, gnf):
    """Name of the reference's convolution cache (src/shear_sim_3d.cpp:
    2812-2816); chi_1 is the normalised scale, gnf gauss_normal_factor_long."""
    human_mu = int(TZ * chi_avg)
    human_sig = int(TZ * chi_1 * gnf + 1)
    pref = "f" if (sim_case == 19 and not prd[0] and not prd[1]) else ("t" if prd[2] else "")
    return f"{pref}cg_{N[0]}_{N[1]}_{N[2]}_mu{human_mu}_sig{human_sig}.dat"


def ref_grf(N, prd, gauss_fac, conv_l, chi_avg, sigma, TZ, sim_case, workdir):
    """The reference's own noise grid and set_random_initial_conditions (run
    in workdir, where it writes its cache) through oracle/_ref."""
    L = ref_lib()
    if L is None:
        raise RuntimeError("reference library not built")
    cut = grf_cut(gauss_fac, conv_l)
    n = int(np.prod(grf_noise_shape(N, prd, cut)))
    noise, chi, out3 = np.zeros(n), np.zeros(int(np.prod(N))), np.zeros(3)
    pr = np.array([1 if p else 0 for p in prd], dtype=np.int32)
    p = np.array([gauss_fac, conv_l, chi_avg, sigma, TZ], dtype=np.float64)
    rc = L.ref_grf(N[0], N[1], N[2], _i(pr), _d(p), int(sim_case), workdir.encode(), _d(noise), n, _d(chi),
                   _d(out3))
    if rc:
        raise RuntimeError(f"ref_grf failed ({rc})")
    return dict(noise=noise, chi=chi, chi_1=float(out3[0]), gnf=float(out3[1]), cut=int(out3[2]),
                files=sorted(os.listdir(workdir)))


# ---------------------------------------------------------------------------
# Output of the reference (numpy restatements): output_cube
# (src/shear_sim_3d.cpp:1963-2161 with Field::fval, src/field.hh:144-178) and
# compute_net_force (src/shear_sim_3d.cpp:1061-1126, src/trans_sim_3d.cpp:
# 1575-1660), in the reference's arithmetic order.

CUBE_NAMES = ["u", "v", "w", "s11", "s12", "s13", "s22", "s23", "s33", "p", "dev", "chi", "tem"]


def cube_values(mode, N, prd, u, sig, chi):
    """The float32 cube of write_files mode `mode` (0-12): x slowest, then y,
    z fastest; velocities carry Nz + 1 z points (the top plane; the periodic
    image of plane 0 when z is periodic)."""
    Nx, Ny, Nz = N
    if mode <= 2:
        nzv = Nz + (0 if prd[2] else 1)
        f = np.asarray(u)[:, mode].reshape(nzv, Ny, Nx)
        if prd[2]:
            f = np.concatenate([f, f[:1]], axis=0)
    else:
        s = np.asarray(sig).reshape(Nz, Ny, Nx, 6)
        s11, s12, s13, s22, s23, s33 = (s[..., q] for q in range(6))
        if 3 <= mode <= 8:
            f = s[..., mode - 3]
        elif mode in (9, 10):
            p = -(s11 + s22 + s33) * (1 / 3.0)
            if mode == 9:
                f = p
            else:
                t11, t22, t33 = s11 + p, s22 + p, s33 + p
                f = np.sqrt(0.5 * (t11 * t11 + t22 * t22 + t33 * t33) + s12 * s12 + s13 * s13 + s23 * s23)
        elif mode == 11:
            f = np.asarray(chi).reshape(Nz, Ny, Nx)
        elif mode == 12:
            f = np.asarray(chi).reshape(Nz, Ny, Nx) * 21000
        else:
            raise ValueError("mode")
    return np.ascontiguousarray(np.transpose(f, (2, 1, 0))).astype(np.float32)


def net_force(N, box, sig, T=None):
    """(ts13, ts23, ts33, bs13, bs23, bs33): the wall tractions of
    compute_net_force.  The bottom 'extrapolation' pairs each cell with its
    +x neighbour (the reference's grido + 1); with T, the physical stress
    T s T^t of the extrapolated transformed stress (the trans override)."""
    Nx, Ny, Nz = N
    dx, dy = (box[1] - box[0]) / Nx, (box[3] - box[2]) / Ny
    s = np.asarray(sig).reshape(Nz, Ny, Nx, 6)

    def comp(a, b):
        v = [1.5 * a[q] - 0.5 * b[q] for q in range(6)]
        if T is None:
            return v[2], v[4], v[5]
        M = np.array([[v[0], v[1], v[2]], [v[1], v[3], v[4]], [v[2], v[4], v[5]]])
        P = (np.asarray(T) @ M) @ np.asarray(T).T
        return P[0, 2], P[1, 2], P[2, 2]

    b = [0.0, 0.0, 0.0]
    t = [0.0, 0.0, 0.0]
    for yy in range(Ny):
        for xx in range(Nx):
            c = comp(s[0, yy, xx], s[0, yy, (xx + 1) % Nx])
            for q in range(3):
                b[q] += c[q]
    for yy in range(Ny):
        for xx in range(Nx):
            c = comp(s[Nz - 1, yy, xx], s[Nz - 2, yy, xx])
            for q in range(3):
                t[q] += c[q]
    f = -dx * dy
    b = [v * f for v in b]
    g = dx * dy
    t = [v * g for v in t]
    return t + b


def l2_comparison(N, prd, box, U, chi_inf, a, b):
    """l2_comparison (src/shear_sim_3d.cpp:3122-3158): the squared,
    normalised differences (velocity, stress, chi) of states a and b (dicts
    with u, sig, chi) on the same grid.  Velocity by the trapezoid rule in z:
    vertex plane 0 and plane Nz (the top plate, or with z periodic the
    periodic image of plane 0) weigh 1/2 (:3134, :3142-3149); stress (the six
    components) and chi over every cell.  Each sum is scaled by
    dx dy dz / volume, the velocity's divided by U^2 and chi's by chi_inf^2
    (:3152-3155).  The reference starts its three accumulators from one
    indeterminate value (local_l2[2] is read before it is set, :3127); here
    they start from zero."""
    Nx, Ny, Nz = N
    nxy = Nx * Ny
    du = (np.asarray(a["u"], dtype=np.float64) - np.asarray(b["u"], dtype=np.float64)).reshape(-1, nxy, 3)
    sq = (du * du).sum(axis=2)
    w = np.ones(Nz)
    w[0] = 0.5
    v = float((sq[:Nz].sum(axis=1) * w).sum()) + 0.5 * float((sq[0] if prd[2] else sq[Nz]).sum())
    ds = np.asarray(a["sig"], dtype=np.float64) - np.asarray(b["sig"], dtype=np.float64)
    dc = np.asarray(a["chi"], dtype=np.float64) - np.asarray(b["chi"], dtype=np.float64)
    Lx, Ly, Lz = box[1] - box[0], box[3] - box[2], box[5] - box[4]
    nv = (Lx / Nx) * (Ly / Ny) * (Lz / Nz) / (Lx * Ly * Lz)
    return [v * nv / (U * U), float((ds * ds).sum()) * nv, float((dc * dc).sum()) * nv / (chi_inf * chi_inf)]


def traction_line(t, vals, trans=False):
    """One line of traction.dat as the reference prints it."""
    if trans:
        return "\t".join("%g" % v for v in vals) + "\n"
    return "\t".join("%.16f" % v for v in [t] + list(vals)) + "\n"
