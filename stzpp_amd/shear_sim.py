"""Host-side mirror of the reference ``shear_sim_3d<sym_mat3>`` QS API.

``ShearSim`` keeps the reference's names (set_initial_conditions,
step_forward_qs, solve_qs, sim_time, n_vcycles, vcycle_time;
src/shear_sim_3d.hh:35-82) and runs every step on the GPU through the C ABI
of include/stz.h.  Fields stay in HBM between steps; ``state()`` copies them
out.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .multigrid import (MultigridError, SOLVE_FP32_GMRES, SOLVE_FP32_GMRES_IR, SOLVE_FP32_IR, SOLVE_REF,
                        SOLVE_RELRES)


def reference_params(Nx=32, Ny=None, Nz=None, **overrides):
    """Parameters of the reference shear_sim driver after nondimensionalisation
    (src/shear_sim.cpp:18-86), as a dict; keyword arguments override them."""
    Ny = Nx if Ny is None else Ny
    Nz = Nx if Nz is None else Nz
    L = _lib.load()
    p = _lib.StzParams()
    L.stz_params_default(ctypes.byref(p), Nx, Ny, Nz)
    d = {k: getattr(p, k) for k, _ in _lib.StzParams._fields_ if k != "prd"}
    d["prd"] = tuple(p.prd)
    d.update(overrides)
    return d


def config0_params(N=32):
    """Config 0: 32^3 cells, mu = 1, K = 2.2, s_y = 1, eps0 = 10, chi_inf = 0.15,
    c0 = 0.4, plates +-0.01; the remaining STZ constants keep the driver's values."""
    return reference_params(N, N, N, mu=1.0, lamb=2.2 - 2.0 / 3.0, sy=1.0, eps0=10.0,
                            chi_inf=0.15, c0=0.4, U=0.01)


class StepError(MultigridError):
    pass


# transform families of trans_sim_3d (src/trans_sim_3d.hh:57-125); include/stz.h STZ_T_*
TRANSFORMS = {"none": 0, "pure_shear": 1, "simple_shear": 2, "homogeneous": 3, "identity": 4, "user": 5}


def transform(family, Ut, t, bz=0.5):
    """(T, T^-1, dT/dt) of a transform family at time t, from the library."""
    L = _lib.load()
    T, Ti, dT = np.zeros(9), np.zeros(9), np.zeros(9)
    fam = TRANSFORMS[family] if isinstance(family, str) else int(family)
    rc = L.stz_transform(fam, Ut, bz, t, _lib.dptr(T), _lib.dptr(Ti), _lib.dptr(dT))
    if rc:
        raise StepError(rc, "unknown transform family")
    return T.reshape(3, 3), Ti.reshape(3, 3), dT.reshape(3, 3)


class ShearSim:
    def __init__(self, N_x, N_y, N_z, params=None, **kw):
        self.L = _lib.load()
        prm = dict(reference_params(N_x, N_y, N_z))
        if params:
            prm.update(params)
        prm.update(kw)
        p = _lib.StzParams()
        self.L.stz_params_default(ctypes.byref(p), N_x, N_y, N_z)
        for k, v in prm.items():
            if k == "prd":
                for a in range(3):
                    p.prd[a] = int(v[a])
            elif k in ("Nx", "Ny", "Nz"):
                continue
            elif k == "transform":
                p.transform = TRANSFORMS[v] if isinstance(v, str) else int(v)
            else:
                setattr(p, k, float(v))
        self.params = prm
        self.N = (N_x, N_y, N_z)
        self.prd = tuple(p.prd)
        h = ctypes.c_void_p()
        rc = self.L.stz_create(ctypes.byref(p), ctypes.byref(h))
        if rc:
            raise StepError(rc, self.L.stz_last_error(None).decode())
        self.h = h
        self.n_vcycles = 0
        self.vcycle_time = 0.0  # hours, as the reference accumulates it
        self.last_stats = None

    def close(self):
        if getattr(self, "h", None):
            self.L.stz_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc:
            raise StepError(rc, self.L.stz_last_error(self.h).decode())

    @property
    def nverts(self):
        Nx, Ny, Nz = self.N
        return Nx * Ny * (Nz + (0 if self.prd[2] else 1))

    @property
    def ncells(self):
        return self.N[0] * self.N[1] * self.N[2]

    def set_initial_conditions(self, chi, u=None, sig=None, t=0.0):
        chi = np.ascontiguousarray(chi, dtype=np.float64).reshape(self.ncells)
        u = np.zeros((self.nverts, 3)) if u is None else np.ascontiguousarray(u, dtype=np.float64)
        sig = np.zeros((self.ncells, 6)) if sig is None else np.ascontiguousarray(sig, dtype=np.float64)
        self._chk(self.L.stz_set_state(self.h, t, _lib.dptr(u), _lib.dptr(sig), _lib.dptr(chi)))

    set_state = set_initial_conditions

    def set_random_initial_conditions(self, noise=None, cache_dir=None, **grf_params):
        """chi <- chi_avg + chi_1 (noise * G) on the GPU
        (set_random_initial_conditions, src/shear_sim_3d.cpp:2793-2934); the
        velocity and stress are left as they are.  grf_params: gauss_fac,
        conv_l, chi_avg, sigma, TZ, sim_case, seed (stzpp_amd.grf).  noise:
        the padded noise grid (default: drawn on the GPU from seed).
        cache_dir: load the reference's cg_*.dat cache from there when it
        exists, else write it."""
        from . import grf as G

        g = G._params(grf_params)
        nz = None
        if noise is not None:
            nz = np.ascontiguousarray(noise, dtype=np.float64).ravel()
            if nz.size != G.geometry(self.N, self.prd, **grf_params)[1]:
                raise ValueError("noise grid has the wrong size")
        cd = None if cache_dir is None else str(cache_dir).encode()
        self._chk(self.L.stz_set_random_initial_conditions(self.h, ctypes.byref(g), _lib.dptr(nz), cd))

    def state(self):
        t = np.zeros(1)
        u = np.zeros((self.nverts, 3))
        sig = np.zeros((self.ncells, 6))
        chi = np.zeros(self.ncells)
        dpl = np.zeros(self.ncells)
        self._chk(self.L.stz_get_state(self.h, _lib.dptr(t), _lib.dptr(u), _lib.dptr(sig),
                                       _lib.dptr(chi), _lib.dptr(dpl)))
        return dict(t=float(t[0]), u=u, sig=sig, chi=chi, dpl=dpl)

    @property
    def sim_time(self):
        return self.state_time()

    def state_time(self):
        t = np.zeros(1)
        self._chk(self.L.stz_get_state(self.h, _lib.dptr(t), None, None, None, None))
        return float(t[0])

    def step_forward_qs(self, dt, nsteps=1, solve=None, rtol=1e-12, inner_cycles=None, use_graph=True):
        """nsteps QS steps.  solve=None runs the reference projection loop;
        SOLVE_RELRES / SOLVE_FP32_IR / SOLVE_FP32_GMRES / SOLVE_FP32_GMRES_IR
        solve each projection to rtol (FP32 V-cycles per refinement step or
        per GMRES iteration: inner_cycles, default 3 and 1).  use_graph=False launches the
        V-cycles kernel by kernel."""
        if inner_cycles is None:
            inner_cycles = 1 if solve in (SOLVE_FP32_GMRES, SOLVE_FP32_GMRES_IR) else 3
        st = _lib.StzStepStats()
        if (solve is None or solve == SOLVE_REF) and use_graph:
            opts = None
        elif solve is None or solve == SOLVE_REF:
            o = _lib.StzmgSolveOpts()
            self.L.stzmg_solve_opts_default(ctypes.byref(o))
            o.mode, o.tol, o.use_graph = SOLVE_REF, 0.0, 0
            opts = ctypes.byref(o)
        else:
            o = _lib.StzmgSolveOpts()
            self.L.stzmg_solve_opts_default(ctypes.byref(o))
            o.mode, o.rtol, o.inner_cycles, o.max_cycles = solve, rtol, inner_cycles, 10000
            o.use_graph = 1 if use_graph else 0
            opts = ctypes.byref(o)
        self._chk(self.L.stz_step_qs(self.h, dt, nsteps, opts, ctypes.byref(st)))
        self.n_vcycles += st.vcycles
        self.vcycle_time += st.solve_ms / 1e3 / 3600.0
        self.last_stats = {k: getattr(st, k) for k, _ in _lib.StzStepStats._fields_}
        return self.last_stats

    def adjoint(self, dt, nsteps, checkpoint_every=5, rtol=1e-12):
        """Loss J = mean (ad_Dpl/dt)^2 after nsteps steps from the current state
        and its gradient: returns (J, dJ/dchi0 per cell, dJ/d(s_y, eps0, chi_inf)).
        Every projection, forward and adjoint, is solved to relative residual
        rtol; states are kept every checkpoint_every steps and segments are
        recomputed in the reverse sweep.  The state afterwards is the state
        after the nsteps forward steps.  On a TransSim this differentiates
        the transformed step (the adjoint solves use the transposed operator
        of each step)."""
        o = _lib.StzAdjointOpts()
        self.L.stz_adjoint_opts_default(ctypes.byref(o))
        o.checkpoint_every, o.rtol = int(checkpoint_every), float(rtol)
        J = ctypes.c_double()
        g = np.zeros(self.ncells)
        gp = np.zeros(3)
        self._chk(self.L.stz_adjoint(self.h, dt, nsteps, ctypes.byref(o), ctypes.byref(J), _lib.dptr(g),
                                     _lib.dptr(gp)))
        return J.value, g, gp

    # ---- output in the reference's formats (include/stz.h) -------------

    def cube_size(self, mode):
        n = self.L.stz_cube_size(self.h, int(mode))
        if n < 0:
            raise StepError(-1, f"unknown cube mode {mode}")
        return n

    def output_cube(self, mode):
        """output_cube's values for write_files mode `mode` (0-12): float32, x
        slowest, then y, z fastest (src/shear_sim_3d.cpp:1963-2161)."""
        n = self.L.stz_cube_size(self.h, int(mode))
        if n < 0:
            self._chk(self.L.stz_output_cube(self.h, int(mode), None, 0))
        out = np.zeros(n, dtype=np.float32)
        self._chk(self.L.stz_output_cube(self.h, int(mode), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n))
        return out

    def time_cube_kernel(self, mode, reps=20):
        """Average ms per launch of the cube kernel of `mode` (device only)."""
        v = np.zeros(1)
        self._chk(self.L.stz_output_time_kernel(self.h, int(mode), int(reps), _lib.dptr(v)))
        return float(v[0])

    def write_cube(self, out_dir, mode, frame):
        """Append the cube to <out_dir>/<name>_grid.<frame>."""
        self._chk(self.L.stz_write_cube(self.h, str(out_dir).encode(), int(mode), int(frame)))

    def write_files(self, frame, out_dir, n_tp, fflags=4096):
        """write_files(frame) (src/shear_sim_3d.cpp:1938-1960) into out_dir;
        the reference's build uses fflags = 4096 (tem)."""
        self._chk(self.L.stz_write_files(self.h, str(out_dir).encode(), int(frame), int(n_tp), int(fflags)))

    def compute_net_force(self):
        """Wall tractions (ts13, ts23, ts33, bs13, bs23, bs33) of
        compute_net_force (src/shear_sim_3d.cpp:1061-1126)."""
        out = np.zeros(6)
        self._chk(self.L.stz_net_force(self.h, _lib.dptr(out)))
        return out

    def write_traction(self, out_dir):
        """Append the traction.dat line (and the trac_grid.bin record)."""
        self._chk(self.L.stz_write_traction(self.h, str(out_dir).encode()))

    @property
    def proj_mag(self):
        """Sum of modsq of the last projection's stress increments."""
        return self.L.stz_proj_mag(self.h)

    def set_output(self, out_dir, traction=True, proj_mag=True):
        """Per-step traction.dat / proj_mag.dat in out_dir during
        step_forward_qs, as the reference's solve_qs writes them."""
        flags = (1 if traction else 0) | (2 if proj_mag else 0)
        self._chk(self.L.stz_set_output(self.h, str(out_dir).encode() if flags else None, flags))

    def solve_qs(self, steps, dt, **kw):
        """The QS time loop (solve_qs, src/shear_sim_3d.cpp:336-377) without
        file output: ``steps`` steps of size dt."""
        return self.step_forward_qs(dt, steps, **kw)

    # ---- direct (explicit dynamics) step (include/stz.h) ----------------

    def direct_dt(self):
        """min(tmult dx^2 / (6 kappa), dx / 4) (solve_direct,
        src/shear_sim_3d.cpp:166-172)."""
        return self.L.stz_direct_dt(self.h)

    def init_ref(self):
        """The reference map at the vertex positions (init_ref,
        src/shear_sim_3d.cpp:1036-1057) and the top plate's at sim_time."""
        self._chk(self.L.stz_init_ref(self.h))

    def set_ref_map(self, xi):
        xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(self.nverts, 3)
        self._chk(self.L.stz_set_ref_map(self.h, _lib.dptr(xi)))

    def ref_map(self):
        """The reference map, (vertices, 3), x fastest."""
        xi = np.zeros((self.nverts, 3))
        self._chk(self.L.stz_get_ref_map(self.h, _lib.dptr(xi)))
        return xi

    def step_forward(self, dt, nsteps=1):
        """nsteps direct Euler steps (step_forward, src/shear_sim_3d.cpp:206-213)."""
        st = _lib.StzStepStats()
        self._chk(self.L.stz_step_direct(self.h, float(dt), int(nsteps), ctypes.byref(st)))
        self.last_stats = {k: getattr(st, k) for k, _ in _lib.StzStepStats._fields_}
        return self.last_stats

    def l2_comparison(self, other):
        """Squared, normalised differences (velocity / U^2, stress, chi /
        chi_inf^2) of this simulation against other on the same grid
        (l2_comparison, src/shear_sim_3d.cpp:3122-3158)."""
        out = np.zeros(3)
        self._chk(self.L.stz_l2_comparison(self.h, other.h, _lib.dptr(out)))
        return out

    def time_direct_kernels(self, dt, reps=20):
        """Average ms per launch of the direct step's kernels (plastic, cell,
        vertex, plate) on the current state, which they leave unchanged."""
        ms = np.zeros(4)
        self._chk(self.L.stz_direct_time_kernels(self.h, float(dt), int(reps), _lib.dptr(ms)))
        return ms

    def solve_direct(self, tf, n_tp, out_dir=None, fflags=4096):
        """The direct time loop (solve_direct, src/shear_sim_3d.cpp:166-203)
        from the current state (reference map at the positions when none is
        set) to tf in n_tp frames, each ending at the last step that stays
        below its target time; with out_dir, write_files at frame 0 and after
        every frame.  Returns the step count."""
        dt = self.direct_dt()
        t0 = self.sim_time
        if self.L.stz_cube_size(self.h, 13) < 0:
            self.init_ref()
        if out_dir is not None:
            self.write_files(0, out_dir, n_tp, fflags)
        steps, t = 0, t0
        for kk in range(1, n_tp + 1):
            target = t0 + (tf - t0) / n_tp * kk
            n = 0
            while t + dt * (1 + 1e-8) < target:
                n += 1
                t += dt
            if n:
                self.step_forward(dt, n)
                t = self.sim_time
            steps += n
            if out_dir is not None:
                self.write_files(kk, out_dir, n_tp, fflags)
        return steps


class TransSim(ShearSim):
    """Host-side mirror of ``trans_sim_3d`` (src/trans_sim_3d.hh:13-36): the QS
    step in transformed coordinates X = T(t)^-1 x.  State is the transformed
    velocity U and stress sigma'; the walls are at rest; T(t) is a transform
    family (pure shear, the reference's compiled-in choice, by default) at
    rate Ut.  The projection operator and its Galerkin hierarchy are rebuilt
    on the GPU every step."""

    def __init__(self, N_x, N_y, N_z, Ut, family="pure_shear", params=None, **kw):
        super().__init__(N_x, N_y, N_z, params, transform=family, Ut=Ut, **kw)
        self.family = family
        self.Ut = Ut

    def matrices(self, t=None):
        """T, T^-1, dT/dt at time t (default: the current simulation time)."""
        t = self.state_time() if t is None else t
        if self.family == "user":
            T = self.user_T + (t - self.user_t0) * self.user_dT
            return T, np.linalg.inv(T), self.user_dT.copy()
        return transform(self.family, self.Ut, t, self.params.get("bz", 0.5))

    def set_transform(self, T, dTdt):
        """An arbitrary transform from the current time t0 on:
        T(t) = T + (t - t0) dTdt (stz_set_transform)."""
        Ta = np.ascontiguousarray(T, dtype=np.float64).reshape(9)
        Da = np.ascontiguousarray(dTdt, dtype=np.float64).reshape(9)
        self._chk(self.L.stz_set_transform(self.h, _lib.dptr(Ta), _lib.dptr(Da)))
        self.family = "user"
        self.user_T, self.user_dT, self.user_t0 = Ta.reshape(3, 3).copy(), Da.reshape(3, 3).copy(), self.state_time()


def shear_compare(dsim, qsim, ncomps, time_interval, qdt, ddt=None, out_path=None, time_scale=1.0, **qs_kw):
    """The comparison loop of shear_compare (src/shear_compare.cpp:182-217):
    a direct simulation dsim against a quasi-static one qsim from their
    current states.  At each of ncomps comparison times t0 + tt time_interval
    dsim steps ddt (default: its direct_dt) while it stays below the target,
    then one shortened step onto it; qsim takes int(time_interval / qdt) QS
    steps; then l2_comparison.  With out_path, writes the reference's
    sc<ii>.dat lines "%g %.12g %.12g %.12g" (time x time_scale and the square
    roots), the first for the initial states.  Returns the rows."""
    import math

    ddt = dsim.direct_dt() if ddt is None else ddt
    qs_steps = int(time_interval / qdt)
    t0 = dsim.sim_time
    rows = []

    def record():
        l2 = dsim.l2_comparison(qsim)
        rows.append((dsim.sim_time * time_scale, math.sqrt(l2[0]), math.sqrt(l2[1]), math.sqrt(l2[2])))

    record()
    for tt in range(1, ncomps + 1):
        target = t0 + tt * time_interval
        t, n = dsim.sim_time, 0
        while t + ddt * (1 + 1e-8) < target:
            n += 1
            t += ddt
        if n:
            dsim.step_forward(ddt, n)
        dsim.step_forward(target - dsim.sim_time)
        if qs_steps:
            qsim.step_forward_qs(qdt, qs_steps, **qs_kw)
        record()
    if out_path is not None:
        with open(out_path, "w") as f:
            for r in rows:
                f.write("%g %.12g %.12g %.12g\n" % r)
    return rows


__all__ = ["ShearSim", "TransSim", "transform", "TRANSFORMS", "reference_params", "config0_params",
           "shear_compare", "SOLVE_REF", "SOLVE_RELRES", "SOLVE_FP32_IR", "SOLVE_FP32_GMRES",
           "SOLVE_FP32_GMRES_IR"]
