"""Write golden vectors for the multigrid path from the compiled reference.

Run in a container that holds the reference tree (oracle/_ref/libstzref.so
built by ``make -C oracle/ref``):

    python tests/golden/gen_golden.py

Each case runs in its own process so that the reference's function-static
state (std_iters, src/shear_sim_3d.cpp:561) cannot leak between cases.
Outputs are small ``.npz`` files of inputs and reference outputs.
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

MU, LAMB = 1.0, 1.5


def mg_case(name, dims, prd, kind, seed):
    m, n, o = dims
    h = 1.0 / m
    if kind == "shear":
        st, fix = O.shear_operator(MU, LAMB, h, h, h, 1.0)
    else:
        T = np.array([[1.0, 0.0, 0.5], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
        st = -O.trans_stencil_blocks(T, MU, LAMB, h, h, h)
        fix = None
    S, B = O.point_operator(dims, prd, st, fix)
    ref = O.RefMG(dims, prd, S, B)
    nl = ref.nlevels
    N = m * n * o
    x = O.det_field(N, seed)
    r = O.det_field(N, seed + 1)
    small = N <= 8192
    out = dict(dims=np.array(dims), prd=np.array(prd), stencil=st,
               fix=np.zeros((3, 3)) if fix is None else fix, seed=np.array(seed),
               nlevels=np.array(nl))
    for l in range(nl):
        out[f"dims{l}"] = np.array(ref.dims(l))
    # Coarse operators, compressed to one stencil per z plane after checking
    # that they do not vary in x and y.
    for l in range(1, nl):
        A = ref.operator(l)
        ml, nl_, ol = ref.dims(l)
        A = A.reshape(ol, nl_ * ml, 27, 3, 3)
        assert np.all(A == A[:, :1]), "coarse operator varies in x/y"
        out[f"rat{l}"] = A[:, 0]
    ref.set_fields(0, x, r)
    out["l2_error_all0"] = np.array(ref.l2_error_all(0))
    if small:
        out["residual"] = ref.residual(0)
        ref.gs(0, 1)
        out["gs1"] = ref.fields(0)[0]
        ref.restrict(1)
        out["restrict_r1"] = ref.fields(1)[1]
        xc = O.det_field(ref.npts(1), seed + 100)
        ref.set_fields(1, xc, None)
        ref.interp(1)
        out["interp_x0"] = ref.fields(0)[0]
    # V-cycles from x = 0: the stopping quantity after each cycle.
    ref.set_fields(0, np.zeros((N, 3)), r)
    norms = []
    for _ in range(4):
        ref.vcycle(1)
        norms.append(ref.l2_error_all(0))
    out["vcycle_norms"] = np.array(norms)
    if small:
        out["vcycle_x"] = ref.fields(0)[0]
    ref.close()
    np.savez_compressed(os.path.join(HERE, f"mg_{name}.npz"), **out)


def trans_case():
    rng = np.random.default_rng(7)
    Ts, outs = [], []
    for q in range(6):
        if q == 0:
            T = np.eye(3)
        elif q == 1:
            T = np.array([[1.0, 0.0, 0.5], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])  # simple shear
        elif q == 2:
            e = np.exp(0.3)
            T = np.diag([e, 1.0, 1.0 / e])  # pure shear (trans_sim_3d.hh:76-94)
        else:
            T = np.eye(3) + 0.3 * rng.uniform(-1, 1, (3, 3))
        Ts.append(T)
        outs.append(O.ref_trans_stencil(T, 1.3, 1.7, 0.11, 0.13, 0.17))
    np.savez_compressed(os.path.join(HERE, "trans_stencil.npz"), T=np.array(Ts),
                        blocks=np.array(outs), mu=1.3, lamb=1.7, h=np.array([0.11, 0.13, 0.17]))


def config0_params(Nx):
    """Config-0 parameters (mu = 1, K = 2.2, s_y = 1, eps0 = 10, chi_inf = 0.15,
    c0 = 0.4, U = 0.01); the remaining constants are the reference driver's,
    src/shear_sim.cpp:18-86, for Nx cells across x in [-1, 1)."""
    TZ, kb = 21000.0, 1.3806503e-23
    E, nu, sy = 101e9, 0.35, 0.85e9
    mu_r, rho = E / (2 * (1 + nu)), 6125.0
    ts = 1e-2 / np.sqrt(mu_r / rho)
    return dict(mu=1.0, lamb=2.2 - 2.0 / 3.0, sy=1.0, tau0=1e-13 / ts, c0=0.4, eps0=10.0,
                delta=8000 / TZ, omega=300 * 26.0 / 17.0 * 1e-30 * sy / (TZ * kb), bath_temp=400 / TZ,
                chi_inf=0.15, ez=1.0, diff_l=1.5 * 2.0 / Nx, U=0.01)


def step_case(name):
    """QS steps of the reference shear_sim_3d<sym_mat3>.

    step_plastic: 16x16x8 cells, config-0 material with eps0 = 0.3 and a
      pre-stress s12 above yield, 3 steps of dt = 0.01 (adaptive plasticity,
      chi diffusion and the in-place ad_Dpl ordering all active).
    step_c0: config 0, 32^3 cells, 10 steps of dt = 0.01 from rest with
      chi0 = 0.09 + 0.01 GRF(seed 0, l = 3); stored as per-plane sums and two
      full z planes of every field.
    """
    import tempfile

    from stzpp_amd import fields

    N = (16, 16, 8) if name == "step_plastic" else (32, 32, 32)
    p = config0_params(N[0])
    Nx, Ny, Nz = N
    nc, nv = Nx * Ny * Nz, Nx * Ny * (Nz + 1)
    chi0 = 0.09 + 0.01 * fields.grf((Nz, Ny, Nx), 3.0, 0).ravel()
    sig0 = np.zeros((nc, 6))
    nsteps = 10
    if name == "step_plastic":
        p["eps0"] = 0.3
        sig0[:, 1] = 1.1 + 0.1 * np.cos(np.arange(nc) * 0.37)
        nsteps = 3
    u0 = np.zeros((nv, 3))
    ref = O.RefShear(N, p, tempfile.mkdtemp())
    dt = 0.01
    ref.setup(dt)
    st, fx = ref.stencil()
    ref.set_state(0.0, u0, sig0, chi0)
    cyc = []
    for _ in range(nsteps):
        before = ref.vcycles
        ref.step(dt, 1)
        cyc.append(ref.vcycles - before)
    s = ref.state()
    out = dict(N=np.array(N), dt=np.array(dt), nsteps=np.array(nsteps), chi0=chi0, sig0=sig0,
               vcycles=np.array(cyc), t=np.array(s["t"]), stencil_sym=st, fix_sym=fx,
               **{f"p_{k}": np.array(v) for k, v in p.items()})
    if name == "step_plastic":
        out.update(u=s["u"], sig=s["sig"], chi=s["chi"], dpl=s["dpl"])
    else:
        def planes(a, nz):
            return a.reshape(nz, Nx * Ny, -1)
        for key, nz in (("u", Nz + 1), ("sig", Nz), ("chi", Nz), ("dpl", Nz)):
            a = planes(s[key], nz)
            out[f"{key}_plane_sum"] = a.sum(axis=1)
            out[f"{key}_plane_sq"] = (a * a).sum(axis=1)
            out[f"{key}_k8"] = a[8]
            out[f"{key}_k20"] = a[20]
    ref.close()
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)


def trans_step_case(name):
    """QS steps of the reference trans_sim_3d (pure-shear transform
    T = diag(e^{Ut t}, 1, e^{-Ut t}), Ut = 0.5).

    trans_step_walled / trans_step_periodic: 16x16x8 cells, z walled or
      periodic; config-0 material with eps0 = 0.3, a pre-stress s12 above
      yield and a small s11, a small initial velocity (walls at rest), 3 steps
      of dt = 0.01.  Also stored: T, T^-1, dT/dt and the stencil at the final
      time (set_up_stencil_and_matrices of the last step).
    """
    import tempfile

    from stzpp_amd import fields

    TZ, kb = 21000.0, 1.3806503e-23
    E, nu, sy = 101e9, 0.35, 0.85e9
    mu_r, rho = E / (2 * (1 + nu)), 6125.0
    ts = 1e-2 / np.sqrt(mu_r / rho)
    N = (16, 16, 8)
    prd = (1, 1, 0) if name == "trans_step_walled" else (1, 1, 1)
    p = dict(mu=1.0, lamb=2.2 - 2.0 / 3.0, sy=1.0, tau0=1e-13 / ts, c0=0.4, eps0=0.3,
             delta=8000 / TZ, omega=300 * 26.0 / 17.0 * 1e-30 * sy / (TZ * kb), bath_temp=400 / TZ,
             chi_inf=0.15, ez=1.0, diff_l=1.5 * 2.0 / N[0], U=0.01)
    Ut, dt, nsteps = 0.5, 0.01, 3
    Nx, Ny, Nz = N
    nc = Nx * Ny * Nz
    nv = Nx * Ny * (Nz + (0 if prd[2] else 1))
    chi0 = 0.09 + 0.01 * fields.grf((Nz, Ny, Nx), 3.0, 0).ravel()
    sig0 = np.zeros((nc, 6))
    sig0[:, 1] = 1.1 + 0.1 * np.cos(np.arange(nc) * 0.37)
    sig0[:, 0] = 0.2 * np.sin(np.arange(nc) * 0.11)
    u0 = 0.001 * np.sin(np.arange(nv * 3) * 0.07).reshape(nv, 3)
    if not prd[2]:
        u0[: Nx * Ny] = 0
        u0[-Nx * Ny:] = 0
    ref = O.RefTrans(N, p, Ut, tempfile.mkdtemp(), prd=prd)
    ref.setup(dt)
    ref.set_state(0.0, u0, sig0, chi0)
    cyc = []
    for _ in range(nsteps):
        before = ref.vcycles
        ref.step(dt, 1)
        cyc.append(ref.vcycles - before)
    s = ref.state()
    T, Ti, dT = ref.matrices()
    st, fx = ref.stencil()
    out = dict(N=np.array(N), prd=np.array(prd), dt=np.array(dt), Ut=np.array(Ut), nsteps=np.array(nsteps),
               u0=u0, chi0=chi0, sig0=sig0, vcycles=np.array(cyc), t=np.array(s["t"]),
               u=s["u"], sig=s["sig"], chi=s["chi"], dpl=s["dpl"], T=T, Tinv=Ti, dTdt=dT,
               stencil=st, fix=fx, **{f"p_{k}": np.array(v) for k, v in p.items()})
    ref.close()
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)


GRF_CASES = {
    # name: N, prd, gauss_fac, conv_l, chi_avg, sigma, TZ, sim_case
    "grf_walled": ((16, 12, 10), (1, 1, 0), 2.0, 2.0, 0.09, 0.01, 21000.0, 0),
    "grf_periodic": ((12, 10, 8), (1, 1, 1), 1.5, 3.0, 0.09, 0.01, 21000.0, 0),
    "grf_open": ((10, 8, 6), (0, 0, 0), 1.0, 2.5, 0.1, 0.02, 21000.0, 19),
    "grf_wrap": ((8, 6, 10), (1, 1, 0), 2.0, 3.0, 0.09, 0.01, 21000.0, 0),
    "grf_c0": ((32, 32, 16), (1, 1, 0), 5.0, 3.0, 0.09, 0.01, 21000.0, 0),
}


def grf_case(name):
    """The reference's random initial chi (set_random_initial_conditions,
    src/shear_sim_3d.cpp:2793-2934) on a noise grid from its own generator
    (glibc rand() from the unseeded state), and the cache file it writes.
    grf_wrap has a cutoff equal to the periodic y size (the largest the
    reference's single wrap handles); grf_c0 uses the driver's gauss_fac = 5
    and a 3-cell correlation length (cut 15)."""
    import tempfile

    N, prd, gf, cl, ca, sg, TZ, sc = GRF_CASES[name]
    d = tempfile.mkdtemp()
    r = O.ref_grf(N, prd, gf, cl, ca, sg, TZ, sc, d)
    assert len(r["files"]) == 1, r["files"]
    raw = np.fromfile(os.path.join(d, r["files"][0]), dtype=np.uint8)
    assert np.array_equal(raw.view(np.float64), r["chi"]), "cache file is not the chi field"
    out = dict(N=np.array(N), prd=np.array(prd), gauss_fac=np.array(gf), conv_l=np.array(cl),
               chi_avg=np.array(ca), sigma=np.array(sg), TZ=np.array(TZ), sim_case=np.array(sc),
               noise=r["noise"], chi=r["chi"], chi_1=np.array(r["chi_1"]), gnf=np.array(r["gnf"]),
               cut=np.array(r["cut"]), cache_name=np.array(r["files"][0]))
    if raw.size <= 1 << 16:
        out["cache_bytes"] = raw
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)


CUBE_NAMES = ["u", "v", "w", "s11", "s12", "s13", "s22", "s23", "s33", "p", "dev", "chi", "tem"]


def output_case(name):
    """The reference's output files for a given state.

    output_walled / output_periodic (shear_sim_3d, 16 x 12 x 8 cells, z walled
      or periodic): the cubes of write_files' modes 0-12 at frame 3
      (output_cube), traction.dat and trac_grid.bin from compute_net_force at
      the set state; walled only: then two QS steps of dt = 0.01 each
      followed by compute_net_force (two more traction lines; proj_mag.dat
      from the projections) and the state after the steps.
    output_trans (trans_sim_3d, pure shear Ut = 0.5, 16 x 12 x 8, z walled):
      traction.dat from compute_net_force at t = 0.3 (physical stress).
    """
    import tempfile

    from stzpp_amd import shear_sim as SS

    N = (16, 12, 8)
    prd = (1, 1, 1) if name == "output_periodic" else (1, 1, 0)
    Nx, Ny, Nz = N
    nc = Nx * Ny * Nz
    nv = Nx * Ny * (Nz + (0 if prd[2] else 1))
    p = SS.config0_params(16)
    p["eps0"] = 0.3
    rng = np.random.default_rng({"output_walled": 21, "output_periodic": 22, "output_trans": 23}[name])
    sig0 = 0.3 * rng.standard_normal((nc, 6))
    sig0[:, 1] += 0.9
    u0 = 1e-3 * rng.standard_normal((nv, 3))
    chi0 = 0.09 + 0.01 * rng.standard_normal(nc)
    t0 = 0.3 if name == "output_trans" else 0.25
    d = tempfile.mkdtemp()
    out = dict(N=np.array(N), prd=np.array(prd), t0=np.array(t0), u0=u0, sig0=sig0, chi0=chi0,
               **{f"p_{k}": np.array(v) for k, v in p.items() if k in O.STEP_PARAM_ORDER})
    if name == "output_trans":
        Ut = 0.5
        ref = O.RefTrans(N, {k: p[k] for k in O.STEP_PARAM_ORDER}, Ut, d, prd=prd)
        ref.setup(0.01)
        ref.set_state(t0, u0, sig0, chi0)
        ref.net_force()
        out["Ut"] = np.array(Ut)
        out["traction_txt"] = np.array(open(os.path.join(d, "traction.dat")).read())
    else:
        ref = O.RefShear(N, {k: p[k] for k in O.STEP_PARAM_ORDER}, d, prd=prd)
        ref.setup(0.01)
        ref.set_state(t0, u0, sig0, chi0)
        for mode, nm in enumerate(CUBE_NAMES):
            ref.output_cube(nm, mode, 3)
            out[f"cube_{nm}"] = np.fromfile(os.path.join(d, f"{nm}_grid.3"), dtype=np.uint8)
        ref.net_force()
        if not prd[2]:
            # (the reference's uncapped loop does not end on the singular
            # periodic-z projection, so only the walled case steps)
            for _ in range(2):
                ref.step(0.01, 1)
                ref.net_force()
            s = ref.state()
            out.update(u1=s["u"], sig1=s["sig"], chi1=s["chi"], dpl1=s["dpl"], t1=np.array(s["t"]))
            out["proj_mag_txt"] = np.array(open(os.path.join(d, "proj_mag.dat")).read())
        out["traction_txt"] = np.array(open(os.path.join(d, "traction.dat")).read())
        out["trac_bin"] = np.fromfile(os.path.join(d, "trac_grid.bin"), dtype=np.uint8)
    ref.close()
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)


def direct_case(name):
    """The reference's direct (explicit dynamics) step (step_forward,
    src/shear_sim_3d.cpp:206-213) from a set state with the reference map
    initialised at the grid positions (init_ref): 16 x 12 x 8 cells, density
    1, viscous damping 0.35 (the driver's values), config-0 material with
    eps0 = 0.3, the reference's time step min(tmult dx^2 / (6 kappa), dx / 4)
    with tmult = 0.5, and 5 steps; the state after step 1 and after step 5.
    direct_walled: z walled, a pre-stress above yield and velocity noise;
    direct_periodic: z periodic, milder noise (the reference's own run of
    the walled case's noise level blows up when z is periodic)."""
    import tempfile

    from stzpp_amd import shear_sim as SS

    N = (16, 12, 8)
    prd = (1, 1, 1) if name == "direct_periodic" else (1, 1, 0)
    p = SS.config0_params(16)
    p["eps0"] = 0.3
    P = {k: p[k] for k in O.STEP_PARAM_ORDER}
    rho, kappa, tmult = 1.0, 0.35, 0.5
    dx = 2.0 / N[0]
    dt = min(tmult * (dx * dx / 3.0 / kappa / 2.0), dx / 4.0)
    nc = N[0] * N[1] * N[2]
    nv = N[0] * N[1] * (N[2] + (0 if prd[2] else 1))
    rng = np.random.default_rng(31 if prd[2] else 30)
    amp_s, amp_u = (0.05, 2e-3) if prd[2] else (0.3, 1e-2)
    sig0 = amp_s * rng.standard_normal((nc, 6))
    sig0[:, 1] += 0.9
    u0 = amp_u * rng.standard_normal((nv, 3))
    chi0 = 0.09 + 0.01 * rng.standard_normal(nc)
    t0 = 0.3
    ref = O.RefShear.direct(N, P, tempfile.mkdtemp(), rho, kappa, prd=prd)
    ref.set_state(t0, u0, sig0, chi0)
    ref.init_ref()
    out = dict(N=np.array(N), prd=np.array(prd), t0=np.array(t0), dt=np.array(dt), rho=np.array(rho),
               kappa=np.array(kappa), tmult=np.array(tmult), nsteps=np.array(5), u0=u0, sig0=sig0, chi0=chi0,
               xi0=ref.ref_map(), **{f"p_{k}": np.array(v) for k, v in P.items()})
    for s in range(5):
        ref.step_direct(dt, 1)
        if s in (0, 4):
            st = ref.state()
            tag = "1" if s == 0 else "5"
            out.update({f"u{tag}": st["u"], f"sig{tag}": st["sig"], f"chi{tag}": st["chi"], f"dpl{tag}": st["dpl"],
                        f"xi{tag}": ref.ref_map(), f"t{tag}": np.array(st["t"])})
    assert all(np.isfinite(out[k]).all() for k in ("u5", "sig5", "chi5", "xi5"))
    ref.close()
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)


def l2_case(name):
    """The reference's l2_comparison (src/shear_sim_3d.cpp:3122-3158) of a
    direct simulation against a quasi-static one from the same state, as
    shear_compare does it (src/shear_compare.cpp:196-205): 16 x 12 x 8 cells,
    z walled or periodic, config-0 material with eps0 = 0.3; the QS run takes
    2 steps of 0.01, the direct run (density 1, damping 0.35) steps of the
    reference's dt up to the same time, the last one shortened to land on
    it (z periodic: the QS run stays at the set state, since the reference's
    projection loop does not end on the singular periodic-z operator).
    Stored: both end states and the three squared differences."""
    import tempfile

    from stzpp_amd import shear_sim as SS

    N = (16, 12, 8)
    prd = (1, 1, 1) if name == "l2_periodic" else (1, 1, 0)
    Nx, Ny, Nz = N
    nc = Nx * Ny * Nz
    nv = Nx * Ny * (Nz + (0 if prd[2] else 1))
    p = SS.config0_params(16)
    p["eps0"] = 0.3
    P = {k: p[k] for k in O.STEP_PARAM_ORDER}
    rng = np.random.default_rng(41 if prd[2] else 40)
    sig0 = 0.05 * rng.standard_normal((nc, 6))
    sig0[:, 1] += 0.9
    u0 = 2e-3 * rng.standard_normal((nv, 3))
    chi0 = 0.09 + 0.01 * rng.standard_normal(nc)
    t0, qdt, nq = 0.3, 0.01, 2
    rho, kappa, tmult = 1.0, 0.35, 0.5
    dx = 2.0 / Nx
    ddt = min(tmult * (dx * dx / 3.0 / kappa / 2.0), dx / 4.0)
    q = O.RefShear(N, P, tempfile.mkdtemp(), prd=prd)
    q.setup(qdt)
    q.set_state(t0, u0, sig0, chi0)
    d = O.RefShear.direct(N, P, tempfile.mkdtemp(), rho, kappa, prd=prd)
    d.set_state(t0, u0, sig0, chi0)
    d.init_ref()
    l2_0 = d.l2_comparison(q)
    target = t0 + nq * qdt
    t, nd = t0, 0
    while t + ddt * (1 + 1e-8) < target:
        d.step_direct(ddt, 1)
        t = d.state()["t"]
        nd += 1
    d.step_direct(target - t, 1)
    if not prd[2]:  # the reference's projection loop does not end on a periodic-z grid
        q.step(qdt, nq)
    sd, sq = d.state(), q.state()
    out = dict(N=np.array(N), prd=np.array(prd), t0=np.array(t0), qdt=np.array(qdt), nq=np.array(nq),
               ddt=np.array(ddt), nd=np.array(nd), rho=np.array(rho), kappa=np.array(kappa),
               tmult=np.array(tmult), u0=u0, sig0=sig0, chi0=chi0, l2_0=l2_0, l2=d.l2_comparison(q),
               **{f"d_{k}": np.asarray(sd[k]) for k in ("t", "u", "sig", "chi")},
               **{f"q_{k}": np.asarray(sq[k]) for k in ("t", "u", "sig", "chi")},
               **{f"p_{k}": np.array(v) for k, v in P.items()})
    q.close()
    d.close()
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)


def adj_case(name):
    """Central finite differences of the reference's QS step
    (step_forward_qs, src/shear_sim_3d.cpp:381-389) for the adjoint's loss
    J = mean over cells of (ad_Dpl/dt)^2 (ad_Dpl as set at :431) after 3
    steps: 32 x 32 x 16 cells, z walled, config-0 material with eps0 = 0.3,
    a yielding band of two planes (sigma_12 = 1.12 against 0.5 elsewhere,
    with noise), t0 = 0.3, dt = 0.01.  Directions: 16 seeded random values
    on 4 x 4 x 4 patches of cells that overlap the band.  For each the
    reference's J at chi0 +- h d for h = 1e-6 and h = 3e-7, so a test can
    tell where J is smooth at that scale (the adaptive plastic sub-steps
    make it piecewise smooth)."""
    import tempfile
    N = (32, 32, 16)
    Nx, Ny, Nz = N
    nc, nv = Nx * Ny * Nz, Nx * Ny * (Nz + 1)
    P = config0_params(Nx)
    P["eps0"] = 0.3
    rng = np.random.default_rng(7)
    sig0 = np.zeros((nc, 6))
    k = np.arange(nc) // (Nx * Ny)
    sig0[:, 1] = np.where((k >= 7) & (k <= 8), 1.12, 0.5) + 0.02 * rng.standard_normal(nc)
    sig0[:, 0] = 0.05 * rng.standard_normal(nc)
    chi0 = 0.09 + 0.005 * rng.standard_normal(nc)
    u0 = 1e-3 * rng.standard_normal((nv, 3))
    t0, dt, nsteps = 0.3, 0.01, 3
    outdir = tempfile.mkdtemp()

    def loss(chi):
        ref = O.RefShear(N, P, outdir)
        ref.setup(dt)
        ref.set_state(t0, u0, sig0, chi)
        ref.step(dt, nsteps)
        d = ref.state()["dpl"]
        ref.close()
        return float(np.mean((d / dt) ** 2))

    drng = np.random.default_rng(12)
    idx, val, jp, jm = [], [], [], []
    for _ in range(16):
        ci, cj, ck = drng.integers(0, Nx), drng.integers(0, Ny), drng.integers(5, 9)
        cells = np.array([((ci + a) % Nx) + Nx * (((cj + b) % Ny) + Ny * (ck + c))
                          for c in range(4) for b in range(4) for a in range(4)])
        v = drng.standard_normal(cells.size)
        d = np.zeros(nc)
        d[cells] = v
        idx.append(cells)
        val.append(v)
        jp.append([loss(chi0 + h * d) for h in (1e-6, 3e-7)])
        jm.append([loss(chi0 - h * d) for h in (1e-6, 3e-7)])
    out = dict(N=np.array(N), t0=np.array(t0), dt=np.array(dt), nsteps=np.array(nsteps), u0=u0, sig0=sig0,
               chi0=chi0, J0=np.array(loss(chi0)), h=np.array([1e-6, 3e-7]), dir_cells=np.array(idx),
               dir_vals=np.array(val), J_plus=np.array(jp), J_minus=np.array(jm),
               **{f"p_{k}": np.array(v) for k, v in P.items()})
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)


CASES = {
    "shear_16x16x17": (("shear_16x16x17", (16, 16, 17), (1, 1, 0), "shear", 11),),
    "trans_16x16x16": (("trans_16x16x16", (16, 16, 16), (1, 1, 1), "trans", 12),),
    "shear_32x32x33": (("shear_32x32x33", (32, 32, 33), (1, 1, 0), "shear", 13),),
}

if __name__ == "__main__":
    if O.ref_lib() is None:
        sys.exit("oracle/_ref/libstzref.so is missing; run make -C oracle/ref")
    if len(sys.argv) > 1:
        which = sys.argv[1]
        if which == "trans":
            trans_case()
        elif which.startswith("step_"):
            step_case(which)
        elif which.startswith("trans_step_"):
            trans_step_case(which)
        elif which.startswith("grf_"):
            grf_case(which)
        elif which.startswith("output_"):
            output_case(which)
        elif which.startswith("direct_"):
            direct_case(which)
        elif which.startswith("l2_"):
            l2_case(which)
        elif which.startswith("adj_"):
            adj_case(which)
        else:
            mg_case(*CASES[which][0])
        sys.exit(0)
    for key in ["trans", *CASES, "step_plastic", "step_c0", "trans_step_walled", "trans_step_periodic",
                *GRF_CASES, "output_walled", "output_periodic", "output_trans", "direct_walled",
                "direct_periodic", "l2_walled", "l2_periodic", "adj_ref_32"]:
        subprocess.run([sys.executable, __file__, key], check=True, cwd=ROOT,
                       stdout=subprocess.DEVNULL)
        print("wrote", key)
