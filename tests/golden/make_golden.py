"""Generate the golden fixtures in tests/golden/ from the reference itself.

Run in a container that has /root/reference:

    make -C oracle ref && python tests/golden/make_golden.py

It drives the reference's own C sources (compiled into oracle/_ref by
oracle/Makefile, reached through oracle/ref_glue.c) on seeded inputs and writes
inputs and outputs as compressed .npz files (plain float64 arrays, no pickles).
The fixtures are data only; the reference cannot travel to the GPU machine,
these vectors do.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from _libs import IFV_MEMBERS, RefLib, ifv_members, make_config  # noqa: E402
import cases  # noqa: E402

EPS = 1e-9


def interface_vectors(ref):
    L, R = cases.interface_pairs(2000, seed=20240607)
    cL = np.sqrt(L[:, 0] * L[:, 6] / L[:, 3])
    cR = np.sqrt(R[:, 0] * R[:, 6] / R[:, 3])
    riem = np.zeros((len(L), 5))
    riem12 = np.zeros((len(L), 5))
    ben = np.zeros((len(L), 5))
    q1d = np.zeros((len(L), 16))
    fx = np.zeros((len(L), 9))
    fy = np.zeros((len(L), 9))
    edir = np.zeros((len(L), 6))
    tau = 2.5e-3
    ref.set_config(make_config())
    for k in range(len(L)):
        riem[k] = ref.riemann([L[k, 0], R[k, 0], L[k, 4], R[k, 4], L[k, 6], R[k, 6], cL[k], cR[k], EPS, EPS], 500)
        riem12[k] = ref.riemann([L[k, 0], R[k, 0], L[k, 4], R[k, 4], L[k, 6], R[k, 6], cL[k], cR[k], 1e-12, 1e-12], 500)
        ben[k] = ref.riemann_ben([L[k, 0], L[k, 4], R[k, 4], L[k, 6], R[k, 6], cL[k], cR[k], EPS, EPS], 50)
        q1d[k] = ref.q1d(L[k], R[k], EPS, EPS)
        rc, out = ref.flux2d(L[k], R[k], tau, EPS, 1.0, 0.0)
        fx[k, 0], fx[k, 1:] = rc, out
        rc, out = ref.flux2d(L[k], R[k], tau, EPS, 0.0, 1.0)
        fy[k, 0], fy[k, 1:] = rc, out
        edir[k] = ref.edir(L[k], R[k], EPS, EPS)
    np.savez_compressed(os.path.join(HERE, "interface_vectors.npz"), L=L, R=R, tau=tau, eps=EPS,
                        riemann=riem, riemann_tol12=riem12, riemann_ben=ben, q1d=q1d,
                        flux_x=fx, flux_y=fy, edir=edir)


def unit_pairs(count, seed):
    """Interface pairs for the per-interface entry points with every argument
    in play: a grid velocity on the left side, a second gas (gamma 1.67) on
    the right side of every other pair."""
    L, R = cases.interface_pairs(count, seed=seed)
    rng = np.random.default_rng(seed + 1)
    c = np.sqrt(L[:, 0] * L[:, 6] / L[:, 3])
    L[:, 1] = rng.uniform(-1, 1, count) * c
    L[:, 2] = rng.uniform(-1, 1, count) * c
    R[1::2, 0] = 1.67
    return L, R


def interface_units(ref):
    """The per-interface solvers called as functions of their own (atc, grid
    velocity, two gases, any normal, side effects on the structs)."""
    L, R = unit_pairs(1000, seed=20240901)
    n = len(L)
    tau = 2.5e-3
    ref.set_config(make_config())
    cL = np.sqrt(L[:, 0] * L[:, 6] / L[:, 3])
    cR = np.sqrt(R[:, 0] * R[:, 6] / R[:, 3])
    riem_in = np.stack([L[:, 0], R[:, 0], L[:, 4], R[:, 4], L[:, 6], R[:, 6], cL, cR,
                        np.full(n, EPS), np.full(n, EPS)], axis=1)
    riem = np.array([ref.riemann(riem_in[k], 500) for k in range(n)])
    atcs = np.array([EPS, np.inf, -0.0])
    q1d = np.array([[ref.q1d(L[k], R[k], EPS, a) for k in range(n)] for a in atcs])
    edir_inf = np.array([ref.edir(L[k], R[k], EPS, np.inf) for k in range(n)])
    normals = np.array([[0.6, 0.8], [-0.28, 0.96], [1.0, 0.0], [0.0, 1.0]])
    flux = np.zeros((len(normals), n, 9))
    for a, (nx, ny) in enumerate(normals):
        for k in range(n):
            rc, o = ref.flux2d(L[k], R[k], tau, EPS, nx, ny)
            flux[a, k, 0], flux[a, k, 1:] = rc, o
    # whole structs: the members the solver does not read hold markers, the
    # flux members too (GRP_2D_flux leaves them when its check fails)
    ns = 300
    rng = np.random.default_rng(7)
    sin_L, sin_R, sout_L, sout_R, src = (np.zeros((ns, 34)), np.zeros((ns, 34)), np.zeros((ns, 34)),
                                         np.zeros((ns, 34)), np.zeros(ns))
    for k in range(ns):
        nx, ny = normals[k % 2]
        extra = {name: rng.uniform(1, 2) for name in ("delta_x", "delta_y", "length", "RHO_int", "P_int",
                                                      "U_int", "V_int", "F_rho", "F_e", "F_u", "F_v",
                                                      "U_rho", "U_e", "U_u", "U_v", "d_e", "t_e")}
        sin_L[k] = ifv_members(L[k], nx, ny, extra)
        sin_R[k] = ifv_members(R[k], nx, ny, extra)
        # the last four fail the solver's checks: a negative density on a
        # supersonic interface (code 1), NaN / inf values (2), NaN slopes (3)
        bad = {ns - 4: {"RHO": -1.0, "P": -1.0, "U": 50.0, "V": 50.0}, ns - 3: {"RHO": np.nan},
               ns - 2: {"U": np.inf}, ns - 1: {"d_rho": np.nan}}
        for name, v in bad.get(k, {}).items():
            sin_L[k, IFV_MEMBERS.index(name)] = v
            if name == "d_rho":
                sin_R[k, IFV_MEMBERS.index(name)] = v
        src[k], sout_L[k], sout_R[k] = ref.flux2d_struct(sin_L[k], sin_R[k], tau, EPS)
    # GuessP on the same pairs: DL DR UL UR PL PR CL CR GammaL GammaR eps
    gp_in = np.stack([L[:, 3], R[:, 3], L[:, 4], R[:, 4], L[:, 6], R[:, 6], cL, cR, L[:, 0], R[:, 0],
                      np.full(n, EPS)], axis=1)
    gp = np.array([ref.guessp(gp_in[k]) for k in range(n)])
    np.savez_compressed(os.path.join(HERE, "interface_units.npz"), L=L, R=R, tau=tau, eps=EPS,
                        guessp_in=gp_in, guessp=gp,
                        riemann_in=riem_in, riemann=riem, atc=atcs, q1d=q1d, edir_inf=edir_inf,
                        normals=normals, flux=flux, struct_in_L=sin_L, struct_in_R=sin_R,
                        struct_out_L=sout_L, struct_out_R=sout_R, struct_rc=src)


STAGE_BCS = ((-4, -4), (-2, -7), (-7, -24), (-24, -2), (-1, -1))


def stage_fields(m, n, seed):
    """Two frames (4, m, n) of smooth positive states, slope sets and ghost
    arrays with every member set (b_f_var: H RHO P U V SRHO SP SU SV TRHO TP TU TV)."""
    rng = np.random.default_rng(seed)
    f = []
    for k in range(2):
        init = cases.smooth_random(m, n, seed + k)
        f.append(np.stack([init["rho"], init["u"], init["v"], init["p"]]))
    s, t = rng.normal(0, 2, (4, m, n)), rng.normal(0, 2, (4, m, n))
    bfv = {"L": rng.uniform(0.5, 1.5, (n, 13)), "R": rng.uniform(0.5, 1.5, (n, 13)),
           "D": rng.uniform(0.5, 1.5, (m, 13)), "U": rng.uniform(0.5, 1.5, (m, 13))}
    return f[0], f[1], s, t, bfv


def stage_cases(ref):
    """bound_cond_slope_limiter_{x,y} and flux_generator_{x,y} on a 9 x 7 grid
    (nt = 1), every boundary condition pair, both limiter forms, with and without
    transversal derivatives, and a flux run with a failing cell."""
    m, n = 9, 7
    out = {}
    f0, f1, s, t, bfv = stage_fields(m, n, 31)
    out.update(f0=f0, f1=f1, s=s, t=t, **{"bfv_" + k: v for k, v in bfv.items()})
    lim_meta, lim_s, lim_bfv = [], [], []
    for bx, by in STAGE_BCS:
        for direction in (0, 1):
            for found, slope in ((0, 1), (1, 1), (0, 0)):
                ref.set_config(make_config(h_x=1 / m, h_y=1 / n, bc_x=bx, bc_y=by))
                r, sl, b = ref.limiter(direction, m, n, f0, f1, t if direction else s, s if direction else t, bfv,
                                       found, slope)
                lim_meta.append([bx, by, direction, found, slope, r])
                lim_s.append(sl)
                lim_bfv.append(np.concatenate([b[k] for k in ("L", "R", "D", "U")]))
    out.update(lim_meta=np.array(lim_meta), lim_s=np.array(lim_s), lim_bfv=np.array(lim_bfv))
    flux_meta, flux_F, flux_I = [], [], []
    tau = 2e-3
    for direction in (0, 1):
        for transversal in (0, 1):
            ref.set_config(make_config(h_x=1 / m, h_y=1 / n))
            lo, hi = (bfv["L"], bfv["R"]) if direction == 0 else (bfv["D"], bfv["U"])
            r, F, I = ref.flux_gen(direction, m, n, f1, s * 0.1, t * 0.1, lo, hi, tau, transversal)
            flux_meta.append([direction, transversal, r])
            flux_F.append(F.reshape(4, -1))
            flux_I.append(I.reshape(4, -1))
    out.update(flux_meta=np.array(flux_meta), flux_tau=tau)
    out["flux_F_x"], out["flux_I_x"] = np.array(flux_F[:2]), np.array(flux_I[:2])
    out["flux_F_y"], out["flux_I_y"] = np.array(flux_F[2:]), np.array(flux_I[2:])
    # a cell with negative pressure: its interfaces fail (return value only; the
    # reference's fluxes there depend on its thread schedule)
    fbad = f1.copy()
    fbad[3, 4, 3] = -0.5
    bad = []
    for direction in (0, 1):
        lo, hi = (bfv["L"], bfv["R"]) if direction == 0 else (bfv["D"], bfv["U"])
        r, F, I = ref.flux_gen(direction, m, n, fbad, s * 0.1, t * 0.1, lo, hi, tau, 1)
        bad.append(r)
    out["fail_f1"], out["fail_rc"] = fbad, np.array(bad)
    np.savez_compressed(os.path.join(HERE, "stage_cases.npz"), **out)


def split2d_cases(ref):
    out = {}
    # four-quadrant 64^2 to t = 0.3 (127 steps), full final state
    m = n = 64
    cfg = cases.four_quadrant_config(m, n)
    fr, npl, tp, c = ref.split2d(m, n, cases.four_quadrant(m, n), cfg)
    out["fq64_final"] = fr[:, npl - 1]
    out["fq64_meta"] = np.array([c[5], npl, tp[npl - 1]])
    # four-quadrant 32 x 48 with 5 output frames
    m, n = 32, 48
    cfg = cases.four_quadrant_config(m, n, t_all=0.2)
    tpin = np.array([0.0, 0.05, 0.1, 0.15, 0.2])
    fr, npl, tp, c = ref.split2d(m, n, cases.four_quadrant(m, n), cfg, time_plot=tpin)
    out["frames_all"] = fr
    out["frames_meta"] = np.concatenate([[c[5], npl], tp])
    # first plot time after t = 0: frame 0 is the live state for a while
    m, n = 24, 20
    init = cases.smooth_random(m, n, 11)
    cfg = make_config(t_all=0.3, h_x=1 / m, h_y=1 / n, bc_x=-4, bc_y=-7, split=1)
    out["late_init"] = np.stack([init[k] for k in ("rho", "u", "v", "p")])
    fr, npl, tp, c = ref.split2d(m, n, init, cfg, time_plot=np.array([0.1, 0.2, 0.3]), N_T=3)
    out["late_frames"] = fr
    out["late_meta"] = np.concatenate([[c[5], npl], tp])
    # every boundary-condition pair on a 12 x 10 smooth field, t = 0.05
    bcs = (-1, -2, -4, -7, -24)
    init = cases.smooth_random(12, 10, 5)
    out["bc_init"] = np.stack([init[k] for k in ("rho", "u", "v", "p")])
    finals, metas = [], []
    for bx in bcs:
        for by in bcs:
            cfg = make_config(t_all=0.05, h_x=1 / 12, h_y=1 / 10, bc_x=bx, bc_y=by, split=1)
            fr, npl, tp, c = ref.split2d(12, 10, init, cfg)
            finals.append(fr[:, npl - 1])
            metas.append([bx, by, c[5], npl, tp[npl - 1]])
    out["bc_final"] = np.stack(finals)
    out["bc_meta"] = np.array(metas, dtype=np.float64)
    # 2D Sod tube 100 x 4, t = 0.2 (the survey's frozen-dt check: 53 steps)
    m, n = 100, 4
    x = (np.arange(m) + 0.5) / m
    init = {"rho": np.repeat(np.where(x < 0.5, 1.0, 0.125)[:, None], n, 1),
            "u": np.zeros((m, n)), "v": np.zeros((m, n)),
            "p": np.repeat(np.where(x < 0.5, 1.0, 0.1)[:, None], n, 1)}
    cfg = make_config(t_all=0.2, h_x=0.01, h_y=0.01, bc_x=-4, bc_y=-4, split=1)
    fr, npl, tp, c = ref.split2d(m, n, init, cfg)
    out["sod2d_final"] = fr[:, npl - 1]
    out["sod2d_meta"] = np.array([c[5], npl, tp[npl - 1]])
    # fixed-tau mode (t_all infinite, config[16] > 0, 7 steps): step counter quirk
    cfg = make_config(N=7, tau=1e-3, h_x=1 / 16, h_y=1 / 16, bc_x=-4, bc_y=-7, split=1)
    cfg[1] = np.inf
    init = cases.smooth_random(16, 16, 9)
    out["fixtau_init"] = np.stack([init[k] for k in ("rho", "u", "v", "p")])
    fr, npl, tp, c = ref.split2d(16, 16, init, cfg, time_plot=np.array([0.0, np.inf]))
    out["fixtau_final"] = fr[:, npl - 1]
    out["fixtau_meta"] = np.array([c[5], npl, tp[npl - 1]])
    # four-quadrant 256^2 to t = 0.3: too big to store whole; per-field sums,
    # L1 norms and probe cells
    m = n = 256
    cfg = cases.four_quadrant_config(m, n)
    fr, npl, tp, c = ref.split2d(m, n, cases.four_quadrant(m, n), cfg)
    fin = fr[:, npl - 1]
    probes = np.array([[3, 5], [100, 200], [128, 128], [200, 31], [255, 255], [77, 160], [160, 77]])
    out["fq256_sum"] = fin.reshape(5, -1).sum(axis=1)
    out["fq256_l1"] = np.abs(fin).reshape(5, -1).sum(axis=1)
    out["fq256_probes"] = probes.astype(np.float64)
    out["fq256_probe_vals"] = fin[:, probes[:, 0], probes[:, 1]]
    out["fq256_rows"] = fin[:, ::37, :]       # every 37th x-line in full
    out["fq256_meta"] = np.array([c[5], npl, tp[npl - 1]])
    np.savez_compressed(os.path.join(HERE, "split2d_cases.npz"), **out)


def split2d_fq1024(ref):
    """The four-quadrant problem at 1024^2 to t = 0.3 (2017 steps), the
    benchmark's size, through the reference's own split solver: per-field sums
    and L1 norms, probe cells and every 97th x-line in full."""
    m = n = 1024
    cfg = cases.four_quadrant_config(m, n)
    fr, npl, tp, c = ref.split2d(m, n, cases.four_quadrant(m, n), cfg)
    fin = fr[:, npl - 1]
    rng = np.random.default_rng(1024)
    probes = np.concatenate([np.array([[0, 0], [511, 511], [512, 512], [1023, 1023], [300, 700], [700, 300]]),
                             rng.integers(0, 1024, (26, 2))])
    out = {"sum": fin.reshape(5, -1).sum(axis=1), "l1": np.abs(fin).reshape(5, -1).sum(axis=1),
           "probes": probes.astype(np.float64), "probe_vals": fin[:, probes[:, 0], probes[:, 1]],
           "rows": fin[:, ::97, :], "meta": np.array([c[5], npl, tp[npl - 1]])}
    np.savez_compressed(os.path.join(HERE, "split2d_fq1024.npz"), **out)


def split2d_fq1024_devpow(_ref):
    """The same 1024^2 whole run through the oracle built with the device's pow
    (oracle/liboracle_devpow.so): the reference's algorithm with the one
    operation the GPU computes differently taken from the GPU's library.  The
    same rows, probes and sums as split2d_fq1024."""
    from _libs import Oracle, ORACLE_DEVPOW_SO
    m = n = 1024
    cfg = cases.four_quadrant_config(m, n)
    fr, npl, tp, c, st = Oracle(ORACLE_DEVPOW_SO).split2d(m, n, cases.four_quadrant(m, n), cfg, N_T=2)
    fin = fr[:, npl - 1]
    probes = np.load(os.path.join(HERE, "split2d_fq1024.npz"), allow_pickle=False)["probes"].astype(int)
    out = {"sum": fin.reshape(5, -1).sum(axis=1), "l1": np.abs(fin).reshape(5, -1).sum(axis=1),
           "probe_vals": fin[:, probes[:, 0], probes[:, 1]], "rows": fin[:, ::97, :],
           "meta": np.array([c[5], npl, tp[npl - 1]])}
    np.savez_compressed(os.path.join(HERE, "split2d_fq1024_devpow.npz"), **out)


def twofluid_vectors(_ref):
    """Two-fluid interfaces through the reference's GRP_2D_flux and Q1D built
    with -DMULTIFLUID_BASICS (oracle/_ref/libhydroref_mf.so): fluid a gamma 1.4
    (config[6]), fluid b gamma 1.67 (config[106]); each side pure a, pure b or
    mixed (volume fraction z_a uniform in (0, 1), gamma from it by the
    reference's mixture rule, src/inter_process_unstruct/cons_qty_calc.c:70),
    weak and strong waves, near-equal states; normals x and y with the split
    path's zero tangential derivatives (+0 in x, -0 in y)."""
    from _libs import RefLibMF
    ref = RefLibMF()
    ga, gb = 1.4, 1.67
    ref.set_config({4: EPS, 6: ga, 106: gb})
    rng = np.random.default_rng(1067)
    n = 1200

    def side(k):
        kind = rng.integers(0, 3)
        z = 1.0 if kind == 0 else (0.0 if kind == 1 else rng.uniform(0.02, 0.98))
        g = ga if z == 1.0 else (gb if z == 0.0 else 1.0 / (z / (ga - 1.0) + (1.0 - z) / (gb - 1.0)) + 1.0)
        rho, p = 10 ** rng.uniform(-1, 1), 10 ** rng.uniform(-1, 1)
        u, v = rng.uniform(-1, 1), rng.uniform(-1, 1)
        d = rng.normal(0, 1, 4)
        dz = rng.normal(0, 1) if 0.0 < z < 1.0 else 0.0
        return [g, 0.0, 0.0, rho, u, v, p, *d, 0.0, 0.0, 0.0, 0.0, z, dz, 0.0, 0.0, 0.0, 0.0]

    L = np.array([side(k) for k in range(n)])
    R = np.array([side(k) for k in range(n)])
    # near-equal states (the acoustic case) for a tenth of the pairs
    near = rng.uniform(size=n) < 0.1
    R[near, 3:7] = L[near, 3:7] * (1 + 1e-11 * rng.normal(size=(near.sum(), 4)))
    tau = 2e-3
    out = {"L": L, "R": R, "tau": tau, "eps": EPS, "gamma_a": ga, "gamma_b": gb}
    for direction, (nx, ny), tz in ((0, (1.0, 0.0), 0.0), (1, (0.0, 1.0), -0.0)):
        Ld, Rd = L.copy(), R.copy()
        Ld[:, 11:15] = tz
        Rd[:, 11:15] = tz
        out[f"flux_{direction}"] = np.array([ref.grp_2d_flux(Ld[k], Rd[k], nx, ny, tau) for k in range(n)])
    out["q1d"] = np.array([ref.q1d(L[k], R[k], EPS, EPS) for k in range(n)])
    np.savez_compressed(os.path.join(HERE, "twofluid_vectors.npz"), **out)


def g2d_vectors(ref):
    """linear_GRP_solver_Edir_G2D (src/riemann_solver/linear_grp_solver_Edir_G2D.c:48-637)
    from the reference's single-fluid build (oracle/_ref/libhydroref.so) and
    its -DMULTIFLUID_BASICS build (libhydroref_mf.so), at atc = eps, INFINITY
    and -0.0.  Pairs: unit_pairs (grid velocity lambda_u on the left side,
    lambda_v on the right one, a second gas on every other right side) with
    near-equal pairs made near-equal in v too (the solver's distance includes
    v), every seventh pair at gamma 3 (the rarefaction's log limit), and the
    multi-fluid members z_a, phi in [0, 1] with normal and tangential slopes."""
    from _libs import RefLibMF
    mf = RefLibMF()
    mf.set_config({4: EPS, 6: 1.4, 106: 1.67})
    L, R = unit_pairs(800, seed=31)
    n = len(L)
    rng = np.random.default_rng(5)
    R[:, 2] = rng.uniform(-1, 1, n) * np.sqrt(R[:, 0] * R[:, 6] / R[:, 3])
    near = np.abs(L[:, 3] - R[:, 3]) < 1e-8
    R[near, 5] = L[near, 5] + 1e-12
    L[::7, 0] = 3.0
    R[::7, 0] = 3.0

    def members():
        return np.concatenate([rng.uniform(0, 1, (n, 1)), rng.normal(0, 1, (n, 2)),
                               rng.uniform(0, 1, (n, 1)), rng.normal(0, 1, (n, 2))], axis=1)

    L = np.concatenate([L, members()], axis=1)
    R = np.concatenate([R, members()], axis=1)
    atcs = np.array([EPS, np.inf, -0.0])
    single = np.array([[ref.g2d(L[k, :15], R[k, :15], EPS, a) for k in range(n)] for a in atcs])
    multi = np.array([[mf.g2d(L[k], R[k], EPS, a) for k in range(n)] for a in atcs])
    np.savez_compressed(os.path.join(HERE, "g2d_vectors.npz"), L=L, R=R, eps=EPS, atc=atcs,
                        single=single, multi=multi)


def unsplit2d_cases(ref):
    """GRP_solver_2D_EUL_source (the unsplit step, config[33] = 0)."""
    out = {}
    m = n = 64
    cfg = cases.four_quadrant_config(m, n, split=0)
    fr, npl, tp, c = ref.unsplit2d(m, n, cases.four_quadrant(m, n), cfg)
    out["fq64_final"] = fr[:, npl - 1]
    out["fq64_meta"] = np.array([c[5], npl, tp[npl - 1]])
    # 32 x 48 with 5 frames
    m, n = 32, 48
    cfg = cases.four_quadrant_config(m, n, t_all=0.2, split=0)
    tpin = np.array([0.0, 0.05, 0.1, 0.15, 0.2])
    fr, npl, tp, c = ref.unsplit2d(m, n, cases.four_quadrant(m, n), cfg, time_plot=tpin)
    out["frames_all"] = fr
    out["frames_meta"] = np.concatenate([[c[5], npl], tp])
    # first plot time after t = 0: frame 0 is the live state for a while, and
    # the CFL bound follows it there
    m, n = 24, 20
    init = cases.smooth_random(m, n, 11)
    cfg = make_config(t_all=0.3, h_x=1 / m, h_y=1 / n, bc_x=-4, bc_y=-7, split=0)
    tpin = np.array([0.1, 0.2, 0.3])
    out["late_init"] = np.stack([init[k] for k in ("rho", "u", "v", "p")])
    fr, npl, tp, c = ref.unsplit2d(m, n, init, cfg, time_plot=tpin, N_T=3)
    out["late_frames"] = fr
    out["late_meta"] = np.concatenate([[c[5], npl], tp])
    # every boundary-condition pair on a 12 x 10 smooth field, t = 0.2
    bcs = (-1, -2, -4, -7, -24)
    init = cases.smooth_random(12, 10, 5)
    out["bc_init"] = np.stack([init[k] for k in ("rho", "u", "v", "p")])
    finals, metas = [], []
    for bx in bcs:
        for by in bcs:
            cfg = make_config(t_all=0.2, h_x=1 / 12, h_y=1 / 10, bc_x=bx, bc_y=by, split=0)
            fr, npl, tp, c = ref.unsplit2d(12, 10, init, cfg)
            finals.append(fr[:, npl - 1])
            metas.append([bx, by, c[5], npl, tp[npl - 1]])
    out["bc_final"] = np.stack(finals)
    out["bc_meta"] = np.array(metas, dtype=np.float64)
    # reflective box with a blast: walls on all four sides
    m, n = 40, 36
    init = cases.blast_box(m, n)
    cfg = make_config(t_all=0.15, h_x=1 / m, h_y=1 / n, bc_x=-2, bc_y=-2, split=0)
    fr, npl, tp, c = ref.unsplit2d(m, n, init, cfg)
    out["blast_final"] = fr[:, npl - 1]
    out["blast_meta"] = np.array([c[5], npl, tp[npl - 1]])
    # fixed-tau mode (t_all infinite, config[16] > 0, 7 steps)
    cfg = make_config(N=7, tau=1e-3, h_x=1 / 16, h_y=1 / 16, bc_x=-24, bc_y=-1, split=0)
    cfg[1] = np.inf
    init = cases.smooth_random(16, 16, 9)
    out["fixtau_init"] = np.stack([init[k] for k in ("rho", "u", "v", "p")])
    fr, npl, tp, c = ref.unsplit2d(16, 16, init, cfg, time_plot=np.array([0.0, np.inf]))
    out["fixtau_final"] = fr[:, npl - 1]
    out["fixtau_meta"] = np.array([c[5], npl, tp[npl - 1]])
    # four-quadrant 256^2 to t = 0.3: sums, L1 norms, probes and every 37th x-line
    m = n = 256
    cfg = cases.four_quadrant_config(m, n, split=0)
    fr, npl, tp, c = ref.unsplit2d(m, n, cases.four_quadrant(m, n), cfg)
    fin = fr[:, npl - 1]
    probes = np.array([[3, 5], [100, 200], [128, 128], [200, 31], [255, 255], [77, 160], [160, 77]])
    out["fq256_sum"] = fin.reshape(5, -1).sum(axis=1)
    out["fq256_l1"] = np.abs(fin).reshape(5, -1).sum(axis=1)
    out["fq256_probes"] = probes.astype(np.float64)
    out["fq256_probe_vals"] = fin[:, probes[:, 0], probes[:, 1]]
    out["fq256_rows"] = fin[:, ::37, :]
    out["fq256_meta"] = np.array([c[5], npl, tp[npl - 1]])
    np.savez_compressed(os.path.join(HERE, "unsplit2d_cases.npz"), **out)


def unsplit2d_g2d_cases(_ref):
    """The unsplit step with GRP_2D_flux's alternative interface solver,
    linear_GRP_solver_Edir_G2D (flux_solver.c:189), from the reference build
    oracle/_ref/libhydroref_g2d.so: four-quadrant 64^2 to t = 0.3, every
    boundary-condition pair on a 12 x 10 smooth field (t = 0.2), a blast in a
    reflective box (t = 0.15)."""
    from _libs import REF_G2D_SO
    ref = RefLib(REF_G2D_SO)
    out = {}
    m = n = 64
    cfg = cases.four_quadrant_config(m, n, split=0)
    fr, npl, tp, c = ref.unsplit2d(m, n, cases.four_quadrant(m, n), cfg)
    out["fq64_final"] = fr[:, npl - 1]
    out["fq64_meta"] = np.array([c[5], npl, tp[npl - 1]])
    bcs = (-1, -2, -4, -7, -24)
    init = cases.smooth_random(12, 10, 5)
    out["bc_init"] = np.stack([init[k] for k in ("rho", "u", "v", "p")])
    finals, metas = [], []
    for bx in bcs:
        for by in bcs:
            cfg = make_config(t_all=0.2, h_x=1 / 12, h_y=1 / 10, bc_x=bx, bc_y=by, split=0)
            fr, npl, tp, c = ref.unsplit2d(12, 10, init, cfg)
            finals.append(fr[:, npl - 1])
            metas.append([bx, by, c[5], npl, tp[npl - 1]])
    out["bc_final"] = np.stack(finals)
    out["bc_meta"] = np.array(metas, dtype=np.float64)
    m, n = 40, 36
    cfg = make_config(t_all=0.15, h_x=1 / m, h_y=1 / n, bc_x=-2, bc_y=-2, split=0)
    fr, npl, tp, c = ref.unsplit2d(m, n, cases.blast_box(m, n), cfg)
    out["blast_final"] = fr[:, npl - 1]
    out["blast_meta"] = np.array([c[5], npl, tp[npl - 1]])
    np.savez_compressed(os.path.join(HERE, "unsplit2d_g2d_cases.npz"), **out)


def grp1d_cases(ref):
    out = {}
    cfg = make_config(t_all=0.2, h_x=1e-3, bc_x=-4, CFL=0.45)
    fr, npl, tp, c = ref.grp1d(1000, cases.sod_1d(1000), cfg)
    out["sod1000_final"] = fr[:, npl - 1]
    out["sod1000_meta"] = np.array([c[5], npl, tp[npl - 1]])
    init = {k: v[:, 0] for k, v in cases.smooth_random(200, 1, 3).items() if k != "v"}
    out["bc_init"] = np.stack([init[k] for k in ("rho", "u", "p")])
    finals, metas = [], []
    for bc in (-1, -2, -4, -7, -24):
        cfg = make_config(t_all=0.1, h_x=1 / 200, bc_x=bc, CFL=0.45)
        fr, npl, tp, c = ref.grp1d(200, init, cfg)
        finals.append(fr[:, npl - 1])
        metas.append([bc, c[5], npl, tp[npl - 1]])
    out["bc_final"] = np.stack(finals)
    out["bc_meta"] = np.array(metas)
    # plot frames: Sod-400 at five plot times, and a first plot time after t = 0
    # (the loop runs in frame 0 until then)
    tp = np.array([0.0, 0.05, 0.1, 0.15, 0.2])
    cfg = make_config(t_all=0.2, h_x=1 / 400, bc_x=-2, CFL=0.45)
    fr, npl, tp, c = ref.grp1d(400, cases.sod_1d(400), cfg, N_T=5, time_plot=tp)
    out["frames_all"] = fr[:, :npl]
    out["frames_meta"] = np.r_[c[5], npl, tp[:npl]]
    init = {k: v[:, 0] for k, v in cases.smooth_random(150, 1, 5).items() if k != "v"}
    out["late_init"] = np.stack([init[k] for k in ("rho", "u", "p")])
    cfg = make_config(t_all=0.16, h_x=1 / 150, bc_x=-7, CFL=0.45)
    fr, npl, tp, c = ref.grp1d(150, init, cfg, N_T=3, time_plot=np.array([0.04, 0.1, 0.16]))
    out["late_frames"] = fr[:, :npl]
    out["late_meta"] = np.r_[c[5], npl, tp[:npl]]
    np.savez_compressed(os.path.join(HERE, "grp1d_cases.npz"), **out)


def main():
    ref = RefLib()
    only = sys.argv[1:]
    for name, fn in (("interface", interface_vectors), ("units", interface_units), ("stages", stage_cases),
                     ("split2d", split2d_cases), ("split2d_fq1024", split2d_fq1024),
                     ("split2d_fq1024_devpow", split2d_fq1024_devpow), ("twofluid", twofluid_vectors),
                     ("g2d", g2d_vectors), ("unsplit2d_g2d", unsplit2d_g2d_cases),
                     ("unsplit2d", unsplit2d_cases), ("grp1d", grp1d_cases)):
        if not only or name in only:
            fn(ref)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
