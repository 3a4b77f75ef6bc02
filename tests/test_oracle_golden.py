"""The oracle (oracle/liboracle.so) against the reference's own outputs.

Fixtures in tests/golden/ were produced by tests/golden/make_golden.py from the
reference sources compiled in oracle/_ref.  Equality is bit-for-bit: the
oracle restates the reference operation for operation.
"""
import numpy as np
import pytest

import cases
from _libs import make_config

EPS = 1e-9


def _eq(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)


def test_interface_vectors_bitexact(oracle, golden_iface):
    g = golden_iface
    L, R, tau = g["L"], g["R"], float(g["tau"])
    bad = []
    for k in range(len(L)):
        cL = np.sqrt(L[k, 0] * L[k, 6] / L[k, 3])
        cR = np.sqrt(R[k, 0] * R[k, 6] / R[k, 3])
        r = oracle.riemann(L[k, 0], R[k, 0], L[k, 4], R[k, 4], L[k, 6], R[k, 6], cL, cR, EPS, EPS, 500)
        if not _eq(r[:5], g["riemann"][k]):
            bad.append(("riemann", k))
        r = oracle.riemann(L[k, 0], R[k, 0], L[k, 4], R[k, 4], L[k, 6], R[k, 6], cL, cR, 1e-12, 1e-12, 500)
        if not _eq(r[:5], g["riemann_tol12"][k]):
            bad.append(("riemann_tol12", k))
        r = oracle.riemann_ben(L[k, 0], L[k, 4], R[k, 4], L[k, 6], R[k, 6], cL, cR, EPS, EPS, 50)
        if not _eq(r[:5], g["riemann_ben"][k]):
            bad.append(("riemann_ben", k))
        if not _eq(oracle.q1d(L[k], R[k], EPS, EPS), g["q1d"][k]):
            bad.append(("q1d", k))
        for key, (nx, ny) in (("flux_x", (1.0, 0.0)), ("flux_y", (0.0, 1.0))):
            rc, out = oracle.flux2d(L[k], R[k], tau, EPS, nx, ny)
            if rc != g[key][k, 0] or (rc == 0 and not _eq(out, g[key][k, 1:])):
                bad.append((key, k))
        if not _eq(oracle.edir(L[k], R[k], EPS, EPS), g["edir"][k]):
            bad.append(("edir", k))
    assert not bad, bad[:10]


def test_interface_units_bitexact(oracle, golden_units):
    """The per-interface solvers with every argument in play (atc INFINITY and
    -0.0, a grid velocity, a second gas, oblique normals): the oracle equals
    the reference bit for bit."""
    g = golden_units
    L, R, tau = g["L"], g["R"], float(g["tau"])
    bad = []
    for k in range(len(L)):
        if not _eq(oracle.riemann(*g["riemann_in"][k], 500)[:5], g["riemann"][k]):
            bad.append(("riemann", k))
        for a, atc in enumerate(g["atc"]):
            if not _eq(oracle.q1d(L[k], R[k], EPS, atc), g["q1d"][a, k]):
                bad.append(("q1d", a, k))
        if not _eq(oracle.edir(L[k], R[k], EPS, np.inf), g["edir_inf"][k]):
            bad.append(("edir_inf", k))
        for a, (nx, ny) in enumerate(g["normals"]):
            rc, out = oracle.flux2d(L[k], R[k], tau, EPS, nx, ny)
            if rc != g["flux"][a, k, 0] or not _eq(out, g["flux"][a, k, 1:]):
                bad.append(("flux", a, k))
    assert not bad, bad[:10]


def test_guessp_bitexact_and_branches(oracle, golden_units):
    """GuessP (the PVRS / two-shock starting pressure) of the oracle equals the
    reference's on 1000 pairs that take each of its three branches; and it is
    what its formulas say: the PVRS value where that branch applies, the
    two-shock root where the collision is strong (checked against the textbook
    two-shock approximation, Toro eq. 9.42, written independently below)."""
    g = golden_units
    gi, want = g["guessp_in"], g["guessp"]
    got = np.array([oracle.guessp(r) for r in gi])
    assert np.array_equal(got, want)
    DL, DR, UL, UR, PL, PR, CL, CR, gL, gR, eps = gi.T
    ppv = np.maximum(0.5 * (PL + PR) + 0.5 * (UL - UR) * 0.25 * (DL + DR) * (CL + CR), eps)
    pmin, pmax = np.minimum(PL, PR), np.maximum(PL, PR)
    pv = ((pmax / pmin <= 2) & (pmin <= ppv) & (ppv <= pmax)) | (ppv < pmin)
    assert pv.sum() > 200 and (~pv).sum() > 200
    assert np.allclose(got[pv], ppv[pv], rtol=1e-14, atol=0)
    AL, AR = 2.0 / ((gL + 1) * DL), 2.0 / ((gR + 1) * DR)
    BL, BR = (gL - 1) / (gL + 1) * PL, (gR - 1) / (gR + 1) * PR
    gel, ger = np.sqrt(AL / (BL + ppv)), np.sqrt(AR / (BR + ppv))
    ts = np.maximum((gel * PL + ger * PR - (UR - UL)) / (gel + ger), eps)
    assert np.allclose(got[~pv], ts[~pv], rtol=1e-12, atol=1e-300)


def test_interface_units_cover_the_arguments(golden_units):
    """atc changes the result (acoustic-only and nonlinear-only differ from the
    default), the grid velocity and the second gas are in play."""
    g = golden_units
    q = g["q1d"]
    assert np.mean(np.any(q[1] != q[0], axis=1)) > 0.5 and np.any(q[2] != q[0])
    assert np.all(g["L"][:, 1] != 0) and np.sum(g["R"][:, 0] != g["L"][:, 0]) == len(g["L"]) // 2
    # the flux's side effects: rotated velocities on both structs, lambda zeroed on the left
    from _libs import IFV_MEMBERS
    iu, iv, ilu = IFV_MEMBERS.index("U"), IFV_MEMBERS.index("V"), IFV_MEMBERS.index("lambda_u")
    assert np.all(g["struct_out_L"][:, ilu] == 0) and np.any(g["struct_in_L"][:, ilu] != 0)
    assert np.any(g["struct_out_R"][:, iu] != g["struct_in_R"][:, iu])
    sp_in = np.hypot(g["struct_in_R"][:, iu], g["struct_in_R"][:, iv])
    sp_out = np.hypot(g["struct_out_R"][:, iu], g["struct_out_R"][:, iv])
    assert np.allclose(sp_in, sp_out, rtol=1e-14)       # a rotation keeps the speed


def test_fixture_covers_every_branch(golden_iface):
    """The vectors exercise shocks and rarefactions on both sides, the acoustic
    shortcut, equal states and supersonic (trivial) interfaces."""
    g = golden_iface
    crw = g["riemann"][:, 2:4]
    assert set(map(tuple, crw.astype(int))) >= {(0, 0), (0, 1), (1, 0), (1, 1)}
    L, R = g["L"], g["R"]
    dist = np.sqrt((L[:, 3] - R[:, 3]) ** 2 + (L[:, 4] - R[:, 4]) ** 2 + (L[:, 6] - R[:, 6]) ** 2)
    assert (dist < EPS).sum() > 50 and (dist == 0).sum() > 50
    ws = g["q1d"][:, :2]
    assert (ws[:, 0] > 0).sum() > 20 and (ws[:, 1] < 0).sum() > 20
    assert (g["flux_x"][:, 0] != 0).sum() >= 0   # failure codes recorded as data


def test_split2d_four_quadrant_64(oracle, golden_2d):
    g = golden_2d
    fr, npl, tp, cfg, st = oracle.split2d(64, 64, cases.four_quadrant(64, 64), cases.four_quadrant_config(64, 64))
    assert st == 0
    assert (cfg[5], npl, tp[npl - 1]) == tuple(g["fq64_meta"])
    assert int(cfg[5]) == 127                     # frozen-dt step count of the survey's run
    assert _eq(fr[:, npl - 1], g["fq64_final"])


def test_split2d_frames(oracle, golden_2d):
    g = golden_2d
    tpin = np.array([0.0, 0.05, 0.1, 0.15, 0.2])
    fr, npl, tp, cfg, st = oracle.split2d(32, 48, cases.four_quadrant(32, 48),
                                          cases.four_quadrant_config(32, 48, t_all=0.2), time_plot=tpin)
    meta = g["frames_meta"]
    assert (cfg[5], npl) == (meta[0], meta[1])
    assert _eq(tp, meta[2:])
    assert _eq(fr, g["frames_all"])


def test_split2d_live_frame_zero(oracle, golden_2d):
    g = golden_2d
    init = dict(zip(("rho", "u", "v", "p"), g["late_init"]))
    cfg = make_config(t_all=0.3, h_x=1 / 24, h_y=1 / 20, bc_x=-4, bc_y=-7, split=1)
    fr, npl, tp, c, st = oracle.split2d(24, 20, init, cfg, time_plot=np.array([0.1, 0.2, 0.3]), N_T=3)
    meta = g["late_meta"]
    assert (c[5], npl) == (meta[0], meta[1]) and _eq(tp, meta[2:])
    assert _eq(fr, g["late_frames"])


def test_split2d_all_boundary_conditions(oracle, golden_2d):
    g = golden_2d
    init = dict(zip(("rho", "u", "v", "p"), g["bc_init"]))
    for k, (bx, by, steps, npl, t_end) in enumerate(g["bc_meta"]):
        cfg = make_config(t_all=0.05, h_x=1 / 12, h_y=1 / 10, bc_x=bx, bc_y=by, split=1)
        fr, n_plot, tp, c, st = oracle.split2d(12, 10, init, cfg)
        assert (c[5], n_plot, tp[n_plot - 1]) == (steps, npl, t_end), (bx, by)
        assert _eq(fr[:, n_plot - 1], g["bc_final"][k]), (bx, by)


def test_split2d_sod_and_fixed_tau(oracle, golden_2d):
    g = golden_2d
    m, n = 100, 4
    x = (np.arange(m) + 0.5) / m
    init = {"rho": np.repeat(np.where(x < 0.5, 1.0, 0.125)[:, None], n, 1), "u": np.zeros((m, n)),
            "v": np.zeros((m, n)), "p": np.repeat(np.where(x < 0.5, 1.0, 0.1)[:, None], n, 1)}
    fr, npl, tp, c, st = oracle.split2d(m, n, init, make_config(t_all=0.2, h_x=0.01, h_y=0.01, bc_x=-4, bc_y=-4, split=1))
    assert int(c[5]) == 53 == int(g["sod2d_meta"][0])
    assert _eq(fr[:, npl - 1], g["sod2d_final"])

    cfg = make_config(N=7, tau=1e-3, h_x=1 / 16, h_y=1 / 16, bc_x=-4, bc_y=-7, split=1)
    cfg[1] = np.inf
    init = dict(zip(("rho", "u", "v", "p"), g["fixtau_init"]))
    fr, npl, tp, c, st = oracle.split2d(16, 16, init, cfg, time_plot=np.array([0.0, np.inf]))
    assert (c[5], npl, tp[npl - 1]) == tuple(g["fixtau_meta"])
    assert int(c[5]) == 8                         # N = 7 steps leave k = N + 1
    assert _eq(fr[:, npl - 1], g["fixtau_final"])


def test_split2d_four_quadrant_256(oracle, golden_2d):
    g = golden_2d
    fr, npl, tp, c, st = oracle.split2d(256, 256, cases.four_quadrant(256, 256), cases.four_quadrant_config(256, 256))
    fin = fr[:, npl - 1]
    assert (c[5], npl, tp[npl - 1]) == tuple(g["fq256_meta"])
    probes = g["fq256_probes"].astype(int)
    assert _eq(fin[:, probes[:, 0], probes[:, 1]], g["fq256_probe_vals"])
    assert _eq(fin[:, ::37, :], g["fq256_rows"])
    assert _eq(fin.reshape(5, -1).sum(axis=1), g["fq256_sum"])


def test_unsplit2d_four_quadrant_and_frames(oracle, golden_unsplit):
    g = golden_unsplit
    fr, npl, tp, cfg, st = oracle.unsplit2d(64, 64, cases.four_quadrant(64, 64),
                                            cases.four_quadrant_config(64, 64, split=0))
    assert st == 0 and (cfg[5], npl, tp[npl - 1]) == tuple(g["fq64_meta"])
    assert _eq(fr[:, npl - 1], g["fq64_final"])
    tpin = np.array([0.0, 0.05, 0.1, 0.15, 0.2])
    fr, npl, tp, cfg, st = oracle.unsplit2d(32, 48, cases.four_quadrant(32, 48),
                                            cases.four_quadrant_config(32, 48, t_all=0.2, split=0), time_plot=tpin)
    meta = g["frames_meta"]
    assert (cfg[5], npl) == (meta[0], meta[1]) and _eq(tp, meta[2:])
    assert _eq(fr, g["frames_all"])


def test_unsplit2d_live_frame_zero(oracle, golden_unsplit):
    """First plot time after t = 0: the CFL bound reads frame 0 while it is the live state."""
    g = golden_unsplit
    init = dict(zip(("rho", "u", "v", "p"), g["late_init"]))
    cfg = make_config(t_all=0.3, h_x=1 / 24, h_y=1 / 20, bc_x=-4, bc_y=-7, split=0)
    fr, npl, tp, c, st = oracle.unsplit2d(24, 20, init, cfg, time_plot=np.array([0.1, 0.2, 0.3]), N_T=3)
    meta = g["late_meta"]
    assert (c[5], npl) == (meta[0], meta[1]) and _eq(tp, meta[2:])
    assert _eq(fr, g["late_frames"])


def test_unsplit2d_all_boundary_conditions(oracle, golden_unsplit):
    g = golden_unsplit
    init = dict(zip(("rho", "u", "v", "p"), g["bc_init"]))
    for k, (bx, by, steps, npl, t_end) in enumerate(g["bc_meta"]):
        cfg = make_config(t_all=0.2, h_x=1 / 12, h_y=1 / 10, bc_x=bx, bc_y=by, split=0)
        fr, n_plot, tp, c, st = oracle.unsplit2d(12, 10, init, cfg)
        assert (c[5], n_plot, tp[n_plot - 1]) == (steps, npl, t_end), (bx, by)
        assert _eq(fr[:, n_plot - 1], g["bc_final"][k]), (bx, by)


def test_unsplit2d_blast_and_fixed_tau(oracle, golden_unsplit):
    g = golden_unsplit
    cfg = make_config(t_all=0.15, h_x=1 / 40, h_y=1 / 36, bc_x=-2, bc_y=-2, split=0)
    fr, npl, tp, c, st = oracle.unsplit2d(40, 36, cases.blast_box(40, 36), cfg)
    assert (c[5], npl, tp[npl - 1]) == tuple(g["blast_meta"])
    assert _eq(fr[:, npl - 1], g["blast_final"])
    cfg = make_config(N=7, tau=1e-3, h_x=1 / 16, h_y=1 / 16, bc_x=-24, bc_y=-1, split=0)
    cfg[1] = np.inf
    init = dict(zip(("rho", "u", "v", "p"), g["fixtau_init"]))
    fr, npl, tp, c, st = oracle.unsplit2d(16, 16, init, cfg, time_plot=np.array([0.0, np.inf]))
    assert (c[5], npl, tp[npl - 1]) == tuple(g["fixtau_meta"]) and int(c[5]) == 8
    assert _eq(fr[:, npl - 1], g["fixtau_final"])


def test_unsplit2d_four_quadrant_256(oracle, golden_unsplit):
    g = golden_unsplit
    fr, npl, tp, c, st = oracle.unsplit2d(256, 256, cases.four_quadrant(256, 256),
                                          cases.four_quadrant_config(256, 256, split=0))
    fin = fr[:, npl - 1]
    assert (c[5], npl, tp[npl - 1]) == tuple(g["fq256_meta"])
    probes = g["fq256_probes"].astype(int)
    assert _eq(fin[:, probes[:, 0], probes[:, 1]], g["fq256_probe_vals"])
    assert _eq(fin[:, ::37, :], g["fq256_rows"])
    assert _eq(fin.reshape(5, -1).sum(axis=1), g["fq256_sum"])


def test_unsplit_differs_from_split(oracle):
    """The transversal terms and the single update are really there: the two
    schemes agree to discretisation error, not to rounding."""
    m = n = 32
    init = cases.four_quadrant(m, n)
    fs = oracle.split2d(m, n, init, cases.four_quadrant_config(m, n, t_all=0.1))[0]
    fu = oracle.unsplit2d(m, n, init, cases.four_quadrant_config(m, n, t_all=0.1, split=0))[0]
    d = np.abs(fs[0, -1] - fu[0, -1]).max()
    assert 1e-4 < d < 0.5


def test_grp1d_sod_and_boundaries(oracle, golden_1d):
    g = golden_1d
    cfg = make_config(t_all=0.2, h_x=1e-3, bc_x=-4, CFL=0.45)
    fr, npl, tp, c, st = oracle.grp1d(1000, cases.sod_1d(1000), cfg)
    assert (c[5], npl, tp[npl - 1]) == tuple(g["sod1000_meta"])
    assert int(c[5]) == 975
    assert _eq(fr[:, npl - 1], g["sod1000_final"])
    init = dict(zip(("rho", "u", "p"), g["bc_init"]))
    for k, (bc, steps, npl0, t_end) in enumerate(g["bc_meta"]):
        cfg = make_config(t_all=0.1, h_x=1 / 200, bc_x=bc, CFL=0.45)
        fr, npl, tp, c, st = oracle.grp1d(200, init, cfg)
        assert (c[5], npl, tp[npl - 1]) == (steps, npl0, t_end), bc
        assert _eq(fr[:, npl - 1], g["bc_final"][k]), bc


def test_grp1d_plot_frames(oracle, golden_1d):
    """Frames at plot times, and a first plot time after t = 0 (the loop runs in
    frame 0 until then): every frame equals the reference's bit for bit."""
    g = golden_1d
    cfg = make_config(t_all=0.2, h_x=1 / 400, bc_x=-2, CFL=0.45)
    fr, npl, tp, c, st = oracle.grp1d(400, cases.sod_1d(400), cfg, N_T=5,
                                      time_plot=np.array([0.0, 0.05, 0.1, 0.15, 0.2]))
    assert np.array_equal(np.r_[c[5], npl, tp[:npl]], g["frames_meta"])
    assert _eq(fr[:, :npl], g["frames_all"])
    init = dict(zip(("rho", "u", "p"), g["late_init"]))
    cfg = make_config(t_all=0.16, h_x=1 / 150, bc_x=-7, CFL=0.45)
    fr, npl, tp, c, st = oracle.grp1d(150, init, cfg, N_T=3, time_plot=np.array([0.04, 0.1, 0.16]))
    assert np.array_equal(np.r_[c[5], npl, tp[:npl]], g["late_meta"])
    assert _eq(fr[:, :npl], g["late_frames"])
    # the frames differ from each other (each is a different time)
    assert not np.array_equal(g["frames_all"][:, 1], g["frames_all"][:, 2])


def test_g2d_bitexact(oracle, golden_g2d):
    """linear_GRP_solver_Edir_G2D of the oracle equals the reference's
    single-fluid build and its -DMULTIFLUID_BASICS build bit for bit, at
    atc = eps, INFINITY and -0.0."""
    g = golden_g2d
    L, R = g["L"], g["R"]
    bad = []
    for a, atc in enumerate(g["atc"]):
        for k in range(len(L)):
            if not _eq(oracle.g2d(L[k, :15], R[k, :15], EPS, atc), g["single"][a, k]):
                bad.append(("single", a, k))
            if not _eq(oracle.g2d(L[k], R[k], EPS, atc), g["multi"][a, k]):
                bad.append(("multi", a, k))
    assert not bad, bad[:10]


def test_unsplit2d_g2d_bitexact(oracle, golden_unsplit_g2d):
    """The unsplit step with GRP_2D_flux's alternative solver (G2D, flux_solver.c:189)
    equals the reference built that way (oracle/_ref/libhydroref_g2d.so) bit for
    bit: four-quadrant 64^2 to t = 0.3, all 25 boundary-condition pairs, a blast
    in a reflective box."""
    g = golden_unsplit_g2d
    m = n = 64
    fo, npo, tpo, co, st = oracle.unsplit2d(m, n, cases.four_quadrant(m, n), cases.four_quadrant_config(m, n, split=0),
                                            g2d=True)
    assert (co[5], npo, tpo[npo - 1]) == tuple(g["fq64_meta"]) and st == 0
    assert _eq(fo[:, npo - 1], g["fq64_final"])
    init = dict(zip(("rho", "u", "v", "p"), g["bc_init"]))
    for k, (bx, by, steps, npl, t_end) in enumerate(g["bc_meta"]):
        cfg = make_config(t_all=0.2, h_x=1 / 12, h_y=1 / 10, bc_x=bx, bc_y=by, split=0)
        fo, npo, tpo, co, st = oracle.unsplit2d(12, 10, init, cfg, g2d=True)
        assert (co[5], npo, tpo[npo - 1]) == (steps, npl, t_end) and _eq(fo[:, npo - 1], g["bc_final"][k]), (bx, by)
    m, n = 40, 36
    cfg = make_config(t_all=0.15, h_x=1 / m, h_y=1 / n, bc_x=-2, bc_y=-2, split=0)
    fo, npo, tpo, co, st = oracle.unsplit2d(m, n, cases.blast_box(m, n), cfg, g2d=True)
    assert (co[5], npo, tpo[npo - 1]) == tuple(g["blast_meta"]) and _eq(fo[:, npo - 1], g["blast_final"])


def test_unsplit2d_g2d_differs_from_q1d(oracle, golden_unsplit, golden_unsplit_g2d):
    """The two interface solvers give two schemes: the G2D run of the
    four-quadrant problem is not the Q1D run (the fixtures are not swapped)."""
    d = np.abs(golden_unsplit_g2d["fq64_final"] - golden_unsplit["fq64_final"]).max()
    assert d > 1e-3


def g2d_branches(L, R, out, atc):
    """Which branch of linear_GRP_solver_Edir_G2D produced each row, read off
    its inputs and outputs: 'upwind', 'acoustic', 'fan', or 'star-<side>-<wave>'
    (the contact's side, L or R, and that side's wave, crw or shock)."""
    lu = L[:, 1]
    cL = np.sqrt(L[:, 0] * L[:, 6] / L[:, 3])
    cR = np.sqrt(R[:, 0] * R[:, 6] / R[:, 3])
    dist = np.sqrt(((L[:, 3:7] - R[:, 3:7]) ** 2).sum(axis=1))
    ws, us = out[:, 0:2], out[:, 14:20]
    crw_L = ws[:, 0] == L[:, 4] - cL
    crw_R = ws[:, 1] == R[:, 4] + cR
    u_star = us[:, 1]
    fan = (crw_L & (u_star - us[:, 4] > lu)) | (crw_R & (u_star + us[:, 5] < lu))
    names = []
    for k in range(len(L)):
        if ws[k, 0] > lu[k] or ws[k, 1] < lu[k]:
            names.append("upwind")
        elif dist[k] < atc:
            names.append("acoustic")
        elif fan[k]:
            names.append("fan")
        else:
            left = not u_star[k] < lu[k]
            crw = crw_L[k] if left else crw_R[k]
            names.append(f"star-{'L' if left else 'R'}-{'crw' if crw else 'shock'}")
    return np.array(names), crw_L, crw_R


def test_g2d_fixture_covers_every_branch(golden_g2d):
    """The G2D vectors reach every branch of the solver in the nonlinear mode
    (atc = eps), the rarefaction's gamma = 3 limit, and the contact-side
    rarefaction on the left with a rarefaction on the right (where the
    reference's D[2] reads the right wave's c_frac)."""
    g = golden_g2d
    L, R = g["L"], g["R"]
    names, crw_L, crw_R = g2d_branches(L, R, g["single"][0], EPS)
    counts = {n: int((names == n).sum()) for n in np.unique(names)}
    print(counts)
    for n in ("upwind", "acoustic", "fan", "star-L-crw", "star-L-shock", "star-R-crw", "star-R-shock"):
        assert counts.get(n, 0) >= 5, n
    assert ((names == "star-L-crw") & crw_R).sum() >= 5
    g3 = L[:, 0] == 3.0
    assert (g3 & np.isin(names, ["fan", "star-L-crw", "star-R-crw"])).sum() >= 5


def test_against_live_reference(oracle, reflib):
    """Fresh random vectors and runs (not in the fixtures) against the reference build."""
    L, R = cases.interface_pairs(3000, seed=99)
    for k in range(len(L)):
        assert _eq(oracle.q1d(L[k], R[k], EPS, EPS), reflib.q1d(L[k], R[k], EPS, EPS)), k
        rc, o = oracle.flux2d(L[k], R[k], 1e-3, EPS, 0.0, 1.0)
        rr, r = reflib.flux2d(L[k], R[k], 1e-3, EPS, 0.0, 1.0)
        assert rc == rr and (rc != 0 or _eq(o, r)), k
        assert _eq(oracle.edir(L[k], R[k], EPS, EPS), reflib.edir(L[k], R[k], EPS, EPS)), k
        assert _eq(oracle.g2d(L[k], R[k], EPS, EPS), reflib.g2d(L[k], R[k], EPS, EPS)), k
    for (m, n, seed) in ((7, 9, 1), (1, 13, 2), (13, 1, 3), (33, 31, 4)):
        cfg = make_config(t_all=0.03, h_x=1 / m, h_y=1 / n, bc_x=-7, bc_y=-2, split=1)
        init = cases.smooth_random(m, n, seed)
        fo, npo, tpo, co, st = oracle.split2d(m, n, init, cfg)
        fr, npr, tpr, cr = reflib.split2d(m, n, init, cfg)
        assert co[5] == cr[5] and _eq(fo[:, :npo], fr[:, :npr]), (m, n)
        cfg = make_config(t_all=0.03, h_x=1 / m, h_y=1 / n, bc_x=-2, bc_y=-7, split=0)
        fo, npo, tpo, co, st = oracle.unsplit2d(m, n, init, cfg)
        fr, npr, tpr, cr = reflib.unsplit2d(m, n, init, cfg)
        assert co[5] == cr[5] and _eq(fo[:, :npo], fr[:, :npr]), ("unsplit", m, n)
