"""GPU parity of the unsplit scheme (GRP_solver_2D_EUL_source, config[33] = 0):
the HIP library against the golden fixtures made by the reference itself, and
against the oracle on the same seeded inputs.

Bar: relative L-infinity difference <= 1e-10 at final time, per cell for rho,
p and E (conftest.parity), as for the split path (the device's FP64 pow may
differ from glibc's in the last bit; everything else follows the reference's
expression order).
"""

import numpy as np
import pytest

import cases
from _libs import make_config
from conftest import parity, rel_linf

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def hg():
    import hydrogrp
    hydrogrp.lib()
    print("device:", hydrogrp.device_name(), flush=True)
    return hydrogrp


def _run(hg, m, n, init, cfg, time_plot=None, N_T=None, g2d=False):
    print(f"unsplit run {m}x{n} bc=({cfg[17]:g},{cfg[18]:g}){' G2D' if g2d else ''}", flush=True)
    return hg.run_unsplit2d(m, n, init, cfg, time_plot=time_plot, N_T=N_T, g2d=g2d)


def test_four_quadrant_64_and_frames(hg, golden_unsplit):
    g = golden_unsplit
    fr, npl, tp, cfg, res = _run(hg, 64, 64, cases.four_quadrant(64, 64), cases.four_quadrant_config(64, 64, split=0))
    assert (cfg[5], npl, tp[npl - 1]) == tuple(g["fq64_meta"]) and res.status == 0
    d = parity(fr[:, npl - 1], g["fq64_final"])
    print("fq64 rel Linf per field:", d, flush=True)
    assert d.max() <= TOL
    tpin = np.array([0.0, 0.05, 0.1, 0.15, 0.2])
    fr, npl, tp, cfg, res = _run(hg, 32, 48, cases.four_quadrant(32, 48),
                                 cases.four_quadrant_config(32, 48, t_all=0.2, split=0), time_plot=tpin)
    meta = g["frames_meta"]
    assert (cfg[5], npl) == (meta[0], meta[1]) and np.array_equal(tp, meta[2:])
    for k in range(npl):
        assert parity(fr[:, k], g["frames_all"][:, k]).max() <= TOL, k


def test_live_frame_zero(hg, golden_unsplit):
    """First plot time after t = 0: the CFL bound follows frame 0 while it is the live state."""
    g = golden_unsplit
    init = dict(zip(("rho", "u", "v", "p"), g["late_init"]))
    cfg = make_config(t_all=0.3, h_x=1 / 24, h_y=1 / 20, bc_x=-4, bc_y=-7, split=0)
    fr, npl, tp, c, res = _run(hg, 24, 20, init, cfg, time_plot=np.array([0.1, 0.2, 0.3]), N_T=3)
    meta = g["late_meta"]
    assert (c[5], npl) == (meta[0], meta[1])
    assert rel_linf(tp, meta[2:]).max() <= 1e-13
    for k in range(npl):
        assert parity(fr[:, k], g["late_frames"][:, k]).max() <= TOL, k


def test_all_boundary_conditions(hg, golden_unsplit):
    g = golden_unsplit
    init = dict(zip(("rho", "u", "v", "p"), g["bc_init"]))
    worst = 0.0
    for k, (bx, by, steps, npl0, t_end) in enumerate(g["bc_meta"]):
        cfg = make_config(t_all=0.2, h_x=1 / 12, h_y=1 / 10, bc_x=bx, bc_y=by, split=0)
        fr, npl, tp, c, res = _run(hg, 12, 10, init, cfg)
        assert (c[5], npl, tp[npl - 1]) == (steps, npl0, t_end), (bx, by)
        d = parity(fr[:, npl - 1], g["bc_final"][k]).max()
        worst = max(worst, d)
        assert d <= TOL, (bx, by, d)
    print("boundary conditions: worst rel Linf", worst, flush=True)


def test_blast_and_fixed_tau(hg, golden_unsplit):
    g = golden_unsplit
    cfg = make_config(t_all=0.15, h_x=1 / 40, h_y=1 / 36, bc_x=-2, bc_y=-2, split=0)
    fr, npl, tp, c, res = _run(hg, 40, 36, cases.blast_box(40, 36), cfg)
    assert (c[5], npl, tp[npl - 1]) == tuple(g["blast_meta"])
    assert parity(fr[:, npl - 1], g["blast_final"]).max() <= TOL
    cfg = make_config(N=7, tau=1e-3, h_x=1 / 16, h_y=1 / 16, bc_x=-24, bc_y=-1, split=0)
    cfg[1] = np.inf
    init = dict(zip(("rho", "u", "v", "p"), g["fixtau_init"]))
    fr, npl, tp, c, res = _run(hg, 16, 16, init, cfg, time_plot=np.array([0.0, np.inf]))
    assert (c[5], npl, tp[npl - 1]) == tuple(g["fixtau_meta"])
    assert parity(fr[:, npl - 1], g["fixtau_final"]).max() <= TOL


def test_four_quadrant_256(hg, golden_unsplit):
    g = golden_unsplit
    fr, npl, tp, c, res = _run(hg, 256, 256, cases.four_quadrant(256, 256),
                               cases.four_quadrant_config(256, 256, split=0))
    fin = fr[:, npl - 1]
    assert (c[5], npl, tp[npl - 1]) == tuple(g["fq256_meta"])
    probes = g["fq256_probes"].astype(int)
    d_rows = parity(fin[:, ::37, :], g["fq256_rows"])
    d_probe = parity(fin[:, probes[:, 0], probes[:, 1]], g["fq256_probe_vals"])
    print("fq256 rel Linf rows", d_rows.max(), "probes", d_probe.max(), flush=True)
    assert d_rows.max() <= TOL and d_probe.max() <= TOL
    assert np.allclose(fin.reshape(5, -1).sum(axis=1), g["fq256_sum"], rtol=1e-12, atol=1e-9)


# tile edges: 62 cells per wave in y, 10 rows per block in x
@pytest.mark.parametrize("g2d", [False, True], ids=["Q1D", "G2D"])
@pytest.mark.parametrize("m,n,bx,by", [(1, 50, -4, -7), (50, 1, -7, -2), (10, 62, -2, -24), (11, 63, -24, -1),
                                       (21, 125, -7, -7), (127, 129, -1, -4), (65, 200, -2, -2)])
def test_odd_shapes_vs_oracle(hg, oracle, m, n, bx, by, g2d):
    init = cases.smooth_random(m, n, m * 1000 + n)
    cfg = make_config(t_all=0.04, h_x=1 / m, h_y=1 / n, bc_x=bx, bc_y=by, split=0)
    fo, npo, tpo, co, st = oracle.unsplit2d(m, n, init, cfg, g2d=g2d)
    fg, npg, tpg, cg, res = _run(hg, m, n, init, cfg, g2d=g2d)
    assert (co[5], npo, tpo[npo - 1]) == (cg[5], npg, tpg[npg - 1])
    assert parity(fg[:, npg - 1], fo[:, npo - 1]).max() <= TOL


@pytest.mark.parametrize("g2d", [False, True], ids=["Q1D", "G2D"])
def test_error_paths_vs_oracle(hg, oracle, g2d):
    m, n = 8, 6
    for mutate, kw in ((("p", (3, 2), 1e-12), {}), (("p", (3, 2), -1.0), {}), (None, {"CFL": 1e-12}),
                       (("rho", (5, 1), 1e-12), {}), (None, {"bc_y": -3}), (("u", (0, 0), np.nan), {})):
        init = cases.smooth_random(m, n, 2)
        if mutate:
            init[mutate[0]][mutate[1]] = mutate[2]
        base = dict(t_all=0.1, h_x=1 / m, h_y=1 / n, bc_x=-4, bc_y=-4, split=0)
        base.update(kw)
        cfg = make_config(**base)
        fo, npo, tpo, co, st = oracle.unsplit2d(m, n, init, cfg, g2d=g2d)
        fg, npg, tpg, cg, res = _run(hg, m, n, init, cfg, g2d=g2d)
        assert (co[5], npo, st) == (cg[5], npg, res.status), (mutate, kw)
        assert tpo[npo - 1] == tpg[npg - 1] or (np.isnan(tpo[npo - 1]) and np.isnan(tpg[npg - 1])), (mutate, kw)
        assert np.array_equal(np.isfinite(fo[:, npo - 1]), np.isfinite(fg[:, npg - 1]))
        ok = np.isfinite(fo[:, npo - 1])
        assert np.max(np.abs(fo[:, npo - 1][ok] - fg[:, npg - 1][ok])) <= 1e-9, (mutate, kw)


def test_conservation_full_size(hg):
    """Periodic in both directions: mass, momentum and energy sums are conserved
    (to rounding) over 24 steps of a 4096 x 4096 four-quadrant state, in the
    default (tile) form."""
    m = n = 4096
    cfg = cases.four_quadrant_config(m, n, split=0, N=24)
    cfg[17] = cfg[18] = -7
    cfg[1] = np.inf
    cfg[16] = 0.45 * (1.0 / m) / 3.0
    init = cases.four_quadrant(m, n)
    s = hg.Solver2D(m, n, cfg)
    try:
        assert s.sweep_kernel(2) == "unsplit_step_kernel<12, 1, 0>"   # the default form
        s.upload(init["rho"], init["u"], init["v"], init["p"])
        r0, u0, v0, p0, e0 = s.download()
        res = s.solve(np.array([0.0, np.inf]), 2)
        r1, u1, v1, p1, e1 = s.download()
    finally:
        s.close()
    assert res.status == 0 and res.k == 25
    for name, a, b in (("mass", r0, r1), ("x-momentum", r0 * u0, r1 * u1), ("y-momentum", r0 * v0, r1 * v1),
                       ("energy", r0 * e0, r1 * e1)):
        s0, s1 = a.sum(), b.sum()
        scale = np.abs(a).sum()
        print(f"{name}: relative drift {(s1 - s0) / scale:.2e}", flush=True)
        assert abs(s1 - s0) <= 1e-11 * scale, name
    assert np.all(np.isfinite(r1)) and r1.min() > 0 and p1.min() > 0


def test_checkpoint_resume_is_bitwise(hg):
    """A checkpoint taken in a plot-frame callback and restored into a fresh
    context continues the unsplit loop bit for bit."""
    m, n = 40, 36
    cfg = cases.four_quadrant_config(m, n, t_all=0.2, split=0)
    init = cases.four_quadrant(m, n)
    tp = np.array([0.0, 0.07, 0.2])
    a = hg.Solver2D(m, n, cfg)
    blob = {}
    try:
        a.upload(init["rho"], init["u"], init["v"], init["p"])

        def grab(nt, ntp, tc):
            if ntp == 1:
                blob["b"] = a.checkpoint()
            return 0

        ra = a.solve(tp, 3, grab)
        full = a.download()
    finally:
        a.close()
    b = hg.Solver2D(m, n, cfg)
    try:
        b.restore(blob["b"])
        rb = b.solve(tp, 3)
        resumed = b.download()
    finally:
        b.close()
    assert (ra.k, ra.nt, ra.status, ra.time_c) == (rb.k, rb.nt, rb.status, rb.time_c)
    assert np.array_equal(np.asarray(full), np.asarray(resumed))


def test_unsplit_rejects_two_fluids(hg):
    """The two-fluid mode runs the split scheme only: an unsplit solve of a
    context with two fluids is refused with an argument error, state untouched."""
    m, n = 16, 12
    cfg = make_config(t_all=0.05, h_x=1 / m, h_y=1 / n, bc_x=-4, bc_y=-2, split=0)
    init = cases.smooth_random(m, n, 4)
    s = hg.Solver2D(m, n, cfg)
    try:
        s.upload(init["rho"], init["u"], init["v"], init["p"])
        s.set_fluids(1.67, np.ones((m, n)))
        before = np.asarray(s.download())
        with pytest.raises(RuntimeError, match="hgrp2d_solve"):
            s.solve(np.array([0.0, 0.05]), 2)
        with pytest.raises(RuntimeError, match="hgrp2d_bench_unsplit"):
            s.bench_unsplit(2)
        assert np.array_equal(before, np.asarray(s.download()))
    finally:
        s.close()


@pytest.mark.parametrize("m,n", [(128, 128), (2048, 4096)])
def test_unsplit_learned_guess_against_oracle(hg, oracle, oracle_devpow, m, n):
    """The unsplit step started from the learned p* guess (Newton tolerance
    1e-12, the converge rule) against the oracle's unsplit step started from the
    reference's guess with the same tolerance and rule (per-cell relative
    L-infinity), and the GPU's plain form against the oracle with the device's
    pow.  128^2 runs to t = 0.3 in the tile form (the default); 2048 x 4096
    takes 3 steps in the strip form."""
    from hydrogrp import mlp
    from conftest import parity
    init = cases.four_quadrant(m, n)
    cfg = cases.four_quadrant_config(m, n, split=0)
    form = None
    if m > 512:
        cfg = cases.four_quadrant_config(m, n, split=0, N=3)
        cfg[1] = np.inf
        form = hg.UNSPLIT_STRIP
        s = hg.Solver2D(m, n, cfg)
        s.set_newton(1e-12, mlp.load(), converge=True)
        s.set_unsplit_form(form)
        assert s.sweep_kernel(2) == "unsplit_strip_kernel<2>"
        s.close()
    tp = np.array([0.0, cfg[1]])
    fl, npl, tpl, cl, rl = hg.run_unsplit2d(m, n, init, cfg, time_plot=tp, newton_tol=1e-12, guess_params=mlp.load(),
                                            converge=True, form=form)
    fo, npo, tpo, co, so = oracle.unsplit2d(m, n, init, cfg, time_plot=tp, newton_tol=1e-12, converge=True)
    assert rl.status == so == 0 and (cl[5], npl, tpl[npl - 1]) == (co[5], npo, tpo[npo - 1])
    d = parity(fl[:, npl - 1], fo[:, npo - 1])
    del fl, fo
    fp, npp, tpp, cp, rp = hg.run_unsplit2d(m, n, init, cfg, time_plot=tp, newton_tol=1e-12, converge=True, form=form)
    fd, npd, tpd, cd, sd = oracle_devpow.unsplit2d(m, n, init, cfg, time_plot=tp, newton_tol=1e-12, converge=True)
    d_dev = parity(fp[:, npp - 1], fd[:, npd - 1])
    print(f"unsplit {m}x{n}, learned (GPU) vs oracle: {d}; plain (GPU) vs device-pow oracle: {d_dev}", flush=True)
    assert d.max() <= 1e-10 and d_dev.max() <= 1e-13


def test_unsplit_g2d_vs_reference(hg, golden_unsplit_g2d):
    """The unsplit step with GRP_2D_flux's alternative interface solver
    (linear_GRP_solver_Edir_G2D, flux_solver.c:189) against the reference
    built that way: four-quadrant 64^2 to t = 0.3, every boundary-condition
    pair, a blast in a reflective box."""
    g = golden_unsplit_g2d
    m = n = 64
    fr, npl, tp, cfg, res = hg.run_unsplit2d(m, n, cases.four_quadrant(m, n), cases.four_quadrant_config(m, n, split=0),
                                             g2d=True)
    assert (cfg[5], npl, tp[npl - 1]) == tuple(g["fq64_meta"]) and res.status == 0
    d = parity(fr[:, npl - 1], g["fq64_final"])
    print("G2D fq64 rel Linf per field:", d, flush=True)
    assert d.max() <= TOL
    init = dict(zip(("rho", "u", "v", "p"), g["bc_init"]))
    worst = 0.0
    for k, (bx, by, steps, npl0, t_end) in enumerate(g["bc_meta"]):
        cfg = make_config(t_all=0.2, h_x=1 / 12, h_y=1 / 10, bc_x=bx, bc_y=by, split=0)
        fr, npl, tp, c, res = hg.run_unsplit2d(12, 10, init, cfg, g2d=True)
        assert (c[5], npl, tp[npl - 1]) == (steps, npl0, t_end), (bx, by)
        worst = max(worst, parity(fr[:, npl - 1], g["bc_final"][k]).max())
    m, n = 40, 36
    cfg = make_config(t_all=0.15, h_x=1 / m, h_y=1 / n, bc_x=-2, bc_y=-2, split=0)
    fr, npl, tp, c, res = hg.run_unsplit2d(m, n, cases.blast_box(m, n), cfg, g2d=True)
    assert (c[5], npl, tp[npl - 1]) == tuple(g["blast_meta"])
    worst = max(worst, parity(fr[:, npl - 1], g["blast_final"]).max())
    print("G2D boundary conditions and blast: worst rel Linf", worst, flush=True)
    assert worst <= TOL


@pytest.mark.parametrize("m,n", [(128, 128), (2048, 4096)])
def test_unsplit_g2d_vs_devpow_oracle(hg, oracle_devpow, m, n):
    """The G2D form against the oracle with the device's pow, per cell: 128^2
    to t = 0.3 in the tile form, 2048 x 4096 for 3 steps in the tile and the
    strip form (the kernel names checked), with the reference's Newton setting
    and, at 128^2, the learned guess under tol 1e-12 and the converge rule
    against the oracle's plain solve with the same setting."""
    from hydrogrp import mlp
    init = cases.four_quadrant(m, n)
    cfg = cases.four_quadrant_config(m, n, split=0)
    if m > 512:
        cfg = cases.four_quadrant_config(m, n, split=0, N=3)
        cfg[1] = np.inf
    tp = np.array([0.0, cfg[1]])
    s = hg.Solver2D(m, n, cfg)
    s.set_interface_solver(True)
    assert s.sweep_kernel(2) == "unsplit_step_g2d_kernel<12, 1, 0>"
    s.set_unsplit_form(hg.UNSPLIT_STRIP)
    assert s.sweep_kernel(2) == "unsplit_strip_g2d_kernel<0>"
    s.close()
    fd, npd, tpd, cd, sd = oracle_devpow.unsplit2d(m, n, init, cfg, time_plot=tp, g2d=True)
    for form in ((None, hg.UNSPLIT_STRIP) if m > 512 else (None,)):
        fg, npg, tpg, cg, rg = hg.run_unsplit2d(m, n, init, cfg, time_plot=tp, g2d=True, form=form)
        assert rg.status == sd == 0 and (cg[5], npg, tpg[npg - 1]) == (cd[5], npd, tpd[npd - 1])
        d = parity(fg[:, npg - 1], fd[:, npd - 1])
        del fg
        print(f"unsplit G2D {m}x{n} ({'strip' if form else 'tile'}) vs device-pow oracle: {d}", flush=True)
        assert d.max() <= 1e-13
    if m <= 512:
        fl, npl, tpl, cl, rl = hg.run_unsplit2d(m, n, init, cfg, time_plot=tp, newton_tol=1e-12,
                                                guess_params=mlp.load(), converge=True, g2d=True)
        fo, npo, tpo, co, so = oracle_devpow.unsplit2d(m, n, init, cfg, time_plot=tp, newton_tol=1e-12,
                                                       converge=True, g2d=True)
        assert rl.status == so == 0 and (cl[5], npl) == (co[5], npo)
        d = parity(fl[:, npl - 1], fo[:, npo - 1])
        print(f"unsplit G2D {m}x{n}, learned guess vs device-pow oracle (tol 1e-12, converge rule): {d}", flush=True)
        assert d.max() <= 1e-10


def test_g2d_solver_refused_for_split(hg):
    """The split sweeps keep the reference's solver: asking a split context
    for G2D is an argument error."""
    m, n = 16, 12
    cfg = make_config(t_all=0.05, h_x=1 / m, h_y=1 / n, bc_x=-4, bc_y=-2, split=1)
    s = hg.Solver2D(m, n, cfg)
    try:
        with pytest.raises(RuntimeError, match="hgrp2d_set_interface_solver"):
            s.set_interface_solver(True)
        s.set_interface_solver(False)
    finally:
        s.close()


@pytest.mark.parametrize("g2d", [False, True], ids=["Q1D", "G2D"])
def test_unsplit_forms_and_tuner_change_no_result(hg, g2d):
    """The tile and strip forms of the unsplit step give the same state bit
    for bit (a 200 x 150 grid: partial tiles and strips), and the tuner picks
    one of them, puts the state back, and the run goes on as without it."""
    m, n = 200, 150
    cfg = cases.four_quadrant_config(m, n, split=0, t_all=0.1)
    init = cases.four_quadrant(m, n)
    tp = np.array([0.0, 0.1])
    outs = []
    for form, want in ((hg.UNSPLIT_TILE, "unsplit_step"), (hg.UNSPLIT_STRIP, "unsplit_strip"), (None, None)):
        s = hg.Solver2D(m, n, cfg)
        try:
            s.set_interface_solver(g2d)
            s.upload(init["rho"], init["u"], init["v"], init["p"])
            if form is None:
                picked, yv, ms, _ = s.tune(3)
                print(f"unsplit tuner ({'G2D' if g2d else 'Q1D'}): form {picked}, {ms:.4f} ms per step", flush=True)
                assert picked in (hg.UNSPLIT_TILE, hg.UNSPLIT_STRIP) and yv == -1 and ms > 0
            else:
                s.set_unsplit_form(form)
                assert s.sweep_kernel(2).startswith(want)
            res = s.solve(tp, 2)
            outs.append((res.k, res.time_c, np.asarray(s.download())))
        finally:
            s.close()
    for k, t, st in outs[1:]:
        assert (k, t) == outs[0][:2] and np.array_equal(st, outs[0][2])
