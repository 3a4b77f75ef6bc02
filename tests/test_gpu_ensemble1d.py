"""GPU parity of the batched 1D GRP engine (hgrp1d_*) with the reference's 1D
path: the golden Sod-1000 and boundary fixtures, ensembles of uneven problems
against the oracle instance by instance, stop/abort paths, and the launch
forms (captured graphs, block sizes, autotuner) giving identical results."""

import numpy as np
import pytest

import cases
from _libs import make_config
from conftest import rel_linf

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def hg():
    import hydrogrp
    hydrogrp.lib()
    return hydrogrp


def _cfg(t_end, bc=-4, **kw):
    return make_config(t_all=t_end, bc_x=bc, CFL=0.45, **kw)


def test_sod_1000(hg, golden_1d):
    g = golden_1d
    cfg = make_config(t_all=0.2, h_x=1e-3, bc_x=-4, CFL=0.45)
    fin, npl, tp, c, st = hg.run_grp1d(1000, cases.sod_1d(1000), cfg)
    assert (c[5], npl, tp[npl - 1]) == tuple(g["sod1000_meta"]) and st == 0
    d = rel_linf(fin, g["sod1000_final"])
    print("Sod-1000 rel Linf", d, flush=True)
    assert d.max() <= TOL


def test_boundary_conditions(hg, golden_1d):
    g = golden_1d
    init = dict(zip(("rho", "u", "p"), g["bc_init"]))
    for k, (bc, steps, npl0, t_end) in enumerate(g["bc_meta"]):
        cfg = make_config(t_all=0.1, h_x=1 / 200, bc_x=bc, CFL=0.45)
        fin, npl, tp, c, st = hg.run_grp1d(200, init, cfg)
        assert (c[5], npl, tp[npl - 1]) == (steps, npl0, t_end), bc
        assert rel_linf(fin, g["bc_final"][k]).max() <= TOL, bc


def _oracle_each(oracle, cells, h, rho, u, p, cfg):
    off = np.concatenate([[0], np.cumsum(cells)])
    out = []
    for i, m in enumerate(cells):
        c = cfg.copy()
        c[10] = h[i]
        sl = slice(off[i], off[i + 1])
        fr, npl, tp, cc, st = oracle.grp1d(int(m), {"rho": rho[sl], "u": u[sl], "p": p[sl]}, c)
        out.append((fr[:, npl - 1], int(cc[5]), st, tp[npl - 1]))
    return off, out


def _final_diff(got, f):
    ok = np.isfinite(f)
    assert np.array_equal(ok, np.isfinite(got))
    if not ok.any():
        return 0.0
    a, b = np.where(ok, got, 0.0), np.where(ok, f, 0.0)
    return rel_linf(a, b).max()


def test_ensemble_matches_oracle_per_instance(hg, oracle, oracle_devpow):
    """Every instance against the oracle (glibc pow, as the reference) and
    against the oracle whose pow is the device's.  The GPU must match the
    second to 1e-12: that pins every operation but pow.  Against the first
    the bar is 1e-10, except where the two oracles themselves differ by more
    (a strong-shock run can amplify last-bit pow differences through limiter
    switches); there the GPU must sit exactly where the device-pow oracle is."""
    cells, h, rho, u, p = cases.ensemble_1d(48, seed=5, n_lo=40, n_hi=400)
    cfg = _cfg(0.15)
    ens = hg.Ensemble1D(cells, h, cfg)
    ens.upload(rho, u, p)
    ms, cu = ens.run()
    fin = ens.download()
    steps, status, t = ens.results()
    ens.close()
    off, ref = _oracle_each(oracle, cells, h, rho, u, p, cfg)
    _, ref_dp = _oracle_each(oracle_devpow, cells, h, rho, u, p, cfg)
    worst, worst_dp, n_stop, amplified = 0.0, 0.0, 0, []
    for i, ((f, k, st, te), (fd, kd, sd, ted)) in enumerate(zip(ref, ref_dp)):
        got = fin[:, off[i]:off[i + 1]]
        assert (steps[i], status[i], t[i]) == (kd, sd, ted), (i, steps[i], status[i], kd, sd)
        d_dp = _final_diff(got, fd)
        worst_dp = max(worst_dp, d_dp)
        assert d_dp <= 1e-12, (i, d_dp)
        n_stop += st != 0
        oracles = _final_diff(fd, f) if (k, st) == (kd, sd) else np.inf
        if oracles > TOL:
            amplified.append((i, oracles))
            continue
        # tau comes from the state every step, so pow's last bit reaches the time too
        assert (steps[i], status[i]) == (k, st) and abs(t[i] - te) <= 1e-12 * te, i
        d = _final_diff(got, f)
        worst = max(worst, d)
        assert d <= TOL, (i, d)
    print(f"48 instances ({cells.sum()} cells, {cu} cell-updates, {ms:.1f} ms): rel Linf vs oracle "
          f"{worst:.2e}, vs device-pow oracle {worst_dp:.2e}; {n_stop} ended by a check; "
          f"amplified pow differences (oracle vs oracle): {amplified}", flush=True)
    assert len(amplified) <= 2


# interface counts (m + 1) on and around the 64-lane waves of the interface
# kernel, cell counts around those of the update kernel
EDGE_SIZES = np.array([1, 2, 62, 63, 64, 65, 126, 127, 128, 129])


@pytest.mark.parametrize("bc", [-4, -7, -2, -24, -1])
def test_wave_edges_against_device_pow_oracle(hg, oracle_devpow, bc):
    """Instances whose sizes put their first and last interfaces and cells on
    the lanes where a wave's neighbours come from loads instead of the next
    lane, for every boundary condition: each instance against the device-pow
    oracle to 1e-12, with equal step counts, status and end time."""
    rng = np.random.default_rng(11 + abs(bc))
    rho, u, p = [], [], []
    for m in EDGE_SIZES:
        x = (np.arange(m) + 0.5) / m
        left = x < 0.5 + 0.1 * rng.uniform(-1, 1)
        rho.append(np.where(left, 1.0, 10 ** rng.uniform(-1, 0)))
        u.append(np.where(left, 0.0, rng.uniform(-0.5, 0.5)))
        p.append(np.where(left, 1.0, 10 ** rng.uniform(-1.5, 0)))
    rho, u, p = (np.concatenate(a) for a in (rho, u, p))
    h = 1.0 / EDGE_SIZES
    cfg = _cfg(0.1, bc=bc)
    ens = hg.Ensemble1D(EDGE_SIZES, h, cfg)
    ens.upload(rho, u, p)
    ens.run()
    fin = ens.download()
    steps, status, t = ens.results()
    ens.close()
    off, ref = _oracle_each(oracle_devpow, EDGE_SIZES, h, rho, u, p, cfg)
    worst = 0.0
    for i, (f, k, st, te) in enumerate(ref):
        assert (steps[i], status[i], t[i]) == (k, st, te), (bc, EDGE_SIZES[i], steps[i], status[i], k, st)
        d = _final_diff(fin[:, off[i]:off[i + 1]], f)
        worst = max(worst, d)
        assert d <= 1e-12, (bc, EDGE_SIZES[i], d)
    print(f"BC {bc}: wave-edge instances vs device-pow oracle {worst:.2e}", flush=True)


def test_launch_forms_agree_bitwise(hg):
    cells, h, rho, u, p = cases.ensemble_1d(24, seed=8, n_lo=100, n_hi=3000)
    cfg = _cfg(0.05)
    outs = []
    for graph, blocks in ((True, None), (False, None), (True, (64, 512)), (True, (512, 64))):
        ens = hg.Ensemble1D(cells, h, cfg)
        if blocks:
            ens.set_blocks(*blocks)
        ens.upload(rho, u, p)
        ens.run(graph=graph)
        outs.append((ens.download(), ens.results()))
        ens.close()
    for fin, res in outs[1:]:
        assert np.array_equal(fin, outs[0][0], equal_nan=True)
        for a, b in zip(res, outs[0][1]):
            assert np.array_equal(a, b)


def test_autotuner(hg):
    cells, h, rho, u, p = cases.ensemble_1d(32, seed=9, n_lo=1000, n_hi=4000)
    cfg = _cfg(0.03)
    ens = hg.Ensemble1D(cells, h, cfg)
    ens.upload(rho, u, p)
    ens.run()
    want = ens.download()
    ens.upload(rho, u, p)
    bi, bu = ens.tune(4)
    print("autotuned block sizes (interface, update):", bi, bu, flush=True)
    assert bi in (64, 128, 256, 512) and bu in (64, 128, 256, 512)
    ens.run()
    assert np.array_equal(ens.download(), want, equal_nan=True)
    ens.close()


def test_error_paths(hg, oracle):
    m = 60
    base = cases.sod_1d(m)
    for mutate, kw in ((("p", 20, -1.0), {}), (None, {"CFL": 1e-12}), (("rho", 30, 1e-12), {})):
        init = {k: v.copy() for k, v in base.items()}
        if mutate:
            init[mutate[0]][mutate[1]] = mutate[2]
        cfg = make_config(t_all=0.1, h_x=1 / m, bc_x=-4, **kw)
        fo, npo, tpo, co, st = oracle.grp1d(m, init, cfg)
        fin, npl, tp, c, sg = hg.run_grp1d(m, init, cfg)
        assert (co[5], st) == (c[5], sg), (mutate, kw)
        assert tpo[npo - 1] == tp[npl - 1]
        ok = np.isfinite(fo[:, npo - 1])
        assert np.array_equal(ok, np.isfinite(fin))
        assert np.max(np.abs(fo[:, npo - 1][ok] - fin[ok])) <= 1e-9
    with pytest.raises(RuntimeError):
        hg.Ensemble1D([10], [0.1], make_config(t_all=0.1, bc_x=-3))


def test_pauses_at_plot_times_leave_results_unchanged(hg):
    """hgrp1d_run_until: instances stop at the start of the step where their
    time first reaches their pause time, continue from there, and end bit for
    bit where an uninterrupted run ends."""
    cells, h, rho, u, p = cases.ensemble_1d(12, seed=77, n_lo=64, n_hi=900)
    cfg = _cfg(0.15)
    ref = hg.Ensemble1D(cells, h, cfg)
    ens = hg.Ensemble1D(cells, h, cfg)
    try:
        ref.upload(rho, u, p)
        ref.run()
        want = ref.download()
        want_res = ref.results()
        ens.upload(rho, u, p)
        pauses = np.linspace(0.01, 0.14, len(cells))
        paused = ens.run_until(pauses)
        assert paused.sum() >= len(cells) - 2        # the others ended (or stopped) before their pause time
        steps, status, t = ens.results()
        assert np.all(t[paused] >= pauses[paused]) and np.all(status[paused] == 0)
        assert np.all(steps[paused] < want_res[0][paused])
        for q in range(3):                            # the unpaused ones are done: their final results
            assert np.array_equal((steps, status, t)[q][~paused], want_res[q][~paused])
        paused2 = ens.run_until(np.full(len(cells), np.inf), graph=False)
        assert not paused2.any()
        got = ens.download()
        got_res = ens.results()
    finally:
        ref.close()
        ens.close()
    assert np.array_equal(got, want)
    for a, b in zip(got_res, want_res):
        assert np.array_equal(a, b)


@pytest.mark.timeout(600)
def test_c4_full_size(hg, oracle_devpow):
    """The C4 ensemble at its full size: 512 instances, N uniform in
    [256, 16384] (4.1 M cells, the largest about 31 600 steps), each to
    t = 0.15 with its own tau.  Instances run alone (a one-instance ensemble,
    the same engine) end bit for bit where they end in the batch, the largest
    among them; the four smallest against the oracle with the device's pow
    to 1e-12, as the 48-instance test holds the engine."""
    cells, h, rho, u, p = cases.ensemble_1d(512, seed=2024)
    cfg = _cfg(0.15)
    ens = hg.Ensemble1D(cells, h, cfg)
    ens.upload(rho, u, p)
    ens.tune(32)
    ms, cu = ens.run()
    fin = ens.download()
    steps, status, t = ens.results()
    ens.close()
    off = np.concatenate([[0], np.cumsum(cells)])
    order = np.argsort(cells)
    pick = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    for i in pick:
        sl = slice(off[i], off[i + 1])
        one = hg.Ensemble1D(cells[i:i + 1], h[i:i + 1], cfg)
        one.upload(rho[sl], u[sl], p[sl])
        one.run()
        f1 = one.download()
        s1, st1, t1 = one.results()
        one.close()
        assert (s1[0], st1[0], t1[0]) == (steps[i], status[i], t[i]), i
        assert np.array_equal(f1, fin[:, sl], equal_nan=True), i
    small = [int(i) for i in order[:4]]
    sub_cells = cells[small]
    sub = [np.concatenate([a[off[i]:off[i + 1]] for i in small]) for a in (rho, u, p)]
    _, ref = _oracle_each(oracle_devpow, sub_cells, h[small], *sub, cfg)
    worst = 0.0
    for (f, k, st, te), i in zip(ref, small):
        assert (steps[i], status[i], t[i]) == (k, st, te), i
        worst = max(worst, _final_diff(fin[:, off[i]:off[i + 1]], f))
    print(f"C4: 512 instances, {int(cells.sum())} cells, {cu} cell-updates in {ms:.0f} ms; instances "
          f"{pick} (N = {list(cells[pick])}) alone: bitwise; the 4 smallest vs the device-pow oracle: "
          f"{worst:.1e}; {int(np.sum(status != 0))} ended by a check; steps up to {int(steps.max())}", flush=True)
    assert worst <= 1e-12
