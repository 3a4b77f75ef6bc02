"""The batch of 2D split problems (hgrp2db_*, hydrogrp.Batch2D) against the
one-grid solver run on each grid alone (hgrp2d_solve with plot times
{0, t_all}, itself pinned to the reference): the same final state bit for
bit, step counter, status and time, for grids of uneven sizes and time steps,
in each Newton mode, with every launch variant and after the tuner, and with
grids whose runs end on a failure (their later sweeps halted on the device,
the other grids unaffected)."""
import numpy as np
import pytest

import cases
from _libs import make_config

pytestmark = pytest.mark.gpu

T_ALL = 0.08
# (m, n, initial data): uneven sizes, each with its own h and tau
GRIDS = [(64, 48, "fq"), (100, 130, "smooth"), (33, 257, "fq"), (200, 64, "smooth"), (1, 50, "smooth"),
         (128, 128, "fq"), (70, 66, "smooth"), (9, 7, "smooth")]


def _init(m, n, kind, seed):
    return cases.four_quadrant(m, n) if kind == "fq" else cases.smooth_random(m, n, seed)


def _grids(spec):
    return [(m, n, 1.0 / m, 1.0 / n) for m, n, _ in spec]


def _single(hg, m, n, init, bx, by, newton):
    cfg = make_config(t_all=T_ALL, h_x=1.0 / m, h_y=1.0 / n, bc_x=bx, bc_y=by, split=1)
    tol, params, conv = newton
    fr, npl, tp, c, res = hg.run_split2d(m, n, init, cfg, time_plot=np.array([0.0, T_ALL]), N_T=2,
                                         newton_tol=tol, guess_params=params, converge=conv)
    return fr[:, npl - 1], int(c[5]), res.status, res.time_c


def _batch(hg, spec, inits, bx, by, newton, sweeps=None, tune=False):
    cfg = make_config(t_all=T_ALL, bc_x=bx, bc_y=by, split=1)
    b = hg.Batch2D(_grids(spec), cfg)
    try:
        tol, params, conv = newton
        if tol or params is not None or conv:
            b.set_newton(tol, params, conv)
        if sweeps is not None:
            b.set_sweeps(*sweeps)
        b.upload(inits)
        if tune:
            b.tune(3)
        ms, cu = b.run()
        out = b.download()
        k, st, t, steps = b.results()
    finally:
        b.close()
    return out, k, st, t


@pytest.fixture(scope="module")
def hg():
    import hydrogrp
    hydrogrp.lib()
    return hydrogrp


@pytest.mark.parametrize("mode", ["reference", "converge", "learned"])
def test_batch_equals_one_grid_runs(hg, mode):
    from hydrogrp import mlp
    newton = {"reference": (0.0, None, False), "converge": (1e-12, None, True),
              "learned": (1e-12, mlp.load(), True)}[mode]
    inits = [_init(m, n, kind, 10 + g) for g, (m, n, kind) in enumerate(GRIDS)]
    for bx, by in ((-4, -4), (-7, -2), (-24, -1)):
        out, k, st, t = _batch(hg, GRIDS, inits, bx, by, newton)
        for g, (m, n, _) in enumerate(GRIDS):
            ref, k1, st1, t1 = _single(hg, m, n, inits[g], bx, by, newton)
            assert (k[g], st[g], t[g]) == (k1, st1, t1), (mode, bx, by, g)
            assert np.array_equal(out[g], ref), (mode, bx, by, g)
        print(f"{mode}, BC {bx}/{by}: {len(GRIDS)} grids bitwise equal to one-grid runs; steps {list(k)}", flush=True)


def test_batch_grids_that_fail(hg):
    """Grids whose runs end on a failed check (a negative pressure, a
    vacuum-like density, a NaN: a check fails in the first sweep) among grids
    that run to the end: each ends where its one-grid run ends (status, step,
    time, state), the others are unaffected."""
    spec = [(64, 48, "fq"), (8, 6, "smooth"), (40, 50, "smooth"), (8, 6, "smooth"), (30, 20, "fq"),
            (8, 6, "smooth")]
    inits = [_init(m, n, kind, 2) for m, n, kind in spec]
    inits[1]["p"][3, 2] = -1.0
    inits[3]["rho"][5, 1] = 1e-12
    inits[5]["rho"][7, 5] = np.nan                  # a non-finite reconstruction at the last interfaces
    newton = (0.0, None, False)
    out, k, st, t = _batch(hg, spec, inits, -4, -4, newton)
    seen = set()
    for g, (m, n, _) in enumerate(spec):
        ref, k1, st1, t1 = _single(hg, m, n, inits[g], -4, -4, newton)
        assert (k[g], st[g], t[g]) == (k1, st1, t1), g
        assert np.array_equal(np.isfinite(out[g]), np.isfinite(ref)), g
        ok = np.isfinite(ref)
        assert np.array_equal(out[g][ok], ref[ok]), g
        seen.add(int(st[g]))
    print(f"statuses {list(st)}, steps {list(k)}", flush=True)
    assert 0 in seen and len(seen) > 1


def test_batch_variants_and_tuner_change_no_result(hg):
    from hydrogrp import mlp
    inits = [_init(m, n, kind, 30 + g) for g, (m, n, kind) in enumerate(GRIDS)]
    newton = (1e-12, mlp.load(), True)
    ref = None
    for sweeps in ((0, 0), (1, 1), (2, 2), (0, 3), (1, 4), (3, 0), (4, 2)):
        out, k, st, t = _batch(hg, GRIDS, inits, -4, -4, newton, sweeps=sweeps)
        if ref is None:
            ref = (out, k, st, t)
        for g in range(len(GRIDS)):
            assert np.array_equal(out[g], ref[0][g]), (sweeps, g)
        assert np.array_equal(k, ref[1]) and np.array_equal(st, ref[2]) and np.array_equal(t, ref[3])
    out, k, st, t = _batch(hg, GRIDS, inits, -4, -4, newton, tune=True)
    for g in range(len(GRIDS)):
        assert np.array_equal(out[g], ref[0][g]), ("tuned", g)
