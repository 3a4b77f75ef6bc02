"""Benchmark: 2D dimension-split Eulerian GRP, FP64, cell-updates per second,
on the learned-guess configuration: the four-quadrant Riemann problem
(data_in/two-dim/RP2D_Positive/Config3 states) at 4096 x 4096, gamma 1.4, CFL
0.45, free boundaries, Newton tolerance 1e-12 with the converge rule, started
from the learned p* guess; beside it the same with the reference's starting
guess ("plain").  One step is one full X(tau/2) Y(tau) X(tau/2) step over the
whole grid with the frozen tau = CFL * h / S_max of the reference; a
cell-update is one cell advanced by one step.  The state is resident in HBM
before the timed region; device time comes from HIP events on the solver's
stream (hgrp2d_bench).

Window: with --steps K, K steps from the state at t = 0.15 (4034 steps of the
reference's own solve, untimed, then checkpointed), each form restarting from
that checkpoint; without --steps, the whole run, 8068 steps from t = 0 to
t = 0.3.  Before timing, the sweep autotuner (hgrp2d_tune) picks each form's
launch variants on the t = 0.15 state.

Prints one JSON line.  `roofline` reports the dominant kernel (the learned
form's X sweep, two launches per step) against HBM peak with the algorithmic
bytes of a sweep (128 B per cell: read and write rho u v p and their four
slopes), its per-launch time from events around every launch of a second run
over the same window (the run that gives `value` carries no such events);
`traffic` is read from the committed rocprofv3 counter summary of that window.
`cpu_baseline` times the reference's own split solver (its sources built in
oracle/_ref, OpenMP; the C99 restatement in oracle/ where that build is absent)
on a few steps of the same 4096^2 workload from the same t = 0.15 state.
`ensemble_1d` runs 512 independent 1D GRP problems of uneven sizes through the
batched engine; `batch_2d` 128 independent 2D problems of uneven sizes through
the batched 2D engine, against the same grids solved one at a time; `two_fluid` runs the shock-bubble interaction (4096 x 1024)
over the whole run to t = 1 with and without the learned guess; `unsplit` runs the four-quadrant workload
at 1024^2 through the unsplit scheme (config[33] = 0) with its own roofline and
CPU baseline, and its learned guess against the reference's at tol 1e-12.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hydrocode-fusion_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "2D GRP FP64 cell-updates/s; learned-p* vs plain GPU at equal L1 error"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
SWEEP_BYTES_PER_CELL = 128.0   # 384 B per cell-update / 3 sweeps
M = N = 4096
T_MID = 0.15                   # start of the --steps window
WHOLE_RUN_STEPS = 8068         # the whole four-quadrant run to t = 0.3 at 4096^2 (frozen tau)
NEWTON_TOL = 1e-12
WORKLOAD = ("2D four-quadrant Riemann problem {m}x{n}, dimension-split GRP (X/2 Y X/2), FP64, learned p* guess, "
            "Newton tol 1e-12 with the converge rule")


def cpu_model():
    """The host CPU's model name and thread count from lscpu."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    info = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            info[k.strip()] = v.strip()
    return {k: info.get(k) for k in ("Model name", "CPU(s)", "Thread(s) per core", "Core(s) per socket", "Socket(s)")}


def measured_traffic(kernel, profile):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary profiles/<profile>.md (a line `traffic_bytes: {...}`),
    or None where no such profile or entry exists."""
    path = os.path.join(ROOT, "profiles", f"{profile}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes:"):
            return json.loads(line.split(":", 1)[1]).get(kernel)
    return None


def c3_leg(steps, warmup):
    """Both forms (learned, plain) on the 4096^2 workload: the window of
    `steps` steps from the t = 0.15 checkpoint, or the whole run (steps None)."""
    import numpy as np
    import hydrogrp
    from hydrogrp import mlp
    import cases
    cfg = cases.four_quadrant_config(M, N)
    init = cases.four_quadrant(M, N)
    s = hydrogrp.Solver2D(M, N, cfg)
    try:
        s.upload(init["rho"], init["u"], init["v"], init["p"])
        tau = cfg[7] * s.h_over_smax()
        k_mid = int(np.floor(T_MID / tau))
        s.bench(k_mid)                                   # the reference's solve to t = 0.15, untimed
        blob = s.checkpoint()
        mid_state = s.download()
        out, finals = {}, {}
        for name, params in (("learned", mlp.load()), ("plain", None)):
            s.set_newton(NEWTON_TOL, params, converge=True)
            s.restore(blob)
            xv, yv, _, _ = s.tune(4)
            if steps:
                s.restore(blob)
                if warmup:
                    s.bench(warmup)
                s.restore(blob)
                ms, _, _ = s.bench(steps)
                finals[name] = s.download()
                s.restore(blob)
                _, ms_x, ms_y = s.bench(steps, 1)
                n_steps = steps
            else:
                s.upload(init["rho"], init["u"], init["v"], init["p"])
                if warmup:
                    s.bench(warmup)
                    s.upload(init["rho"], init["u"], init["v"], init["p"])
                ms, ms_x, ms_y = s.bench(WHOLE_RUN_STEPS, 8)
                finals[name] = s.download()
                n_steps = WHOLE_RUN_STEPS
            out[name] = {"value": M * N * n_steps / (ms / 1000.0), "ms_per_step": ms / n_steps,
                         "sweeps": {"x": s.sweep_kernel(0), "y": s.sweep_kernel(1), "variants": [xv, yv]},
                         "ms_per_launch": {"sweep_x": ms_x, "sweep_y": ms_y}}
    finally:
        s.close()
    a, b = finals["learned"], finals["plain"]
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(b)), "non-finite state after the benchmark"
    rel_l1 = float((np.abs(a - b).sum(axis=(1, 2)) / np.abs(b).sum(axis=(1, 2))).max())
    window = (f"{steps} steps from t = {k_mid * tau:.4f} ({k_mid} steps of the reference's solve, untimed)"
              if steps else f"the whole run to t = 0.3, {WHOLE_RUN_STEPS} steps from t = 0")
    return out, rel_l1, window, (k_mid, mid_state, tau)


def newton_iterations():
    """Newton iterations of both guesses under the converge rule on held-out
    states of the training distribution, and the largest p* difference."""
    import numpy as np
    from hydrogrp import mlp
    _, _, st, ok = mlp.data(777, 1 << 16, NEWTON_TOL)
    plain = mlp.riemann_guess_batch(st[ok], tol=NEWTON_TOL, converge=True)
    learned = mlp.riemann_guess_batch(st[ok], tol=NEWTON_TOL, params=mlp.load(), converge=True)
    dp = np.abs(learned[:, 1] - plain[:, 1]) / plain[:, 1]
    return {"states": int(ok.sum()), "plain": float(plain[:, 4].mean()), "learned": float(learned[:, 4].mean()),
            "p_star_max_rel_difference": float(dp.max())}


ENSEMBLE_BYTES_PER_CELL_UPDATE = 96.0   # 1D: rho u p and their slopes, read and written


def measured_bytes_per_cell_update():
    """HBM bytes per 1D cell-update of the ensemble's two wide kernels from the
    committed rocprofv3 summary (profiles/grp1d_ensemble512_trace.md), or None."""
    path = os.path.join(ROOT, "profiles", "grp1d_ensemble512_trace.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_cell_update:"):
            return float(line.split(":", 1)[1])
    return None


def ensemble_leg(n_inst=512, seed=2024):
    """The batched 1D workload: n_inst independent Riemann problems, cell counts
    uniform in [256, 16384], states from the learned-guess training
    distribution, each to t = 0.15 with its own CFL time step.  Device time of
    the whole run after the autotuner has picked the block sizes."""
    import numpy as np
    import hydrogrp
    import cases
    from _libs import make_config
    cells, h, rho, u, p = cases.ensemble_1d(n_inst, seed=seed)
    cfg = make_config(t_all=0.15, bc_x=-4, CFL=0.45)
    ens = hydrogrp.Ensemble1D(cells, h, cfg)
    ens.upload(rho, u, p)
    bs = ens.tune(32)
    ms, cu = ens.run()
    steps, status, _ = ens.results()
    ens.close()
    achieved = ENSEMBLE_BYTES_PER_CELL_UPDATE * cu / (ms / 1000.0) / 1e9
    return {"workload": f"{n_inst} independent 1D GRP Riemann problems, {int(cells.sum())} cells "
                        f"(N uniform in [256, 16384], seed {seed}), FP64, t_end 0.15, per-instance CFL",
            "value": cu / (ms / 1000.0), "unit": "cell-updates/s", "cell_updates": int(cu), "ms": ms,
            "max_steps": int(steps.max()), "block_sizes": {"iface": bs[0], "update": bs[1]},
            "ended_by_check": int(np.sum(status != 0)),
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "bytes_per_cell_update": ENSEMBLE_BYTES_PER_CELL_UPDATE,
                         "traffic_bytes_per_cell_update": measured_bytes_per_cell_update()}}


def batch2d_leg(n_grids=128, seed=7, t_all=0.1):
    """Independent 2D problems of uneven sizes: n_grids four-quadrant grids
    with m, n uniform in [64, 512] (seeded), each to t_all with its own frozen
    time step, learned guess, Newton tolerance 1e-12 with the converge rule.
    The batch (hgrp2db_*, one launch per sweep for every grid, launch variants
    from its tuner) against the same grids solved one after another by the
    one-grid solver (hgrp2d_solve); both rates count cells x completed steps
    over the wall time of the step loops."""
    import numpy as np
    import hydrogrp
    from hydrogrp import mlp
    import cases
    from _libs import make_config
    rng = np.random.default_rng(seed)
    sizes = [(int(a), int(b)) for a, b in rng.integers(64, 513, (n_grids, 2))]
    inits = [cases.four_quadrant(m, n) for m, n in sizes]
    params = mlp.load()
    cfg = make_config(t_all=t_all, bc_x=-4, bc_y=-4, split=1)
    b = hydrogrp.Batch2D([(m, n, 1.0 / m, 1.0 / n) for m, n in sizes], cfg)
    try:
        b.set_newton(NEWTON_TOL, params, converge=True)
        b.upload(inits)
        xv, yv, _, _ = b.tune(4)
        t0 = time.perf_counter()
        ms, cu = b.run()
        wall = time.perf_counter() - t0
        k, st, _, steps = b.results()
    finally:
        b.close()
    cu_one, secs_one = 0, 0.0
    for (m, n), init in zip(sizes, inits):
        c1 = make_config(t_all=t_all, h_x=1.0 / m, h_y=1.0 / n, bc_x=-4, bc_y=-4, split=1)
        s = hydrogrp.Solver2D(m, n, c1)
        try:
            s.set_newton(NEWTON_TOL, params, converge=True)
            s.upload(init["rho"], init["u"], init["v"], init["p"])
            t0 = time.perf_counter()
            res = _quiet(s.solve, np.array([0.0, t_all]), 2)
            secs_one += time.perf_counter() - t0
            cu_one += m * n * res.k                      # k = the steps taken (config[5])
        finally:
            s.close()
    return {"workload": f"{n_grids} four-quadrant grids, m and n uniform in [64, 512] (seed {seed}), "
                        f"{sum(m * n for m, n in sizes)} cells, to t = {t_all} each with its own tau, FP64, "
                        "learned guess, Newton tol 1e-12 with the converge rule",
            "unit": "cell-updates/s", "value": cu / wall, "device_ms": ms, "cell_updates": int(cu),
            "variants": [xv, yv], "max_steps": int(steps.max()), "ended_by_check": int(np.sum(st != 0)),
            "one_grid_at_a_time": {"value": cu_one / secs_one, "seconds": secs_one},
            "speedup": (cu / wall) / (cu_one / secs_one)}


TF_SWEEP_BYTES_PER_CELL = 160.0   # rho u v p, the volume fraction and their five slopes, read and written


def two_fluid_leg(m=4096, n=1024, steps=None):
    """The two-fluid shock-bubble workload (4 x 1 domain, Mach 1.22 shock, gamma
    1.4 / 1.67 bubble, reflective top and bottom), Newton tolerance 1e-12, the
    reference's starting guess against the learned one, `steps` steps from t = 0
    each; by default the whole run to t = 1, ceil(1 / tau) steps of the frozen
    tau (about 7900).  The learned and plain final states are compared in
    relative L1, beside the plain run from densities one ulp apart (how far
    this flow carries a rounding difference over the same steps)."""
    import numpy as np
    import hydrogrp
    from hydrogrp import mlp
    import cases
    from _libs import make_config
    init, frac = cases.shock_bubble(m, n)
    cfg = make_config(t_all=1.0, h_x=4.0 / m, h_y=1.0 / n, bc_x=-4, bc_y=-2, split=1)
    init_ulp = dict(init, rho=np.nextafter(init["rho"], np.inf))
    out, finals = {}, {}
    for name, p, data in (("plain", None, init), ("learned", mlp.load(), init), ("plain_ulp", None, init_ulp)):
        s = hydrogrp.Solver2D(m, n, cfg)
        s.set_newton(1e-12, p, converge=True)
        s.upload(data["rho"], data["u"], data["v"], data["p"])
        s.set_fluids(1.67, frac)
        if name == "plain_ulp":                  # the state only, untimed
            s.bench(steps)
            finals[name] = s.download()
            s.close()
            continue
        s.bench(2)
        s.upload(data["rho"], data["u"], data["v"], data["p"])
        s.set_fluids(1.67, frac)
        if steps is None:
            steps = int(np.ceil(cfg[1] / (cfg[7] * s.h_over_smax())))
        ms, ms_x, ms_y = s.bench(steps, 8)
        kx, ky = s.sweep_kernel(0), s.sweep_kernel(1)
        h_s = s.h_over_smax()
        state = s.download()
        s.close()
        assert np.all(np.isfinite(state)), "non-finite two-fluid state"
        finals[name] = state
        out[name] = {"value": m * n * steps / (ms / 1000.0), "ms_per_step": ms / steps,
                     "ms_per_launch": {kx: ms_x, ky: ms_y}, "kernels": [kx, ky]}
    def rel_l1(a, b):
        return float((np.abs(a - b).sum(axis=(1, 2)) / np.abs(b).sum(axis=(1, 2))).max())
    d_learned, d_ulp = rel_l1(finals["learned"], finals["plain"]), rel_l1(finals["plain_ulp"], finals["plain"])
    del finals
    kx = out["learned"]["kernels"][0]
    achieved = TF_SWEEP_BYTES_PER_CELL * m * n / (out["learned"]["ms_per_launch"][kx] / 1000.0) / 1e9
    return {"workload": f"two-fluid shock-bubble {m}x{n} (4x1 domain), Newton tol 1e-12, {steps} steps from t=0 "
                        f"(to t = {steps * cfg[7] * h_s:.3f} with the frozen tau)",
            "unit": "cell-updates/s", "plain": out["plain"], "learned": out["learned"],
            "speedup": out["learned"]["value"] / out["plain"]["value"],
            "rel_l1_difference": d_learned, "rel_l1_plain_from_data_one_ulp_apart": d_ulp,
            "roofline": {"bound": "hbm", "kernel": kx, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "bytes_per_cell": TF_SWEEP_BYTES_PER_CELL,
                         "traffic": measured_traffic(kx, f"sweep_tf_kernels_shockbubble{m}x{n}_trace")}}


UNSPLIT_BYTES_PER_CELL = 208.0   # read and write rho u v p E and both slope sets (26 doubles)


def unsplit_leg(steps=2017, cpu_seconds=8.0, with_cpu=True, size=1024):
    """The unsplit scheme (GRP_solver_2D_EUL_source, config[33] = 0) on the
    four-quadrant workload at size^2: `steps` steps of one fused launch each
    with tau = CFL * h / S_max from t = 0 (2017 steps = the whole run to t = 0.3
    at 1024^2), device time from HIP events; roofline of the step kernel
    against HBM peak with 208 B per cell (rho u v p E and the x- and y-slopes,
    read and written); the learned guess against the reference's, both at Newton
    tolerance 1e-12 with the converge rule, over the same steps; the same steps
    with GRP_2D_flux's alternative solver (G2D); and the reference's unsplit
    loop on the host cores on a bounded number of steps."""
    import numpy as np
    import hydrogrp
    import cases
    from _libs import make_config
    m = n = size
    cfg = cases.four_quadrant_config(m, n, split=0)
    init = cases.four_quadrant(m, n)
    s = hydrogrp.Solver2D(m, n, cfg)
    s.upload(init["rho"], init["u"], init["v"], init["p"])
    s.bench_unsplit(5)
    s.upload(init["rho"], init["u"], init["v"], init["p"])
    ms, ms_step = s.bench_unsplit(steps, 8)
    kernel = s.sweep_kernel(2)
    state = s.download()
    # the learned guess against the reference's guess, both at tol 1e-12 with the converge rule
    from hydrogrp import mlp
    lvp = {}
    for name, params in (("learned", mlp.load()), ("plain", None)):
        s.set_newton(NEWTON_TOL, params, converge=True)
        s.upload(init["rho"], init["u"], init["v"], init["p"])
        ms_f, ms_step_f = s.bench_unsplit(steps, 8)
        lvp[name] = {"value": m * n * steps / (ms_f / 1000.0), "ms_per_step": ms_f / steps,
                     "ms_per_launch": ms_step_f, "kernel": s.sweep_kernel(2), "state": s.download()}
    # GRP_2D_flux with the reference's alternative solver (G2D), its own Newton settings
    s.set_newton(0.0, None, converge=False)
    s.set_interface_solver(True)
    s.upload(init["rho"], init["u"], init["v"], init["p"])
    ms_g, ms_step_g = s.bench_unsplit(steps, 8)
    g2d = {"solver": "linear_GRP_solver_Edir_G2D (flux_solver.c:189), the reference's Newton setting",
           "value": m * n * steps / (ms_g / 1000.0), "ms_per_step": ms_g / steps, "ms_per_launch": ms_step_g,
           "kernel": s.sweep_kernel(2)}
    s.close()
    assert np.all(np.isfinite(state)), "non-finite state after the unsplit benchmark"
    a, b = lvp["learned"].pop("state"), lvp["plain"].pop("state")
    learned_vs_plain = {"newton": "tol 1e-12, converge rule", "learned": lvp["learned"], "plain": lvp["plain"],
                        "speedup": lvp["learned"]["value"] / lvp["plain"]["value"],
                        "rel_l1_difference": float((np.abs(a - b).sum(axis=(1, 2))
                                                    / np.abs(b).sum(axis=(1, 2))).max())}
    achieved = UNSPLIT_BYTES_PER_CELL * m * n / (ms_step / 1000.0) / 1e9
    out = {"workload": f"2D four-quadrant Riemann problem {m}x{n}, unsplit GRP (one step per tau), FP64, "
                       f"{steps} steps from t=0",
           "value": m * n * steps / (ms / 1000.0), "unit": "cell-updates/s", "ms_per_step": ms / steps,
           "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": measured_traffic(kernel, f"unsplit_step_fourquadrant{m}_trace"),
                        "ms_per_launch": ms_step},
           "learned_vs_plain": learned_vs_plain, "g2d": g2d}
    if with_cpu:
        run2d, kind, label, threads = _cpu_solver(split=False)
        h = 1.0 / m
        c = np.sqrt(1.4 * init["p"] / init["rho"])
        tau = 0.45 * np.min(h / (c + np.abs(init["u"]) + np.abs(init["v"])))

        def run(k):
            cc = make_config(N=k, tau=tau, h_x=h, h_y=h, bc_x=-4, bc_y=-4, split=0)
            cc[1] = np.inf
            t0 = time.perf_counter()
            run2d(m, n, init, cc, time_plot=np.array([0.0, np.inf]))
            return time.perf_counter() - t0

        k = max(2, int(cpu_seconds / max(run(1), 1e-3)))
        t = run(k)
        out["cpu_baseline"] = {"value": m * n * k / t, "unit": "cell-updates/s", "cores": threads, "kind": kind,
                               "sample": f"{k} unsplit steps of the same {m}x{n} workload ({t:.1f} s), {label}, "
                                         f"{threads} threads"}
    return out


def _quiet(fn, *args, **kw):
    """fn(*args) with the process's C-level stdout sent to /dev/null, so that
    stdout carries the one JSON line only: the reference's solvers draw a
    progress bar on every step, and the one-grid solver prints the
    reference's end-of-run lines ("Time is up at time step k.")."""
    import ctypes
    sys.stdout.flush()
    ctypes.CDLL(None).fflush(None)
    saved, devnull = os.dup(1), os.open(os.devnull, os.O_WRONLY)
    try:
        os.dup2(devnull, 1)
        return fn(*args, **kw)
    finally:
        ctypes.CDLL(None).fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)


def _cpu_solver(split):
    """The CPU path timed as the baseline: the reference's own sources built in
    oracle/_ref where that build is present (kind "reference"), else the oracle
    (kind "port").  Returns (run2d(m, n, init, cfg, time_plot), kind, label, threads)."""
    from _libs import Oracle, RefLib
    orc = Oracle()
    threads = int(orc.lib.orc_threads())       # the same OpenMP runtime and environment
    try:
        ref = RefLib()
        fn = ref.split2d if split else ref.unsplit2d
        return (lambda *a, **kw: _quiet(fn, *a, **kw)), "reference", \
            "oracle/_ref/libhydroref.so (the reference's own sources, gcc -O2 -fopenmp)", threads
    except OSError:
        fn = orc.split2d if split else orc.unsplit2d
        return fn, "port", "oracle/liboracle.so (C99 restatement, gcc -O2 -fopenmp)", threads


def cpu_leg(mid, target_seconds=12.0):
    """The reference's CPU split solver on the host cores, fixed-tau mode
    (exactly the frozen tau), from the same t = 0.15 state of the same 4096^2
    workload, a bounded number of steps; and the same on one thread."""
    import numpy as np
    from _libs import make_config
    k_mid, state, tau = mid
    run2d, kind, label, threads = _cpu_solver(split=True)
    init = {k: state[q] for q, k in enumerate(("rho", "u", "v", "p"))}
    h = 1.0 / M

    def run(k):
        cfg = make_config(N=k, tau=tau, h_x=h, h_y=h, bc_x=-4, bc_y=-4, split=1)
        cfg[1] = np.inf
        t0 = time.perf_counter()
        run2d(M, N, init, cfg, time_plot=np.array([0.0, np.inf]))
        return time.perf_counter() - t0

    t1 = run(1)
    k = max(2, int(target_seconds / max(t1, 1e-3)))
    t = run(k)
    cpu = cpu_model()
    where = f"{cpu['Model name']}, {cpu['CPU(s)']} logical CPUs" if cpu and cpu.get("Model name") else "host CPU"
    out = {"value": M * N * k / t, "unit": "cell-updates/s", "cores": threads, "kind": kind, "cpu": cpu,
           "sample": f"{k} steps of the same {M}x{N} four-quadrant workload from t = {k_mid * tau:.4f} ({t:.1f} s), "
                     f"{label}, {threads} threads on {where}"}
    # the same on one thread (OpenMP's thread count set through its runtime, then restored)
    single = None
    try:
        import ctypes
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        try:
            t = run(1)
        finally:
            gomp.omp_set_num_threads(threads)
        single = {"value": M * N / t, "unit": "cell-updates/s", "cores": 1, "kind": kind,
                  "sample": f"1 step of the same workload ({t:.1f} s), {label}, 1 thread"}
    except OSError:
        pass
    return out, single


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="time this many steps from the t = 0.15 state (default: the whole run, 8068 steps)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched 1D ensemble")
    ap.add_argument("--no-two-fluid", action="store_true", help="skip the two-fluid shock-bubble")
    ap.add_argument("--no-unsplit", action="store_true", help="skip the unsplit scheme")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not built)", file=sys.stderr)
        sys.exit(2)
    if args.steps is not None and args.steps < 1:
        ap.error("--steps must be positive")

    import hydrogrp
    t0 = time.perf_counter()

    def note(what):
        print(f"bench.py: {what} done ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)

    forms, rel_l1, window, mid = c3_leg(args.steps, args.warmup)
    note("learned / plain 4096^2 leg")
    learned, plain = forms["learned"], forms["plain"]
    steps = args.steps or WHOLE_RUN_STEPS
    kernel = learned["sweeps"]["x"]
    ms_x = learned["ms_per_launch"]["sweep_x"]
    achieved = SWEEP_BYTES_PER_CELL * M * N / (ms_x / 1000.0) / 1e9
    nbytes = 1 << 30
    copy_gbs = 2.0 * nbytes / (hydrogrp.copy_probe(nbytes, 10) / 1000.0) / 1e9
    # HBM bytes per launch are set by the grid and the kernel, not by the time
    # window: the counter profile of the window from t = 0.15 serves both forms
    tag = f"sweep_kernels_fourquadrant{M}_t015_hbm_traffic"
    line = {
        "metric": METRIC,
        "value": learned["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": learned["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (four-quadrant initial data generated in-process)",
        "config": {"workload": WORKLOAD.format(m=M, n=N) + "; " + window, "n_x": M, "n_y": N, "gamma": 1.4,
                   "cfl": 0.45, "boundary": "free (-4, -4)", "newton": "tol 1e-12, converge rule, learned guess",
                   "device": hydrogrp.device_name()},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kernel, tag), "traffic_profile": f"profiles/{tag}.md",
                     "copy_kernel_gbs": copy_gbs,
                     "ms_per_launch": {kernel: ms_x, learned["sweeps"]["y"]: learned["ms_per_launch"]["sweep_y"]}},
        "learned_vs_plain": {"window": window, "learned": learned, "plain": plain,
                             "speedup": learned["value"] / plain["value"], "rel_l1_difference": rel_l1,
                             "newton_iterations_held_out": newton_iterations()},
    }
    note("held-out Newton iterations")
    if not args.no_ensemble:
        line["ensemble_1d"] = ensemble_leg()
        note("1D ensemble leg")
    if not args.no_ensemble:
        line["batch_2d"] = batch2d_leg()
        note("batched 2D leg")
    if not args.no_two_fluid:
        line["two_fluid"] = two_fluid_leg()
        note("two-fluid leg")
    if not args.no_unsplit:
        line["unsplit"] = unsplit_leg(with_cpu=not args.no_cpu)
        note("unsplit leg")
    if not args.no_cpu:
        line["cpu_baseline"], single = cpu_leg(mid)
        if single:
            line["cpu_baseline_1thread"] = single
        note("CPU baseline")
    line["wall_s"] = time.perf_counter() - t0
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
