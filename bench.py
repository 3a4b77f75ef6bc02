"""Benchmark: 2D dimension-split Eulerian GRP, FP64, cell-updates per second.

Workload: the 2D four-quadrant Riemann problem (data_in/two-dim/RP2D_Positive/Config3
states) at 1024 x 1024, gamma 1.4, CFL 0.45, free boundaries.  One step is one
full X(tau/2) Y(tau) X(tau/2) step over the whole grid with the frozen
tau = CFL * h / S_max of the reference; a cell-update is one cell advanced by
one step.  The state is resident in HBM before the timed region; device time
comes from HIP events on the solver's stream (hgrp2d_bench).

Prints one JSON line (rank 0).  `roofline` reports the dominant kernel (the
x-sweep, two launches per step) against HBM peak with the algorithmic bytes of
a sweep (128 B per cell: read and write rho u v p and their four slopes);
`cpu_baseline` times the reference's own split solver (its sources built in
oracle/_ref, OpenMP; the C99 restatement in oracle/ where that build is absent)
on a bounded number of steps of the same workload on the host's cores.
`learned_guess` compares, at 4096^2 with Newton tolerance 1e-12, the plain
solve with the one started from the learned p* guess (rate, and the relative L1
difference of the two final states).  `ensemble_1d` runs 512 independent 1D
GRP problems of uneven sizes through the batched engine; `two_fluid` runs the
shock-bubble interaction (4096 x 1024) with and without the learned guess;
`unsplit` runs the same four-quadrant workload through the unsplit scheme
(config[33] = 0) with its own roofline and CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hydrocode-fusion_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
SWEEP_BYTES_PER_CELL = 128.0   # 384 B per cell-update / 3 sweeps
M = N = 1024
WHOLE_RUN_STEPS = 2017     # the whole four-quadrant run to t = 0.3 at 1024^2 (frozen tau)
WORKLOAD = "2D four-quadrant Riemann problem {m}x{n}, dimension-split GRP (X/2 Y X/2), FP64"


def measured_traffic(kernel, m, n):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/*_hbm_traffic.md, a line
    `traffic_bytes: {...}`), or None where no such profile exists for this
    kernel and size."""
    path = os.path.join(ROOT, "profiles", f"sweep_kernels_fourquadrant{m}_hbm_traffic.md")
    if m != n or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes:"):
            return json.loads(line.split(":", 1)[1]).get(kernel)
    return None


def gpu_leg(steps, warmup):
    import numpy as np
    import hydrogrp
    import cases
    cfg = cases.four_quadrant_config(M, N)
    s = hydrogrp.Solver2D(M, N, cfg)
    init = cases.four_quadrant(M, N)
    s.upload(init["rho"], init["u"], init["v"], init["p"])
    if warmup:
        s.bench(warmup)
        s.upload(init["rho"], init["u"], init["v"], init["p"])    # timed run starts from t = 0
    ms_total, ms_x, ms_y = s.bench(steps)
    kernels = (s.sweep_kernel(0), s.sweep_kernel(1))
    state = s.download()
    s.close()
    assert np.all(np.isfinite(state)), "non-finite state after the benchmark"
    # a plain copy of 1 GiB: the measured HBM ceiling for streaming kernels
    nbytes = 1 << 30
    copy_gbs = 2.0 * nbytes / (hydrogrp.copy_probe(nbytes, 10) / 1000.0) / 1e9
    return ms_total, ms_x, ms_y, kernels, hydrogrp.device_name(), copy_gbs


def learned_leg(size=4096, steps=200):
    """The learned-guess workload: the same four-quadrant problem at size^2 with
    Newton tolerance 1e-12, plain (the reference's starting guess) against the
    learned guess, from t = 0 for `steps` steps each; plus Newton iteration
    counts of both on held-out states of the training distribution."""
    import numpy as np
    import hydrogrp
    from hydrogrp import mlp
    import cases
    params = mlp.load()
    cfg = cases.four_quadrant_config(size, size)
    init = cases.four_quadrant(size, size)
    out, states = {}, {}
    for name, p in (("plain", None), ("learned", params)):
        s = hydrogrp.Solver2D(size, size, cfg)
        s.set_newton(1e-12, p)
        s.upload(init["rho"], init["u"], init["v"], init["p"])
        s.bench(3)
        s.upload(init["rho"], init["u"], init["v"], init["p"])
        ms, ms_x, ms_y = s.bench(steps)
        states[name] = s.download()
        s.close()
        out[name] = {"value": size * size * steps / (ms / 1000.0), "ms_per_step": ms / steps,
                     "ms_per_launch": {"sweep_x": ms_x, "sweep_y": ms_y}}
    a, b = states["learned"], states["plain"]
    rel_l1 = float((np.abs(a - b).sum(axis=(1, 2)) / np.abs(b).sum(axis=(1, 2))).max())
    # the developed flow: the whole run to t = 0.3 at dev^2 (plain, tol 1e-12), then the same
    # steps_dev steps from its final state with either guess (engine checkpoint between them)
    dev, steps_dev = size // 2, 60
    cfg_d = cases.four_quadrant_config(dev, dev)
    init_d = cases.four_quadrant(dev, dev)
    s = hydrogrp.Solver2D(dev, dev, cfg_d)
    s.set_newton(1e-12, None)
    s.upload(init_d["rho"], init_d["u"], init_d["v"], init_d["p"])
    _quiet(s.solve, np.array([0.0, cfg_d[1]]), 2)      # the solver prints its end-of-run lines
    blob = s.checkpoint()
    developed = {}
    for name, p in (("plain", None), ("learned", params)):
        s.restore(blob)
        s.set_newton(1e-12, p)
        ms, ms_x, ms_y = s.bench(steps_dev)
        developed[name] = {"value": dev * dev * steps_dev / (ms / 1000.0), "ms_per_step": ms / steps_dev,
                           "ms_per_launch": {"sweep_x": ms_x, "sweep_y": ms_y}}
    s.close()
    developed["workload"] = (f"2D four-quadrant {dev}x{dev}, Newton tol 1e-12, {steps_dev} steps from the "
                             "state at t = 0.3 (the whole run done first, plain)")
    developed["speedup"] = developed["learned"]["value"] / developed["plain"]["value"]
    # the whole run at the benchmark's 1024^2 with either guess
    whole, finals = {}, {}
    cfg_w = cases.four_quadrant_config(M, N)
    init_w = cases.four_quadrant(M, N)
    for name, p in (("plain", None), ("learned", params)):
        s = hydrogrp.Solver2D(M, N, cfg_w)
        s.set_newton(1e-12, p)
        s.upload(init_w["rho"], init_w["u"], init_w["v"], init_w["p"])
        s.bench(3)
        s.upload(init_w["rho"], init_w["u"], init_w["v"], init_w["p"])
        ms, ms_x, ms_y = s.bench(WHOLE_RUN_STEPS)
        finals[name] = s.download()
        s.close()
        whole[name] = {"value": M * N * WHOLE_RUN_STEPS / (ms / 1000.0), "ms_per_step": ms / WHOLE_RUN_STEPS,
                       "ms_per_launch": {"sweep_x": ms_x, "sweep_y": ms_y}}
    a, b = finals["learned"], finals["plain"]
    whole["workload"] = (f"2D four-quadrant {M}x{N}, Newton tol 1e-12, the whole run to t = 0.3 "
                         f"({WHOLE_RUN_STEPS} steps from t=0)")
    whole["speedup"] = whole["learned"]["value"] / whole["plain"]["value"]
    whole["rel_l1_difference"] = float((np.abs(a - b).sum(axis=(1, 2)) / np.abs(b).sum(axis=(1, 2))).max())
    _, _, st, ok = mlp.data(777, 1 << 16, 1e-12)
    it_p = mlp.riemann_guess_batch(st[ok], tol=1e-12)[:, 4]
    it_l = mlp.riemann_guess_batch(st[ok], tol=1e-12, params=params)[:, 4]
    return {"workload": f"2D four-quadrant {size}x{size}, Newton tol 1e-12, {steps} steps from t=0",
            "unit": "cell-updates/s", "plain": out["plain"], "learned": out["learned"],
            "speedup": out["learned"]["value"] / out["plain"]["value"], "rel_l1_difference": rel_l1,
            "developed_flow": developed, "whole_run": whole,
            "newton_iterations_held_out": {"states": int(ok.sum()), "plain": float(it_p.mean()),
                                           "learned": float(it_l.mean()),
                                           "fallback_share": float(np.mean(it_l > 3))}}


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
    return {"workload": f"{n_inst} independent 1D GRP Riemann problems, {int(cells.sum())} cells "
                        f"(N uniform in [256, 16384], seed {seed}), FP64, t_end 0.15, per-instance CFL",
            "value": cu / (ms / 1000.0), "unit": "cell-updates/s", "cell_updates": int(cu), "ms": ms,
            "max_steps": int(steps.max()), "block_sizes": {"iface": bs[0], "update": bs[1]},
            "ended_by_check": int(np.sum(status != 0))}


def two_fluid_leg(m=4096, n=1024, steps=300):
    """The two-fluid shock-bubble workload (4 x 1 domain, Mach 1.22 shock, gamma
    1.4 / 1.67 bubble, reflective top and bottom), Newton tolerance 1e-12, the
    reference's starting guess against the learned one, `steps` steps from t = 0
    each (the whole run to t = 1 is about 7900 steps)."""
    import numpy as np
    import hydrogrp
    from hydrogrp import mlp
    import cases
    from _libs import make_config
    init, frac = cases.shock_bubble(m, n)
    cfg = make_config(t_all=1.0, h_x=4.0 / m, h_y=1.0 / n, bc_x=-4, bc_y=-2, split=1)
    out = {}
    for name, p in (("plain", None), ("learned", mlp.load())):
        s = hydrogrp.Solver2D(m, n, cfg)
        s.set_newton(1e-12, p)
        s.upload(init["rho"], init["u"], init["v"], init["p"])
        s.set_fluids(1.67, frac)
        s.bench(2)
        s.upload(init["rho"], init["u"], init["v"], init["p"])
        s.set_fluids(1.67, frac)
        ms, ms_x, ms_y = s.bench(steps)
        state = s.download()
        s.close()
        assert np.all(np.isfinite(state)), "non-finite two-fluid state"
        out[name] = {"value": m * n * steps / (ms / 1000.0), "ms_per_step": ms / steps,
                     "ms_per_launch": {"sweep_x_tf": ms_x, "sweep_y_tf": ms_y}}
    return {"workload": f"two-fluid shock-bubble {m}x{n} (4x1 domain), Newton tol 1e-12, {steps} steps from t=0",
            "unit": "cell-updates/s", "plain": out["plain"], "learned": out["learned"],
            "speedup": out["learned"]["value"] / out["plain"]["value"]}


UNSPLIT_BYTES_PER_CELL = 208.0   # read and write rho u v p E and both slope sets (26 doubles)


def unsplit_leg(steps=2017, cpu_seconds=8.0, with_cpu=True):
    """The unsplit scheme (GRP_solver_2D_EUL_source, config[33] = 0) on the same
    four-quadrant workload: `steps` steps of one fused launch each with
    tau = CFL * h / S_max from t = 0 (2017 steps = the whole run to t = 0.3),
    device time from HIP events; roofline of the step kernel against HBM peak
    with 208 B per cell (rho u v p E and the x- and y-slopes, read and
    written); and the oracle's unsplit loop on the host cores on a bounded
    number of steps."""
    import numpy as np
    import hydrogrp
    import cases
    from _libs import make_config
    cfg = cases.four_quadrant_config(M, N, split=0)
    init = cases.four_quadrant(M, N)
    s = hydrogrp.Solver2D(M, N, cfg)
    s.upload(init["rho"], init["u"], init["v"], init["p"])
    s.bench_unsplit(5)
    s.upload(init["rho"], init["u"], init["v"], init["p"])
    ms, ms_step = s.bench_unsplit(steps)
    kernel = s.sweep_kernel(2)
    state = s.download()
    s.close()
    assert np.all(np.isfinite(state)), "non-finite state after the unsplit benchmark"
    achieved = UNSPLIT_BYTES_PER_CELL * M * N / (ms_step / 1000.0) / 1e9
    out = {"workload": f"2D four-quadrant Riemann problem {M}x{N}, unsplit GRP (one step per tau), FP64, "
                       f"{steps} steps from t=0",
           "value": M * N * steps / (ms / 1000.0), "unit": "cell-updates/s", "ms_per_step": ms / steps,
           "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kernel, M, N),
                        "ms_per_launch": ms_step}}
    if with_cpu:
        run2d, kind, label, threads = _cpu_solver(split=False)
        h = 1.0 / M
        c = np.sqrt(1.4 * init["p"] / init["rho"])
        tau = 0.45 * np.min(h / (c + np.abs(init["u"]) + np.abs(init["v"])))

        def run(k):
            cc = make_config(N=k, tau=tau, h_x=h, h_y=h, bc_x=-4, bc_y=-4, split=0)
            cc[1] = np.inf
            t0 = time.perf_counter()
            run2d(M, N, init, cc, time_plot=np.array([0.0, np.inf]))
            return time.perf_counter() - t0

        k = max(2, int(cpu_seconds / max(run(1), 1e-3)))
        t = run(k)
        out["cpu_baseline"] = {"value": M * N * k / t, "unit": "cell-updates/s", "cores": threads, "kind": kind,
                               "sample": f"{k} unsplit steps of the same {M}x{N} workload ({t:.1f} s), {label}, "
                                         f"{threads} threads"}
    return out


def _quiet(fn, *args, **kw):
    """fn(*args) with the process's C-level stdout sent to /dev/null: the
    reference's solvers draw a progress bar on every step."""
    import ctypes
    sys.stdout.flush()
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


def cpu_leg(target_seconds=12.0):
    """The reference's CPU solver on the host cores, fixed-tau mode (exactly the
    frozen tau), bounded sample."""
    import numpy as np
    import cases
    from _libs import make_config
    run2d, kind, label, threads = _cpu_solver(split=True)
    init = cases.four_quadrant(M, N)
    h = 1.0 / M
    c = np.sqrt(1.4 * init["p"] / init["rho"])
    tau = 0.45 * np.min(h / (c + np.abs(init["u"]) + np.abs(init["v"])))

    def run(k):
        cfg = make_config(N=k, tau=tau, h_x=h, h_y=h, bc_x=-4, bc_y=-4, split=1)
        cfg[1] = np.inf
        t0 = time.perf_counter()
        run2d(M, N, init, cfg, time_plot=np.array([0.0, np.inf]))
        return time.perf_counter() - t0

    t1 = run(1)
    k = max(2, int(target_seconds / max(t1, 1e-3)))
    t = run(k)
    out = {"value": M * N * k / t, "unit": "cell-updates/s", "cores": threads, "kind": kind,
           "sample": f"{k} steps of the same {M}x{N} four-quadrant workload ({t:.1f} s), {label}, {threads} threads"}
    # the same on one thread (OpenMP's thread count set through its runtime, then restored)
    single = None
    try:
        import ctypes
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        try:
            k1 = max(2, int(0.5 * target_seconds / max(run(1), 1e-3)))
            t = run(k1)
        finally:
            gomp.omp_set_num_threads(threads)
        single = {"value": M * N * k1 / t, "unit": "cell-updates/s", "cores": 1, "kind": kind,
                  "sample": f"{k1} steps of the same workload ({t:.1f} s), {label}, 1 thread"}
    except OSError:
        pass
    return out, single


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 2017 steps = the whole run to t = 0.3 at the frozen tau (tau = 0.45 h / 2.9544)
    ap.add_argument("--steps", type=int, default=2017)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--size", type=int, default=1024, help="grid cells per side (default: the 1024^2 workload)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-guess comparison at 4096^2")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched 1D ensemble")
    ap.add_argument("--no-two-fluid", action="store_true", help="skip the two-fluid shock-bubble")
    ap.add_argument("--no-unsplit", action="store_true", help="skip the unsplit scheme")
    args = ap.parse_args()
    global M, N
    M = N = args.size
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    ms_total, ms_x, ms_y, kernels, device, copy_gbs = gpu_leg(args.steps, args.warmup)
    wall = time.perf_counter() - t0
    value = M * N * args.steps / (ms_total / 1000.0)
    achieved = SWEEP_BYTES_PER_CELL * M * N / (ms_x / 1000.0) / 1e9
    line = {
        "metric": "2D GRP FP64 cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (four-quadrant initial data generated in-process)",
        "config": {"workload": WORKLOAD.format(m=M, n=N), "n_x": M, "n_y": N, "gamma": 1.4, "cfl": 0.45,
                   "boundary": "free (-4, -4)", "device": device},
        "roofline": {"bound": "hbm", "kernel": kernels[0], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kernels[0], M, N),
                     "copy_kernel_gbs": copy_gbs,
                     "ms_per_launch": {kernels[0]: ms_x, kernels[1]: ms_y}},
        "wall_s": wall,
    }
    def note(what):
        print(f"bench.py: {what} done ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)

    note("main leg")
    if not args.no_learned:
        line["learned_guess"] = learned_leg()
        note("learned-guess leg")
    if not args.no_ensemble:
        line["ensemble_1d"] = ensemble_leg()
        note("1D ensemble leg")
    if not args.no_two_fluid:
        line["two_fluid"] = two_fluid_leg()
        note("two-fluid leg")
    if not args.no_unsplit:
        line["unsplit"] = unsplit_leg(with_cpu=not args.no_cpu)
        note("unsplit leg")
    if not args.no_cpu:
        line["cpu_baseline"], single = cpu_leg()
        if single:
            line["cpu_baseline_1thread"] = single
        note("CPU baseline")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
