"""GPU tests of the two-fluid split path (hgrp2d_set_fluids): it has no
structured counterpart in the reference, so it is pinned by its single-fluid
limit (bit for bit against the reference-parity path), Abgrall's uniform
pressure/velocity invariant, conservation, and the exact two-gamma Riemann
solution (tests/exact_riemann.py)."""
import numpy as np
import pytest

import cases
import exact_riemann
from _libs import make_config
from conftest import rel_linf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hg():
    import hydrogrp
    hydrogrp.lib()
    return hydrogrp


def _run(hg, m, n, init, frac, cfg, gamma_b=1.67, time_plot=None):
    s = hg.Solver2D(m, n, cfg)
    try:
        s.upload(init["rho"], init["u"], init["v"], init["p"])
        if frac is not None:
            s.set_fluids(gamma_b, frac)
        tp = np.array([0.0, cfg[1]]) if time_plot is None else time_plot
        res = s.solve(tp, 2)
        out = s.download()
        fa = s.download_fraction() if frac is not None else None
    finally:
        s.close()
    return out, fa, res


def test_single_fluid_limit_is_bitwise(hg):
    m, n = 96, 80
    init = cases.four_quadrant(m, n)
    cfg = cases.four_quadrant_config(m, n, t_all=0.2)
    one, _, r1 = _run(hg, m, n, init, None, cfg)
    two, fa, r2 = _run(hg, m, n, init, np.ones((m, n)), cfg)
    assert (r1.k, r1.status, r1.time_c) == (r2.k, r2.status, r2.time_c)
    assert np.array_equal(one.view(np.int64), two.view(np.int64))
    assert np.all(fa == 1.0)


# the X sweep's blocks update 11 rows: heights 1, 11, 12, 22, 23 and 34 end
# in a full, a one-row and a partial block
@pytest.mark.parametrize("m,n,bx", [(1, 40, -4), (11, 20, -7), (12, 33, -2), (22, 65, -24), (23, 17, -2),
                                    (34, 64, -4)])
def test_single_fluid_limit_at_block_edges(hg, m, n, bx):
    """a = 1 everywhere: the two-fluid sweeps equal the single-fluid ones bit for
    bit on heights around the X sweep's block edges, for each x boundary the
    two-fluid sweeps take (-2, -4, -7, -24)."""
    init = cases.smooth_random(m, n, 100 * m + n)
    cfg = make_config(t_all=0.03, h_x=1 / max(m, 2), h_y=1 / n, bc_x=bx, bc_y=-4, split=1)
    one, _, r1 = _run(hg, m, n, init, None, cfg)
    two, fa, r2 = _run(hg, m, n, init, np.ones((m, n)), cfg)
    assert r1.status == 0 and (r1.k, r1.status, r1.time_c) == (r2.k, r2.status, r2.time_c)
    assert np.array_equal(one.view(np.int64), two.view(np.int64))


def test_abgrall_pressure_and_velocity_stay_uniform(hg):
    m = n = 64
    init, frac = cases.material_interface(m, n)
    cfg = make_config(t_all=0.2, h_x=1 / m, h_y=1 / n, bc_x=-7, bc_y=-7, split=1)
    out, fa, res = _run(hg, m, n, init, frac, cfg)
    assert res.status == 0
    dp = np.max(np.abs(out[3] - 1.0))
    du = max(np.max(np.abs(out[1] - 0.5)), np.max(np.abs(out[2] - 0.25)))
    moved = np.abs(fa - frac).max()
    print(f"material interface after {res.k} steps: max |p-1| {dp:.2e}, max |u-u0| {du:.2e}, "
          f"volume fraction moved by up to {moved:.2f}", flush=True)
    assert moved > 0.5                       # the blob has travelled
    assert dp <= 1e-11 and du <= 1e-11


def test_conservation_with_periodic_boundaries(hg):
    m, n = 80, 64
    init, frac = cases.material_interface(m, n, u0=0.3, v0=-0.2)
    rng = np.random.default_rng(3)
    init["p"] = init["p"] + 0.3 * np.sin(2 * np.pi * (np.arange(m)[:, None] + 0.5) / m) * np.ones((1, n))
    init["u"] = init["u"] + 0.1 * rng.uniform(-1, 1, (m, n))
    cfg = make_config(t_all=0.1, h_x=1 / m, h_y=1 / n, bc_x=-7, bc_y=-7, split=1)
    s = hg.Solver2D(m, n, cfg)
    s.upload(init["rho"], init["u"], init["v"], init["p"])
    s.set_fluids(1.67, frac)
    before = s.download()
    res = s.solve(np.array([0.0, 0.1]), 2)
    after = s.download()
    s.close()
    assert res.status == 0

    def totals(st):
        rho, u, v, p, e = st
        return np.array([rho.sum(), (rho * u).sum(), (rho * v).sum(), (rho * e).sum()])
    t0, t1 = totals(before), totals(after)
    rel = np.abs(t1 - t0) / np.maximum(np.abs(t0), 1.0)
    print("mass, momentum x/y, energy drift:", rel, flush=True)
    assert rel.max() <= 1e-12


def test_two_gamma_shock_tube_converges_to_exact(hg):
    errs = []
    for m in (200, 400, 800):
        n = 4
        x = (np.arange(m) + 0.5) / m
        left = np.repeat((x < 0.5)[:, None], n, axis=1)
        init = {"rho": np.where(left, 1.0, 0.125), "u": np.zeros((m, n)), "v": np.zeros((m, n)),
                "p": np.where(left, 1.0, 0.1)}
        frac = np.where(left, 1.0, 0.0)
        cfg = make_config(t_all=0.2, h_x=1 / m, h_y=1 / m, bc_x=-4, bc_y=-4, split=1)
        out, fa, res = _run(hg, m, n, init, frac, cfg, gamma_b=1.67)
        assert res.status == 0
        er, eu, ep = exact_riemann.sample(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, (x - 0.5) / res.time_c, 1.4, 1.67)
        errs.append(np.mean(np.abs(out[0][:, 0] - er)))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print("two-gamma shock tube: L1(rho)", errs, "orders", rates, flush=True)
    assert errs[-1] < 5e-3 and np.all(rates > 0.6)


def test_shock_bubble_runs_and_stays_bounded(hg):
    m, n = 256, 64
    init, frac = cases.shock_bubble(m, n)
    cfg = make_config(t_all=0.4, h_x=4.0 / m, h_y=1.0 / n, bc_x=-4, bc_y=-2, split=1)
    out, fa, res = _run(hg, m, n, init, frac, cfg)
    print(f"shock-bubble {m}x{n}: {res.k} steps, status {res.status}, t {res.time_c:.3f}", flush=True)
    assert res.status == 0 and np.all(np.isfinite(out))
    assert out[3].min() > 0 and out[0].min() > 0
    assert fa.min() >= 0.0 and fa.max() <= 1.0
    x = (np.arange(m) + 0.5) * 4.0 / m
    w0, w1 = 1.0 - frac, 1.0 - fa
    c0 = (w0 * x[:, None]).sum() / w0.sum()
    c1 = (w1 * x[:, None]).sum() / w1.sum()
    assert c1 > c0 + 0.01                     # the shocked bubble is pushed downstream


def test_two_gamma_riemann_vectors(hg, oracle):
    rng = np.random.default_rng(17)
    rows, ok = [], 0
    for _ in range(300):
        gL, gR = rng.choice([1.4, 1.67, 1.2]), rng.choice([1.4, 1.67, 3.0])
        rL, rR = 10 ** rng.uniform(-1, 1), 10 ** rng.uniform(-1, 1)
        pL, pR = 10 ** rng.uniform(-1, 1), 10 ** rng.uniform(-1, 1)
        cL, cR = np.sqrt(gL * pL / rL), np.sqrt(gR * pR / rR)
        uL, uR = 0.0, rng.uniform(-2, 1) * min(cL, cR)
        rows.append([gL, gR, uL, uR, pL, pR, cL, cR, 1e-9, 1e-12])
    inp = np.array(rows)
    out = hg.riemann_batch(inp, 500)
    for k, r in enumerate(rows):
        us, ps, c0, c1, gap, it = oracle.riemann(*r[:10], 500)
        assert (out[k, 2], out[k, 3]) == (c0, c1)
        assert abs(out[k, 1] - ps) <= 1e-12 * ps
        p_ex, u_ex = exact_riemann.star_state(r[4] / r[6] ** 2 * r[0], 0.0, r[4], r[5] / r[7] ** 2 * r[1], r[3], r[5],
                                              r[0], r[1])
        assert abs(out[k, 1] - p_ex) <= 1e-9 * p_ex


def test_learned_guess_in_two_fluid_sweeps(hg):
    """Shock-bubble with Newton tolerance 1e-12 and the converge rule, the
    learned guess against the reference's guess: the same solution."""
    from hydrogrp import mlp
    params = mlp.load()
    m, n = 256, 64
    init, frac = cases.shock_bubble(m, n)
    cfg = make_config(t_all=0.3, h_x=4.0 / m, h_y=1.0 / n, bc_x=-4, bc_y=-2, split=1)
    outs = []
    for p in (None, params):
        s = hg.Solver2D(m, n, cfg)
        s.upload(init["rho"], init["u"], init["v"], init["p"])
        s.set_fluids(1.67, frac)
        s.set_newton(1e-12, p, converge=True)
        res = s.solve(np.array([0.0, 0.3]), 2)
        assert res.status == 0
        outs.append((s.download(), s.download_fraction(), res.k))
        s.close()
    (a, fa, ka), (b, fb, kb) = outs
    assert ka == kb
    d = rel_linf(b, a)
    l1 = np.abs(b - a).sum(axis=(1, 2)) / np.abs(a).sum(axis=(1, 2))
    print("two-fluid learned vs plain (converge rule): rel Linf", d, "rel L1", l1, flush=True)
    assert l1.max() <= 1e-11 and d.max() <= 1e-10
    assert np.abs(fb - fa).max() <= 1e-10


def test_c5_full_size_bounded_steps(hg):
    """The shock-bubble configuration at its full size, 4096 x 1024 (4 x 1
    domain, free ends in x, reflective top and bottom, learned guess, Newton
    tolerance 1e-12 with the converge rule), for 300 steps: every value finite
    and positive, the volume fraction within [0, 1], and, around the bubble that
    the shock has not reached yet, pressure and velocity uniform across the
    material interface (Abgrall's condition) as far as the single-fluid scheme
    keeps them: the reference's GRP scheme itself (here the device's
    single-fluid path, which equals it to rounding, on the same densities)
    grows rounding noise at the smeared rim to ~2e-9 in these 300 steps."""
    from hydrogrp import mlp
    m, n = 4096, 1024
    init, frac = cases.shock_bubble(m, n)
    cfg = make_config(t_all=1.0, N=300, h_x=4.0 / m, h_y=1.0 / n, bc_x=-4, bc_y=-2, split=1)
    x = (np.arange(m) + 0.5) * 4.0 / m
    ahead = (x > 0.7) & (x < 1.3)                               # the shock is near x = 0.56
    dev = {}
    for name, two in (("two-fluid", True), ("single-fluid", False)):
        s = hg.Solver2D(m, n, cfg)
        try:
            s.set_newton(1e-12, mlp.load(), converge=True)
            s.upload(init["rho"], init["u"], init["v"], init["p"])
            if two:
                s.set_fluids(1.67, frac)
            res = s.solve(np.array([0.0, cfg[1]]), 2)
            out = s.download()
            fa = s.download_fraction() if two else None
        finally:
            s.close()
        assert res.status == 0 and res.k == 301
        assert np.all(np.isfinite(out)) and out[0].min() > 0 and out[3].min() > 0
        dev[name] = (np.abs(out[3][ahead] - 1.0).max(), max(np.abs(out[1][ahead]).max(), np.abs(out[2][ahead]).max()))
        if two:
            assert fa.min() >= 0.0 and fa.max() <= 1.0
            assert ((fa > 1e-6) & (fa < 1 - 1e-6)).sum() > 1000    # the bubble's rim
    print(f"C5 4096x1024, 300 steps, around the bubble: |p - 1|, max |u|, |v|: {dev}", flush=True)
    (dp2, du2), (dp1, du1) = dev["two-fluid"], dev["single-fluid"]
    assert dp2 <= max(10.0 * dp1, 1e-12) and du2 <= max(10.0 * du1, 1e-12) and dp2 <= 1e-7


def test_c5_full_size_conservation(hg):
    """The same grid with periodic ends in x (reflective top and bottom): mass,
    x-momentum and energy sums are conserved to rounding over 100 steps, with
    the shock passing through the bubble's two fluids."""
    m, n = 4096, 1024
    init, frac = cases.shock_bubble(m, n)
    cfg = make_config(t_all=np.inf, N=100, h_x=4.0 / m, h_y=1.0 / n, bc_x=-7, bc_y=-2, split=1)
    s = hg.Solver2D(m, n, cfg)
    try:
        s.upload(init["rho"], init["u"], init["v"], init["p"])
        s.set_fluids(1.67, frac)
        r0, u0, v0, p0, e0 = s.download()
        res = s.solve(np.array([0.0, np.inf]), 2)
        r1, u1, v1, p1, e1 = s.download()
    finally:
        s.close()
    assert res.status == 0 and res.k == 101
    for name, a, b in (("mass", r0, r1), ("x-momentum", r0 * u0, r1 * u1), ("energy", r0 * e0, r1 * e1)):
        s0, s1 = a.sum(), b.sum()
        scale = np.abs(a).sum()
        print(f"C5 grid, periodic in x: {name} relative drift {(s1 - s0) / scale:.2e}", flush=True)
        assert abs(s1 - s0) <= 1e-11 * scale, name


def test_c5_whole_run_learned_within_the_rounding_spread(hg):
    """C5 to t = 1 (4096 x 1024, about 7900 steps), tol 1e-12 with the converge
    rule: the learned and the plain run take the same steps and end no farther
    apart in relative L1 than the plain run ends from the plain run on
    densities one ulp apart (the shock-bubble's mixing layer carries any
    rounding difference that far over the run)."""
    from hydrogrp import mlp
    m, n = 4096, 1024
    init, frac = cases.shock_bubble(m, n)
    cfg = make_config(t_all=1.0, h_x=4.0 / m, h_y=1.0 / n, bc_x=-4, bc_y=-2, split=1)
    init_ulp = dict(init, rho=np.nextafter(init["rho"], np.inf))
    out = {}
    for name, data, params in (("plain", init, None), ("learned", init, mlp.load()), ("ulp", init_ulp, None)):
        s = hg.Solver2D(m, n, cfg)
        try:
            s.set_newton(1e-12, params, converge=True)
            s.upload(data["rho"], data["u"], data["v"], data["p"])
            s.set_fluids(1.67, frac)
            res = s.solve(np.array([0.0, 1.0]), 2)
            out[name] = (s.download(), res)
        finally:
            s.close()
        assert out[name][1].status == 0 and np.all(np.isfinite(out[name][0])), name

    def rel_l1(a, b):
        return float((np.abs(a - b).sum(axis=(1, 2)) / np.abs(b).sum(axis=(1, 2))).max())
    assert out["learned"][1].k == out["plain"][1].k
    d, d_ulp = rel_l1(out["learned"][0], out["plain"][0]), rel_l1(out["ulp"][0], out["plain"][0])
    print(f"C5 to t = 1 ({out['plain'][1].k} steps): learned vs plain rel L1 {d:.2e}; plain from data "
          f"one ulp apart vs plain {d_ulp:.2e}", flush=True)
    assert d <= 2.0 * d_ulp and d <= 1e-5


def test_two_fluid_flux_against_multifluid_reference(hg, golden_twofluid):
    """On the GPU: the two-fluid sweeps' interface solve (hgrp_grp2d_flux_tf_batch)
    against the reference's GRP_2D_flux built with -DMULTIFLUID_BASICS
    (tests/golden/twofluid_vectors.npz), both normals."""
    g = golden_twofluid
    L, R, tau = g["L"], g["R"], float(g["tau"])
    for direction in (0, 1):
        Ld, Rd = L.copy(), R.copy()
        Ld[:, 11:15] = Rd[:, 11:15] = 0.0 if direction == 0 else -0.0
        out = hg.flux_tf_batch(Ld, Rd, tau, 1e-9, direction, float(g["gamma_a"]), float(g["gamma_b"]))
        ref = g[f"flux_{direction}"]
        assert np.array_equal(out[:, 0], ref[:, 0])
        want, got = ref[:, 1:10], out[:, 1:10]
        d = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
        print(f"two-fluid flux dir {direction} (GPU): {np.mean(np.all(got == want, axis=1)) * 100:.1f}% rows "
              f"bit-identical, max rel {d.max():.2e}", flush=True)
        assert d.max() <= 1e-10
