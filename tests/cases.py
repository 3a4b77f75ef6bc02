"""Synthetic inputs shared by the tests, bench.py and the fixture generator.

Every generator is deterministic for a given seed.  Numbers follow the
reference's case definitions:
* four_quadrant: the 2D Riemann problem of data_in/two-dim/RP2D_Positive/Config3
  (same states and split at (0.5, 0.5); value_start.m:1-4,27-48), any n_x x n_y.
* sod_1d: Toro's Sod tube, (1, 0, 1) | (0.125, 0, 0.1), split at x = 0.5.
"""
import math

import numpy as np

# (rho, u, v, p) of the four quadrants
FQ_NE = (1.5, 0.0, 0.0, 1.5)
FQ_NW = (0.5323, 1.206, 0.0, 0.3)
FQ_SW = (0.138, 1.206, 1.206, 0.029)
FQ_SE = (0.5323, 0.0, 1.206, 0.3)


def four_quadrant(m, n):
    """Cell (j, i) = (x index, y index); x < 0.5 <=> j < m/2, y < 0.5 <=> i < n/2."""
    out = {k: np.empty((m, n)) for k in ("rho", "u", "v", "p")}
    west = (np.arange(m) + 0.5) / m < 0.5
    south = (np.arange(n) + 0.5) / n < 0.5
    for wj in (True, False):
        for si in (True, False):
            st = {(True, True): FQ_SW, (True, False): FQ_NW,
                  (False, True): FQ_SE, (False, False): FQ_NE}[(wj, si)]
            mask = np.outer(west == wj, south == si)
            for k, val in zip(("rho", "u", "v", "p"), st):
                out[k][mask] = val
    return out


def four_quadrant_config(m, n, t_all=0.3, split=1, **kw):
    from _libs import make_config
    return make_config(t_all=t_all, h_x=1.0 / m, h_y=1.0 / n, bc_x=-4, bc_y=-4, split=split, **kw)


def blast_box(m, n):
    """High-pressure disc (p = 10 inside r^2 < 0.02 around (0.4, 0.55), 0.1 outside)
    in a unit box of gas at rest with rho = 1."""
    x = (np.arange(m)[:, None] + 0.5) / m
    y = (np.arange(n)[None, :] + 0.5) / n
    r2 = (x - 0.4) ** 2 + (y - 0.55) ** 2
    return {"rho": np.ones((m, n)), "u": np.zeros((m, n)), "v": np.zeros((m, n)),
            "p": np.where(r2 < 0.02, 10.0, 0.1)}


def smooth_random(m, n, seed, amp=0.2):
    """Smooth positive random data (sums of a few sine modes) for BC tests."""
    rng = np.random.default_rng(seed)
    x = (np.arange(m) + 0.5) / m
    y = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, y, indexing="ij")
    out = {}
    for k, base in (("rho", 1.0), ("u", 0.0), ("v", 0.0), ("p", 1.0)):
        f = np.full((m, n), base)
        for _ in range(3):
            a, b = rng.uniform(-amp, amp, 2)
            kx, ky = rng.integers(1, 4, 2)
            ph = rng.uniform(0, 2 * np.pi)
            f = f + a * np.sin(2 * np.pi * kx * X + ph) * np.cos(2 * np.pi * ky * Y + b)
        out[k] = f
    return out


def sod_1d(m):
    x = (np.arange(m) + 0.5) / m
    left = x < 0.5
    return {"rho": np.where(left, 1.0, 0.125), "u": np.zeros(m), "p": np.where(left, 1.0, 0.1)}


def interface_pairs(count, seed, gamma=1.4):
    """Random left/right interface states covering every branch of the GRP solvers.

    Returns (L, R) arrays of shape (count, 15) in orc_ifv order: gamma, lambda_u,
    lambda_v, rho, u, v, p, d_rho, d_u, d_v, d_p, t_rho, t_u, t_v, t_p.
    Mixes: strong jumps (density ratio 1e-3..1e3, pressure ratio 1e-4..1e4,
    velocity jump +-5 c_L), supersonic drifts (trivial / sonic-fan branches),
    near-identical states (acoustic branch), identical states, and weak jumps.
    """
    rng = np.random.default_rng(seed)
    L = np.zeros((count, 15))
    R = np.zeros((count, 15))
    L[:, 0] = R[:, 0] = gamma
    kind = rng.integers(0, 6, count)
    rhoL = 10 ** rng.uniform(-1, 1, count)
    pL = 10 ** rng.uniform(-1, 1, count)
    cL = np.sqrt(gamma * pL / rhoL)
    rratio = 10 ** rng.uniform(-3, 3, count)
    pratio = 10 ** rng.uniform(-4, 4, count)
    du = rng.uniform(-5, 5, count) * cL
    drift = rng.uniform(-3, 3, count) * cL
    weak = 10 ** rng.uniform(-8, -2, count)
    tiny = 10 ** rng.uniform(-13, -9.5, count)
    rhoR = np.where(kind <= 1, rhoL / rratio, rhoL * (1 + weak * rng.uniform(-1, 1, count)))
    pR = np.where(kind <= 1, pL / pratio, pL * (1 + weak * rng.uniform(-1, 1, count)))
    uL = np.where(kind == 1, drift, rng.uniform(-0.5, 0.5, count) * cL)
    uR = np.where(kind <= 1, uL + du, uL + weak * cL * rng.uniform(-1, 1, count))
    # kind 3: acoustic (dist < 1e-9); kind 4: identical; kind 5: weak + drift
    near = kind == 3
    rhoR = np.where(near, rhoL + tiny, rhoR)
    pR = np.where(near, pL - tiny, pR)
    uR = np.where(near, uL + tiny, uR)
    same = kind == 4
    rhoR = np.where(same, rhoL, rhoR)
    pR = np.where(same, pL, pR)
    uR = np.where(same, uL, uR)
    uL = np.where(kind == 5, uL + drift, uL)
    uR = np.where(kind == 5, uR + drift, uR)
    L[:, 3], L[:, 4], L[:, 6] = rhoL, uL, pL
    R[:, 3], R[:, 4], R[:, 6] = rhoR, uR, pR
    L[:, 5] = rng.uniform(-1, 1, count)
    R[:, 5] = rng.uniform(-1, 1, count)
    # slopes: normal derivatives always, tangential ones on a third of the cases
    for A in (L, R):
        A[:, 7:11] = rng.normal(0, 1, (count, 4)) * np.where(rng.random((count, 1)) < 0.8, 1.0, 0.0)
        tmask = rng.random((count, 1)) < 0.33
        A[:, 11:15] = rng.normal(0, 1, (count, 4)) * tmask
    return L, R


def ensemble_1d(n_inst, seed, n_lo=256, n_hi=16384, gamma=1.4):
    """Independent 1D Riemann problems on [0, 1] (jump at x = 0.5) with cell
    counts uniform in [n_lo, n_hi] and states from the learned-guess training
    distribution: rho_L/rho_R log-uniform 1e-3..1e3, p_L/p_R log-uniform
    1e-4..1e4, (u_R - u_L)/c_L uniform -5..5; each instance is scaled so that
    its larger sound speed is 1 (rho_L = 1, u_L = 0).
    Returns (cells, h, rho, u, p) with the instances concatenated."""
    rng = np.random.default_rng(seed)
    cells = rng.integers(n_lo, n_hi + 1, n_inst)
    rho, u, p = [], [], []
    for m in cells:
        r_ratio = 10 ** rng.uniform(-3, 3)
        p_ratio = 10 ** rng.uniform(-4, 4)
        du = rng.uniform(-5, 5)
        rho_L, p_L = 1.0, 1.0
        rho_R, p_R = rho_L / r_ratio, p_L / p_ratio
        c_max = max(np.sqrt(gamma * p_L / rho_L), np.sqrt(gamma * p_R / rho_R))
        p_L, p_R = p_L / c_max ** 2, p_R / c_max ** 2
        c_L = np.sqrt(gamma * p_L / rho_L)
        x = (np.arange(m) + 0.5) / m
        left = x < 0.5
        rho.append(np.where(left, rho_L, rho_R))
        u.append(np.where(left, 0.0, du * c_L))
        p.append(np.where(left, p_L, p_R))
    return cells, 1.0 / cells, np.concatenate(rho), np.concatenate(u), np.concatenate(p)


def _disc_fraction(m, n, lx, ly, cx, cy, r, sub=8):
    """Fraction of each cell (domain [0,lx]x[0,ly], m x n cells) inside the
    circle of radius r at (cx, cy), by sub x sub point sampling."""
    hx, hy = lx / m, ly / n
    off = (np.arange(sub) + 0.5) / sub
    xs = (np.arange(m)[:, None] + off[None, :]) * hx          # m x sub
    ys = (np.arange(n)[:, None] + off[None, :]) * hy          # n x sub
    inside = ((xs[:, None, :, None] - cx) ** 2 + (ys[None, :, None, :] - cy) ** 2) <= r * r
    return inside.mean(axis=(2, 3))


def shock_bubble(m, n, mach=1.22, gamma_a=1.4, gamma_b=1.67, lx=4.0, ly=1.0):
    """Two-fluid shock-bubble interaction: ambient fluid a (rho 1, p 1) on
    [0,lx]x[0,ly], a bubble of fluid b (rho 0.138, p 1) of radius 0.25 at
    (1, 0.5), a planar shock of Mach `mach` at x = 0.5 moving into the ambient
    gas (post-shock state from the Rankine-Hugoniot relations).  Returns
    (init dict rho u v p, volume fraction of fluid a)."""
    g = gamma_a
    M2 = mach * mach
    p2 = 1.0 + 2.0 * g / (g + 1.0) * (M2 - 1.0)
    r2 = (g + 1.0) * M2 / ((g - 1.0) * M2 + 2.0)
    u2 = math.sqrt(g) * 2.0 / (g + 1.0) * (mach - 1.0 / mach)
    x = (np.arange(m) + 0.5) * lx / m
    post = np.repeat((x < 0.5)[:, None], n, axis=1)
    inside = _disc_fraction(m, n, lx, ly, 1.0, 0.5, 0.25)
    frac_a = 1.0 - inside
    rho = np.where(post, r2, frac_a * 1.0 + inside * 0.138)
    u = np.where(post, u2, 0.0)
    p = np.where(post, p2, 1.0)
    return {"rho": rho, "u": u, "v": np.zeros((m, n)), "p": p}, frac_a


def material_interface(m, n, u0=0.5, v0=0.25):
    """Uniform pressure and velocity with a circular blob of a second fluid
    (Abgrall's test): only the contact moves, p and u must stay uniform."""
    inside = _disc_fraction(m, n, 1.0, 1.0, 0.5, 0.5, 0.25)
    frac_a = 1.0 - inside
    rho = frac_a * 1.0 + inside * 0.138
    return ({"rho": rho, "u": np.full((m, n), u0), "v": np.full((m, n), v0), "p": np.ones((m, n))}, frac_a)


def c3_states(count, seed, near_vacuum=0):
    """Interface states of the learned-guess distribution at gamma 1.4: left
    (1, 0, 1); rho_L/rho_R log-uniform in [1e-3, 1e3], p_L/p_R in [1e-4, 1e4],
    (u_R - u_L)/c_L uniform in [-5, 5]; states that would open a vacuum are
    dropped.  near_vacuum: that many more with (u_R - u_L)/c_L in [4.5, 4.98]
    and p_L/p_R in [1e2, 1e4] (double rarefactions with p* down to ~1e-16).
    Rows: rho_L u_L p_L rho_R u_R p_R."""
    g = 1.4
    rng = np.random.default_rng(seed)
    rows = []
    for k in range(count + near_vacuum):
        rr = 1.0 / 10 ** rng.uniform(-3, 3)
        if k < count:
            pr = 1.0 / 10 ** rng.uniform(-4, 4)
            du = rng.uniform(-5, 5)
        else:
            pr = 1.0 / 10 ** rng.uniform(2, 4)
            du = rng.uniform(4.5, 4.98)
        cl, cr = np.sqrt(g), np.sqrt(g * pr / rr)
        ur = du * cl
        if 2 * (cl + cr) / (g - 1) <= ur:
            continue
        rows.append((1.0, 0.0, 1.0, rr, ur, pr))
    return np.array(rows)


def write_ensemble_bin(path, n_inst=512, seed=2024):
    """ensemble_1d(n_inst, seed) in the file format of grp1d_bench.out:
    int32 n_inst, int32 cells[], then float64 rho[], u[], p[] (little-endian)."""
    cells, _, rho, u, p = ensemble_1d(n_inst, seed)
    with open(path, "wb") as f:
        np.array([n_inst], "<i4").tofile(f)
        cells.astype("<i4").tofile(f)
        for a in (rho, u, p):
            np.asarray(a, "<f8").tofile(f)


def write_grid_bin(path, config, init, frac=None, gamma_b=None):
    """A 2D workload in the file format of grp2d_file_bench.out: int32 m, n;
    float64 config[400]; float64 rho, u, v, p (m x n, row j at j*n); int32
    two_fluid, then (if 1) float64 gamma_b and the volume fraction (m x n)."""
    m, n = np.asarray(init["rho"]).shape
    with open(path, "wb") as f:
        np.array([m, n], "<i4").tofile(f)
        np.asarray(config, "<f8").tofile(f)
        for k in ("rho", "u", "v", "p"):
            np.ascontiguousarray(init[k], "<f8").tofile(f)
        np.array([0 if frac is None else 1], "<i4").tofile(f)
        if frac is not None:
            np.array([gamma_b], "<f8").tofile(f)
            np.ascontiguousarray(frac, "<f8").tofile(f)
