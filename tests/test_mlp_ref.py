"""CPU checks of the learned-guess reference (tests/mlp_ref.py): its gradient
against finite differences, and the two-rarefaction estimate against the
oracle's exact solver where both waves are rarefactions (there it is exact)."""
import numpy as np

import mlp_ref


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    P = rng.normal(0, 0.4, mlp_ref.NP)
    X = rng.normal(0, 1, (64, 6))
    T = rng.normal(0, 1, 64)
    _, g = mlp_ref.loss_grad(P, X, T, quantize=False)
    idx = rng.choice(mlp_ref.NP, 60, replace=False)
    idx = np.concatenate([idx, [0, 191, 192, 223, 224, 1247, 1248, 1279, 1280, 1311, 1312]])
    h = 1e-6
    for k in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        fd = (mlp_ref.loss_grad(Pp, X, T, False)[0] - mlp_ref.loss_grad(Pm, X, T, False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-6 * abs(fd), (k, fd, g[k])


def test_quantized_forward_is_close_to_plain():
    rng = np.random.default_rng(6)
    P = rng.normal(0, 0.3, mlp_ref.NP)
    X = rng.normal(0, 1, (256, 6))
    yq, _ = mlp_ref.forward(P, X, True)
    yp, _ = mlp_ref.forward(P, X, False)
    assert np.max(np.abs(yq - yp)) < 0.05 * (1 + np.max(np.abs(yp)))


def test_two_rarefaction_estimate_is_exact_for_two_rarefactions(oracle):
    rng = np.random.default_rng(7)
    g = 1.4
    worst, n = 0.0, 0
    for _ in range(400):
        rho_L, p_L = 10 ** rng.uniform(-1, 1), 10 ** rng.uniform(-1, 1)
        rho_R, p_R = 10 ** rng.uniform(-1, 1), 10 ** rng.uniform(-1, 1)
        c_L, c_R = np.sqrt(g * p_L / rho_L), np.sqrt(g * p_R / rho_R)
        u_L, u_R = 0.0, rng.uniform(0.0, 4.0) * min(c_L, c_R)
        us, ps, crw0, crw1, gap, it = oracle.riemann(g, g, u_L, u_R, p_L, p_R, c_L, c_R, 1e-9, 1e-13, 500)
        if ps >= min(p_L, p_R):
            continue                        # not two rarefactions
        n += 1
        _, p_tr = mlp_ref.features(rho_L, u_L, p_L, rho_R, u_R, p_R, g)
        worst = max(worst, abs(p_tr - ps) / ps)
    assert n > 100
    assert worst < 1e-9, worst


def test_two_gamma_two_rarefaction_estimate(oracle):
    """With two gases the estimate takes two Newton steps from the mean-exponent
    closed form: close to the exact star pressure of two rarefactions."""
    rng = np.random.default_rng(8)
    worst, n = 0.0, 0
    for _ in range(400):
        gL, gR = rng.uniform(1.2, 1.8, 2)
        rho_L, p_L, rho_R, p_R = 10 ** rng.uniform(-1, 1, 4)
        c_L, c_R = np.sqrt(gL * p_L / rho_L), np.sqrt(gR * p_R / rho_R)
        u_R = rng.uniform(0.0, 3.0) * min(c_L, c_R)
        us, ps, *_ = oracle.riemann(gL, gR, 0.0, u_R, p_L, p_R, c_L, c_R, 1e-9, 1e-13, 500)
        if ps >= min(p_L, p_R):
            continue
        n += 1
        _, p_tr = mlp_ref.features(rho_L, 0.0, p_L, rho_R, u_R, p_R, gL, gR)
        worst = max(worst, abs(p_tr - ps) / ps)
    assert n > 100
    assert worst < 5e-4, worst            # two Newton steps from the mean-exponent form
