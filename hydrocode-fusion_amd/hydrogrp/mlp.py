"""The learned starting guess for p* (hgrp_mlp_* entry points of libhydrogrp.so).

A 6-32-32-1 GELU network predicts log(p*/p_TR), the correction to the
two-rarefaction estimate p_TR; Newton on the exact solver's velocity gap starts
from p_TR * exp(prediction), takes at most 3 steps from there and otherwise
restarts from the reference's own guess.  Training, inference (BF16 MFMA) and
the guessed solves all run on the GPU; this module marshals arrays and keeps
the trained parameters (pstar_mlp.npz beside this file, written by
train_pstar.py).
"""
import ctypes as C
import os

import numpy as np

from . import _check, _p, lib

NP = 1313
N_IN, N_H = 6, 32
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pstar_mlp.npz")
_FP = C.POINTER(C.c_float)


def _f(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_FP)


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (FP32 master vector)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NP, np.float32)
    lim1, lim2, lim3 = np.sqrt(6 / (N_IN + N_H)), np.sqrt(6 / (2 * N_H)), np.sqrt(6 / (N_H + 1))
    P[0:192] = rng.uniform(-lim1, lim1, 192)
    P[224:1248] = rng.uniform(-lim2, lim2, 1024)
    P[1280:1312] = rng.uniform(-lim3, lim3, 32)
    return P


def data(seed, count, tol=1e-12):
    """Training-distribution samples: (features[count,6], target[count],
    states[count,9] = rho_L u_L p_L rho_R u_R p_R gamma_L gamma_R p*, ok[count])."""
    feat = np.zeros((count, N_IN), np.float32)
    target = np.zeros(count, np.float32)
    states = np.zeros((count, 9))
    ok = np.zeros(count, np.int32)
    _check(lib().hgrp_mlp_data(seed, count, tol, _f(feat), _f(target), _p(states),
                               ok.ctypes.data_as(C.POINTER(C.c_int))), "hgrp_mlp_data")
    return feat, target, states, ok.astype(bool)


def loss_grad(params, feat, target):
    params = np.ascontiguousarray(params, np.float32)
    feat = np.ascontiguousarray(feat, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    grad = np.zeros(NP, np.float32)
    loss = C.c_double()
    _check(lib().hgrp_mlp_loss_grad(_f(params), _f(feat), _f(target), len(target), C.byref(loss), _f(grad)),
           "hgrp_mlp_loss_grad")
    return loss.value, grad


def infer(params, feat):
    params = np.ascontiguousarray(params, np.float32)
    feat = np.ascontiguousarray(feat, np.float32)
    out = np.zeros(len(feat), np.float32)
    _check(lib().hgrp_mlp_infer(_f(params), _f(feat), len(feat), _f(out)), "hgrp_mlp_infer")
    return out


def train(params, seed=1, n_samples=1 << 20, epochs=4, batch=8192, lr=3e-3, tol=1e-12):
    """Trains on the GPU; returns (params, per-epoch mean squared error)."""
    params = np.array(params, np.float32)
    losses = np.zeros(epochs, np.float32)
    _check(lib().hgrp_mlp_train(seed, n_samples, epochs, batch, lr, tol, _f(params), _f(losses)),
           "hgrp_mlp_train")
    return params, losses


def riemann_guess_batch(states, gamma=1.4, eps=1e-9, tol=1e-12, params=None, N=500, converge=False):
    """Exact solves of (rho_L u_L p_L rho_R u_R p_R [gamma_L gamma_R ...]) rows (without
    the gamma columns every row has `gamma`); params None: the reference's guess;
    converge: the converge stopping rule (else the reference's).
    Returns rows (u*, p*, CRW0, CRW1, iterations, starting guess or -1)."""
    states = np.asarray(states, np.float64)
    inp = np.zeros((len(states), 10))
    inp[:, :6] = states[:, :6]
    if states.shape[1] >= 8:
        inp[:, 6:8] = states[:, 6:8]
    else:
        inp[:, 6:8] = gamma
    inp[:, 8], inp[:, 9] = eps, tol
    out = np.zeros((len(states), 6))
    fp = None if params is None else _f(np.ascontiguousarray(params, np.float32))
    rule = 1 if converge else 0
    _check(lib().hgrp_riemann_guess_batch(_p(inp), len(inp), N, rule, fp, _p(out)), "hgrp_riemann_guess_batch")
    return out


def load(path=WEIGHTS):
    """Trained master parameters (FP32) from pstar_mlp.npz."""
    with np.load(path, allow_pickle=False) as z:
        return np.array(z["params"], np.float32)
