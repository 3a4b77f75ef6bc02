"""The allocation-free flux batch, hgrp_grp2d_flux_batch_noalloc, against the
allocating hgrp_grp2d_flux_batch through the C ABI: the same kernel on the same
sides, so the results are equal bit for bit (check codes, fluxes, interface
values).  Two calls in a row share the device's fixed staging array, the
second with fewer interfaces than the first, so stale entries of the first
call would show; a batch larger than the array is refused."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_DP = C.POINTER(C.c_double)
MAX = 1 << 16


def _p(a):
    return a.ctypes.data_as(_DP)


def _lib():
    import hydrogrp
    L = hydrogrp.lib()
    f = L.hgrp_grp2d_flux_batch_noalloc
    f.restype = C.c_int
    f.argtypes = [_DP, _DP, C.c_int, C.c_double, C.c_double, C.c_int, _DP]
    g = L.hgrp_grp2d_flux_batch
    g.restype = C.c_int
    g.argtypes = [_DP, _DP, C.c_int, C.c_double, C.c_double, C.c_int, _DP]
    return f, g


def _random_sides(count, seed):
    """Sides of 15 doubles (gamma, lambda_u, lambda_v, rho u v p, d_rho d_u d_v
    d_p, t_rho t_u t_v t_p): strong and weak jumps, a few vacuum-like states."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(2):
        s = np.zeros((count, 15))
        s[:, 0] = 1.4
        s[:, 3] = 10 ** rng.uniform(-3, 1, count)
        s[:, 4] = rng.uniform(-3, 3, count)
        s[:, 5] = rng.uniform(-1, 1, count)
        s[:, 6] = 10 ** rng.uniform(-4, 1, count)
        s[:, 7:11] = rng.normal(0, 1, (count, 4))
        out.append(s)
    L, R = out
    R[::97] = L[::97]                       # equal states: the acoustic shortcut
    R[::89, 6] = 1e-13                      # near-vacuum sides: failing checks
    return np.ascontiguousarray(L), np.ascontiguousarray(R)


def _both(f, g, L, R, tau, eps, d):
    n = len(L)
    a = np.full((n, 9), np.nan)
    b = np.full((n, 9), -1.0)
    assert g(_p(L), _p(R), n, tau, eps, d, _p(a)) == 0
    assert f(_p(L), _p(R), n, tau, eps, d, _p(b)) == 0
    return a, b


def test_noalloc_flux_batch_equals_allocating_one():
    f, g = _lib()
    here = os.path.dirname(__file__)
    gold = np.load(os.path.join(here, "golden", "interface_vectors.npz"), allow_pickle=False)
    cases = [(_random_sides(50000, 11), 2e-3, 0), (_random_sides(50000, 11), 2e-3, 1)]
    gl, gr = gold["L"].copy(), gold["R"].copy()
    gl[:, 11:] = 0.0
    gr[:, 11:] = 0.0
    cases.append(((np.ascontiguousarray(gl[:1500]), np.ascontiguousarray(gr[:1500])), float(gold["tau"]), 0))
    for (L, R), tau, d in cases:                 # 50000, 50000 in y, then 1500 interfaces
        a, b = _both(f, g, L, R, tau, 1e-9, d)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (len(L), d)
        codes = np.unique(a[:, 0])
        print(f"{len(L)} interfaces, dir {d}: bitwise equal; check codes {codes}", flush=True)
    assert set(np.unique(a[:, 0])) <= {0.0, 1.0, 2.0, 3.0}


def test_noalloc_flux_batch_refuses_more_than_its_array():
    f, _ = _lib()
    L, R = _random_sides(MAX + 1, 3)
    out = np.zeros((MAX + 1, 9))
    assert f(_p(L), _p(R), MAX + 1, 1e-3, 1e-9, 0, _p(out)) == -1002
    assert not out.any()
    # the largest batch it takes
    a, b = _both(f, _lib()[1], L[:MAX], R[:MAX], 1e-3, 1e-9, 0)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
