"""GPU parity of the 2D step's stages as functions of their own:
bound_cond_slope_limiter_x / _y and flux_generator_x / _y from libhydrocode2d
(CV frames, CV[0]'s slope / flux / interface arrays and the b_f_var ghost
arrays through ctypes) against the reference's own outputs in
tests/golden/stage_cases.npz (9 x 7 grid, nt = 1, every boundary condition
pair, both limiter forms, with and without transversal derivatives).

Bar: slopes and ghost arrays bit for bit (copies, negations, one division and
the minmod choice); fluxes and interface values <= 1e-10 relative (scale
max(|reference|, 1)); return values equal.
"""
import ctypes as C
import os

import numpy as np
import pytest

from _libs import DP, DPP, ROOT, CellVarStru, make_config

pytestmark = pytest.mark.gpu
HOST_SO = os.path.join(ROOT, "hydrocode-fusion_amd", "lib", "libhydrocode2d.so")
STAGE_BCS = ((-4, -4), (-2, -7), (-7, -24), (-24, -2), (-1, -1))


@pytest.fixture(scope="module")
def host():
    L = C.CDLL(HOST_SO)
    P = C.POINTER(CellVarStru)
    for name in ("bound_cond_slope_limiter_x", "bound_cond_slope_limiter_y"):
        fn = getattr(L, name)
        fn.restype = C.c_bool
        fn.argtypes = [C.c_int, C.c_int, C.c_int, P] + [C.c_void_p] * 4 + [C.c_bool, C.c_bool, C.c_double]
    for name in ("flux_generator_x", "flux_generator_y"):
        fn = getattr(L, name)
        fn.restype = C.c_int
        fn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, P, C.c_void_p, C.c_void_p, C.c_bool]
    L.cfg = (C.c_double * 400).in_dll(L, "config")
    return L


class Grid:
    """CV[0], CV[1] with row pointers into numpy arrays (fields rho u v p)."""

    def __init__(self, m, n):
        self.m, self.n = m, n
        self.CV = (CellVarStru * 2)()
        self.keep = []
        self.frames = np.zeros((2, 4, m, n))
        self.s, self.t = np.zeros((4, m, n)), np.zeros((4, m, n))
        self.F, self.Ix = np.zeros((4, m + 1, n)), np.zeros((4, m + 1, n))
        self.G, self.Iy = np.zeros((4, m, n + 1)), np.zeros((4, m, n + 1))
        for k in range(2):
            for q, name in enumerate(("RHO", "U", "V", "P")):
                setattr(self.CV[k], name, self._rows(self.frames[k, q]))
        for q, (sn, tn) in enumerate((("s_rho", "t_rho"), ("s_u", "t_u"), ("s_v", "t_v"), ("s_p", "t_p"))):
            setattr(self.CV[0], sn, self._rows(self.s[q]))
            setattr(self.CV[0], tn, self._rows(self.t[q]))
        for q, (fx, ix, gy, iy) in enumerate((("F_rho", "rhoIx", "G_rho", "rhoIy"), ("F_u", "uIx", "G_u", "uIy"),
                                              ("F_v", "vIx", "G_v", "vIy"), ("F_e", "pIx", "G_e", "pIy"))):
            setattr(self.CV[0], fx, self._rows(self.F[q]))
            setattr(self.CV[0], ix, self._rows(self.Ix[q]))
            setattr(self.CV[0], gy, self._rows(self.G[q]))
            setattr(self.CV[0], iy, self._rows(self.Iy[q]))

    def _rows(self, a):
        rows = (DP * a.shape[0])(*[a[j].ctypes.data_as(DP) for j in range(a.shape[0])])
        self.keep.append(rows)
        return C.cast(rows, DPP)


def _bfv(arr):
    a = np.ascontiguousarray(arr, dtype=np.float64).copy()
    return a, a.ctypes.data_as(C.c_void_p)


def _set_config(host, **kw):
    cfg = make_config(**kw)
    for k in range(400):
        host.cfg[k] = float(cfg[k])


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)


def test_limiters_every_boundary_pair(host, golden_stages):
    g = golden_stages
    m, n = g["f0"].shape[1:]
    k = 0
    for bx, by in STAGE_BCS:
        for direction in (0, 1):
            for found, slope in ((0, 1), (1, 1), (0, 0)):
                meta = g["lim_meta"][k]
                assert tuple(meta[:5]) == (bx, by, direction, found, slope)
                _set_config(host, h_x=1 / m, h_y=1 / n, bc_x=bx, bc_y=by)
                grid = Grid(m, n)
                grid.frames[0], grid.frames[1] = g["f0"], g["f1"]
                grid.s[:], grid.t[:] = g["s"], g["t"]
                b = {key: _bfv(g["bfv_" + key]) for key in ("L", "R", "D", "U")}
                fn = host.bound_cond_slope_limiter_y if direction else host.bound_cond_slope_limiter_x
                print(f"limiter dir {direction} bc ({bx}, {by}) found {found} slope {slope}", flush=True)
                r = fn(m, n, 1, grid.CV, *(b[key][1] for key in ("L", "R", "D", "U")), bool(found), bool(slope), 0.0)
                assert int(r) == meta[5], k
                got_s = grid.t if direction else grid.s
                assert np.array_equal(got_s, g["lim_s"][k]), k
                other = grid.s if direction else grid.t
                assert np.array_equal(other, g["s"] if direction else g["t"]), k    # untouched
                got_b = np.concatenate([b[key][0] for key in ("L", "R", "D", "U")])
                assert np.array_equal(got_b, g["lim_bfv"][k]), k
                k += 1


@pytest.mark.parametrize("direction", [0, 1], ids=["x", "y"])
@pytest.mark.parametrize("transversal", [0, 1])
def test_flux_generators(host, golden_stages, direction, transversal):
    g = golden_stages
    m, n = g["f0"].shape[1:]
    _set_config(host, h_x=1 / m, h_y=1 / n)
    host.cfg[17] = host.cfg[18] = -4
    grid = Grid(m, n)
    grid.frames[1] = g["f1"]
    grid.s[:], grid.t[:] = g["s"] * 0.1, g["t"] * 0.1
    lo, hi = (_bfv(g["bfv_L"]), _bfv(g["bfv_R"])) if direction == 0 else (_bfv(g["bfv_D"]), _bfv(g["bfv_U"]))
    fn = host.flux_generator_y if direction else host.flux_generator_x
    r = fn(m, n, 1, float(g["flux_tau"]), grid.CV, lo[1], hi[1], bool(transversal))
    meta = g["flux_meta"][2 * direction + transversal]
    assert tuple(meta[:2]) == (direction, transversal) and r == meta[2]
    key = "y" if direction else "x"
    F = (grid.G if direction else grid.F).reshape(4, -1)
    I = (grid.Iy if direction else grid.Ix).reshape(4, -1)
    dF = _rel(F, g["flux_F_" + key][transversal])
    dI = _rel(I, g["flux_I_" + key][transversal])
    print(f"flux_generator_{key} transversal={transversal}: max rel F {dF.max():.2e} I {dI.max():.2e}", flush=True)
    assert dF.max() <= 1e-10 and dI.max() <= 1e-10
    # the ghost arrays are read only
    assert np.array_equal(lo[0], g["bfv_L" if direction == 0 else "bfv_D"])


def test_flux_generator_failing_cell(host, golden_stages):
    """A cell with negative pressure: the return value is the reference's, the
    failing interfaces hold NaN, every other interface its flux."""
    g = golden_stages
    m, n = g["f0"].shape[1:]
    _set_config(host, h_x=1 / m, h_y=1 / n)
    for direction in (0, 1):
        grid = Grid(m, n)
        grid.frames[1] = g["fail_f1"]
        grid.s[:], grid.t[:] = g["s"] * 0.1, g["t"] * 0.1
        lo, hi = (_bfv(g["bfv_L"]), _bfv(g["bfv_R"])) if direction == 0 else (_bfv(g["bfv_D"]), _bfv(g["bfv_U"]))
        fn = host.flux_generator_y if direction else host.flux_generator_x
        r = fn(m, n, 1, float(g["flux_tau"]), grid.CV, lo[1], hi[1], True)
        assert r == g["fail_rc"][direction]
        F = grid.G if direction else grid.F
        bad = ~np.isfinite(F[0])
        assert bad.sum() == 2          # the two interfaces of the failing cell in this direction
