"""ctypes bindings used by the tests.

* ``Oracle``  -- oracle/liboracle.so, the CPU restatement of the reference
  (test infrastructure: the checker, never the thing measured).
* ``RefLib``  -- oracle/_ref/libhydroref.so, the reference's own sources
  compiled in this container (absent on machines without /root/reference).
* ``hydrogrp`` -- the product library (HIP), imported from the package.
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
ORACLE_DEVPOW_SO = os.path.join(ROOT, "oracle", "liboracle_devpow.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libhydroref.so")
REF_MF_SO = os.path.join(ROOT, "oracle", "_ref", "libhydroref_mf.so")
REF_G2D_SO = os.path.join(ROOT, "oracle", "_ref", "libhydroref_g2d.so")
PKG_DIR = os.path.join(ROOT, "hydrocode-fusion_amd")
if PKG_DIR not in sys.path:
    sys.path.insert(0, PKG_DIR)

D = C.c_double
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(DP)


class OrcIfv(C.Structure):
    _fields_ = [(n, D) for n in (
        "gamma", "lambda_u", "lambda_v", "rho", "u", "v", "p",
        "d_rho", "d_u", "d_v", "d_p", "t_rho", "t_u", "t_v", "t_p")]


def ifv_from(vec):
    f = OrcIfv()
    for k, (name, _) in enumerate(OrcIfv._fields_):
        setattr(f, name, float(vec[k]))
    return f


N_CONF = 400


def make_config(**kw):
    """config[] with the reference's defaults for the keys this path reads."""
    cfg = np.full(N_CONF, np.inf)
    cfg[0] = 2.0
    named = dict(t_all=1, eps=4, N=5, gamma=6, CFL=7, h_x=10, h_y=11, n_x=13, n_y=14,
                 tau=16, bc_x=17, bc_y=18, split=33, alpha=41)
    defaults = dict(eps=1e-9, gamma=1.4, CFL=0.45, alpha=1.9)
    defaults.update(kw)
    for k, v in defaults.items():
        cfg[named[k]] = v
    if np.isfinite(cfg[1]) and not np.isfinite(cfg[5]):
        cfg[5] = float(2**31 - 1)
    return cfg


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.lib = C.CDLL(path)
        L.orc_riemann_exact.restype = D
        L.orc_riemann_exact.argtypes = [DP, DP, D, D, D, D, D, D, D, D, IP, D, D, C.c_int, IP]
        L.orc_riemann_exact_rule.restype = D
        L.orc_riemann_exact_rule.argtypes = [DP, DP, D, D, D, D, D, D, D, D, IP, D, D, C.c_int, IP, C.c_int]
        L.orc_set_newton.argtypes = [D, C.c_int]
        L.orc_set_flux_solver.argtypes = [C.c_int]
        L.orc_riemann_exact_ben.restype = D
        L.orc_riemann_exact_ben.argtypes = [DP, DP, D, D, D, D, D, D, D, IP, D, D, C.c_int, IP]
        L.orc_grp_q1d.restype = None
        L.orc_grp_q1d.argtypes = [DP, DP, DP, DP, C.POINTER(OrcIfv), C.POINTER(OrcIfv), D, D]
        L.orc_grp_2d_flux.restype = C.c_int
        L.orc_grp_2d_flux.argtypes = [OrcIfv, OrcIfv, D, D, D, D, DP, DP]
        L.orc_grp_g2d.restype = None
        L.orc_grp_g2d.argtypes = [DP, DP, DP, DP, C.POINTER(OrcIfv), C.POINTER(OrcIfv), DP, DP, D, D]
        L.orc_guessp.restype = D
        L.orc_guessp.argtypes = [D] * 11
        L.orc_grp_edir.restype = None
        L.orc_grp_edir.argtypes = [DP, DP, C.POINTER(OrcIfv), C.POINTER(OrcIfv), D, D]
        PP = C.POINTER(DP)
        for fn in (L.orc_split2d, L.orc_unsplit2d):
            fn.restype = C.c_int
            fn.argtypes = [C.c_int, C.c_int, PP, PP, PP, PP, PP, C.c_int, IP, DP, DP, DP]
        L.orc_grp1d.restype = C.c_int
        L.orc_grp1d.argtypes = [C.c_int, PP, PP, PP, PP, IP, DP, DP, DP]
        L.orc_pow_many.argtypes = [DP, DP, C.c_int, DP]

    def libm_pow(self, x, y):
        """glibc's pow elementwise (numpy's power is a different implementation)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.empty_like(x)
        self.lib.orc_pow_many(dptr(x), dptr(y), len(x), dptr(out))
        return out

    def riemann(self, gL, gR, uL, uR, pL, pR, cL, cR, eps, tol, N):
        us, ps = D(), D()
        crw = (C.c_int * 2)()
        nit = C.c_int()
        gap = self.lib.orc_riemann_exact(C.byref(us), C.byref(ps), gL, gR, uL, uR, pL, pR, cL, cR,
                                         crw, eps, tol, N, C.byref(nit))
        return us.value, ps.value, crw[0], crw[1], gap, nit.value

    def riemann_rule(self, gL, gR, uL, uR, pL, pR, cL, cR, eps, tol, N, rule):
        """riemann() with a stopping rule (0 the reference's, 1 the converge rule)."""
        us, ps = D(), D()
        crw = (C.c_int * 2)()
        nit = C.c_int()
        gap = self.lib.orc_riemann_exact_rule(C.byref(us), C.byref(ps), gL, gR, uL, uR, pL, pR, cL, cR,
                                              crw, eps, tol, N, C.byref(nit), rule)
        return us.value, ps.value, crw[0], crw[1], gap, nit.value

    def riemann_ben(self, g, uL, uR, pL, pR, cL, cR, eps, tol, N):
        us, ps = D(), D()
        crw = (C.c_int * 2)()
        nit = C.c_int()
        gap = self.lib.orc_riemann_exact_ben(C.byref(us), C.byref(ps), g, uL, uR, pL, pR, cL, cR,
                                             crw, eps, tol, N, C.byref(nit))
        return us.value, ps.value, crw[0], crw[1], gap, nit.value

    def guessp(self, row):
        return self.lib.orc_guessp(*[float(x) for x in row])

    def q1d(self, Lv, Rv, eps, atc):
        ws, Dd, U, Us = (D * 2)(), (D * 4)(), (D * 4)(), (D * 6)()
        self.lib.orc_grp_q1d(ws, Dd, U, Us, C.byref(ifv_from(Lv)), C.byref(ifv_from(Rv)), eps, atc)
        return np.array(list(ws) + list(Dd) + list(U) + list(Us))

    def g2d(self, Lv, Rv, eps, atc):
        """linear_GRP_solver_Edir_G2D on sides of 15 doubles (the single-fluid
        build) or 21 (the multi-fluid members z_a d_z_a t_z_a phi d_phi t_phi
        appended).  Returns wave_speed[2], D[6], U[6], U_star[6]."""
        ws, Dd, U, Us = (D * 2)(), (D * 6)(), (D * 6)(), (D * 6)()
        zf = [None, None]
        if len(Lv) == 21:
            zf = [(D * 6)(*[float(x) for x in Lv[15:21]]), (D * 6)(*[float(x) for x in Rv[15:21]])]
        self.lib.orc_grp_g2d(ws, Dd, U, Us, C.byref(ifv_from(Lv)), C.byref(ifv_from(Rv)), zf[0], zf[1], eps, atc)
        return np.array(list(ws) + list(Dd) + list(U) + list(Us))

    def flux2d(self, Lv, Rv, tau, eps, nx, ny):
        F, I = (D * 4)(), (D * 4)()
        rc = self.lib.orc_grp_2d_flux(ifv_from(Lv), ifv_from(Rv), tau, eps, nx, ny, F, I)
        return rc, np.array(list(F) + list(I))

    def edir(self, Lv, Rv, eps, atc):
        Dd, U = (D * 3)(), (D * 3)()
        self.lib.orc_grp_edir(Dd, U, C.byref(ifv_from(Lv)), C.byref(ifv_from(Rv)), eps, atc)
        return np.array(list(Dd) + list(U))

    def split2d(self, m, n, init, config, time_plot=None, N_T=None, newton_tol=0.0, converge=False):
        """init: dict rho,u,v,p (m x n arrays, [j, i]), optionally e (else formed as the
        reference's main does); returns (frames, N_plot, time_plot, config, status).
        newton_tol (<= 0: eps) / converge: the Newton settings of the solves (the
        defaults are the reference's)."""
        return self._run2d(self.lib.orc_split2d, m, n, init, config, time_plot, N_T, newton_tol, converge)

    def unsplit2d(self, m, n, init, config, time_plot=None, N_T=None, newton_tol=0.0, converge=False, g2d=False):
        """GRP_solver_2D_EUL_source (the unsplit step); arguments and results as split2d.
        g2d: GRP_2D_flux with linear_GRP_solver_Edir_G2D (its alternative solver)."""
        return self._run2d(self.lib.orc_unsplit2d, m, n, init, config, time_plot, N_T, newton_tol, converge, g2d)

    def flux2d_g2d(self, Lv, Rv, tau, eps, nx, ny):
        """flux2d with GRP_2D_flux's alternative solver, linear_GRP_solver_Edir_G2D."""
        self.lib.orc_set_flux_solver(1)
        try:
            return self.flux2d(Lv, Rv, tau, eps, nx, ny)
        finally:
            self.lib.orc_set_flux_solver(0)

    def _run2d(self, fn, m, n, init, config, time_plot, N_T, newton_tol=0.0, converge=False, g2d=False):
        if time_plot is None:
            time_plot = np.array([0.0, config[1]])
        time_plot = np.ascontiguousarray(time_plot, dtype=np.float64)
        N_plot = len(time_plot)
        if N_T is None:
            N_T = min(5, N_plot)
        gamma = config[6]
        frames = np.zeros((5, N_T, m, n))
        r, u, v, p = (np.asarray(init[k], dtype=np.float64) for k in ("rho", "u", "v", "p"))
        frames[0, 0], frames[1, 0], frames[2, 0], frames[3, 0] = r, u, v, p
        frames[4, 0] = init["e"] if init.get("e") is not None else initial_energy(r, u, v, p, gamma)
        ptrs = [(DP * N_T)(*[frames[f, k].ctypes.data_as(DP) for k in range(N_T)]) for f in range(5)]
        cfg = np.array(config, dtype=np.float64)
        npl = C.c_int(N_plot)
        cpu = np.zeros(N_T)
        self.lib.orc_set_newton(newton_tol, 1 if converge else 0)
        self.lib.orc_set_flux_solver(1 if g2d else 0)
        try:
            st = fn(m, n, *ptrs, N_T, C.byref(npl), dptr(time_plot), dptr(cpu), dptr(cfg))
        finally:
            self.lib.orc_set_newton(0.0, 0)
            self.lib.orc_set_flux_solver(0)
        return frames, npl.value, time_plot, cfg, st

    def grp1d(self, m, init, config, N_T=2, time_plot=None):
        if time_plot is None:
            time_plot = np.array([0.0, config[1]])
        time_plot = np.ascontiguousarray(time_plot, dtype=np.float64)
        gamma = config[6]
        frames = np.zeros((4, N_T, m))
        r, u, p = (np.asarray(init[k], dtype=np.float64) for k in ("rho", "u", "p"))
        frames[0, 0], frames[1, 0], frames[2, 0] = r, u, p
        frames[3, 0] = 0.5 * u * u + p / (gamma - 1.0) / r
        ptrs = [(DP * N_T)(*[frames[f, k].ctypes.data_as(DP) for k in range(N_T)]) for f in range(4)]
        cfg = np.array(config, dtype=np.float64)
        npl = C.c_int(N_T)
        cpu = np.zeros(N_T)
        st = self.lib.orc_grp1d(m, *ptrs, C.byref(npl), dptr(time_plot), dptr(cpu), dptr(cfg))
        return frames, npl.value, time_plot, cfg, st


def initial_energy(r, u, v, p, gamma):
    """E exactly as src/hydrocode_2D/hydrocode.c:251-252 forms it."""
    e = 0.5 * u * u + p / (gamma - 1.0) / r
    e = e + 0.5 * v * v
    return e


# struct i_f_var of a single-fluid build, member order (src/include/var_struc.h:116-131)
IFV_MEMBERS = ("n_x", "n_y", "delta_x", "delta_y", "length", "lambda_u", "lambda_v", "gamma",
               "RHO", "P", "U", "V", "RHO_int", "P_int", "U_int", "V_int", "F_rho", "F_e", "F_u", "F_v",
               "U_rho", "U_e", "U_u", "U_v", "d_rho", "d_e", "d_u", "d_v", "d_p", "t_rho", "t_e", "t_u", "t_v", "t_p")
# the 15-double side vectors (gamma, lambda_u, lambda_v, rho, u, v, p, d_*, t_*) as member names
SIDE15_MEMBERS = ("gamma", "lambda_u", "lambda_v", "RHO", "U", "V", "P", "d_rho", "d_u", "d_v", "d_p",
                  "t_rho", "t_u", "t_v", "t_p")


class IfvStruct(C.Structure):
    """struct i_f_var (single-fluid members) as include/var_struc.h lays it out."""
    _fields_ = [(name, C.c_double) for name in IFV_MEMBERS]


def ifv_members(side15, n_x=0.0, n_y=0.0, extra=None):
    """34-member vector from a 15-double side; other members from `extra` (dict) or 0."""
    a = np.zeros(34)
    for k, name in enumerate(SIDE15_MEMBERS):
        a[IFV_MEMBERS.index(name)] = side15[k]
    a[0], a[1] = n_x, n_y
    for name, v in (extra or {}).items():
        a[IFV_MEMBERS.index(name)] = v
    return a


class RefLib:
    """The reference's own sources built in oracle/_ref (checker only)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        L.ref_config_reset.restype = None
        L.ref_config_set.argtypes = [C.c_int, D]
        L.ref_config_get.argtypes = [C.c_int]
        L.ref_config_get.restype = D
        L.ref_riemann_exact.argtypes = [DP, C.c_int, DP]
        L.ref_riemann_exact_ben.argtypes = [DP, C.c_int, DP]
        L.ref_q1d.argtypes = [DP, DP, D, D, DP]
        L.ref_g2d.argtypes = [DP, DP, D, D, DP]
        L.ref_grp_2d_flux.argtypes = [DP, DP, D, D, D, DP]
        L.ref_grp_2d_flux.restype = C.c_int
        L.ref_grp_edir.argtypes = [DP, DP, D, D, DP]
        L.ref_grp_2d_flux_struct.argtypes = [DP, DP, D, DP, DP]
        L.ref_grp_2d_flux_struct.restype = C.c_int
        L.ref_guessp.argtypes = [DP]
        L.ref_guessp.restype = D
        L.ref_limiter.argtypes = [C.c_int, C.c_int, C.c_int, DP, DP, DP, DP, DP, DP, DP, DP, C.c_int, C.c_int]
        L.ref_limiter.restype = C.c_int
        L.ref_flux_gen.argtypes = [C.c_int, C.c_int, C.c_int, DP, DP, DP, DP, DP, D, C.c_int, DP, DP]
        L.ref_flux_gen.restype = C.c_int
        L.ref_split2d.argtypes = [C.c_int, C.c_int, DP, C.c_int, IP, DP]
        L.ref_unsplit2d.argtypes = [C.c_int, C.c_int, DP, C.c_int, IP, DP]
        L.ref_grp1d.argtypes = [C.c_int, DP, C.c_int, IP, DP]

    def set_config(self, cfg):
        self.lib.ref_config_reset()
        for k, v in enumerate(cfg):
            self.lib.ref_config_set(k, float(v))

    def riemann(self, inp, N):
        inp = np.ascontiguousarray(inp, dtype=np.float64)
        out = np.zeros(5)
        self.lib.ref_riemann_exact(dptr(inp), N, dptr(out))
        return out

    def riemann_ben(self, inp, N):
        inp = np.ascontiguousarray(inp, dtype=np.float64)
        out = np.zeros(5)
        self.lib.ref_riemann_exact_ben(dptr(inp), N, dptr(out))
        return out

    def q1d(self, Lv, Rv, eps, atc):
        Lv = np.ascontiguousarray(Lv, dtype=np.float64)
        Rv = np.ascontiguousarray(Rv, dtype=np.float64)
        out = np.zeros(16)
        self.lib.ref_q1d(dptr(Lv), dptr(Rv), eps, atc, dptr(out))
        return out

    def g2d(self, Lv, Rv, eps, atc):
        out = np.zeros(20)
        self.lib.ref_g2d(dptr(np.ascontiguousarray(Lv, dtype=np.float64)),
                         dptr(np.ascontiguousarray(Rv, dtype=np.float64)), eps, atc, dptr(out))
        return out

    def flux2d(self, Lv, Rv, tau, eps, nx, ny):
        self.lib.ref_config_set(4, eps)
        Lv = np.ascontiguousarray(Lv, dtype=np.float64)
        Rv = np.ascontiguousarray(Rv, dtype=np.float64)
        out = np.zeros(8)
        rc = self.lib.ref_grp_2d_flux(dptr(Lv), dptr(Rv), nx, ny, tau, dptr(out))
        return rc, out

    def edir(self, Lv, Rv, eps, atc):
        Lv = np.ascontiguousarray(Lv, dtype=np.float64)
        Rv = np.ascontiguousarray(Rv, dtype=np.float64)
        out = np.zeros(6)
        self.lib.ref_grp_edir(dptr(Lv), dptr(Rv), eps, atc, dptr(out))
        return out

    def flux2d_struct(self, Lm, Rm, tau, eps):
        """GRP_2D_flux on two whole structs (IFV_MEMBERS order, 34 doubles each):
        returns (code, L after, R after)."""
        self.lib.ref_config_set(4, eps)
        Lm = np.ascontiguousarray(Lm, dtype=np.float64)
        Rm = np.ascontiguousarray(Rm, dtype=np.float64)
        Lo, Ro = np.zeros(34), np.zeros(34)
        rc = self.lib.ref_grp_2d_flux_struct(dptr(Lm), dptr(Rm), tau, dptr(Lo), dptr(Ro))
        return rc, Lo, Ro

    def guessp(self, row):
        """GuessP (riemann_solver_starPU.c:20-58): row = DL DR UL UR PL PR CL CR GammaL GammaR eps."""
        row = np.ascontiguousarray(row, dtype=np.float64)
        return self.lib.ref_guessp(dptr(row))

    def limiter(self, direction, m, n, f0, f1, slopes, cross, bfv, found, slope=1):
        """bound_cond_slope_limiter_{x,y} with nt = 1: f0, f1 (4, m, n) frames; slopes
        (4, m, n) of the direction (previous values in); cross: the other direction's;
        bfv: dict L, R (n, 13), D, U (m, 13).  Returns (value, slopes, bfv) after the call."""
        sl = np.ascontiguousarray(slopes, dtype=np.float64).copy()
        cr = np.ascontiguousarray(cross, dtype=np.float64).copy()
        b = {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in bfv.items()}
        f0 = np.ascontiguousarray(f0, dtype=np.float64).copy()
        f1 = np.ascontiguousarray(f1, dtype=np.float64).copy()
        r = self.lib.ref_limiter(direction, m, n, dptr(f0), dptr(f1), dptr(sl), dptr(cr), dptr(b["L"]), dptr(b["R"]),
                                 dptr(b["D"]), dptr(b["U"]), int(found), int(slope))
        return r, sl, b

    def flux_gen(self, direction, m, n, f1, s, t, bfv_lo, bfv_hi, tau, transversal):
        """flux_generator_{x,y} with nt = 1: returns (value, F (4, rows, cols), I (4, rows, cols))."""
        rows, cols = (m, n + 1) if direction else (m + 1, n)
        F, I = np.zeros((4, rows, cols)), np.zeros((4, rows, cols))
        arrs = [np.ascontiguousarray(a, dtype=np.float64).copy() for a in (f1, s, t, bfv_lo, bfv_hi)]
        r = self.lib.ref_flux_gen(direction, m, n, *(dptr(a) for a in arrs), tau, int(transversal), dptr(F), dptr(I))
        return r, F, I

    def split2d(self, m, n, init, config, time_plot=None, N_T=None):
        return self._run2d(self.lib.ref_split2d, m, n, init, config, time_plot, N_T)

    def unsplit2d(self, m, n, init, config, time_plot=None, N_T=None):
        return self._run2d(self.lib.ref_unsplit2d, m, n, init, config, time_plot, N_T)

    def _run2d(self, fn, m, n, init, config, time_plot, N_T):
        if time_plot is None:
            time_plot = np.array([0.0, config[1]])
        time_plot = np.ascontiguousarray(time_plot, dtype=np.float64)
        N_plot = len(time_plot)
        if N_T is None:
            N_T = min(5, N_plot)
        gamma = config[6]
        frames = np.zeros((5, N_T, m, n))
        r, u, v, p = (np.asarray(init[k], dtype=np.float64) for k in ("rho", "u", "v", "p"))
        frames[0, 0], frames[1, 0], frames[2, 0], frames[3, 0] = r, u, v, p
        frames[4, 0] = initial_energy(r, u, v, p, gamma)
        self.set_config(config)
        npl = C.c_int(N_plot)
        fn(m, n, dptr(frames), N_T, C.byref(npl), dptr(time_plot))
        cfg = np.array([self.lib.ref_config_get(k) for k in range(N_CONF)])
        return frames, npl.value, time_plot, cfg

    def grp1d(self, m, init, config, N_T=2, time_plot=None):
        if time_plot is None:
            time_plot = np.array([0.0, config[1]])
        time_plot = np.ascontiguousarray(time_plot, dtype=np.float64)
        gamma = config[6]
        frames = np.zeros((4, N_T, m))
        r, u, p = (np.asarray(init[k], dtype=np.float64) for k in ("rho", "u", "p"))
        frames[0, 0], frames[1, 0], frames[2, 0] = r, u, p
        frames[3, 0] = 0.5 * u * u + p / (gamma - 1.0) / r
        cfg = np.array(config, dtype=np.float64)
        cfg[0] = 1.0
        self.set_config(cfg)
        npl = C.c_int(N_T)
        self.lib.ref_grp1d(m, dptr(frames), N_T, C.byref(npl), dptr(time_plot))
        cfg = np.array([self.lib.ref_config_get(k) for k in range(N_CONF)])
        return frames, npl.value, time_plot, cfg


class RefLibMF:
    """The reference's interface solvers built with -DMULTIFLUID_BASICS in
    oracle/_ref/libhydroref_mf.so (checker only); sides of 21 doubles
    (oracle/ref_glue_mf.c)."""

    def __init__(self, path=REF_MF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        L.refmf_config_reset.restype = None
        L.refmf_config_set.argtypes = [C.c_int, D]
        L.refmf_q1d.argtypes = [DP, DP, D, D, DP]
        L.refmf_grp_2d_flux.argtypes = [DP, DP, D, D, D, DP]
        L.refmf_g2d.argtypes = [DP, DP, D, D, DP]

    def set_config(self, values):
        self.lib.refmf_config_reset()
        for k, v in values.items():
            self.lib.refmf_config_set(k, float(v))

    def q1d(self, Lv, Rv, eps, atc):
        out = np.zeros(20)
        self.lib.refmf_q1d(dptr(np.ascontiguousarray(Lv, dtype=np.float64)),
                           dptr(np.ascontiguousarray(Rv, dtype=np.float64)), eps, atc, dptr(out))
        return out

    def g2d(self, Lv, Rv, eps, atc):
        out = np.zeros(20)
        self.lib.refmf_g2d(dptr(np.ascontiguousarray(Lv, dtype=np.float64)),
                           dptr(np.ascontiguousarray(Rv, dtype=np.float64)), eps, atc, dptr(out))
        return out

    def grp_2d_flux(self, Lv, Rv, n_x, n_y, tau):
        out = np.zeros(13)
        self.lib.refmf_grp_2d_flux(dptr(np.ascontiguousarray(Lv, dtype=np.float64)),
                                   dptr(np.ascontiguousarray(Rv, dtype=np.float64)), n_x, n_y, tau, dptr(out))
        return out


# ---- the reference's struct layouts through ctypes (include/var_struc.h) -----------
DPP = C.POINTER(DP)


class CellVarStru(C.Structure):
    _fields_ = ([(n, DPP) for n in ("E", "RHO", "U", "V", "P")]
                + [(n, DP) for n in ("d_rho", "d_u", "d_v", "d_p")]
                + [(n, DPP) for n in ("s_rho", "s_u", "s_v", "s_p", "t_rho", "t_u", "t_v", "t_p",
                                      "rhoIx", "uIx", "vIx", "pIx", "rhoIy", "uIy", "vIy", "pIy",
                                      "F_rho", "F_e", "F_u", "F_v", "G_rho", "G_e", "G_u", "G_v")])


class FluVar(C.Structure):
    _fields_ = [(n, DP) for n in ("RHO", "U", "V", "P")]


class CVFrames:
    """N_T frames of rho u v p e as row-pointer arrays (CV[k].X[j][i]) backed by numpy."""

    ORDER = ("RHO", "U", "V", "P", "E")

    def __init__(self, m, n, N_T, h_x=1.0, h_y=1.0):
        self.m, self.n, self.N_T = m, n, N_T
        self.data = np.zeros((5, N_T, m, n))
        self.CV = (CellVarStru * N_T)()
        self._rows = []
        for k in range(N_T):
            for f, name in enumerate(self.ORDER):
                rows = (DP * m)(*[self.data[f, k, j].ctypes.data_as(DP) for j in range(m)])
                self._rows.append(rows)
                setattr(self.CV[k], name, C.cast(rows, DPP))
        self.Xa = np.array([[j * h_x for _ in range(n + 1)] for j in range(m + 1)])
        self.Ya = np.array([[i * h_y for i in range(n + 1)] for _ in range(m + 1)])
        self.X = (DP * (m + 1))(*[self.Xa[j].ctypes.data_as(DP) for j in range(m + 1)])
        self.Y = (DP * (m + 1))(*[self.Ya[j].ctypes.data_as(DP) for j in range(m + 1)])

    def set_initial(self, init, gamma):
        r, u, v, p = (np.asarray(init[k], dtype=np.float64) for k in ("rho", "u", "v", "p"))
        self.data[0, 0], self.data[1, 0], self.data[2, 0], self.data[3, 0] = r, u, v, p
        self.data[4, 0] = initial_energy(r, u, v, p, gamma)

    def frames(self):
        """(5, N_T, m, n) in the oracle's field order rho u v p e."""
        return self.data
