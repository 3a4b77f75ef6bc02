"""Python access to libhydrogrp.so (the gfx950 Eulerian GRP library) via ctypes.

The library's C ABI is declared in include/hydrocode_grp.h and
include/hydrocode_dropin.h; this module only marshals numpy arrays.  There is
no CPU fallback: if the shared library is missing, importing ``lib()`` raises.
"""
import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libhydrogrp.so")
N_CONF = 400

_D, _DP, _IP = C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int)
# stopping rules of the Newton solve for p* (HGRP_NEWTON_* in include/hydrocode_grp.h)
NEWTON_REFERENCE, NEWTON_CONVERGE = 0, 1
UNSPLIT_AUTO, UNSPLIT_TILE, UNSPLIT_STRIP = 0, 1, 2     # include/hydrocode_grp.h HGRP_UNSPLIT_*


class Result(C.Structure):
    _fields_ = [("k", C.c_int), ("nt", C.c_int), ("status", C.c_int), ("sweeps", C.c_int),
                ("time_c", _D), ("tau", _D), ("seconds", _D)]


FRAME_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, _D)

_lib = None


def lib():
    """Load libhydrogrp.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C hydrocode-fusion_amd` "
                           "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    L.hgrp2d_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, _DP]
    L.hgrp2d_destroy.argtypes = [C.c_void_p]
    L.hgrp2d_upload.argtypes = [C.c_void_p, _DP, _DP, _DP, _DP, _DP]
    L.hgrp2d_download.argtypes = [C.c_void_p, _DP, _DP, _DP, _DP, _DP]
    L.hgrp2d_solve.argtypes = [C.c_void_p, _DP, C.c_int, C.c_int, FRAME_FN, C.c_void_p, C.POINTER(Result)]
    L.hgrp2d_bench.argtypes = [C.c_void_p, C.c_int, C.c_int, _DP, _DP, _DP]
    L.hgrp2d_bench_unsplit.argtypes = [C.c_void_p, C.c_int, C.c_int, _DP, _DP]
    L.hgrp2d_set_sweeps.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.hgrp2d_set_interface_solver.argtypes = [C.c_void_p, C.c_int]
    L.hgrp2d_set_unsplit_form.argtypes = [C.c_void_p, C.c_int]
    L.hgrp2d_tune.argtypes = [C.c_void_p, C.c_int, _IP, _IP, _DP]
    L.hgrp2d_h_over_smax.argtypes = [C.c_void_p, _DP]
    L.hgrp_grp2d_flux_batch.argtypes = [_DP, _DP, C.c_int, _D, _D, C.c_int, _DP]
    L.hgrp_riemann_batch.argtypes = [_DP, C.c_int, C.c_int, _DP]
    L.hgrp_grp2d_flux_tf_batch.argtypes = [_DP, _DP, C.c_int, _D, _D, C.c_int, _D, _D, _DP]
    L.hgrp_pow_batch.argtypes = [_DP, _DP, C.c_int, _DP]
    L.hgrp_q1d_batch.argtypes = [_DP, _DP, C.c_int, _D, _D, _DP]
    L.hgrp_g2d_batch.argtypes = [_DP, _DP, C.c_int, _D, _D, _DP]
    L.hgrp_guessp_batch.argtypes = [_DP, C.c_int, _DP]
    L.hgrp_riemann_pvrs_batch.argtypes = [_DP, C.c_int, C.c_int, _DP]
    L.hgrp_grp2d_flux_general_batch.argtypes = [_DP, _DP, C.c_int, _D, _D, _D, _D, _DP]
    L.hgrp_edir_batch.argtypes = [_DP, _DP, C.c_int, _D, _D, _DP]
    L.hgrp2db_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, _IP, _IP, _DP, _DP, _DP]
    L.hgrp2db_destroy.argtypes = [C.c_void_p]
    L.hgrp2db_upload.argtypes = [C.c_void_p, _DP, _DP, _DP, _DP]
    L.hgrp2db_set_newton.argtypes = [C.c_void_p, _D, C.c_int, C.POINTER(C.c_float)]
    L.hgrp2db_set_sweeps.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.hgrp2db_tune.argtypes = [C.c_void_p, C.c_int, _IP, _IP, _DP]
    L.hgrp2db_run.argtypes = [C.c_void_p, _DP, C.POINTER(C.c_longlong)]
    L.hgrp2db_results.argtypes = [C.c_void_p, _IP, _IP, _DP, _IP]
    L.hgrp2db_download.argtypes = [C.c_void_p, _DP, _DP, _DP, _DP, _DP]
    L.hgrp_device_name.argtypes = [C.c_char_p, C.c_int]
    L.hgrp_copy_probe.argtypes = [C.c_size_t, C.c_int, _DP]
    _FP = C.POINTER(C.c_float)
    L.hgrp2d_set_newton.argtypes = [C.c_void_p, _D, C.c_int, _FP]
    L.hgrp2d_set_fluids.argtypes = [C.c_void_p, _D, _DP]
    L.hgrp2d_download_fraction.argtypes = [C.c_void_p, _DP]
    L.hgrp2d_checkpoint_size.argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
    L.hgrp2d_checkpoint.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.hgrp2d_restore.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.hgrp2d_sweep_kernel.argtypes = [C.c_void_p, C.c_int]
    L.hgrp2d_sweep_kernel.restype = C.c_char_p
    L.hgrp_mlp_data.argtypes = [C.c_ulonglong, C.c_int, _D, _FP, _FP, _DP, _IP]
    L.hgrp_mlp_loss_grad.argtypes = [_FP, _FP, _FP, C.c_int, _DP, _FP]
    L.hgrp_mlp_infer.argtypes = [_FP, _FP, C.c_int, _FP]
    L.hgrp_mlp_train.argtypes = [C.c_ulonglong, C.c_int, C.c_int, C.c_int, _D, _D, _FP, _FP]
    L.hgrp_riemann_guess_batch.argtypes = [_DP, C.c_int, C.c_int, C.c_int, _FP, _DP]
    L.hgrp1d_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, _IP, _DP, _DP]
    L.hgrp1d_destroy.argtypes = [C.c_void_p]
    L.hgrp1d_upload.argtypes = [C.c_void_p, _DP, _DP, _DP, _DP]
    L.hgrp1d_download.argtypes = [C.c_void_p, _DP, _DP, _DP, _DP]
    L.hgrp1d_run.argtypes = [C.c_void_p, C.c_int, _DP, C.POINTER(C.c_longlong)]
    L.hgrp1d_results.argtypes = [C.c_void_p, _IP, _IP, _DP]
    L.hgrp1d_run_until.argtypes = [C.c_void_p, C.c_int, _DP, _IP]
    L.hgrp1d_tune.argtypes = [C.c_void_p, C.c_int, _IP, _IP]
    L.hgrp1d_set_blocks.argtypes = [C.c_void_p, C.c_int, C.c_int]
    _lib = L
    return L


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")


def copy_probe(nbytes, reps=20):
    """Device time (ms) of one plain copy of nbytes (read nbytes + write nbytes)."""
    ms = _D()
    _check(lib().hgrp_copy_probe(nbytes, reps, C.byref(ms)), "copy probe")
    return ms.value


def device_name():
    buf = C.create_string_buffer(256)
    lib().hgrp_device_name(buf, 256)
    return buf.value.decode()


class Solver2D:
    """Dimension-split 2D GRP solver resident on the GPU (hgrp2d_* entry points)."""

    def __init__(self, m, n, config):
        self.m, self.n = m, n
        self.config = np.ascontiguousarray(config, dtype=np.float64)
        assert self.config.size == N_CONF
        self._h = C.c_void_p()
        _check(lib().hgrp2d_create(C.byref(self._h), m, n, _p(self.config)), "hgrp2d_create")

    def close(self):
        if self._h:
            lib().hgrp2d_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, rho, u, v, p, e=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(self.m, self.n) for a in (rho, u, v, p)]
        ee = None if e is None else np.ascontiguousarray(e, dtype=np.float64).reshape(self.m, self.n)
        _check(lib().hgrp2d_upload(self._h, *[_p(a) for a in arrs], _p(ee)), "hgrp2d_upload")

    def download(self):
        out = np.empty((5, self.m, self.n))
        _check(lib().hgrp2d_download(self._h, *[_p(out[f]) for f in range(5)]), "hgrp2d_download")
        return out

    def h_over_smax(self):
        h = _D()
        _check(lib().hgrp2d_h_over_smax(self._h, C.byref(h)), "hgrp2d_h_over_smax")
        return h.value

    def solve(self, time_plot, N_T=None, on_frame=None):
        """Run the reference's time loop; on_frame(nt, nt_plot, time_c) sees each plot frame."""
        tp = np.ascontiguousarray(time_plot, dtype=np.float64)
        if N_T is None:
            N_T = min(5, len(tp))
        res = Result()
        cb = FRAME_FN(lambda user, nt, ntp, tc: int(bool(on_frame(nt, ntp, tc))) if on_frame else 0)
        _check(lib().hgrp2d_solve(self._h, _p(tp), len(tp), N_T, cb, None, C.byref(res)), "hgrp2d_solve")
        return res

    def sweep_kernel(self, direction):
        """Name of the kernel launched for an x (0) or y (1) sweep of this grid."""
        return lib().hgrp2d_sweep_kernel(self._h, direction).decode()

    def set_newton(self, tol=0.0, params=None, converge=False):
        """Newton tolerance for p* (<= 0: eps, the reference's), the learned
        starting guess (params: master vector from hydrogrp.mlp, None: off) and
        the stopping rule (converge: NEWTON_CONVERGE, else the reference's)."""
        fp = None
        if params is not None:
            params = np.ascontiguousarray(params, dtype=np.float32)
            fp = params.ctypes.data_as(C.POINTER(C.c_float))
        rule = NEWTON_CONVERGE if converge else NEWTON_REFERENCE
        _check(lib().hgrp2d_set_newton(self._h, tol, rule, fp), "hgrp2d_set_newton")

    def set_interface_solver(self, g2d):
        """GRP_2D_flux's solver in the unsplit step: g2d True for
        linear_GRP_solver_Edir_G2D (the reference's alternative), False for Q1D."""
        _check(lib().hgrp2d_set_interface_solver(self._h, 1 if g2d else 0), "hgrp2d_set_interface_solver")

    def set_unsplit_form(self, form):
        """Launch form of the unsplit step: UNSPLIT_AUTO (by grid size), UNSPLIT_TILE
        or UNSPLIT_STRIP; tune() on an unsplit context picks the faster."""
        _check(lib().hgrp2d_set_unsplit_form(self._h, form), "hgrp2d_set_unsplit_form")

    def set_fluids(self, gamma_b, frac_a):
        """Two fluids: volume fraction of fluid a (gamma = config[6]); fluid b has gamma_b."""
        fa = np.ascontiguousarray(frac_a, dtype=np.float64)
        assert fa.shape == (self.m, self.n)
        _check(lib().hgrp2d_set_fluids(self._h, gamma_b, _p(fa)), "hgrp2d_set_fluids")

    def download_fraction(self):
        out = np.empty((self.m, self.n))
        _check(lib().hgrp2d_download_fraction(self._h, _p(out)), "hgrp2d_download_fraction")
        return out

    def checkpoint(self):
        """Full-precision device state (and loop position inside a frame callback) as bytes."""
        n = C.c_size_t()
        _check(lib().hgrp2d_checkpoint_size(self._h, C.byref(n)), "hgrp2d_checkpoint_size")
        buf = C.create_string_buffer(n.value)
        _check(lib().hgrp2d_checkpoint(self._h, buf, n.value), "hgrp2d_checkpoint")
        return buf.raw

    def restore(self, blob):
        buf = C.create_string_buffer(blob, len(blob))
        _check(lib().hgrp2d_restore(self._h, buf, len(blob)), "hgrp2d_restore")

    def bench(self, steps, sample_every=0):
        """(ms of all steps, average ms of one X and of one Y sweep launch, from
        events around every sample_every-th step; NaN for sample_every 0)."""
        t, x, y = _D(), _D(), _D()
        _check(lib().hgrp2d_bench(self._h, steps, sample_every, C.byref(t), C.byref(x), C.byref(y)), "hgrp2d_bench")
        return t.value, x.value, y.value

    def bench_unsplit(self, steps, sample_every=0):
        """(ms of all steps, average ms of one unsplit step launch)."""
        t, s = _D(), _D()
        _check(lib().hgrp2d_bench_unsplit(self._h, steps, sample_every, C.byref(t), C.byref(s)),
               "hgrp2d_bench_unsplit")
        return t.value, s.value

    def set_sweeps(self, x_variant, y_variant):
        _check(lib().hgrp2d_set_sweeps(self._h, x_variant, y_variant), "hgrp2d_set_sweeps")

    def tune(self, steps=4):
        """Times the sweep launch variants on the current state (which it
        restores); returns (x_variant, y_variant, ms per X launch, ms per Y launch).
        On an unsplit context: (form, -1, ms per step, nan)."""
        xv, yv, ms = C.c_int(), C.c_int(), (C.c_double * 2)()
        _check(lib().hgrp2d_tune(self._h, steps, C.byref(xv), C.byref(yv), ms), "hgrp2d_tune")
        return xv.value, yv.value, ms[0], ms[1]


def run_split2d(m, n, init, config, time_plot=None, N_T=None, newton_tol=0.0, guess_params=None,
                converge=False, sweeps=None):
    """Whole run like the oracle's split2d(): returns (frames, N_plot, time_plot, config, status).
    newton_tol / guess_params / converge: see Solver2D.set_newton (defaults: the
    reference's solve); sweeps: (x_variant, y_variant) for Solver2D.set_sweeps."""
    config = np.array(config, dtype=np.float64)
    config[33] = 1.0
    return _run2d(m, n, init, config, time_plot, N_T, newton_tol, guess_params, converge, sweeps)


def run_unsplit2d(m, n, init, config, time_plot=None, N_T=None, newton_tol=0.0, guess_params=None,
                  converge=False, g2d=False, form=None):
    """The unsplit scheme (GRP_solver_2D_EUL_source, config[33] = 0), like the
    oracle's unsplit2d(); results as run_split2d.  g2d: GRP_2D_flux with
    linear_GRP_solver_Edir_G2D (the reference's alternative solver); form: the
    launch form (UNSPLIT_TILE, UNSPLIT_STRIP; None: the default)."""
    config = np.array(config, dtype=np.float64)
    config[33] = 0.0
    return _run2d(m, n, init, config, time_plot, N_T, newton_tol, guess_params, converge, g2d=g2d, form=form)


def _run2d(m, n, init, config, time_plot, N_T, newton_tol, guess_params, converge=False, sweeps=None, g2d=False,
           form=None):
    if time_plot is None:
        time_plot = np.array([0.0, config[1]])
    time_plot = np.array(time_plot, dtype=np.float64)
    if N_T is None:
        N_T = min(5, len(time_plot))
    s = Solver2D(m, n, config)
    try:
        if newton_tol or guess_params is not None or converge:
            s.set_newton(newton_tol, guess_params, converge)
        if sweeps is not None:
            s.set_sweeps(*sweeps)
        if g2d:
            s.set_interface_solver(True)
        if form is not None:
            s.set_unsplit_form(form)
        s.upload(init["rho"], init["u"], init["v"], init["p"], init.get("e"))
        frames = np.zeros((5, N_T, m, n))
        frames[:, 0] = s.download()

        def on_frame(nt, ntp, tc):
            frames[:, nt] = s.download()
            return 0

        res = s.solve(time_plot, N_T, on_frame)
        frames[:, res.nt] = s.download()
    finally:
        s.close()
    config[5] = res.k
    time_plot[res.nt] = res.time_c
    return frames, res.nt + 1, time_plot, config, res


class Batch2D:
    """Independent 2D split problems of uneven sizes advanced together
    (hgrp2db_* entry points).  grids: list of (m, n, h_x, h_y)."""

    def __init__(self, grids, config):
        self.m = np.ascontiguousarray([g[0] for g in grids], dtype=np.int32)
        self.n = np.ascontiguousarray([g[1] for g in grids], dtype=np.int32)
        hx = np.ascontiguousarray([g[2] for g in grids], dtype=np.float64)
        hy = np.ascontiguousarray([g[3] for g in grids], dtype=np.float64)
        self.config = np.ascontiguousarray(config, dtype=np.float64)
        assert self.config.size == N_CONF
        self.offsets = np.concatenate([[0], np.cumsum(self.m.astype(np.int64) * self.n)])
        self._h = C.c_void_p()
        _check(lib().hgrp2db_create(C.byref(self._h), len(self.m), self.m.ctypes.data_as(_IP),
                                    self.n.ctypes.data_as(_IP), _p(hx), _p(hy), _p(self.config)), "hgrp2db_create")

    def close(self):
        if self._h:
            lib().hgrp2db_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, states):
        """states: one dict of rho, u, v, p (m x n arrays) per grid."""
        cat = [np.ascontiguousarray(np.concatenate([np.asarray(s[f], dtype=np.float64).ravel() for s in states]))
               for f in ("rho", "u", "v", "p")]
        assert cat[0].size == self.offsets[-1]
        _check(lib().hgrp2db_upload(self._h, *(_p(a) for a in cat)), "hgrp2db_upload")

    def set_newton(self, tol, params=None, converge=False):
        rule = NEWTON_CONVERGE if converge or params is not None else NEWTON_REFERENCE
        fp = None
        if params is not None:
            pa = np.ascontiguousarray(params, dtype=np.float32)
            self._params = pa
            fp = pa.ctypes.data_as(C.POINTER(C.c_float))
        _check(lib().hgrp2db_set_newton(self._h, tol, rule, fp), "hgrp2db_set_newton")

    def set_sweeps(self, x_variant, y_variant):
        _check(lib().hgrp2db_set_sweeps(self._h, x_variant, y_variant), "hgrp2db_set_sweeps")

    def tune(self, steps=4):
        xv, yv, ms = C.c_int(), C.c_int(), np.zeros(2)
        _check(lib().hgrp2db_tune(self._h, steps, C.byref(xv), C.byref(yv), _p(ms)), "hgrp2db_tune")
        return xv.value, yv.value, ms[0], ms[1]

    def run(self):
        ms, cu = _D(), C.c_longlong()
        _check(lib().hgrp2db_run(self._h, C.byref(ms), C.byref(cu)), "hgrp2db_run")
        return ms.value, cu.value

    def results(self):
        g = len(self.m)
        k, st, steps, t = (np.zeros(g, np.int32), np.zeros(g, np.int32), np.zeros(g, np.int32), np.zeros(g))
        _check(lib().hgrp2db_results(self._h, k.ctypes.data_as(_IP), st.ctypes.data_as(_IP), _p(t),
                                     steps.ctypes.data_as(_IP)), "hgrp2db_results")
        return k, st, t, steps

    def download(self):
        """The final states: a list of (5, m, n) arrays (rho u v p e)."""
        tot = int(self.offsets[-1])
        out = np.zeros((5, tot))
        _check(lib().hgrp2db_download(self._h, *(_p(out[f]) for f in range(5))), "hgrp2db_download")
        return [out[:, a:b].reshape(5, int(m), int(n))
                for a, b, m, n in zip(self.offsets[:-1], self.offsets[1:], self.m, self.n)]


class Ensemble1D:
    """Independent 1D GRP problems advanced together (hgrp1d_* entry points)."""

    def __init__(self, cells, h, config):
        self.cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.h = np.ascontiguousarray(h, dtype=np.float64)
        self.config = np.ascontiguousarray(config, dtype=np.float64)
        assert self.config.size == N_CONF and len(self.h) == len(self.cells)
        self.offsets = np.concatenate([[0], np.cumsum(self.cells)])
        self._h = C.c_void_p()
        _check(lib().hgrp1d_create(C.byref(self._h), len(self.cells), self.cells.ctypes.data_as(_IP),
                                   _p(self.h), _p(self.config)), "hgrp1d_create")

    def close(self):
        if self._h:
            lib().hgrp1d_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, rho, u, p, e=None):
        arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, p)]
        e = None if e is None else np.ascontiguousarray(e, dtype=np.float64)
        _check(lib().hgrp1d_upload(self._h, *(_p(a) for a in arr), _p(e)), "hgrp1d_upload")

    def download(self):
        n = int(self.offsets[-1])
        out = np.zeros((4, n))
        _check(lib().hgrp1d_download(self._h, *(_p(out[q]) for q in range(4))), "hgrp1d_download")
        return out

    def run(self, graph=True):
        ms, cu = _D(), C.c_longlong()
        _check(lib().hgrp1d_run(self._h, int(graph), C.byref(ms), C.byref(cu)), "hgrp1d_run")
        return ms.value, cu.value

    def run_until(self, t_pause, graph=True):
        """Advances every unfinished instance until it finishes or, at the start
        of a step, its time has reached t_pause[i]; returns the paused flags."""
        t = np.ascontiguousarray(np.broadcast_to(t_pause, (len(self.cells),)), dtype=np.float64)
        paused = np.zeros(len(self.cells), np.int32)
        _check(lib().hgrp1d_run_until(self._h, int(graph), _p(t), paused.ctypes.data_as(_IP)), "hgrp1d_run_until")
        return paused.astype(bool)

    def results(self):
        n = len(self.cells)
        steps, status, t = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n)
        _check(lib().hgrp1d_results(self._h, steps.ctypes.data_as(_IP), status.ctypes.data_as(_IP), _p(t)),
               "hgrp1d_results")
        return steps, status, t

    def tune(self, probe_steps=4):
        a, b = C.c_int(), C.c_int()
        _check(lib().hgrp1d_tune(self._h, probe_steps, C.byref(a), C.byref(b)), "hgrp1d_tune")
        return a.value, b.value

    def set_blocks(self, bs_iface, bs_update):
        """Block sizes of the interface and update kernels (what tune() picks)."""
        _check(lib().hgrp1d_set_blocks(self._h, bs_iface, bs_update), "hgrp1d_set_blocks")


def run_grp1d(m, init, config):
    """One 1D problem through the batched engine, shaped like the oracle's
    grp1d(): (final rho u p e [4, m], N_plot, time_plot, config, status)."""
    config = np.array(config, dtype=np.float64)
    ens = Ensemble1D([m], [config[10]], config)
    try:
        ens.upload(init["rho"], init["u"], init["p"], init.get("e"))
        ens.run()
        fin = ens.download()
        steps, status, t = ens.results()
    finally:
        ens.close()
    config[5] = steps[0]
    return fin, 2, np.array([0.0, t[0]]), config, int(status[0])


def flux_batch(L, R, tau, eps, direction):
    L = np.ascontiguousarray(L, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64)
    out = np.zeros((len(L), 9))
    _check(lib().hgrp_grp2d_flux_batch(_p(L), _p(R), len(L), tau, eps, direction, _p(out)), "flux batch")
    return out


def flux_tf_batch(L, R, tau, eps, direction, gamma_a, gamma_b):
    """GRP_2D_flux with the two-fluid members (hgrp_grp2d_flux_tf_batch): rows of
    21-double sides; returns rows (code, F x4, interface values x4, z_a at t + tau)."""
    L = np.ascontiguousarray(L, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64)
    assert L.shape == R.shape and L.shape[1] == 21
    out = np.zeros((len(L), 10))
    _check(lib().hgrp_grp2d_flux_tf_batch(_p(L), _p(R), len(L), tau, eps, direction, gamma_a, gamma_b, _p(out)),
           "hgrp_grp2d_flux_tf_batch")
    return out


def riemann_batch(inp, N):
    """Riemann_solver_exact per row of inp (gammaL gammaR u_L u_R p_L p_R c_L c_R
    eps tol); rows of U_star P_star CRW0 CRW1 gap Newton-iterations."""
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    out = np.zeros((len(inp), 6))
    _check(lib().hgrp_riemann_batch(_p(inp), len(inp), N, _p(out)), "riemann batch")
    return out


def guessp_batch(rows):
    """GuessP (the reference's PVRS / two-shock starting pressure) per row of
    DL DR UL UR PL PR CL CR GammaL GammaR eps."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    out = np.zeros(len(rows))
    _check(lib().hgrp_guessp_batch(_p(rows), len(rows), _p(out)), "guessp batch")
    return out


def riemann_pvrs_batch(rows, N=500):
    """The exact solve from GuessP per row of rho_L u_L p_L rho_R u_R p_R gamma_L
    gamma_R eps tol: rows of u* p* CRW0 CRW1 iterations guess."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    out = np.zeros((len(rows), 6))
    _check(lib().hgrp_riemann_pvrs_batch(_p(rows), len(rows), N, _p(out)), "riemann pvrs batch")
    return out


def _sides(L, R, width=15):
    L = np.ascontiguousarray(L, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64)
    if L.ndim != 2 or L.shape[1] != width or R.shape != L.shape:
        raise ValueError(f"interface sides are (count, {width}) arrays of equal shape")
    return L, R


# the single-fluid build's constants for z_a d_z_a t_z_a phi d_phi t_phi
G2D_SINGLE_FLUID = (0.0, -0.0, -0.0, 0.0, -0.0, -0.0)


def g2d_batch(L, R, eps, atc):
    """linear_GRP_solver_Edir_G2D per pair (the genuinely-2D GRP solver, default
    build).  Sides: (count, 21) = the 15 of q1d_batch, then z_a d_z_a t_z_a phi
    d_phi t_phi; (count, 15) sides get the single-fluid constants.  Rows of
    wave_speed[2], D[6], U[6] (rho u v p z_a phi), U_star[6]."""
    L = np.asarray(L, dtype=np.float64)
    R = np.asarray(R, dtype=np.float64)
    if L.ndim == 2 and L.shape[1] == 15:
        L = np.concatenate([L, np.tile(G2D_SINGLE_FLUID, (len(L), 1))], axis=1)
    if R.ndim == 2 and R.shape[1] == 15:
        R = np.concatenate([R, np.tile(G2D_SINGLE_FLUID, (len(R), 1))], axis=1)
    L, R = _sides(L, R, 21)
    out = np.zeros((len(L), 20))
    _check(lib().hgrp_g2d_batch(_p(L), _p(R), len(L), eps, atc, _p(out)), "g2d batch")
    return out


def q1d_batch(L, R, eps, atc):
    """linear_GRP_solver_Edir_Q1D per pair: rows of wave_speed[2], D[4], U[4], U_star[6]."""
    L, R = _sides(L, R)
    out = np.zeros((len(L), 16))
    _check(lib().hgrp_q1d_batch(_p(L), _p(R), len(L), eps, atc, _p(out)), "q1d batch")
    return out


def flux_general_batch(L, R, tau, eps, n_x, n_y):
    """GRP_2D_flux per pair at the normal (n_x, n_y): rows of code, F_rho F_u F_v F_e,
    RHO_int U_int V_int P_int."""
    L, R = _sides(L, R)
    out = np.zeros((len(L), 9))
    _check(lib().hgrp_grp2d_flux_general_batch(_p(L), _p(R), len(L), tau, eps, n_x, n_y, _p(out)),
           "general flux batch")
    return out


def edir_batch(L, R, eps, atc):
    """linear_GRP_solver_Edir per pair: rows of D[3], U[3] (rho u p)."""
    L, R = _sides(L, R)
    out = np.zeros((len(L), 6))
    _check(lib().hgrp_edir_batch(_p(L), _p(R), len(L), eps, atc, _p(out)), "edir batch")
    return out


def pow_batch(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(len(x))
    _check(lib().hgrp_pow_batch(_p(x), _p(y), len(x), _p(out)), "pow batch")
    return out
