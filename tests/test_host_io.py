"""The C99 host side (lib/libhydrocode2d.so): data_in reading, config
precedence and data_out writers, against the reference's own I/O modules
(oracle/_ref/libhydroref_io.so) byte for byte.  CPU only: no GPU call here."""
import ctypes as C
import os

import numpy as np
import pytest

import cases
from _libs import CVFrames, FluVar, ROOT

HOST_SO = os.path.join(ROOT, "hydrocode-fusion_amd", "lib", "libhydrocode2d.so")
REF_IO_SO = os.path.join(ROOT, "oracle", "_ref", "libhydroref_io.so")


def _load(path):
    L = C.CDLL(path)
    L.arg_preprocess.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_char_p), C.c_char_p]
    L.initialize_2D.restype = FluVar
    L.initialize_2D.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.POINTER(C.c_double))]
    L.file_2D_write.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.POINTER(C.c_double), C.c_char_p, C.POINTER(C.c_double)]
    L.file_2D_write_POINT_TEC.argtypes = L.file_2D_write.argtypes
    L.cfg = (C.c_double * 400).in_dll(L, "config")
    return L


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOST_SO):
        pytest.fail(f"{HOST_SO} missing: run __graft_entry__.build()")
    return _load(HOST_SO)


@pytest.fixture(scope="module")
def refio():
    if not os.path.exists(REF_IO_SO):
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    return _load(REF_IO_SO)


def _tree(tmp, name):
    run = tmp / name / "run" / "here"
    run.mkdir(parents=True)
    return tmp / name, run


def _write_case(base, m, n):
    import hydrogrp.datain as datain
    init = cases.smooth_random(m, n, 17, amp=0.3)
    init["u"][2, 3] = -1.25e-7
    init["p"][4, 1] = 3.0e-300
    folder = base / "data_in" / "two-dim" / "smallcase"
    datain.write_case_2d(str(folder), init, [(1, 0.05), (4, 1e-9), (6, 1.4), (10, 1.0 / m), (11, 1.0 / n),
                                              (17, -7), (18, -2)], time_plot=[0.03, 0.01])
    with open(folder / "config.txt", "a") as f:
        f.write("# comment line\n\n7 0.3\n")
    return init


def _read(lib, run, argv):
    cwd = os.getcwd()
    os.chdir(run)
    try:
        for k in range(1, 400):
            lib.cfg[k] = float("inf")
        arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
        lib.arg_preprocess(4, len(argv), arr, None)
        lib.cfg[0] = 2.0
        N, Np = C.c_int(), C.c_int()
        tp = C.POINTER(C.c_double)()
        fv = lib.initialize_2D(argv[1].encode(), C.byref(N), C.byref(Np), C.byref(tp))
        nx, ny = int(lib.cfg[13]), int(lib.cfg[14])
        fields = np.array([np.ctypeslib.as_array(getattr(fv, k), (nx * ny,)).copy() for k in ("RHO", "U", "V", "P")])
        return fields, np.array(lib.cfg[:]), N.value, Np.value, np.ctypeslib.as_array(tp, (Np.value,)).copy()
    finally:
        os.chdir(cwd)


def test_datain_roundtrip(host, tmp_path):
    base, run = _tree(tmp_path, "ours")
    init = _write_case(base, 9, 7)
    fields, cfg, N, Np, tp = _read(host, run, ["prog", "smallcase", "out", "2_GRP", "EUL", "33=1", "7=0.4"])
    for f, k in enumerate(("rho", "u", "v", "p")):
        assert np.array_equal(fields[f], init[k].T.reshape(-1))    # file line = y index
    assert (cfg[13], cfg[14], cfg[3], cfg[9], cfg[33]) == (9, 7, 63, 2, 1)
    assert cfg[7] == 0.4                                          # CLI beats config.txt
    assert (N, Np) == (4, 4) and np.array_equal(tp, [0.0, 0.01, 0.03, 0.05])
    assert cfg[41] == 1.9 and cfg[5] == 2**31 - 1                 # defaults


def test_reader_matches_reference(host, refio, tmp_path):
    argv = ["prog", "smallcase", "out", "2_GRP", "EUL", "33=1", "7=0.4", "41=1.5"]
    got = []
    for name, lib in (("ours", host), ("ref", refio)):
        base, run = _tree(tmp_path, name)
        _write_case(base, 9, 7)
        got.append(_read(lib, run, argv))
    (f0, c0, N0, P0, t0), (f1, c1, N1, P1, t1) = got
    assert np.array_equal(f0, f1) and (N0, P0) == (N1, P1) and np.array_equal(t0, t1)
    assert np.array_equal(c0, c1)


def _write_outputs(lib, run, frames, time_plot, cfg_items):
    cwd = os.getcwd()
    os.chdir(run)
    try:
        for k in range(400):
            lib.cfg[k] = float("inf")
        for k, v in cfg_items:
            lib.cfg[k] = v
        N = frames.N_T
        cpu = (C.c_double * N)()
        tp = (C.c_double * N)(*time_plot)
        args = (frames.m, frames.n, N, C.cast(frames.CV, C.c_void_p), C.cast(frames.X, C.c_void_p),
                C.cast(frames.Y, C.c_void_p), cpu, b"outcase", tp)
        lib.file_2D_write(*args)
        lib.file_2D_write_POINT_TEC(*args)
    finally:
        os.chdir(cwd)


def test_writers_match_reference_byte_for_byte(host, refio, tmp_path):
    m, n, N = 6, 5, 3
    rng = np.random.default_rng(4)
    outs = []
    for name, lib in (("ours", host), ("ref", refio)):
        fr = CVFrames(m, n, N, h_x=0.1, h_y=0.2)
        fr.data[:] = rng.normal(size=fr.data.shape) * 10.0 ** rng.integers(-12, 12, size=fr.data.shape)
        rng = np.random.default_rng(4)    # same values for both
        fr.data[0, 0, 1, 1] = -0.0
        fr.data[1, 1, 2, 2] = 1e-300
        base, run = _tree(tmp_path, name)
        _write_outputs(lib, run, fr, [0.0, 0.125, 0.3], [(0, 2), (1, 0.3), (3, m * n), (4, 1e-9), (5, 77), (6, 1.4),
                                                         (7, 0.45), (8, 0), (9, 2), (10, 0.1), (11, 0.2),
                                                         (17, -4), (18, -7)])
        outdir = base / "data_out" / "two-dim" / "EUL_2_order" / "outcase"
        outs.append({p.name: p.read_bytes() for p in sorted(outdir.iterdir())})
    assert set(outs[0]) == set(outs[1]) and len(outs[0]) == 10, sorted(outs[0])
    for name in outs[0]:
        assert outs[0][name] == outs[1][name], name


def test_full_precision_frames_roundtrip(host, tmp_path):
    """FLU_VAR.f64 (the full-precision frames written beside the reference's
    outputs): every double comes back bit for bit, headers are checked."""
    rng = np.random.default_rng(4)
    m, n, N = 7, 5, 3
    cv = CVFrames(m, n, N)
    cv.data[:] = rng.normal(size=cv.data.shape) * 10.0 ** rng.integers(-300, 300, cv.data.shape)
    cv.data[0, 0, 0, 0] = -0.0
    tp = np.ascontiguousarray([0.0, 0.1, 0.2 + 1e-17])
    DP = C.POINTER(C.c_double)
    host.file_2D_write_f64.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p, DP]
    host.file_2D_read_f64.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(DP), C.POINTER(DP)]
    f = str(tmp_path / "FLU_VAR.f64").encode()
    assert host.file_2D_write_f64(f, m, n, N, cv.CV, tp.ctypes.data_as(DP)) == 0
    assert os.path.getsize(f) == 8 + 16 + 8 * N + 8 * N * 5 * m * n
    rx, ry, rn = C.c_int(), C.c_int(), C.c_int()
    tpp, frp = DP(), DP()
    assert host.file_2D_read_f64(f, C.byref(rx), C.byref(ry), C.byref(rn), C.byref(tpp), C.byref(frp)) == 0
    assert (rx.value, ry.value, rn.value) == (m, n, N)
    back = np.ctypeslib.as_array(frp, (N * 5 * m * n,)).reshape(N, 5, m, n)
    assert np.array_equal(np.ctypeslib.as_array(tpp, (N,)).view(np.int64), tp.view(np.int64))
    assert np.array_equal(back.transpose(1, 0, 2, 3).view(np.int64), cv.data.view(np.int64))
    bad = tmp_path / "bad.f64"
    bad.write_bytes(b"NOTAFILE" + bytes(64))
    assert host.file_2D_read_f64(str(bad).encode(), C.byref(rx), C.byref(ry), C.byref(rn), C.byref(tpp),
                                 C.byref(frp)) == 2
    assert host.file_2D_read_f64(str(tmp_path / "missing.f64").encode(), C.byref(rx), C.byref(ry), C.byref(rn),
                                 C.byref(tpp), C.byref(frp)) == 1
