"""The device GRP solvers (csrc/grp_device.hpp, csrc/grp1d_device.hpp) built for
the host as plain C++ (tests/native/grp_host.cpp with a host stand-in for
<hip/hip_runtime.h>), checked on the CPU against the reference's vectors and
against the oracle built with the same pow: with pow equal, every other
operation must agree to the last bit."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from _libs import ROOT

SRC = os.path.join(ROOT, "tests", "native", "grp_host.cpp")
INC = [os.path.join(ROOT, "tests", "native", "shim"), os.path.join(ROOT, "hydrocode-fusion_amd", "csrc")]
EPS = 1e-9


@pytest.fixture(scope="module")
def dev(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("grphost") / "grp_host.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
                    *[f"-I{d}" for d in INC], SRC, "-o", so], check=True)
    return C.CDLL(so)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_edir_against_reference_and_devpow_oracle(dev, golden_iface, oracle_devpow):
    L, R = np.ascontiguousarray(golden_iface["L"]), np.ascontiguousarray(golden_iface["R"])
    out = np.zeros((len(L), 6))
    dev.hg_edir_many(_dp(L), _dp(R), len(L), C.c_double(EPS), _dp(out))
    ref = golden_iface["edir"]
    d = np.abs(out - ref) / np.maximum(np.abs(ref), 1.0)
    print(f"Edir: {np.mean(np.all(out == ref, axis=1)) * 100:.2f}% rows bit-identical to the reference, "
          f"max rel {d.max():.2e}")
    assert d.max() <= 1e-12
    same = np.array([oracle_devpow.edir(L[k], R[k], EPS, EPS) for k in range(len(L))])
    assert np.array_equal(out, same)


def test_flux_against_reference_and_devpow_oracle(dev, golden_iface, oracle_devpow):
    L, R, tau = golden_iface["L"].copy(), golden_iface["R"].copy(), float(golden_iface["tau"])
    L[:, 11:15] = 0.0
    R[:, 11:15] = 0.0
    for direction, nxy, key in ((0, (1.0, 0.0), "flux_x"), (1, (0.0, 1.0), "flux_y")):
        Ld, Rd = np.ascontiguousarray(L), np.ascontiguousarray(R)
        if direction == 1:
            Ld[:, 11:15] = -0.0
            Rd[:, 11:15] = -0.0
        out = np.zeros((len(L), 9))
        dev.hg_flux_many(_dp(Ld), _dp(Rd), len(L), C.c_double(tau), C.c_double(EPS), direction, _dp(out))
        want = np.array([[rc, *o] for rc, o in (oracle_devpow.flux2d(Ld[k], Rd[k], tau, EPS, *nxy)
                                                 for k in range(len(L)))])
        ok = want[:, 0] == 0
        assert np.array_equal(out[:, 0] != 0, want[:, 0] != 0)
        assert np.array_equal(out[ok], want[ok]), direction
