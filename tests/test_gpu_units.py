"""GPU parity of the per-interface solvers as functions of their own: the
batched entry points (hgrp_riemann_batch, hgrp_q1d_batch, hgrp_g2d_batch,
hgrp_grp2d_flux_general_batch, hgrp_edir_batch) and the drop-ins with the
reference's names and signatures in libhydrocode2d (Riemann_solver_exact,
linear_GRP_solver_Edir_Q1D, linear_GRP_solver_Edir_G2D, linear_GRP_solver_Edir,
GRP_2D_flux), against the reference's own outputs in
tests/golden/interface_units.npz and g2d_vectors.npz.

Bars: the batched solvers per element relative to the value (floor 1e-6 of
the column's largest magnitude), 1e-13 from the oracle built with the device's
pow and 1e-12 from the reference; the one-call drop-ins 1e-10 (scale
max(|reference|, 1)); the exact Riemann solve 1e-9 (its absolute Newton stop);
wave types, check codes and every struct member the solver only moves
(rotations, the zeroed grid velocity, untouched members) bit for bit.
"""
import ctypes as C
import os

import numpy as np
import pytest

from _libs import IFV_MEMBERS, ROOT, IfvStruct, ifv_members

pytestmark = pytest.mark.gpu
EPS = 1e-9
TOL = 1e-10
HOST_SO = os.path.join(ROOT, "hydrocode-fusion_amd", "lib", "libhydrocode2d.so")


@pytest.fixture(scope="module")
def hg():
    import hydrogrp
    hydrogrp.lib()
    print("device:", hydrogrp.device_name(), flush=True)
    return hydrogrp


@pytest.fixture(scope="module")
def host():
    L = C.CDLL(HOST_SO)
    P = C.POINTER(IfvStruct)
    DP = C.POINTER(C.c_double)
    L.Riemann_solver_exact.restype = C.c_double
    L.Riemann_solver_exact.argtypes = [DP, DP] + [C.c_double] * 8 + [C.POINTER(C.c_bool), C.c_double, C.c_double,
                                                                      C.c_int]
    L.linear_GRP_solver_Edir_Q1D.restype = None
    L.linear_GRP_solver_Edir_Q1D.argtypes = [DP, DP, DP, DP, P, P, C.c_double, C.c_double]
    L.linear_GRP_solver_Edir.restype = None
    L.linear_GRP_solver_Edir.argtypes = [DP, DP, P, P, C.c_double, C.c_double]
    L.linear_GRP_solver_Edir_G2D.restype = None
    L.linear_GRP_solver_Edir_G2D.argtypes = [DP, DP, DP, DP, P, P, C.c_double, C.c_double]
    L.GRP_2D_flux.restype = C.c_int
    L.GRP_2D_flux.argtypes = [P, P, C.c_double]
    L.cfg = (C.c_double * 400).in_dll(L, "config")
    return L


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)


def _rel_col(a, b):
    """|a - b| relative to |b|, floored at 1e-6 of each column's largest magnitude."""
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-6 * np.abs(b).max(axis=0) + 1e-300)


def _struct(vec):
    return IfvStruct(*[float(x) for x in vec])


def _members(s):
    return np.array([getattr(s, name) for name in IFV_MEMBERS])


def test_riemann_batch_with_gap(hg, golden_units):
    g = golden_units
    out = hg.riemann_batch(g["riemann_in"], 500)
    ref = g["riemann"]
    assert np.array_equal(out[:, 2:4], ref[:, 2:4])
    assert _rel(out[:, :2], ref[:, :2]).max() <= 1e-9        # the Newton stop rule is an absolute 1e-9 gap
    same = out[:, 4] == ref[:, 4]
    print(f"riemann gap: {np.mean(same) * 100:.2f}% bit-identical", flush=True)
    assert np.mean(same) >= 0.95
    assert np.array_equal(out[:, 4] <= EPS, ref[:, 4] <= EPS)   # converged exactly where the reference is


@pytest.mark.parametrize("a", [0, 1, 2], ids=["atc=eps", "atc=inf", "atc=-0"])
def test_q1d_batch(hg, golden_units, oracle_devpow, a):
    """Q1D per element relative to the value (floor 1e-6 of the column's
    largest magnitude): 1e-13 from the oracle with the device's pow, 1e-12
    from the reference (the two oracles differ by at most 8.6e-16 here)."""
    g = golden_units
    atc = float(g["atc"][a])
    out = hg.q1d_batch(g["L"], g["R"], EPS, atc)
    ref = g["q1d"][a]
    same = np.array([oracle_devpow.q1d(g["L"][k], g["R"][k], EPS, atc) for k in range(len(out))])
    dd, d = _rel_col(out, same), _rel_col(out, ref)
    print(f"Q1D atc={atc}: {np.mean(np.all(out == same, axis=1)) * 100:.2f}% rows bit-identical to the device-pow "
          f"oracle (max rel {dd.max():.2e}), {np.mean(np.all(out == ref, axis=1)) * 100:.2f}% to the reference "
          f"(max rel {d.max():.2e})", flush=True)
    assert dd.max() <= 1e-13 and d.max() <= 1e-12


@pytest.mark.parametrize("a", [0, 1, 2], ids=["atc=eps", "atc=inf", "atc=-0"])
def test_g2d_batch(hg, golden_g2d, oracle_devpow, a):
    """linear_GRP_solver_Edir_G2D (hgrp_g2d_batch) on the reference's
    single-fluid and multi-fluid vectors: to 1e-13 of the oracle with the
    device's pow and to 1e-12 of the reference (the two oracles differ by at
    most 5.1e-16 here), NaN rows where the reference has them."""
    g = golden_g2d
    atc = float(g["atc"][a])
    for kind in ("single", "multi"):
        L, R = g["L"], g["R"]
        if kind == "single":
            L, R = L[:, :15], R[:, :15]
        out = hg.g2d_batch(L, R, EPS, atc)
        ref = g[kind][a]
        ok = np.isfinite(ref).all(axis=1)
        assert np.array_equal(ok, np.isfinite(out).all(axis=1))
        same = np.array([oracle_devpow.g2d(L[k], R[k], EPS, atc) for k in range(len(L))])
        dd = _rel_col(out[ok], same[ok])
        d = _rel_col(out[ok], ref[ok])
        print(f"G2D {kind} atc={atc}: against the device-pow oracle {np.mean(np.all(dd == 0, axis=1)) * 100:.2f}% "
              f"rows bit-identical, max rel {dd.max():.2e}; against the reference max rel {d.max():.2e}",
              flush=True)
        assert dd.max() <= 1e-13 and d.max() <= 1e-12, (kind, dd.max(), d.max())


def test_dropin_g2d(host, hg, golden_g2d):
    """linear_GRP_solver_Edir_G2D with the reference's signature (single-fluid
    structs, one interface per call) returns the batch entry's results bit
    for bit, all six entries of D and U."""
    g = golden_g2d
    L, R = g["L"][:, :15], g["R"][:, :15]
    for a, atc in enumerate(g["atc"]):
        batch = hg.g2d_batch(L, R, EPS, float(atc))
        for k in range(0, len(L), 20):
            sl, sr = _struct(ifv_members(L[k])), _struct(ifv_members(R[k]))
            ws, D, U, Us = (C.c_double * 2)(), (C.c_double * 6)(), (C.c_double * 6)(), (C.c_double * 6)()
            host.linear_GRP_solver_Edir_G2D(ws, D, U, Us, C.byref(sl), C.byref(sr), EPS, float(atc))
            got = np.array(list(ws) + list(D) + list(U) + list(Us))
            assert np.array_equal(got, batch[k], equal_nan=True), (k, atc)


def test_flux_general_batch(hg, golden_units, oracle_devpow):
    """GRP_2D_flux at four normals, per element as test_q1d_batch: 1e-13 from
    the device-pow oracle, 1e-12 from the reference (the oracles: <=2.3e-14)."""
    g = golden_units
    L, R, tau = g["L"], g["R"], float(g["tau"])
    for a, (nx, ny) in enumerate(g["normals"]):
        out = hg.flux_general_batch(L, R, tau, EPS, nx, ny)
        ref = g["flux"][a]
        assert np.array_equal(out[:, 0], ref[:, 0])
        ok = ref[:, 0] == 0
        same = np.array([[rc, *o] for rc, o in (oracle_devpow.flux2d(L[k], R[k], tau, EPS, nx, ny)
                                                 for k in range(len(L)))])
        dd, d = _rel_col(out[ok, 1:], same[ok, 1:]), _rel_col(out[ok, 1:], ref[ok, 1:])
        print(f"GRP_2D_flux normal ({nx}, {ny}): max rel {dd.max():.2e} from the device-pow oracle, "
              f"{d.max():.2e} from the reference", flush=True)
        assert dd.max() <= 1e-13 and d.max() <= 1e-12


def test_edir_batch_acoustic_only(hg, golden_units, oracle_devpow):
    g = golden_units
    out = hg.edir_batch(g["L"], g["R"], EPS, np.inf)
    same = np.array([oracle_devpow.edir(g["L"][k], g["R"][k], EPS, np.inf) for k in range(len(out))])
    assert _rel_col(out, same).max() <= 1e-13 and _rel_col(out, g["edir_inf"]).max() <= 1e-12


def test_dropin_riemann_and_solvers(host, golden_units):
    """The reference-signature functions, one interface per call."""
    g = golden_units
    L, R = g["L"], g["R"]
    for k in range(0, len(L), 25):
        inp = g["riemann_in"][k]
        us, ps = C.c_double(), C.c_double()
        crw = (C.c_bool * 2)()
        gap = host.Riemann_solver_exact(C.byref(us), C.byref(ps), *[float(x) for x in inp[:8]], crw,
                                        float(inp[8]), float(inp[9]), 500)
        ref = g["riemann"][k]
        assert (crw[0], crw[1]) == (bool(ref[2]), bool(ref[3])), k
        assert _rel(np.array([us.value, ps.value]), ref[:2]).max() <= 1e-9, k
        assert (gap <= EPS) == (ref[4] <= EPS), k
        sl, sr = _struct(ifv_members(L[k])), _struct(ifv_members(R[k]))
        for a, atc in enumerate(g["atc"]):
            ws, D, U, Us = (C.c_double * 2)(), (C.c_double * 4)(), (C.c_double * 4)(), (C.c_double * 6)()
            host.linear_GRP_solver_Edir_Q1D(ws, D, U, Us, C.byref(sl), C.byref(sr), EPS, float(atc))
            got = np.array(list(ws) + list(D) + list(U) + list(Us))
            assert _rel(got, g["q1d"][a, k]).max() <= TOL, (k, atc)
        D3, U3 = (C.c_double * 3)(), (C.c_double * 3)()
        host.linear_GRP_solver_Edir(D3, U3, C.byref(sl), C.byref(sr), EPS, np.inf)
        assert _rel(np.array(list(D3) + list(U3)), g["edir_inf"][k]).max() <= TOL, k


def test_dropin_grp_2d_flux_structs(host, golden_units):
    """GRP_2D_flux on whole structs: return code, results, and its side effects
    on both sides (velocities and their derivatives rotated into the normal
    frame, the left grid velocity zeroed, everything else left as it was)."""
    g = golden_units
    host.cfg[4] = EPS
    results = [IFV_MEMBERS.index(n) for n in ("RHO_int", "P_int", "U_int", "V_int", "F_rho", "F_e", "F_u", "F_v")]
    moved = [k for k in range(34) if k not in results]
    codes = []
    for k in range(len(g["struct_rc"])):
        sl, sr = _struct(g["struct_in_L"][k]), _struct(g["struct_in_R"][k])
        rc = host.GRP_2D_flux(C.byref(sl), C.byref(sr), float(g["tau"]))
        codes.append(rc)
        assert rc == g["struct_rc"][k], k
        outL, outR = _members(sl), _members(sr)
        assert np.array_equal(outL[moved], g["struct_out_L"][k, moved], equal_nan=True), k
        assert np.array_equal(outR, g["struct_out_R"][k], equal_nan=True), k
        if rc:
            assert np.array_equal(outL[results], g["struct_out_L"][k, results]), k     # untouched on failure
        else:
            assert _rel(outL[results], g["struct_out_L"][k, results]).max() <= TOL, k
    assert set(codes) == {0, 1, 2, 3}


def test_guessp_batch(hg, golden_units):
    """GuessP on the GPU equals the reference's bit for bit."""
    out = hg.guessp_batch(golden_units["guessp_in"])
    assert np.array_equal(out, golden_units["guessp"])


def test_riemann_from_guessp(hg, golden_units):
    """The exact solve started from GuessP ends at the reference's answer (to
    the Newton tolerance) with the same wave types; the Newton iterations it
    takes against the reference's halved-tangent guess are reported."""
    from hydrogrp import mlp
    g = golden_units
    Lv, Rv = g["L"], g["R"]
    rows = np.stack([Lv[:, 3], Lv[:, 4], Lv[:, 6], Rv[:, 3], Rv[:, 4], Rv[:, 6], Lv[:, 0], Rv[:, 0],
                     np.full(len(Lv), EPS), np.full(len(Lv), EPS)], axis=1)
    out = hg.riemann_pvrs_batch(rows, 500)
    ref = g["riemann"]
    assert np.array_equal(out[:, 2:4], ref[:, 2:4])
    scale = np.maximum(np.abs(ref[:, :2]), 1.0)
    assert (np.abs(out[:, :2] - ref[:, :2]) / scale).max() <= 1e-8
    assert np.all(out[:, 5] >= EPS)
    plain = mlp.riemann_guess_batch(rows[:, :8], eps=EPS, tol=EPS)
    nz = plain[:, 4] > 0
    print(f"Newton iterations over {nz.sum()} solved interfaces: halved tangents {plain[nz, 4].mean():.2f}, "
          f"GuessP {out[nz, 4].mean():.2f}", flush=True)


def test_copy_probe(hg):
    """The calibration copy: sizes that are not a multiple of the grid's
    stride run, the argument checks reject a zero size, and the rate is that
    of a streaming kernel (well above 1 TB/s for 256 MiB read + written)."""
    for nbytes in (16, 4096 + 48, (1 << 20) + 16 * 7):
        assert hg.copy_probe(nbytes, 1) > 0.0
    with pytest.raises(RuntimeError):
        hg.copy_probe(0, 1)
    nbytes = 256 << 20
    ms = hg.copy_probe(nbytes, 5)
    gbs = 2.0 * nbytes / (ms * 1e-3) / 1e9
    print(f"copy probe: {ms:.4f} ms per 256 MiB copy, {gbs:.0f} GB/s", flush=True)
    assert gbs > 1000.0
