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


def test_transversal_flux_against_reference_and_devpow_oracle(dev, golden_iface, oracle_devpow):
    """The unsplit path's GRP_2D_flux (tangential derivatives on a third of the
    golden cases, both normals) against the reference's own outputs and, bit for
    bit, against the oracle with the device's pow."""
    L, R, tau = (np.ascontiguousarray(golden_iface["L"]), np.ascontiguousarray(golden_iface["R"]),
                 float(golden_iface["tau"]))
    assert np.mean(np.any(L[:, 11:15] != 0, axis=1)) > 0.25
    for direction, nxy, key in ((0, (1.0, 0.0), "flux_x"), (1, (0.0, 1.0), "flux_y")):
        out = np.zeros((len(L), 9))
        dev.hg_flux_t_many(_dp(L), _dp(R), len(L), C.c_double(tau), C.c_double(EPS), direction, _dp(out))
        ref = golden_iface[key]
        ok = ref[:, 0] == 0
        assert np.array_equal(out[:, 0] != 0, ref[:, 0] != 0)
        d = np.abs(out[ok, 1:] - ref[ok, 1:]) / np.maximum(np.abs(ref[ok, 1:]), 1.0)
        print(f"{key}: {np.mean(np.all(out[ok] == ref[ok], axis=1)) * 100:.2f}% rows bit-identical, max rel {d.max():.2e}")
        assert d.max() <= 1e-8
        want = np.array([[rc, *o] for rc, o in (oracle_devpow.flux2d(L[k], R[k], tau, EPS, *nxy)
                                                 for k in range(len(L)))])
        assert np.array_equal(out[ok], want[ok]), direction


def test_reciprocal_division_is_ieee_division(dev):
    """div_by (a / b from a precomputed reciprocal, Markstein's correction)
    against IEEE division, bit for bit, signs of zero included."""
    rng = np.random.default_rng(12)
    n = 2_000_000
    mant = rng.uniform(1.0, 2.0, n) * rng.choice([-1.0, 1.0], n)
    a = np.ldexp(mant, rng.integers(-1070, 1020, n))
    a[:1000] = np.ldexp(rng.uniform(1, 2, 1000), rng.integers(-60, 60, 1000))
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 2.2250738585072014e-308,
                        1.7976931348623157e308, -1.7976931348623157e308, 1.0, -1.0])
    a = np.ascontiguousarray(np.concatenate([a, special]))
    for b in (1 / 1024, 1 / 4096, 1e-3, 1 / 12, 1 / 10, 0.7, 3.0, 1e-5, 1 / 3, 2.0 ** -40, -0.37, 1e-200, 0.0):
        out = np.zeros_like(a)
        dev.hg_div_by_many(_dp(a), len(a), C.c_double(b), _dp(out))
        with np.errstate(divide="ignore", invalid="ignore"):
            want = a / b
        same = (out.view(np.int64) == want.view(np.int64)) | (np.isnan(out) & np.isnan(want))
        assert same.all(), (b, a[~same][:5], out[~same][:5], want[~same][:5])


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)


def test_interface_units_against_reference_and_devpow_oracle(dev, golden_units, oracle_devpow):
    """The per-interface entry points' forms (Riemann gap, Q1D with atc / grid
    velocity / two gases, GRP_2D_flux at any normal, Edir with atc = INFINITY):
    close to the reference, and bit for bit the oracle with the device's pow."""
    g = golden_units
    L, R, tau = np.ascontiguousarray(g["L"]), np.ascontiguousarray(g["R"]), float(g["tau"])
    n = len(L)
    rin = np.ascontiguousarray(g["riemann_in"])
    out = np.zeros((n, 6))
    dev.hg_riemann_many(_dp(rin), n, 500, _dp(out))
    same = np.array([oracle_devpow.riemann(*rin[k], 500) for k in range(n)])
    assert np.array_equal(out, same)
    assert np.array_equal(out[:, 2:4], g["riemann"][:, 2:4]) and _rel(out[:, :2], g["riemann"][:, :2]).max() <= 1e-9
    for a, atc in enumerate(g["atc"]):
        out = np.zeros((n, 16))
        dev.hg_q1d_many(_dp(L), _dp(R), n, C.c_double(EPS), C.c_double(atc), _dp(out))
        same = np.array([oracle_devpow.q1d(L[k], R[k], EPS, atc) for k in range(n)])
        assert np.array_equal(out, same), atc
        d = _rel(out, g["q1d"][a])
        print(f"Q1D atc={atc}: {np.mean(np.all(out == g['q1d'][a], axis=1)) * 100:.2f}% rows bit-identical "
              f"to the reference, max rel {d.max():.2e}")
        assert d.max() <= 1e-9
    for a, (nx, ny) in enumerate(g["normals"]):
        out = np.zeros((n, 9))
        dev.hg_flux_g_many(_dp(L), _dp(R), n, C.c_double(tau), C.c_double(EPS), C.c_double(nx), C.c_double(ny),
                           _dp(out))
        want = np.array([[rc, *o] for rc, o in (oracle_devpow.flux2d(L[k], R[k], tau, EPS, nx, ny)
                                                 for k in range(n))])
        assert np.array_equal(out, want), (nx, ny)
        assert _rel(out, g["flux"][a]).max() <= 1e-9
    out = np.zeros((n, 6))
    dev.hg_edir_atc_many(_dp(L), _dp(R), n, C.c_double(EPS), C.c_double(np.inf), _dp(out))
    same = np.array([oracle_devpow.edir(L[k], R[k], EPS, np.inf) for k in range(n)])
    assert np.array_equal(out, same)
    assert _rel(out, g["edir_inf"]).max() <= 1e-12


def test_guessp_against_reference(dev, golden_units):
    """The device GuessP (no pow: sqrt and divisions only) equals the reference's bit for bit."""
    gi = np.ascontiguousarray(golden_units["guessp_in"])
    out = np.zeros(len(gi))
    dev.hg_guessp_many(_dp(gi), len(gi), _dp(out))
    assert np.array_equal(out, golden_units["guessp"])


def test_riemann_starting_guess_paths(dev, golden_units):
    """The exact solve from a supplied starting pressure: from p* itself it
    stops at once with p*; from a guess 1e8 times too high it restarts after
    three Newton steps and then ends bit for bit where the reference's path
    ends (three more iterations counted)."""
    rin = np.ascontiguousarray(golden_units["riemann_in"])
    n = len(rin)
    plain = np.zeros((n, 6))
    dev.hg_riemann_many(_dp(rin), n, 500, _dp(plain))
    solved = plain[:, 5] > 0
    out = np.zeros((n, 6))
    g = np.ascontiguousarray(np.where(solved, plain[:, 1], -1.0))
    dev.hg_riemann_guess_many(_dp(rin), n, 500, _dp(g), _dp(out))
    assert np.array_equal(out[solved, 1], plain[solved, 1]) and np.all(out[solved, 5] == 0)
    assert np.array_equal(out[~solved], plain[~solved])
    g = np.ascontiguousarray(np.where(solved, 1e8 * plain[:, 1], -1.0))
    dev.hg_riemann_guess_many(_dp(rin), n, 500, _dp(g), _dp(out))
    restarted = solved & (out[:, 5] == plain[:, 5] + 3)
    assert restarted.sum() >= 0.9 * solved.sum()
    assert np.array_equal(out[restarted, :5], plain[restarted, :5])


def test_converge_rule_against_devpow_oracle_and_guess_independence(dev, oracle_devpow):
    """The device's converge rule (riemann_exact<true>) built for the host equals
    the device-pow oracle's bit for bit from the reference's guess; started from
    guesses off the root by ~3 % (the trained network's typical error), it ends
    within 1e-12 of that answer, which the reference's rule does not."""
    import cases
    states = cases.c3_states(3000, seed=43, near_vacuum=300)
    g = 1.4
    cl = np.sqrt(g * states[:, 2] / states[:, 0])
    cr = np.sqrt(g * states[:, 5] / states[:, 3])
    rin = np.ascontiguousarray(np.stack([np.full(len(states), g), np.full(len(states), g), states[:, 1], states[:, 4],
                                         states[:, 2], states[:, 5], cl, cr, np.full(len(states), EPS),
                                         np.full(len(states), 1e-12)], axis=1))
    n = len(rin)
    none = np.full(n, -1.0)
    out = {}
    for rule in (0, 1):
        o = np.zeros((n, 6))
        dev.hg_riemann_rule_many(_dp(rin), n, 500, rule, _dp(none), _dp(o))
        out[rule] = o
    want = np.array([oracle_devpow.riemann_rule(*r, 500, 1) for r in rin])
    assert np.array_equal(out[1][:, [0, 1, 2, 3, 5]], want[:, [0, 1, 2, 3, 5]])
    # guesses off by lognormal(0.03) from the converged root
    rng = np.random.default_rng(44)
    guess = np.ascontiguousarray(out[1][:, 1] * np.exp(rng.normal(0.0, 0.03, n)))
    for rule in (0, 1):
        o = np.zeros((n, 6))
        dev.hg_riemann_rule_many(_dp(rin), n, 500, rule, _dp(guess), _dp(o))
        d = np.abs(o[:, 1] - out[rule][:, 1]) / out[rule][:, 1]
        print(f"rule {rule}: p* from the perturbed guesses vs the reference's guess: max rel {d.max():.1e}, "
              f"{np.sum(d > 1e-12)} of {n} over 1e-12; iterations {out[rule][:, 5].mean():.3f} -> {o[:, 5].mean():.3f}")
        if rule == 1:
            assert d.max() <= 1e-12
        else:
            assert np.sum(d > 1e-12) > 0


def test_two_fluid_flux_against_multifluid_reference(dev):
    """The two-fluid interface solve of the split sweeps (grp_2d_flux_tf:
    Q1D with a gas per side, the volume fraction carried with the contact, the
    energy flux with gamma from z_a at t + tau/2) against the reference's own
    GRP_2D_flux built with -DMULTIFLUID_BASICS (tests/golden/twofluid_vectors.npz):
    fluxes, interface values and z_a at t + tau."""
    from conftest import load_golden
    g = load_golden("twofluid_vectors.npz")
    L, R = np.ascontiguousarray(g["L"]), np.ascontiguousarray(g["R"])
    n, tau = len(L), float(g["tau"])
    for direction in (0, 1):
        Ld, Rd = L.copy(), R.copy()
        Ld[:, 11:15] = Rd[:, 11:15] = 0.0 if direction == 0 else -0.0
        out = np.zeros((n, 10))
        dev.hg_flux_tf_many(_dp(Ld), _dp(Rd), n, C.c_double(tau), C.c_double(EPS), direction,
                            C.c_double(float(g["gamma_a"])), C.c_double(float(g["gamma_b"])), _dp(out))
        ref = g[f"flux_{direction}"]
        assert np.array_equal(out[:, 0], ref[:, 0])
        want = ref[:, [1, 2, 3, 4, 5, 6, 7, 8, 9]]
        got = out[:, 1:10]
        d = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
        mixed = (L[:, 15] > 0) & (L[:, 15] < 1) | (R[:, 15] > 0) & (R[:, 15] < 1)
        print(f"two-fluid flux dir {direction}: {np.mean(np.all(got == want, axis=1)) * 100:.1f}% rows bit-identical, "
              f"max rel {d.max():.2e} (mixed sides: {d[mixed].max():.2e})", flush=True)
        assert d.max() <= 1e-12


def test_flux_check_codes_on_invalid_states(dev, golden_units):
    """GRP_2D_flux on the reference's struct cases, among them negative and
    non-finite states: the device code built for the host returns the
    reference's check code on every case (a NaN velocity gap ends the Newton
    loop, as the reference's `while (gap > tol && n != N)` does), and the
    reference's results where the code is 0."""
    from _libs import IFV_MEMBERS, SIDE15_MEMBERS
    g = golden_units
    idx = [IFV_MEMBERS.index(name) for name in SIDE15_MEMBERS]
    L = np.ascontiguousarray(g["struct_in_L"][:, idx])
    R = np.ascontiguousarray(g["struct_in_R"][:, idx])
    nx, ny = g["struct_in_L"][:, 0], g["struct_in_L"][:, 1]
    tau = float(g["tau"])
    res = ("F_rho", "F_u", "F_v", "F_e", "RHO_int", "U_int", "V_int", "P_int")
    want_cols = [IFV_MEMBERS.index(name) for name in res]
    bad = 0
    for k in range(len(L)):
        out = np.zeros(9)
        dev.hg_flux_g_many(_dp(L[k:k + 1].copy()), _dp(R[k:k + 1].copy()), 1, C.c_double(tau), C.c_double(EPS),
                           C.c_double(nx[k]), C.c_double(ny[k]), _dp(out))
        assert out[0] == g["struct_rc"][k], k
        bad += out[0] != 0
        if out[0] == 0:
            want = g["struct_out_L"][k, want_cols]
            assert np.max(np.abs(out[1:] - want) / np.maximum(np.abs(want), 1.0)) <= 1e-10, k
    assert bad > 0


def test_g2d_against_reference_and_devpow_oracle(dev, golden_g2d, oracle_devpow):
    """grp_g2d (csrc/g2d_device.hpp) built for the host: bit for bit the oracle
    with the device's pow, and the reference's single-fluid and multi-fluid
    builds to the last bits of glibc's pow."""
    g = golden_g2d
    n = len(g["L"])
    single = np.array([0.0, -0.0, -0.0, 0.0, -0.0, -0.0])
    for kind in ("single", "multi"):
        L, R = g["L"].copy(), g["R"].copy()
        if kind == "single":
            L[:, 15:] = single
            R[:, 15:] = single
        L, R = np.ascontiguousarray(L), np.ascontiguousarray(R)
        for a, atc in enumerate(g["atc"]):
            out = np.zeros((n, 20))
            dev.hg_g2d_many(_dp(L), _dp(R), n, C.c_double(EPS), C.c_double(atc), _dp(out))
            w = 21 if kind == "multi" else 15
            same = np.array([oracle_devpow.g2d(L[k, :w], R[k, :w], EPS, atc) for k in range(n)])
            assert np.array_equal(out, same, equal_nan=True), (kind, atc)
            ref = g[kind][a]
            ok = np.isfinite(ref).all(axis=1)
            assert np.array_equal(ok, np.isfinite(out).all(axis=1))
            d = np.abs(out[ok] - ref[ok]) / np.maximum(np.abs(ref[ok]), 1e-6 * np.abs(ref[ok]).max(axis=0) + 1e-300)
            print(f"G2D {kind} atc={atc}: {np.mean(np.all(out[ok] == ref[ok], axis=1)) * 100:.2f}% rows "
                  f"bit-identical to the reference, max rel {d.max():.2e}")
            assert d.max() <= 1e-14
