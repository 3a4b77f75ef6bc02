import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, os.path.join(ROOT, "hydrocode-fusion_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from _libs import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def oracle_devpow():
    """The oracle with pow() taken from the device library (oracle/liboracle_devpow.so)."""
    from _libs import Oracle, ORACLE_DEVPOW_SO
    return Oracle(ORACLE_DEVPOW_SO)


@pytest.fixture(scope="session")
def reflib():
    from _libs import RefLib, REF_SO
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (needs /root/reference): `make -C oracle ref`")
    return RefLib()


def load_golden(name):
    return np.load(os.path.join(HERE, "golden", name), allow_pickle=False)


@pytest.fixture(scope="session")
def golden_iface():
    return load_golden("interface_vectors.npz")


@pytest.fixture(scope="session")
def golden_units():
    return load_golden("interface_units.npz")


@pytest.fixture(scope="session")
def golden_unsplit_g2d():
    return load_golden("unsplit2d_g2d_cases.npz")


@pytest.fixture(scope="session")
def golden_g2d():
    return load_golden("g2d_vectors.npz")


@pytest.fixture(scope="session")
def golden_stages():
    return load_golden("stage_cases.npz")


@pytest.fixture(scope="session")
def golden_2d():
    return load_golden("split2d_cases.npz")


@pytest.fixture(scope="session")
def golden_unsplit():
    return load_golden("unsplit2d_cases.npz")


@pytest.fixture(scope="session")
def golden_twofluid():
    return load_golden("twofluid_vectors.npz")


@pytest.fixture(scope="session")
def golden_1d():
    return load_golden("grp1d_cases.npz")


def rel_linf(a, b):
    """Per-field relative L-infinity difference: max|a - b| / max|b| (fields on axis 0)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    out = []
    for f in range(a.shape[0]):
        scale = np.max(np.abs(b[f]))
        d = np.max(np.abs(a[f] - b[f])) if a[f].size else 0.0
        out.append(d / scale if scale > 0 else d)
    return np.array(out)


def parity(a, b):
    """Relative L-infinity difference per field (fields on axis 0: rho u v p E):
    per cell for the positive fields rho, p, E (max |a - b| / |b|), against the
    field's largest magnitude for the velocities, which pass through zero."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    out = []
    for f in range(a.shape[0]):
        if a[f].size == 0:
            out.append(0.0)
        elif f in (0, 3, 4):
            out.append(float(np.max(np.abs(a[f] - b[f]) / np.abs(b[f]))))
        else:
            scale = np.max(np.abs(b[f]))
            d = np.max(np.abs(a[f] - b[f]))
            out.append(d / scale if scale > 0 else d)
    return np.array(out)
