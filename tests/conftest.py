import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU reference run")


@pytest.fixture(scope="session")
def t0_inputs():
    from solar_amd import CONFIGS, make_inputs
    cfg = CONFIGS["t0"]
    return (cfg,) + make_inputs(cfg)


def rel_l2(a, b):
    """Relative L2 error of a against b; 0 if both are zero."""
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    if den == 0:
        return 0.0 if num == 0 else np.inf
    return num / den


def per_field_rel_l2(a, b):
    """Relative L2 per field for arrays shaped (6, 5, nr, n, n)."""
    return [rel_l2(a[:, f], b[:, f]) for f in range(5)]


def cube_points(p, xi, eta):
    """Unit vectors of panel p at (xi, eta), the panel geometry of metric.f90:97-158."""
    a, b = np.tan(xi), np.tan(eta)
    o = np.ones_like(a)
    v = {0: (o, a, b), 1: (-a, o, b), 2: (-o, -a, b), 3: (a, -o, b), 4: (-b, a, o), 5: (b, a, -o)}[p]
    v = np.stack(v, -1)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def sph_basis(x):
    th = np.arccos(np.clip(x[..., 2], -1, 1))
    ph = np.arctan2(x[..., 1], x[..., 0])
    eth = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)], -1)
    eph = np.stack([-np.sin(ph), np.cos(ph), 0 * ph], -1)
    er = x
    return eth, eph, er


def theta_phi_to_xieta(p, xi, eta, vt, vp):
    """Panel-basis components from spherical ones: the inverse conversion of
    transform.f90:63-111, written from the panel geometry."""
    tx, te = np.tan(xi), np.tan(eta)
    C, D = np.sqrt(1 + tx ** 2), np.sqrt(1 + te ** 2)
    dl = 1 + tx ** 2 + te ** 2
    sd = np.sqrt(dl)
    if p < 4:
        return C * D / sd * vp, -vt + tx * te / sd * vp
    fac = 1 / np.sqrt(dl - 1)
    if p == 4:
        return fac * (D * tx * vt - D * te / sd * vp), fac * (C * te * vt + C * tx / sd * vp)
    return fac * (-D * tx * vt + D * te / sd * vp), -fac * (C * te * vt + C * tx / sd * vp)


def grid(n):
    g = -np.pi / 4 + np.arange(n) * (np.pi / 2) / (n - 1)
    return g, g[1] - g[0]
