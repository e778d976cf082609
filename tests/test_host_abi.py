"""CPU tests of the product library's host side: the C ABI loads and exports
every entry point of include/solar_amd.h, the host tables equal the oracle's
bit for bit, the seeded inputs are deterministic, and the handle fails loudly
without a GPU."""
import os
import re

import numpy as np
import pytest

import oracle
import solar_amd
from solar_amd import CONFIGS, background, generate_ic, host_tables, load_library
from solar_amd.lib import EXPORTS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = load_library()
    hdr = open(os.path.join(ROOT, "include", "solar_amd.h")).read()
    names = set(re.findall(r"\b(solar_\w+)\s*\(", hdr))
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(lib, n), n
    # the Python binding declares every exported entry point
    assert names == {e[0] for e in EXPORTS}


@pytest.mark.parametrize("n,nr,kind", [(16, 24, 1), (32, 64, 0), (48, 20, 1), (96, 40, 1), (20, 24, 1), (100, 30, 1),
                                       (24, 1, 1)])
def test_host_tables_equal_oracle_tables(n, nr, kind):
    # nr = 1: the 2-D case, one shell (its radius is replaced by 1)
    gr, c2, rho0, g0 = background(kind, nr) if nr > 1 else (np.array([v]) for v in (0.85, 1.3, 0.9, 0.4))
    L = max(1, round(0.08 * nr))
    o = oracle.Oracle(n, nr, gr, c2, rho0, g0, L, 0.1)
    t = o.tables()
    h = host_tables(n, nr, gr, c2, rho0, g0, L, 0.1)
    for k in t:
        assert np.array_equal(t[k], h[k]), k
    assert np.array_equal(h["sigma"], [o.sponge_sigma(k + 1) for k in range(nr)])
    # DDR(rho0) of compute_rhs_acoustic.f90:51 on the radial profile
    d = o.ddr(rho0[:, None])[:, 0]
    assert np.array_equal(d, h["drho0dr"])


def test_invalid_sizes_rejected():
    gr, c2, rho0, g0 = background(0, 24)
    for n in (15, 513):
        with pytest.raises(solar_amd.SolarError, match=r"\[16, 512\]"):
            host_tables(n, 24, gr, c2, rho0, g0, 2, 0.1)
    for nr in (2, 11):
        with pytest.raises(solar_amd.SolarError, match="1 or at least 12"):
            host_tables(16, nr, gr[:nr], c2[:nr], rho0[:nr], g0[:nr], 1, 0.1)
    with pytest.raises(solar_amd.SolarError, match="sponge"):
        host_tables(16, 24, gr, c2, rho0, g0, 20, 0.1)


def test_background_hydrostatic_and_units():
    gr, c2, rho0, g0 = background(1, 200)
    assert gr[0] == 0.7 and abs(gr[-1] - 1.0) < 1e-15
    assert c2[0] == 1.0 and rho0[0] == 1.0  # nondimensionalised at the first point
    p0 = rho0 * c2 / (5.0 / 3.0)
    dp = np.gradient(p0, gr, edge_order=2)
    assert np.abs(dp + rho0 * g0).max() < 1e-3 * np.abs(rho0 * g0).max()


def test_ic_seeded_and_deterministic():
    gr, c2, rho0, g0 = background(1, 24)
    a, ia = generate_ic(1, 16, 24, gr, c2, 7, [1e-3, 1e-6, 20])
    b, ib = generate_ic(1, 16, 24, gr, c2, 7, [1e-3, 1e-6, 20])
    c, ic = generate_ic(1, 16, 24, gr, c2, 8, [1e-3, 1e-6, 20])
    assert np.array_equal(a, b) and np.array_equal(ia, ib)
    assert not np.array_equal(a, c)
    assert 1 <= ia[0] <= 20 and abs(ia[1]) <= ia[0]
    assert np.all(a[:, :4] == 0) and np.abs(a[:, 4]).max() > 1e-4
    q, info = generate_ic(0, 32, 64, *background(0, 64)[:2], 1234, [1e-3, 0.05, 0])
    assert abs(q[:, 4].max() - 1e-3) < 2e-4 and np.array_equal(q[:, 0], q[:, 4])


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_init_without_gpu_fails_loudly():
    cfg = CONFIGS["t0"]
    gr, c2, rho0, g0 = background(1, cfg.nr)
    with pytest.raises(solar_amd.SolarError, match="HIP"):
        solar_amd.Solver(cfg.n, cfg.nr, gr, c2, rho0, g0, cfg.sponge_len)
