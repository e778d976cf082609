"""Panel sizes the tiles do not divide.

The reference takes any nx = ny (header:4-10; its production grid is 128^2 and
the post-processing data hold 192^2, 200^2 and 204^2 panels).  For n not a
multiple of 16 the stage and filter kernels run partial tiles at the E and N
panel edges.  Checked against the CPU restatement bit for bit: the RHS with
BC and sponge, FILTER_EXP, the panel exchange, and several RK4 steps, in FP64
and with FP32 storage.  The adjoint and the ensemble, which tile the panel
exactly, refuse such n with an error.
"""
import numpy as np
import pytest

import oracle
from solar_amd import RHS_BC, RHS_SPONGE, SolarError, Solver, background, generate_ic
from tests.conftest import per_field_rel_l2

pytestmark = pytest.mark.gpu

# n: 20 (one partial tile per row), 100 (16x16 tiles, 4 columns past the edge),
# 200 and 204 (the reference's post-processing grids; 204 takes 32x8 tiles)
SIZES = [(20, 16), (100, 14), (200, 13), (204, 12)]


def _inputs(n, nr, seed):
    gr, c2, rho0, g0 = background(1, nr)
    q, _ = generate_ic(1, n, nr, gr, c2, seed, [1e-3, 1e-6, 6.0])
    q = q + 1e-4 * np.random.default_rng(seed).standard_normal(q.shape)
    L = max(1, round(0.08 * nr))
    dt = 0.5 * min(gr[1] - gr[0], gr[0] * (np.pi / 2) / (n - 1))
    return (gr, c2, rho0, g0), q, L, dt


def _same(got, want, what):
    ndiff = int(np.count_nonzero(got != want))
    errs = per_field_rel_l2(got, want)
    print(f"[sizes] {what}: differing={ndiff} of {got.size} rel_l2={['%.1e' % e for e in errs]}", flush=True)
    assert ndiff == 0, what


@pytest.mark.parametrize("n,nr", SIZES)
def test_any_panel_size_bit_exact(n, nr):
    print(f"[case] sizes n={n} nr={nr}", flush=True)
    bg, q, L, dt = _inputs(n, nr, n)
    o = oracle.Oracle(n, nr, *bg, L, 0.1)
    with Solver(n, nr, *bg, L, 0.1) as s:
        f = np.random.default_rng(n + 1).standard_normal((6, nr, n, n))
        gx, gy = s.scalar_ghosts(f)
        ox, oy = o.scalar_ghosts(f)
        assert np.array_equal(gx, ox) and np.array_equal(gy, oy)
        _same(s.compute_rhs(q, RHS_BC | RHS_SPONGE), o.stage_rhs(q), f"n={n} rhs")
        _same(s.filter_exp(q), o.filter_exp(q), f"n={n} filter")
        s.set_state(q)
        s.rk4(3, dt)
        got = s.get_state()
    want, _ = o.rk4(q, 3, dt)
    _same(got, want, f"n={n} rk4 x3")


def test_any_panel_size_fp32_state():
    """FP32 storage at n = 100: bit-identical to the restatement that rounds
    its stored state to float after every stage and after the filter."""
    from tests.test_gpu_fp32 import _float_storage_steps
    n, nr = 100, 14
    bg, q, L, dt = _inputs(n, nr, 7)
    o = oracle.Oracle(n, nr, *bg, L, 0.1)
    want = _float_storage_steps(o, q, 2, dt)
    with Solver(n, nr, *bg, L, 0.1, state_bits=32) as s:
        s.set_state(q)
        s.rk4(2, dt)
        got = s.get_state()
    _same(got, want, "n=100 FP32 state x2")


def test_adjoint_and_ensemble_refuse_untiled_sizes():
    from solar_amd import Ensemble
    bg, q, L, dt = _inputs(100, 14, 3)
    with Solver(100, 14, *bg, L, 0.1) as s:
        with pytest.raises(SolarError, match="multiple of 16"):
            s.adj_rhs(q)
    gr, c2, rho0, g0 = bg
    member = dict(n=100, nr=14, gr=gr, c2=c2, rho0=rho0, g0=g0, sponge_len=L, a_sponge=0.1, dt=dt)
    with pytest.raises(SolarError, match="multiple of 16"):
        Ensemble([member])


@pytest.mark.parametrize("n", [32, 100])
def test_two_dimensional_case_bit_exact(n):
    """nr = 1, the reference's 2-D case: no radial derivative, FILTER_R, BC_RHS
    or SPONGE, and r = 1.  RHS, FILTER_EXP and RK4 steps against the CPU
    restatement, bit for bit (n = 100 also runs partial tiles)."""
    gr, c2, rho0, g0 = (np.array([v]) for v in (0.85, 1.3, 0.9, 0.4))
    rng = np.random.default_rng(n)
    q = rng.standard_normal((6, 5, 1, n, n))
    dt = 0.5 * (np.pi / 2) / (n - 1) / np.sqrt(1.3)
    print(f"[case] 2-D n={n} nr=1", flush=True)
    o = oracle.Oracle(n, 1, gr, c2, rho0, g0, 1, 0.1)
    with Solver(n, 1, gr, c2, rho0, g0, 1, 0.1) as s:
        _same(s.compute_rhs(q, RHS_BC | RHS_SPONGE), o.stage_rhs(q), f"2-D n={n} rhs")
        _same(s.filter_exp(q), o.filter_exp(q), f"2-D n={n} filter")
        s.set_state(q)
        s.rk4(5, dt)
        got = s.get_state()
        with pytest.raises(SolarError, match="multiple of 16|nr >= 12"):
            s.adj_rhs(q)
    want, _ = o.rk4(q, 5, dt)
    _same(got, want, f"2-D n={n} rk4 x5")
