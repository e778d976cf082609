"""FP32-state variant: state and 2N accumulator stored as float, all
arithmetic, metric terms, interpolation weights and ghost strips in FP64.
Compared with the FP64 CPU restatement on the same inputs (the error is the
float rounding of the stored state, stated per field in the assertion), and
bit for bit with the restatement when it rounds its stored state the same way."""
import numpy as np
import pytest

import oracle
from solar_amd import CONFIGS, Solver, make_inputs
from tests.conftest import per_field_rel_l2

pytestmark = pytest.mark.gpu


def test_fp32_state_c1_full_run():
    cfg = CONFIGS["c1"]
    bg, q0, dt, _ = make_inputs(cfg)
    with Solver(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge, state_bits=32) as s:
        s.set_state(q0)
        s.rk4(cfg.steps, dt)
        got = s.get_state()
    want, _ = oracle.Oracle(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge).rk4(q0, cfg.steps, dt)
    errs = per_field_rel_l2(got, want)
    print(f"[fp32] c1 x{cfg.steps}: rel_l2 per field vs FP64 {['%.2e' % e for e in errs]}", flush=True)
    # relative L2 per field against the FP64 solution after 200 steps
    assert max(errs) < 1e-5, errs
    assert errs[4] > 1e-9  # it really is FP32 storage


def test_fp32_rhs_and_filter_match_fp64_on_float_inputs():
    cfg = CONFIGS["t0"]
    bg, q0, dt, _ = make_inputs(cfg)
    q = np.random.default_rng(11).standard_normal(q0.shape).astype(np.float32).astype(np.float64)
    o = oracle.Oracle(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge)
    with Solver(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge, state_bits=32) as s:
        rhs = s.compute_rhs(q, 3)
        qf = s.filter_exp(q)
    # inputs are exactly representable, compute is FP64, only the output is rounded to float
    for got, want in ((rhs, o.stage_rhs(q)), (qf, o.filter_exp(q))):
        rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
        assert np.all((rel <= 2 ** -24) | (np.abs(got - want) <= 1e-38)), rel.max()


def _float_storage_steps(o, q0, steps, dt):
    """RK4 steps of the CPU restatement with the state and the 2N accumulator
    rounded to float after every stage and after the filter, as the FP32-state
    kernels store them: the RHS, the 2N update and the filter are evaluated
    in FP64 on the float-valued inputs, in the same order (timestep.f90:61-75)."""
    alpha, beta = oracle.rk_coeffs()
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    q = f32(q0)
    for _ in range(steps):
        w = np.zeros_like(q)
        for s in range(5):
            wn = alpha[s] * w + dt * o.stage_rhs(q)
            q = f32(q + beta[s] * wn)
            w = f32(wn)
        q = f32(o.filter_exp(q))
    return q


@pytest.mark.parametrize("name,n,nr,steps", [("t0", None, None, 3), ("tiles", 48, 20, 2), ("chunks", 32, 131, 1),
                                              ("wide", 352, 12, 1)])
def test_fp32_steps_equal_float_rounded_restatement(name, n, nr, steps):
    """The FP32-state steps are bit-identical to the CPU restatement with the
    same float rounding of the stored state (at the smoke grid, a grid of
    16x16 tiles, one with uneven radial chunks, and one whose gathers hold two
    strip points per thread)."""
    from solar_amd import background, generate_ic
    if name == "t0":
        cfg = CONFIGS["t0"]
        bg, q0, dt, _ = make_inputs(cfg)
        n, nr, L = cfg.n, cfg.nr, cfg.sponge_len
    else:
        bg = background(1, nr)
        L = max(1, round(0.08 * nr))
        q0, _ = generate_ic(1, n, nr, bg[0], bg[1], 17, [1e-3, 1e-6, 8])
        q0 = q0 + 1e-4 * np.random.default_rng(n).standard_normal(q0.shape)
        dt = 0.5 * min(bg[0][1] - bg[0][0], bg[0][0] * (np.pi / 2) / (n - 1))
    print(f"[case] fp32 exact {name} n={n} nr={nr} steps={steps}", flush=True)
    o = oracle.Oracle(n, nr, *bg, L, 0.1)
    want = _float_storage_steps(o, q0, steps, dt)
    with Solver(n, nr, *bg, L, 0.1, state_bits=32) as s:
        s.set_state(q0)
        s.rk4(steps, dt)
        got = s.get_state()
    ndiff = int(np.count_nonzero(got != want))
    print(f"[fp32] {name}: differing elements {ndiff} of {got.size}", flush=True)
    assert ndiff == 0
