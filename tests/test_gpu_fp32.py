"""FP32-state variant: state and 2N accumulator stored as float, all
arithmetic, metric terms, interpolation weights and ghost strips in FP64.
Compared with the FP64 CPU restatement on the same inputs; the error is the
float rounding of the stored state, so the bound is a few ulp of float growth
over the run, stated per field in the assertion."""
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
