"""Discrete adjoint of the RK4 step (sensitivities; the reference has none).

Pinned by what the mathematics fixes rather than by a second implementation:
  - transpose identities <A x, y> = <x, A^T y> for A = L (one right-hand side
    with BC_RHS and SPONGE), F (FILTER_EXP) and S (one RK4 step), on random
    inputs, radial and 3-D c^2 fields, both tile shapes;
  - the misfit gradient dJ/dc^2 against central finite differences of J along
    random directions, and its independence of the checkpoint interval.
The forward path with a 3-D c^2 field is checked bit for bit against the CPU
restatement, whose c2(i,j,k) array mirrors compute_rhs_acoustic.f90:134.
"""
import numpy as np
import pytest

import oracle
from solar_amd import RHS_BC, RHS_SPONGE, Config, Solver, background, generate_ic

pytestmark = pytest.mark.gpu

# (n, nr, background kind): n = 32 takes the 32x8 adjoint tiles, 16 and 48 the 16x16 ones;
# nr = 100 splits the radial march (4 stage chunks of 25 planes, 2 filter chunks of 50)
CASES = [(16, 24, 1), (32, 40, 1), (48, 32, 0), (16, 100, 1)]


def _setup(n, nr, kind, steps=1):
    cfg = Config("adj", n, nr, kind, 0, steps)
    bg = background(kind, nr)
    return cfg, bg, cfg.dt(bg[0], bg[1])


def _c2_field(bg, n, nr, seed):
    """Positive 3-D c^2: the radial profile times a random +-10% modulation."""
    rng = np.random.default_rng(seed)
    return bg[1][None, :, None, None] * (1.0 + 0.1 * rng.uniform(-1, 1, (6, nr, n, n)))


def _transpose_gap(ax, y, x, aty):
    """|<Ax,y> - <x,A^T y>| relative to |Ax||y| (rounding of FP64 sums is ~1e-16 of that)."""
    lhs, rhs = float(np.vdot(ax, y)), float(np.vdot(x, aty))
    return abs(lhs - rhs) / (np.linalg.norm(ax) * np.linalg.norm(y)), lhs, rhs


@pytest.mark.parametrize("n,nr,kind", CASES)
@pytest.mark.parametrize("field", [False, True])
def test_rhs_transpose(n, nr, kind, field):
    cfg, bg, _ = _setup(n, nr, kind)
    rng = np.random.default_rng(n * 100 + nr + field)
    x = rng.standard_normal((6, 5, nr, n, n))
    w = rng.standard_normal((6, 5, nr, n, n))
    with Solver(n, nr, *bg, cfg.sponge_len, cfg.a_sponge) as s:
        if field:
            s.set_c2_field(_c2_field(bg, n, nr, 7))
        lx = s.compute_rhs(x, RHS_BC | RHS_SPONGE)
        ltw = s.adj_rhs(w)
    gap, a, b = _transpose_gap(lx, w, x, ltw)
    print(f"[adj] L n={n} nr={nr} field={field}: <Lx,w>={a:.15e} <x,LTw>={b:.15e} gap={gap:.2e}", flush=True)
    assert gap < 1e-13


@pytest.mark.parametrize("n,nr,kind", CASES)
def test_filter_transpose(n, nr, kind):
    cfg, bg, _ = _setup(n, nr, kind)
    rng = np.random.default_rng(n + 3 * nr)
    x = rng.standard_normal((6, 5, nr, n, n))
    o = rng.standard_normal((6, 5, nr, n, n))
    with Solver(n, nr, *bg, cfg.sponge_len, cfg.a_sponge) as s:
        fx = s.filter_exp(x)
        fto = s.adj_filter(o)
    gap, a, b = _transpose_gap(fx, o, x, fto)
    print(f"[adj] F n={n} nr={nr}: gap={gap:.2e}", flush=True)
    assert gap < 1e-13


@pytest.mark.parametrize("n,nr", [(176, 13), (352, 12), (512, 12)])
def test_transposes_wide_panels(n, nr):
    """L, F and S transposes on wide panels: the transposed gathers run one strip
    point per thread up to 1024 threads (n = 176), and beyond n = 341 three
    (vector strips) or two (scalar strips) per thread (n = 352, 512); at
    n = 512 the staged lines take more than 64 KiB of LDS."""
    cfg, bg, _ = _setup(n, nr, 1)
    rng = np.random.default_rng(n + nr)
    x = rng.standard_normal((6, 5, nr, n, n))
    w = rng.standard_normal((6, 5, nr, n, n))
    with Solver(n, nr, *bg, cfg.sponge_len, cfg.a_sponge) as s:
        gap, a, b = _transpose_gap(s.compute_rhs(x, RHS_BC | RHS_SPONGE), w, x, s.adj_rhs(w))
        print(f"[adj] L n={n} nr={nr}: gap={gap:.2e}", flush=True)
        assert gap < 1e-13
        gap, a, b = _transpose_gap(s.filter_exp(x), w, x, s.adj_filter(w))
        print(f"[adj] F n={n} nr={nr}: gap={gap:.2e}", flush=True)
        assert gap < 1e-13
        # the whole step, whose transposed gathers also carry the W update
        dt = _setup(n, nr, 1)[2]
        s.set_state(x)
        s.rk4(1, dt)
        gap, a, b = _transpose_gap(s.get_state(), w, x, s.adj_step(w, dt))
        print(f"[adj] S n={n} nr={nr}: gap={gap:.2e}", flush=True)
        assert gap < 1e-13


@pytest.mark.parametrize("n,nr,kind", CASES)
def test_step_transpose(n, nr, kind):
    cfg, bg, dt = _setup(n, nr, kind)
    rng = np.random.default_rng(5 * n + nr)
    x = rng.standard_normal((6, 5, nr, n, n))
    lam = rng.standard_normal((6, 5, nr, n, n))
    with Solver(n, nr, *bg, cfg.sponge_len, cfg.a_sponge) as s:
        s.set_c2_field(_c2_field(bg, n, nr, 9))
        s.set_state(x)
        s.rk4(1, dt)
        sx = s.get_state()
        stl = s.adj_step(lam, dt)
    gap, a, b = _transpose_gap(sx, lam, x, stl)
    print(f"[adj] S n={n} nr={nr}: gap={gap:.2e}", flush=True)
    assert gap < 1e-13


def test_c2_field_forward_matches_oracle_bitwise():
    n, nr, kind = 16, 24, 1
    cfg, bg, dt = _setup(n, nr, kind)
    q0, _ = generate_ic(1, n, nr, bg[0], bg[1], 321, [1e-3, 1e-6, 20.0])
    c2f = _c2_field(bg, n, nr, 11)
    o = oracle.Oracle(n, nr, *bg, cfg.sponge_len, cfg.a_sponge)
    o.set_c2_field(c2f)
    want, _ = o.rk4(q0, 5, dt)
    with Solver(n, nr, *bg, cfg.sponge_len, cfg.a_sponge) as s:
        s.set_c2_field(c2f)
        s.set_state(q0)
        s.rk4(5, dt)
        got = s.get_state()
        # the radial profile broadcast to a field is the radial run, bit for bit
        s.set_c2_field(np.broadcast_to(bg[1][None, :, None, None], (6, nr, n, n)))
        s.set_state(q0)
        s.rk4(5, dt)
        bcast = s.get_state()
        s.set_c2_field(None)
        s.set_state(q0)
        s.rk4(5, dt)
        radial = s.get_state()
    assert np.array_equal(got, want)
    assert np.array_equal(bcast, radial)
    assert not np.array_equal(got, radial)


def _gradient_problem(n=16, nr=24, steps=24, nrec=8, seed=2):
    cfg, bg, dt = _setup(n, nr, 1, steps)
    q0, _ = generate_ic(0, n, nr, bg[0], bg[1], 77, [1e-3, 0.05, 0.0])
    rng = np.random.default_rng(seed)
    rec = np.stack([rng.integers(0, 6, nrec), rng.integers(4, nr - 4, nrec), rng.integers(0, n, nrec),
                    rng.integers(0, n, nrec)], axis=1)
    # "true" model: the radial profile with a 2% Gaussian blob
    xyz_k = np.arange(nr)[None, :, None, None]
    blob = np.exp(-((xyz_k - nr / 2) / 4.0) ** 2) * np.ones((6, nr, n, n))
    c2_0 = np.broadcast_to(bg[1][None, :, None, None], (6, nr, n, n)).copy()
    c2_true = c2_0 * (1.0 + 0.02 * blob)
    return cfg, bg, dt, q0, rec, c2_0, c2_true


def test_misfit_gradient_matches_finite_differences():
    cfg, bg, dt, q0, rec, c2_0, c2_true = _gradient_problem()
    steps = cfg.steps
    with Solver(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge) as s:
        s.set_receivers(rec)
        s.set_c2_field(c2_true)
        s.set_state(q0)
        data = s.record(steps, dt)

        def J(c2):
            s.set_c2_field(c2)
            s.set_state(q0)
            d = s.record(steps, dt) - data
            return float(np.sum(d * d))

        s.set_c2_field(c2_0)
        J0, grad, seis = s.misfit_gradient(q0, steps, dt, data, checkpoint_every=5)
        assert J0 > 0
        # same traces; only the order of the final sum differs (host loop vs numpy)
        assert abs(J0 - J(c2_0)) <= 1e-13 * J0
        print(f"[adj] J0={J0:.6e}", flush=True)
        # no atomics anywhere on the path: the same call gives the same bits
        _, grad_again, _ = s.misfit_gradient(q0, steps, dt, data, checkpoint_every=5)
        assert np.array_equal(grad, grad_again)
        # checkpoints restart the forward from the exact state, so any segmentation gives the same bits
        _, grad_one_segment, _ = s.misfit_gradient(q0, steps, dt, data, checkpoint_every=1000)
        assert np.array_equal(grad, grad_one_segment)
        # the divergence cache holds every segment here (nothing recomputed); without it the four
        # earlier segments of 5 steps are recomputed from their checkpoints: the same bits
        assert s.recomputed_steps == 0
        s.set_adjoint_cache(False)
        _, grad_recomputed, _ = s.misfit_gradient(q0, steps, dt, data, checkpoint_every=5)
        assert s.recomputed_steps == steps - 5
        s.set_adjoint_cache(True)
        assert np.array_equal(grad, grad_recomputed)
        # FP32 storage of the stage divergences: the gradient carries their rounding only
        s.set_adjoint_options(32)
        _, grad32, _ = s.misfit_gradient(q0, steps, dt, data, checkpoint_every=5)
        s.set_adjoint_options(64)
        rel32 = np.linalg.norm(grad32 - grad) / np.linalg.norm(grad)
        print(f"[adj] FP32 divergences: rel L2 vs FP64 {rel32:.2e}", flush=True)
        assert 0.0 < rel32 < 1e-6

        rng = np.random.default_rng(4)
        errs = []
        for t in range(8):
            d = rng.standard_normal(c2_0.shape) * c2_0
            eps = 1e-5
            fd = (J(c2_0 + eps * d) - J(c2_0 - eps * d)) / (2 * eps)
            ad = float(np.vdot(grad, d))
            errs.append(abs(fd - ad) / abs(fd))
            print(f"[adj] grad dir {t}: fd={fd:.12e} adj={ad:.12e} rel={errs[-1]:.2e}", flush=True)
    assert max(errs) < 1e-6, errs


def test_autograd_harness():
    import torch

    from solar_amd.autograd import MisfitC2, RK4Step
    cfg, bg, dt, q0, rec, c2_0, c2_true = _gradient_problem(steps=10)
    n, nr = cfg.n, cfg.nr
    rng = np.random.default_rng(12)
    with Solver(n, nr, *bg, cfg.sponge_len, cfg.a_sponge) as s:
        # two chained steps: d<S S x, lam>/dx = S^T S^T lam
        x = torch.tensor(rng.standard_normal((6, 5, nr, n, n)), requires_grad=True)
        lam = torch.tensor(rng.standard_normal((6, 5, nr, n, n)))
        y = RK4Step.apply(RK4Step.apply(x, s, dt), s, dt)
        (y * lam).sum().backward()
        want = s.adj_step(s.adj_step(lam.numpy(), dt), dt)
        # autograd chains the two transposed steps itself, so compare to rounding
        assert np.allclose(x.grad.numpy(), want, rtol=0, atol=1e-13 * np.abs(want).max())
        lhs = float((y.detach() * lam).sum())
        rhs = float((x.detach() * x.grad).sum())
        assert abs(lhs - rhs) <= 1e-13 * float(y.norm() * lam.norm())

        s.set_receivers(rec)
        s.set_c2_field(c2_true)
        s.set_state(q0)
        data = s.record(cfg.steps, dt)
        c2 = torch.tensor(c2_0, requires_grad=True)
        J = MisfitC2.apply(c2, s, torch.tensor(q0), cfg.steps, dt, data, 4)
        (3.0 * J).backward()
        J2, grad, _ = s.misfit_gradient(q0, cfg.steps, dt, data, checkpoint_every=4)
    assert abs(float(J) - J2) <= 1e-14 * J2
    assert np.allclose(c2.grad.numpy(), 3.0 * grad, rtol=1e-12, atol=1e-12 * np.abs(grad).max())


def test_tangent_linear_dot_product():
    """<2 (seis - d), TL dc2> = <dJ/dc2, dc2>: the tangent-linear traces and the
    adjoint gradient agree to rounding (no finite-difference error); the
    tangent-linear traces also match central differences of the traces."""
    cfg, bg, dt, q0, rec, c2_0, c2_true = _gradient_problem()
    steps = cfg.steps
    with Solver(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge) as s:
        s.set_receivers(rec)
        s.set_c2_field(c2_true)
        s.set_state(q0)
        data = s.record(steps, dt)
        s.set_c2_field(c2_0)
        J, grad, seis = s.misfit_gradient(q0, steps, dt, data, checkpoint_every=5)
        rng = np.random.default_rng(21)
        for t in range(3):
            dc2 = rng.standard_normal(c2_0.shape) * c2_0
            s.set_c2_field(c2_0)
            dseis = s.tangent_linear(q0, steps, dt, dc2)
            lhs = float(np.sum(2.0 * (seis - data) * dseis))
            rhs = float(np.vdot(grad, dc2))
            print(f"[adj] TL dot test {t}: <2r, TL dc2>={lhs:.15e} <grad, dc2>={rhs:.15e} "
                  f"rel={abs(lhs - rhs) / abs(rhs):.2e}", flush=True)
            assert abs(lhs - rhs) <= 1e-10 * abs(rhs)
            if t == 0:
                eps = 1e-6
                s.set_c2_field(c2_0 + eps * dc2)
                s.set_state(q0)
                sp = s.record(steps, dt)
                s.set_c2_field(c2_0 - eps * dc2)
                s.set_state(q0)
                sm = s.record(steps, dt)
                fd = (sp - sm) / (2 * eps)
                rel = np.linalg.norm(fd - dseis) / np.linalg.norm(dseis)
                print(f"[adj] TL traces vs central differences: rel L2 {rel:.2e}", flush=True)
                assert rel < 1e-5


def test_receiver_buffer_replaced_not_leaked():
    """Growing receiver sets replace the device buffer: the handle's device
    bytes are its base plus the largest set so far (8 B per receiver)."""
    cfg, bg, dt = _setup(16, 24, 1)
    rng = np.random.default_rng(9)
    with Solver(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge) as s:
        base = s.device_bytes()
        for count, cap in ((4, 4), (8, 8), (16, 16), (2, 16), (16, 16)):
            rec = np.stack([rng.integers(0, 6, count), rng.integers(0, cfg.nr, count),
                            rng.integers(0, cfg.n, count), rng.integers(0, cfg.n, count)], axis=1)
            s.set_receivers(rec)
            assert s.device_bytes() == base + 8 * cap, (count, s.device_bytes() - base)
        s.set_state(np.zeros((6, 5, cfg.nr, cfg.n, cfg.n)))
        assert s.record(2, dt).shape == (2, 16)
