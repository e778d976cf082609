"""Device-state output: the handle's writers produce the same files as the
host writers on the downloaded state, restarts round-trip through the device,
and the norm diagnostic matches its closed form on the host copy."""
import numpy as np
import pytest

from solar_amd import CONFIGS, Solver, make_inputs
from solar_amd.lib import radial_trace_write, visual_write

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bits", [64, 32])
def test_handle_writers_and_restart_round_trip(tmp_path, bits):
    cfg = CONFIGS["t0"]
    bg, q0, dt, _ = make_inputs(cfg)
    with Solver(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge, state_bits=bits) as s:
        s.set_state(q0)
        s.rk4(3, dt)
        q = s.get_state()
        s.write_restart(tmp_path / "dev", 2, 3 * dt, 3)
        s.write_visual(tmp_path / "dev", 2, 7)
        s.write_radial_trace(tmp_path / "dev")
        visual_write(tmp_path / "host", q, 2, 7)
        radial_trace_write(tmp_path / "host", q)
        for p in range(6):
            assert (tmp_path / "dev" / "Visual" / "002" / f"vis_{p}").read_bytes() == \
                (tmp_path / "host" / "Visual" / "002" / f"vis_{p}").read_bytes()
            assert (tmp_path / "dev" / f"trace_{p + 1}").read_text() == \
                (tmp_path / "host" / f"trace_{p + 1}").read_text()
        s.set_state(np.zeros_like(q0))
        t, st = s.read_restart(tmp_path / "dev", 2)
        assert t == 3 * dt and st == 3
        assert np.array_equal(s.get_state(), q)
        # auxiliary.f90:122 as written, and the RMS
        for f in range(5):
            ss = float(np.sum(q[:, f].astype(np.float64) ** 2))
            want_ref = (ss / (cfg.n * cfg.n * cfg.nr) * 6.0) ** 0.5
            want_rms = (ss / (6 * cfg.n * cfg.n * cfg.nr)) ** 0.5
            assert s.norm(f, True) == pytest.approx(want_ref, rel=1e-12, abs=1e-300)
            assert s.norm(f, False) == pytest.approx(want_rms, rel=1e-12, abs=1e-300)
        assert s.norm(3) == pytest.approx(s.norm_vr(), rel=1e-14)
