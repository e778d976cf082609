"""The Fortran drivers, run as child processes: cube_gpu's final state is the
Python path's bit for bit (same seeded inputs, same dt expression), its
every-6th-step output files exist in the reference layouts, and adjoint_gpu's
adjoint directional derivative agrees with its central difference."""
import os
import re
import subprocess

import numpy as np
import pytest

from solar_amd import Config, Solver, background, generate_ic

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CUBE = os.path.join(ROOT, "driver", "cube_gpu")
ADJ = os.path.join(ROOT, "driver", "adjoint_gpu")


@pytest.mark.skipif(not os.path.exists(CUBE), reason="Fortran driver not built (amdflang absent)")
def test_cube_gpu_matches_python_path(tmp_path):
    n, nr, steps, seed = 16, 24, 13, 1234
    out = tmp_path / "q.bin"
    r = subprocess.run([CUBE, str(n), str(nr), "1", "1", str(steps), str(seed), str(out), str(tmp_path / "o")],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    got = np.fromfile(out, dtype=np.float64).reshape(6, 5, nr, n, n)
    gr, c2, rho0, g0 = background(1, nr)
    q0, _ = generate_ic(1, n, nr, gr, c2, seed, [1e-3, 1e-6, 20.0])
    cfg = Config("drv", n, nr, 1, 1, steps)
    with Solver(n, nr, gr, c2, rho0, g0, cfg.sponge_len, 0.1) as s:
        s.set_state(q0)
        s.rk4(steps, cfg.dt(gr, c2))
        want = s.get_state()
    assert np.array_equal(got, want)
    # WRITE_VISUAL / WRITE_RADIAL_TRACE at steps 0, 6, 12
    for cue in ("000", "001", "002"):
        for p in range(6):
            assert (tmp_path / "o" / "Visual" / cue / f"vis_{p}").stat().st_size == n * n * 5 * 8 + 8
    for z in range(1, 7):
        assert len((tmp_path / "o" / f"trace_{z}").read_text().splitlines()) == 3 * nr


@pytest.mark.skipif(not os.path.exists(ADJ), reason="Fortran driver not built (amdflang absent)")
def test_adjoint_gpu_gradient_matches_central_difference():
    r = subprocess.run([ADJ, "16", "24", "30"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    m = re.search(r"J =\s*([0-9.Ee+-]+)", r.stdout)
    assert m and float(m.group(1)) > 0.0, r.stdout  # the receivers see the c^2 change
    m = re.search(r"rel =\s*([0-9.Ee+-]+)", r.stdout)
    assert m, r.stdout
    assert float(m.group(1)) < 1e-6
