"""On-disk formats of the step-adjacent output (output.f90), checked by
parsing the files independently of the writer: byte layout of the MPI-IO
restart (header, view offset 28, Fortran order), the unformatted sequential
record of WRITE_VISUAL, the (1x, 5(ES13.5)) trace with its format reversion,
WRITE_GRID's (1x, 7i5), and o_rad's tie rule.  Host-only (no GPU)."""
import os

import numpy as np
import pytest

from solar_amd import SolarError
from solar_amd.lib import (grid_write, output_radius, radial_trace_write, restart_read, restart_write,
                           visual_write)

N, NR = 16, 12


def _q(seed=0):
    return np.random.default_rng(seed).standard_normal((6, 5, NR, N, N))


def test_restart_layout_and_round_trip(tmp_path):
    q = _q(1)
    restart_write(tmp_path, q, 7, 0.125, 42)
    for p in range(6):
        f = tmp_path / "Restart" / f"restart.007.{p + 1}"
        raw = f.read_bytes()
        assert len(raw) == 28 + 8 * q[p].size
        assert np.frombuffer(raw[:16], "<i4").tolist() == [N, N, NR, 5]
        assert np.frombuffer(raw[16:24], "<f8")[0] == 0.125
        assert np.frombuffer(raw[24:28], "<i4")[0] == 42
        # q(nx,ny,nr,5) in Fortran order = [field][k][j][i] in C order
        data = np.frombuffer(raw[28:], "<f8").reshape(5, NR, N, N)
        assert np.array_equal(data, q[p])
    back, t, st = restart_read(tmp_path, N, NR, 7)
    assert np.array_equal(back, q) and t == 0.125 and st == 42
    with pytest.raises(SolarError):
        restart_read(tmp_path, N, NR + 1, 7)  # header sizes are checked
    with pytest.raises(SolarError):
        restart_read(tmp_path, N, NR, 8)  # missing cue
    restart_write(tmp_path, q, 123, 1.0, 1)
    assert (tmp_path / "Restart" / "restart.123.6").exists()


def test_visual_record(tmp_path):
    q = _q(2)
    visual_write(tmp_path, q, 3, 5)
    for p in range(6):
        raw = (tmp_path / "Visual" / "003" / f"vis_{p}").read_bytes()
        nbytes = N * N * 5 * 8
        assert len(raw) == nbytes + 8
        assert np.frombuffer(raw[:4], "<i4")[0] == nbytes
        assert np.frombuffer(raw[-4:], "<i4")[0] == nbytes
        # (((q_out(i,j,l), i), j), l) with q_out = q(:,:,o_rad,:)
        assert np.array_equal(np.frombuffer(raw[4:-4], "<f8").reshape(5, N, N), q[p, :, 4])
    with pytest.raises(SolarError):
        visual_write(tmp_path, q, 3, NR + 1)


def test_radial_trace_format(tmp_path):
    q = _q(3)
    q[0, 0, 0, 0, 0] = 0.0
    q[0, 0, 1, 0, 0] = -2.5e-120  # three-digit exponent: gfortran drops the E
    q[0, 0, 2, 0, 0] = 123456.7
    radial_trace_write(tmp_path, q)
    radial_trace_write(tmp_path, q)  # appended
    for p in range(6):
        lines = (tmp_path / f"trace_{p + 1}").read_text().splitlines()
        assert len(lines) == 2 * NR  # nr*5 values, 5 per record
        for rec in range(2):
            block = lines[rec * NR:(rec + 1) * NR]
            assert block[0].startswith(" ") and len(block[0]) == 1 + 65
            assert all(len(s) == 65 for s in block[1:])  # reversion to 5(ES13.5): no 1x
            fields = [block[0][1:][c:c + 13] for c in range(0, 65, 13)]
            for s in block[1:]:
                fields += [s[c:c + 13] for c in range(0, 65, 13)]
            want = [q[p, l, k, 0, 0] for l in range(5) for k in range(NR)]
            for got, w in zip(fields, want):
                if w == -2.5e-120:
                    assert got == " -2.50000-120"
                else:
                    assert got == "%13.5E" % w
    assert (tmp_path / "trace_1").read_text().splitlines()[0][:27] == " " + "  0.00000E+00" + " -2.50000-120"


def test_grid_file(tmp_path):
    grid_write(tmp_path, 32, 64)
    for p in range(6):
        assert (tmp_path / "Visual" / "grid" / f"grid_{p}").read_text() == \
            " " + "".join("%5d" % v for v in (0, 31, 0, 31, 0, 63, p + 1)) + "\n"


def test_output_radius_rule():
    gr = np.linspace(0.7, 1.0, 31)
    assert output_radius(gr, 0.98) == int(np.argmin(np.abs(gr - 0.98))) + 1
    assert output_radius(np.array([0.0, 1.0, 2.0]), 0.5) == 2  # .LE. keeps the later of a tie
    with pytest.raises(SolarError):
        output_radius(np.array([5.0, 6.0]), 0.0)


def test_writers_create_missing_directories(tmp_path):
    d = tmp_path / "a" / "b"
    restart_write(d, _q(4), 0, 0.0, 0)
    assert os.path.exists(d / "Restart" / "restart.000.1")
