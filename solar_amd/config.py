"""Workload configurations of the acoustic RK4 hot path.

Each entry fixes grid, background, initial condition, seed, sponge and the
time step, so the same inputs reach the GPU path and the CPU checker.
dt = cfl * min(dr, r_min*dxi) / c_max is the fixed time-step formula (the
reference discards its own CFL estimate and hard-codes dt, timestep.f90:33-35).
"""
import math
from dataclasses import dataclass

from .lib import background, generate_ic


@dataclass(frozen=True)
class Config:
    name: str
    n: int           # angular points per panel edge
    nr: int          # radial points
    background: int  # 0 uniform, 1 stratified
    ic: int          # 0 Gaussian pulse, 1 spherical-harmonic mode
    steps: int
    seed: int = 1234
    cfl: float = 0.5
    a_sponge: float = 0.1
    sponge_frac: float = 0.08  # sponge over the outer 8% of the radial points at each end

    @property
    def sponge_len(self):
        return max(1, int(round(self.sponge_frac * self.nr)))

    @property
    def points(self):
        return 6 * self.n * self.n * self.nr

    def dt(self, gr, c2):
        dxi = (math.pi / 2) / (self.n - 1)
        h = min(gr[1] - gr[0], gr[0] * dxi)
        return self.cfl * h / math.sqrt(max(c2))


CONFIGS = {
    # tiny parity case
    "t0": Config("t0", 16, 24, 1, 0, 10),
    # Gaussian pulse, uniform background, 6x32x32x64, 200 steps
    "c1": Config("c1", 32, 64, 0, 0, 200),
    # stratified, spherical-harmonic mode + noise, 6x128x128x256, 1000 steps
    "c2": Config("c2", 128, 256, 1, 1, 1000),
    # the same at 6x256x256x384, 2000 steps
    "c3": Config("c3", 256, 384, 1, 1, 2000),
    # forward of the differentiation case, 6x96x96x192
    "c5": Config("c5", 96, 192, 1, 1, 2000),
}


def make_inputs(cfg):
    """(gr, c2, rho0, g0), q0, dt for a configuration."""
    gr, c2, rho0, g0 = background(cfg.background, cfg.nr)
    if cfg.ic == 0:
        q0, info = generate_ic(0, cfg.n, cfg.nr, gr, c2, cfg.seed, [1e-3, 0.05, 0.0])
    else:
        q0, info = generate_ic(1, cfg.n, cfg.nr, gr, c2, cfg.seed, [1e-3, 1e-6, 20.0])
    return (gr, c2, rho0, g0), q0, cfg.dt(gr, c2), info
