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


def ensemble_spec(count=48, seed=20240):
    """The batched workload: count independent Gaussian-pulse simulations on
    the uniform background, angular n from {32,48,64,96}, radial nr from
    {64,128,192}, pulse widths in [0.02,0.08] and source positions from the
    seed.  Returns (members, spec): members are the Ensemble() dicts, and
    spec[m] = (n, nr, width, ic_seed) gives member m's initial state through
    ensemble_member_ic."""
    import numpy as np
    rng = np.random.default_rng(seed)
    spec = [(int(rng.choice([32, 48, 64, 96])), int(rng.choice([64, 128, 192])), float(rng.uniform(0.02, 0.08)),
             int(rng.integers(1, 2**31))) for _ in range(count)]
    members = []
    for n, nr, width, ic_seed in spec:
        cfg = Config("ens", n, nr, 0, 0, 0)
        gr, c2, rho0, g0 = background(0, nr)
        members.append(dict(n=n, nr=nr, gr=gr, c2=c2, rho0=rho0, g0=g0, sponge_len=cfg.sponge_len,
                            a_sponge=cfg.a_sponge, dt=cfg.dt(gr, c2)))
    return members, spec


def ensemble_member_ic(member, spec_entry):
    n, nr, width, ic_seed = spec_entry
    q0, _ = generate_ic(0, n, nr, member["gr"], member["c2"], ic_seed, [1e-3, width, 0.0])
    return q0


def misfit_problem(cfg=None, nrec=16, blob_amp=0.02, blob_width=0.08):
    """The differentiation workload on a configuration (default c5,
    6x96x96x192): nrec p' receivers at the radial point closest to r = 0.98
    with seed-drawn panels and angular indices, and a "true" c^2 field carrying
    a Gaussian blob of relative amplitude blob_amp at a seed-drawn position
    (r = 0.85) from which the data are made.  Returns a dict with cfg, bg, q0,
    dt, rec (nrec x (panel, k, j, i)), blob (6, nr, n, n) and c2_true."""
    import numpy as np
    from .lib import panel_points
    cfg = cfg or CONFIGS["c5"]
    bg, q0, dt, _ = make_inputs(cfg)
    gr, c2 = bg[0], bg[1]
    n = cfg.n
    rng = np.random.default_rng(cfg.seed)
    kr = int(np.argmin(np.abs(gr - 0.98)))
    rec = np.array([[int(rng.integers(0, 6)), kr, int(rng.integers(0, n)), int(rng.integers(0, n))]
                    for _ in range(nrec)])
    xyz = panel_points(n)  # (6, n, n, 3) unit vectors
    centre = rng.standard_normal(3)
    centre /= np.linalg.norm(centre)
    pos = gr[None, :, None, None, None] * xyz[:, None, :, :, :]
    blob = np.exp(-np.sum((pos - 0.85 * centre) ** 2, axis=-1) / (2 * blob_width ** 2))
    c2_true = c2[None, :, None, None] * (1.0 + blob_amp * blob)
    return dict(cfg=cfg, bg=bg, q0=q0, dt=dt, rec=rec, blob=blob, c2_true=c2_true, receiver_radius=float(gr[kr]))
