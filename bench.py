"""Benchmark of the acoustic RK4 hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2|c1|c5]
                    [--cpu-sample-steps S] [--no-cpu]
    python bench.py --only copy     # plain device copy, for counter calibration

A step is one RK4 call (timestep.f90:41-86): five fused stages plus the
explicit filter over the whole six-panel grid, state resident in HBM.
value = 6*n*n*nr*K / (wall seconds of K steps), grid-point RK4 steps/s.

The headline workload is the 6x256x256x384 FP64 grid (c3); the line also
carries the 6x128x128x256 grid (c2) as a nested object, and (fp32_state) the
same two grids with an FP32 state: step, stage and filter times and the step
speed-up over the FP64 state.

Printed as one JSON line, with
  roofline     : the dominant kernel (the fused stage kernel), algorithmic
                 bytes per launch / its mean launch time (HIP events on the
                 library's stream), against the 8 TB/s HBM peak; plus the whole
                 step against 800 B per point per step (step_frac, the 2N
                 scheme storing every stage) and against the bytes the step
                 moves (step_frac_moved: 720 B in FP64)
  cpu_baseline : the CPU restatement of the reference (oracle/) on this host,
                 on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-point RK4 steps/s (FP64, % HBM roofline) at 1 GPU; fwd+adjoint vs plain fwd cost"
HBM_PEAK = 8.0e12
# The roofline basis of the metric: the 2N step storing q and qw after every
# stage (FP64 state of 5 fields = 40 B per point):
#   stage 1 and 5: read q + write q and qw / read q and qw + write q -> 120 B
#   stages 2-4: read q and qw, write q and qw -> 160 B;  filter: 80 B
# (half for a 32-bit state).  The library's FP64 step moves less (it stores q
# after stages 2, 4 and 5 only: 80, 160, 120, 160, 120 + 80 = 720 B); the
# dominant kernel's roofline uses the bytes its launches really move, which
# the handle reports (Solver.algorithmic_bytes).
STEP_BYTES = 800

DESCR = {
    "c1": "c1: Gaussian pulse, uniform background, 6x32x32x64, FP64",
    "c2": "c2: stratified background, spherical-harmonic mode + noise, sponge, 6x128x128x256, FP64",
    "c3": "c3: stratified background, spherical-harmonic mode + noise, sponge, 6x256x256x384, FP64",
    "c5": "c5: stratified background, spherical-harmonic mode + noise, sponge, 6x96x96x192, FP64",
}


def _cpu_threads():
    """OpenMP threads of the CPU sample: OMP_NUM_THREADS when the host sets it
    (the process's CPU share on a shared machine), else every CPU this process
    may run on."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else avail
    os.environ["OMP_NUM_THREADS"] = str(threads)
    return threads, avail, os.cpu_count() or 1


def _cpu_baseline(cfg, bg, q0, dt, sample_steps):
    threads, avail, nproc = _cpu_threads()
    import oracle
    o = oracle.Oracle(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge)
    o.rk4(q0, 1, dt)  # first touch / warm-up
    t = time.perf_counter()
    o.rk4(q0, sample_steps, dt)
    el = time.perf_counter() - t
    return {"value": cfg.points * sample_steps / el, "unit": "grid-point RK4 steps/s", "cores": threads,
            "kind": "port", "nproc": nproc, "affinity_cpus": avail,
            "sample": f"{sample_steps} RK4 step(s) of the {cfg.name} workload ({cfg.points} points) "
                      f"after one warm-up step, C restatement of the reference with OpenMP on {threads} "
                      f"threads (host: {nproc} CPUs, {avail} in this process's affinity), {el:.1f} s"}


def _copy_calibration(reps=4, nbytes=1 << 30):
    """A plain device copy of a known byte count (torch's copy of a 1 GiB FP64
    buffer, reps times): the transfer against which the FETCH_SIZE and
    WRITE_SIZE counters are calibrated (each copy reads and writes nbytes)."""
    import torch
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    return {"copy": {"reps": reps, "bytes_read_per_copy": nbytes, "bytes_written_per_copy": nbytes,
                     "GB_per_s": 2 * nbytes * reps / el / 1e9}}


def _adjoint_bench(fd_check, steps=None, checkpoint_every=0):
    """Sensitivity workload: misfit of 16 p' receivers at r = 0.98 (seed-drawn
    angular positions) over the c5 run (6x96x96x192, 2000 steps), data from a
    model whose c^2 carries a 2% Gaussian blob at a seed-drawn position;
    gradient dJ/dc^2 by the checkpointed discrete adjoint.  Reports the wall
    time of misfit_gradient over that of the plain forward run: with the
    stage divergences kept in FP64 (the exact discrete gradient) with
    memory-sized segments, with checkpoints every 50 steps (with and without
    the divergence cache of trailing segments), and with FP32 divergences;
    each gradient is checked by a central difference of J.  The headline
    ratio is checkpoints every 50 steps, FP64, with the library's default
    cache."""
    import numpy as np
    from solar_amd import Solver, misfit_problem
    prob = misfit_problem()
    cfg, bg, q0, dt, rec, blob, c2_true = (prob[k] for k in ("cfg", "bg", "q0", "dt", "rec", "blob", "c2_true"))
    c2 = bg[1]
    n, nr = cfg.n, cfg.nr
    steps = steps or cfg.steps
    out = {"workload": "c5: 6x96x96x192, stratified, 2000 RK4 steps, FP64", "steps": steps,
           "receivers": int(rec.shape[0]), "receiver_radius": prob["receiver_radius"]}
    with Solver(n, nr, *bg, cfg.sponge_len, cfg.a_sponge) as s:
        s.set_receivers(rec)
        s.set_c2_field(c2_true)
        s.set_state(q0)
        data = s.record(steps, dt)
        s.set_c2_field(None)
        s.set_state(q0)
        s.rk4(5, dt)
        s.set_state(q0)
        s.sync()
        t = time.perf_counter()
        s.rk4(steps, dt)
        fwd = time.perf_counter() - t
        out["fwd_s"] = fwd
        dirn = c2[None, :, None, None] * blob
        # (name, divergence bits, checkpoint spacing, divergence cache)
        runs = [("fp64_div", 64, checkpoint_every, True), ("fp64_div_ckpt50", 64, 50, True),
                ("fp64_div_ckpt50_nocache", 64, 50, False), ("fp32_div", 32, checkpoint_every, True)]
        grads = {}
        for name, bits, ck, cache in runs:
            s.set_adjoint_options(bits)
            s.set_adjoint_cache(cache)
            t = time.perf_counter()
            J, grad, seis = s.misfit_gradient(q0, steps, dt, data, ck)
            tot = time.perf_counter() - t
            out[name] = {"fwd_adj_s": tot, "ratio": tot / fwd, "segment_steps": s.segment_used,
                         "recomputed_steps": s.recomputed_steps, "divergence_cache": cache,
                         "grad_finite": bool(np.isfinite(grad).all()),
                         "phases_s": {k: round(v, 4) for k, v in s.adjoint_times().items()}}
            grads[name] = float(np.vdot(grad, dirn))
        s.set_adjoint_options(64)
        s.set_adjoint_cache(True)
        out["J"] = J
        # the configuration as stated: FP64 divergences, checkpoints every 50 steps (the library's
        # default divergence cache keeps trailing segments' divergences in free memory)
        out["ratio"] = out["fp64_div_ckpt50"]["ratio"]
        out["ratio_ckpt50_nocache"] = out["fp64_div_ckpt50_nocache"]["ratio"]
        out["ratio_memory_segments"] = out["fp64_div"]["ratio"]  # FP64, segments sized to free memory
        out["ratio_fp32_div"] = out["fp32_div"]["ratio"]  # stage divergences stored in FP32
        if fd_check:
            # central difference of J along the blob direction (step small enough that the
            # O(eps^2) truncation of the difference stays far below the 1e-6 bar)
            eps = 1e-5

            def J_at(field):
                s.set_c2_field(field)
                s.set_state(q0)
                r = s.record(steps, dt) - data
                return float(np.sum(r * r))
            base = np.broadcast_to(c2[None, :, None, None], dirn.shape)
            fd = (J_at(base + eps * dirn) - J_at(base - eps * dirn)) / (2 * eps)
            out["fd_dJ"] = fd
            for name in grads:
                out[name]["adj_dJ"] = grads[name]
                out[name]["grad_rel_err"] = abs(fd - grads[name]) / abs(fd)
    return out


def _workload_descr(name, bits):
    d = DESCR[name]
    return d if bits == 64 else d.replace("FP64", "FP32 state, FP64 arithmetic")


def _grid_bench(name, steps, cpu_steps, bits=64):
    """A second grid (by default the 6x128x128x256 one) in the same line: GPU
    rate, stage-kernel roofline, and a CPU sample of cpu_steps steps of the
    same grid (after a warm-up step); cpu_steps = 0 skips the CPU sample."""
    from solar_amd import CONFIGS, Solver, make_inputs
    cfg = CONFIGS[name]
    bg, q0, dt, _ = make_inputs(cfg)
    with Solver(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge, state_bits=bits) as s:
        s.set_state(q0)
        s.rk4(5, dt)
        s.sync()
        t = time.perf_counter()
        s.rk4_async(steps, dt)
        s.sync()
        wall = time.perf_counter() - t
        ok = s.check_finite()
    os.environ["SOLAR_PROFILE"] = "1"
    p = Solver(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge, state_bits=bits)
    os.environ.pop("SOLAR_PROFILE")
    p.set_state(q0)
    p.rk4(2, dt)
    p.rk4(20, dt)
    kt = p.kernel_times()
    ab = p.algorithmic_bytes()
    p.close()
    rate = cfg.points * steps / wall
    scale = bits / 64.0
    stage_bytes = sum(ab[:5]) / 5.0 * cfg.points
    out = {"workload": _workload_descr(name, bits), "steps": steps, "value": rate, "unit": "grid-point RK4 steps/s",
           "ms_per_step": wall / steps * 1e3, "step_frac": rate * STEP_BYTES * scale / HBM_PEAK,
           "step_bytes_per_point": sum(ab), "step_frac_moved": rate * sum(ab) / HBM_PEAK,
           "stage_frac": stage_bytes / (kt["stage_ms"] * 1e-3) / HBM_PEAK, "state_finite": ok,
           "kernel_ms_all": {k: round(v, 5) for k, v in kt.items()}}
    if cpu_steps:
        c = _cpu_baseline(cfg, bg, q0, dt, cpu_steps)
        out["cpu_baseline"] = c
        out["gpu_over_cpu"] = rate / c["value"]
    return out


def _ensemble_bench(steps, reps):
    """Batched workload: 48 independent Gaussian-pulse simulations, uniform
    background, angular n and radial nr drawn from {32,48,64,96} x {64,128,192},
    pulse widths in [0.02, 0.08] and source positions from the seed; FP64;
    every step one hipGraph replay with one launch per kernel class and tile
    shape across all members, tiles chosen by the autotuner."""
    import numpy as np
    from solar_amd import Ensemble, ensemble_member_ic, ensemble_spec
    members, spec = ensemble_spec(48, 20240)
    with Ensemble(members) as ens:
        t = time.perf_counter()
        ens.autotune(reps)
        tune_s = time.perf_counter() - t
        for k in range(len(members)):
            ens.set_state(k, ensemble_member_ic(members[k], spec[k]))
        ens.rk4(3)
        ens.sync()
        t = time.perf_counter()
        ens.rk4_async(steps)
        ens.sync()
        wall = time.perf_counter() - t
        info = ens.info()
        small = min(range(48), key=lambda k: spec[k][0] ** 2 * spec[k][1])
        finite = bool(np.isfinite(ens.get_state(small)).all())
        tiles = sorted({(spec[k][0], spec[k][1]) + ens.tiles(k) for k in range(48)})
    rate = info["points"] * steps / wall
    return {"workload": "48 Gaussian-pulse simulations, n in {32,48,64,96}, nr in {64,128,192}, FP64",
            "members": 48, "points": info["points"], "steps": steps, "ms_per_step": wall / steps * 1e3,
            "value": rate, "unit": "grid-point RK4 steps/s", "step_frac": rate * STEP_BYTES / HBM_PEAK,
            "launches_per_step": info["launches_per_step"], "autotune_s": tune_s, "state_finite": finite,
            "tiles_by_class": [{"n": c[0], "nr": c[1], "stage": [c[2], c[3], c[4]], "filter": [c[5], c[6], c[7]]}
                               for c in tiles]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--cpu-sample-steps", type=int, default=None,
                    help="CPU sample steps of the headline grid (default: 1 for c3, 4 otherwise)")
    ap.add_argument("--second", default="c2",
                    help="second grid reported in a nested object ('none' to skip)")
    ap.add_argument("--no-c3", action="store_true", help="same as --second none")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--state-bits", type=int, default=64, choices=(64, 32),
                    help="storage of the state and 2N accumulator (arithmetic is FP64 either way)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the fwd+adjoint cost measurement")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-state measurement of the FP64 line")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--no-adjoint-fd", action="store_true", help="skip its finite-difference gradient check")
    ap.add_argument("--adjoint-steps", type=int, default=None, help="steps of the adjoint workload (default 2000)")
    ap.add_argument("--checkpoint-every", type=int, default=0,
                    help="segment length of the adjoint run (0: the longest that fits in device memory)")
    ap.add_argument("--only", choices=("adjoint", "ensemble", "copy"), default=None,
                    help="run one extra section alone (for profiling); prints its JSON object")
    ap.add_argument("--precision-study", type=int, default=0, metavar="STEPS",
                    help="also run STEPS steps in FP64 and FP32 storage and report the relative L2 error "
                         "of each field of the FP32 run against the FP64 run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this library runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding the panels across GPUs is a design only; no code runs it)", file=sys.stderr)
        sys.exit(2)

    if args.only == "adjoint":
        print(json.dumps({"adjoint": _adjoint_bench(not args.no_adjoint_fd, args.adjoint_steps,
                                                    args.checkpoint_every)}), flush=True)
        return
    if args.only == "ensemble":
        print(json.dumps({"ensemble": _ensemble_bench(args.ensemble_steps, 10)}), flush=True)
        return
    if args.only == "copy":
        print(json.dumps(_copy_calibration()), flush=True)
        return

    from solar_amd import CONFIGS, Solver, make_inputs
    cfg = CONFIGS[args.workload]
    # default: the configuration's step count, at most 400 steps of timed work
    steps = args.steps if args.steps is not None else min(cfg.steps, 400)
    bg, q0, dt, info = make_inputs(cfg)

    bits = args.state_bits
    s = Solver(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge, state_bits=bits)
    s.set_state(q0)
    s.rk4(args.warmup, dt)
    s.sync()
    t0 = time.perf_counter()
    s.rk4_async(steps, dt)
    s.sync()
    wall = time.perf_counter() - t0
    ok = s.check_finite()
    s.close()
    rate = cfg.points * steps / wall

    # Kernel times with HIP events bracketing every launch on the library's
    # stream (a separate handle: SOLAR_PROFILE is read at init).
    os.environ["SOLAR_PROFILE"] = "1"
    p = Solver(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge, state_bits=bits)
    os.environ.pop("SOLAR_PROFILE")
    p.set_state(q0)
    p.rk4(2, dt)
    p.rk4(args.profile_steps, dt)
    kt = p.kernel_times()
    ab = p.algorithmic_bytes()  # per point: five stages, filter
    p.close()
    scale = bits / 64.0  # FP32 storage halves the roofline basis
    stage_bytes = sum(ab[:5]) / 5.0 * cfg.points
    stage_s = kt["stage_ms"] * 1e-3
    achieved = stage_bytes / stage_s
    traffic = None
    # counter-measured HBM bytes of the stage kernel, for this workload and storage width
    suffix = "" if bits == 64 else "_fp32state"
    tpath = os.path.join(ROOT, "profiles", f"traffic_stage_{cfg.name}{suffix}.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": traffic,
                "kernel": "k_stage (fused RK stage)", "kernel_ms": kt["stage_ms"],
                "algorithmic_bytes_per_launch": stage_bytes,
                "stage_bytes_per_point": list(ab[:5]),
                "step_frac": rate * STEP_BYTES * scale / HBM_PEAK,
                "step_bytes_per_point": sum(ab),
                "step_frac_moved": rate * sum(ab) / HBM_PEAK,
                "kernel_ms_all": {k: round(v, 5) for k, v in kt.items()}}

    precision = None
    if args.precision_study:
        import numpy as np
        fin = {}
        for b in (64, 32):
            with Solver(cfg.n, cfg.nr, *bg, cfg.sponge_len, cfg.a_sponge, state_bits=b) as x:
                x.set_state(q0)
                x.rk4(args.precision_study, dt)
                fin[b] = x.get_state()
        errs = [float(np.linalg.norm(fin[32][:, f] - fin[64][:, f]) / np.linalg.norm(fin[64][:, f]))
                for f in range(5)]
        precision = {"steps": args.precision_study, "rel_l2_fp32_vs_fp64": dict(zip(
            ["rho", "v_xi", "v_eta", "v_r", "p"], errs))}
        del fin

    second = None
    if not args.no_c3 and args.second not in ("none", args.workload):
        second = _grid_bench(args.second, 200, 0 if args.no_cpu or bits != 64 else 1, bits)

    fp32 = None
    if not args.no_fp32 and bits == 64:
        fp32 = {cfg.name: _grid_bench(cfg.name, 200, 0, 32)}
        fp32[cfg.name]["over_fp64"] = (wall / steps * 1e3) / fp32[cfg.name]["ms_per_step"]
        if second:
            fp32[args.second] = _grid_bench(args.second, 200, 0, 32)
            fp32[args.second]["over_fp64"] = second["ms_per_step"] / fp32[args.second]["ms_per_step"]

    ensemble = None
    if not args.no_ensemble and bits == 64:
        ensemble = _ensemble_bench(args.ensemble_steps, 10)

    adjoint = None
    if not args.no_adjoint and bits == 64:
        adjoint = _adjoint_bench(not args.no_adjoint_fd, args.adjoint_steps, args.checkpoint_every)

    cpu = None
    if not args.no_cpu:
        cpu_steps = args.cpu_sample_steps or (1 if cfg.points > 100_000_000 else 4)
        cpu = _cpu_baseline(cfg, bg, q0, dt, cpu_steps)

    out = {"metric": METRIC, "value": rate, "unit": "grid-point RK4 steps/s", "n_gpus": 1, "steps": steps,
           "warmup": args.warmup, "ms_per_step": wall / steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64" if bits == 64 else "f32-state/f64-arith", "data": "synthetic",
           "config": {"workload": _workload_descr(cfg.name, bits), "n": cfg.n, "nr": cfg.nr, "panels": 6, "points": cfg.points,
                      "dt": dt, "seed": cfg.seed, "state_finite": ok},
           "roofline": roofline, "cpu_baseline": cpu}
    if precision:
        out["precision"] = precision
    if adjoint:
        out["adjoint"] = adjoint
    if ensemble:
        out["ensemble"] = ensemble
    if fp32:
        out["fp32_state"] = fp32
    if second:
        out[args.second] = second
    if cpu:
        out["gpu_over_cpu"] = rate / cpu["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
