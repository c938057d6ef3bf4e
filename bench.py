"""Benchmark: the 4096x2048 fp32 fluid step on one MI355X.

Workload (plain GPU drop-in): one 4096x2048 domain, walls + 400 seeded disks
scaled to 5% solid, 8-mode sinusoidal velocity, 1M tracers in the fluid;
each timestep is Simulation::step (diffusion, RK2 Catmull-Rom advection,
divergence, 2 MG V-cycles at 2 pre / 2 post sweeps, coarse solve of 5 sweeps
on the 16x8 level, velocity correction, BCs, *_current copy) followed by one
RK2 tracer step.  Inputs are resident in HBM before the timed region.

Prints ONE JSON line (rank 0): value = timesteps/s; roofline of the dominant
kernel (level-0 red/black smoother sweep, 13 algorithmic bytes per cell);
cpu_baseline = the CPU oracle (oracle/, OpenMP) on a bounded sample of the
same workload on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SMOOTH_BYTES_PER_CELL = 13  # p r/w 8 + f 4 + flag 1 per red+black sweep
STEP_BYTES_PER_CELL_22 = 278.6  # step at 2/2 smoothing, 2 V-cycles (algorithmic)
TRACER_BYTES = 24  # xy r/w 16 + age r/w 8 per point


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--width", type=int, default=4096)
    ap.add_argument("--height", type=int, default=2048)
    ap.add_argument("--tracers", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-graph", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build measures one GPU: the fluid step of one domain "
                          "runs on one MI355X; multi-GPU row-strip sharding is a design note only "
                          "(run with --gpus 1)"}))
        sys.exit(2)

    import numpy as np

    import ubootgl_amd as U
    from ubootgl_amd import workloads

    if U.device_count() < 1:
        raise SystemExit("bench.py needs a GPU: libubg.so sees no HIP device")

    t0 = time.time()
    if (args.width, args.height) == (4096, 2048):
        d = workloads.config1()
    else:
        g = workloads.rng(1)
        flag = workloads.disks_to_fraction(args.width, args.height, g, 400, 8.0, 128.0, 0.05)
        vx, vy = workloads.velocity(flag, g, 8, 6)
        d = workloads.Domain(flag, vx, vy)
    xy = workloads.tracers_in_fluid(d, args.tracers, seed=2) if args.tracers else None
    log(f"workload {d.W}x{d.H} solid={workloads.solid_fraction(d.flag.astype(np.float64)):.4f} "
        f"dt={float(d.dt):.3e} generated in {time.time() - t0:.1f}s")

    params = U.Params(nu_pre=2, nu_post=2, nu_coarse=5, vcycles_per_step=2,
                      use_graph=0 if args.no_graph else 1)
    vx, vy = d.init_simulation_fields()
    sim = U.Simulation(d.flag, d.pwidth, d.mu, params=params)
    for name, arr in (("vx", vx), ("vx_back", vx), ("vy", vy), ("vy_back", vy)):
        sim.set(name, arr)
    if xy is not None:
        sim.set_tracers(xy)
    dt = float(d.dt)

    sim.run(dt, args.warmup)
    sim.sync()
    t1 = time.perf_counter()
    sim.run(dt, args.steps)
    sim.sync()
    t2 = time.perf_counter()
    ms_per_step = (t2 - t1) * 1e3 / args.steps
    steps_per_s = args.steps / (t2 - t1)
    log(f"timed: {args.steps} steps in {(t2 - t1):.3f}s -> {steps_per_s:.1f} steps/s")

    # dominant kernel: HIP events on the simulation's stream around every
    # level-0 smoother launch, over an instrumented pass of the same steps
    ninstr = max(10, min(args.steps, 50))
    sim.timer_enable(True)
    sim.run(dt, ninstr)
    sim.sync()
    kms, klaunch = sim.timer_read()
    sim.timer_enable(False)
    cells = d.W * d.H
    kavg_ms = kms / max(klaunch, 1)
    achieved = SMOOTH_BYTES_PER_CELL * cells / (kavg_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "rbgs_l0_4096x2048_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    log(f"smoother: {klaunch} launches, avg {kavg_ms * 1e3:.2f} us -> {achieved:.0f} GB/s algorithmic")

    vxg = sim.get("vx")
    finite = bool(np.isfinite(vxg).all())

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(d, xy, args.cpu_steps)

    step_gb = (STEP_BYTES_PER_CELL_22 * cells + TRACER_BYTES * (args.tracers or 0)) / 1e9
    out = {
        "metric": "timesteps/s + HBM GB/s @4096x2048 fp32; batched solves/s & p99 @1e-6 residual",
        "value": round(steps_per_s, 3),
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"fluid step {d.W}x{d.H} fp32, 5% disks + walls, 8 modes, "
                        f"{args.tracers} tracers, MG 2 V-cycles/step at 2/2 smoothing, "
                        f"coarse 5 sweeps, {'graph' if not args.no_graph else 'eager'}",
            "width": d.W, "height": d.H, "tracers": args.tracers,
            "step_algorithmic_GB": round(step_gb, 4),
            "step_algorithmic_GBps": round(step_gb / (ms_per_step * 1e-3), 1),
            "fields_finite": finite,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "rbgs red+black sweep, level 0",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "avg_launch_us": round(kavg_ms * 1e3, 3),
            "launches": klaunch,
            "bytes_per_launch": SMOOTH_BYTES_PER_CELL * cells,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(d, xy, nsteps):
    """The CPU oracle (oracle/_build/libubg_oracle.so, OpenMP) on the same
    workload: `nsteps` steps after one warm-up step, with the tracers."""
    import numpy as np

    from oracle import pyoracle as O

    threads = int(os.environ.get("UBG_CPU_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    O.set_threads(threads)
    vx, vy = d.init_simulation_fields()
    ref = O.OracleSim(d.flag, float(d.pwidth), float(d.mu), nu_pre=2, nu_post=2, nu_coarse=5,
                      vcycles=2)
    for fid, arr in ((O.VX, vx), (O.VX_BACK, vx), (O.VY, vy), (O.VY_BACK, vy)):
        ref.set(fid, arr)
    pts = xy.copy() if xy is not None else np.zeros((0, 2), np.float32)
    age = np.zeros(pts.shape[0], np.float32)
    dt = float(d.dt)
    ref.step(dt)
    t = time.perf_counter()
    for k in range(nsteps):
        ref.step(dt)
        if pts.shape[0]:
            ref.advect_tracers(pts, age, dt, k + 1)
    el = time.perf_counter() - t
    log(f"cpu baseline: {nsteps} steps in {el:.2f}s on {threads} threads")
    return {"value": round(nsteps / el, 4), "unit": "timesteps/s", "cores": threads,
            "kind": "port",
            "sample": f"{nsteps} timesteps (+1 warm-up) of the same {d.W}x{d.H} workload "
                      f"with tracers, oracle/ OpenMP build, {el:.1f}s"}


if __name__ == "__main__":
    main()
