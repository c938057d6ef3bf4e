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
SMOOTH_BYTES_PER_CELL = 13  # one red+black sweep: p r/w 8 + f 4 + flag 1
RR_BYTES_PER_CELL = 10      # residual + restriction (p 4, f 4, flag 1, coarse rhs 1)
PC_BYTES_PER_CELL = 10.25   # prolongation + correction (coarse 1, coarse flag 0.25, flag 1, p r/w 8)
# compulsory HBM bytes per cell of the two fused level-0 passes
PRE_PASS_BYTES = 15.0       # p r 4 + f 4 + flag 1 + p w 4 + coarse rhs w 1 + coarse p w 1
POST_PASS_BYTES = 14.25     # p r 4 + f 4 + flag 1 + coarse p 1 + coarse flag 0.25 + p w 4
STEP_BYTES_PER_CELL_22 = 278.6  # step at 2/2 smoothing, 2 V-cycles (algorithmic)
TRACER_BYTES = 24  # xy r/w 16 + age r/w 8 per point
ADVECT_BYTES_PER_CELL = 17  # vx, vy r 8 + flag 1 + vx, vy w 8
# VALU issue peak in wave64 instructions/s from MI355X's published fp32
# vector peak: 157.3 TFLOP/s / (64 lanes x 2 flops of an FMA)
VALU_PEAK_GINST = 157.3e3 / 128
KERNEL_NAMES = {"sweep": "level-0 rbgs sweep (red + black, one launch)",
                "pre": "level-0 fused pass: rbgs sweep + residual + restriction",
                "post": "level-0 fused pass: prolongation + correction + rbgs sweep",
                "advect": "advect: RK2 backtrace + 3 Catmull-Rom samples per face (vx and vy)"}


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
    ap.add_argument("--cpu-steps", type=int, default=40,
                    help="timesteps per CPU-baseline run (3 runs, the median reported)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-hbm-sweep", action="store_true",
                    help="skip the level-0 sweep timed on 8192x4096 (working set > 256 MB)")
    ap.add_argument("--copies", type=int, default=0,
                    help="step: after the run, N saveCurrentVelocityFields stages (plain "
                         "33.5 MB device copies; counter calibration)")
    ap.add_argument("--workload", choices=["step", "ir", "batch", "serve", "sims"],
                    default="step",
                    help="step: the 4096x2048 fluid step (default); ir: the 8192x4096 "
                         "FP64-refined solve to 1e-10; batch: 256 uneven domains to 1e-6; "
                         "serve: Poisson arrivals at --rate of 3-timestep solve requests; "
                         "sims: --sims simulations stepped with their solves batched")
    ap.add_argument("--sims", type=int, default=64)
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--duration", type=float, default=1.0)
    ap.add_argument("--templates", type=int, default=48)
    ap.add_argument("--steps-per-turn", type=int, default=1,
                    help="serve: refinement steps the worker enqueues per turn")
    ap.add_argument("--seq-requests", type=int, default=200,
                    help="serve: requests of the stream timed as sequential solves")
    ap.add_argument("--domains", type=int, default=256)
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
    if args.workload == "ir":
        return bench_ir(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "sims":
        return bench_sims(args)

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

    # dominant kernels: HIP events on the simulation's stream around every
    # level-0 launch of the three multigrid kernels (plain rbgs sweep; fused
    # sweep+residual+restriction; fused prolongation+correction+sweep) over an
    # instrumented pass of the same steps
    from ubootgl_amd import _lib as LL
    ninstr = max(10, min(args.steps, 50))
    timers = ((LL.TIMER_SMOOTH_L0, "sweep", SMOOTH_BYTES_PER_CELL, SMOOTH_BYTES_PER_CELL),
              (LL.TIMER_L0_PRE, "pre", PRE_PASS_BYTES, SMOOTH_BYTES_PER_CELL + RR_BYTES_PER_CELL),
              (LL.TIMER_L0_POST, "post", POST_PASS_BYTES, PC_BYTES_PER_CELL + SMOOTH_BYTES_PER_CELL),
              (LL.TIMER_ADVECT, "advect", ADVECT_BYTES_PER_CELL, ADVECT_BYTES_PER_CELL))
    for t, *_ in timers:
        sim.timer_enable(t, True)
    sim.run(dt, ninstr)
    sim.sync()
    cells = d.W * d.H
    kern = {}
    for t, name, pass_b, alg_b in timers:
        ms, n = sim.timer_read(t)
        sim.timer_enable(t, False)
        if n == 0:
            continue
        avg = ms / n
        kern[name] = dict(total_ms=ms, launches=n, avg_us=avg * 1e3,
                          pass_GBps=pass_b * cells / (avg * 1e-3) / 1e9,
                          unfused_equiv_GBps=alg_b * cells / (avg * 1e-3) / 1e9,
                          pass_bytes=pass_b * cells)
        log(f"{name}: {n} launches, avg {avg * 1e3:.2f} us -> "
            f"{kern[name]['pass_GBps']:.0f} GB/s compulsory, "
            f"{kern[name]['unfused_equiv_GBps']:.0f} GB/s unfused-equivalent")
    # the dominant kernel: largest total time of the timed kernels
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    K = kern[dom]
    achieved = K["pass_GBps"]
    # HBM bytes (and VALU instructions) per launch of the same kernels from
    # rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE, SQ_INSTS_VALU), the byte
    # counters scaled by what they read for plain 33.5 MB device copies
    prof = {}
    tpath = os.path.join(ROOT, "profiles", "step_4096x2048_traffic.json")
    if (d.W, d.H) == (4096, 2048) and os.path.exists(tpath):
        with open(tpath) as fh:
            prof = json.load(fh)
    traffic = prof.get(dom, {}).get("bytes_per_launch")
    # these counter figures come from the committed profile, not from this run
    prof_source = None if not prof else (
        f"not measured in this run: profiles/step_4096x2048_traffic.json, rocprofv3 --pmc "
        f"FETCH_SIZE / WRITE_SIZE of the same workload at commit {prof.get('commit', '?')}, "
        f"scaled by the counters' reading of plain device copies")
    valu = None
    vi = prof.get(dom, {}).get("valu_wave_instructions_per_launch")
    if vi:
        rate = vi / (K["avg_us"] * 1e-6) / 1e9
        valu = {"wave_instructions_per_launch": vi, "achieved": round(rate, 1),
                "peak": VALU_PEAK_GINST, "unit": "G wave-instructions/s",
                "frac": round(rate / VALU_PEAK_GINST, 4),
                "source": "instructions per launch from the same profile (SQ_INSTS_VALU), "
                          "divided by this run's launch time"}
    sw = kern.get("sweep")
    smoother = None if sw is None else {
        "kernel": KERNEL_NAMES["sweep"], "achieved": round(sw["pass_GBps"], 1),
        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(sw["pass_GBps"] / HBM_PEAK_GBS, 4),
        "avg_launch_us": round(sw["avg_us"], 3),
        "traffic": prof.get("sweep", {}).get("bytes_per_launch")}

    for _ in range(args.copies):
        sim.stage(LL.STAGE_SAVE_CURRENT, dt)
    sim.sync()
    vxg = sim.get("vx")
    finite = bool(np.isfinite(vxg).all())
    sim.close()
    smoother_big = None if args.no_hbm_sweep else smoother_out_of_cache()

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
            "kernel": KERNEL_NAMES[dom],
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": prof_source,
            "avg_launch_us": round(K["avg_us"], 3),
            "launches": K["launches"],
            "bytes_per_launch": K["pass_bytes"],
            "unfused_equiv_GBps": round(K["unfused_equiv_GBps"], 1),
            "valu": valu,
            "step_achieved": round(step_gb / (ms_per_step * 1e-3), 1),
            "step_frac": round(step_gb / (ms_per_step * 1e-3) / HBM_PEAK_GBS, 4),
            "smoother": smoother,
            "smoother_8192x4096": smoother_big,
            "other_kernels_us": {k: round(v["avg_us"], 3) for k, v in kern.items() if k != dom},
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def smoother_out_of_cache(nsweeps=40):
    """The level-0 red/black sweep on an 8192x4096 grid: p, p_alt and f
    (134 MB each) plus the mask byte make a 436 MB working set, above the
    256 MB Infinity Cache, so the achieved rate is HBM's.  Timed with the
    dispatch-attached events over `nsweeps` back-to-back sweeps."""
    import numpy as np

    import ubootgl_amd as U
    from ubootgl_amd import _lib as LL

    W, H = 8192, 4096
    g = np.random.default_rng(5)
    flag = np.ones((H, W), np.float32)
    flag[0, :] = flag[-1, :] = flag[:, 0] = flag[:, -1] = 0
    for _ in range(200):  # disks, ~3% solid
        cx, cy, r = g.uniform(64, W - 64), g.uniform(64, H - 64), g.uniform(8, 48)
        y0, y1, x0, x1 = int(cy - r), int(cy + r) + 1, int(cx - r), int(cx + r) + 1
        yy, xx = np.mgrid[y0:y1, x0:x1]
        flag[y0:y1, x0:x1][(xx + 0.5 - cx) ** 2 + (yy + 0.5 - cy) ** 2 <= r * r] = 0
    sim = U.Simulation(flag, (W - 1) / W, 0.0)
    sim.mg.set_level(0, LL.LEVEL_P, (g.standard_normal((H, W)) * flag).astype(np.float32))
    sim.mg.set_level(0, LL.LEVEL_F, g.standard_normal((H, W)).astype(np.float32))
    h = float(np.float32(1.0 / W))
    sim.mg.rbgs(h, 1.0, 4)  # warm-up
    sim.sync()
    sim.timer_enable(LL.TIMER_SMOOTH_L0, True)
    t = time.perf_counter()
    sim.mg.rbgs(h, 1.0, nsweeps)
    sim.sync()
    wall = time.perf_counter() - t
    ms, n = sim.timer_read(LL.TIMER_SMOOTH_L0)
    sim.timer_enable(LL.TIMER_SMOOTH_L0, False)
    sim.close()
    avg = ms / max(n, 1)
    gbs = SMOOTH_BYTES_PER_CELL * W * H / (avg * 1e-3) / 1e9
    log(f"sweep 8192x4096: {n} launches, avg {avg * 1e3:.2f} us -> {gbs:.0f} GB/s "
        f"({wall * 1e6 / nsweeps:.2f} us per sweep by wall clock)")
    return {"kernel": KERNEL_NAMES["sweep"], "grid": f"{W}x{H}",
            "working_set_MB": round(13 * W * H / 1e6, 1),  # p, p_alt, f + mask
            "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "avg_launch_us": round(avg * 1e3, 3),
            "launches": n, "wall_us_per_sweep": round(wall * 1e6 / nsweeps, 3)}


def bench_ir(args):
    """8192x4096 Poisson solve (800 disks, 3% solid, rhs = -div of a 16-mode
    field) by FP64 iterative refinement around FP32 V-cycles (3/3 sweeps,
    coarse 5, 4 V-cycles per refinement step) to 1e-10 relative residual,
    from a zero guess each time.  value = solves/s."""
    import ubootgl_amd as U
    from ubootgl_amd import workloads

    t0 = time.time()
    P = workloads.config3()
    log(f"problem {P.W}x{P.H} solid={workloads.solid_fraction(P.flag.astype('f8')):.4f} "
        f"generated in {time.time() - t0:.1f}s")
    mg = U.MG(P.flag)
    mg.ir_set_rhs(P.f)
    steps, warm = max(1, min(args.steps, 10)), max(1, min(args.warmup, 2))
    res = None
    for _ in range(warm):
        mg.ir_set_guess(None)
        res = mg.ir_solve(float(P.h), tol=1e-10, vcycles_per_iter=4)
    mg.sync()
    t1 = time.perf_counter()
    for _ in range(steps):
        mg.ir_set_guess(None)
        res = mg.ir_solve(float(P.h), tol=1e-10, vcycles_per_iter=4)
    mg.sync()
    el = time.perf_counter() - t1
    log(f"timed: {steps} solves in {el:.3f}s; {res['iterations']} refinement steps, "
        f"{res['vcycles']} V-cycles, rel residual {res['rel_residual']:.2e}")
    cells = P.cells
    # algorithmic bytes: per V-cycle (3/3, all levels) 131 B/cell, per
    # refinement step the fused update pass 33 B/cell
    alg = (131.0 * res["vcycles"] + 33.0 * (res["iterations"] + 1)) * cells
    out = {
        "metric": "mixed-precision solves/s @8192x4096 to 1e-10 (FP64 refinement, FP32 V-cycles)",
        "value": round(steps / el, 4), "unit": "solves/s", "n_gpus": 1, "steps": steps,
        "warmup": warm, "ms_per_step": round(el * 1e3 / steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "poisson 8192x4096, 800 disks -> 3% solid, rhs -div(16 modes), "
                               "Dirichlet ring, tol 1e-10, 4 FP32 V-cycles (3/3/5) per step",
                   "iterations": res["iterations"], "vcycles": res["vcycles"],
                   "rel_residual": res["rel_residual"],
                   "algorithmic_GBps": round(alg / (el / steps) / 1e9, 1)},
    }
    print(json.dumps(out), flush=True)


def bench_batch(args):
    """256 independent solve domains of uneven size (W log-uniform in
    128..2048, multiples of 32, aspect 1..4, 0-8% disks, own 8-mode field),
    each to 1e-6 relative residual by FP64 refinement around FP32 V-cycles
    (3/3/5, 4 per step): one batched solve (autotuned, graph-replayed) against
    the same solves run one domain after another through MG.ir_solve.
    value = batched solves/s."""
    import ubootgl_amd as U
    from ubootgl_amd import batch as B
    from ubootgl_amd import workloads

    t0 = time.time()
    probs = workloads.config2(args.domains)
    cells = sum(P.cells for P in probs)
    log(f"{len(probs)} domains, {cells / 1e6:.1f} M cells, generated in {time.time() - t0:.1f}s")
    bt = B.Batch([(P.W, P.H) for P in probs])
    for d, P in enumerate(probs):
        bt.set_problem(d, P)
    tune = bt.autotune()
    log(f"autotune: {tune}")
    steps, warm = max(1, min(args.steps, 5)), max(1, min(args.warmup, 2))
    el = 0.0
    for it in range(warm + steps):
        for d in range(len(probs)):
            bt.set_guess(d)
        t1 = time.perf_counter()
        reps = bt.solve(tol=1e-6, vcycles_per_iter=4)
        t2 = time.perf_counter()
        if it >= warm:
            el += t2 - t1
    res = [bt.result(d) for d in range(len(probs))]
    batched = len(probs) * steps / el
    log(f"batched: {steps} x {len(probs)} solves in {el:.3f}s -> {batched:.1f} solves/s, "
        f"{reps} graph replays, iterations {min(r['iterations'] for r in res)}.."
        f"{max(r['iterations'] for r in res)}")
    # sequential plain-GPU solves of the same problems
    mgs = []
    for P in probs:
        mg = U.MG(P.flag)
        mg.ir_set_rhs(P.f)
        mgs.append(mg)
    for mg, P in zip(mgs[:8], probs[:8]):  # warm-up
        mg.ir_set_guess(None)
        mg.ir_solve(float(P.h), tol=1e-6, vcycles_per_iter=4)
    seq = 0.0
    for mg, P in zip(mgs, probs):
        mg.ir_set_guess(None)
        t1 = time.perf_counter()
        mg.ir_solve(float(P.h), tol=1e-6, vcycles_per_iter=4)
        seq += time.perf_counter() - t1
    sequential = len(probs) / seq
    log(f"sequential: {len(probs)} solves in {seq:.3f}s -> {sequential:.1f} solves/s")
    alg = sum((131.0 * r["vcycles"] + 33.0 * (r["iterations"] + 1)) * P.cells
              for r, P in zip(res, probs))
    out = {
        "metric": "batched solves/s @1e-6 residual, 256 uneven domains",
        "value": round(batched, 3), "unit": "solves/s", "n_gpus": 1, "steps": steps,
        "warmup": warm, "ms_per_step": round(el * 1e3 / steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": f"{len(probs)} poisson domains W 128..2048 (log-uniform, x32), "
                               f"aspect 1..4, 0-8% disks, tol 1e-6, 4 FP32 V-cycles (3/3/5) "
                               f"per FP64 refinement step",
                   "cells": cells, "tuning": tune,
                   "sequential_solves_per_s": round(sequential, 3),
                   "speedup_vs_sequential": round(batched / sequential, 2),
                   "algorithmic_GBps": round(alg / (el / steps) / 1e9, 1)},
    }
    print(json.dumps(out), flush=True)


def bench_sims(args):
    """--sims independent simulations (W log-uniform in 128..1024, multiples
    of 32, aspect 1..4, walls + disks, own 4-mode field; MG 2/2/5, 2 V-cycles
    per step as the 4096x2048 workload): timesteps of all of them with their
    pressure solves batched (one launch per multigrid stage for all, the rest
    of each step on its own stream) against the same simulations stepped one
    after another (each a graph replay).  value = simulation timesteps/s."""
    import numpy as np

    import ubootgl_amd as U
    from ubootgl_amd import batch as B
    from ubootgl_amd import workloads

    sizes = workloads.batch_sizes(args.sims, seed=11, wmin=128, wmax=1024)
    P = U.Params(nu_pre=2, nu_post=2, nu_coarse=5, vcycles_per_step=2)
    pairs, dts = [], []
    for i, (w, h) in enumerate(sizes):
        d = workloads.small_domain(w, h, seed=500 + i, ndisks=max(1, w * h // 8192))
        vx, vy = d.init_simulation_fields()
        two = []
        for _ in range(2):
            s = U.Simulation(d.flag, d.pwidth, d.mu, params=P)
            for name, a in (("vx", vx), ("vx_back", vx), ("vy", vy), ("vy_back", vy)):
                s.set(name, a)
            two.append(s)
        pairs.append(two)
        dts.append(float(d.dt))
    cells = sum(w * h for w, h in sizes)
    alone = [a for a, _ in pairs]
    sb = B.SimBatch([b for _, b in pairs])
    steps, warm = max(1, args.steps // 5), max(1, args.warmup)

    def run(k, batched):
        for _ in range(k):
            if batched:
                sb.step(dts)
            else:
                for s, dt in zip(alone, dts):
                    s.step(dt)
        for s in (b for _, b in pairs) if batched else alone:
            s.sync()

    run(warm, False)
    run(warm, True)
    t1 = time.perf_counter()
    run(steps, False)
    t_alone = time.perf_counter() - t1
    t1 = time.perf_counter()
    run(steps, True)
    t_batch = time.perf_counter() - t1
    same = all(np.array_equal(a.get("p"), b.get("p")) and np.array_equal(a.get("vx"), b.get("vx"))
               for a, b in pairs)
    rate_b = len(pairs) * steps / t_batch
    rate_a = len(pairs) * steps / t_alone
    log(f"sims: {len(pairs)} simulations, {cells / 1e6:.2f} M cells; {steps} steps: batched "
        f"{rate_b:.0f} sim-steps/s, one by one {rate_a:.0f} sim-steps/s; fields equal: {same}")
    out = {
        "metric": "simulation timesteps/s, pressure solves of several simulations batched",
        "value": round(rate_b, 1), "unit": "sim-timesteps/s", "n_gpus": 1, "steps": steps,
        "warmup": warm, "ms_per_step": round(t_batch * 1e3 / steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(pairs)} simulations W 128..1024 (log-uniform, x32), "
                               f"aspect 1..4, walls + disks, MG 2/2/5, 2 V-cycles per step",
                   "cells": cells, "one_by_one_sim_steps_per_s": round(rate_a, 1),
                   "speedup_vs_one_by_one": round(rate_b / rate_a, 2),
                   "fields_bit_identical": bool(same)},
    }
    print(json.dumps(out), flush=True)


def bench_serve(args):
    """Requests arrive as a Poisson stream (seed 7) at --rate per second for
    --duration seconds; each is 3 consecutive projection solves (to 1e-6
    relative to ||F f||, warm-started) on one of --templates domains (W
    log-uniform 64..4096, H log-uniform 64..2048, multiples of 32, 0-8% disks,
    8-mode velocity whose phases advance per timestep), drawn uniformly.  The
    server admits by a 200 GB HBM budget and batches continuously.
    value = solves/s over the run; p50/p99 = submit-to-done latency."""
    import numpy as np

    from ubootgl_amd import serve as S
    from ubootgl_amd import workloads

    t0 = time.time()
    tpls = workloads.serving_templates(args.templates, seed=7)
    log(f"{len(tpls)} templates, mean {np.mean([t.cells for t in tpls]) / 1e3:.0f} k cells, "
        f"generated in {time.time() - t0:.1f}s")
    srv = S.Server(hbm_budget=200 * 10**9, tol=1e-6, vcycles_per_iter=4,
                   steps_per_turn=args.steps_per_turn)
    gids = [srv.add_geometry(t.flag, t.h) for t in tpls]
    # warm-up: one request per template
    for g, t in zip(gids, tpls):
        srv.wait(srv.submit(g, t.rhs, t.fnorm))
    arr = workloads.arrivals(args.rate, args.duration, seed=7)
    pick = workloads.rng(8).integers(0, len(tpls), arr.size)
    base = srv.clock()
    rids = []
    tw = time.perf_counter()
    for ta, i in zip(arr, pick):
        while time.perf_counter() - tw < ta:
            time.sleep(min(0.0005, max(0.0, ta - (time.perf_counter() - tw))))
        rids.append(srv.submit(gids[i], tpls[i].rhs, tpls[i].fnorm))
    infos = [srv.wait(r) for r in rids]
    st = srv.stats()
    srv.close()
    # sequential plain-GPU solves of the same requests (the first --seq-requests
    # of the stream): one single-domain refinement solve at a time, each
    # timestep warm-started from the previous one and stopped at the same
    # tol * ||F f||; only the solve calls are timed (geometry setup and rhs
    # upload are not)
    from ubootgl_amd import simulation as Usim
    mgs = {}
    seq_t, seq_solves = 0.0, 0
    for i in pick[:args.seq_requests]:
        t = tpls[i]
        if i not in mgs:
            mgs[i] = Usim.MG(t.flag)
        mg = mgs[i]
        for k in range(len(t.rhs)):
            mg.ir_set_rhs(t.rhs[k])  # host checks and upload: not timed
            if k == 0:
                mg.ir_set_guess(None)
            t1 = time.perf_counter()
            mg.ir_solve(float(t.h), tol=1e-6, vcycles_per_iter=4, ref_norm=float(t.fnorm[k]))
            seq_t += time.perf_counter() - t1
            seq_solves += 1
    for mg in mgs.values():
        mg.close()
    sequential = seq_solves / seq_t
    log(f"sequential: {seq_solves} solves in {seq_t:.3f}s -> {sequential:.1f} solves/s")
    lat = np.array([i["t_done"] - i["t_submit"] for i in infos])
    ok = sum(1 for i in infos if i["state"] == "done" and i["converged"])
    span = max(i["t_done"] for i in infos) - min(i["t_submit"] for i in infos)
    solves = 3 * len(infos)
    log(f"served {len(infos)} requests ({ok} converged) in {span:.2f}s: "
        f"{solves / span:.1f} solves/s, p50 {np.percentile(lat, 50) * 1e3:.1f} ms, "
        f"p99 {np.percentile(lat, 99) * 1e3:.1f} ms, peak active {st['peak_active']}, "
        f"peak {st['peak_bytes'] / 1e9:.2f} GB")
    out = {
        "metric": "served solves/s with p50/p99 latency @1e-6, Poisson arrivals",
        "value": round(solves / span, 3), "unit": "solves/s", "n_gpus": 1,
        "steps": len(infos), "warmup": len(tpls), "ms_per_step": round(span * 1e3 / len(infos), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"Poisson arrivals {args.rate:g}/s for {args.duration:g}s (seed 7), "
                               f"3 warm-started solves per request to 1e-6, {len(tpls)} domains "
                               f"64x64..4096x2048, 200 GB budget",
                   "requests": len(infos), "converged": ok,
                   "requests_per_s": round(len(infos) / span, 2),
                   "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
                   "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
                   "peak_active": st["peak_active"],
                   "peak_GB": round(st["peak_bytes"] / 1e9, 3),
                   "worker_steps": st["steps"], "worker_busy_s": round(st["busy_s"], 3),
                   "worker_s": {k: round(st[k + "_s"], 3)
                                for k in ("admit", "setup", "wait", "retire")},
                   "sequential_solves_per_s": round(sequential, 3),
                   "sequential_sample": f"the first {min(args.seq_requests, len(pick))} "
                                        f"requests of the stream, one solve at a time",
                   "speedup_vs_sequential": round(solves / span / sequential, 2)},
    }
    print(json.dumps(out), flush=True)


def host_cpu():
    """The host as lscpu reports it, plus the CPUs this process may run on."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket":
                "cores_per_socket", "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys and keys[k.strip()] not in info:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except Exception as e:  # lscpu missing: report what Python knows
        info["lscpu_error"] = str(e)
    info["affinity_cpus"] = len(os.sched_getaffinity(0))
    return info


def timing_build(host):
    """oracle/'s timing build (the reference's -Ofast -march=native flags,
    eight-wide advect) for THIS host: rebuilt when it was built on another CPU
    model (-march=native)."""
    import subprocess

    from oracle import pyoracle as O
    stamp = os.path.join(ROOT, "oracle", "_build", "native.cpu")
    model = str(host.get("model", "unknown"))
    fresh = os.path.exists(O.NATIVE_LIB_PATH) and os.path.exists(stamp) and \
        open(stamp).read().strip() == model
    if not fresh:
        t = time.time()
        out = subprocess.run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "native"],
                             capture_output=True, text=True)
        if out.returncode != 0:
            raise RuntimeError("oracle timing build failed:\n" + out.stderr[-2000:])
        with open(stamp, "w") as fh:
            fh.write(model + "\n")
        log(f"cpu baseline: built the timing oracle for {model} in {time.time() - t:.1f}s")
    return O.NATIVE_LIB_PATH


def cpu_baseline(d, xy, nsteps, runs=3):
    """The CPU restatement of the step (oracle/, OpenMP) in its timing build
    -- the reference's compiler flags -std=c++17 -Ofast -march=native
    -fopenmp -DNDEBUG (Makefile:12) and an eight-wide AVX2 advect as the
    reference's -- on the same workload with the tracers: `runs` runs of
    `nsteps` steps after one warm-up step; value = the median run.  Threads:
    every CPU this process may use, capped by OMP_NUM_THREADS when set (the
    GPU box gives each GPU a share of the host's cores), bound close, one per
    core."""
    import numpy as np

    from oracle import pyoracle as O

    host = host_cpu()
    lib = timing_build(host)
    threads = int(os.environ.get("UBG_CPU_THREADS", "0"))
    if not threads:
        threads = host["affinity_cpus"]
        if os.environ.get("OMP_NUM_THREADS", "").isdigit():
            threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    # read by the OpenMP runtime when the library loads (nothing in this
    # process has loaded it before)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    O.use_timing_build(lib)
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
    times = []
    k = 0
    for _ in range(runs):
        t = time.perf_counter()
        for _ in range(nsteps):
            ref.step(dt)
            if pts.shape[0]:
                ref.advect_tracers(pts, age, dt, k + 1)
            k += 1
        times.append(time.perf_counter() - t)
    med = float(np.median(times))
    log(f"cpu baseline: {runs} x {nsteps} steps, {', '.join(f'{t:.2f}' for t in times)} s "
        f"on {threads} threads ({host.get('model')})")
    return {"value": round(nsteps / med, 4), "unit": "timesteps/s", "cores": threads,
            "kind": "port",
            "sample": f"median of {runs} runs of {nsteps} timesteps (+1 warm-up) of the same "
                      f"{d.W}x{d.H} workload with its {pts.shape[0]} tracers; oracle/ timing "
                      f"build (g++ -std=c++17 -Ofast -march=native -fopenmp -DNDEBUG, the "
                      f"reference's flags; eight-wide AVX2 advect); OMP_PROC_BIND="
                      f"{os.environ.get('OMP_PROC_BIND')}, OMP_PLACES="
                      f"{os.environ.get('OMP_PLACES')}",
            "runs_s": [round(t, 3) for t in times],
            "host": host}


if __name__ == "__main__":
    main()
