"""Throughput of the creature simulator on one MI355X.

Workload (the configuration the headline metric is quoted on): 3,000,000
creatures from seed 2 with 3-12 nodes and at most 24 bones + muscles, flat
ground, Earth gravity, calm air, 20 s trials at 60 Hz x 8 substeps, FP32
state. One "step" of this benchmark is one whole generation: every creature
simulated from its start pose to the end of its trial (a fall, or 1,200
steps). Inputs are resident in HBM before the timed region.

The metric is creature-substeps per second: the substeps actually simulated
(steps run x 8, summed over creatures, as examples/p2_speed.rs:56-72 counts
them) divided by the timed wall seconds.

Prints one JSON line. `--gpus N` with N > 1 is not supported: every box here
has one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FP32 operations per creature-substep of this workload, counted
# by the oracle compiled with a counting scalar (oracle/sim_count.cpp) on a
# 4,096-creature sample of the same workload; tests/test_flops.py re-measures
# it. Convention: add/sub/mul/div/min/max/floor/compare = 1, FMA = 2,
# sqrt = 1, cos = 14, atan2 = 32.
FLOPS_PER_CREATURE_SUBSTEP = 2107.6
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32, 256 CUs at 2.4 GHz


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--creatures", type=int, default=3_000_000)
    ap.add_argument("--duration", type=float, default=20.0)
    ap.add_argument("--cpu-sample", type=int, default=16384)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def cpu_baseline(batch, world, sample: int, threads: int):
    """The FP32 oracle (a port of the reference's kernel to C) on host cores,
    over an evenly strided sample of the same batch."""
    import oracle  # the checker; timed here only as the CPU baseline

    from evolution_simulator_gpu_amd import run_for

    idx = np.linspace(0, batch.count - 1, num=min(sample, batch.count)).astype(np.int64)
    sub = batch.subset(idx)
    p, run = run_for(sub, world)
    t0 = time.perf_counter()
    _, _, steps = oracle.run_f32(sub, p, run, threads=threads)
    dt = time.perf_counter() - t0
    substeps = steps * world.substeps
    return {"value": substeps / dt, "unit": "creature-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{sub.count} creatures strided over the same 3M batch, 20 s trials, "
                      f"{substeps:.3e} substeps in {dt:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported on these boxes.",
              file=sys.stderr)
        sys.exit(2)
    import torch

    import evolution_simulator_gpu_amd as E
    from evolution_simulator_gpu_amd import _native as N

    t0 = time.perf_counter()
    spec = E.synth_spec(1, args.creatures)
    batch = E.synth(spec)
    batch.validate()
    world = E.world(duration=args.duration, substeps=8, rate=60)
    p, run = E.run_for(batch, world)
    gen_s = time.perf_counter() - t0
    db = E.DeviceBatch(batch, "cuda")
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        db.launch(p, run, stream)
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_begin = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        db.launch(p, run, stream)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_begin
    kernel_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]

    res = db.results_numpy()
    steps_run = (res["ground_hi"].view(np.uint32) >> 16).astype(np.int64)
    substeps_per_gen = int(steps_run.sum()) * world.substeps
    value = substeps_per_gen * args.steps / wall
    avg_launch_s = float(np.mean(kernel_ms)) / 1e3
    flops_per_launch = FLOPS_PER_CREATURE_SUBSTEP * substeps_per_gen
    achieved = flops_per_launch / avg_launch_s / 1e12

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(batch, world, args.cpu_sample, threads)

    line = {
        "metric": "creature-substeps/s at 3M creatures, 20 s trials",
        "value": value,
        "unit": "creature-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain forward: 3,000,000 creatures from seed 2, 3-12 nodes, bones + muscles <= 24, "
                        "flat ground, 20 s trials at 60 Hz x 8 substeps",
            "creatures": batch.count,
            "substeps_per_generation": substeps_per_gen,
            "fell_share": float((res["fall_time"] > 0).mean()),
            "generate_s": gen_s,
            "kernel": N.lib().evo_build_info().decode(),
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None,
            "flops_per_creature_substep": FLOPS_PER_CREATURE_SUBSTEP,
            "avg_launch_ms": avg_launch_s * 1e3,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
