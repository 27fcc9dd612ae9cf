"""The algorithmic operation count bench.py prices its roofline with.

bench.py states 2,148.7 FP32 operations per creature-substep for the
3,000,000-creature, 20 s workload, counted by the oracle built with a counting
scalar (oracle/sim_count.cpp). This re-counts a smaller strided sample of the
same batch and checks the convention on a case with a known count.
"""
import importlib.util
from pathlib import Path

import numpy as np

import evolution_simulator_gpu_amd as E
import oracle

ROOT = Path(__file__).resolve().parent.parent


def _bench():
    spec = importlib.util.spec_from_file_location("bench", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_bench_flop_figure_matches_a_recount():
    batch = E.synth(E.synth_spec(1, 3_000_000))
    idx = np.linspace(0, batch.count - 1, num=384).astype(np.int64)
    sub = batch.subset(idx)
    p, run = E.run_for(sub, E.world(duration=20.0, substeps=8, rate=60))
    ops, steps = oracle.count_ops(sub, p, run)
    per = ops / (steps * 8)
    assert abs(per / _bench().FLOPS_PER_CREATURE_SUBSTEP - 1.0) < 0.05, per


def test_counts_grow_with_the_body():
    """More muscles and nodes cost more per substep, and the count per step
    does not depend on the trial's length."""
    counts = []
    for preset in (0, 1, 2):
        b = E.synth(E.synth_spec(preset, 64))
        p, run = E.run_for(b, E.world(duration=1.0, substeps=8, ground=0, gravity=0.0))
        ops, steps = oracle.count_ops(b, p, run)
        counts.append(ops / (steps * 8))
    assert counts[0] < counts[1] < counts[2]
    b = E.synth(E.synth_spec(0, 32))
    w1 = E.world(duration=1.0, substeps=8, ground=0, gravity=0.0)
    w2 = E.world(duration=2.0, substeps=8, ground=0, gravity=0.0)
    o1, s1 = oracle.count_ops(b, *E.run_for(b, w1))
    o2, s2 = oracle.count_ops(b, *E.run_for(b, w2))
    assert s2 > 1.5 * s1
    assert abs((o2 / s2) / (o1 / s1) - 1.0) < 0.02
