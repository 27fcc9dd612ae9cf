"""GPU against the FP32 oracle on every world effect, the rung rule and the
early screen, recordings, repeats and batch invariance.

Re-expresses the reference's GPU invariants: tests/gpu_repeatability.rs:20-140
(repeat runs bit-identical, a replay's score equals the trial's score),
tests/cuda_physics.rs:134-230 (a creature scores the same alone or in a
mixed batch) and tests/cuda_effects.rs:21-169 (every effect changes the
trial), each checked bit for bit against the oracle.
"""
import numpy as np
import pytest

import evolution_simulator_gpu_amd as E
import oracle
from evolution_simulator_gpu_amd import _native as N
from tests.helpers import assert_bit_exact, bits

pytestmark = pytest.mark.gpu

EFFECTS = {
    "terrain": dict(terrain=0.08),
    "slope": dict(slope=0.08),
    "gaps": dict(gaps=0.35),
    "hurdles": dict(hurdles=0.08),
    "quake": dict(quake=0.05),
    "mud": dict(mud=0.05),
    "water": dict(water=0.35),
    "ice": dict(patches=0.8),
    "wind": dict(wind=-3.0),
    "air": dict(air_retention=0.985),
    "brambles": dict(brambles=8.0),
    "gravity": dict(gravity=19.6),
    "heat_drought": dict(muscle_energy=0.5, muscle_recovery=0.3),
    "hard": dict(terrain=0.25, gravity=19.6, ground_friction=0.6, patches=0.8, gaps=0.35, hurdles=0.08,
                 quake=0.05),
    "no_ground": dict(ground=0),
}


def _check(batch, w, label, threads=8):
    p, run = E.run_for(batch, w)
    want, _, _ = oracle.run_f32(batch, p, run, threads=threads)
    got = E.simulate(batch, w)
    assert_bit_exact(got, want, label)
    return got


@pytest.mark.parametrize("name", sorted(EFFECTS))
def test_world_effect_is_bit_exact(gpu, name):
    batch = E.synth(E.synth_spec(0, 512), worlds=None)
    w = E.world(duration=2.0, substeps=8, **EFFECTS[name])
    got = _check(batch, w, name)
    # the effect changes some body's distance (tests/cuda_effects.rs:21-169)
    calm = E.simulate(batch, E.world(duration=2.0, substeps=8))
    if name != "heat_drought":
        assert np.abs(got["fitness"] - calm["fitness"]).max() > 5e-3


def test_large_bodies_in_a_hard_world(gpu):
    batch = E.synth(E.synth_spec(2, 256))
    w = E.world(duration=2.0, substeps=16, **EFFECTS["hard"], water=0.12, wind=-1.0, mud=0.02)
    _check(batch, w, "large bodies, hard world")


def test_rungs_and_screen_are_bit_exact(gpu):
    """Both early rungs armed and the 5 s screen, with every creature flag
    (src/rungs.rs:18-35): the stops, the trace words and the end codes."""
    batch = E.synth(E.synth_spec(0, 1024))
    flags = np.array([0, 1, 2, 4, 8, 16, 32, 3], np.uint32)[np.arange(batch.count) % 8]
    batch.heads[:, 7] = (batch.heads[:, 7] & 0xFFFF) | (flags << 16)
    r1 = N.Rung((0.9, 0.1, 0.0, -0.01, 0.2, 0.0), -0.05, 0b10000000)
    r2 = N.Rung((0.8, 0.2, 0.1, 0.0, 0.0, -0.1), 0.02, 0b00000010)
    w = E.world(duration=6.0, substeps=8, screen_seconds=5.0, screen_bar=0.05, screen_bar_young=-0.05,
                screen_bar_reshaped=0.0, r1=r1, r2=r2)
    got = _check(batch, w, "rungs and screen")
    code = got["ground_hi"].view(np.uint32) & 0xFFFF
    assert ((code >> 6) & 3 == 1).any() and ((code >> 6) & 3 == 2).any()  # both rungs stopped someone
    assert (got["screened"] > 0).any() and (got["screen_x"] != 0).any()
    assert ((code >> 14) & 1 == 1).any()  # audit creatures carry their bit


def test_recording_frames_and_score(gpu):
    """RECORD: every frame equals the oracle's, and the recorded result is the
    trial's own result bit for bit (tests/gpu_repeatability.rs:76-140)."""
    cases = [(1, E.world(duration=1.5, substeps=8)),  # unrolled recording class
             (1, E.world(duration=1.0, substeps=16, rate=120, **EFFECTS["hard"])),
             (2, E.world(duration=1.0, substeps=8, **EFFECTS["hard"]))]  # looped recording class
    for preset, w in cases:
        batch = E.synth(E.synth_spec(preset, 24))
        p, run = E.run_for(batch, w, record=True)
        want, wframes, _ = oracle.run_f32(batch, p, run, threads=4, frames=True)
        got, frames = E.simulate(batch, w, record=True)
        assert_bit_exact(got, want, "recording result")
        plain = E.simulate(batch, w)
        assert np.array_equal(bits(got), bits(plain))
        a = frames.view(np.uint32)
        b = wframes.view(np.uint32)
        nan = np.isnan(frames) & np.isnan(wframes)
        assert np.array_equal(a[~nan], b[~nan])


def test_recording_bodies_with_tendons_on_flat_ground(gpu):
    """Flat ground with tendons records through the any-world unrolled
    instance (flat tendon-free batches take their own): frames and result
    equal the oracle's."""
    from evolution_simulator_gpu_amd import genome as G

    rng = np.random.default_rng(23)
    w = E.world(duration=1.5, substeps=8)
    cs = [G.random_creature(rng, int(rng.integers(3, 13)), int(rng.integers(2, 15)), id=300 + i) for i in range(24)]
    batch = G.pack(cs, w)
    assert not batch.tendon_free() and batch.nodes().max() <= 12 and batch.muscle_counts().max() <= 24
    p, run = E.run_for(batch, w, record=True)
    want, wframes, _ = oracle.run_f32(batch, p, run, threads=4, frames=True)
    got, frames = E.simulate(batch, w, record=True)
    assert_bit_exact(got, want, "recording with tendons, flat ground")
    nan = np.isnan(frames) & np.isnan(wframes)
    assert np.array_equal(frames.view(np.uint32)[~nan], wframes.view(np.uint32)[~nan])


def test_repeat_runs_are_bit_identical(gpu):
    batch = E.synth(E.synth_spec(1, 4096))
    w = E.world(duration=3.0, substeps=8)
    a = E.simulate(batch, w)
    b = E.simulate(batch, w)
    assert np.array_equal(bits(a), bits(b))


def test_a_creature_scores_the_same_in_any_batch(gpu):
    """tests/cuda_physics.rs:134-230: alone, or among bodies of other sizes in
    another order, a creature's result is the same bit pattern."""
    big = E.synth(E.synth_spec(2, 2048))
    w = E.world(duration=2.0, substeps=8)
    together = E.simulate(big, w)
    rng = np.random.default_rng(5)
    picks = rng.choice(big.count, 24, replace=False)
    for c in picks[:8]:
        alone = E.simulate(big.subset([c]), w)
        assert np.array_equal(bits(alone)[0], bits(together)[c])
    shuffled = big.subset(picks[::-1])
    again = E.simulate(shuffled, w)
    assert np.array_equal(bits(again), bits(together)[picks[::-1]])


@pytest.mark.parametrize("preset", [1, 2])
def test_top_k_replay_capture(gpu, preset):
    """The k fittest creatures recorded in one launch through gathered head
    words: the replayed result equals the trial's result bit for bit
    (tests/gpu_repeatability.rs:76-140) and every frame equals the oracle's
    recording of that creature. The recording class follows the picked
    bodies' sizes, so a batch of 3-32 node bodies replays its fittest in
    whichever class they fit."""
    import torch

    batch = E.synth(E.synth_spec(preset, 3000))
    w = E.world(duration=1.5, substeps=8)
    p, run = E.run_for(batch, w)
    db = E.DeviceBatch(batch)
    db.launch(p, run)
    pick, frames, res = db.capture_top(16, p, run)
    torch.cuda.synchronize()
    pick = pick.cpu().numpy()
    full = db.results_numpy()
    assert np.all(full["fitness"][pick] >= np.sort(full["fitness"])[-16])
    got = res.cpu().numpy().view(N.RESULT_DTYPE)
    assert np.array_equal(bits(got), bits(full)[pick])
    sub = batch.subset(pick)
    q, r = E.run_for(sub, w, record=True)
    q.stride = p.stride
    _, wframes, _ = oracle.run_f32(sub, q, r, threads=4, frames=True)
    f = frames.cpu().numpy()
    nan = np.isnan(f) & np.isnan(wframes)
    assert np.array_equal(f.view(np.uint32)[~nan], wframes.view(np.uint32)[~nan])
