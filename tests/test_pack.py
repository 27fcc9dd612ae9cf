"""Genome packing (src/physics.rs:182-296, src/physics2.rs:190-427,
src/kernel.rs:336-593) on the reference's own fixture
(tests/fixtures/warp_walker.json, copied to tests/golden/) and on random
canonical creatures, pinned by what the model must conserve: mass, bone
lengths, rest angles, the start pose's centre and ground contact, the muscle
record's derived fields, the packing order and the slot map."""
import numpy as np
import pytest

import evolution_simulator_gpu_amd as E
import oracle
from evolution_simulator_gpu_amd import _native as N
from evolution_simulator_gpu_amd import genome as G
from tests.helpers import assert_bit_exact

from pathlib import Path

WALKER = (Path(__file__).resolve().parent / "golden" / "warp_walker.json").read_text()


def _walker(id=0):
    c = G.Creature.from_json(WALKER)
    c.id = id
    return c


def _unpack(batch, c):
    h = batch.heads[c]
    n, m = int(h[0] & 255), int(h[1])
    rec = batch.records[h[3]:h[3] + n * 6].reshape(n, 6)
    bones = batch.records[h[3] + n * 6:h[3] + n * 6 + (n - 1) * 6].reshape(n - 1, 6)
    mus = batch.muscles[h[4]:h[4] + m * 20].reshape(m, 20)
    return rec, bones, mus


def _order(creature):
    return [0] + [b["b"] for b in creature.bones]


def test_walker_mass_is_conserved():
    """Node masses (0.1 kg at 8 cm, with the square of the diameter), half of
    each bone's 4 kg/m^2 * L^2 at either end, organs, and 0.05 kg + 1 kg/m
    of slack length per muscle: the packed nodes carry the sum."""
    w = _walker()
    batch = G.pack([w], E.world())
    rec, _, mus = _unpack(batch, 0)
    d = np.array([n["diameter"] for n in w.nodes])
    nodes = np.clip(0.1 * (d / 0.08) ** 2, 0.02, 10.0).sum()
    bones = sum(4.0 * b["rest_length"] ** 2 + b["organ_mass"] for b in w.bones)
    muscles = sum(0.05 + float(mus[k, 7]) for k in range(len(w.muscles)))
    total = rec[:, 0].view(np.float32).astype(np.float64).sum()
    assert abs(total - (nodes + bones + muscles)) < 1e-5 * total
    assert batch.heads[0, 5:6].view(np.float32)[0] == np.float32(total) or \
        abs(batch.heads[0, 5:6].view(np.float32)[0] - total) < 1e-5 * total
    # slack: the longest length or the start span, whichever is longer
    assert np.all(mus[:, 7] >= np.array([m["long"] for m in w.muscles], np.float32))


@pytest.mark.parametrize("world", [dict(), dict(terrain=0.08), dict(quake=0.12, terrain=0.03), dict(ground=0)])
def test_walker_start_pose(world):
    """Bones keep their rest lengths, joints their rest angles; the centre
    of mass sits over x = 0 and the lowest surface on the ground
    (src/physics2.rs:365-398)."""
    w = _walker(id=77)
    cfg = E.world(**world)
    batch = G.pack([w], cfg)
    rec, bones, _ = _unpack(batch, 0)
    x = rec[:, 3].view(np.float32).astype(np.float64)
    y = rec[:, 4].view(np.float32).astype(np.float64)
    m = rec[:, 0].view(np.float32).astype(np.float64)
    r = rec[:, 1].view(np.float32).astype(np.float64)
    assert abs((m * x).sum() / m.sum()) < 1e-6
    pivot = bones[:, 0]
    length = bones[:, 1].view(np.float32)
    for j in range(len(pivot)):
        assert abs(np.hypot(x[j + 1] - x[pivot[j]], y[j + 1] - y[pivot[j]]) - length[j]) < 1e-5
    # absolute bone angles from the pose; relative to the parent bone they are
    # the rest angles (low end of the range minus the gene's min angle)
    ang = np.arctan2(y[1:] - y[pivot], x[1:] - x[pivot])
    lo = bones[:, 3].view(np.float32)
    genes = w.bones
    order = _order(w)
    for j in range(1, len(pivot)):
        parent = bones[j, 2]
        rest = float(lo[j]) - genes[j]["min_angle"]
        rel = (ang[j] - ang[parent] + np.pi) % (2 * np.pi) - np.pi
        assert abs(((rel - rest + np.pi) % (2 * np.pi)) - np.pi) < 1e-5
    if world.get("ground", 1):
        lib = N.lib()
        amp = cfg.terrain + (cfg.quake * (0.6 + ((lib.evo_quake_hash(77) >> 16) & 0xFFFF) * (0.8 / 65536))
                             if cfg.quake > 0 else 0.0)
        phase = (lib.evo_quake_hash(77) & 0xFFFF) / 65536.0 if cfg.quake > 0 else 0.0
        h = np.array([lib.evo_ground_height(cfg, float(xi), amp, phase) for xi in x])
        assert abs((y - r - h).min()) < 1e-5
    else:
        assert abs(y[0] - w.nodes[0]["y"]) < 1e-6  # not lowered without a ground
    assert len(order) == len(w.nodes)


def test_walker_feet_and_muscle_records():
    """The one foot ends the leg (src/kernel.rs:345-351); muscle records
    carry the genes' rhythm, the renumbered sensor node and the strength
    bounds (src/kernel.rs:370-409)."""
    w = _walker()
    batch = G.pack([w], E.world())
    rec, _, mus = _unpack(batch, 0)
    assert list(rec[:, 5]) == [0, 0, 0, 0, 0, 1]
    order = _order(w)
    renumber = {node: r for r, node in enumerate(order)}
    f32 = np.float32
    for k, g in enumerate(w.muscles):
        row = mus[k]
        assert row[9] == f32(1.0) / f32(g["period"])
        assert row[10] == f32(g["phase"]) and row[11] == f32(g["duty"]) and row[14] == f32(g["stiffness"])
        cap = f32(f32(f32(f32(48.0) * f32(g["period"])) * min(f32(g["duty"]), f32(1.0) - f32(g["duty"])))) / f32(np.pi)
        assert row[8] == min(f32(g["long"]) - f32(g["short"]), cap)
        strength = row[4] / 200.0
        assert 0.0 < strength <= 1.0
        assert abs(row[5] - 1.0 / (120.0 * strength)) < 1e-6 * row[5]
        assert row[3] == f32(1.0) / (f32(8.0) * max(f32(g["long"]), f32(0.05)))
        sensor = row[16:17].view(np.uint32)[0]
        if g["sensor"] == 255:
            assert sensor == N.NONE
        else:
            ba, bb = w.bones[g["bone_a"]], w.bones[g["bone_b"]]
            assert sensor == renumber[[ba["a"], ba["b"], bb["a"], bb["b"]][g["sensor"]]]
        nodes = row[0:1].view(np.uint32)[0]
        assert (nodes >> 8) & 255 == g["bone_a"] + 1 and nodes >> 24 == g["bone_b"] + 1
    # muscles on one pair of bones on one side share their strength: at most
    # as strong together as one alone could be
    same = {}
    for k, g in enumerate(w.muscles):
        same.setdefault((min(g["bone_a"], g["bone_b"]), max(g["bone_a"], g["bone_b"])), []).append(mus[k, 4])
    assert any(len(v) > 1 for v in same.values())


def test_random_creatures_pack_once_in_order():
    """src/kernel.rs:643-656: every creature packs once; packed order is by
    (muscles, nodes, index) and the slots map back."""
    rng = np.random.default_rng(3)
    cs = [G.random_creature(rng, int(rng.integers(2, 14)), int(rng.integers(0, 20)), id=i) for i in range(300)]
    batch = G.pack(cs, E.world(), threads=4)
    batch.validate()
    assert np.array_equal(np.sort(batch.slots), np.arange(300))
    key = batch.muscle_counts().astype(np.int64) * 1000 + batch.nodes() * 1 + 0
    assert np.all(np.diff(key) >= 0)
    for c in range(0, 300, 37):
        src = cs[batch.slots[c]]
        assert batch.nodes()[c] == len(src.nodes) and batch.muscle_counts()[c] == len(src.muscles)
        assert batch.heads[c, 2] == N.lib().evo_quake_hash(src.id)
    again = G.pack(cs, E.world(), threads=1)
    assert np.array_equal(again.records, batch.records) and np.array_equal(again.heads, batch.heads)


def test_subset_copies_each_creature_and_renumbers_offsets():
    """Batch.subset (evo_subset): each picked creature's node, bone and
    muscle records land word for word at offsets that follow the pick order;
    the other head words and the slots come along; an empty pick is an empty
    batch; a pick of the whole batch in order is the batch."""
    rng = np.random.default_rng(7)
    batch = E.synth(E.synth_spec(4, 5000))
    for pick in (rng.choice(batch.count, 1200, replace=False), np.sort(rng.choice(batch.count, 700, replace=False)),
                 np.arange(batch.count), np.array([17]), np.array([], np.int64)):
        sub = batch.subset(pick)
        assert sub.count == len(pick)
        rw = mf = 0
        for i, c in enumerate(pick):
            h = batch.heads[c]
            n, m = int(h[0] & 255), int(h[1])
            words = n * N.NODE_WORDS + (n - 1) * N.BONE_WORDS
            assert np.array_equal(sub.records[rw:rw + words], batch.records[h[3]:h[3] + words])
            got = sub.muscles[mf:mf + m * N.MUSCLE_WORDS].view(np.uint32)
            assert np.array_equal(got, batch.muscles[h[4]:h[4] + m * N.MUSCLE_WORDS].view(np.uint32))
            assert sub.heads[i, 3] == rw and sub.heads[i, 4] == mf
            assert np.array_equal(np.delete(sub.heads[i], [3, 4]), np.delete(h, [3, 4]))
            rw += words
            mf += m * N.MUSCLE_WORDS
        assert np.array_equal(sub.slots, batch.slots[pick.astype(np.int64)])
        if len(pick):
            sub.validate()
    whole = batch.subset(np.arange(batch.count))
    assert np.array_equal(whole.records, batch.records) and np.array_equal(whole.heads, batch.heads)
    # the C entry point refuses an index past the batch
    out = np.zeros(1024, np.uint32)
    bad = np.array([batch.count], np.uint32)
    assert N.lib().evo_subset(N.ptr(batch.records), N.ptr(batch.muscles), N.ptr(batch.heads), batch.count,
                              N.ptr(bad), 1, N.ptr(out), out.size, N.ptr(out.view(np.float32)), out.size,
                              N.ptr(out)) == N.EVO_ERR_INVALID
    # and a picked head whose counts are outside the simulator's limits
    one = np.array([0], np.uint32)
    for word, value in [(0, 0), (0, 1), (0, 33), (1, 97)]:
        heads = batch.heads.copy()
        heads[0, word] = (heads[0, word] & ~np.uint32(255)) | value if word == 0 else value
        assert N.lib().evo_subset(N.ptr(batch.records), N.ptr(batch.muscles), N.ptr(heads), batch.count,
                                  N.ptr(one), 1, N.ptr(out), out.size, N.ptr(out.view(np.float32)), out.size,
                                  N.ptr(out)) == N.EVO_ERR_INVALID, (word, value)


def test_non_canonical_genomes_are_refused():
    w = _walker()
    bad = G.Creature(w.nodes, [dict(b) for b in w.bones], w.muscles)
    bad.bones[2]["a"] = 4  # node 4 is reached later: not parent-first
    with pytest.raises(ValueError):
        G.pack([bad], E.world())
    loop = G.Creature(w.nodes, [dict(b) for b in w.bones], w.muscles)
    loop.bones[4]["b"] = 1  # a node reached twice
    with pytest.raises(ValueError):
        G.pack([loop], E.world())


@pytest.mark.gpu
@pytest.mark.parametrize("world", [dict(duration=4.0), dict(duration=4.0, quake=0.05, terrain=0.03)])
def test_packed_walkers_and_random_bodies_on_the_gpu(gpu, world):
    """Eight walkers with their own ids (tests/cuda_effects.rs:31-43) and
    random bodies, packed here, run bit for bit as the oracle runs them."""
    rng = np.random.default_rng(5)
    cs = [_walker(id=i) for i in range(8)]
    cs += [G.random_creature(rng, int(rng.integers(2, 20)), int(rng.integers(0, 40)), id=100 + i) for i in range(200)]
    w = E.world(substeps=8, **world)
    batch = G.pack(cs, w)
    p, run = E.run_for(batch, w)
    want, _, _ = oracle.run_f32(batch, p, run, threads=8)
    got = E.simulate(batch, w)
    assert_bit_exact(got, want, "packed genomes")


def test_recording_decodes_to_the_creature_numbering():
    """engine::recorded / engine::thin (src/engine.rs:1062-1127, :127-152) on
    the oracle's recording of the walker: frame 0 is the packed start pose in
    the creature's own node numbering, bones keep close to their lengths,
    energies stay in [0, 1], the last frame's centre of mass is the score,
    and thinning keeps every n-th frame and the last."""
    w = _walker(id=3)
    cfg = E.world(duration=2.0, substeps=8)
    batch = G.pack([w], cfg)
    p, run = E.run_for(batch, cfg, record=True)
    res, frames, _ = oracle.run_f32(batch, p, run, frames=True)
    full = G.decode_recording(frames, p.stride, w)
    count = frames.size // (2 * p.stride)
    assert full["positions"].shape == (count, len(w.nodes), 2)
    rec, _, _ = _unpack(batch, 0)
    order = G.node_order(w)
    start = np.stack([rec[:, 3].view(np.float32), rec[:, 4].view(np.float32)], axis=1)
    assert np.array_equal(full["positions"][0][order], start)
    pos = full["positions"].astype(np.float64)
    for b in w.bones:
        d = np.hypot(*(pos[:, b["b"]] - pos[:, b["a"]]).T)
        assert np.abs(d - b["rest_length"]).max() < 0.05 * b["rest_length"] + 1e-3
    assert full["energy"].min() >= 0.0 and full["energy"].max() <= 1.0
    mass = rec[:, 0].view(np.float32).astype(np.float64)
    # the score is the centre of mass after the trial's last step (at a fall
    # the recording goes on limp): frame settle + steps
    steps = int(res["ground_hi"].view(np.uint32)[0] >> 16)
    at = N.lib().evo_settle_frames(cfg.rate) + steps
    com = (pos[at][order][:, 0] * mass).sum() / mass.sum()
    assert abs(com - res["fitness"][0]) < 1e-5
    thin = G.decode_recording(frames, p.stride, w, every=7)
    keep = [t for t in range(count) if t % 7 == 0 or t == count - 1]
    assert len(thin["positions"]) == len(keep)
    for key in ("positions", "energy", "ground", "friction", "broken"):
        assert np.array_equal(thin[key], full[key][keep])
    assert np.all(full["ground"] >= 0.0)
