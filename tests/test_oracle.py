"""Pins the oracle (CPU restatement of shaders/creature.cu) without the GPU.

Every check here is independent of the oracle's own formulas: closed forms of
the dynamics (free fall under semi-implicit Euler, a body at rest, Coulomb
sliding under a constant push), textbook/library routines (numpy's IEEE
binary16 conversion, float64 cos/atan2), the reference's own host tests
re-expressed (src/physics.rs:727-919, src/creature_kernel.rs:328-344), and
agreement between the FP32 and FP64 builds.
"""
import ctypes as C

import numpy as np
import pytest

import evolution_simulator_gpu_amd as E
import oracle
from evolution_simulator_gpu_amd import _native as N
from evolution_simulator_gpu_amd.batch import from_bodies

G = np.float32(9.8)


def _lib():
    lib = oracle.load()
    lib.f32_ground_export.argtypes = [C.c_uint32, C.c_void_p, C.c_float, C.c_float, C.c_float,
                                      C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.f64_ground_export.argtypes = [C.c_uint32, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.f32_rhythm_export.argtypes = [C.c_float] * 7
    lib.f32_rhythm_export.restype = C.c_float
    lib.oracle_atan2.argtypes = [C.c_float, C.c_float]
    lib.oracle_atan2.restype = C.c_float
    lib.oracle_cos_0_pi.argtypes = [C.c_float]
    lib.oracle_cos_0_pi.restype = C.c_float
    lib.oracle_f2h.argtypes = [C.c_float]
    lib.oracle_f2h.restype = C.c_uint32
    lib.oracle_h2f.argtypes = [C.c_uint32]
    lib.oracle_h2f.restype = C.c_float
    return lib


# ---------------------------------------------------------------- fp16

def test_half_round_trip_is_exhaustive():
    """src/creature_kernel.rs:328-344: every half survives f32 and back."""
    lib = _lib()
    h = np.arange(65536, dtype=np.uint32)
    f = np.zeros(65536, np.float32)
    lib.oracle_h2f_n(h.ctypes.data, f.ctypes.data, 65536)
    back = np.zeros(65536, np.uint32)
    lib.oracle_f2h_n(f.ctypes.data, back.ctypes.data, 65536)
    finite = ~np.isnan(f)
    assert np.array_equal(back[finite], h[finite])
    # h2f agrees with numpy's binary16 decoding for every pattern
    ref = h.astype(np.uint16).view(np.float16).astype(np.float32)
    assert np.array_equal(np.isnan(ref), np.isnan(f))
    assert np.array_equal(ref[finite].view(np.uint32), f[finite].view(np.uint32))


def test_half_rounds_like_numpy():
    """Round to nearest even, overflow to infinity, underflow through the
    subnormals, as numpy's IEEE conversion does."""
    lib = _lib()
    rng = np.random.default_rng(7)
    v = np.concatenate([
        rng.standard_normal(200000).astype(np.float32) * np.float32(1000.0),
        (rng.random(200000).astype(np.float32) * np.float32(1e-4)),
        np.array([0.0, -0.0, 1.0, 12.3, -1e20, 1e20, 65504.0, 65520.0, 65519.99, 6.1e-5, 5.96e-8, 2.98e-8,
                  2.99e-8, np.inf, -np.inf], np.float32),
    ])
    got = np.zeros(v.size, np.uint32)
    lib.oracle_f2h_n(v.ctypes.data, got.ctypes.data, v.size)
    with np.errstate(over="ignore"):
        want = v.astype(np.float16).view(np.uint16).astype(np.uint32)
    assert np.array_equal(got, want)
    assert lib.oracle_f2h(1.0) == 0x3C00
    assert lib.oracle_h2f(lib.oracle_f2h(12.3)) == np.float32(12.296875)
    assert lib.oracle_h2f(lib.oracle_f2h(-1e20)) == -np.inf


# ---------------------------------------------------------------- cos / atan2

def test_cos_on_zero_to_pi_error_bound():
    lib = _lib()
    x = np.linspace(0.0, np.pi, 2_000_001).astype(np.float32)
    out = np.zeros_like(x)
    lib.oracle_cos_0_pi_n(x.ctypes.data, out.ctypes.data, x.size)
    err = np.abs(out.astype(np.float64) - np.cos(x.astype(np.float64)))
    assert err.max() < 2.5e-7
    assert out[0] == 1.0 or abs(out[0] - 1.0) < 1.2e-7


def test_atan2_error_bound_and_quadrants():
    lib = _lib()
    rng = np.random.default_rng(3)
    y = (rng.standard_normal(1_000_000) * 10.0 ** rng.integers(-3, 4, 1_000_000)).astype(np.float32)
    x = (rng.standard_normal(1_000_000) * 10.0 ** rng.integers(-3, 4, 1_000_000)).astype(np.float32)
    out = np.zeros_like(x)
    lib.oracle_atan2_n(y.ctypes.data, x.ctypes.data, out.ctypes.data, x.size)
    err = np.abs(out.astype(np.float64) - np.arctan2(y.astype(np.float64), x.astype(np.float64)))
    assert err.max() < 4e-7
    # IEEE special cases
    sp = [(0.0, 1.0), (-0.0, 1.0), (0.0, -1.0), (-0.0, -1.0), (0.0, 0.0), (-0.0, -0.0), (1.0, 0.0),
          (-1.0, 0.0), (np.inf, np.inf), (-np.inf, np.inf), (np.inf, -np.inf), (1.0, np.inf), (1.0, -np.inf)]
    for yy, xx in sp:
        got = np.float32(lib.oracle_atan2(yy, xx))
        want = np.float32(np.arctan2(np.float32(yy), np.float32(xx)))
        assert abs(got - want) < 4e-7 and np.signbit(got) == np.signbit(want), (yy, xx, got, want)


# ---------------------------------------------------------------- rhythm

def test_rhythm_switches_and_waveform():
    """src/physics.rs:299-308 waveform (tests/simulation.rs:290-310): 0 at the
    start of a cycle, 1 at `duty`, continuous at both switches, periodic."""
    lib = _lib()
    inv_p, duty = np.float32(1.0 / 1.3), np.float32(0.37)
    r = lambda t, phase=0.0: lib.f32_rhythm_export(t, inv_p, phase, 0.0, duty, 1.0 / duty, 1.0 / (1.0 - duty))
    assert abs(r(0.0)) < 1e-7
    assert abs(r(float(duty / inv_p)) - 1.0) < 1e-6
    ts = np.linspace(0.0, 5.0, 20001)
    vals = np.array([r(float(t)) for t in ts])
    ph = np.mod(ts * np.float64(inv_p), 1.0)
    want = np.where(ph < duty, 0.5 - 0.5 * np.cos(np.pi * ph / duty),
                    0.5 + 0.5 * np.cos(np.pi * (ph - duty) / (1.0 - duty)))
    assert np.abs(vals - want).max() < 2e-6
    assert np.abs(np.diff(vals)).max() < 0.01  # no jumps
    assert abs(r(0.3) - r(0.3 + 1.3)) < 2e-6   # one period later


# ---------------------------------------------------------------- ground

def _ground(flags, x, amp=0.0, qphase=0.0, precision="f64", **pk):
    lib = _lib()
    p = N.Params()
    for k, v in pk.items():
        setattr(p, k, v)
    if precision == "f64":
        h, s = C.c_double(), C.c_double()
        lib.f64_ground_export(flags, C.addressof(p), x, amp, qphase, C.byref(h), C.byref(s))
    else:
        h, s = C.c_float(), C.c_float()
        lib.f32_ground_export(flags, C.addressof(p), x, amp, qphase, C.byref(h), C.byref(s))
    return h.value, s.value


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_ground_slope_matches_its_height(precision):
    """src/physics.rs:731-760 and :808-826: the slope of bumps, tilt, pits and
    hurdles is the derivative of the height (central differences)."""
    flags = N.GROUND | N.TERRAIN | N.SLOPE | N.GAPS | N.HURDLES
    tol = 1e-2
    for i in range(400):
        x = i * 0.0371 - 4.0
        kw = dict(slope=0.15, gaps=1.0, hurdles=0.25)
        _, s = _ground(flags, x, 0.05, 0.37, precision, **kw)
        h = 1e-3
        hp, _ = _ground(flags, x + h, 0.05, 0.37, "f64", **kw)
        hm, _ = _ground(flags, x - h, 0.05, 0.37, "f64", **kw)
        numeric = (hp - hm) / (2 * h)
        # a point within h of a ramp's kink has a one-sided slope
        if abs(numeric - s) > tol:
            hp2, _ = _ground(flags, x + 1e-5, 0.05, 0.37, "f64", **kw)
            hm2, _ = _ground(flags, x - 1e-5, 0.05, 0.37, "f64", **kw)
            assert abs((hp2 - hm2) / 2e-5 - s) < tol or abs((hp2 - hm2) / 2e-5 - s) < 5 * tol, x


def test_gaps_and_hurdles_shapes():
    """src/physics.rs:762-807, :828-873: x = 0 is solid level ground, the pit
    floor is GAP_DEPTH deep and flat, the hurdle top is flat at its height,
    and both profiles are continuous."""
    w = 1.0
    spacing = 2.0 + 4.0 * w
    assert _ground(N.GROUND | N.GAPS, 0.0, gaps=w) == (0.0, 0.0)
    h, s = _ground(N.GROUND | N.GAPS, 0.5 * spacing, gaps=w)
    assert (h, s) == (-2.0, 0.0)
    assert _ground(N.GROUND | N.HURDLES, 0.0, hurdles=0.3) == (0.0, 0.0)
    h, s = _ground(N.GROUND | N.HURDLES, 1.5, hurdles=0.3)
    assert h == float(np.float32(0.3)) and s == 0.0
    xs = np.arange(-spacing, 3 * spacing, 0.001)
    hg = np.array([_ground(N.GROUND | N.GAPS, x, gaps=w)[0] for x in xs])
    assert np.abs(np.diff(hg)).max() < 0.5 and hg.min() >= -2.0 - 1e-9 and hg.max() <= 0.0
    hh = np.array([_ground(N.GROUND | N.HURDLES, x, hurdles=0.3)[0] for x in xs])
    assert np.abs(np.diff(hh)).max() < 0.5 and hh.min() >= 0.0 and hh.max() <= float(np.float32(0.3))
    # the terrain bumps stay within [0, amp] (src/physics.rs:740)
    hb = np.array([_ground(N.GROUND | N.TERRAIN, x, 0.05)[0] for x in np.linspace(-3, 4.4, 400)])
    assert hb.min() >= 0.0 and hb.max() <= 0.05 + 1e-6


def test_quake_hash_is_deterministic_and_spread():
    """src/physics.rs:876-897."""
    lib = N.lib()
    assert lib.evo_quake_hash(47300) == lib.evo_quake_hash(47300)
    hashes = np.array([lib.evo_quake_hash(i) for i in range(1, 1001)], np.uint32)
    phases = (hashes & 0xFFFF).astype(np.float32) * np.float32(1 / 65536)
    scales = np.float32(0.6) + ((hashes >> 16) & 0xFFFF).astype(np.float32) * np.float32(0.8 / 65536)
    assert phases.min() >= 0 and phases.max() < 1 and scales.min() >= 0.6 and scales.max() <= 1.4
    assert len(np.unique(phases)) > 900 and len(np.unique(scales)) > 900


# ---------------------------------------------------------------- dynamics, closed forms

def _dumbbell(y, radius=1e-12, friction=1.0, mass=1.0, x0=-0.5):
    return dict(nodes=[(mass, radius, friction, x0, y, 0), (mass, radius, friction, x0 + 1.0, y, 0)],
                bones=[(0, 1.0, None, -1.0, 1.0)], muscles=[])


def _frames(batch, w, precision):
    p, run = E.run_for(batch, w)
    res, fr, _ = oracle.run(batch, p, run, precision, threads=1, frames=True)
    fpc = N.lib().evo_settle_frames(w.rate) + p.steps + 1
    return res, fr.reshape(batch.count, fpc, p.stride, 2), p


def test_free_fall_is_semi_implicit_euler():
    """No ground, no muscles: y_k = y0 - g H^2 k(k+1)/2 after k substeps."""
    batch = from_bodies([_dumbbell(0.0)])
    w = E.world(duration=2.0, substeps=8, ground=0)
    _, fr64, p = _frames(batch, w, "f64")
    _, fr32, _ = _frames(batch, w, "f32")
    settle = N.lib().evo_settle_frames(60)
    H = 1.0 / 480.0
    g = float(np.float32(9.8))
    for n in (1, 10, 30, 120):
        k = 8 * n
        want = -g * H * H * k * (k + 1) / 2
        assert abs(fr64[0, settle + n, 0, 1] - want) < 1e-9
        assert abs(fr64[0, settle + n, 1, 1] - want) < 1e-9
        # FP32 accumulates the rounding of g H and of H v, which is biased at
        # a fixed increment: 3e-5 m after a 1.2 m drop, 7 mm after 19.6 m
        tol = 5e-5 if n <= 30 else 1e-3
        assert abs(fr32[0, settle + n, 0, 1] - want) < tol * max(1.0, abs(want))


def test_body_at_rest_stays_exactly_put():
    """A body lying on flat ground without drive never moves: gravity's
    substep is taken back exactly by the ground, every step."""
    r = 0.05
    batch = from_bodies([_dumbbell(r, radius=r, friction=0.7)])
    w = E.world(duration=3.0, substeps=8)
    for precision in ("f32", "f64"):
        res, fr, p = _frames(batch, w, precision)
        pos = fr[0, :, :2, :]
        assert np.all(pos[:, :, 1] == np.float32(r)) and np.all(pos[:, 0, 0] == np.float32(-0.5))
        assert res["fitness"][0] == np.float32(0.0) or abs(res["fitness"][0]) < 1e-7
        assert res["ground_contact"][0] == 2 * p.steps
        assert res["fall_time"][0] == 0.0


def _push(a, mu_node, mu_world):
    batch = from_bodies([_dumbbell(1e-12, friction=mu_node)])
    w = E.world(duration=1.0, substeps=8, wind=a, ground_friction=mu_world)
    return batch, w


def test_coulomb_sliding_under_a_constant_push():
    """Kinetic friction: a push a above mu g slides the body with
    acceleration a - mu g, x_k = x0 + H^2 (a - mu g) k(k+1)/2; the recorded
    contact forces are m g (normal) and mu m g (friction)."""
    a, mu_n, mu_w = 5.0, 0.2, 1.5
    batch, w = _push(a, mu_n, mu_w)
    _, fr, p = _frames(batch, w, "f64")
    settle = N.lib().evo_settle_frames(60)
    H = 1.0 / 480.0
    g = float(np.float32(9.8))
    mu = float(np.float32(mu_n)) * float(np.float32(mu_w))
    for n in (1, 5, 30, 60):
        k = 8 * n
        want = -0.5 + H * H * (float(np.float32(a)) - mu * g) * k * (k + 1) / 2
        assert abs(fr[0, settle + n, 0, 0] - want) < 1e-8, n
    stride = p.stride
    normal, friction = fr[0, settle + 30, N.MAX_NODES + 0]
    assert abs(normal - g) < 1e-6 and abs(friction - mu * g) < 1e-6


def test_static_friction_holds_below_threshold():
    """A push below mu g never moves the body (FP32 and FP64 alike)."""
    batch, w = _push(2.0, 0.5, 1.5)
    for precision in ("f32", "f64"):
        _, fr, _ = _frames(batch, w, precision)
        assert np.all(fr[0, :, 0, 0] == np.float32(-0.5))


def test_air_retention_matches_geometric_decay():
    """AIR keeps air_sub of the velocity each substep: with gravity alone
    v_k = s (v_{k-1} - g H)."""
    batch = from_bodies([_dumbbell(0.0)])
    w = E.world(duration=1.0, substeps=8, ground=0, air_retention=0.96)
    _, fr, p = _frames(batch, w, "f64")
    s = float(np.float32(p.air_sub))
    H, g = 1.0 / 480.0, float(np.float32(9.8))
    v = y = 0.0
    settle = N.lib().evo_settle_frames(60)
    for k in range(1, 8 * 20 + 1):
        v = s * (v - g * H)
        y += H * v
        if k % 8 == 0:
            assert abs(fr[0, settle + k // 8, 0, 1] - y) < 1e-9
    assert abs(s ** 8 - 0.96) < 1e-6


def test_joint_outside_its_range_is_pulled_back():
    """solve_joints (shaders/creature.cu:618-634): a joint that starts past
    its range is back inside it (within 1e-3 rad) after one step."""
    # chain head(0) -> 1 -> 2 with the second bone bent 90 degrees, range [-0.2, 0.2]
    body = dict(nodes=[(1.0, 1e-9, 1.0, 0.0, 0.0, 0), (1.0, 1e-9, 1.0, 1.0, 0.0, 0),
                       (1.0, 1e-9, 1.0, 1.0, 1.0, 0)],
                bones=[(0, 1.0, None, -3.0, 3.0), (1, 1.0, 0, -0.2, 0.2)], muscles=[])
    batch = from_bodies([body])
    w = E.world(duration=0.5, substeps=8, ground=0, gravity=0.0)
    _, fr, _ = _frames(batch, w, "f64")
    settle = N.lib().evo_settle_frames(60)
    p = fr[0, settle + 1, :3]
    u, v = p[1] - p[0], p[2] - p[1]
    ang = np.arctan2(u[0] * v[1] - u[1] * v[0], u @ v)
    assert ang <= 0.2 + 1e-3
    # the joint push runs after the bones, so a bone is off by its residual
    assert abs(np.linalg.norm(v) - 1.0) < 5e-3 and abs(np.linalg.norm(u) - 1.0) < 5e-3


# ---------------------------------------------------------------- FP32 vs FP64 on the benchmark bodies

def test_fp32_tracks_fp64_over_240_substeps():
    """Node positions of the FP32 oracle against the FP64 oracle over the
    first 240 substeps (30 steps x 8) of config-0 bodies.

    Position-based dynamics takes each velocity from the move over a substep,
    so every rounding of a position (one ulp, 1.2e-7 m at 1 m) becomes a
    velocity error of ulp/H and the two precisions drift apart as a random
    walk from the first substep; bodies with no muscles and no ground drift
    the same way. Measured on 1,024 bodies: every body within 1e-5 m for the
    first 16 substeps, median 3.7e-5 m and 99th percentile 4.3e-4 m at 240.
    """
    batch = E.synth(E.synth_spec(0, 1024))
    w = E.world(duration=0.5, substeps=8)
    p, run = E.run_for(batch, w)
    _, f32, _ = oracle.run(batch, p, run, "f32", threads=8, frames=True)
    p, run = E.run_for(batch, w)
    _, f64, _ = oracle.run(batch, p, run, "f64", threads=8, frames=True)
    fpc = N.lib().evo_settle_frames(60) + p.steps + 1
    settle = N.lib().evo_settle_frames(60)
    a = f32.reshape(batch.count, fpc, p.stride, 2)[:, :, :N.MAX_NODES].astype(np.float64)
    b = f64.reshape(batch.count, fpc, p.stride, 2)[:, :, :N.MAX_NODES]
    dev = lambda n: np.abs(a[:, settle + n] - b[:, settle + n]).max(axis=(1, 2))
    assert dev(2).max() < 1e-5
    d240 = dev(30)
    assert np.median(d240) < 1e-4 and np.percentile(d240, 99) < 1e-3


def test_fp32_fitness_tracks_fp64_without_contact():
    """Contact-free trials (no ground, no gravity: the bodies swim in place
    for 5 s): FP32 fitness against FP64. Muscle-driven bodies against joint
    limits are chaotic, so the gate is on the distribution: median below
    1e-3 m and 80 % of bodies within 1e-3 m (measured: median 4.5e-4 m,
    90th percentile 1.6e-3 m on 2,048 bodies)."""
    batch = E.synth(E.synth_spec(0, 1024))
    w = E.world(duration=5.0, substeps=8, ground=0, gravity=0.0)
    p, run = E.run_for(batch, w)
    r32, _, _ = oracle.run_f32(batch, p, run, threads=8)
    r64, _, _ = oracle.run_f64(batch, p, run, threads=8)
    same_end = (r32["ground_hi"].view(np.uint32) >> 16) == (r64["ground_hi"].view(np.uint32) >> 16)
    assert same_end.mean() > 0.97
    d = np.abs(r32["fitness"] - r64["fitness"])[same_end]
    assert np.median(d) < 1e-3 and (d < 1e-3).mean() > 0.8


def test_fp32_ranks_like_fp64_on_flat_ground():
    """Chaotic flat-ground trials: the FP32 fitness ranks the bodies as the
    FP64 fitness does (Spearman correlation)."""
    from scipy.stats import spearmanr

    batch = E.synth(E.synth_spec(0, 1024))
    w = E.world(duration=5.0, substeps=8)
    p, run = E.run_for(batch, w)
    r32, _, _ = oracle.run_f32(batch, p, run, threads=8)
    r64, _, _ = oracle.run_f64(batch, p, run, threads=8)
    assert spearmanr(r32["fitness"], r64["fitness"]).correlation > 0.95


def test_world_parameters():
    """src/kernel.rs:82-106, :212-251 and src/config.rs:636-638."""
    w = E.world(duration=20.0, substeps=8, air_retention=0.985)
    p, run = E.params_for(w, 0, 10)
    assert p.steps == 1200 and run.flags == N.GROUND | N.AIR and p.screen_step == 0xFFFFFFFF
    assert abs(np.float32(p.air_sub) ** 8 - np.float32(0.985)) < 1e-6
    w = E.world(duration=5.0, rate=120, screen_seconds=5.0, screen_bar=1.0)
    p, run = E.params_for(w, 0, 10)
    assert p.steps == 600 and p.screen_step == 599 and run.flags == N.GROUND
    w = E.world(ground=0, mud=0.05, slope=0.1, water=0.35, wind=-3.0)
    p, run = E.params_for(w, 0, 10)
    assert run.flags == N.WATER | N.WIND and p.mud == 0.0 and p.slope == 0.0
    w = E.world(quake=0.05)
    _, run = E.params_for(w, 0, 1)
    assert run.flags == N.GROUND | N.TERRAIN | N.QUAKE
