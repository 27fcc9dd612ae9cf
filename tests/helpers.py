"""Shared helpers of the tests: batches, worlds and bit-exact comparison."""
from __future__ import annotations

import numpy as np

import evolution_simulator_gpu_amd as E
from evolution_simulator_gpu_amd import _native as N


def bits(results: np.ndarray) -> np.ndarray:
    """The 19 words of each result as uint32, with every NaN made one pattern."""
    words = results.view(np.float32).reshape(-1, 19).copy()
    nan = np.isnan(words)
    out = words.view(np.uint32).copy()
    out[nan] = 0x7FC00000
    return out


def assert_bit_exact(got: np.ndarray, want: np.ndarray, label: str = "") -> None:
    g, w = bits(got), bits(want)
    if not np.array_equal(g, w):
        rows, cols = np.nonzero(g != w)
        first = rows[0]
        fields = sorted({N.RESULT_FIELDS[c] for c in cols[rows == first]})
        raise AssertionError(
            f"{label}: {len(np.unique(rows))} of {len(g)} results differ; first creature {first} in {fields}: "
            f"got {got[first]} want {want[first]}")


def config0_batch(count: int | None = None):
    spec = E.synth_spec(0, count)
    return E.synth(spec)


def flat_world(duration: float = 5.0, substeps: int = 8, rate: int = 60, **kw):
    return E.world(duration=duration, substeps=substeps, rate=rate, **kw)
