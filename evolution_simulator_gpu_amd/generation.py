"""One generation over many worlds and body sizes.

The reference runs the wild islands as one unit per world, each with a
kernel compiled for that world, and splits every unit's launches by body size
(src/ring.rs:117-160, src/cuda_engine.rs:889-947, :2171-2311). Here a
generation is one class-major packed batch in which every creature reads its
own world from a table; each body-size class runs as one persistent kernel,
the classes run concurrently (they share the CUs), and the whole generation
is one HIP graph. The share of the GPU each class gets is tuned by timing a
few grid splits on the generation's own batch with short trials.

`PerShapeGeneration` runs the same creatures the reference's way, one launch
per (world, body-size class) over up to 8 streams, as the baseline the
batched generation is measured against.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _native as N
from .batch import Batch, DeviceBatch, Graph, WorldTable, body_classes, class_ranges, params_for, run_for


def _events(torch, n):
    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


class IslandGeneration:
    """A class-major batch (synthetic spec sorted = 2) whose packed creature c
    runs in worlds[world_of[c]]."""

    def __init__(self, batch: Batch, worlds: list, world_of: np.ndarray, device: str = "cuda"):
        import torch

        self.batch = batch
        self.worlds = worlds
        self.world_of = np.asarray(world_of)
        self.table = WorldTable(worlds, self.world_of, batch.count, 0)
        self.db = DeviceBatch(batch, device)
        self.db.set_worlds(self.table)
        self.p, self.run = self.table.run_for(batch)
        self.ranges = class_ranges(batch)
        # one world for every creature: the world's parameters go to the
        # kernels as launch arguments (the any-world instances), not through
        # the table (1.13x on the full-stack batch, results identical)
        self.table_worlds = len(worlds) > 1
        self.stream = torch.cuda.Stream()
        self.graph: Graph | None = None
        self.grids = tuple(0 for _ in self.ranges)

    def _ranges(self, grids):
        return [r[:5] + (int(g),) for r, g in zip(self.ranges, grids)]

    def capacity(self) -> list[int]:
        """Blocks of each class's kernel that fit on the GPU at once."""
        out = []
        for r in self.ranges:
            q = N.Run.from_buffer_copy(self.run)
            q.max_nodes, q.max_muscles, q.no_tendons = r[2], r[3], r[4]
            out.append(int(N.lib().evo_kernel_capacity(C.byref(q), int(self.table_worlds))))
        return out

    def launch(self, grids=None, p: N.Params | None = None) -> None:
        self.db.launch_ranges(p or self.p, self.run, self._ranges(grids or self.grids), worlds=self.table_worlds,
                              stream=self.stream)

    def tune(self, seconds: float = 2.0, candidates=None, progress=None) -> dict:
        """Times each candidate split of the GPU between the classes on this
        batch with trials cut to `seconds`, keeps the fastest, and returns
        the timings. A candidate is one grid (blocks) per class; 0 means as
        many as fit."""
        import torch

        caps = self.capacity()
        if candidates is None:
            if len(caps) == 1:
                candidates = [(0,)]
            elif len(caps) == 2:
                m, b = caps
                candidates = [(0, 0), (m // 2, b // 2), (m // 4, (3 * b) // 4), ((3 * m) // 4, b // 4),
                              (m // 2, b), (m // 4, b), (m, b // 2)]
            else:
                # shares of the GPU, each class's grid that share of the
                # blocks of it that fit alone; all 0: every kernel as large
                # as fits, one after another as blocks free up
                k = len(caps)
                shares = [tuple(1 / k for _ in caps)]
                for big in range(k):
                    shares.append(tuple(0.5 if i == big else 0.5 / (k - 1) for i in range(k)))
                candidates = [tuple(0 for _ in caps)]
                candidates += [tuple(max(1, int(f * c)) for f, c in zip(share, caps)) for share in shares]
        short = [N.WorldConfig.from_buffer_copy(w) for w in self.worlds]
        for w in short:
            w.duration = seconds
        # every creature reads its trial length from its world's record, so
        # the cut trials run from a table of their own
        short_table = WorldTable(short, self.world_of, self.batch.count, 0)
        p_short, _ = short_table.run_for(self.batch)
        self.db.set_worlds(short_table)
        timings = {}
        try:
            for cand in candidates:
                self.launch(cand, p_short)  # warm the instance
                e0, e1 = _events(torch, 2)
                e0.record(self.stream)
                self.launch(cand, p_short)
                e1.record(self.stream)
                torch.cuda.synchronize()
                timings[tuple(int(x) for x in cand)] = e0.elapsed_time(e1)
                if progress is not None:
                    progress(f"grid split {tuple(int(x) for x in cand)}: {timings[tuple(int(x) for x in cand)]:.0f} ms")
        finally:
            torch.cuda.synchronize()
            self.db.set_worlds(self.table)
        best = min(timings, key=timings.get)
        self.grids = best
        self.graph = None
        return {"capacity": caps, "seconds": seconds, "ms": {str(k): v for k, v in timings.items()}, "best": best}

    def capture(self) -> None:
        """Records the generation's launches into a HIP graph."""
        import torch

        self.launch()  # creates the side streams before the capture
        torch.cuda.synchronize()
        g = Graph()
        g.capture(lambda s: self.db.launch_ranges(self.p, self.run, self._ranges(self.grids),
                                                  worlds=self.table_worlds, stream=s), self.stream)
        self.graph = g

    def step(self) -> None:
        if self.graph is not None:
            self.graph.replay(self.stream)
        else:
            self.launch()

    def results(self) -> np.ndarray:
        return self.db.results_numpy()


class PerShapeGeneration:
    """The same creatures as one launch per (world, body-size class), each
    with its world's launch parameters, spread over `streams` streams."""

    def __init__(self, batch: Batch, worlds: list, world_of: np.ndarray, device: str = "cuda", streams: int = 8):
        import torch

        world_of = np.asarray(world_of)
        n = batch.nodes().astype(np.int64)
        m = batch.muscle_counts().astype(np.int64)
        # the body-size classes of class_ranges, bodies over 12 nodes as one
        # shape: a per-world launch of the few bodies over 20 nodes would be
        # a latency-bound launch of its own
        cls = np.minimum(body_classes(n, m), 2)
        # packed order within each (world, class) group is kept
        self.order = np.lexsort((np.arange(batch.count), cls, world_of))
        self.batch = batch.subset(self.order)
        self.db = DeviceBatch(self.batch, device)
        self.db.counter = torch.zeros(streams, dtype=torch.int32, device=device)
        wo, cl = world_of[self.order], cls[self.order]
        key = wo * 3 + cl
        cuts = np.concatenate([[0], np.nonzero(np.diff(key))[0] + 1, [batch.count]])
        stride = 0
        self.launches = []
        nn, mm = self.batch.nodes(), self.batch.muscle_counts()
        for a, b in zip(cuts[:-1], cuts[1:]):
            w = worlds[int(wo[a])]
            p, run = params_for(w, int(a), int(b - a), stride)
            run.max_nodes, run.max_muscles = int(nn[a:b].max()), int(mm[a:b].max())
            lo = int(self.batch.heads[a, 4])
            hi = int(self.batch.heads[b - 1, 4]) + int(mm[b - 1]) * N.MUSCLE_WORDS
            run.no_tendons = int(np.all(self.batch.muscles[lo:hi].reshape(-1, N.MUSCLE_WORDS)[:, 6] == 0.0))
            self.launches.append((p, run))
        self.streams = [torch.cuda.Stream() for _ in range(streams)]

    def step(self, progress=None) -> None:
        """Queues every launch; with `progress`, waits for them, reporting
        every 20 s how many have finished (without serializing them)."""
        import time as _time

        import torch

        lib = N.lib()
        marks = []
        for k, (p, run) in enumerate(self.launches):
            i = k % len(self.streams)
            rc = lib.evo_advance_device(self.db.records.data_ptr(), self.db.muscles.data_ptr(),
                                        self.db.heads.data_ptr(), self.db.results.data_ptr(),
                                        self.db.counter.data_ptr() + 4 * i, C.byref(p), C.byref(run), None,
                                        C.c_void_p(self.streams[i].cuda_stream))
            if rc != N.EVO_OK:
                raise RuntimeError(f"evo_advance_device failed: {rc}")
            if progress is not None:
                ev = torch.cuda.Event()
                ev.record(self.streams[i])
                marks.append(ev)
        if progress is not None:
            last = _time.perf_counter()
            while True:
                done = sum(1 for ev in marks if ev.query())
                if done == len(marks):
                    break
                if _time.perf_counter() - last > 20.0:
                    progress(f"per-shape launches finished: {done} of {len(marks)}")
                    last = _time.perf_counter()
                _time.sleep(0.05)

    def results_in_packed_order(self) -> np.ndarray:
        """Results mapped back to the order of the batch it was built from."""
        res = self.db.results_numpy()
        out = np.empty_like(res)
        out[self.order] = res
        return out


def timed(torch, fn, steps: int) -> float:
    """Seconds per call of fn over `steps` calls, device-synchronized."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps
