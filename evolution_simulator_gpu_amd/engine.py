"""The submission engine (include/evo_engine.h, evo_open .. evo_poll): the
device side of the reference's `Device` trait (src/engine.rs:430-467) as a
Python object over the C ABI. Submissions copy their pack before returning;
results arrive, in the caller's order, in the array given at submit time
when the poll that reports the ticket returns."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .batch import Batch


class EngineError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{message} (code {code})")
        self.code = code

    @property
    def out_of_memory(self) -> bool:
        return self.code == N.EVO_ERR_OOM


def _pack(batch: Batch, order: bool = True) -> N.Pack:
    p = N.Pack()
    p.records, p.record_words = batch.records.ctypes.data, batch.records.size
    p.muscles, p.muscle_floats = batch.muscles.ctypes.data, batch.muscles.size
    p.heads, p.count = batch.heads.ctypes.data, batch.count
    p.slots = batch.slots.ctypes.data if order else None
    return p


class Engine:
    def __init__(self, device: int = 0, standard_slots: int = 24, big_unit: int = 8192, big_slots: int = 4,
                 device_budget: int = 0):
        cfg = N.EngineConfig()
        N.lib().evo_engine_default(C.byref(cfg))
        cfg.device, cfg.standard_slots, cfg.big_unit, cfg.big_slots = device, standard_slots, big_unit, big_slots
        cfg.device_budget = device_budget
        self.handle = C.c_void_p(None)
        rc = N.lib().evo_open(C.byref(cfg), C.byref(self.handle))
        if rc != N.EVO_OK:
            raise EngineError(rc, "evo_open failed")
        self._pending: dict[int, tuple] = {}

    def _check(self, rc: int, what: str) -> None:
        if rc != N.EVO_OK:
            raise EngineError(rc, f"{what}: {N.lib().evo_last_error(self.handle).decode()}")

    def free_slots(self, creatures: int | None = None) -> int:
        lib = N.lib()
        return lib.evo_free_slots(self.handle) if creatures is None else lib.evo_free_slots_for(self.handle, creatures)

    def confirm_free(self) -> bool:
        return bool(N.lib().evo_confirm_free(self.handle))

    def replay_free(self) -> bool:
        return bool(N.lib().evo_replay_free(self.handle))

    def submit(self, batch: Batch, world: N.WorldConfig, confirm: bool = False) -> int:
        """Queues the unit; returns its ticket. The results array (caller
        order: batch.slots) comes back from poll."""
        out = np.zeros(batch.count, N.RESULT_DTYPE)
        ticket = C.c_uint64(0)
        rc = N.lib().evo_submit(self.handle, C.byref(_pack(batch)), C.byref(world),
                                N.SLOT_CONFIRM if confirm else N.SLOT_STANDARD, out.ctypes.data, C.byref(ticket))
        self._check(rc, "evo_submit")
        self._pending[ticket.value] = (out, None)
        return ticket.value

    def record(self, batch: Batch, world: N.WorldConfig) -> int:
        pack = _pack(batch)
        n = N.lib().evo_record_frame_floats(C.byref(pack), C.byref(world))
        out = np.zeros(batch.count, N.RESULT_DTYPE)
        frames = np.zeros(n, np.float32)
        ticket = C.c_uint64(0)
        rc = N.lib().evo_record(self.handle, C.byref(pack), C.byref(world), out.ctypes.data, frames.ctypes.data, n,
                                C.byref(ticket))
        self._check(rc, "evo_record")
        self._pending[ticket.value] = (out, frames)
        return ticket.value

    def poll(self, timeout_us: int = 0):
        """(ticket, results, frames or None, gpu_seconds) of a finished
        submission, or None when none finished within the timeout."""
        done = N.Completion()
        rc = N.lib().evo_poll(self.handle, timeout_us, C.byref(done))
        if rc == N.EVO_ERR_NOT_READY:
            return None
        self._check(rc, "evo_poll")
        out, frames = self._pending.pop(done.ticket)
        return done.ticket, out, frames, done.gpu_seconds

    def release_idle(self) -> int:
        return N.lib().evo_release_idle(self.handle)

    def allocated_bytes(self) -> int:
        return N.lib().evo_allocated_bytes(self.handle)

    def close(self) -> None:
        if self.handle and self.handle.value:
            N.lib().evo_close(self.handle)
            self.handle = C.c_void_p(None)

    def __del__(self):
        self.close()
