"""MI355X creature simulator: the trial physics of
Schnellverstellzirkel/evolution_simulator_gpu (shaders/creature.cu) as
hand-written HIP kernels for gfx950 behind a C ABI (include/evo_engine.h).
"""
from . import _native
from .batch import Batch, DeviceBatch, frame_stride, params_for, run_for, simulate, synth, synth_spec, world

__all__ = ["Batch", "DeviceBatch", "frame_stride", "params_for", "run_for", "simulate", "synth", "synth_spec",
           "world", "_native"]
