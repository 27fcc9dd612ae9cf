"""MI355X creature simulator: the trial physics of
Schnellverstellzirkel/evolution_simulator_gpu (shaders/creature.cu) as
hand-written HIP kernels for gfx950 behind a C ABI (include/evo_engine.h).
"""
from . import _native
from .batch import (Batch, DeviceBatch, Graph, WorldTable, class_ranges, frame_stride, island_worlds, params_for,
                    run_for, simulate, synth, synth_spec, wild_levels, wild_world, world)

__all__ = ["Batch", "DeviceBatch", "Graph", "WorldTable", "class_ranges", "frame_stride", "island_worlds",
           "params_for", "run_for", "simulate", "synth", "synth_spec", "wild_levels", "wild_world", "world", "_native"]
