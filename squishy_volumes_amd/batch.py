"""Batched scenes: many independent scenes advanced together by one engine.

The reference runs one simulation per engine.  Here a batch of scenes that
share the node spacing, the time step, the domain and the engine options
(their particles, materials and sizes differ) runs as one: each scene gets
its own copy of the domain in a periodic layout (copy (i, j, k) shifted by
(i, j, k) * period), the engine's grid covers all copies, and every kernel
of a substep is one launch over all of them.  The sort's Morton block list
spans the scenes, so work is balanced per block, whatever the scene sizes.
The cull, walls, plane and sphere act on positions within each particle's
own copy (mpm_consts.scene_period / scene_repeat), and the copies are
separated by a gap wider than the stencil, so scenes never interact.

The offset table maps scene s to its particle range [start, start + count)
in the batch and to the shift of its copy.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass

import numpy as np

from .io_state import FrameInput, IoState, Particles
from .scenes import Scene, concat


@dataclass
class SceneSlot:
    start: int
    count: int
    shift: np.ndarray  # (3,) added to the scene's positions in the batch


class Batch:
    def __init__(self, scenes: list[Scene], gap_cells: int = 16):
        if not scenes:
            raise ValueError("empty batch")
        first = scenes[0]
        for sc in scenes[1:]:
            if (sc.consts != first.consts or sc.time_step != first.time_step or sc.options != first.options
                    or sc.io_state.time != first.io_state.time):
                raise ValueError("batched scenes must share consts, time step, options and time")
        self.scenes = scenes
        h = first.consts.scaled_grid_node_size()
        lo = np.asarray(first.consts.scaled_domain_min(), np.float64)
        hi = np.asarray(first.consts.scaled_domain_max(), np.float64)
        # the period: a whole number of 4-node blocks wider than the domain by
        # the gap, so that every copy has the same block alignment
        width = hi - lo + gap_cells * h
        self.period = np.ceil(width / (4 * h)) * 4 * h
        side = max(1, math.ceil(len(scenes) ** (1.0 / 3.0) - 1e-9))
        self.repeat = (side, side, math.ceil(len(scenes) / (side * side)))
        self.slots: list[SceneSlot] = []
        parts = []
        start = 0
        for i, sc in enumerate(scenes):
            cell = np.array([i % side, (i // side) % side, i // (side * side)], np.float64)
            shift = cell * self.period
            p = dataclasses.replace(sc.io_state.particles)
            p.positions = (sc.io_state.particles.positions.astype(np.float64) + shift).astype(np.float32)
            p.initial_positions = (sc.io_state.particles.initial_positions.astype(np.float64) + shift).astype(np.float32)
            parts.append(p)
            self.slots.append(SceneSlot(start, p.count, shift))
            start += p.count
        particles = concat(parts)
        self.frame_input = self._frame_input([sc.frame_input for sc in scenes])
        options = dataclasses.replace(first.options, scene_period=tuple(float(v) for v in self.period),
                                      scene_repeat=tuple(int(v) for v in self.repeat))
        self.scene = dataclasses.replace(first, name=f"batch_of_{len(scenes)}", options=options,
                                         io_state=IoState(time=first.io_state.time, particles=particles),
                                         frame_input=self.frame_input,
                                         info={"scenes": len(scenes), "particles": particles.count,
                                               "counts": [s.count for s in self.slots]})

    def _frame_input(self, inputs: list[FrameInput]) -> FrameInput:
        """One FrameInput for the batch: the scenes' gravity (which must agree)
        and their per-particle flags and goals, concatenated, goals shifted
        into each scene's copy."""
        f0 = inputs[0]
        for f in inputs[1:]:
            if f.frame != f0.frame or f.gravity_a != f0.gravity_a or f.gravity_b != f0.gravity_b:
                raise ValueError("batched scenes must share the frame and gravity")
        out = dataclasses.replace(f0)
        for name in ("particle_flags_a", "particle_flags_b", "goal_positions_a", "goal_positions_b"):
            arrs = [getattr(f, name) for f in inputs]
            if all(a is None for a in arrs):
                continue
            if any(a is None for a in arrs):
                raise ValueError(f"{name} is given for some scenes of the batch only")
            if name.startswith("goal"):
                arrs = [(np.asarray(a, np.float64).reshape(-1, 3) + slot.shift).astype(np.float32)
                        for a, slot in zip(arrs, self.slots)]
            setattr(out, name, np.ascontiguousarray(np.concatenate(arrs)))
        return out

    @property
    def count(self) -> int:
        return self.scene.io_state.particles.count

    def split(self, particles: Particles) -> list[Particles]:
        """The batch's particles back as one Particles per scene, in scene coordinates."""
        out = []
        for slot in self.slots:
            sl = slice(slot.start, slot.start + slot.count)
            p = Particles.empty(slot.count, with_jp=particles.plastic_jp is not None)
            for name in p.__dataclass_fields__:
                arr = getattr(particles, name)
                if arr is not None:
                    setattr(p, name, np.ascontiguousarray(arr[sl]))
            p.positions = (p.positions.astype(np.float64) - slot.shift).astype(np.float32)
            p.initial_positions = (p.initial_positions.astype(np.float64) - slot.shift).astype(np.float32)
            out.append(p)
        return out
