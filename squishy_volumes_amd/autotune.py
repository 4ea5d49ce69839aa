"""Autotuner of the engine's launch knobs, measured on the workload itself.

The knobs are those of mpm_tuning (include/mpm_engine.h): the re-sort
interval (substeps between particle sorts: a longer interval saves sorts, a
shorter one keeps particles inside their blocks' staged node regions), the
work-group count of the persistent transfer kernels (all resident work-groups
by default; fewer leave CUs to the grid update's tail, more over-subscribe
the CUs and balance uneven blocks), the particles each lane takes per pass of
a block (fewer make more, smaller passes: less idle tail, more flushes), the
work-group size of the grid update (64, 128 or 256 threads over the same 16
waves per CU: smaller groups retire unevenly loaded waves independently) and
whether whole sort cycles run as captured HIP graphs (they measure alike at
these sizes, so the default search keeps them on).  Each candidate runs the scene on a fresh
engine for a few warm-up and a few timed substeps; the fastest wins.  Results
are cached in a JSON file keyed by the workload's signature (device, particle
count, node spacing, number of batched scenes), so a workload is tuned once.
The reference has no tuner (its kernels have fixed launch shapes).
"""
from __future__ import annotations

import itertools
import json
import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional

RESORT_INTERVALS = (4, 8, 16)
PERSISTENT_BLOCKS = (0, 2048)
PARTICLES_PER_LANE = (1, 2, 3)
GRID_UPDATE_THREADS = (64, 256)
GRAPH = (0,)  # 0: sort cycles as HIP graphs (the engine's default), -1: every kernel launched


@dataclass
class TuneResult:
    best: dict
    ms_per_substep: float
    trials: list = field(default_factory=list)  # (tuning, ms per substep)
    cached: bool = False


def signature(scene, device_name: str = "") -> str:
    n = scene.io_state.particles.count
    reps = scene.options.scene_repeat or (1, 1, 1)
    return (f"{device_name}|n={n}|h={scene.consts.scaled_grid_node_size():.6g}|dt={scene.time_step:.6g}"
            f"|repeat={tuple(int(r) for r in reps)}")


def time_tuning(scene, tuning: dict, warmup: int, substeps: int, device: int = 0) -> float:
    """Milliseconds per substep of `scene` under `tuning` (wall clock around
    step_to, which returns after the engine's stream has drained)."""
    from . import scenes as scenes_mod
    from .gpu_state import GpuState

    g = GpuState.from_io_state(device, scene.consts, scene.time_step, scene.io_state, scene.options, tuning)
    try:
        fi = scene.frame_input
        if warmup:
            g.step_to(scenes_mod.advance_time(g.time, scene.time_step, warmup), fi)
        t0 = time.perf_counter()
        st = g.step_to(scenes_mod.advance_time(g.time, scene.time_step, substeps), fi)
        el = time.perf_counter() - t0
        if st.result != 0:
            raise RuntimeError(f"step failed under tuning {tuning}: result {st.result}")
        return el / substeps * 1e3
    finally:
        g.close()


def autotune(scene, *, resort_intervals=RESORT_INTERVALS, persistent_blocks=PERSISTENT_BLOCKS,
             particles_per_lane=PARTICLES_PER_LANE, grid_update_threads=GRID_UPDATE_THREADS, graph=GRAPH,
             warmup: int = 8, substeps: int = 48,
             cache: Optional[Path] = None, device: int = 0, log=None) -> TuneResult:
    """The fastest (resort_interval, persistent_blocks, particles_per_lane,
    grid_update_threads, graph) for `scene`: a grid search over the
    candidates, timing each on the scene itself."""
    device_name = ""
    try:
        import torch

        device_name = torch.cuda.get_device_name(device)
    except Exception:
        pass
    key = signature(scene, device_name)
    if cache is not None and Path(cache).exists():
        table = json.loads(Path(cache).read_text())
        if key in table:
            e = table[key]
            return TuneResult(e["best"], e["ms_per_substep"], e.get("trials", []), cached=True)
    trials = []
    for ri, pb, pl, gu, gr in itertools.product(resort_intervals, persistent_blocks, particles_per_lane,
                                                 grid_update_threads, graph):
        tuning = {"resort_interval": ri, "persistent_blocks": pb, "particles_per_lane": pl,
                  "grid_update_threads": gu, "graph": gr}
        ms = time_tuning(scene, tuning, warmup, substeps, device)
        trials.append((tuning, ms))
        if log:
            log(f"autotune: {tuning} {ms:.4f} ms/substep")
    best, ms = min(trials, key=lambda t: t[1])
    if cache is not None:
        p = Path(cache)
        table = json.loads(p.read_text()) if p.exists() else {}
        table[key] = {"best": best, "ms_per_substep": ms, "trials": trials}
        p.write_text(json.dumps(table, indent=1))
    return TuneResult(best, ms, trials)
