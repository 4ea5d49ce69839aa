"""Batched scenes on the GPU: every scene of a batch advances as it does alone.

Stated tolerances: per scene, x / v / F after 120 substeps within 1e-4
(norm-wise relative) of the same scene run alone on the engine and of the
FP64 oracle, Jp within rtol 1e-4; culling in each scene's own copy of the
domain matches the single run (up to 2 particles that sit within float
round-off of the boundary, since a scene's positions carry the copy's
shift)."""
import numpy as np
import pytest

from squishy_volumes_amd import abi, scenes
from squishy_volumes_amd.batch import Batch
from squishy_volumes_amd.gpu_state import GpuRunParameters
from tests.helpers import rel_norm_err
from tests.test_batch import uneven
from tests.test_gpu_step import compare_states, engine, oracles

pytestmark = pytest.mark.gpu

SUBSTEPS = 120


def test_uneven_snow_scenes_in_one_batch_match_their_single_runs():
    scs = uneven([500, 1728, 4000, 216, 1000])
    b = Batch(scs)
    T = scenes.advance_time(0.0, b.scene.time_step, SUBSTEPS)
    g = engine(b.scene)
    st, err = g.produce_next_state(b.frame_input, GpuRunParameters(target_time=T))
    assert err is None
    parts = b.split(st.particles)
    for i, (sc, p) in enumerate(zip(scs, parts)):
        alone, err = engine(sc).produce_next_state(sc.frame_input, GpuRunParameters(target_time=T))
        assert err is None
        errs = {k: rel_norm_err(a, r) for k, a, r in (("x", p.positions, alone.particles.positions),
                                                      ("v", p.velocities, alone.particles.velocities),
                                                      ("F", p.position_gradients, alone.particles.position_gradients))}
        print(f"scene {i} ({sc.io_state.particles.count} particles) vs alone", {k: f"{v:.2e}" for k, v in errs.items()})
        assert max(errs.values()) <= 1e-4
        assert np.allclose(p.plastic_jp, alone.particles.plastic_jp, rtol=1e-4)
        # the scene moved: the sphere and walls act in each copy
        assert np.abs(p.positions - sc.io_state.particles.positions).max() > 1e-2
    for i in (0, 3):
        _, o64 = oracles(scs[i])
        o64.step_to(T, scs[i].frame_input)
        compare_states(parts[i], o64.particles_f64(), scs[i].consts.grid_node_size, f"batched scene {i} vs fp64")


def test_culling_acts_in_each_scene_copy():
    vel = [(30.0, 0.0, 0.0), (0.0, -30.0, 0.0), (0.0, 0.0, 30.0), (-30.0, 0.0, 0.0)]
    corner = [(56, 28, 28), (28, 4, 28), (28, 28, 56), (4, 28, 28)]
    scs = [scenes.cube_scene(n_cells=4 + i, corner_cells=c, velocity=v, dt=1e-4, seed=i, gravity=(0.0, 0.0, 0.0))
           for i, (c, v) in enumerate(zip(corner, vel))]
    b = Batch(scs)
    T = scenes.advance_time(0.0, b.scene.time_step, 40)
    g = engine(b.scene)
    st, err = g.produce_next_state(b.frame_input, GpuRunParameters(target_time=T))
    assert err is None
    parts = b.split(st.particles)
    for i, (sc, p) in enumerate(zip(scs, parts)):
        alone, _ = engine(sc).produce_next_state(sc.frame_input, GpuRunParameters(target_time=T))
        dead = (p.flags & abi.TOMBSTONED) != 0
        dead1 = (alone.particles.flags & abi.TOMBSTONED) != 0
        print(f"scene {i}: {int(dead.sum())} of {p.count} culled (alone {int(dead1.sum())})")
        assert dead1.sum() > 0 and (~dead1).sum() > 0
        assert int((dead != dead1).sum()) <= 2
        live = ~dead & ~dead1
        assert rel_norm_err(p.positions[live], alone.particles.positions[live]) <= 1e-4


def test_autotuner_picks_the_fastest_measured_tuning_and_caches_it(tmp_path):
    from squishy_volumes_amd import autotune

    b = Batch(uneven([500, 4000, 1728]))
    cache = tmp_path / "tune.json"
    r = autotune.autotune(b.scene, resort_intervals=(4, 8), persistent_blocks=(0, 64), particles_per_lane=(3,),
                          grid_update_threads=(64, 256), graph=(0,), warmup=2, substeps=6, cache=cache)
    assert len(r.trials) == 8 and not r.cached
    assert r.ms_per_substep == min(ms for _, ms in r.trials)
    assert r.best in [t for t, _ in r.trials]
    again = autotune.autotune(b.scene, cache=cache)
    assert again.cached and again.best == r.best
    # any tuning gives the same physics up to round-off
    T = scenes.advance_time(0.0, b.scene.time_step, 30)
    runs = []
    for tuning in ({"resort_interval": 4, "persistent_blocks": 64, "grid_update_threads": 64},
                   {"resort_interval": 16, "persistent_blocks": 2048, "grid_update_threads": 128}):
        st, err = engine(b.scene, tuning=tuning).produce_next_state(b.frame_input, GpuRunParameters(target_time=T))
        assert err is None
        runs.append(st.particles)
    assert rel_norm_err(runs[0].positions, runs[1].positions) <= 1e-5
    assert rel_norm_err(runs[0].velocities, runs[1].velocities) <= 1e-4


def test_batched_scenes_with_one_learned_closure_match_their_single_runs():
    """The combined form at small scale: one network (trained on pairs the
    engine generates) shared by a batch of uneven snow scenes gives each
    scene what it gives the scene alone (x, v, F within 1e-4), and online
    fine-tuning from the batch's own pairs keeps the judged error small."""
    import torch

    from squishy_volumes_amd import closure

    scs = uneven([500, 1728, 216])
    b = Batch(scs)
    hard = b.scene.options.snow_hardening
    g = engine(b.scene)
    g.step_to(scenes.advance_time(0.0, b.scene.time_step, 40), b.frame_input)
    x, t = g.closure_pairs(stride=1)
    net = closure.Mlp(seed=4)
    net.fit(x, t, epochs=150, batch=512, lr=3e-3, lr_final=1e-4)
    hull = closure.ClosureModel.hull_of(x)
    hull = (hull[0] - 0.5, hull[1] + 0.5)
    T = scenes.advance_time(0.0, b.scene.time_step, 60)
    gb = engine(b.scene)
    gb.set_closure(net.device_model(hull), judge_stride=2)
    st, err = gb.produce_next_state(b.frame_input, GpuRunParameters(target_time=T))
    assert err is None
    stats = gb.closure_stats()
    print("batch", stats)
    assert stats["evaluated"] > 0 and stats["relative_stress_error"] <= 1e-2
    for i, (sc, p) in enumerate(zip(scs, b.split(st.particles))):
        e = engine(sc)
        e.set_closure(net.device_model(hull), judge_stride=2)
        alone, err = e.produce_next_state(sc.frame_input, GpuRunParameters(target_time=T))
        assert err is None
        errs = [rel_norm_err(a, r) for a, r in ((p.positions, alone.particles.positions),
                                                (p.velocities, alone.particles.velocities),
                                                (p.position_gradients, alone.particles.position_gradients))]
        print(f"scene {i} with the closure, batch vs alone: x {errs[0]:.1e} v {errs[1]:.1e} F {errs[2]:.1e}")
        assert max(errs) <= 1e-4
    # one round of online fine-tuning on the batch's current pairs
    x2, t2 = gb.closure_pairs(stride=1)
    net.fit(x2, t2, epochs=5, batch=512, lr=3e-4, lr_final=1e-4, seed=1)
    gb.set_closure(net.device_model(closure.ClosureModel.hull_union(hull, closure.ClosureModel.hull_of(x2))),
                   judge_stride=2)
    gb.step_to(scenes.advance_time(gb.time, b.scene.time_step, 20), b.frame_input)
    stats2 = gb.closure_stats()
    print("after fine-tuning", stats2)
    assert stats2["evaluated"] > 0 and stats2["relative_stress_error"] <= 1e-2
    torch.cuda.synchronize()
