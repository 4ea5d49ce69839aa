"""Batched scenes, host side: the layout of the copies, the offset table, the
round trip through split, and the engine's check of the period (which runs
before any device call, so it is tested without a GPU)."""
import ctypes as C

import numpy as np
import pytest

from squishy_volumes_amd import abi, scenes
from squishy_volumes_amd.batch import Batch
from squishy_volumes_amd.io_state import make_consts


def uneven(sizes, h=1.0 / 64, dt=2e-4):
    out = []
    for i, n in enumerate(sizes):
        sc = scenes.snow_block_on_sphere(h=h, target_particles=n, dt=dt, seed=i, E=1.4e5 * (1 + 0.1 * i), nu=0.2)
        out.append(sc)
    return out


def test_layout_and_offsets():
    scs = uneven([500, 1728, 4000, 216, 1000])
    b = Batch(scs)
    h = scs[0].consts.grid_node_size
    assert b.repeat == (2, 2, 2)
    # the period: whole 4-node blocks, at least the domain plus the gap
    assert np.allclose(np.round(b.period / (4 * h)), b.period / (4 * h))
    assert np.all(b.period >= 1.0 + 16 * h)
    counts = [sc.io_state.particles.count for sc in scs]
    assert [s.count for s in b.slots] == counts
    assert [s.start for s in b.slots] == list(np.cumsum([0] + counts[:-1]))
    assert b.count == sum(counts)
    shifts = {tuple(np.round(s.shift / b.period).astype(int)) for s in b.slots}
    assert len(shifts) == len(scs)  # every scene in its own copy
    c = make_consts(b.scene.consts, b.scene.time_step, b.scene.options)
    assert list(c.scene_repeat) == [2, 2, 2]
    assert np.allclose(list(c.scene_period), b.period)


def test_split_inverts_the_combination():
    scs = uneven([216, 1000, 512])
    b = Batch(scs)
    parts = b.split(b.scene.io_state.particles)
    for sc, p in zip(scs, parts):
        q = sc.io_state.particles
        assert np.array_equal(p.flags, q.flags)
        assert np.array_equal(p.parameters, q.parameters)
        assert np.array_equal(p.velocities, q.velocities)
        # positions went through float32 at the shifted place and back
        assert np.abs(p.positions - q.positions).max() <= 4 * np.finfo(np.float32).eps * 4.0


def test_scenes_must_share_the_grid():
    a = scenes.snow_block_on_sphere(h=1.0 / 64, target_particles=500)
    b = scenes.snow_block_on_sphere(h=1.0 / 32, target_particles=500)
    with pytest.raises(ValueError):
        Batch([a, b])


def test_engine_rejects_a_period_that_overlaps_the_copies():
    lib = abi.load_library()
    sc = scenes.snow_block_on_sphere(h=1.0 / 64, target_particles=500)
    c = make_consts(sc.consts, sc.time_step, sc.options)
    c.scene_repeat[:] = [2, 1, 1]
    c.scene_period[:] = [1.0 + 2.0 / 64, 0.0, 0.0]  # a gap of 2 nodes: too small
    view = sc.io_state.particles.view()
    t = abi.MpmTuning()
    h = C.c_void_p()
    assert lib.mpm_create(C.byref(c), C.byref(view), C.byref(t), C.byref(h)) == abi.ERR_INVALID_ARGUMENT
    assert not h.value


def test_uneven_scene_set_sizes_and_placement():
    scs = scenes.uneven_scene_set(count=12, lo=2000, hi=60000, seed=5, h=1.0 / 64)
    for sc in scs:
        p = sc.io_state.particles
        assert 0.5 * 2000 <= p.count <= 1.5 * 60000
        wall = sc.options.box_walls
        assert np.all(p.positions > np.asarray(wall[0])) and np.all(p.positions < np.asarray(wall[1]))
        E = 2 * p.parameters["mu"] * 1.3
        assert np.all((E > 0.99e4) & (E < 1.01e6))
    b = Batch(scs)
    assert b.repeat == (3, 3, 2)
