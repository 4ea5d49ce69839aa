"""Frame files of a stepped engine (IoState::write / IoState::read,
rust/crates/file_frame/src/io_state.rs:31-75), through the C ABI.

A cube dropped onto a floor mesh is stepped until its particles carry
collider bits, velocity gradients and elastic energies; its state, read
with mpm_read_particles and mpm_read_grid, is written as the reference's
frame file and decoded back: every particle field, the time and the grid
nodes come back bit for bit.  An engine created from the decoded state then
continues as the stepped engine does (x, v, F within 1e-5: the two differ
only in the order of the scatter's atomic sums).
"""
import dataclasses

import numpy as np
import pytest

from squishy_volumes_amd import scenes
from squishy_volumes_amd.gpu_state import GpuRunParameters, GpuState
from squishy_volumes_amd.io_state import IoState
from tests.helpers import rel_norm_err
from tests.test_gpu_colliders import DT, consts, cube_parts, floor

pytestmark = pytest.mark.gpu


def test_frame_file_of_a_stepped_engine_round_trips(tmp_path):
    parts = cube_parts((12, 12, 10), 5, (0.3, 0.0, -2.0))
    mesh, fi = floor(0.2, friction=0.6)
    c = consts()
    g = GpuState.from_io_state(0, c, DT, IoState(0.0, parts))
    g.set_collider_mesh(mesh)
    T = scenes.advance_time(0.0, DT, 100)
    st, err = g.produce_next_state(fi, GpuRunParameters(target_time=T, store_grid=True))
    assert err is None, err
    p = st.particles
    assert (p.collider_bits != 0).sum() > 50 and np.abs(p.velocity_gradients).max() > 0
    assert np.abs(p.elastic_energies).max() > 0 and st.grid_nodes is not None and st.grid_nodes.masses.size > 0

    path = tmp_path / "frame_000100.bin"
    size = st.write(path)
    assert path.stat().st_size == size
    back = IoState.read(path)
    assert back.time == st.time == g.time
    for f in dataclasses.fields(p):
        a, b = getattr(p, f.name), getattr(back.particles, f.name)
        if a is None:
            assert b is None, f.name
            continue
        assert a.dtype == b.dtype and a.shape == b.shape, f.name
        assert a.tobytes() == b.tobytes(), f.name
    for f in dataclasses.fields(st.grid_nodes):
        a, b = getattr(st.grid_nodes, f.name), getattr(back.grid_nodes, f.name)
        assert a.tobytes() == b.tobytes(), f.name

    # an engine created from the decoded frame continues as the stepped one
    h = GpuState.from_io_state(0, c, DT, back)
    h.set_collider_mesh(mesh)
    T2 = scenes.advance_time(T, DT, 20)
    s1, e1 = g.produce_next_state(fi, GpuRunParameters(target_time=T2))
    s2, e2 = h.produce_next_state(fi, GpuRunParameters(target_time=T2))
    assert e1 is None and e2 is None and s1.time == s2.time
    for k, a, b in (("x", s1.particles.positions, s2.particles.positions),
                    ("v", s1.particles.velocities, s2.particles.velocities),
                    ("F", s1.particles.matrix_F(), s2.particles.matrix_F())):
        err = rel_norm_err(a, b.astype(np.float64))
        print(f"continued from the frame file, {k}: {err:.2e}")
        assert err <= 1e-5, (k, err)
    assert np.array_equal(s1.particles.collider_bits, s2.particles.collider_bits)
