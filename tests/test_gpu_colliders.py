"""Mesh colliders on the GPU against the oracle's restatement of the collide
phase and the twin-node meld (rust/crates/cpu/src/phase/collide.rs:21-206,
meld_grid.rs:16-69), through the C ABI.

  * closed forms on a lone particle (the oracle's pins, tests/test_oracle_colliders.py):
    the bits set within the accept distance, and a crossing's push-back with
    friction and damping;
  * a cube dropped onto a floor mesh, through the contact, against the FP64
    and the float oracle (x, v, F; collider bits equal to the float oracle's
    except for particles within float round-off of the 2h / 2.2h thresholds);
  * two blocks thrown at each other through a thin sheet: the sheet keeps
    them apart, as in the oracle;
  * a moving mesh (frames a and b differ) with friction and damping;
  * a table of twin nodes too small for the scene: the call is redone with a
    larger one and matches the oracle.
Stated tolerance: 1e-4 norm-wise relative on x, v, F against FP64 (the
substep tolerance of the other parity tests), measured errors printed.
"""
import dataclasses

import numpy as np
import pytest

from oracle import pyoracle
from squishy_volumes_amd import abi, scenes
from squishy_volumes_amd.gpu_state import GpuRunParameters, GpuState
from squishy_volumes_amd.io_state import FrameInput, InputConsts, IoState, Particles
from tests.helpers import rel_norm_err
from tests.test_oracle_colliders import lone

pytestmark = pytest.mark.gpu

H = 1.0 / 32
DT = 1e-3


def consts(lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), h=H, fps=1):
    return InputConsts(grid_node_size=h, frames_per_second=fps, domain_min=lo, domain_max=hi)


def run_both(c, parts, mesh, fi, substeps, dt=DT, precision_pairs=(8, 4)):
    """The engine and the oracles over `substeps` substeps (one call each)."""
    T = scenes.advance_time(0.0, dt, substeps)
    g = GpuState.from_io_state(0, c, dt, IoState(0.0, parts))
    g.set_collider_mesh(mesh)
    st, err = g.produce_next_state(fi, GpuRunParameters(target_time=T))
    assert err is None, err
    refs = []
    for prec in precision_pairs:
        o = pyoracle.Oracle(c, dt, IoState(0.0, parts), None, prec)
        o.set_collider_mesh(mesh)
        o.step_to(T, fi)
        refs.append((o.particles(), o.particles_f64()))
        o.close()
    return g, st, refs


def report(label, st, ref64):
    p = st.particles
    errs = {"x": rel_norm_err(p.positions, ref64["x"]), "v": rel_norm_err(p.velocities, ref64["v"]),
            "F": rel_norm_err(p.matrix_F(), ref64["F"])}
    print(label, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    return errs


def floor(z, friction=0.0, damping=0.0, subdivisions=6):
    mesh, verts = scenes.quad_mesh([(-0.5, -0.5, z), (1.5, -0.5, z), (1.5, 1.5, z), (-0.5, 1.5, z)],
                                   subdivisions=subdivisions)
    fi = scenes.mesh_frame(FrameInput(frame=0, gravity_a=(0.0, 0.0, -9.81)), verts, friction=friction, damping=damping,
                           num_triangles=mesh.triangles.shape[0])
    return mesh, fi


@pytest.mark.parametrize("d,bits", [(0.15, 0x10001), (-0.15, 0x10000), (0.21, 0)])
def test_lone_particle_bits(d, bits):
    c = InputConsts(grid_node_size=0.1, frames_per_second=24, domain_min=(-10, -10, -10), domain_max=(10, 10, 10))
    mesh, verts = scenes.quad_mesh([(-3, -3, 0), (3, -3, 0), (3, 3, 0), (-3, 3, 0)], subdivisions=4)
    fi = scenes.mesh_frame(FrameInput(frame=0, gravity_a=(0.0, 0.0, 0.0)), verts, num_triangles=mesh.triangles.shape[0])
    g = GpuState.from_io_state(0, c, DT, IoState(0.0, lone((0.013, 0.021, d))))
    g.set_collider_mesh(mesh)
    st, err = g.produce_next_state(fi, GpuRunParameters(scenes.advance_time(0.0, DT, 1)))
    assert err is None and st.particles.collider_bits[0] == bits


@pytest.mark.parametrize("friction,damping", [(0.0, 0.0), (0.5, 0.3), (200.0, 2.0)])
def test_lone_particle_crossing_push_back(friction, damping):
    c = InputConsts(grid_node_size=0.1, frames_per_second=24, domain_min=(-10, -10, -10), domain_max=(10, 10, 10))
    mesh, verts = scenes.quad_mesh([(-3, -3, 0), (3, -3, 0), (3, 3, 0), (-3, 3, 0)], subdivisions=4)
    fi = scenes.mesh_frame(FrameInput(frame=0, gravity_a=(0.0, 0.0, 0.0)), verts, friction=friction, damping=damping,
                           num_triangles=mesh.triangles.shape[0])
    d = float(np.float32(-0.04))
    v = np.array([0.7, -0.2, -1.5], np.float32).astype(np.float64)
    g = GpuState.from_io_state(0, c, DT, IoState(0.0, lone((0.013, 0.021, d), vel=v, bits=0x10001)))
    g.set_collider_mesh(mesh)
    st, err = g.produce_next_state(fi, GpuRunParameters(scenes.advance_time(0.0, DT, 1)))
    assert err is None and st.particles.collider_bits[0] == 0x10001
    fr, dm = float(np.float32(friction)), float(np.float32(damping))
    n = np.array([0.0, 0.0, -1.0])
    vn = n * (v @ n)
    vt = v - vn
    tn = np.linalg.norm(vt)
    want = v - (vt / tn) * min(fr * abs(d) / np.float32(DT), tn) - min(dm, 1.0) * vn - np.array([0, 0, d]) / np.float32(DT)
    assert np.allclose(st.particles.velocities[0], want, rtol=1e-5, atol=1e-4), (st.particles.velocities[0], want)


def cube_parts(corner_cells, n_cells, velocity, seed=0, E=2e4):
    sc = scenes.cube_scene(n_cells=n_cells, h=H, corner_cells=corner_cells, E=E, nu=0.3, velocity=velocity, dt=DT,
                           seed=seed)
    return sc.io_state.particles


@pytest.mark.parametrize("friction", [0.0, 0.6])
def test_cube_dropped_on_a_floor_mesh(friction):
    parts = cube_parts((12, 12, 10), 5, (0.3, 0.0, -2.0))
    mesh, fi = floor(0.2, friction=friction)
    g, st, [(o64, r64), (o32, r32)] = run_both(consts(), parts, mesh, fi, 120)
    p = st.particles
    near = (r64["x"][:, 2] < 0.2 + 2 * H)
    assert near.sum() > 50, "the cube reached the floor"
    assert (o64.collider_bits != 0).sum() > 50
    # no particle went through the floor
    assert p.positions[:, 2].min() > 0.2 - 2 * H
    mism = int((p.collider_bits != o32.collider_bits).sum())
    errs = report(f"cube on floor (friction {friction}), 120 substeps vs fp64, {mism} bit mismatches", st, r64)
    assert mism <= 3
    assert max(errs.values()) <= 1e-4


def test_thin_sheet_keeps_two_blocks_apart():
    parts_a = cube_parts((10, 12, 12), 4, (1.5, 0.0, 0.0), seed=1)
    parts_b = cube_parts((18, 12, 12), 4, (-1.5, 0.0, 0.0), seed=2)
    parts = scenes.concat([parts_a, parts_b])
    x0 = (10 + 4 + 18) / 2 * H  # the sheet halfway between the blocks
    mesh, verts = scenes.quad_mesh([(x0, -0.5, -0.5), (x0, 1.5, -0.5), (x0, 1.5, 1.5), (x0, -0.5, 1.5)], subdivisions=4)
    fi = scenes.mesh_frame(FrameInput(frame=0, gravity_a=(0.0, 0.0, 0.0)), verts, num_triangles=mesh.triangles.shape[0])
    g, st, [(o64, r64), (o32, r32)] = run_both(consts(), parts, mesh, fi, 100)
    p = st.particles
    na = parts_a.count
    assert p.positions[:na, 0].max() < x0 + H and p.positions[na:, 0].min() > x0 - H
    errs = report("two blocks and a sheet, 100 substeps vs fp64", st, r64)
    assert max(errs.values()) <= 1e-4


def test_moving_mesh_with_friction_and_damping():
    """A floor rising at 0.48 m/s (frames a and b differ: vertex velocities
    (b - a) * fps) under a resting cube."""
    parts = cube_parts((12, 12, 8), 4, (0.5, 0.0, -0.5))
    mesh, va = scenes.quad_mesh([(-0.5, -0.5, 0.2), (1.5, -0.5, 0.2), (1.5, 1.5, 0.2), (-0.5, 1.5, 0.2)], subdivisions=6)
    vb = va.copy()
    vb[:, 2] += 0.02
    fi = scenes.mesh_frame(FrameInput(frame=0, gravity_a=(0.0, 0.0, -9.81), gravity_b=(0.0, 0.0, -9.81)), va, vb,
                           friction=0.4, damping=0.5, num_triangles=mesh.triangles.shape[0])
    g, st, [(o64, r64), (o32, r32)] = run_both(consts(fps=24), parts, mesh, fi, 40)
    assert (o64.collider_bits != 0).sum() > 20
    errs = report("moving floor, 40 substeps vs fp64", st, r64)
    assert max(errs.values()) <= 1e-4


def separate_particles_on_a_floor():
    """2,000 particles 3 cells apart, all near a floor: 27 twins each."""
    rng = np.random.default_rng(7)
    ij = np.array([(i, j) for i in range(50) for j in range(40)])
    pos = np.zeros((ij.shape[0], 3), np.float32)
    pos[:, 0] = (3 * ij[:, 0] + 1.3 + rng.uniform(0, 0.4, ij.shape[0])) * H
    pos[:, 1] = (3 * ij[:, 1] + 1.3 + rng.uniform(0, 0.4, ij.shape[0])) * H
    pos[:, 2] = 0.2 + 1.2 * H
    p = Particles.empty(pos.shape[0])
    p.positions[:] = pos
    p.initial_positions[:] = pos
    p.position_gradients[:] = np.eye(3, dtype=np.float32)
    p.velocities[:] = np.array([0.0, 0.0, -1.0], np.float32)
    p.flags[:] = abi.IS_SOLID
    p.parameters["mass"] = 1e-3
    p.parameters["initial_volume"] = 1e-6
    p.parameters["mu"] = 10.0
    p.parameters["lambda"] = 10.0
    c = consts(lo=(0.0, 0.0, 0.0), hi=(5.0, 4.0, 1.0))
    mesh, verts = scenes.quad_mesh([(-0.5, -0.5, 0.2), (5.5, -0.5, 0.2), (5.5, 4.5, 0.2), (-0.5, 4.5, 0.2)], subdivisions=8)
    fi = scenes.mesh_frame(FrameInput(frame=0, gravity_a=(0.0, 0.0, -9.81)), verts, num_triangles=mesh.triangles.shape[0])
    return c, p, mesh, fi


def test_twin_table_overflow_redoes_the_call():
    """More twins than the engine's first table holds: the call is redone
    with a larger table and matches the oracle."""
    c, p, mesh, fi = separate_particles_on_a_floor()
    g, st, [(o64, r64), (o32, r32)] = run_both(c, p, mesh, fi, 30)
    assert np.all(st.particles.collider_bits != 0)
    errs = report("2000 separate particles on a floor, 30 substeps vs fp64", st, r64)
    assert max(errs.values()) <= 1e-4
