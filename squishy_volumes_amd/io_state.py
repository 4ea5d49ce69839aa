"""The data that crosses the backend boundary, as numpy arrays.

Mirrors IoState (rust/crates/file_frame/src/io_state.rs:11-18), Particles
(rust/crates/file_frame/src/particles.rs:92-109), GridNodes
(rust/crates/file_frame/src/grid_nodes.rs:9-15), the InputConsts the substep
reads (rust/crates/file_input/src/header.rs:9-69) and the per-frame inputs of
FrameInput (rust/crates/xpu/src/frame_input.rs:33-278).

Layout: vectors are (N, 3) float32, matrices (N, 3, 3) float32 indexed
[particle, column, row] - the memory order of [[f32; 3]; 3] in IoState, whose
outer index is the column.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import abi


def lame_mu(E, nu):
    """mu, rust/crates/util/src/elastic.rs:52-57."""
    return E / 2.0 / (1.0 + nu)


def lame_lambda(E, nu):
    """lambda, rust/crates/util/src/elastic.rs:59-64."""
    return E * nu / (1.0 + nu) / (1.0 - 2.0 * nu)


@dataclass
class Particles:
    flags: np.ndarray
    parameters: np.ndarray  # abi.PARAMS_DTYPE records
    elastic_energies: np.ndarray
    collider_bits: np.ndarray
    positions: np.ndarray
    position_gradients: np.ndarray
    velocities: np.ndarray
    velocity_gradients: np.ndarray
    initial_positions: np.ndarray
    plastic_jp: Optional[np.ndarray] = None

    @property
    def count(self) -> int:
        return int(self.flags.shape[0])

    @staticmethod
    def empty(n: int, with_jp: bool = False) -> "Particles":
        return Particles(
            flags=np.zeros(n, np.uint32),
            parameters=np.zeros(n, abi.PARAMS_DTYPE),
            elastic_energies=np.zeros(n, np.float32),
            collider_bits=np.zeros(n, np.uint32),
            positions=np.zeros((n, 3), np.float32),
            position_gradients=np.zeros((n, 3, 3), np.float32),
            velocities=np.zeros((n, 3), np.float32),
            velocity_gradients=np.zeros((n, 3, 3), np.float32),
            initial_positions=np.zeros((n, 3), np.float32),
            plastic_jp=np.ones(n, np.float32) if with_jp else None,
        )

    def contiguous(self) -> "Particles":
        for name in self.__dataclass_fields__:
            a = getattr(self, name)
            if a is not None:
                setattr(self, name, np.ascontiguousarray(a))
        return self

    def copy(self) -> "Particles":
        return Particles(**{k: (None if getattr(self, k) is None else getattr(self, k).copy()) for k in self.__dataclass_fields__})

    def view(self) -> abi.MpmParticlesView:
        """An MpmParticlesView over these arrays (they must stay alive)."""
        self.contiguous()
        v = abi.MpmParticlesView()
        v.count = self.count
        v.time = 0.0
        v.flags = abi.ptr(self.flags)
        v.parameters = abi.ptr(self.parameters)
        v.elastic_energies = abi.ptr(self.elastic_energies)
        v.collider_bits = abi.ptr(self.collider_bits)
        v.positions = abi.ptr(self.positions)
        v.position_gradients = abi.ptr(self.position_gradients)
        v.velocities = abi.ptr(self.velocities)
        v.velocity_gradients = abi.ptr(self.velocity_gradients)
        v.initial_positions = abi.ptr(self.initial_positions)
        v.plastic_jp = abi.ptr(self.plastic_jp)
        return v

    def matrix_F(self) -> np.ndarray:
        """F as ordinary (N, row, col) matrices."""
        return np.transpose(self.position_gradients, (0, 2, 1))

    def matrix_C(self) -> np.ndarray:
        return np.transpose(self.velocity_gradients, (0, 2, 1))


@dataclass
class GridNodes:
    node_ids: np.ndarray
    collider_bits: np.ndarray
    masses: np.ndarray
    velocities: np.ndarray


@dataclass
class IoState:
    time: float
    particles: Particles
    grid_nodes: Optional[GridNodes] = None

    def write(self, path) -> int:
        """IoState::write (rust/crates/file_frame/src/io_state.rs:31-62): the
        reference's frame file, written to a temporary file beside `path` and
        moved into place; returns the bytes written.  Encoded by the engine
        library's mpm_frame_encode (host code)."""
        import os
        from pathlib import Path

        lib = abi.load_library()
        parts = self.particles.contiguous()
        view = parts.view()
        view.time = self.time
        gview, keep = _grid_view(self.grid_nodes)
        gptr = C.byref(gview) if gview is not None else None
        size = int(lib.mpm_frame_file_size(C.byref(view), gptr))
        if size == 0:
            raise ValueError("particles or grid nodes are incomplete")
        buf = np.empty(size, np.uint8)
        written = C.c_uint64()
        rc = lib.mpm_frame_encode(C.byref(view), gptr, buf.ctypes.data, size, C.byref(written))
        if rc != 0:
            raise RuntimeError(f"mpm_frame_encode: {lib.mpm_error_string(rc).decode()}")
        path = Path(path)
        temp = path.parent / "temp.bin"
        buf.tofile(temp)
        os.replace(temp, path)
        return size

    @staticmethod
    def read(path) -> "IoState":
        """IoState::read (rust/crates/file_frame/src/io_state.rs:64-74)."""
        lib = abi.load_library()
        data = np.fromfile(path, dtype=np.uint8)
        n, m = C.c_uint64(), C.c_uint64()
        rc = lib.mpm_frame_decode_counts(data.ctypes.data, data.size, C.byref(n), C.byref(m))
        if rc != 0:
            raise ValueError(f"{path} is not a frame file of this format version")
        parts = Particles.empty(int(n.value))
        view = parts.view()
        grid = None
        gview = None
        if m.value != 2**64 - 1:
            k = int(m.value)
            grid = GridNodes(np.zeros((k, 3), np.int32), np.zeros(k, np.uint32), np.zeros(k, np.float32),
                             np.zeros((k, 3), np.float32))
            gview, keep = _grid_view(grid)
            gview.capacity = k
        rc = lib.mpm_frame_decode(data.ctypes.data, data.size, C.byref(view), C.byref(gview) if gview is not None else None)
        if rc != 0:
            raise ValueError(f"{path}: {lib.mpm_error_string(rc).decode()}")
        return IoState(time=view.time, particles=parts, grid_nodes=grid)


def _grid_view(grid: Optional[GridNodes]):
    if grid is None:
        return None, None
    arrays = [np.ascontiguousarray(grid.node_ids, np.int32), np.ascontiguousarray(grid.collider_bits, np.uint32),
              np.ascontiguousarray(grid.masses, np.float32), np.ascontiguousarray(grid.velocities, np.float32)]
    g = abi.MpmGridView()
    g.capacity = g.count = int(arrays[2].shape[0])
    g.node_ids, g.collider_bits, g.masses, g.velocities = (abi.ptr(a) for a in arrays)
    return g, arrays


@dataclass
class InputConsts:
    """InputConsts, rust/crates/file_input/src/header.rs:9-69."""

    grid_node_size: float
    frames_per_second: int = 24
    simulation_scale: float = 1.0
    domain_min: tuple = (-100.0, -100.0, -100.0)
    domain_max: tuple = (100.0, 100.0, 100.0)

    def scaled_grid_node_size(self) -> float:
        return self.grid_node_size / self.simulation_scale

    def scaled_domain_min(self):
        return tuple(c / self.simulation_scale for c in self.domain_min)

    def scaled_domain_max(self):
        return tuple(c / self.simulation_scale for c in self.domain_max)

    def seconds_per_frame(self) -> float:
        return 1.0 / self.frames_per_second


@dataclass
class EngineOptions:
    """Grid forces, boundary conditions and snow constants (no reference counterpart)."""

    grid_gravity: bool = False
    sticky_plane_z: Optional[float] = None
    box_walls: Optional[tuple] = None  # ((xmin, ymin, zmin), (xmax, ymax, zmax))
    sphere: Optional[tuple] = None  # ((cx, cy, cz), radius): separating slip on the sphere
    scene_period: Optional[tuple] = None  # batched scenes: period of the repeated domain per axis
    scene_repeat: Optional[tuple] = None  # and the number of copies per axis
    snow_compression: float = 2.5e-2
    snow_stretch: float = 4.5e-3
    snow_hardening: float = 10.0


def make_consts(consts: InputConsts, time_step: float, options: EngineOptions | None = None, device: int = 0) -> abi.MpmConsts:
    options = options or EngineOptions()
    c = abi.MpmConsts()
    c.grid_node_size = consts.scaled_grid_node_size()
    c.time_step = time_step
    c.domain_min[:] = consts.scaled_domain_min()
    c.domain_max[:] = consts.scaled_domain_max()
    c.frames_per_second = consts.frames_per_second
    opt = 0
    if options.grid_gravity:
        opt |= abi.OPT_GRID_GRAVITY
    if options.sticky_plane_z is not None:
        opt |= abi.OPT_STICKY_PLANE
        c.plane_z = options.sticky_plane_z
    if options.box_walls is not None:
        opt |= abi.OPT_BOX_WALLS
        c.wall_min[:] = options.box_walls[0]
        c.wall_max[:] = options.box_walls[1]
    if options.sphere is not None:
        opt |= abi.OPT_SPHERE
        c.sphere_center[:] = options.sphere[0]
        c.sphere_radius = options.sphere[1]
    if options.scene_repeat is not None:
        c.scene_period[:] = options.scene_period
        c.scene_repeat[:] = options.scene_repeat
    c.options = opt
    c.snow_compression = options.snow_compression
    c.snow_stretch = options.snow_stretch
    c.snow_hardening = options.snow_hardening
    c.device = device
    return c


@dataclass
class FrameInput:
    """The substep's view of FrameInput: frame a, optional frame b, gravity and goals."""

    frame: int = 0
    gravity_a: tuple = (0.0, 0.0, -9.81)
    gravity_b: Optional[tuple] = None
    particle_flags_a: Optional[np.ndarray] = None
    particle_flags_b: Optional[np.ndarray] = None
    goal_positions_a: Optional[np.ndarray] = None
    goal_positions_b: Optional[np.ndarray] = None
    # the collider mesh's frames (InputFrame vertex positions, triangle frictions and dampings)
    vertex_positions_a: Optional[np.ndarray] = None
    vertex_positions_b: Optional[np.ndarray] = None
    triangle_frictions_a: Optional[np.ndarray] = None
    triangle_frictions_b: Optional[np.ndarray] = None
    triangle_dampings_a: Optional[np.ndarray] = None
    triangle_dampings_b: Optional[np.ndarray] = None

    def frame_factor(self, time: float, fps: int) -> float:
        """FrameInput::frame_factor, rust/crates/xpu/src/frame_input.rs:266-278."""
        frame_time = time * fps
        low = int(np.floor(frame_time))
        if low != self.frame:
            raise ValueError(f"Wanted to interpolate from {low}, but {self.frame} is loaded")
        return float(np.float32(frame_time % 1.0))

    def to_c(self) -> tuple[abi.MpmFrameInput, list]:
        fi = abi.MpmFrameInput()
        keep = []
        fi.frame = self.frame
        fi.gravity_a[:] = self.gravity_a
        fi.has_b = 1 if self.gravity_b is not None else 0
        fi.gravity_b[:] = self.gravity_b if self.gravity_b is not None else self.gravity_a
        for name in ("particle_flags_a", "particle_flags_b", "goal_positions_a", "goal_positions_b",
                     "vertex_positions_a", "vertex_positions_b", "triangle_frictions_a", "triangle_frictions_b",
                     "triangle_dampings_a", "triangle_dampings_b"):
            a = getattr(self, name)
            if a is not None:
                a = np.ascontiguousarray(a, dtype=np.uint32 if "flags" in name else np.float32)
                keep.append(a)
                setattr(fi, name, a.ctypes.data)
        return fi, keep


@dataclass
class ColliderMesh:
    """The collider mesh's topology (Topology, rust/crates/mesh_util/src/mesh.rs):
    triangles as vertex index triples and the collider (0..15) of each triangle."""

    num_vertices: int
    triangles: np.ndarray            # (num_triangles, 3) uint32
    triangle_colliders: np.ndarray   # (num_triangles,) uint32

    def to_c(self) -> tuple[abi.MpmColliderMesh, list]:
        tri = np.ascontiguousarray(self.triangles, np.uint32)
        col = np.ascontiguousarray(self.triangle_colliders, np.uint32)
        m = abi.MpmColliderMesh()
        m.num_vertices = int(self.num_vertices)
        m.num_triangles = int(tri.shape[0])
        m.triangles = tri.ctypes.data
        m.triangle_colliders = col.ctypes.data
        return m, [tri, col]
