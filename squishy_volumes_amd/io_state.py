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
        for name in ("particle_flags_a", "particle_flags_b", "goal_positions_a", "goal_positions_b"):
            a = getattr(self, name)
            if a is not None:
                a = np.ascontiguousarray(a, dtype=np.uint32 if "flags" in name else np.float32)
                keep.append(a)
                setattr(fi, name, a.ctypes.data)
        return fi, keep
