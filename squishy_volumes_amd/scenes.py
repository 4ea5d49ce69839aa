"""Seeded synthetic scenes for the parity tests and the benchmark.

Every scene is generated from a seed with numpy's PCG64, so the engine and the
CPU oracle receive byte-identical inputs.  Particles are sampled at 8 per cell
as a 2x2x2 stratified jitter inside each cell; V0 = h^3 / 8 and m = rho V0
(initialize_io_state, rust/crates/core/src/initialization.rs:195-262, with
size = h / 2 per particle).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import abi
from .io_state import ColliderMesh, EngineOptions, FrameInput, InputConsts, IoState, Particles, lame_lambda, lame_mu


@dataclass
class Scene:
    name: str
    consts: InputConsts
    time_step: float
    options: EngineOptions
    io_state: IoState
    frame_input: FrameInput
    substeps: int
    info: dict = field(default_factory=dict)


def advance_time(time: float, dt: float, substeps: int) -> float:
    """The time after `substeps` steps, accumulated as the substep loop does
    (f32 dt added to an f64 time, rust/crates/cpu/src/cpu_state.rs:187-190)."""
    t = float(time)
    dtf = float(np.float32(dt))
    for _ in range(substeps):
        t += dtf
    return t


def jittered_cells(cells: np.ndarray, h: float, rng: np.random.Generator, ppc: int = 8) -> np.ndarray:
    """Stratified jitter: ppc = 8 points per cell, one per octant."""
    assert ppc == 8
    octs = np.array([[i, j, k] for i in range(2) for j in range(2) for k in range(2)], np.float64)
    base = cells[:, None, :].astype(np.float64) + 0.5 * octs[None, :, :]
    jit = rng.random((cells.shape[0], 8, 3)) * 0.5
    return ((base + jit) * h).reshape(-1, 3)


def box_cells(lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    """Integer cells c with lo <= c < hi (per axis)."""
    axes = [np.arange(lo[a], hi[a]) for a in range(3)]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    return g


def sphere_points(center, radius, h, rng) -> np.ndarray:
    c = np.asarray(center, np.float64)
    lo = np.floor((c - radius) / h).astype(np.int64)
    hi = np.ceil((c + radius) / h).astype(np.int64)
    cells = box_cells(lo, hi)
    pts = jittered_cells(cells, h, rng)
    keep = np.sum((pts - c) ** 2, axis=1) < radius * radius
    return pts[keep]


def particles_from(points: np.ndarray, flags: int, rho: float, h: float, velocity, *, E=None, nu=None,
                   bulk_modulus=None, exponent=7, with_jp=False) -> Particles:
    n = points.shape[0]
    p = Particles.empty(n, with_jp=with_jp)
    p.positions[:] = points.astype(np.float32)
    p.initial_positions[:] = p.positions
    p.position_gradients[:] = np.eye(3, dtype=np.float32)
    p.velocities[:] = np.asarray(velocity, np.float32)
    p.flags[:] = flags
    vol = np.float32(h ** 3 / 8.0)
    p.parameters["initial_volume"] = vol
    p.parameters["mass"] = np.float32(vol * np.float32(rho))
    if flags & abi.IS_FLUID:
        p.parameters["bulk_modulus"] = bulk_modulus
        p.parameters["exponent"] = exponent
    else:
        p.parameters["mu"] = np.float32(lame_mu(np.float32(E), np.float32(nu)))
        p.parameters["lambda"] = np.float32(lame_lambda(np.float32(E), np.float32(nu)))
    return p


def concat(parts: list[Particles]) -> Particles:
    out = Particles.empty(0, with_jp=any(p.plastic_jp is not None for p in parts))
    for name in out.__dataclass_fields__:
        arrs = [getattr(p, name) for p in parts]
        if name == "plastic_jp":
            if out.plastic_jp is None:
                continue
            arrs = [a if a is not None else np.ones(p.count, np.float32) for a, p in zip(arrs, parts)]
        setattr(out, name, np.ascontiguousarray(np.concatenate(arrs)))
    return out


def cube_scene(*, n_cells: int = 16, h: float = 1.0 / 64, corner_cells=(24, 24, 16), E: float = 5e4, nu: float = 0.3,
               rho: float = 1000.0, velocity=(0.0, 0.0, -2.0), gravity=(0.0, 0.0, -9.81), dt: float = 1e-4,
               substeps: int = 1000, model: str = "neo_hookean", sticky_plane_z: Optional[float] = None,
               domain=((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)), fps: int = 24, seed: int = 0) -> Scene:
    """One elastic cube of n_cells^3 cells at 8 particles per cell."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(corner_cells, np.int64)
    pts = jittered_cells(box_cells(lo, lo + n_cells), h, rng)
    flags = abi.IS_SOLID | (abi.FIXED_COROTATED if model == "fixed_corotated" else 0)
    parts = particles_from(pts, flags, rho, h, velocity, E=E, nu=nu)
    consts = InputConsts(grid_node_size=h, frames_per_second=fps, domain_min=tuple(domain[0]), domain_max=tuple(domain[1]))
    opts = EngineOptions(sticky_plane_z=sticky_plane_z)
    return Scene(
        name=f"cube{n_cells}_{model}",
        consts=consts,
        time_step=dt,
        options=opts,
        io_state=IoState(time=0.0, particles=parts),
        frame_input=FrameInput(frame=0, gravity_a=tuple(gravity)),
        substeps=substeps,
        info={"particles": parts.count, "h": h},
    )


def plain_cube_parity(**kw) -> Scene:
    """The reference-expressible form of the 64^3 cube: Neo-Hookean, no plane,
    a domain large enough that no particle is culled."""
    kw.setdefault("domain", ((-0.5, -0.5, -0.5), (1.5, 1.5, 1.5)))
    return cube_scene(model="neo_hookean", sticky_plane_z=None, **kw)


def plain_cube_stated(**kw) -> Scene:
    """The 64^3 cube as stated: fixed-corotated, sticky ground plane at z = 0.1."""
    return cube_scene(model="fixed_corotated", sticky_plane_z=0.1, **kw)


def spheres_and_slab(*, h: float = 1.0 / 256, target_particles: int = 4_200_000, seed: int = 0,
                     dt: float = 5e-5, substeps: int = 500, fps: int = 1) -> Scene:
    """Three Neo-Hookean spheres (r = 0.08, 0.115, 0.15; E log-uniform in
    [1e4, 1e6]; nu = 0.3) above a weakly compressible fluid slab (K = 1e5),
    in a unit box with separating-slip walls, 8 particles per cell."""
    rng = np.random.default_rng(seed)
    spheres = [((0.30, 0.30, 0.45), 0.15), ((0.70, 0.35, 0.40), 0.115), ((0.50, 0.72, 0.50), 0.08)]
    Es = np.exp(rng.uniform(np.log(1e4), np.log(1e6), size=3))
    parts = []
    for (c, r), E in zip(spheres, Es):
        pts = sphere_points(c, r, h, rng)
        parts.append(particles_from(pts, abi.IS_SOLID, 1000.0, h, (0.0, 0.0, -1.0), E=float(E), nu=0.3))
    n_spheres = sum(p.count for p in parts)
    wall = 4 * h
    # A slab three cells thick, centred in x and y, whose side is chosen so
    # that the scene holds about target_particles.
    layers = 3
    side = int(round(np.sqrt(max(target_particles - n_spheres, 8 * layers) / (8 * layers))))
    side = min(side, int(np.floor(0.92 / h)))
    lo_xy = int(round(0.5 / h)) - side // 2
    hi_xy = lo_xy + side
    z0 = int(np.ceil(wall / h)) + 1
    cells = box_cells(np.array([lo_xy, lo_xy, z0]), np.array([hi_xy, hi_xy, z0 + layers]))
    pts = jittered_cells(cells, h, rng)
    parts.append(particles_from(pts, abi.IS_FLUID, 1000.0, h, (0.0, 0.0, 0.0), bulk_modulus=1e5, exponent=7))
    allp = concat(parts)
    consts = InputConsts(grid_node_size=h, frames_per_second=fps, domain_min=(0.0, 0.0, 0.0), domain_max=(1.0, 1.0, 1.0))
    opts = EngineOptions(box_walls=((wall, wall, wall), (1.0 - wall, 1.0 - wall, 1.0 - wall)))
    return Scene(
        name="spheres_and_slab",
        consts=consts,
        time_step=dt,
        options=opts,
        io_state=IoState(time=0.0, particles=allp),
        frame_input=FrameInput(frame=0, gravity_a=(0.0, 0.0, -9.81)),
        substeps=substeps,
        info={"particles": allp.count, "sphere_particles": n_spheres, "slab_layers": layers, "E": Es.tolist(), "h": h},
    )


def snow_block_on_sphere(*, h: float = 1.0 / 128, target_particles: int = 1_000_000, seed: int = 0,
                         dt: float = 1e-4, substeps: int = 1000, fps: int = 1, E: float = 1.4e5, nu: float = 0.2,
                         rho: float = 400.0) -> Scene:
    """A snow block (fixed-corotated elasticity with the singular values of F
    clamped to [1 - 2.5e-2, 1 + 4.5e-3] and hardening 10) falling at 3 m/s
    onto a sphere collider of radius 0.15, in a unit box with separating-slip
    walls, 8 particles per cell."""
    rng = np.random.default_rng(seed)
    side = int(round((target_particles / 8) ** (1.0 / 3.0)))
    lo_xy = int(round(0.5 / h)) - side // 2
    z0 = int(round(0.42 / h))
    cells = box_cells(np.array([lo_xy, lo_xy, z0]), np.array([lo_xy + side, lo_xy + side, z0 + side]))
    pts = jittered_cells(cells, h, rng)
    p = particles_from(pts, abi.IS_SOLID | abi.SNOW, rho, h, (0.0, 0.0, -3.0), E=E, nu=nu, with_jp=True)
    p.plastic_jp[:] = 1.0
    wall = 4 * h
    consts = InputConsts(grid_node_size=h, frames_per_second=fps, domain_min=(0.0, 0.0, 0.0), domain_max=(1.0, 1.0, 1.0))
    opts = EngineOptions(box_walls=((wall, wall, wall), (1.0 - wall, 1.0 - wall, 1.0 - wall)),
                         sphere=((0.5, 0.5, 0.22), 0.15))
    return Scene(
        name="snow_block_on_sphere",
        consts=consts,
        time_step=dt,
        options=opts,
        io_state=IoState(time=0.0, particles=p),
        frame_input=FrameInput(frame=0, gravity_a=(0.0, 0.0, -9.81)),
        substeps=substeps,
        info={"particles": p.count, "block_cells": side, "E": E, "nu": nu, "h": h},
    )


def random_objects_scene(target_particles: int, *, seed: int = 0, h: float = 1.0 / 128, dt: float = 5e-5,
                         substeps: int = 200, fps: int = 1, max_objects: int = 4,
                         materials: tuple = ("elastic",)) -> Scene:
    """1 to max_objects cubes and spheres that together hold about
    target_particles, placed without overlap in a unit box with
    separating-slip walls, 8 particles per cell, random velocities up to
    1 m/s.  Each object's material is drawn from `materials`:
      elastic  fixed-corotated, rho 1000, E log-uniform in [1e4, 1e6], nu 0.3;
      snow     the elastoplastic snow model, rho 400, E log-uniform in
               [1e5, 5e5], nu 0.2;
      fluid    weakly compressible, rho 1000, K 1e5, gamma 7."""
    rng = np.random.default_rng(seed)
    wall = 4 * h
    k = int(rng.integers(1, max_objects + 1))
    shares = rng.dirichlet(np.ones(k)) * target_particles
    placed = []  # (centre, bounding radius)
    parts = []
    for share in shares:
        cells = max(share / 8.0, 8.0)
        cube = bool(rng.random() < 0.5)
        if cube:
            side = max(2, int(round(cells ** (1.0 / 3.0))))
            half = 0.5 * side * h
            bound = half * np.sqrt(3.0)
        else:
            radius = (3.0 * cells / (4.0 * np.pi)) ** (1.0 / 3.0) * h
            bound = radius
        lo, hi = wall + 2 * h + bound, 1.0 - wall - 2 * h - bound
        if lo >= hi:
            continue
        for _ in range(50):
            c = rng.uniform(lo, hi, 3)
            if all(np.linalg.norm(c - c2) > bound + b2 + 2 * h for c2, b2 in placed):
                break
        else:
            continue
        placed.append((c, bound))
        if cube:
            c0 = np.floor(c / h - 0.5 * side).astype(np.int64)
            pts = jittered_cells(box_cells(c0, c0 + side), h, rng)
        else:
            pts = sphere_points(c, radius, h, rng)
        material = materials[int(rng.integers(len(materials)))]
        vel = rng.uniform(-1.0, 1.0, 3)
        vel *= rng.uniform(0.0, 1.0) / max(np.linalg.norm(vel), 1e-9)
        if material == "elastic":
            E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
            parts.append(particles_from(pts, abi.IS_SOLID | abi.FIXED_COROTATED, 1000.0, h, vel, E=E, nu=0.3))
        elif material == "snow":
            E = float(np.exp(rng.uniform(np.log(1e5), np.log(5e5))))
            parts.append(particles_from(pts, abi.IS_SOLID | abi.SNOW, 400.0, h, vel, E=E, nu=0.2, with_jp=True))
        elif material == "fluid":
            parts.append(particles_from(pts, abi.IS_FLUID, 1000.0, h, vel, bulk_modulus=1e5, exponent=7))
        else:
            raise ValueError(f"unknown material {material!r}")
    allp = concat(parts)
    consts = InputConsts(grid_node_size=h, frames_per_second=fps, domain_min=(0.0, 0.0, 0.0), domain_max=(1.0, 1.0, 1.0))
    opts = EngineOptions(box_walls=((wall, wall, wall), (1.0 - wall, 1.0 - wall, 1.0 - wall)))
    return Scene(
        name=f"random_objects_{seed}",
        consts=consts,
        time_step=dt,
        options=opts,
        io_state=IoState(time=0.0, particles=allp),
        frame_input=FrameInput(frame=0, gravity_a=(0.0, 0.0, -9.81)),
        substeps=substeps,
        info={"particles": allp.count, "objects": len(parts), "h": h},
    )


def uneven_scene_set(count: int = 64, lo: int = 5_000, hi: int = 500_000, seed: int = 0,
                     total: Optional[int] = None, **kw) -> list[Scene]:
    """`count` random-object scenes whose particle counts are drawn
    log-uniform in [lo, hi] and, with `total`, scaled (and clipped to
    [lo, hi]) so that they add up to about `total` (the batched workload: 64
    scenes, 8M particles in all)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), count))
    if total is not None:
        for _ in range(20):
            sizes = np.clip(sizes * (total / sizes.sum()), lo, hi)
    return [random_objects_scene(int(s), seed=seed * 1000 + i, **kw) for i, s in enumerate(sizes)]


def mixed_scene_set(count: int = 128, lo: int = 10_000, hi: int = 1_000_000, seed: int = 1,
                    total: Optional[int] = 16_000_000, **kw) -> list[Scene]:
    """The combined workload: `count` scenes of elastic, snow and fluid
    objects, log-uniform sizes in [lo, hi] scaled to about `total`."""
    return uneven_scene_set(count, lo, hi, seed, total, materials=("elastic", "snow", "fluid"), **kw)


def quad_mesh(corners, collider: int = 0, subdivisions: int = 1) -> tuple[ColliderMesh, np.ndarray]:
    """A planar quadrilateral (corners in counter-clockwise order seen from
    the side its normal points to) cut into subdivisions^2 cells of two
    triangles each: the topology and the vertex positions (float32)."""
    c = np.asarray(corners, np.float64)
    k = subdivisions
    uu, vv = np.meshgrid(np.linspace(0, 1, k + 1), np.linspace(0, 1, k + 1), indexing="ij")
    uu, vv = uu.ravel(), vv.ravel()
    verts = ((1 - uu) * (1 - vv))[:, None] * c[0] + (uu * (1 - vv))[:, None] * c[1] + (uu * vv)[:, None] * c[2] + \
            ((1 - uu) * vv)[:, None] * c[3]
    idx = lambda i, j: i * (k + 1) + j
    tris = []
    for i in range(k):
        for j in range(k):
            tris.append((idx(i, j), idx(i + 1, j), idx(i + 1, j + 1)))
            tris.append((idx(i, j), idx(i + 1, j + 1), idx(i, j + 1)))
    tris = np.array(tris, np.uint32)
    return ColliderMesh(verts.shape[0], tris, np.full(tris.shape[0], collider, np.uint32)), verts.astype(np.float32)


def box_mesh(lo, hi, collider: int = 0) -> tuple[ColliderMesh, np.ndarray]:
    """The closed surface of an axis-aligned box (12 triangles, outward normals)."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    verts = np.array([[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1], lo[2] if i & 4 == 0 else hi[2]]
                      for i in range(8)], np.float32)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tris = []
    for a, b, c, d in quads:
        tris.append((a, b, c))
        tris.append((a, c, d))
    tris = np.array(tris, np.uint32)
    return ColliderMesh(8, tris, np.full(12, collider, np.uint32)), verts


def mesh_frame(fi: FrameInput, vertices_a, vertices_b=None, friction: float = 0.0, damping: float = 0.0,
               num_triangles: int = 0) -> FrameInput:
    """fi with a collider mesh's frames: vertex positions of frames a (and b),
    and a uniform friction and damping per triangle."""
    import dataclasses

    fa = np.full(num_triangles, friction, np.float32)
    da = np.full(num_triangles, damping, np.float32)
    return dataclasses.replace(
        fi, vertex_positions_a=np.asarray(vertices_a, np.float32),
        vertex_positions_b=None if vertices_b is None else np.asarray(vertices_b, np.float32),
        triangle_frictions_a=fa, triangle_frictions_b=fa if vertices_b is not None else None,
        triangle_dampings_a=da, triangle_dampings_b=da if vertices_b is not None else None)
