"""Pins of the oracle's substep by invariants of the MLS-MPM method.

Each test is chosen so that a plausible mistake in the restatement of
rust/crates/cpu/src/phase/*.rs fails it: a dropped or mis-scaled stress term
(the grid force must be minus the energy gradient), a transposed affine term
(APIC must carry a non-symmetric affine field exactly), a wrong weight moment,
gravity placement, culling strictness, goal velocities, the energy-error path.
"""
import numpy as np
import scipy.linalg
import pytest

from oracle import pyoracle
from squishy_volumes_amd import abi, scenes
from squishy_volumes_amd.io_state import EngineOptions, FrameInput, InputConsts, IoState, Particles
from tests.helpers import reference_fixtures, rel_norm_err


def lattice_particles(h=0.5, n_side=8, seed=0):
    """The reference's 512 lattice points (many_positions,
    rust/crates/gpu/src/test_util.rs:112-631) as solid particles."""
    pos = np.array(reference_fixtures()["many_positions"], np.float32)
    n = len(pos)
    p = Particles.empty(n)
    p.positions[:] = pos
    p.initial_positions[:] = pos
    p.position_gradients[:] = np.eye(3, dtype=np.float32)
    p.flags[:] = abi.IS_SOLID
    p.parameters["mass"] = 0.7
    p.parameters["initial_volume"] = 0.3
    return p


def run_one(parts, h, dt, gravity=(0.0, 0.0, 0.0), options=None, precision=8, frame_input=None, substeps=1,
            domain=((-100, -100, -100), (100, 100, 100))):
    consts = InputConsts(grid_node_size=h, frames_per_second=24, domain_min=domain[0], domain_max=domain[1])
    o = pyoracle.Oracle(consts, dt, IoState(0.0, parts), options, precision)
    fi = frame_input or FrameInput(frame=0, gravity_a=gravity)
    st = o.step_to(scenes.advance_time(0.0, dt, substeps), fi)
    return o, st


def test_p2g_conserves_mass_and_momentum():
    parts = lattice_particles()
    rng = np.random.default_rng(5)
    parts.velocities[:] = rng.uniform(-1, 1, (parts.count, 3))
    parts.velocity_gradients[:] = rng.uniform(-1, 1, (parts.count, 3, 3))
    F = np.eye(3) + 0.2 * rng.uniform(-1, 1, (parts.count, 3, 3))
    parts.position_gradients[:] = np.transpose(F, (0, 2, 1))
    parts.parameters["mu"] = 1e3
    parts.parameters["lambda"] = 2e3
    consts = InputConsts(grid_node_size=0.5, frames_per_second=24)
    g = (0.0, 0.0, -9.81)
    o = pyoracle.Oracle(consts, 1e-3, IoState(0.0, parts), None, 8)
    assert o.run_to_scatter(FrameInput(frame=0, gravity_a=g)) == 0
    ids, m, mom = o.grid_f64()
    pm = parts.parameters["mass"].astype(np.float64)
    assert np.isclose(m.sum(), pm.sum(), rtol=1e-12)
    v_after_gravity = parts.velocities.astype(np.float64) + np.float32(1e-3) * np.array(np.float32(g), np.float64)
    expect = (pm[:, None] * v_after_gravity).sum(0)
    # The affine and stress terms carry no net momentum: sum_i w (x_i - x_p) = 0.
    assert np.allclose(mom.sum(0), expect, rtol=1e-9, atol=1e-9)


def test_grid_force_is_minus_energy_gradient():
    """dt * f_i from P2G equals -dt dE/dx_i, E = V0 Psi(F(x)) with the MLS
    deformation update dF/dx_i = 4/h^2 w_ip (x_i - x_p)^T F."""
    h, dt = 0.5, 1e-3
    parts = Particles.empty(1)
    xp = np.array([0.3, -0.2, 0.45])
    parts.positions[0] = xp
    F = np.array([[1.1, 0.2, -0.1], [0.05, 0.9, 0.15], [-0.1, 0.1, 1.2]])
    parts.position_gradients[0] = F.T
    parts.flags[:] = abi.IS_SOLID
    parts.parameters["mass"] = 0.0  # isolate the stress term
    parts.parameters["initial_volume"] = 0.4
    parts.parameters["mu"] = 300.0
    parts.parameters["lambda"] = 500.0
    consts = InputConsts(grid_node_size=h, frames_per_second=24)
    o = pyoracle.Oracle(consts, dt, IoState(0.0, parts), None, 8)
    o.run_to_scatter(FrameInput(frame=0, gravity_a=(0, 0, 0)))
    ids, m, mom = o.grid_f64()
    xp = parts.positions[0].astype(np.float64)
    F = F.astype(np.float32).astype(np.float64)
    V0 = 0.4
    p1 = np.zeros(1, abi.PARAMS_DTYPE)
    p1["mu"], p1["lambda"] = 300.0, 500.0

    def energy(Fm):
        return V0 * pyoracle.stress(Fm[None], np.array([abi.IS_SOLID], np.uint32), p1)[2][0]

    for k, node in enumerate(ids):
        d = node * h - xp
        w = np.prod([pyoracle.kernel_quadratic(c) for c in node - xp / h])
        grad = np.zeros(3)
        eps = 1e-6
        for a in range(3):
            e = np.zeros(3)
            e[a] = eps
            dFp = (np.eye(3) + np.outer(e, 4 / h**2 * w * d)) @ F
            dFm = (np.eye(3) - np.outer(e, 4 / h**2 * w * d)) @ F
            grad[a] = (energy(dFp) - energy(dFm)) / (2 * eps)
        assert np.allclose(mom[k], -dt * grad, rtol=1e-5, atol=1e-9), (node, mom[k], -dt * grad)


def test_apic_carries_affine_fields_exactly():
    """v_p = v0 + B x_p, C_p = B (B non-symmetric), no stiffness, no gravity:
    after one substep v = v0 + B x_p, C = B, x' = x + dt v, F' = (I + dt B) F."""
    parts = lattice_particles()
    # Dyadic values so that the float32 inputs hold the affine field exactly.
    v0 = np.array([0.25, -0.75, 0.125])
    B = np.array([[0.125, 0.5, -0.3125], [-0.1875, 0.0625, 0.375], [0.25, -0.625, -0.125]])
    x = parts.positions.astype(np.float64)
    parts.velocities[:] = v0 + x @ B.T
    parts.velocity_gradients[:] = B.T  # (col, row) layout
    dt = 1e-3
    o, st = run_one(parts, 0.5, dt)
    assert st.result == 0 and st.substeps == 1
    r = o.particles_f64()
    v = v0 + x @ B.T
    assert np.allclose(r["v"], v, atol=1e-10)
    assert np.allclose(r["C"], B[None], atol=1e-10)
    assert np.allclose(r["x"], x + np.float32(dt) * v, atol=1e-10)
    assert np.allclose(r["F"], (np.eye(3) + np.float32(dt) * B)[None], atol=1e-10)


def test_free_fall():
    parts = lattice_particles()
    parts.velocities[:] = (0.0, 0.0, -2.0)
    dt, n = 1e-3, 7
    g = (0.0, 0.0, -9.81)
    o, st = run_one(parts, 0.5, dt, gravity=g, substeps=n)
    assert st.substeps == n
    r = o.particles_f64()
    dtf = float(np.float32(dt))
    gz = float(np.float32(-9.81))
    vz = -2.0 + n * dtf * gz
    assert np.allclose(r["v"][:, 2], vz, atol=1e-9) and np.allclose(r["v"][:, :2], 0, atol=1e-12)
    # semi-implicit Euler: x_n = x_0 + dt sum_k v_k
    zs = parts.positions[:, 2].astype(np.float64) + dtf * sum(-2.0 + k * dtf * gz for k in range(1, n + 1))
    assert np.allclose(r["x"][:, 2], zs, atol=1e-9)
    assert np.allclose(r["C"], 0, atol=1e-9) and np.allclose(r["F"], np.eye(3)[None], atol=1e-12)


def test_grid_gravity_equals_particle_gravity_in_exact_arithmetic():
    sc = scenes.plain_cube_parity(n_cells=4)
    T = scenes.advance_time(0.0, sc.time_step, 5)
    a = pyoracle.Oracle(sc.consts, sc.time_step, sc.io_state, EngineOptions(), 8)
    b = pyoracle.Oracle(sc.consts, sc.time_step, sc.io_state, EngineOptions(grid_gravity=True), 8)
    a.step_to(T, sc.frame_input)
    b.step_to(T, sc.frame_input)
    ra, rb = a.particles_f64(), b.particles_f64()
    for k in ra:
        assert rel_norm_err(rb[k], ra[k]) < 1e-9 or np.abs(rb[k] - ra[k]).max() < 1e-9, k


def test_cull_is_strict():
    parts = lattice_particles()
    parts.velocities[:] = 0
    # Particles exactly on the domain boundary after the step are culled (strict >, <).
    lo = float(parts.positions[:, 0].min())
    o, st = run_one(parts, 0.5, 1e-3, domain=((lo, -100, -100), (100, 100, 100)))
    out = o.particles()
    on_boundary = parts.positions[:, 0] == lo
    assert on_boundary.any()
    assert np.all(out.flags[on_boundary] & abi.TOMBSTONED)
    assert not np.any(out.flags[~on_boundary] & abi.TOMBSTONED)


def test_energy_error_flags_failed_particle():
    parts = lattice_particles()
    parts.parameters["mu"] = 10.0
    parts.parameters["lambda"] = 10.0
    F = np.eye(3)
    F[2, 2] = -1.0
    parts.position_gradients[5] = F.T
    o, st = run_one(parts, 0.5, 1e-3)
    assert st.result == abi.ERR_ENERGY
    out = o.particles()
    assert out.flags[5] & abi.FAILED
    assert np.count_nonzero(out.flags & abi.FAILED) >= 1


def test_goal_velocity():
    parts = lattice_particles()
    n = parts.count
    goal = parts.positions + np.array([0.01, -0.02, 0.005], np.float32)
    flags = np.zeros(n, np.uint32)
    flags[:] = abi.HAS_GOAL
    fi = FrameInput(frame=0, gravity_a=(0.0, 0.0, -9.81), particle_flags_a=flags, goal_positions_a=goal)
    o, st = run_one(parts, 0.5, 1e-3, frame_input=fi)
    r = o.particles_f64()
    dt = float(np.float32(1e-3))
    # Every particle has a goal: a uniform field (goal - x)/dt, reproduced exactly.
    assert np.allclose(r["v"], (goal.astype(np.float64) - parts.positions) / dt, rtol=1e-5)
    assert np.allclose(r["x"], goal, atol=1e-6)


def test_sticky_plane_and_box_walls():
    sc = scenes.cube_scene(n_cells=4, corner_cells=(30, 30, 8), velocity=(1.0, -1.0, -3.0))
    opts = EngineOptions(sticky_plane_z=0.16, box_walls=((0.49, 0.0, 0.0), (0.52, 1.0, 1.0)))
    o = pyoracle.Oracle(sc.consts, sc.time_step, sc.io_state, opts, 8)
    o.step_to(scenes.advance_time(0.0, sc.time_step, 1), sc.frame_input)
    ids, m, v = o.grid_f64()
    x = ids * sc.consts.grid_node_size
    live = m > 0
    below = live & (x[:, 2] < 0.16)
    assert below.any() and np.all(v[below] == 0)
    outside_lo = live & (x[:, 0] < 0.49)
    outside_hi = live & (x[:, 0] > 0.52)
    assert outside_lo.any() and outside_hi.any()
    assert np.all(v[outside_lo, 0] >= 0) and np.all(v[outside_hi, 0] <= 0)


def test_fluid_projection_keeps_only_volume():
    parts = lattice_particles()
    parts.flags[:] = abi.IS_FLUID
    parts.parameters["bulk_modulus"] = 100.0
    parts.parameters["exponent"] = 7
    rng = np.random.default_rng(2)
    parts.velocity_gradients[:] = rng.uniform(-1, 1, (parts.count, 3, 3))
    dt = 1e-2
    o, st = run_one(parts, 0.5, dt)
    r = o.particles_f64()
    F = r["F"]
    J = np.linalg.det(F)
    # F = J^(1/3) R  =>  F F^T = J^(2/3) I
    assert np.allclose(F @ np.transpose(F, (0, 2, 1)), (J ** (2 / 3))[:, None, None] * np.eye(3), atol=1e-10)
    assert np.any(np.abs(J - 1) > 1e-4)
    # the projection keeps the advanced volume: J = det(F0 + dt C F0), F0 = I,
    # with C the gathered velocity gradient the advance used
    F_pre = np.eye(3) + float(np.float32(dt)) * r["C"]  # the step is an f32
    assert np.allclose(J, np.linalg.det(F_pre), rtol=1e-12, atol=0)
    # and the rotation: F / J^(1/3) is the orthogonal polar factor of F_pre
    Q = np.array([scipy.linalg.polar(f)[0] for f in F_pre])
    assert np.allclose(F / np.cbrt(J)[:, None, None], Q, atol=1e-12)


def test_snow_clamps_singular_values_and_tracks_plastic_volume():
    parts = lattice_particles()
    parts.flags[:] = abi.IS_SOLID | abi.SNOW
    parts.plastic_jp = np.ones(parts.count, np.float32)
    parts.parameters["mu"] = 100.0
    parts.parameters["lambda"] = 100.0
    rng = np.random.default_rng(4)
    B = rng.uniform(-3, 3, (parts.count, 3, 3))
    parts.velocity_gradients[:] = np.transpose(B, (0, 2, 1))
    x = parts.positions.astype(np.float64)
    parts.velocities[:] = 0
    dt = 1e-2
    opts = EngineOptions(snow_compression=2.5e-2, snow_stretch=4.5e-3, snow_hardening=10.0)
    o, st = run_one(parts, 0.5, dt, options=opts)
    out = o.particles()
    r = o.particles_f64()
    s = np.linalg.svd(r["F"], compute_uv=False)
    assert s.max() <= 1 + 4.5e-3 + 1e-6 and s.min() >= 1 - 2.5e-2 - 1e-6
    # det(F_E) * Jp = det(F_trial): the clamp moves volume between the parts.
    Jtrial = np.linalg.det(np.eye(3)[None] + np.float32(dt) * r["C"])
    assert np.allclose(np.linalg.det(r["F"]) * out.plastic_jp, Jtrial, rtol=2e-6)


def test_float_and_double_oracles_agree_on_the_cube():
    sc = scenes.plain_cube_parity(n_cells=4)
    T = scenes.advance_time(0.0, sc.time_step, 20)
    a = pyoracle.Oracle(sc.consts, sc.time_step, sc.io_state, sc.options, 4)
    b = pyoracle.Oracle(sc.consts, sc.time_step, sc.io_state, sc.options, 8)
    a.step_to(T, sc.frame_input)
    b.step_to(T, sc.frame_input)
    ra, rb = a.particles_f64(), b.particles_f64()
    assert rel_norm_err(ra["x"], rb["x"]) < 1e-6
    assert rel_norm_err(ra["v"], rb["v"]) < 1e-5
    assert rel_norm_err(ra["F"], rb["F"]) < 1e-6
