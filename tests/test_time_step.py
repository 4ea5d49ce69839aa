"""The adaptive time-step limits (limit_time_step.rs:25-223,
rust/crates/cpu/src/phase/): the oracle pinned by closed forms (CPU), and
the engine's device reductions against the oracle (GPU).

Stated tolerances: closed forms to 1e-12 (FP64 oracle); engine against the
oracle 1e-4 relative for the velocity, deformation and isolated-particle
limits, 1e-3 for the sound-speed limit (an FP32 SVD and the divided
differences of the stress between close singular values)."""
import numpy as np
import pytest

from oracle import time_step_oracle as tso
from squishy_volumes_amd import abi, scenes
from squishy_volumes_amd.gpu_state import GpuRunParameters
from squishy_volumes_amd.io_state import Particles


def record(**kw):
    p = np.zeros(1, abi.PARAMS_DTYPE)[0]
    for k, v in kw.items():
        p[k] = v
    return p


def test_undeformed_solid_sounds_at_the_p_wave_speed():
    mu, lam, rho, h = 1.9e4, 2.9e4, 1000.0, 1.0 / 64
    p = record(mass=rho * 2e-6, initial_volume=2e-6, mu=mu, **{"lambda": lam})
    mu, lam, rho = float(p["mu"]), float(p["lambda"]), float(p["mass"]) / float(p["initial_volume"])  # as stored
    c = np.sqrt((lam + 2 * mu) / rho)
    assert tso.sound_limit(False, p, np.eye(3), h) == pytest.approx(h / c, rel=1e-12)
    # a rotation changes nothing
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    assert tso.sound_limit(False, p, R, h) == pytest.approx(h / c, rel=1e-12)


def test_fluid_at_rest_sounds_at_the_tait_speed():
    K, gamma, rho, h = 1e5, 7, 1000.0, 1.0 / 128
    p = record(mass=rho * 1e-6, initial_volume=1e-6, bulk_modulus=K, exponent=gamma)
    rho = float(p["mass"]) / float(p["initial_volume"])  # as stored
    assert tso.sound_limit(True, p, np.eye(3), h) == pytest.approx(h / np.sqrt(gamma * K / rho), rel=1e-12)
    # isolated fluid particle at J = 1: h sqrt(rho / (6 * 3 * gamma K))
    assert tso.isolated_limit(True, p, np.eye(3), h) == pytest.approx(h * np.sqrt(rho / (18 * gamma * K)), rel=1e-12)


def test_isolated_solid_velocity_and_deformation_closed_forms():
    h = 0.5
    p = record(mass=2.0, initial_volume=0.5, mu=3.0, **{"lambda": 4.0})
    assert tso.isolated_limit(False, p, np.eye(3), h) == pytest.approx(
        np.sqrt(2.0 / (0.5 * (3 / h**2) * 0.5 * (3.0 + 1.5 * 4.0))), rel=1e-12)
    flags = np.array([abi.IS_SOLID, abi.IS_SOLID, abi.IS_SOLID | abi.TOMBSTONED], np.uint32)
    params = np.array([p, p, p])
    F = np.repeat(np.eye(3)[None], 3, 0)
    v = np.array([[1.0, 2.0, 2.0], [0.0, 0.0, -1.0], [100.0, 0.0, 0.0]])  # the tombstoned one is ignored
    C = np.zeros((3, 3, 3))
    C[1, 0, 2] = -4.0
    C[2, 1, 1] = 1e3
    lim = tso.limits(flags, params, F, v, C, h)
    assert lim["by_velocity"] == pytest.approx(0.5 * h / 3.0)
    assert lim["by_deformation"] == pytest.approx(0.2 / 4.0)


def mixed_state(n_cells=5, seed=7):
    """A perturbed cube, a third of it fluid, one particle tombstoned, some F
    with two equal singular values (the separation branches)."""
    from tests.test_gpu_step import perturbed_cube

    sc = perturbed_cube(n_cells=n_cells, seed=seed)
    p = sc.io_state.particles
    rng = np.random.default_rng(seed)
    fl = rng.random(p.count) < 1 / 3
    p.flags[fl] = abi.IS_FLUID
    p.parameters["bulk_modulus"][fl] = 1e4
    p.parameters["exponent"][fl] = 7
    F = p.matrix_F().astype(np.float64)
    for i in range(0, p.count, 17):  # U diag(a, a, b) V^T
        U, _, Vt = np.linalg.svd(F[i])
        F[i] = U @ np.diag([1.02, 1.02, 0.97]) @ Vt
    Jf = np.linalg.det(F[fl]) ** (1 / 3)
    U, _, Vt = np.linalg.svd(F[fl])
    F[fl] = (U @ Vt) * Jf[:, None, None]  # fluids: F = J^(1/3) R
    p.position_gradients[:] = np.transpose(F, (0, 2, 1)).astype(np.float32)
    p.flags[3] |= abi.TOMBSTONED
    p.velocities[3] = (1e3, 0, 0)
    return sc


def oracle_limits(parts: Particles, h: float):
    return tso.limits(parts.flags, parts.parameters, parts.matrix_F().astype(np.float64),
                      parts.velocities.astype(np.float64), parts.matrix_C().astype(np.float64), h)


def compare(eng, ref, label):
    print(label, {k: (f"{eng[k]:.6g}", f"{ref[k]:.6g}") for k in ref})
    for k, tol in (("by_velocity", 1e-4), ("by_deformation", 1e-4), ("by_isolated", 1e-4), ("by_sound", 1e-3)):
        assert eng[k] == pytest.approx(ref[k], rel=tol), (label, k)


@pytest.mark.gpu
def test_engine_limits_match_the_oracle():
    from tests.test_gpu_step import engine

    sc = mixed_state()
    h = sc.consts.grid_node_size
    g = engine(sc)
    compare(g.time_step_limits(), oracle_limits(sc.io_state.particles, h), "uploaded state")
    st, err = g.produce_next_state(sc.frame_input, GpuRunParameters(scenes.advance_time(0.0, sc.time_step, 15)))
    assert err is None
    compare(g.time_step_limits(), oracle_limits(st.particles, h), "after 15 substeps")


@pytest.mark.gpu
def test_adapting_the_step_between_calls():
    from tests.test_gpu_step import engine

    sc = mixed_state()
    g = engine(sc)
    lim = g.time_step_limits()
    dt = g.adapt_time_step(max_time_step=1.0)
    assert dt == pytest.approx(min(lim.values()), rel=1e-6) and dt < sc.time_step * 10
    t = g.time
    target = t + 10.5 * dt
    st, err = g.produce_next_state(sc.frame_input, GpuRunParameters(target_time=target))
    assert err is None and g.last_status.substeps == 11
    assert g.adapt_time_step(max_time_step=1e-5) == pytest.approx(1e-5)
    with pytest.raises(Exception):
        g.set_time_step(float("nan"))
    # through the run parameters, as CpuRunParameters carries them
    lim = min(g.time_step_limits().values())
    target = g.time + 5.5 * lim
    st, err = g.produce_next_state(sc.frame_input, GpuRunParameters(target_time=target, max_time_step=1.0,
                                                                    adaptive_time_steps=True))
    assert err is None and g.last_status.substeps == 6
