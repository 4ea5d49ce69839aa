"""Per-substep adaptive time steps on the GPU against the oracle's adaptive
mode, through the C ABI (mpm_set_adaptive_time_steps).

CpuState::produce_next_state with adaptive_time_steps (rust/crates/cpu/src/
cpu_state.rs:138-197) takes the limits of limit_time_step.rs:25-223 twice per
substep and steps by its AdaptiveTimeStepState's allowed step
(adaptive_time_step_state.rs:13-70).  The oracle's adaptive mode is pinned on
the CPU (tests/test_oracle_adaptive.py); here the engine's device reductions
and its host-side state follow it:
  * one substep per call: the step of every substep against the FP64
    oracle's, on a scene where the isolated-particle limit binds (a third of
    the particles fluid) and one where the velocity limit binds;
  * long calls: the same number of substeps, the same time and state;
  * mesh colliders with adaptive steps, and the twin-table redo of a substep;
  * a particle that comes close to inversion stops the call after its
    substep; a mid-call cancel leaves the state of an uncancelled run.
Stated tolerances: each step 1e-4 relative to the FP64 oracle's (measured
errors printed), x, v, F 1e-4 norm-wise relative as in the fixed-step
parity tests.
"""
import numpy as np
import pytest

from oracle import pyoracle
from squishy_volumes_amd import abi, scenes
from squishy_volumes_amd.gpu_state import GpuError, GpuRunParameters, GpuState
from squishy_volumes_amd.io_state import IoState
from tests.helpers import rel_norm_err
from tests.test_oracle_adaptive import thrown_up
from tests.test_time_step import mixed_state

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-4
TOL_STATE = 1e-4


def scene(name):
    return thrown_up() if name == "thrown_up" else mixed_state(n_cells=4)


def params(target, max_dt):
    return GpuRunParameters(target_time=target, max_time_step=max_dt, adaptive_time_steps=True)


def state_errors(st, ref64):
    p = st.particles
    return {"x": rel_norm_err(p.positions, ref64["x"]), "v": rel_norm_err(p.velocities, ref64["v"]),
            "F": rel_norm_err(p.matrix_F(), ref64["F"])}


def check_state(label, st, ref64, tol=TOL_STATE):
    errs = state_errors(st, ref64)
    print(label, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= tol, (label, errs)


@pytest.mark.parametrize("name", ["mixed", "thrown_up"])
def test_each_step_matches_the_oracle(name):
    sc = scene(name)
    max_dt = 0.01
    g = GpuState.from_io_state(0, sc.consts, sc.time_step, sc.io_state, sc.options)
    o = pyoracle.Oracle(sc.consts, sc.time_step, sc.io_state, sc.options, 8)
    o.set_adaptive(True, max_dt)
    steps = []
    for k in range(24):
        t0, to = g.time, o.time
        st, err = g.produce_next_state(sc.frame_input, params(t0 + 1e-12, max_dt))
        assert err is None and g.last_status.substeps == 1, err
        s = o.step_to(to + 1e-12, sc.frame_input)
        assert s.result == 0 and s.substeps == 1
        steps.append((g.time - t0, o.time - to))
    got, want = np.array(steps).T
    err = np.max(np.abs(got - want) / want)
    print(f"{name}: 24 steps {want.min():.4g}..{want.max():.4g} s, max relative step error {err:.2e}", flush=True)
    assert want.max() > want.min() * (1 + 1e-4)  # the steps adapt
    assert err <= TOL_STEP
    check_state(f"{name} after 24 adaptive substeps vs fp64", st, o.particles_f64())
    o.close()


@pytest.mark.parametrize("name,target", [("mixed", 0.0123), ("thrown_up", 0.0917)])
def test_a_long_call_matches_the_oracle(name, target):
    sc = scene(name)
    max_dt = 0.01
    g = GpuState.from_io_state(0, sc.consts, sc.time_step, sc.io_state, sc.options)
    st, err = g.produce_next_state(sc.frame_input, params(target, max_dt))
    assert err is None, err
    o = pyoracle.Oracle(sc.consts, sc.time_step, sc.io_state, sc.options, 8)
    o.set_adaptive(True, max_dt)
    s = o.step_to(target, sc.frame_input)
    assert s.result == 0
    print(f"{name}: {g.last_status.substeps} substeps (oracle {s.substeps}) to t = {g.time:.6g} (oracle {o.time:.6g})")
    assert g.last_status.substeps == s.substeps > 10
    assert g.time == pytest.approx(o.time, rel=TOL_STEP)
    check_state(f"{name}: one call to t = {target}", st, o.particles_f64())
    # a fixed step after the adaptive calls: the engine takes consts' step again
    st2, err = g.produce_next_state(sc.frame_input, GpuRunParameters(target_time=g.time + 2.5 * sc.time_step))
    assert err is None and g.last_status.substeps == 3
    o.close()


def run_with_mesh(c, parts, mesh, fi, target, max_dt):
    g = GpuState.from_io_state(0, c, 1e-3, IoState(0.0, parts))
    g.set_collider_mesh(mesh)
    st, err = g.produce_next_state(fi, params(target, max_dt))
    assert err is None, err
    o = pyoracle.Oracle(c, 1e-3, IoState(0.0, parts), None, 8)
    o.set_adaptive(True, max_dt)
    o.set_collider_mesh(mesh)
    s = o.step_to(target, fi)
    assert s.result == 0
    print(f"{g.last_status.substeps} substeps (oracle {s.substeps})", flush=True)
    assert g.last_status.substeps == s.substeps
    return g, st, o


def test_cube_on_a_floor_mesh_with_adaptive_steps():
    from tests.test_gpu_colliders import consts, cube_parts, floor

    parts = cube_parts((12, 12, 10), 5, (0.3, 0.0, -2.0))
    mesh, fi = floor(0.2, friction=0.6)
    c = consts()
    g, st, o = run_with_mesh(c, parts, mesh, fi, 0.15, 2e-3)
    o32 = o.particles()
    assert (o32.collider_bits != 0).sum() > 50
    assert st.particles.positions[:, 2].min() > 0.2 - 2 * c.grid_node_size
    assert int((st.particles.collider_bits != o32.collider_bits).sum()) <= 3
    check_state("cube on a floor mesh, adaptive, vs fp64", st, o.particles_f64())
    o.close()


def test_twin_table_redo_with_adaptive_steps():
    """The separate particles on a floor of the twin-overflow test: the first
    substep's table overflows and the substep is redone with the adaptive
    state as it was before it."""
    from tests.test_gpu_colliders import separate_particles_on_a_floor

    c, parts, mesh, fi = separate_particles_on_a_floor()
    g, st, o = run_with_mesh(c, parts, mesh, fi, 0.02, 1e-3)
    assert np.all(st.particles.collider_bits != 0)
    check_state("2000 particles on a floor, adaptive, vs fp64", st, o.particles_f64())
    o.close()


def test_close_to_inversion_stops_an_adaptive_call():
    """The lonely particle of the failure tests with F0_zz = 2e-7 and
    C_zz = -1500: the deformation limit 0.2 / 1500 binds, so every advance
    scales F_zz by 1 - 0.2 = 0.8, and the fourth leaves det F = 8.2e-8 below
    1e-7: the call stops after 4 substeps with the particle FAILED."""
    from tests.test_gpu_failures import cube_with_lonely_particle

    sc = cube_with_lonely_particle(2e-7, -1500.0)
    g = GpuState.from_io_state(0, sc.consts, sc.time_step, sc.io_state, sc.options)
    st, err = g.produce_next_state(sc.frame_input, params(0.05, 1e-3))
    assert err is not None and err.code == abi.ERR_CLOSE_TO_INVERTED, err
    assert g.last_status.substeps == 4
    p = st.particles
    assert p.flags[-1] & abi.FAILED
    det = np.linalg.det(p.matrix_F()[-1].astype(np.float64))
    assert det == pytest.approx(2e-7 * 0.8**4, rel=1e-3)
    dt = float(np.float32(0.2) / np.float32(1500.0))
    assert st.time == pytest.approx(4 * dt, rel=1e-5)


def test_cancel_mid_call_with_adaptive_steps():
    """A cancel before substep 9 of an adaptive call: the state equals nine
    single-substep calls' (the adaptive state carries over between calls)."""
    sc = scene("thrown_up")
    max_dt = 0.01
    g = GpuState.from_io_state(0, sc.consts, sc.time_step, sc.io_state, sc.options)
    g.debug_cancel_after(9)
    with pytest.raises(GpuError) as e:
        g.produce_next_state(sc.frame_input, params(0.2, max_dt))
    assert e.value.code == abi.ERR_CANCELED and g.last_status.substeps == 9
    st = g.to_io_state()
    r = GpuState.from_io_state(0, sc.consts, sc.time_step, sc.io_state, sc.options)
    for _ in range(9):
        ref, err = r.produce_next_state(sc.frame_input, params(r.time + 1e-12, max_dt))
        assert err is None
    assert g.time == pytest.approx(r.time, rel=1e-6)
    assert rel_norm_err(st.particles.positions, ref.particles.positions) <= 1e-6
    assert rel_norm_err(st.particles.velocities, ref.particles.velocities) <= 1e-5
    g.debug_cancel_after(-1)


def test_adaptive_arguments():
    sc = scene("thrown_up")
    g = GpuState.from_io_state(0, sc.consts, sc.time_step, sc.io_state, sc.options)
    for bad in (0.0, -1.0, float("nan")):
        with pytest.raises(GpuError):
            g.set_adaptive_time_steps(True, bad)
    g.set_adaptive_time_steps(True, float("inf"))
    g.set_adaptive_time_steps(False)


def test_a_call_past_its_frame_stops_where_the_oracle_stops():
    """The thrown-up cube at 24 frames per second with frame 0's input and a
    target past the frame's end: the substep that would start in frame 1
    fails with MPM_ERR_WRONG_FRAME (FrameInput::frame_factor,
    rust/crates/xpu/src/frame_input.rs:266-278), after the same substeps and
    at the same time as the oracle."""
    sc = scene("thrown_up")
    from squishy_volumes_amd.io_state import InputConsts

    c = sc.consts
    consts24 = InputConsts(grid_node_size=c.grid_node_size, frames_per_second=24, domain_min=c.domain_min,
                           domain_max=c.domain_max)
    g = GpuState.from_io_state(0, consts24, sc.time_step, sc.io_state, sc.options)
    with pytest.raises(GpuError) as e:
        g.produce_next_state(sc.frame_input, params(0.08, 0.01))
    assert e.value.code == abi.ERR_WRONG_FRAME
    o = pyoracle.Oracle(consts24, sc.time_step, sc.io_state, sc.options, 8)
    o.set_adaptive(True, 0.01)
    s = o.step_to(0.08, sc.frame_input)
    assert s.result == abi.ERR_WRONG_FRAME
    print(f"stopped at t = {g.time:.6g} (oracle {o.time:.6g}) after frame 0 ends at {1 / 24:.6g}")
    assert g.time == pytest.approx(o.time, rel=1e-5) and g.time >= 1 / 24 > g.time - 0.01
    o.close()
