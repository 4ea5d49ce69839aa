"""The substep loop with whole sort cycles replayed as HIP graphs against the
same loop with every kernel launched on its own.

Each cycle's launches read their substep's parameters from the device table,
so a replay differs from the eager launches in nothing but the launch path;
what remains is the run-to-run order of the float atomics.  Stated
tolerance: x, v, F within 1e-5 (norm-wise relative) of the eager run after
every call, over calls of uneven lengths (cycles that straddle calls, both
particle buffers, gravity that changes every substep through the frame
factor, goals switched on after graphs were captured; velocities there within
1e-3, since a goal velocity (goal - x) / dt magnifies the positions'
round-off by 1 / dt), and the FP64 oracle's 1e-4 at the end."""
import numpy as np
import pytest

from squishy_volumes_amd import abi, scenes
from squishy_volumes_amd.gpu_state import GpuRunParameters
from squishy_volumes_amd.io_state import FrameInput
from tests.helpers import rel_norm_err
from tests.test_gpu_step import compare_states, engine, oracles, perturbed_cube

pytestmark = pytest.mark.gpu

CALLS = (37, 50, 5, 64, 23)


def close(a, b, label, tol=1e-5, tol_v=None):
    errs = {k: rel_norm_err(getattr(a, f), getattr(b, f))
            for k, f in (("x", "positions"), ("v", "velocities"), ("F", "position_gradients"))}
    print(label, {k: f"{v:.1e}" for k, v in errs.items()})
    assert errs["x"] <= tol and errs["F"] <= tol, (label, errs)
    assert errs["v"] <= (tol_v or tol), (label, errs)


@pytest.mark.parametrize("interval", [1, 3, 8])
def test_graph_cycles_match_eager_launches(interval):
    sc = perturbed_cube(n_cells=6)
    fi = FrameInput(frame=0, gravity_a=(0.0, 0.0, -9.81), gravity_b=(4.0, -2.0, -3.0))
    g = engine(sc, tuning={"resort_interval": interval})
    e = engine(sc, tuning={"resort_interval": interval, "graph": -1})
    t, launched = 0.0, 0
    for k in CALLS:
        t = scenes.advance_time(t, sc.time_step, k)
        sg, err = g.produce_next_state(fi, GpuRunParameters(target_time=t))
        assert err is None and g.last_status.substeps == k
        launched += int(g.last_status.graph_launches)
        se, err = e.produce_next_state(fi, GpuRunParameters(target_time=t))
        assert err is None and e.last_status.graph_launches == 0
        close(sg.particles, se.particles, f"interval {interval}, after {k} substeps")
    # every whole cycle strictly inside a call after its first is a graph launch
    expected = sum(max(0, (k - 1) // interval - 1) for k in CALLS)
    assert launched == expected, (launched, expected)
    _, o64 = oracles(sc)
    o64.step_to(t, fi)
    compare_states(sg.particles, o64.particles_f64(), sc.consts.grid_node_size, f"graph loop, interval {interval}")


def test_goals_after_captured_cycles():
    sc = perturbed_cube(n_cells=4)
    p = sc.io_state.particles
    g = engine(sc, tuning={"resort_interval": 4})
    e = engine(sc, tuning={"resort_interval": 4, "graph": -1})
    plain = FrameInput(frame=0, gravity_a=(0.0, 0.0, -9.81))
    t = scenes.advance_time(0.0, sc.time_step, 30)
    for eng in (g, e):
        eng.produce_next_state(plain, GpuRunParameters(target_time=t))
    assert g.last_status.graph_launches > 0
    fl = np.zeros(p.count, np.uint32)
    fl[::3] = abi.HAS_GOAL
    goal = p.positions + np.float32(0.01)
    fi = FrameInput(frame=0, gravity_a=(0.0, 0.0, -9.81), gravity_b=(0.0, 0.0, -9.81), particle_flags_a=fl,
                    particle_flags_b=fl, goal_positions_a=goal, goal_positions_b=goal)
    t = scenes.advance_time(t, sc.time_step, 30)
    sg, _ = g.produce_next_state(fi, GpuRunParameters(target_time=t))
    se, _ = e.produce_next_state(fi, GpuRunParameters(target_time=t))
    assert g.last_status.graph_launches > 0
    # A goal velocity (goal - x) / dt divides the positions' round-off by dt.
    close(sg.particles, se.particles, "goals after captured cycles", tol_v=1e-3)
