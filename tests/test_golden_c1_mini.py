"""The small plain case against its committed FP64 golden vectors
(tests/golden/c1_mini_fp64.npz, written by tests/golden/make_c1_mini.py):
the oracle reproduces them (CPU), and the engine holds the substep tolerance
against them (GPU)."""
import numpy as np
import pytest

from tests.golden import make_c1_mini as gold
from tests.helpers import rel_norm_err

GOLDEN = np.load(gold.OUT)


def test_inputs_are_the_seeded_scene():
    sc = gold.scene()
    assert np.array_equal(sc.io_state.particles.positions, GOLDEN["x0"])


def test_oracle_reproduces_the_golden_vectors():
    _, T, ref = gold.run()
    assert T == float(GOLDEN["time"])
    for k in ("x", "v", "F", "C"):
        assert rel_norm_err(ref[k], GOLDEN[k]) <= 1e-12, k


@pytest.mark.gpu
def test_engine_matches_the_golden_vectors():
    """x, v and F within the substep tolerance. The cube falls as a rigid
    body, so its C is near zero: C is held to 1e-3 of its norm plus the FP32
    round-off floor of a velocity gradient (4/h * 1e-6 * |v|)."""
    from squishy_volumes_amd.gpu_state import GpuRunParameters
    from tests.test_gpu_step import compare_states, engine

    sc = gold.scene()
    g = engine(sc)
    st, err = g.produce_next_state(sc.frame_input, GpuRunParameters(target_time=float(GOLDEN["time"])))
    assert err is None
    compare_states(st.particles, {k: GOLDEN[k] for k in ("x", "v", "F", "C")}, sc.consts.grid_node_size,
                   "engine vs golden fp64")
