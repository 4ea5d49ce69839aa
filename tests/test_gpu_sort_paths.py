"""The engine's two re-sorts, through the C ABI.

The re-sort orders the particles by Morton block key (the reference sorts to
optimise memory access only, rust/crates/cpu/src/phase/sort.rs:16).  Keys
computed from the positions (after create) take the stable radix sort (which
the radix-sort tests hold bit-exact against a stable CPU sort); keys that a
transfer wrote take a counting sort over every key when the key range has at
most 2^22 keys: the particles that stayed in their block keep their order,
the ones that moved in follow them.  The same moving cube in a 1 m domain
(15-bit keys: counting sorts) and in a 9 m long one (24-bit keys: radix
sorts only) must step alike: the two orders differ only within a block, which
changes nothing but the order of the scatter's atomic sums.  Stated
tolerance: x, v, F 1e-5 norm-wise relative between the two runs (the order of
the sums), and both within 1e-4 of the FP64 oracle.  A block of 3008
particles, sorted before every substep, against the FP64 oracle too.
"""
import numpy as np
import pytest

from oracle import pyoracle
from squishy_volumes_amd import abi, scenes
from squishy_volumes_amd.gpu_state import GpuRunParameters, GpuState
from squishy_volumes_amd.io_state import FrameInput, InputConsts, IoState
from tests.helpers import rel_norm_err

pytestmark = pytest.mark.gpu


def moving_cube(domain_max_x):
    return scenes.cube_scene(n_cells=8, h=1.0 / 64, corner_cells=(24, 24, 16), E=2e4, velocity=(3.0, -2.0, 1.0),
                             dt=2e-4, domain=((0.0, 0.0, 0.0), (domain_max_x, 1.0, 1.0)), seed=11)


def run(sc, substeps):
    g = GpuState.from_io_state(0, sc.consts, sc.time_step, sc.io_state, sc.options,
                               {"resort_interval": 2, "grid_window": -1})
    status = g.step_to(scenes.advance_time(0.0, sc.time_step, substeps), sc.frame_input)
    return status.grid_blocks, g.to_io_state()


def test_counting_sort_and_radix_sort_step_alike():
    small, long = moving_cube(1.0), moving_cube(9.0)
    blocks1, s1 = run(small, 40)
    blocks2, s2 = run(long, 40)
    # the grids share their y and z extents: the long one's x extent from its block count
    nb = int(round(blocks1 ** (1 / 3)))
    assert nb**3 == blocks1 and blocks2 % (nb * nb) == 0
    bits = [3 * int(np.ceil(np.log2(n))) for n in (nb, blocks2 // (nb * nb))]
    print("key bits", bits, flush=True)
    assert bits[0] <= 22 < bits[1]
    p1, p2 = s1.particles, s2.particles
    for k, a, b in (("x", p1.positions, p2.positions), ("v", p1.velocities, p2.velocities),
                    ("F", p1.matrix_F(), p2.matrix_F())):
        err = rel_norm_err(a, b.astype(np.float64))
        print(f"counting vs radix sort, {k}: {err:.2e}", flush=True)
        assert err <= 1e-5, (k, err)
    o = pyoracle.Oracle(small.consts, small.time_step, small.io_state, small.options, 8)
    o.step_to(scenes.advance_time(0.0, small.time_step, 40), small.frame_input)
    ref = o.particles_f64()
    for label, p in (("counting", p1), ("radix", p2)):
        errs = {"x": rel_norm_err(p.positions, ref["x"]), "v": rel_norm_err(p.velocities, ref["v"]),
                "F": rel_norm_err(p.matrix_F(), ref["F"])}
        print(f"{label} sort vs fp64", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert max(errs.values()) <= 1e-4, (label, errs)
    o.close()


def test_a_block_of_3008_particles_matches_the_oracle():
    """3008 particles in the 64 base cells of one 4^3 block (47 per cell), a
    sort before every substep: ten substeps against the FP64 oracle."""
    h = 1.0 / 64
    rng = np.random.default_rng(2)
    pts = rng.uniform(24.6, 28.4, (3008, 3)) * h
    parts = scenes.particles_from(pts, abi.IS_SOLID, 1000.0, h, (0.3, -0.2, 0.1), E=1e4, nu=0.3)
    consts = InputConsts(grid_node_size=h, frames_per_second=24, domain_min=(0.0, 0.0, 0.0),
                         domain_max=(1.0, 1.0, 1.0))
    fi = FrameInput(frame=0, gravity_a=(0.0, 0.0, -9.81))
    g = GpuState.from_io_state(0, consts, 1e-4, IoState(0.0, parts), None, {"resort_interval": 1})
    T = scenes.advance_time(0.0, 1e-4, 10)
    st, err = g.produce_next_state(fi, GpuRunParameters(target_time=T))
    assert err is None, err
    assert g.last_status.active_blocks == 1
    o = pyoracle.Oracle(consts, 1e-4, IoState(0.0, parts), None, 8)
    o.step_to(T, fi)
    ref = o.particles_f64()
    p = st.particles
    errs = {"x": rel_norm_err(p.positions, ref["x"]), "v": rel_norm_err(p.velocities, ref["v"]),
            "F": rel_norm_err(p.matrix_F(), ref["F"])}
    print("3008 particles in one block vs fp64", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= 1e-4, errs
    o.close()
