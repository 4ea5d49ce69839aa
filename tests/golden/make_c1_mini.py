"""Golden vectors of the small plain case: the FP64 oracle's state after 10
substeps of a 4^3-cell Neo-Hookean cube at 8 particles per cell (512
particles, h = 1/64, dt = 1e-4, v0 = (0, 0, -2), no plane; the parity form of
the 32,768-particle cube scaled down), with the inputs regenerated from the
seeded scene generator.  Run from the repository root:

    python tests/golden/make_c1_mini.py

The oracle is pinned by the closed forms and the reference fixtures its own
tests check; these vectors freeze its output so that a later change to the
oracle, or to the scene generator, shows up as a difference."""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import pyoracle  # noqa: E402
from squishy_volumes_amd import scenes  # noqa: E402

SUBSTEPS = 10
OUT = Path(__file__).resolve().parent / "c1_mini_fp64.npz"


def scene():
    return scenes.plain_cube_parity(n_cells=4)


def run():
    sc = scene()
    T = scenes.advance_time(0.0, sc.time_step, SUBSTEPS)
    o = pyoracle.Oracle(sc.consts, sc.time_step, sc.io_state, sc.options, 8)
    o.step_to(T, sc.frame_input)
    ref = o.particles_f64()
    o.close()
    return sc, T, ref


if __name__ == "__main__":
    sc, T, ref = run()
    np.savez_compressed(OUT, time=T, x0=sc.io_state.particles.positions, **{k: ref[k] for k in ("x", "v", "F", "C")})
    print(f"wrote {OUT} ({sc.io_state.particles.count} particles, {SUBSTEPS} substeps)")
