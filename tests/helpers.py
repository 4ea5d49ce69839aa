"""Shared helpers of the test suite."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def reference_fixtures() -> dict:
    return json.loads((GOLDEN / "reference_fixtures.json").read_text())


def random_position_gradients(n: int, seed: int = 42) -> np.ndarray:
    """F with entries U[0, 1), |det F| in (0.1, 10), sign-fixed to det > 0: the
    distribution of test_position_gradients_random
    (rust/crates/gpu/src/test_data/particles.rs:32-56), from numpy's PCG64."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        F = rng.random((3, 3))
        d = abs(np.linalg.det(F))
        if 1e-1 < d < 1e1:
            if np.linalg.det(F) < 0:
                F = -F
            out.append(F)
    return np.array(out)


def rel_norm_err(a, b) -> float:
    """||a - b|| / max(||b||, tiny) over whole arrays."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def check_by_norm(a, b, epsilon=1e-6, max_relative=0.01):
    """check_iters_by_norm (rust/crates/gpu/src/test_util.rs:2738-2764)."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    scale = max(np.linalg.norm(a), np.linalg.norm(b))
    err = np.linalg.norm(a - b)
    assert err <= epsilon + max_relative * scale, f"error={err}, tolerance={epsilon + max_relative * scale}"


def rotation(axis, angle) -> np.ndarray:
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K
