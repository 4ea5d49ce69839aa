"""CPU oracle of the adaptive time-step limits (TEST INFRASTRUCTURE ONLY).

Only the tests may use this module; the engine never does.

A numpy FP64 restatement of CpuState's limiters,
rust/crates/cpu/src/phase/limit_time_step.rs:
  limit_time_step_by_speed_of_sound      :36-130
  limit_time_step_by_isolated_particles  :132-185
  limit_time_step_by_velocity            :199-208
  limit_time_step_by_deformation         :210-222
with the stress derivatives in diagonal space of
rust/crates/util/src/elastic.rs:97-141 (invariants), :276-295 and :314-339
(Neo-Hookean), :555-570 and :630-656 (inviscid), and
SINGULAR_VALUE_SEPARATION = 1e-5 (rust/crates/util/src/consts.rs:14).
Singular values are non-negative and sorted in descending order, as
nalgebra 0.33's svd() returns them (numpy's svd does the same).

Parity pins (tests/test_time_step.py): for an undeformed solid the sound
speed is the P-wave speed sqrt((lambda + 2 mu) / rho), for a fluid at J = 1
the Tait speed sqrt(gamma K / rho); the isolated-particle, velocity and
deformation limits are checked against their closed forms.

Particles are given as in the engine's state: flags, parameters (the
mpm_particle_parameters records), F [n][3][3] (row, col), v [n][3], C [n][3][3].
"""
from __future__ import annotations

import numpy as np

IS_FLUID = 1 << 1
TOMBSTONED = 1 << 5
SEPARATION = 1e-5


def _first_second(fluid: bool, p, s):
    x, y, z = s
    J = x * y * z
    g = np.array([y * z, x * z, x * y])
    H = np.array([[0.0, z, y], [z, 0.0, x], [y, x, 0.0]])
    if fluid:
        K, gamma = float(p["bulk_modulus"]), int(p["exponent"])
        d1 = K * (1.0 - 1.0 / J**gamma)
        d2 = gamma * K / J ** (gamma + 1)
        first = d1 * g
        second = d2 * np.outer(g, g) + d1 * H
    else:
        mu, lam = float(p["mu"]), float(p["lambda"])
        d1 = (lam * np.log(J) - mu) / J
        d2 = (lam * (1.0 - np.log(J)) + mu) / J**2
        first = (mu / 2.0) * 2.0 * s + d1 * g
        second = np.eye(3) * (mu / 2.0 * 2.0) + d2 * np.outer(g, g) + d1 * H
    return J, first, second


def sound_limit(fluid: bool, p, F, h: float) -> float:
    s = np.linalg.svd(np.asarray(F, np.float64), compute_uv=False)
    J, P, M = _first_second(fluid, p, s)
    x, y, z = s
    # nalgebra m11 = M[0, 0], m21 = M[1, 0], m32 = M[2, 1], m13 = M[0, 2]
    kx = (P[0] + x * M[0, 0] - y * M[1, 0]) / 2 / x if abs(x - y) < SEPARATION else (x * P[0] - y * P[1]) / (x * x - y * y)
    ky = (P[1] + y * M[1, 1] - z * M[2, 1]) / 2 / y if abs(y - z) < SEPARATION else (y * P[1] - z * P[2]) / (y * y - z * z)
    kz = (P[2] + x * M[2, 2] - x * M[0, 2]) / 2 / z if abs(z - x) < SEPARATION else (z * P[2] - x * P[0]) / (z * z - x * x)
    kappa = max(x * x * M[0, 0], y * y * M[1, 1], z * z * M[2, 2], y * y * kx, y * z * ky, z * x * kz) / J
    rho = float(p["mass"]) / float(p["initial_volume"]) / J
    c2 = kappa / rho
    return h / np.sqrt(c2) if c2 > 0 else np.inf


def isolated_limit(fluid: bool, p, F, h: float) -> float:
    m, vol = float(p["mass"]), float(p["initial_volume"])
    if not fluid:
        xi = 3.0 / h / h
        return float(np.sqrt(m / (vol * xi * (1.0 - 1.0 / 2.0) * (float(p["mu"]) + 3.0 / 2.0 * float(p["lambda"])))))
    rho0 = m / vol
    j = float(np.linalg.det(np.asarray(F, np.float64)))
    K, gamma = float(p["bulk_modulus"]), int(p["exponent"])
    first = K * (1.0 - 1.0 / j**gamma)
    if abs(j - 1.0) > SEPARATION:
        return h / j * float(np.sqrt(rho0 * (j - 1.0) / (6.0 * first * 3.0)))
    second = gamma * K / j ** (gamma + 1)
    return h * float(np.sqrt(rho0 / (6.0 * second * 3.0)))


def limits(flags, params, F, v, C, h: float) -> dict:
    live = (np.asarray(flags) & TOMBSTONED) == 0
    idx = np.nonzero(live)[0]
    snd = min((sound_limit(bool(flags[i] & IS_FLUID), params[i], F[i], h) for i in idx), default=np.inf)
    iso = min((isolated_limit(bool(flags[i] & IS_FLUID), params[i], F[i], h) for i in idx), default=np.inf)
    vmax = float(np.linalg.norm(np.asarray(v, np.float64)[live], axis=1).max()) if idx.size else 0.0
    vel = 0.5 * h / vmax if vmax != 0.0 else np.inf
    Cl = np.abs(np.asarray(C, np.float64)[live]).reshape(-1)
    dfm = float((0.2 / np.maximum(Cl, 1e-8)).min()) if idx.size else np.inf
    return {"by_velocity": vel, "by_deformation": dfm, "by_sound": snd, "by_isolated": iso}
