"""The learned constitutive closure: training and evaluation on the GPU.

The reference has no learned stress model.  Here a 6-64-64-3 ReLU perceptron
maps the sorted singular values of F and the material parameters of a snow
particle to its principal Kirchhoff stresses, standing in for the analytic
fixed-corotated snow stress (tau = U diag(2 mu (s - 1) s + lambda (J - 1) J) U^T
with mu, lambda hardened by exp(xi (1 - Jp)); tau_fixed_corotated_svd in
csrc/device_math.hpp).  Forward, backward and Adam run in the engine's HIP
kernels (csrc/closure.hip, BF16 matrix-core products with FP32 accumulation)
through the C ABI; this module holds the parameters in device memory and
drives them.  There is no CPU fallback: the engine library must be built.

Features (raw -> network input):
    s1 >= s2 >= s3 (singular values)   (s - 1) / SIGMA_SCALE
    Young's modulus E                    log10(E / 1e5)
    Poisson's ratio nu                   (nu - 0.3) / 0.1
    plastic Jp                           (Jp - 1) / JP_SCALE
Outputs: the principal stresses minus their linear-elastic part,
(tau_i - h (2 mu (s_i - 1) + lambda tr(s - 1))) / (E h STRESS_SCALE m) with
h = exp(xi (1 - Jp)) and m = max(|s - 1| / SIGMA_SCALE, 1e-6), in the order
of s.  The closure adds the linear part in FP32, exactly; the network learns
the remainder, which is second order in the strain, divided by the strain
magnitude (STRESS_SCALE = 0.025^2 makes it order one at the compression
limit), in units of the hardened modulus, so that it learns the dependence
on s and nu, not the modulus scale.  The error of the network (and of its
BF16 arithmetic) then scales with the strain times the remainder's share of
the stress, not with the stress.
The engine evaluates the same inputs and outputs (include/mpm_engine.h,
mpm_set_closure).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi

N_IN, N_H, N_OUT = 6, 64, 3
SIGMA_SCALE = 0.025
JP_SCALE = 0.1
STRESS_SCALE = 0.025 * 0.025


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def lame_from_young(E, nu):
    E = np.asarray(E, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def young_from_lame(mu, lam):
    mu = np.asarray(mu, dtype=np.float64)
    lam = np.asarray(lam, dtype=np.float64)
    return mu * (3 * lam + 2 * mu) / (lam + mu), lam / (2 * (lam + mu))


def analytic_principal_stress(s, E, nu, jp, hardening: float) -> np.ndarray:
    """Principal Kirchhoff stresses of the fixed-corotated snow model for
    singular values s [n][3] (tau_fixed_corotated_svd, csrc/device_math.hpp)."""
    s = np.asarray(s, dtype=np.float64)
    mu, lam = lame_from_young(E, nu)
    h = np.exp(hardening * (1.0 - np.asarray(jp, dtype=np.float64)))
    J = s.prod(axis=1)
    return 2 * (mu * h)[:, None] * (s - 1) * s + ((lam * h) * (J - 1) * J)[:, None]


def features(s, E, nu, jp) -> np.ndarray:
    """Network inputs [n][6] from sorted singular values and parameters."""
    s = np.asarray(s, dtype=np.float64)
    x = np.empty((s.shape[0], N_IN), dtype=np.float32)
    x[:, :3] = (s - 1.0) / SIGMA_SCALE
    x[:, 3] = np.log10(np.asarray(E, dtype=np.float64) / 1e5)
    x[:, 4] = (np.asarray(nu, dtype=np.float64) - 0.3) / 0.1
    x[:, 5] = (np.asarray(jp, dtype=np.float64) - 1.0) / JP_SCALE
    return x


def stress_unit(s, E, jp, hardening: float) -> np.ndarray:
    """E h STRESS_SCALE m: the unit of the network's outputs."""
    s = np.asarray(s, dtype=np.float64)
    m = np.maximum(np.linalg.norm(s - 1.0, axis=1) / SIGMA_SCALE, 1e-6)
    return (np.asarray(E, dtype=np.float64) * np.exp(hardening * (1.0 - np.asarray(jp, dtype=np.float64)))
            * STRESS_SCALE * m)


def linear_principal_stress(s, E, nu, jp, hardening: float) -> np.ndarray:
    """The linear-elastic part h (2 mu (s - 1) + lambda tr(s - 1)) of the
    principal stresses, which the closure computes exactly."""
    s = np.asarray(s, dtype=np.float64)
    mu, lam = lame_from_young(E, nu)
    h = np.exp(hardening * (1.0 - np.asarray(jp, dtype=np.float64)))
    e = s - 1.0
    return (2 * (mu * h))[:, None] * e + ((lam * h) * e.sum(axis=1))[:, None]


def targets(tau, s, E, nu, jp, hardening: float) -> np.ndarray:
    """Network targets of principal stresses tau [n][3]."""
    rem = np.asarray(tau, dtype=np.float64) - linear_principal_stress(s, E, nu, jp, hardening)
    return (rem / stress_unit(s, E, jp, hardening)[:, None]).astype(np.float32)


def stress_from_outputs(y, s, E, nu, jp, hardening: float) -> np.ndarray:
    """Principal stresses from network outputs y [n][3], as the closure forms them."""
    return (np.asarray(y, dtype=np.float64) * stress_unit(s, E, jp, hardening)[:, None]
            + linear_principal_stress(s, E, nu, jp, hardening))


@dataclass
class Pairs:
    """Training pairs: network inputs x [n][6], targets t [n][3] and the raw
    values they come from."""
    x: np.ndarray
    t: np.ndarray
    s: np.ndarray
    E: np.ndarray
    nu: np.ndarray
    jp: np.ndarray

    def __len__(self):
        return self.x.shape[0]

    @staticmethod
    def concat(parts):
        return Pairs(*(np.concatenate([getattr(p, f) for p in parts]) for f in ("x", "t", "s", "E", "nu", "jp")))

    def subset(self, idx):
        return Pairs(self.x[idx], self.t[idx], self.s[idx], self.E[idx], self.nu[idx], self.jp[idx])


def pairs_from_particles(particles, hardening: float, flags_mask: int = abi.SNOW) -> Pairs:
    """(F, stress) pairs of the live snow particles of an IoState's particles:
    the singular values of F (the elastic F that the engine stores), E, nu
    and Jp, with the analytic stress as the target."""
    live = ((particles.flags & flags_mask) != 0) & ((particles.flags & abi.TOMBSTONED) == 0)
    F = particles.matrix_F()[live].astype(np.float64)
    s = np.linalg.svd(F, compute_uv=False)
    mu = particles.parameters["mu"][live].astype(np.float64)
    lam = particles.parameters["lambda"][live].astype(np.float64)
    E, nu = young_from_lame(mu, lam)
    jp = particles.plastic_jp[live].astype(np.float64) if particles.plastic_jp is not None else np.ones(int(live.sum()))
    tau = analytic_principal_stress(s, E, nu, jp, hardening)
    return Pairs(features(s, E, nu, jp), targets(tau, s, E, nu, jp, hardening), s, E, nu, jp)


@dataclass
class ClosureModel:
    """Trained parameters (padded FP32 layout) and the training hull."""
    params: np.ndarray
    hull_lo: np.ndarray
    hull_hi: np.ndarray

    @staticmethod
    def hull_of(x: np.ndarray, margin: float = 0.01, floor: float = 1e-3):
        """Per-input range of the training inputs, widened by margin of its
        width and at least floor (an input that is constant in the training
        set, such as E of a one-material scene, is computed on the device in
        FP32 and may differ from the host's value in the last bits)."""
        if hasattr(x, "amin"):  # a device tensor: the range on the device
            lo, hi = x.amin(0).double().cpu().numpy(), x.amax(0).double().cpu().numpy()
        else:
            lo, hi = x.min(axis=0).astype(np.float64), x.max(axis=0).astype(np.float64)
        pad = np.maximum(margin * (hi - lo), floor)
        return (lo - pad).astype(np.float32), (hi + pad).astype(np.float32)

    @staticmethod
    def hull_union(a, b):
        """The smallest box holding the hulls a = (lo, hi) and b."""
        return np.minimum(a[0], b[0]), np.maximum(a[1], b[1])

    def save(self, path):
        np.savez(path, params=self.params, hull_lo=self.hull_lo, hull_hi=self.hull_hi)

    @staticmethod
    def load(path) -> "ClosureModel":
        d = np.load(path)
        return ClosureModel(d["params"], d["hull_lo"], d["hull_hi"])


@dataclass
class TrainLog:
    epochs: list = field(default_factory=list)  # mean squared error per epoch


class Mlp:
    """Perceptron parameters in device memory (FP32 master copy, padded
    layout of include/mpm_engine.h) with Adam state."""

    def __init__(self, seed: int = 0):
        import torch

        self.torch = torch
        self.lib = abi.load_library()
        self.n_param = int(self.lib.mpm_mlp_param_count())
        dev = torch.device("cuda", torch.cuda.current_device())
        self.params = torch.from_numpy(self.initial_params(seed)).to(dev)
        self.grad = torch.zeros(self.n_param, dtype=torch.float32, device=dev)
        self.m = torch.zeros_like(self.grad)
        self.v = torch.zeros_like(self.grad)
        self.loss = torch.zeros(1, dtype=torch.float32, device=dev)
        self.step = 0

    @staticmethod
    def initial_params(seed: int) -> np.ndarray:
        """He-uniform weights, zero biases, zero padding."""
        rng = np.random.default_rng(seed)
        p = np.zeros(64 * 16 + 64 + 64 * 64 + 64 + 16 * 64 + 16, dtype=np.float32)
        o_b1 = 64 * 16
        o_w2 = o_b1 + 64
        o_b2 = o_w2 + 64 * 64
        o_w3 = o_b2 + 64
        p[:o_b1].reshape(64, 16)[:, :N_IN] = rng.uniform(-1, 1, (64, N_IN)) * np.sqrt(6.0 / N_IN)
        p[o_w2:o_b2] = (rng.uniform(-1, 1, (64, 64)) * np.sqrt(6.0 / 64)).ravel()
        p[o_w3:o_w3 + 16 * 64].reshape(16, 64)[:N_OUT] = rng.uniform(-1, 1, (N_OUT, 64)) * np.sqrt(3.0 / 64)
        return p

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"closure kernel failed: {self.lib.mpm_error_string(rc).decode()}")

    def forward(self, x):
        """Outputs [n][3] for inputs x [n][6] (float32 device tensor)."""
        torch = self.torch
        x = x.contiguous()
        y = torch.empty((x.shape[0], N_OUT), dtype=torch.float32, device=x.device)
        self._check(self.lib.mpm_mlp_forward(abi.ptr(self.params), abi.ptr(x), abi.ptr(y), x.shape[0], _stream(torch)))
        return y

    def accumulate(self, x, t, scale: float):
        """Adds scale * d/dp sum((y - t)^2) to self.grad and the sum to self.loss."""
        self._check(
            self.lib.mpm_mlp_grad(abi.ptr(self.params), abi.ptr(x.contiguous()), abi.ptr(t.contiguous()), x.shape[0],
                                  float(scale), abi.ptr(self.grad), abi.ptr(self.loss), _stream(self.torch))
        )

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        self.step += 1
        self._check(
            self.lib.mpm_mlp_adam(abi.ptr(self.params), abi.ptr(self.grad), abi.ptr(self.m), abi.ptr(self.v), float(lr),
                                  float(beta1), float(beta2), float(eps), self.step, _stream(self.torch))
        )

    def model(self, x_train) -> ClosureModel:
        lo, hi = ClosureModel.hull_of(x_train)
        return ClosureModel(self.params.cpu().numpy().copy(), lo, hi)

    def device_model(self, hull) -> ClosureModel:
        """The current parameters (a device copy, handed to the engine without
        a trip through host memory) with the training hull `hull` = (lo, hi)."""
        return ClosureModel(self.params.detach().clone(), np.asarray(hull[0], np.float32),
                            np.asarray(hull[1], np.float32))

    def fit(self, x, t, epochs: int, batch: int = 4096, lr: float = 2e-3, lr_final: float = 1e-4, seed: int = 0) -> TrainLog:
        """Mini-batch Adam on (x, t) (device tensors), learning rate decaying
        geometrically from lr to lr_final; returns the mean squared error of
        each epoch."""
        torch = self.torch
        n = x.shape[0]
        g = torch.Generator(device=x.device)
        g.manual_seed(seed)
        steps = epochs * ((n + batch - 1) // batch)
        decay = (lr_final / lr) ** (1.0 / max(steps - 1, 1))
        log = TrainLog()
        k = 0
        for _ in range(epochs):
            perm = torch.randperm(n, device=x.device, generator=g)
            self.loss.zero_()
            for b0 in range(0, n, batch):
                idx = perm[b0:b0 + batch]
                xb, tb = x[idx], t[idx]
                self.accumulate(xb, tb, 1.0 / (xb.shape[0] * N_OUT))
                self.adam(lr * decay**k)
                k += 1
            log.epochs.append(float(self.loss.item()) / (n * N_OUT))
        return log
