"""The learned stress model: its CPU oracle (pinned by finite differences and
torch's BF16 conversion) and the HIP kernels against the oracle.

Stated tolerances (norm-wise relative error over whole arrays):
  * forward outputs vs the oracle with the kernels' BF16 rounding points: 1e-4
    (FP32 against FP64 accumulation, and the rare BF16 rounding of a value
    that lies within FP32 round-off of a tie);
  * loss gradient vs the same oracle: 1e-3 per parameter block, loss 1e-5;
  * one Adam step: 1e-6.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo
from squishy_volumes_amd import closure
from tests.helpers import rel_norm_err


def random_params(seed=0, scale=1.0):
    p = closure.Mlp.initial_params(seed).astype(np.float64) * scale
    rng = np.random.default_rng(seed + 100)
    # Non-zero biases so that every term of the gradient is exercised.
    p[mo.OFF_B1:mo.OFF_W2] = rng.uniform(-0.2, 0.2, mo.N_H)
    p[mo.OFF_B2:mo.OFF_W3] = rng.uniform(-0.2, 0.2, mo.N_H)
    p[mo.OFF_B3:mo.OFF_B3 + 3] = rng.uniform(-0.2, 0.2, 3)
    return p.astype(np.float32)


def random_samples(n, seed=1):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 1, (n, 6)).astype(np.float32), rng.normal(0, 1, (n, 3)).astype(np.float32)


# ------------------------------------------------------------------ oracle pins


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-3, 2000), rng.normal(0, 1e4, 2000)]).astype(np.float32)
    # exact ties: the low 16 bits 0x8000, with even and odd kept parts
    ties = (((np.arange(0, 128, dtype=np.uint32) << 16) + 0x3F800000) | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(mo.bf16(x), ref)


def test_oracle_gradient_matches_finite_differences():
    """Unrounded FP64 backpropagation against central differences of the loss."""
    p = random_params(3).astype(np.float64)
    x, t = random_samples(40, 5)
    scale = 1.0 / (40 * 3)
    _, g = mo.loss_grad(p, x, t, scale, rounded=False)
    rng = np.random.default_rng(7)
    idx = np.concatenate([rng.integers(0, mo.N_PARAM, 60), [mo.OFF_B1, mo.OFF_B3, mo.OFF_B3 + 2, mo.OFF_W3 + 5]])
    eps = 1e-6
    for k in idx:
        dp = np.zeros_like(p)
        dp[k] = eps
        lp, _ = mo.loss_grad(p + dp, x, t, scale, rounded=False)
        lm, _ = mo.loss_grad(p - dp, x, t, scale, rounded=False)
        fd = scale * (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
    # padded entries (input columns 6..15 of W1, output rows 3..15 of W3) get no gradient
    assert not np.any(g[mo.OFF_W1:mo.OFF_B1].reshape(64, 16)[:, 6:])
    assert not np.any(g[mo.OFF_W3:mo.OFF_B3].reshape(16, 64)[3:])


def test_oracle_adam_first_step_is_signed_learning_rate():
    """After one bias-corrected step, m_hat = g and v_hat = g^2: p moves by lr * g / (|g| + eps)."""
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    p, m, v = mo.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), 0.01, 0.9, 0.999, 1e-8, 1)
    assert np.allclose(p, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)


# ------------------------------------------------------------------ GPU kernels

def device_tensors(*arrays):
    import torch

    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


@pytest.mark.gpu
def test_mlp_forward_matches_oracle():
    p = random_params(11)
    for n in (1, 16, 1003, 70000):
        x, _ = random_samples(n, n)
        net = closure.Mlp()
        net.params.copy_(device_tensors(p)[0])
        y = net.forward(device_tensors(x)[0]).cpu().numpy()
        ref, _ = mo.forward(p, x, rounded=True)
        e = rel_norm_err(y, ref)
        e64 = rel_norm_err(y, mo.forward(p, x, rounded=False)[0])
        print(f"n={n}: forward vs BF16 oracle {e:.2e}, vs unrounded FP64 {e64:.2e}")
        assert e <= 1e-4, (n, e)


@pytest.mark.gpu
def test_mlp_gradient_matches_oracle():
    p = random_params(21)
    n = 5003
    x, t = random_samples(n, 9)
    scale = 1.0 / (n * 3)
    net = closure.Mlp()
    xd, td, pd = device_tensors(x, t, p)
    net.params.copy_(pd)
    net.accumulate(xd, td, scale)
    g = net.grad.cpu().numpy()
    loss = float(net.loss.item())
    ref_loss, ref_g = mo.loss_grad(p, x, t, scale, rounded=True)
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss, (loss, ref_loss)
    blocks = {"W1": (mo.OFF_W1, mo.OFF_B1), "b1": (mo.OFF_B1, mo.OFF_W2), "W2": (mo.OFF_W2, mo.OFF_B2),
              "b2": (mo.OFF_B2, mo.OFF_W3), "W3": (mo.OFF_W3, mo.OFF_B3), "b3": (mo.OFF_B3, mo.N_PARAM)}
    for name, (a, b) in blocks.items():
        e = rel_norm_err(g[a:b], ref_g[a:b])
        print(f"gradient {name}: {e:.2e}")
        assert e <= 1e-3, (name, e)
    # padding stays exactly zero
    assert not np.any(g[mo.OFF_W1:mo.OFF_B1].reshape(64, 16)[:, 6:])
    assert not np.any(g[mo.OFF_W3:mo.OFF_B3].reshape(16, 64)[3:])


@pytest.mark.gpu
def test_mlp_adam_matches_oracle():
    rng = np.random.default_rng(4)
    p, g = random_params(2), rng.normal(0, 1, mo.N_PARAM).astype(np.float32)
    m, v = rng.normal(0, 0.1, mo.N_PARAM).astype(np.float32), rng.uniform(0, 0.1, mo.N_PARAM).astype(np.float32)
    net = closure.Mlp()
    for dst, src in zip((net.params, net.grad, net.m, net.v), device_tensors(p, g, m, v)):
        dst.copy_(src)
    net.step = 6
    net.adam(1e-3, 0.9, 0.999, 1e-8)
    rp, rm, rv = mo.adam(p.astype(np.float64), g, m, v, 1e-3, 0.9, 0.999, 1e-8, 7)
    assert rel_norm_err(net.params.cpu().numpy(), rp) <= 1e-6
    assert rel_norm_err(net.m.cpu().numpy(), rm) <= 1e-6
    assert rel_norm_err(net.v.cpu().numpy(), rv) <= 1e-6
    assert not np.any(net.grad.cpu().numpy())


def snow_samples(n, seed):
    """Singular values inside the snow model's clamp, E, nu, Jp as a snow scene has them."""
    rng = np.random.default_rng(seed)
    s = -np.sort(-rng.uniform(1 - 2.5e-2, 1 + 4.5e-3, (n, 3)), axis=1)
    E = 10 ** rng.uniform(4, 6, n)
    nu = rng.uniform(0.2, 0.4, n)
    jp = rng.uniform(0.8, 1.05, n)
    return s, E, nu, jp


@pytest.mark.gpu
def test_training_learns_the_snow_stress():
    import torch

    s, E, nu, jp = snow_samples(200000, 0)
    tau = closure.analytic_principal_stress(s, E, nu, jp, 10.0)
    x, t = device_tensors(closure.features(s, E, nu, jp), closure.targets(tau, s, E, nu, jp, 10.0))
    net = closure.Mlp(seed=1)
    before = rel_norm_err(net.forward(x).cpu().numpy(), t.cpu().numpy())
    log = net.fit(x, t, epochs=60, batch=4096, lr=3e-3, lr_final=1e-4)
    sv, Ev, nuv, jpv = snow_samples(50000, 1)
    tv = closure.analytic_principal_stress(sv, Ev, nuv, jpv, 10.0)
    yv = net.forward(device_tensors(closure.features(sv, Ev, nuv, jpv))[0]).cpu().numpy()
    err = rel_norm_err(closure.stress_from_outputs(yv, sv, Ev, nuv, jpv, 10.0), tv)
    print(f"relative stress error: before {before:.3f}, after {err:.2e}; epoch losses {log.epochs[0]:.3e} .. {log.epochs[-1]:.3e}")
    assert log.epochs[-1] < 0.01 * log.epochs[0]
    assert err <= 2.5e-2
    torch.cuda.synchronize()


def test_targets_invert_and_the_remainder_is_first_order_in_the_strain():
    """The closure's outputs: targets() and stress_from_outputs() are inverse,
    the linear part alone is the stress to first order, and the target (the
    second-order remainder over the strain magnitude) shrinks with the strain."""
    s, E, nu, jp = snow_samples(2000, 3)
    tau = closure.analytic_principal_stress(s, E, nu, jp, 10.0)
    t = closure.targets(tau, s, E, nu, jp, 10.0)
    back = closure.stress_from_outputs(t, s, E, nu, jp, 10.0)
    assert rel_norm_err(back, tau) <= 1e-6
    for k in (0.1, 0.01):
        sk = 1.0 + k * (s - 1.0)
        tk = closure.targets(closure.analytic_principal_stress(sk, E, nu, jp, 10.0), sk, E, nu, jp, 10.0)
        ratio = np.linalg.norm(tk) / np.linalg.norm(t)
        assert 0.5 * k <= ratio <= 2.0 * k, (k, ratio)
        lin = closure.linear_principal_stress(sk, E, nu, jp, 10.0)
        tau_k = closure.analytic_principal_stress(sk, E, nu, jp, 10.0)
        assert rel_norm_err(lin, tau_k) <= 0.1 * k
