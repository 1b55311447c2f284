"""Pins of the FP64 twist-network oracle (oracle/twistnet.py), on the CPU.

- the backward pass equals float64 central finite differences of the loss;
- the convolution equals a direct (loop) evaluation of the 3x3 stride-2 formula;
- the parameter layout matches the library's (660,310 parameters).
"""
import numpy as np
import pytest

import merge_fusion_amd as mf
from oracle import twistnet as tn


def _params(seed):
    lay, total = tn.layout()
    rng = np.random.default_rng(seed)
    p = np.zeros(total)
    for w, s, b, n in lay:
        fan_in = int(np.prod(s[1:]))
        p[w:w + int(np.prod(s))] = rng.standard_normal(int(np.prod(s))) * np.sqrt(2.0 / fan_in)
        p[b:b + n] = rng.standard_normal(n) * 0.05
    return p


def _inputs(seed, B):
    rng = np.random.default_rng(seed)
    x = rng.random((B, 2, tn.IN_H, tn.IN_W)) * 1.2
    x[rng.random(x.shape) < 0.05] = 0.0
    t = rng.standard_normal((B, 6))
    return x, t


def test_layout_matches_library():
    lay, total = tn.layout()
    assert total == 660310
    assert mf.load_hip().kifi_twistnet_param_count() == total


def test_conv_matches_direct_formula():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 3, 9, 10))
    W = rng.standard_normal((4, 3, 3, 3))
    c = tn._im2col(x)
    z = c @ W.reshape(4, -1).T                       # (1, Ho, Wo, 4)
    Ho, Wo = tn.conv_out(9), tn.conv_out(10)
    ref = np.zeros((Ho, Wo, 4))
    for co in range(4):
        for ho in range(Ho):
            for wo in range(Wo):
                s = 0.0
                for ci in range(3):
                    for ky in range(3):
                        for kx in range(3):
                            hi, wi = 2 * ho - 1 + ky, 2 * wo - 1 + kx
                            if 0 <= hi < 9 and 0 <= wi < 10:
                                s += x[0, ci, hi, wi] * W[co, ci, ky, kx]
                ref[ho, wo, co] = s
    assert np.allclose(z[0], ref, atol=1e-12)
    # col2im is the adjoint of im2col: <im2col(x), d> == <x, col2im(d)>
    d = rng.standard_normal(c.shape)
    assert np.isclose((c * d).sum(), (x * tn._col2im(d, x.shape)).sum(), rtol=1e-12)


@pytest.mark.parametrize("layer", range(7))
def test_backward_equals_central_differences(layer):
    p = _params(3)
    x, t = _inputs(4, 2)
    loss, g = tn.loss_grad(p, x, t)
    lay, _ = tn.layout()
    w, s, b, n = lay[layer]
    rng = np.random.default_rng(layer)
    idx = list(rng.choice(int(np.prod(s)), 4, replace=False) + w) + [b + int(rng.integers(n))]
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = 0.5 * np.sum((tn.forward(pp, x) - t) ** 2) / 2
        lm = 0.5 * np.sum((tn.forward(pm, x) - t) ** 2) / 2
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (layer, i, fd, g[i])
    assert np.isclose(loss, 0.5 * np.sum((tn.forward(p, x) - t) ** 2) / 2)


def test_bf16_round_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0], np.float32)
    r = tn.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -7) and r[3] == -2.5 and r[4] == 0.0


# ---- Adam (Kingma & Ba 2015, Algorithm 1) ----------------------------------
def test_adam_first_step_is_lr_sign():
    """m = v = 0: the bias corrections make m^ = g and v^ = g^2, so the first
    step is -lr g / (|g| + eps): -lr sign(g) for |g| >> eps."""
    g = np.array([3.0, -2e-3, 1e-6, -7.5, 0.0])
    p0 = np.array([0.5, -1.0, 2.0, 0.0, 1.0])
    p1, m, v = tn.adam_step(p0, g, np.zeros(5), np.zeros(5), 1, 1e-3)
    assert np.allclose(p1 - p0, -1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_adam_constant_gradient_keeps_the_step():
    """A constant gradient gives m^ = g and v^ = g^2 at every step (the bias
    corrections exactly undo the moments' warm-up): every step equals the first."""
    g = np.array([0.3, -4.0, 2e-5])
    p, m, v = np.zeros(3), np.zeros(3), np.zeros(3)
    for t in range(1, 8):
        q, m, v = tn.adam_step(p, g, m, v, t, 2e-3)
        assert np.allclose(q - p, -2e-3 * g / (np.abs(g) + 1e-8), rtol=1e-10, atol=0)
        p = q


def test_adam_moments_follow_their_recurrences():
    """Alternating gradients: the first moment decays geometrically toward the
    alternating mean while the second tracks g^2 (closed forms of the two
    linear recurrences)."""
    p, m, v = np.zeros(1), np.zeros(1), np.zeros(1)
    for t in range(1, 11):
        g = np.array([1.0 if t % 2 else -1.0])
        p, m, v = tn.adam_step(p, g, m, v, t, 1e-3)
        # m_t = (1 - b1) sum_k b1^(t-k) g_k, v_t = (1 - b2) sum_k b2^(t-k)
        mt = 0.1 * sum(0.9 ** (t - k) * (1.0 if k % 2 else -1.0) for k in range(1, t + 1))
        vt = 0.001 * sum(0.999 ** (t - k) for k in range(1, t + 1))
        assert np.isclose(m[0], mt, rtol=1e-12) and np.isclose(v[0], vt, rtol=1e-12)
