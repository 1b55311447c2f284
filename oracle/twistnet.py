"""FP64 restatement of the twist-prediction CNN (test infrastructure only).

The parity checker for kifi_twistnet_* in libkifi_hip.so: the same network and
parameter layout, forward and backward in float64 numpy.  Only tests/ load it.
No reference counterpart (the reference has no learned part).

Network: input (B, 2, 106, 128); five 3x3 convolutions, stride 2, padding 1,
channels 2-16-32-64-64-128, ReLU; flatten (C, H, W) order to 2048; FC 256 +
ReLU; FC 6.  Loss 0.5/B * sum (y - t)^2.  Parameters flat, per layer: weights
(conv [Co][Ci][3][3], fc [out][in]) then bias.

The backward pass is pinned by central finite differences in float64
(tests/test_oracle_twistnet.py).
"""
import numpy as np

CONV_C = [2, 16, 32, 64, 64, 128]
IN_H, IN_W = 106, 128
FC1, OUT = 256, 6


def conv_out(n):
    return (n - 1) // 2 + 1


def layout():
    """[(w_offset, w_shape, b_offset, n_out)] for the 7 layers, and the total count."""
    out, o = [], 0
    for l in range(5):
        shape = (CONV_C[l + 1], CONV_C[l], 3, 3)
        w = o
        o += int(np.prod(shape))
        out.append((w, shape, o, CONV_C[l + 1]))
        o += CONV_C[l + 1]
    fin = CONV_C[5] * 4 * 4
    for shape in ((FC1, fin), (OUT, FC1)):
        w = o
        o += int(np.prod(shape))
        out.append((w, shape, o, shape[0]))
        o += shape[0]
    return out, o


def unpack(params):
    lay, total = layout()
    p = np.asarray(params, np.float64)
    assert p.size == total
    return [(p[w:w + int(np.prod(s))].reshape(s), p[b:b + n]) for w, s, b, n in lay]


def _im2col(x):
    """(B, C, H, W) -> cols (B, Ho, Wo, C*9) for k3 s2 p1; column = c*9 + ky*3 + kx."""
    B, C, H, W = x.shape
    Ho, Wo = conv_out(H), conv_out(W)
    xp = np.zeros((B, C, H + 2, W + 2), x.dtype)
    xp[:, :, 1:H + 1, 1:W + 1] = x
    cols = np.empty((B, Ho, Wo, C, 3, 3), x.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, :, ky, kx] = xp[:, :, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2].transpose(0, 2, 3, 1)
    return cols.reshape(B, Ho, Wo, C * 9)


def _col2im(dcols, shape):
    B, C, H, W = shape
    Ho, Wo = conv_out(H), conv_out(W)
    d = dcols.reshape(B, Ho, Wo, C, 3, 3)
    xp = np.zeros((B, C, H + 2, W + 2), dcols.dtype)
    for ky in range(3):
        for kx in range(3):
            xp[:, :, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2] += d[:, :, :, :, ky, kx].transpose(0, 3, 1, 2)
    return xp[:, :, 1:H + 1, 1:W + 1]


def _keep(a):
    return a


def _bf16_64(a):
    return bf16_round(np.asarray(a, np.float32)).astype(np.float64)


def forward(params, x, keep=False, store=None):
    """store: rounding applied to every stored activation (None: exact float64;
    bf16_store: the bfloat16 storage of the device kernels)."""
    r = store or _keep
    layers = unpack(params)
    a = np.asarray(x, np.float64).reshape(-1, 2, IN_H, IN_W)
    acts, cols = [a], []
    for l in range(5):
        W, b = layers[l]
        c = _im2col(a)
        z = c @ W.reshape(W.shape[0], -1).T + b            # (B, Ho, Wo, Co)
        a = r(np.maximum(z, 0.0).transpose(0, 3, 1, 2))
        acts.append(a)
        cols.append(c)
    f = a.reshape(a.shape[0], -1)
    W5, b5 = layers[5]
    h = r(np.maximum(f @ W5.T + b5, 0.0))
    W6, b6 = layers[6]
    y = h @ W6.T + b6
    if keep:
        return y, (acts, cols, f, h)
    return y


def loss_grad(params, x, t, store=None):
    """Loss and flat gradient.  store as in forward(); it also rounds the stored
    output and activation gradients (dy, and each layer's input gradient)."""
    r = store or _keep
    layers = unpack(params)
    lay, total = layout()
    y, (acts, cols, f, h) = forward(params, x, keep=True, store=store)
    t = np.asarray(t, np.float64).reshape(y.shape)
    B = y.shape[0]
    loss = 0.5 * np.sum((y - t) ** 2) / B
    g = np.zeros(total)

    def put(l, dW, db):
        w, s, b, n = lay[l]
        g[w:w + dW.size] = dW.reshape(-1)
        g[b:b + n] = db

    dy = r((y - t) / B)
    W6, _ = layers[6]
    put(6, dy.T @ h, dy.sum(0))
    dh = r(dy @ W6) * (h > 0)
    W5, _ = layers[5]
    put(5, dh.T @ f, dh.sum(0))
    da = (r(dh @ W5) * (f > 0)).reshape(acts[5].shape)
    for l in range(4, -1, -1):
        W, _ = layers[l]
        dz = da.transpose(0, 2, 3, 1)                       # (B, Ho, Wo, Co); ReLU' already applied
        c = cols[l]
        put(l, dz.reshape(-1, dz.shape[-1]).T @ c.reshape(-1, c.shape[-1]), dz.sum((0, 1, 2)))
        if l == 0:
            break
        dcols = dz @ W.reshape(W.shape[0], -1)
        da = r(_col2im(dcols, acts[l].shape)) * (acts[l] > 0)
    return loss, g


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_store(a):
    """Float64 values rounded through binary32 to bfloat16, as the kernels store them."""
    return _bf16_64(a)


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) in float64: returns the
    new parameters and moments for step t >= 1.
    m_t = b1 m + (1 - b1) g; v_t = b2 v + (1 - b2) g^2;
    p_t = p - lr (m_t / (1 - b1^t)) / (sqrt(v_t / (1 - b2^t)) + eps)."""
    p, g, m, v = (np.asarray(a, np.float64) for a in (p, g, m, v))
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v

