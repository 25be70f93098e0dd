"""cnn_oracle.py -- numpy restatement of the learned preconditioner's network,
its gradients, its Adam step and the preconditioner / harvest maps.

TEST INFRASTRUCTURE ONLY: the checker the HIP kernels (dft_amd/csrc/cnn.hip,
harvest.hip) are compared against.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import it; the product path never does.

The network has no reference counterpart (the reference preconditions
Davidson with the diagonal of H only, EigenvalueSolverMolecule.cpp:483-555);
it is pinned in tests/test_cnn_oracle.py by an independent conv3d (torch, CPU,
float64), by central finite differences of the loss, by a hand-computed Adam
step and by bf16 round-to-nearest-even special cases.

Conventions (shared with the kernels):
  parameters, for layer l = 0..3 with (cin, cout) = (2,C), (C,C), (C,C), (C,1):
    W_l[cout][cin][27] (tap = kz*9 + ky*3 + kx, the output at v reads the
    input at v + (kx-1, ky-1, kz-1), zero outside), then b_l[cout];
  activations [item][channel][z][y][x]; SiLU after layers 0-2;
  loss = mean((out - target)^2).
"""
import numpy as np


def layout(C=16):
    """[(cin, cout, w_offset, b_offset)] per layer and the total count."""
    cins, couts = (2, C, C, C), (C, C, C, 1)
    off, out = 0, []
    for ci, co in zip(cins, couts):
        w = off
        off += co * ci * 27
        b = off
        off += co
        out.append((ci, co, w, b))
    return out, off


def bf16_round(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _taps():
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                yield kz * 9 + ky * 3 + kx, kx - 1, ky - 1, kz - 1


def _shift(x, dx, dy, dz):
    """y[..., z, y, x] = x[..., z+dz, y+dy, x+dx] with zeros outside."""
    out = np.zeros_like(x)
    nz, ny, nx = x.shape[-3:]
    zs, ys, xs = slice(max(0, -dz), nz - max(0, dz)), slice(max(0, -dy), ny - max(0, dy)), slice(max(0, -dx), nx - max(0, dx))
    zd, yd, xd = slice(max(0, dz), nz + min(0, dz)), slice(max(0, dy), ny + min(0, dy)), slice(max(0, dx), nx + min(0, dx))
    out[..., zs, ys, xs] = x[..., zd, yd, xd]
    return out


def conv(x, W, b=None):
    """x: (B, cin, nz, ny, nx); W: (cout, cin, 27) -> (B, cout, nz, ny, nx), float64."""
    x = np.asarray(x, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    B, ci = x.shape[:2]
    co = W.shape[0]
    y = np.zeros((B, co) + x.shape[2:])
    for t, dx, dy, dz in _taps():
        xs = _shift(x, dx, dy, dz)
        y += np.einsum("oi,bi...->bo...", W[:, :, t], xs)
    if b is not None:
        y += np.asarray(b, dtype=np.float64)[None, :, None, None, None]
    return y


def conv_transpose(d, W):
    """Adjoint of conv in x: d (B, cout, ...) -> (B, cin, ...)."""
    W = np.asarray(W, dtype=np.float64)
    out = np.zeros((d.shape[0], W.shape[1]) + d.shape[2:])
    for t, dx, dy, dz in _taps():
        out += _shift(np.einsum("oi,bo...->bi...", W[:, :, t], d), -dx, -dy, -dz)
    return out


def conv_wgrad(x, d):
    """dL/dW for y = conv(x, W): x (B, cin, ...), d (B, cout, ...) -> (cout, cin, 27)."""
    x = np.asarray(x, dtype=np.float64)
    B = x.shape[0]
    dd = d.reshape(B, d.shape[1], -1)
    gW = np.zeros((d.shape[1], x.shape[1], 27))
    for t, dx, dy, dz in _taps():
        gW[:, :, t] = np.einsum("bon,bin->oi", dd, _shift(x, dx, dy, dz).reshape(B, x.shape[1], -1))
    return gW


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(params, C=16):
    lay, _ = layout(C)
    p = np.asarray(params, dtype=np.float64)
    return [(p[w:w + co * ci * 27].reshape(co, ci, 27), p[b:b + co]) for ci, co, w, b in lay]


def forward(params, x, C=16):
    """params: effective (bf16-rounded) values; x (B, 2, nz, ny, nx).
    Returns out (B, nz, ny, nx) and the pre-activations [z0, z1, z2]."""
    Ws = unpack(params, C)
    zs = []
    a = np.asarray(x, dtype=np.float64)
    for l, (W, b) in enumerate(Ws):
        z = conv(a, W, b)
        if l < 3:
            zs.append(z)
            a = silu(z)
        else:
            return z[:, 0], zs


def loss_and_grad(params, x, target, C=16):
    """Mean-squared-error loss and its gradient (flat, kernel layout)."""
    Ws = unpack(params, C)
    out, zs = forward(params, x, C)
    M = out.size
    e = out - np.asarray(target, dtype=np.float64)
    loss = float(np.sum(e * e) / M)
    d = (2.0 / M) * e[:, None]
    ins = [np.asarray(x, dtype=np.float64)] + [silu(z) for z in zs]
    grads = [None] * 4
    for l in range(3, -1, -1):
        W, _ = Ws[l]
        grads[l] = (conv_wgrad(ins[l], d), d.sum(axis=(0, 2, 3, 4)))
        if l > 0:
            d = conv_transpose(d, W) * dsilu(zs[l - 1])
    flat = np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads])
    return loss, flat


def adam_step(w, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba, Algorithm 1), float64."""
    w, g, m, v = (np.asarray(a, dtype=np.float64) for a in (w, g, m, v))
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def he_params(C=16, seed=0):
    """He-uniform weights, zero biases (last layer scaled by 0.1), float32."""
    lay, n = layout(C)
    rng = np.random.default_rng(seed)
    p = np.zeros(n, dtype=np.float32)
    for l, (ci, co, w, b) in enumerate(lay):
        bound = np.sqrt(6.0 / (ci * 27)) * (0.1 if l == 3 else 1.0)
        p[w:w + co * ci * 27] = rng.uniform(-bound, bound, co * ci * 27)
    return p


def prepare(r, veff, theta, h):
    """Network input of one residual column: {r / max|r|, (V - theta) h2}."""
    h2 = 1.0 / sum(1.0 / hh ** 2 for hh in h)
    s = np.max(np.abs(r))
    return np.stack([r / s if s > 0 else np.zeros_like(r), (veff - theta) * h2]), s, h2
