"""CPU pins of the learned preconditioner's oracle (oracle/cnn_oracle.py).

The network has no counterpart in the reference, so its oracle is pinned by
independent means: torch's conv3d and autograd on the CPU in float64, central
finite differences of the loss, the adjoint identity of the convolution, the
closed form of Adam's first step, and bfloat16 rounding special cases."""
import numpy as np
import pytest
import torch

from oracle import cnn_oracle as O


def torch_forward(p, x, C):
    lay, _ = O.layout(C)
    a = torch.as_tensor(x, dtype=torch.float64)
    for l, (ci, co, w, b) in enumerate(lay):
        W = p[w:w + co * ci * 27].reshape(co, ci, 3, 3, 3)
        a = torch.nn.functional.conv3d(a, W, p[b:b + co], padding=1)
        if l < 3:
            a = torch.nn.functional.silu(a)
    return a[:, 0]


@pytest.mark.parametrize("C,shape", [(4, (2, 2, 5, 6, 7)), (16, (1, 2, 4, 3, 9))])
def test_forward_matches_torch_conv3d(C, shape):
    rng = np.random.default_rng(C)
    p = O.he_params(C, seed=3)
    x = rng.standard_normal(shape)
    out, zs = O.forward(p, x, C)
    ref = torch_forward(torch.tensor(p.astype(np.float64)), x, C).numpy()
    assert out.shape == ref.shape
    assert np.max(np.abs(out - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    assert len(zs) == 3 and zs[0].shape == (shape[0], C) + shape[2:]


def test_gradient_matches_autograd():
    C = 4
    rng = np.random.default_rng(1)
    p = O.he_params(C, seed=5)
    x = rng.standard_normal((2, 2, 5, 4, 6))
    t = rng.standard_normal((2, 5, 4, 6))
    loss, g = O.loss_and_grad(p, x, t, C)
    pt = torch.tensor(p.astype(np.float64), requires_grad=True)
    L = ((torch_forward(pt, x, C) - torch.tensor(t)) ** 2).mean()
    L.backward()
    assert abs(L.item() - loss) <= 1e-14 * abs(loss)
    assert np.max(np.abs(pt.grad.numpy() - g)) <= 1e-12 * np.max(np.abs(g))


def test_gradient_finite_difference():
    C = 4
    rng = np.random.default_rng(2)
    p = O.he_params(C, seed=6).astype(np.float64)
    x = rng.standard_normal((1, 2, 4, 4, 5))
    t = rng.standard_normal((1, 4, 4, 5))
    _, g = O.loss_and_grad(p, x, t, C)
    lay, n = O.layout(C)
    # a weight and a bias of every layer
    idx = [w + 7 for _, _, w, _ in lay] + [b for _, _, _, b in lay]
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (O.loss_and_grad(pp, x, t, C)[0] - O.loss_and_grad(pm, x, t, C)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_conv_transpose_is_adjoint():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 3, 4, 5, 6))
    y = rng.standard_normal((2, 5, 4, 5, 6))
    W = rng.standard_normal((5, 3, 27))
    lhs = np.sum(O.conv(x, W) * y)
    rhs = np.sum(x * O.conv_transpose(y, W))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_adam_first_step_closed_form():
    # step 1 from m = v = 0: m_hat = g, v_hat = g^2 -> w - lr g / (|g| + eps)
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    w = np.array([1.0, 1.0, 1.0, 1.0])
    w1, m1, v1 = O.adam_step(w, g, np.zeros(4), np.zeros(4), 1, lr=0.01, eps=1e-8)
    assert np.allclose(w1, w - 0.01 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 0.001 * g * g)


def test_bf16_round_to_nearest_even():
    one = np.float32(1.0)
    cases = {
        1.0: 0x3F80,
        -2.0: 0xC000,
        float(one + np.float32(2.0 ** -8)): 0x3F80,       # halfway, ties to the even 1.0
        float(one + np.float32(3 * 2.0 ** -8)): 0x3F82,   # halfway, ties to the even 1 + 2^-6
        float(one + np.float32(2.0 ** -8 + 2.0 ** -20)): 0x3F81,  # above halfway rounds up
        0.0: 0x0000,
    }
    for v, bits in cases.items():
        got = int(O.bf16_bits(np.array([v], dtype=np.float32))[0])
        assert got == bits, (v, hex(got))
    rng = np.random.default_rng(4)
    x = (rng.standard_normal(4096) * 10.0 ** rng.uniform(-8, 8, 4096)).astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(O.bf16_round(x), ref)


def test_layout_counts():
    lay, n = O.layout(16)
    assert n == 16 * 2 * 27 + 16 + 2 * (16 * 16 * 27 + 16) + 16 * 27 + 1
    assert [l[:2] for l in lay] == [(2, 16), (16, 16), (16, 16), (16, 1)]
