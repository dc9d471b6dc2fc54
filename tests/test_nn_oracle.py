"""Pins the network oracle's own calculus (CPU): with the BF16 rounding
switched off, its gradients equal central finite differences of its loss,
and its data-gradient convolution is the adjoint of the forward one."""
import numpy as np

import nn_oracle as nno


def test_conv_adjoint():
    rs = np.random.RandomState(0)
    X = rs.randn(4, 3, 3, 8)
    dY = rs.randn(4, 3, 3, 16)
    W = rs.randn(27, 16, 8)
    lhs = np.sum(nno.conv(X, W) * dY)
    rhs = np.sum(X * nno.conv_data_grad(dY, W))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # weight gradient: d/dW sum(conv(X, W) * dY)
    G = nno.conv_weight_grad(X, dY)
    E = rs.randn(*W.shape)
    assert abs(np.sum(G * E) - np.sum(nno.conv(X, E) * dY)) <= 1e-10 * abs(np.sum(G * E))


def test_gradient_is_finite_difference_of_loss():
    nno.ROUND = False
    try:
        rs = np.random.RandomState(1)
        Ls, L = 1, 3
        shape = (2, 4, L, L, L, Ls)
        src = rs.randn(*shape) + 1j * rs.randn(*shape)
        sol = rs.randn(*shape) + 1j * rs.randn(*shape)
        w = rs.randn(27 * 16 * 16 * 2 + 27 * 16 * 16 * 2) * 0.2
        loss, g = nno.loss_and_grad(src, sol, nno.unpack_weights(w, Ls), Ls)
        eps = 1e-6
        for i in rs.choice(w.size, 12, replace=False):
            wp, wm = w.copy(), w.copy()
            wp[i] += eps
            wm[i] -= eps
            lp, _ = nno.loss_and_grad(src, sol, nno.unpack_weights(wp, Ls), Ls)
            lm, _ = nno.loss_and_grad(src, sol, nno.unpack_weights(wm, Ls), Ls)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (i, fd, g[i])
    finally:
        nno.ROUND = True


def test_bf16_rounding_matches_torch():
    import torch
    rs = np.random.RandomState(2)
    a = np.concatenate([rs.randn(10000).astype(np.float32) * 10.0 ** rs.randint(-5, 5, 10000),
                        np.array([1.00390625, 1.01171875, -1.00390625, 0.0], np.float32)])
    want = torch.from_numpy(a).to(torch.bfloat16).float().numpy()
    assert np.array_equal(nno.bf16(a), want)
