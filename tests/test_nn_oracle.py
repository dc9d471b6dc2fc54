"""Pins the network oracle (CPU): its convolutions equal torch's conv3d with
circular padding (an independent implementation), its transposed layers are
the adjoints of the forward ones, its gradients equal central finite
differences of its loss (BF16 rounding switched off), P is symmetric positive
definite, and its BF16 rounding is torch's."""
import numpy as np
import pytest

import nn_oracle as nno

torch = pytest.importorskip("torch")
F = torch.nn.functional


def torch_conv3(X, W):
    """conv3 through torch: X (Lt,L,L,K) -> (Lt,L,L,N), W[tap][n][k] with
    tap = (dx+1) + 3 (dy+1) + 9 (dt+1) -> weight[n, k, dt+1, dy+1, dx+1]."""
    x = torch.from_numpy(np.ascontiguousarray(X.transpose(3, 0, 1, 2)))[None].double()
    wt = torch.from_numpy(np.ascontiguousarray(W.reshape(3, 3, 3, W.shape[1], W.shape[2]).transpose(3, 4, 0, 1, 2)))
    y = F.conv3d(F.pad(x, (1, 1, 1, 1, 1, 1), mode="circular"), wt.double())
    return y[0].numpy().transpose(1, 2, 3, 0)


def test_conv3_equals_torch_conv3d_circular():
    rs = np.random.RandomState(0)
    for (Lt, L, K, N) in [(4, 3, 8, 16), (5, 6, 16, 16), (2, 2, 16, 8)]:
        X = rs.randn(Lt, L, L, K)
        W = rs.randn(27, N, K)
        got = nno.conv3(X, W)
        want = torch_conv3(X, W)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), (Lt, L)


def test_pw_equals_torch_1x1x1():
    rs = np.random.RandomState(1)
    X = rs.randn(3, 4, 4, 32)
    W1 = rs.randn(16, 32)
    x = torch.from_numpy(np.ascontiguousarray(X.transpose(3, 0, 1, 2)))[None]
    want = F.conv3d(x, torch.from_numpy(W1)[:, :, None, None, None])[0].numpy().transpose(1, 2, 3, 0)
    assert np.abs(nno.pw(X, W1) - want).max() <= 1e-12 * np.abs(want).max()


def test_transposed_layers_are_adjoints():
    rs = np.random.RandomState(2)
    X = rs.randn(4, 3, 3, 16)
    Y = rs.randn(4, 3, 3, 16)
    W = rs.randn(27, 16, 16)
    assert abs(np.sum(nno.conv3(X, W) * Y) - np.sum(X * nno.conv3_T(Y, W))) <= 1e-10 * np.sum(np.abs(X))
    Xs = rs.randn(4, 3, 3, 32)
    W1 = rs.randn(16, 32)
    assert abs(np.sum(nno.pw(Xs, W1) * Y) - np.sum(Xs * nno.pw_T(Y, W1))) <= 1e-10 * np.sum(np.abs(Xs))
    # weight gradients: d/dW sum(conv(B, W) * A)
    E = rs.randn(*W.shape)
    G = nno.conv3_weight_grad(Y, X)
    assert abs(np.sum(G * E) - np.sum(nno.conv3(X, E) * Y)) <= 1e-10 * abs(np.sum(G * E))
    E1 = rs.randn(*W1.shape)
    G1 = nno.pw_weight_grad(Y, Xs)
    assert abs(np.sum(G1 * E1) - np.sum(nno.pw(Xs, E1) * Y)) <= 1e-10 * abs(np.sum(G1 * E1))


def _spd_operator(rs, n):
    Q = rs.randn(n, n)
    return Q @ Q.T / n + 0.5 * np.eye(n)


def test_P_is_symmetric_positive_definite():
    nno.ROUND = False
    try:
        rs = np.random.RandomState(3)
        Ls, L, Lt = 2, 3, 4
        w = rs.randn(nno.num_weights(Ls)) * 0.3
        shape = (2, 4, Lt, L, L, Ls)
        a = rs.randn(*shape) + 1j * rs.randn(*shape)
        b = rs.randn(*shape) + 1j * rs.randn(*shape)
        # P without the 1/rms input scaling: P(v) = forward(v) * rms(v)
        def P(v):
            Z, c = nno.forward(v, w, Ls)
            return Z / c["s"]
        Ua = nno.spinor_to_channels(nno.native_from_reference(a))
        Ub = nno.spinor_to_channels(nno.native_from_reference(b))
        lhs, rhs = np.sum(Ua * P(b)), np.sum(P(a) * Ub)
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
        assert np.sum(Ua * P(a)) > np.exp(w[-1]) * np.sum(Ua * Ua) * (1 - 1e-12)
    finally:
        nno.ROUND = True


def test_gradient_is_finite_difference_of_loss():
    nno.ROUND = False
    try:
        rs = np.random.RandomState(4)
        Ls, L, Lt = 1, 3, 3
        shape = (2, 4, Lt, L, L, Ls)
        n = 2 * int(np.prod(shape))
        Amat = _spd_operator(rs, n)

        def A(v):
            r = Amat @ np.concatenate([v.real.reshape(-1), v.imag.reshape(-1)])
            return (r[:n // 2] + 1j * r[n // 2:]).reshape(v.shape)

        src = rs.randn(*shape) + 1j * rs.randn(*shape)
        sol = rs.randn(*shape) + 1j * rs.randn(*shape)
        w = rs.randn(nno.num_weights(Ls)) * 0.2
        loss, g = nno.loss_and_grad(src, sol, w, Ls, A)
        eps = 1e-6
        picks = list(rs.choice(w.size - 1, 12, replace=False)) + [w.size - 1]
        for i in picks:
            wp, wm = w.copy(), w.copy()
            wp[i] += eps
            wm[i] -= eps
            lp, _ = nno.loss_and_grad(src, sol, wp, Ls, A)
            lm, _ = nno.loss_and_grad(src, sol, wm, Ls, A)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (i, fd, g[i])
    finally:
        nno.ROUND = True


def test_bf16_rounding_matches_torch():
    rs = np.random.RandomState(2)
    a = np.concatenate([rs.randn(10000).astype(np.float32) * 10.0 ** rs.randint(-5, 5, 10000),
                        np.array([1.00390625, 1.01171875, -1.00390625, 0.0], np.float32)])
    want = torch.from_numpy(a).to(torch.bfloat16).float().numpy()
    assert np.array_equal(nno.bf16(a), want)
