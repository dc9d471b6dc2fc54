"""CPU oracle of the learned preconditioner's network -- TEST INFRASTRUCTURE
ONLY (tests/ use it to check the GPU forward pass and gradients).

The network (chiralheisenberg_amd/csrc/chs_learn.h) is linear and symmetric
positive definite:
    P(r) = c u + C1^T C2^T C2 C1 u,   u = r / rms(r),   c = exp(theta),
C1 a 1x1x1 convolution from the 16 Ls spinor channels to 16, C2 a 3x3x3
convolution over (x, y, t) with periodic padding, 16 -> 16; channel
(s*8 + plane)*2 + (re, im).  Parameters: W1[n][k] (16 x 16 Ls),
W2[tap][n][k] (27 x 16 x 16, tap = (dx+1) + 3 (dy+1) + 9 (dt+1)), theta.

The GPU multiplies BF16 operands and accumulates in FP32; this restatement
rounds the same operands to BF16 (round to nearest even) at the same points
and accumulates in FP64, so the two agree to FP32 accumulation-order effects.
Activations are arrays (Lt, L, L, C) with axes t, y, x, channel.

The loss of a pair (src, sol) is the relative A-norm of the error,
<w, A w> / <u, x>, u = src / rms(src), x = sol / rms(src), w = P(u) - x,
with A = M^dagger M supplied by the caller (a function on reference-layout
complex spinors).
"""
import numpy as np


ROUND = True   # tests of the oracle's own calculus switch the rounding off
H = 16


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    if not ROUND:
        return np.asarray(a, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def f32(a):
    return np.asarray(a, dtype=np.float32) if ROUND else np.asarray(a, dtype=np.float64)


def taps():
    out = []
    for t in range(27):
        out.append((t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1))  # dx, dy, dt
    return out


def shift(X, d):
    """X[x + delta] for X of shape (Lt, L, L, C) (axes t, y, x)."""
    dx, dy, dt = d
    return np.roll(X, shift=(-dt, -dy, -dx), axis=(0, 1, 2))


def conv3(X, W):
    """Y[x][n] = sum_t sum_k W[t][n][k] X[x + delta_t][k]; X (Lt,L,L,K), W (27, N, K)."""
    Y = 0.0
    for t, d in enumerate(taps()):
        Y = Y + np.einsum("tyxk,nk->tyxn", shift(X, d).astype(np.float64), W[t].astype(np.float64))
    return Y


def conv3_T(Y, W):
    """The adjoint of conv3: X[x][k] = sum_t sum_n W[t][n][k] Y[x - delta_t][n]."""
    X = 0.0
    for t, d in enumerate(taps()):
        nd = (-d[0], -d[1], -d[2])
        X = X + np.einsum("tyxn,nk->tyxk", shift(Y, nd).astype(np.float64), W[t].astype(np.float64))
    return X


def conv3_weight_grad(A, B):
    """dW[t][n][k] = sum_x A[x][n] B[x + delta_t][k]: the gradient of
    sum(conv3(B, W) * A) with respect to W."""
    return np.stack([np.einsum("tyxn,tyxk->nk", A.astype(np.float64), shift(B, d).astype(np.float64))
                     for d in taps()])


def pw(X, W1):
    """Y[x][n] = sum_k W1[n][k] X[x][k] (the 1x1x1 layer)."""
    return np.einsum("tyxk,nk->tyxn", X.astype(np.float64), W1.astype(np.float64))


def pw_T(Y, W1):
    """The adjoint of pw: X[x][k] = sum_n W1[n][k] Y[x][n]."""
    return np.einsum("tyxn,nk->tyxk", Y.astype(np.float64), W1.astype(np.float64))


def pw_weight_grad(A, B):
    """dW1[n][k] = sum_x A[x][n] B[x][k]."""
    return np.einsum("tyxn,tyxk->nk", A.astype(np.float64), B.astype(np.float64))


def spinor_to_channels(sp_native):
    """complex native spinor (2, 4, Ls, Lt, L, L) [f, d, s, t, y, x] -> (Lt, L, L, 16 Ls)."""
    f, d, Ls, Lt, L, _ = sp_native.shape
    p = sp_native.reshape(8, Ls, Lt, L, L)                      # plane = d + 4 f
    ch = np.stack([p.real, p.imag], axis=-1)                    # (8, Ls, Lt, L, L, 2)
    return ch.transpose(2, 3, 4, 1, 0, 5).reshape(Lt, L, L, Ls * 16)   # (t,y,x, s, plane, ri)


def channels_to_spinor(ch, Ls):
    Lt, L = ch.shape[0], ch.shape[1]
    c = ch.reshape(Lt, L, L, Ls, 8, 2).transpose(4, 3, 0, 1, 2, 5)   # (plane, s, t, y, x, ri)
    return (c[..., 0] + 1j * c[..., 1]).reshape(2, 4, Ls, Lt, L, L)


def native_from_reference(sp_ref):
    """reference layout (2, 4, Lt, L, L, Ls) -> native (2, 4, Ls, Lt, L, L)."""
    return np.ascontiguousarray(np.moveaxis(sp_ref, 5, 2))


def reference_from_native(sp_nat):
    return np.ascontiguousarray(np.moveaxis(sp_nat, 2, 5))


def unpack_weights(w, Ls):
    Cs = 16 * Ls
    dt = np.float32 if ROUND else np.float64
    w = np.asarray(w)
    n1, n2 = H * Cs, 27 * H * H
    W1 = np.asarray(w[:n1], dtype=dt).reshape(H, Cs)
    W2 = np.asarray(w[n1:n1 + n2], dtype=dt).reshape(27, H, H)
    return W1, W2, float(w[n1 + n2])


def num_weights(Ls):
    return H * 16 * Ls + 27 * H * H + 1


def _scale(nat):
    """1 / rms over the real components, as the GPU forms it (FP64 sum of the
    FP32 values, rounded to FP32)."""
    v = nat.astype(np.complex128)
    rms = np.sqrt(np.sum(np.abs(v) ** 2) / (2.0 * v.size))
    return np.float32(1.0 / rms) if ROUND else 1.0 / rms


def forward(src_ref, w, Ls):
    """P(src / rms(src)) on a reference-layout spinor (FP32-valued); returns
    the output channels (Lt, L, L, 16 Ls) float64 and a cache for backward."""
    W1, W2, theta = unpack_weights(w, Ls)
    W1b, W2b = bf16(W1), bf16(W2)
    nat = native_from_reference(src_ref.astype(np.complex64 if ROUND else np.complex128))
    s = _scale(nat)
    U = f32(spinor_to_channels(nat))
    X0 = bf16(f32(U * s))
    h1 = bf16(f32(pw(X0, W1b)))
    h2 = bf16(f32(conv3(h1, W2b)))
    h3 = bf16(f32(conv3_T(h2, W2b)))
    coef = np.float32(np.exp(np.float32(theta)) * s) if ROUND else np.exp(theta) * s
    Z = pw_T(h3, W1b) + np.float64(coef) * U.astype(np.float64)
    return Z, dict(W1b=W1b, W2b=W2b, theta=theta, s=s, U=U, X0=X0, h1=h1, h2=h2, h3=h3)


def loss_and_grad(src_ref, sol_ref, w, Ls, A):
    """Relative A-norm loss of one pair and the gradient of all parameters
    (concatenated like the library's parameter vector).  A: M^dagger M on
    reference-layout complex spinors."""
    Z, c = forward(src_ref, w, Ls)
    s = c["s"]
    natx = native_from_reference(sol_ref.astype(np.complex64 if ROUND else np.complex128))
    Xs = f32(spinor_to_channels(natx))
    Wc = f32(Z) - f32(Xs * s)                                   # w = P(u) - s sol
    Lt, L = Wc.shape[0], Wc.shape[1]
    Aw_ref = A(reference_from_native(channels_to_spinor(Wc.astype(np.float64), Ls)))
    Aw = spinor_to_channels(native_from_reference(Aw_ref))
    den = float(s) ** 2 * np.sum(c["U"].astype(np.float64) * Xs.astype(np.float64))
    loss = np.sum(Wc.astype(np.float64) * Aw) / den
    G = f32(2.0 / den * Aw)
    Gb = bf16(G)
    dh3 = bf16(f32(pw(Gb, c["W1b"])))
    dh2 = bf16(f32(conv3(dh3, c["W2b"])))
    dh1 = bf16(f32(conv3_T(dh2, c["W2b"])))
    dW1 = pw_weight_grad(dh1, c["X0"]) + pw_weight_grad(c["h3"], Gb)
    dW2 = conv3_weight_grad(dh2, c["h1"]) + conv3_weight_grad(c["h2"], dh3)
    dth = np.exp(c["theta"]) * float(s) * np.sum(G.astype(np.float64) * c["U"].astype(np.float64))
    return loss, np.concatenate([dW1.reshape(-1), dW2.reshape(-1), [dth]])
