"""CPU oracle of the learned preconditioner's network -- TEST INFRASTRUCTURE
ONLY (tests/ use it to check the GPU forward pass and gradients).

The network (chiralheisenberg_amd/csrc/chs_learn.h): 4 layers of 3x3x3
convolutions over (x, y, t) with periodic padding, channels
Cin -> H -> H -> H -> Cin with Cin = 16 Ls, tanh after layers 1-3, no bias,
plus the identity path (output = u + conv4(...), u = spinor / rms(spinor));
channel (s*8 + plane)*2 + (re, im).  Weights are
W[tap][cout][cin], tap = (dx+1) + 3 (dy+1) + 9 (dt+1).

The GPU multiplies BF16 operands and accumulates in FP32; this restatement
rounds the same operands to BF16 (round to nearest even) at the same points
and accumulates in FP64, so the two agree to FP32 accumulation-order effects.
The loss is the mean squared error against the target x / rms(x).
"""
import numpy as np


ROUND = True   # tests of the oracle's own calculus switch the rounding off


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    if not ROUND:
        return np.asarray(a, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def taps():
    out = []
    for t in range(27):
        out.append((t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1))  # dx, dy, dt
    return out


def shift(X, d):
    """X[x + delta] for X of shape (Lt, L, L, C) (axes t, y, x)."""
    dx, dy, dt = d
    return np.roll(X, shift=(-dt, -dy, -dx), axis=(0, 1, 2))


def conv(X, W):
    """Y[x] = sum_t W[t] X[x + delta_t]; X (Lt,L,L,K), W (27, N, K) -> (Lt,L,L,N)."""
    Y = 0.0
    for t, d in enumerate(taps()):
        Y = Y + np.einsum("tyxk,nk->tyxn", shift(X, d).astype(np.float64), W[t].astype(np.float64))
    return Y


def conv_data_grad(dY, W):
    """dX[x] = sum_t W[t]^T dY[x - delta_t]."""
    dX = 0.0
    for t, d in enumerate(taps()):
        nd = (-d[0], -d[1], -d[2])
        dX = dX + np.einsum("tyxn,nk->tyxk", shift(dY, nd).astype(np.float64), W[t].astype(np.float64))
    return dX


def conv_weight_grad(X, dZ):
    """dW[t][n][k] = sum_x X[x + delta_t][k] dZ[x][n]."""
    return np.stack([np.einsum("tyxk,tyxn->nk", shift(X, d).astype(np.float64), dZ.astype(np.float64))
                     for d in taps()])


def spinor_to_channels(sp_native):
    """complex native spinor (2, 4, Ls, Lt, L, L) [f, d, s, t, y, x] -> (Lt, L, L, 16 Ls) float32."""
    f, d, Ls, Lt, L, _ = sp_native.shape
    p = sp_native.reshape(8, Ls, Lt, L, L)                      # plane = d + 4 f
    ch = np.stack([p.real, p.imag], axis=-1)                    # (8, Ls, Lt, L, L, 2)
    ch = ch.transpose(2, 3, 4, 1, 0, 5).reshape(Lt, L, L, Ls * 16)  # (t,y,x, s, plane, ri)
    return ch.astype(np.float32)


def channels_to_spinor(ch, Ls):
    Lt, L = ch.shape[0], ch.shape[1]
    c = ch.reshape(Lt, L, L, Ls, 8, 2).transpose(4, 3, 0, 1, 2, 5)   # (plane, s, t, y, x, ri)
    return (c[..., 0] + 1j * c[..., 1]).reshape(2, 4, Ls, Lt, L, L)


def native_from_reference(sp_ref):
    """reference layout (2, 4, Lt, L, L, Ls) -> native (2, 4, Ls, Lt, L, L)."""
    return np.ascontiguousarray(np.moveaxis(sp_ref, 5, 2))


def unpack_weights(w, Ls, H=16):
    Cin = 16 * Ls
    shapes = [(27, H, Cin), (27, H, H), (27, H, H), (27, Cin, H)]
    out, o = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(np.asarray(w[o:o + n], dtype=np.float32 if ROUND else np.float64).reshape(s))
        o += n
    return out


def rms(a):
    a = np.asarray(a)
    return np.sqrt(np.mean(np.abs(a) ** 2) / 2.0)


def forward(src_ref, weights, Ls):
    """Network output on a reference-layout FP32-valued spinor; returns
    (output channels (Lt,L,L,Cin) float64, cache for backward)."""
    W = [bf16(w) for w in weights]
    nat = native_from_reference(src_ref.astype(np.complex64 if ROUND else np.complex128))
    isc = np.float32(1.0 / rms(nat.astype(np.complex128))) if ROUND else 1.0 / rms(nat)
    U = spinor_to_channels(nat).astype(np.float32 if ROUND else np.float64) * isc
    X0 = bf16(U)
    ft = np.float32 if ROUND else np.float64
    A1 = bf16(np.tanh(conv(X0, W[0]).astype(ft)))
    A2 = bf16(np.tanh(conv(A1, W[1]).astype(ft)))
    A3 = bf16(np.tanh(conv(A2, W[2]).astype(ft)))
    Y = conv(A3, W[3]) + U.astype(np.float64)
    return Y, (W, X0, A1, A2, A3)


def backward(cache, G4):
    """Gradients of all weights (concatenated like the library's weight
    vector) from the gradient G4 (Lt,L,L,Cin) of a loss with respect to the
    network output channels."""
    W, X0, A1, A2, A3 = cache
    dW4 = conv_weight_grad(A3, bf16(G4))
    dA3 = bf16(conv_data_grad(bf16(G4), W[3]))
    dZ3 = bf16(dA3 * (1 - A3 * A3))
    dW3 = conv_weight_grad(A2, dZ3)
    dA2 = bf16(conv_data_grad(dZ3, W[2]))
    dZ2 = bf16(dA2 * (1 - A2 * A2))
    dW2 = conv_weight_grad(A1, dZ2)
    dA1 = bf16(conv_data_grad(dZ2, W[1]))
    dZ1 = bf16(dA1 * (1 - A1 * A1))
    dW1 = conv_weight_grad(X0, dZ1)
    return np.concatenate([g.reshape(-1) for g in (dW1, dW2, dW3, dW4)])


def loss_and_grad(src_ref, sol_ref, weights, Ls):
    """MSE loss against sol / rms(sol) and the gradients of all weights
    (concatenated like the library's weight vector)."""
    Y, cache = forward(src_ref, weights, Ls)
    natx = native_from_reference(sol_ref.astype(np.complex64 if ROUND else np.complex128))
    tsc = np.float32(1.0 / rms(natx.astype(np.complex128))) if ROUND else 1.0 / rms(natx)
    T = spinor_to_channels(natx).astype(np.float64) * tsc
    nreal = T.size
    diff = (Y.astype(np.float32).astype(np.float64) if ROUND else Y) - T
    loss = np.sum(diff * diff) / nreal
    G4 = (2.0 / nreal * diff).astype(np.float32 if ROUND else np.float64)
    return loss, backward(cache, G4)
