// chs_learn.h -- device kernels of the learned preconditioner: a 3-D
// convolutional network on the Ls-stacked spinor field, computed with BF16
// MFMA (v_mfma_f32_16x16x16_bf16, FP32 accumulation) for inference and
// training.
//
// Network: 4 layers of 3x3x3 convolutions over (x, y, t) with periodic
// padding, channels Cin -> H -> H -> H -> Cin, Cin = 16*Ls (s, spin-flavour
// plane, re/im), H = 16 hidden channels, tanh between layers, no bias, plus
// an identity path: N(r) = u + conv4(...conv1(u)), u = r / rms(r).  The H
// channel bottleneck alone would confine the output to a subspace of
// dimension H*V out of 16*Ls*V, in which a Krylov solver stalls; with the
// identity path the network learns a correction to plain CG's direction.
// The map is odd and scale-invariant: N(a r) = sign(a) N(r).
//
// Activations are BF16, channel-last: act[site * C + c], site in the native
// order (x fastest).  Channel c of a spinor is (s*8 + plane)*2 + (re, im), so
// the first and last layers read and write the FP32 native-layout spinor
// directly (plane p = d + 4 f at p*n5 + s*kvol + site).
//
// One conv kernel serves the forward pass and the data gradient:
//   out[x][n] = sum_t sum_k B_t(k, n) in[x + delta_t][k],
//   forward: B_t(k, n) = W[t][n][k]; data gradient: B_t(k, n) = W[26-t][k][n]
// (weights stored W[tap][cout][cin]).  A wavefront computes a 16-site x
// 16-channel output tile; lane l holds rows l%16 of A and columns l%16 of B
// with the four k-values 4*(l/16)..+3 (the 16x16x16 MFMA lane layout).
#pragma once
#include "chs_device.h"

namespace chs {
namespace nn {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf(float x) { return (unsigned short)f2bf(x); }
__device__ __forceinline__ float unbf(unsigned short b) { return __uint_as_float(((uint32_t)b) << 16); }

struct Geo {
    int L, Lt, kvol, Ls;
    long n5;   // Ls * kvol (spinor plane length)
};

// site of x + delta_t, t = (dx+1) + 3 (dy+1) + 9 (dt+1), periodic
__device__ __forceinline__ int shift_site(const Geo &g, int site, int t) {
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dt = t / 9 - 1;
    const int L = g.L, Lt = g.Lt;
    int x = site % L, r = site / L;
    int y = r % L, z = r / L;
    x += dx;
    x = x < 0 ? x + L : (x >= L ? x - L : x);
    y += dy;
    y = y < 0 ? y + L : (y >= L ? y - L : y);
    z += dt;
    z = z < 0 ? z + Lt : (z >= Lt ? z - Lt : z);
    return x + L * (y + L * z);
}

// spinor channel c -> element offset of its FP32 component
__device__ __forceinline__ long spinor_elem(const Geo &g, int c, int site) {
    const int ri = c & 1, sp = c >> 1;
    const int p = sp & 7, s = sp >> 3;
    return 2 * ((long)p * g.n5 + (long)s * g.kvol + site) + ri;
}

enum InMode { IN_ACT = 0, IN_SPINOR = 1 };
enum OutMode { OUT_ACT = 0, OUT_SPINOR = 1 };

struct ConvArgs {
    Geo g;
    int K, N;                    // input / output channels of this pass
    const void *in;              // IN_ACT: bf16 [site][K]; IN_SPINOR: float native spinor
    const float *in_scale;       // IN_SPINOR: device scalar multiplying the input
    const unsigned short *W;     // bf16 weights W[tap][cout][cin] of the forward layer
    int wflip;                   // 0 forward, 1 data gradient (transposed, flipped taps)
    void *out;                   // OUT_ACT: bf16 [site][N]; OUT_SPINOR: float native spinor
    const float *out_scale;      // OUT_SPINOR: device scalar multiplying the output
    const float *res;            // OUT_SPINOR: if not NULL, res * *res_scale is added (identity path)
    const float *res_scale;
    int act;                     // OUT_ACT: 1 = tanh
    const int *skip;             // if not NULL and *skip != 0 the launch does nothing
};

template <int INM, int OUTM>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
    if (a.skip && *a.skip) return;
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ntile_n = a.N / 16;
    const long ntile_s = (a.g.kvol + 15) / 16;
    if (wave >= ntile_s * ntile_n) return;
    const int tn = (int)(wave % ntile_n);
    const long ts = wave / ntile_n;
    const int row = lane & 15, quad = lane >> 4;
    const int site0 = (int)(ts * 16);
    const int my_site = site0 + row < a.g.kvol ? site0 + row : 0;
    const int n = tn * 16 + row;   // B column held by this lane
    const float isc = (INM == IN_SPINOR) ? *a.in_scale : 1.0f;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 27; t++) {
        const int nb = shift_site(a.g, my_site, t);
        const int tw = a.wflip ? 26 - t : t;
        for (int k0 = 0; k0 < a.K; k0 += 16) {
            const int k = k0 + 4 * quad;
            s16x4 av, bv;
            if (INM == IN_ACT) {
                const unsigned short *ip = (const unsigned short *)a.in + (long)nb * a.K + k;
                const uint2 raw = *reinterpret_cast<const uint2 *>(ip);
                av[0] = (short)(raw.x & 0xffff);
                av[1] = (short)(raw.x >> 16);
                av[2] = (short)(raw.y & 0xffff);
                av[3] = (short)(raw.y >> 16);
            } else {
                // channels k..k+3 = planes p, p+1 (re, im) at one s
                const float *sp = (const float *)a.in;
                const float2 v0 = *reinterpret_cast<const float2 *>(sp + spinor_elem(a.g, k, nb));
                const float2 v1 = *reinterpret_cast<const float2 *>(sp + spinor_elem(a.g, k + 2, nb));
                av[0] = (short)bf(v0.x * isc);
                av[1] = (short)bf(v0.y * isc);
                av[2] = (short)bf(v1.x * isc);
                av[3] = (short)bf(v1.y * isc);
            }
            if (!a.wflip) {
                // B(k, n) = W[t][n][k]: four consecutive cin
                const unsigned short *wp = a.W + ((long)tw * a.N + n) * a.K + k;
                const uint2 raw = *reinterpret_cast<const uint2 *>(wp);
                bv[0] = (short)(raw.x & 0xffff);
                bv[1] = (short)(raw.x >> 16);
                bv[2] = (short)(raw.y & 0xffff);
                bv[3] = (short)(raw.y >> 16);
            } else {
                // B(k, n) = W[26-t][k][n]: forward cout = k (K of this pass), cin = n
#pragma unroll
                for (int i = 0; i < 4; i++) bv[i] = (short)a.W[((long)tw * a.K + k + i) * a.N + n];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
    }
    // lane holds rows 4*quad + i, column row (= lane % 16)
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int site = site0 + 4 * quad + i;
        if (site >= a.g.kvol) continue;
        const int c = tn * 16 + row;
        float v = acc[i];
        if (OUTM == OUT_ACT) {
            if (a.act == 1) v = tanhf(v);
            ((unsigned short *)a.out)[(long)site * a.N + c] = bf(v);
        } else {
            const long e = spinor_elem(a.g, c, site);
            if (a.res) v += a.res[e] * *a.res_scale;
            ((float *)a.out)[e] = v * *a.out_scale;
        }
    }
}

// Weight gradient of one layer, split over site chunks:
//   part[chunk][t][cout][cin] = sum_{x in chunk} X[x + delta_t][cin] dZ[x][cout]
// X: the layer input (bf16 act [site][Cin], or the scaled FP32 spinor);
// dZ: gradient at the layer's pre-activation (bf16 act [site][Cout], or an
// FP32 native spinor for the output layer).  One wavefront per
// (t, cin-tile, cout-tile, chunk): M = cin, N = cout, reduction over sites.
struct WgradArgs {
    Geo g;
    int Cin, Cout, chunk;   // chunk: sites per chunk (multiple of 16)
    const void *X;
    const float *x_scale;
    const void *dZ;
    float *part;
};

template <int XM, int DZM>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int tci = a.Cin / 16, tco = a.Cout / 16;
    const int nchunk = (a.g.kvol + a.chunk - 1) / a.chunk;
    const long total = 27L * tci * tco * nchunk;
    if (wave >= total) return;
    long w = wave;
    const int ch = (int)(w % nchunk);
    w /= nchunk;
    const int co_t = (int)(w % tco);
    w /= tco;
    const int ci_t = (int)(w % tci);
    const int t = (int)(w / tci);
    const int row = lane & 15, quad = lane >> 4;
    const int ci = ci_t * 16 + row;   // A row
    const int co = co_t * 16 + row;   // B column
    const float xs = (XM == IN_SPINOR) ? *a.x_scale : 1.0f;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const int s_begin = ch * a.chunk;
    const int s_end = min(a.g.kvol, s_begin + a.chunk);
    for (int s0 = s_begin; s0 < s_end; s0 += 16) {
        s16x4 av, bv;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int site = s0 + 4 * quad + i;
            float xv = 0.f, dv = 0.f;
            if (site < s_end) {
                const int nb = shift_site(a.g, site, t);
                if (XM == IN_ACT) xv = unbf(((const unsigned short *)a.X)[(long)nb * a.Cin + ci]);
                else xv = ((const float *)a.X)[spinor_elem(a.g, ci, nb)] * xs;
                if (DZM == OUT_ACT) dv = unbf(((const unsigned short *)a.dZ)[(long)site * a.Cout + co]);
                else dv = ((const float *)a.dZ)[spinor_elem(a.g, co, site)];
            }
            av[i] = (short)bf(xv);
            bv[i] = (short)bf(dv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    // D rows = cin (4*quad + i), column = cout (lane % 16)
    float *dst = a.part + (long)ch * 27 * a.Cout * a.Cin;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int cin = ci_t * 16 + 4 * quad + i;
        dst[((long)t * a.Cout + co) * a.Cin + cin] = acc[i];
    }
}

// grad[i] = sum over chunks of part[chunk][i], in chunk order
__global__ __launch_bounds__(256) void k_wgrad_sum(long nw, int nchunk, const float *__restrict__ part,
                                                   float *__restrict__ grad) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= nw) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; c++) s += part[(long)c * nw + i];
    grad[i] = s;
}

// dZ = dY * (1 - Y^2) for the tanh layers, bf16 in and out
__global__ __launch_bounds__(256) void k_tanh_back(long n, const unsigned short *__restrict__ Y,
                                                   const unsigned short *__restrict__ dY,
                                                   unsigned short *__restrict__ dZ) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float y = unbf(Y[i]);
    dZ[i] = bf(unbf(dY[i]) * (1.f - y * y));
}

// Adam step on FP32 master weights, refreshed bf16 copy
__global__ __launch_bounds__(256) void k_adam(long n, float *__restrict__ w, float *__restrict__ m,
                                              float *__restrict__ v, const float *__restrict__ g, float lr, float b1,
                                              float b2, float c1, float c2, float eps,
                                              unsigned short *__restrict__ wbf) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    wbf[i] = bf(wi);
}

__global__ __launch_bounds__(256) void k_to_bf16(long n, const float *__restrict__ w, unsigned short *__restrict__ o) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) o[i] = bf(w[i]);
}

}  // namespace nn
}  // namespace chs
