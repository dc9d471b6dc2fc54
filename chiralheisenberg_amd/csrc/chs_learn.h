// chs_learn.h -- device kernels of the learned preconditioner: a linear,
// symmetric positive definite 3-D convolutional network on the Ls-stacked
// spinor field, on gfx950's v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32
// accumulation), for inference and training.
//
// Network (4 convolutions, H = 16 hidden channels):
//   P(r) = c u + C1^T C2^T C2 C1 u,   u = r / rms(r),   c = exp(theta) > 0,
// C1: 1x1x1 convolution, 16 Ls spinor channels -> H (a per-site channel map),
// C2: 3x3x3 convolution over (x, y, t), periodic, H -> H.  With K = C2 C1,
// P = c I + K^T K is symmetric positive definite for any weights, so it is a
// valid preconditioner of CG on M^dagger M (which a non-linear or
// non-symmetric network is not: it breaks CG's conjugacy).  The scale of P
// does not matter to preconditioned CG.
//
// Spinor channel c = (s*8 + plane)*2 + (re, im); an FP32 native-layout
// spinor holds channel c of a site at float index 2 (plane n5 + s kvol +
// site) + ri.  Activations are BF16, channel-last: act[site * 16 + ch].
//
// Every layer is one GEMM tile per wavefront, D = W X: rows of D are output
// channels, columns are 16 sites, the reduction runs over (tap, input
// channel) in steps of 32.  MFMA lane layout (16x16x32): lane l supplies
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j = 0..7, and
// receives D[4 (l / 16) + i][l % 16], i = 0..3.  With sites as the columns,
// every lane's D values are 4 consecutive channels of one site: one 8-byte
// activation store, or two float2 spinor stores that 16 lanes make
// contiguous per plane.
//
// The 3x3x3 layers (little arithmetic, 27-fold input reuse) stage their input
// in LDS: a workgroup owns a band of R rows of one t-plane and loads the
// rows y0-1 .. y0+R of the planes t-1, t, t+1 once; all 27 taps read LDS.
// The 1x1x1 layers read each input element exactly once (no reuse to stage).
#pragma once
#include "chs_device.h"

namespace chs {
namespace nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr int H = 16;      // hidden channels
constexpr int kRows = 4;   // rows per workgroup band of the 3x3x3 layers

__device__ __forceinline__ unsigned short bf(float x) { return (unsigned short)f2bf(x); }
__device__ __forceinline__ float unbf(unsigned short b) { return __uint_as_float(((uint32_t)b) << 16); }

__device__ __forceinline__ f32x4 mfma32(const u16x8 &a, const u16x8 &b, f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                    acc, 0, 0, 0);
}

__device__ __forceinline__ u16x8 ld8(const unsigned short *p) {
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(u16x8, v);
}

struct Geo {
    int L, Lt, kvol, Ls;
    long n5;   // Ls * kvol (spinor plane length)
};

// float index of channel c (re/im component) of site in an FP32 native spinor
__device__ __forceinline__ long spinor_elem(const Geo &g, int c, int site) {
    const int ri = c & 1, sp = c >> 1;
    const int p = sp & 7, s = sp >> 3;
    return 2 * ((long)p * g.n5 + (long)s * g.kvol + site) + ri;
}

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

// ---- C1: act[site][n] = bf16( sum_k W[n][k] bf16(in[k][site] * scale) ) ----
// One wavefront per 16-site tile; the reduction runs over the Cs = 16 Ls
// spinor channels, 32 per MFMA (lane quad q reads the 4 planes p0..p0+3 of
// one slice: channels 8q .. 8q+7 of the step).  scale: device scalar (NULL:
// 1).  skip: if not NULL and *skip != 0 the launch does nothing.
__global__ __launch_bounds__(256) void k_pw_in(Geo g, const float *__restrict__ in, const float *scale,
                                               const unsigned short *__restrict__ W, unsigned short *__restrict__ act,
                                               const int *skip) {
    if (skip && *skip) return;
    const int lane = threadIdx.x & 63;
    const long tile = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long ntile = (g.kvol + 15) / 16;
    if (tile >= ntile) return;
    const int col = lane & 15, q = lane >> 4;
    const int site = (int)(tile * 16) + col;
    const int sc = site < g.kvol ? site : 0;
    const float sf = scale ? *scale : 1.0f;
    const int Cs = 16 * g.Ls;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < Cs; kk += 32) {
        const int k = kk + 8 * q;
        u16x8 a, b;
        if (k < Cs) {   // odd Ls: the last step's upper half is zero
            const int s = k >> 4, p0 = (k & 15) >> 1;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float2 v =
                    *reinterpret_cast<const float2 *>(in + 2 * ((long)(p0 + j) * g.n5 + (long)s * g.kvol + sc));
                b[2 * j] = bf(v.x * sf);
                b[2 * j + 1] = bf(v.y * sf);
            }
            a = ld8(W + (long)col * Cs + k);
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) a[j] = b[j] = 0;
        }
        acc = mfma32(a, b, acc);
    }
    if (site < g.kvol) {
        ushort4 o;
        o.x = bf(acc[0]);
        o.y = bf(acc[1]);
        o.z = bf(acc[2]);
        o.w = bf(acc[3]);
        *reinterpret_cast<ushort4 *>(act + (long)site * H + 4 * q) = o;
    }
}

// ---- C1^T with the identity path ----------------------------------------------
//   z[c][site] = oscale * ( sum_n WT[c][n] act[site][n] + coef r[c][site] ),
//   coef = exp(*theta) * (*rscale)   (rscale NULL: 1; theta NULL: coef = 0)
// WT[c][n] = W1[n][c] (the transposed BF16 copy).  One wavefront per
// (16-site tile, 16-channel tile); K = 16 is padded to the MFMA's 32 with
// zeros (lane quads 2, 3).  oscale: device scalar (NULL: 1).
__global__ __launch_bounds__(256) void k_pw_out(Geo g, const unsigned short *__restrict__ act,
                                                const unsigned short *__restrict__ WT, const float *__restrict__ r,
                                                const float *rscale, const float *theta, const float *oscale,
                                                float *__restrict__ z, const int *skip) {
    if (skip && *skip) return;
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nct = g.Ls;   // 16 Ls channels / 16
    const long ntile = (g.kvol + 15) / 16;
    if (wave >= ntile * nct) return;
    const int ct = (int)(wave % nct);
    const long tile = wave / nct;
    const int col = lane & 15, q = lane >> 4;
    const int site = (int)(tile * 16) + col;
    const int sc = site < g.kvol ? site : 0;
    const int c0 = ct * 16;
    u16x8 a, b;
    if (q < 2) {
        a = ld8(WT + (long)(c0 + col) * H + 8 * q);
        b = ld8(act + (long)sc * H + 8 * q);
    } else {
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = b[j] = 0;
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma32(a, b, acc);
    if (site >= g.kvol) return;
    const float coef = theta ? expf(*theta) * (rscale ? *rscale : 1.0f) : 0.0f;
    const float os = oscale ? *oscale : 1.0f;
    // rows c0 + 4q .. +3: planes p, p + 1 (re, im) of one slice
    const long e0 = spinor_elem(g, c0 + 4 * q, site), e1 = spinor_elem(g, c0 + 4 * q + 2, site);
    float2 r0 = {0.f, 0.f}, r1 = {0.f, 0.f};
    if (theta) {
        r0 = *reinterpret_cast<const float2 *>(r + e0);
        r1 = *reinterpret_cast<const float2 *>(r + e1);
    }
    *reinterpret_cast<float2 *>(z + e0) = make_float2(os * (acc[0] + coef * r0.x), os * (acc[1] + coef * r0.y));
    *reinterpret_cast<float2 *>(z + e1) = make_float2(os * (acc[2] + coef * r1.x), os * (acc[3] + coef * r1.y));
}

// ---- C2 / C2^T: 3x3x3, 16 -> 16, LDS-staged ------------------------------------
//   out[x][n] = bf16( sum_t sum_k W[t][n][k] in[x + delta_t][k] )
// W: the forward weights W2[t][n][k], or for C2^T the flipped transpose
// W2T[t][n][k] = W2[26 - t][k][n].  One workgroup per (t-plane, band of kRows
// rows); the reduction's 27 taps x 16 channels = 432 run as 14 MFMA steps of
// two taps (the 28th tap is zero).
__host__ __device__ inline int c3_bands(int L) { return (L + kRows - 1) / kRows; }
__host__ __device__ inline int c3_lds_bytes(int L) { return 3 * (kRows + 2) * L * H * 2; }

__global__ __launch_bounds__(256) void k_c3(Geo g, const unsigned short *__restrict__ in,
                                            const unsigned short *__restrict__ W, unsigned short *__restrict__ out,
                                            const int *skip) {
    if (skip && *skip) return;
    extern __shared__ unsigned char smem[];
    uint4 *sh = reinterpret_cast<uint4 *>(smem);   // staged sites, 2 x 16 bytes each
    const int L = g.L, Lt = g.Lt;
    const int nb = c3_bands(L);
    const int it = blockIdx.x / nb, band = blockIdx.x - it * nb;
    const int y0 = band * kRows, Rb = min(kRows, L - y0);
    const int rows = Rb + 2;
    // stage rows y0-1 .. y0+Rb of planes it-1, it, it+1
    const int nstage = 3 * rows * L * 2;
    for (int i = threadIdx.x; i < nstage; i += blockDim.x) {
        const int half = i & 1, sidx = i >> 1;
        const int x = sidx % L, rr = sidx / L;
        const int j = rr % rows, kp = rr / rows;
        int y = y0 - 1 + j;
        y = y < 0 ? y + L : (y >= L ? y - L : y);
        int t = it - 1 + kp;
        t = t < 0 ? t + Lt : (t >= Lt ? t - Lt : t);
        const long site = x + (long)L * (y + (long)L * t);
        sh[i] = *reinterpret_cast<const uint4 *>(in + site * H + 8 * half);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int col = lane & 15, q = lane >> 4;
    const int nsite = Rb * L, ntile = (nsite + 15) / 16;
    for (int tile = wv; tile < ntile; tile += 4) {
        const int bs = tile * 16 + col;
        const int bsc = bs < nsite ? bs : 0;
        const int x = bsc % L, r = bsc / L;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
        for (int step = 0; step < 14; step++) {
            const int tap = 2 * step + (q >> 1);
            u16x8 a, b;
            if (tap < 27) {
                const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dt = tap / 9 - 1;
                int xx = x + dx;
                xx = xx < 0 ? xx + L : (xx >= L ? xx - L : xx);
                const int sidx = ((1 + dt) * rows + (r + 1 + dy)) * L + xx;
                b = __builtin_bit_cast(u16x8, sh[2 * sidx + (q & 1)]);
                a = ld8(W + ((long)tap * H + col) * H + 8 * (q & 1));
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) a[j] = b[j] = 0;
            }
            acc = mfma32(a, b, acc);
        }
        if (bs < nsite) {
            const int y = y0 + r;
            const long site = x + (long)L * (y + (long)L * it);
            ushort4 o;
            o.x = bf(acc[0]);
            o.y = bf(acc[1]);
            o.z = bf(acc[2]);
            o.w = bf(acc[3]);
            *reinterpret_cast<ushort4 *>(out + site * H + 4 * q) = o;
        }
    }
}

// ---- weight gradients ------------------------------------------------------
// Pointwise layer W1 (16 x Cs), two operand pairs accumulated:
//   part[chunk][n][k] = sum_{x in chunk} A1[x][n] bf16(s1 B1[k][x]) + A2[x][n] bf16(B2[k][x])
// A1, A2: activations; B1, B2: FP32 native spinors (s1: device scalar).  The
// MFMA reduction runs over sites, 32 per step: lane quad q takes the sites
// 8q .. 8q+7 of the step.  One wavefront per (chunk, 16-channel tile).
__global__ __launch_bounds__(256) void k_wg_pw(Geo g, int chunk, const unsigned short *__restrict__ A1,
                                               const float *__restrict__ B1, const float *s1,
                                               const unsigned short *__restrict__ A2,
                                               const float *__restrict__ B2, float *__restrict__ part) {
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nkt = g.Ls;   // Cs / 16
    const int nchunk = (g.kvol + chunk - 1) / chunk;
    if (wave >= (long)nchunk * nkt) return;
    const int kt = (int)(wave % nkt), ch = (int)(wave / nkt);
    const int col = lane & 15, q = lane >> 4;
    const int k = kt * 16 + col;   // B column: spinor channel
    const float sc1 = s1 ? *s1 : 1.0f;
    const int x0 = ch * chunk, x1 = min(g.kvol, x0 + chunk);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int xs = x0; xs < x1; xs += 32) {
        u16x8 a1, b1, a2, b2;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int x = xs + 8 * q + j;
            const bool ok = x < x1;
            const int xc = ok ? x : x0;
            const long e = spinor_elem(g, k, xc);
            a1[j] = ok ? A1[(long)xc * H + col] : 0;
            a2[j] = ok ? A2[(long)xc * H + col] : 0;
            b1[j] = ok ? bf(B1[e] * sc1) : 0;
            b2[j] = ok ? bf(B2[e]) : 0;
        }
        acc = mfma32(a1, b1, acc);
        acc = mfma32(a2, b2, acc);
    }
    const int Cs = 16 * g.Ls;
    float *dst = part + (long)ch * H * Cs;
#pragma unroll
    for (int i = 0; i < 4; i++) dst[(long)(4 * q + i) * Cs + k] = acc[i];
}

// 3x3x3 layer W2 (27 x 16 x 16), two operand pairs accumulated:
//   part[chunk][t][n][k] = sum_{x in chunk} A1[x][n] B1[x + delta_t][k] + A2[x][n] B2[x + delta_t][k]
// One wavefront per (chunk, tap).
__global__ __launch_bounds__(256) void k_wg_c3(Geo g, int chunk, const unsigned short *__restrict__ A1,
                                               const unsigned short *__restrict__ B1,
                                               const unsigned short *__restrict__ A2,
                                               const unsigned short *__restrict__ B2, float *__restrict__ part) {
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nchunk = (g.kvol + chunk - 1) / chunk;
    if (wave >= 27L * nchunk) return;
    const int t = (int)(wave % 27), ch = (int)(wave / 27);
    const int col = lane & 15, q = lane >> 4;
    const int x0 = ch * chunk, x1 = min(g.kvol, x0 + chunk);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int xs = x0; xs < x1; xs += 32) {
        u16x8 a1, b1, a2, b2;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int x = xs + 8 * q + j;
            const bool ok = x < x1;
            const int xc = ok ? x : x0;
            const long nb = shift_site(g, xc, t);
            a1[j] = ok ? A1[(long)xc * H + col] : 0;
            a2[j] = ok ? A2[(long)xc * H + col] : 0;
            b1[j] = ok ? B1[nb * H + col] : 0;
            b2[j] = ok ? B2[nb * H + col] : 0;
        }
        acc = mfma32(a1, b1, acc);
        acc = mfma32(a2, b2, acc);
    }
    float *dst = part + ((long)ch * 27 + t) * H * H;
#pragma unroll
    for (int i = 0; i < 4; i++) dst[(4 * q + i) * H + col] = acc[i];
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

// BF16 copies of the FP32 weights w = [W1 (16 x Cs) | W2 (27 x 16 x 16) | theta]:
// W1, its transpose W1T (Cs x 16), W2 and its flipped transpose
// W2T[t][n][k] = W2[26 - t][k][n].  One thread per element of W1 or W2.
__global__ __launch_bounds__(256) void k_weights_bf16(int Cs, const float *__restrict__ w,
                                                      unsigned short *__restrict__ w1, unsigned short *__restrict__ w1t,
                                                      unsigned short *__restrict__ w2,
                                                      unsigned short *__restrict__ w2t) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const long n1 = (long)H * Cs;
    if (i < n1) {
        const int n = (int)(i / Cs), k = (int)(i % Cs);
        const unsigned short b = bf(w[i]);
        w1[i] = b;
        w1t[(long)k * H + n] = b;
    } else if (i < n1 + 27 * H * H) {
        const int j = (int)(i - n1);
        const int t = j / (H * H), n = (j / H) % H, k = j % H;
        const unsigned short b = bf(w[i]);
        w2[j] = b;
        w2t[((26 - t) * H + k) * H + n] = b;
    }
}

// Adam step on the FP32 master weights
__global__ __launch_bounds__(256) void k_adam(long n, float *__restrict__ w, float *__restrict__ m,
                                              float *__restrict__ v, const float *__restrict__ g, float lr, float b1,
                                              float b2, float c1, float c2, float eps) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace nn
}  // namespace chs
