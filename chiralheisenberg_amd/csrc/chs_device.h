// chs_device.h -- gfx950 device code of libchs: the domain-wall operator
// (rhmc/cheisdwf.f:2634-3501), its fused solver epilogues, the multi-shift
// Lanczos vector updates (rhmc/cheisdwf.f:655-695), CG vector updates
// (rhmc/cheisdwf.f:1098-1148), the wall-only force (rhmc/cheisdwf.f:733-849)
// and deterministic FP64 reductions.
//
// Spinors are 8 spin-flavour planes (d + 4 f) of Ls*kvol complex numbers;
// the order inside a plane (the "native" layout, s outermost) and the stencil
// itself are in chs_stencil.h.  Everything here is elementwise over planes or
// acts on the scalar state of the solvers.
//
// Storage precisions: FP64 (double2), FP32 (float2), BF16 and FP16 (re, im
// packed in one 32-bit word).  FP64 storage computes in FP64; the others
// compute and accumulate in FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chs {

constexpr int kBlock = 256;     // 4 wavefronts of 64
constexpr int kMaxShift = 32;   // CHS_MAX_SHIFTS

struct Lat {
    int L, Lt, Ls;
    int kvol;       // L*L*Lt
    long n5;        // Ls*kvol: complex numbers per spin-flavour plane
    long ps;        // plane stride in elements (>= n5)
    double diag;    // 4 - M5
    int apbc;       // antiperiodic in t (ibound == -1)
};

// ---- complex arithmetic in the compute type ------------------------------
template <typename T> struct cx { T r, i; };
template <typename T> __device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.r + b.r, a.i + b.i}; }
template <typename T> __device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.r - b.r, a.i - b.i}; }
template <typename T> __device__ __forceinline__ cx<T> cscale(T s, cx<T> a) { return {s * a.r, s * a.i}; }
// e * i * a, e = +1 or -1
template <typename T> __device__ __forceinline__ cx<T> cmuli(T e, cx<T> a) { return {-e * a.i, e * a.r}; }
template <typename T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}

// ---- storage traits --------------------------------------------------------
struct F64 {
    typedef double2 st;
    typedef double rt;
    static constexpr bool scaled = false;
    static __device__ __forceinline__ cx<double> ld(const st *p, long i) {
        double2 v = p[i];
        return {v.x, v.y};
    }
    static __device__ __forceinline__ void sto(st *p, long i, cx<double> v) { p[i] = make_double2(v.r, v.i); }
};
struct F32 {
    typedef float2 st;
    typedef float rt;
    static constexpr bool scaled = false;
    static __device__ __forceinline__ cx<float> ld(const st *p, long i) {
        float2 v = p[i];
        return {v.x, v.y};
    }
    static __device__ __forceinline__ void sto(st *p, long i, cx<float> v) { p[i] = make_float2(v.r, v.i); }
    // elements i, i+1 (i even) in one 16-byte access
    typedef float4 st2;
    static __device__ __forceinline__ void ld2(const st *p, long i, cx<float> &a, cx<float> &b) {
        const float4 v = *reinterpret_cast<const float4 *>(p + i);
        a = {v.x, v.y};
        b = {v.z, v.w};
    }
    static __device__ __forceinline__ st2 pack2(cx<float> a, cx<float> b) { return make_float4(a.r, a.i, b.r, b.i); }
};
__device__ __forceinline__ uint32_t f2bf(float x) {
    // round to nearest even; inputs are finite
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
struct B16 {
    typedef uint32_t st;
    typedef float rt;
    static constexpr bool scaled = false;
    static __device__ __forceinline__ cx<float> ld(const st *p, long i) {
        uint32_t v = p[i];
        return {__uint_as_float(v << 16), __uint_as_float(v & 0xffff0000u)};
    }
    static __device__ __forceinline__ void sto(st *p, long i, cx<float> v) { p[i] = f2bf(v.r) | (f2bf(v.i) << 16); }
    // elements i, i+1 (i even) in one 8-byte access
    typedef uint2 st2;
    static __device__ __forceinline__ void ld2(const st *p, long i, cx<float> &a, cx<float> &b) {
        const uint2 v = *reinterpret_cast<const uint2 *>(p + i);
        a = {__uint_as_float(v.x << 16), __uint_as_float(v.x & 0xffff0000u)};
        b = {__uint_as_float(v.y << 16), __uint_as_float(v.y & 0xffff0000u)};
    }
    static __device__ __forceinline__ st2 pack2(cx<float> a, cx<float> b) {
        return make_uint2(f2bf(a.r) | (f2bf(a.i) << 16), f2bf(b.r) | (f2bf(b.i) << 16));
    }
};

// FP16 storage (re, im as IEEE binary16 in one 32-bit word, round to nearest
// even), FP32 compute.  FP16 keeps 11 significant bits (BF16: 8) but only 5
// exponent bits, so the CG that stores p and M p this way keeps them scaled
// by the current residual's rms (scaled = true: CgState.sigma).
struct H16 {
    typedef uint32_t st;
    typedef float rt;
    static constexpr bool scaled = true;
    static __device__ __forceinline__ float h2f(uint32_t b) {
        return (float)__builtin_bit_cast(_Float16, (unsigned short)b);
    }
    static __device__ __forceinline__ uint32_t f2h(float x) {
        return (uint32_t)__builtin_bit_cast(unsigned short, (_Float16)x);
    }
    static __device__ __forceinline__ cx<float> ld(const st *p, long i) {
        const uint32_t v = p[i];
        return {h2f(v & 0xffffu), h2f(v >> 16)};
    }
    static __device__ __forceinline__ void sto(st *p, long i, cx<float> v) { p[i] = f2h(v.r) | (f2h(v.i) << 16); }
};

// Blocks are handed to the 8 XCDs round-robin; give each XCD a contiguous
// range of logical blocks so that the x/y/t neighbours a workgroup reads were
// brought into the same XCD's L2 by its neighbouring workgroups.
__device__ __forceinline__ long logical_block() {
    const unsigned b = blockIdx.x, nb = gridDim.x;
    if ((nb & 7u) == 0u) return (long)(b & 7u) * (nb >> 3) + (b >> 3);
    return b;
}

// Deterministic block sum (xor butterfly inside each wavefront, then the 4
// wavefront sums in fixed order).  Result valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __shared__ double sh[kBlock / 64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int k = 0; k < kBlock / 64; k++) t += sh[k];
    }
    __syncthreads();
    return t;
}

// W psi for the wall block W = c * (phi.sigma [+ hpin sigma3 at site 0]):
// imass=1: c = 1 on both walls (rhmc/cheisdwf.f:2912-2995);
// imass=3: c = -i on the s=Ls wall, +i on the s=1 wall, no pinning
// (rhmc/cheisdwf.f:2996-3066).  M^dagger uses the same blocks on the
// opposite projectors (rhmc/cheisdwf.f:3345-3498).
template <typename T>
__device__ __forceinline__ void wall_apply(cx<T> &o0, cx<T> &o1, cx<T> v0, cx<T> v1, T c1, T c2, T c3,
                                           int imass, bool last_wall) {
    // (phi.sigma) v = (c3 v0 + (c1 - i c2) v1, (c1 + i c2) v0 - c3 v1)
    cx<T> w0 = {c3 * v0.r + c1 * v1.r + c2 * v1.i, c3 * v0.i + c1 * v1.i - c2 * v1.r};
    cx<T> w1 = {c1 * v0.r - c2 * v0.i - c3 * v1.r, c1 * v0.i + c2 * v0.r - c3 * v1.i};
    if (imass == 3) {
        const T e = last_wall ? T(-1) : T(1);
        w0 = cmuli(e, w0);
        w1 = cmuli(e, w1);
    }
    o0 = cadd(o0, w0);
    o1 = cadd(o1, w1);
}

// ---- epilogues of the fused stencil kernels (chs_stencil.h) ---------------
enum Epi {
    EPI_STORE = 0,   // out = M in
    EPI_NORM = 1,    // out = M in, partial += |out|^2
    EPI_SCALEQ = 2,  // in = r: out = (M r)/beta, aux0 = q = r/beta (Lanczos, scal[0] = beta)
    EPI_DOT = 3,     // out = M in, partial += Re <aux0, out>
    EPI_CG = 4,      // no store: aux1 (x) += alpha aux0 (p), aux2 (r) -= alpha (M in),
                     // partial += |r|^2; alpha = scal[0]; flag bit0: skip x update
    EPI_CGQ = 5,     // EPI_CG that also stores out = M in (flexible CG keeps A p)
    EPI_CGW = 6      // EPI_CG with alpha formed inside the launch from the previous
                     // launch's |M p|^2 partials (cgs = CgState: cgw_publish / cgw_wait)
};

struct EpiArgs {
    double *partial;
    const double *scal;
    void *aux0, *aux1, *aux2;
    const int *skip;   // if not NULL and *skip != 0 the launch does nothing
    int flag;
    void *cgs;         // EPI_CGW: the CgState that receives and publishes alpha
};

// ---- deterministic FP64 reductions -----------------------------------------
// One workgroup sums n partials in a fixed order: thread t adds part[t],
// part[t + kBlock], ... in that order, then block_sum.  The sum is
// latency-bound: up to kU loads per thread are issued (predicated on the
// bound) before the first addition, so n <= kU * kBlock partials cost one
// memory round trip.  kU changes the batching of the loads, not the order of
// the additions, so every kU gives the same bits.  16 keeps the registers of
// the stencils that inline it (the in-launch alpha) low; the one-workgroup
// kernels over the Lanczos update's one partial per 256 elements (14,156 at
// L=24, Ls=32) use 64.
template <int kU = 16>
__device__ __forceinline__ double sum_partials(const double *__restrict__ part, int n) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += kU * kBlock) {
        double t[kU];
#pragma unroll
        for (int k = 0; k < kU; k++) {
            const int j = i + k * kBlock;
            t[k] = j < n ? part[j] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kU; k++)
            if (i + k * kBlock < n) v += t[k];
    }
    return block_sum(v);
}

// |a|^2 (mode 0) or Re <a, b> (mode 1) over n complex FP64 numbers, or
// sum a_i^2 over n reals (mode 2); one partial per workgroup.
__global__ __launch_bounds__(kBlock) void k_reduce(long n, const double2 *__restrict__ a,
                                                   const double2 *__restrict__ b, int mode,
                                                   double *__restrict__ partial) {
    const long lb = logical_block();
    const long stride = (long)gridDim.x * kBlock;
    double v = 0.0;
    if (mode == 2) {
        const double *ar = (const double *)a;
        for (long i = lb * kBlock + threadIdx.x; i < n; i += stride) v += ar[i] * ar[i];
    } else {
        for (long i = lb * kBlock + threadIdx.x; i < n; i += stride) {
            const double2 x = a[i];
            if (mode == 0) v += x.x * x.x + x.y * x.y;
            else {
                const double2 y = b[i];
                v += x.x * y.x + x.y * y.y;
            }
        }
    }
    const double t = block_sum(v);
    if (threadIdx.x == 0) partial[lb] = t;
}

__global__ __launch_bounds__(kBlock) void k_finish_sum(const double *__restrict__ partial, int n,
                                                       double *__restrict__ out) {
    const double t = sum_partials(partial, n);
    if (threadIdx.x == 0) out[0] = t;
}

// ---- elementwise kernels (FP64 complex unless stated) -------------------
// y = a x + b y over n complex numbers, with a and b host scalars
__global__ __launch_bounds__(kBlock) void k_axpby(long n, double a, const double2 *__restrict__ x, double b,
                                                  double2 *__restrict__ y) {
    const long i = logical_block() * kBlock + threadIdx.x;
    if (i < n) {
        const double2 xv = x[i];
        double2 yv = (b == 0.0) ? make_double2(0.0, 0.0) : y[i];
        yv.x = a * xv.x + b * yv.x;
        yv.y = a * xv.y + b * yv.y;
        y[i] = yv;
    }
}

template <class SD, class SS>
__global__ __launch_bounds__(kBlock) void k_convert(long n, typename SD::st *__restrict__ dst,
                                                    const typename SS::st *__restrict__ src) {
    const long i = logical_block() * kBlock + threadIdx.x;
    if (i < n) {
        const cx<typename SS::rt> v = SS::ld(src, i);
        SD::sto(dst, i, cx<typename SD::rt>{(typename SD::rt)v.r, (typename SD::rt)v.i});
    }
}

// Fill n complex numbers (or n reals when REAL) with pseudo-random values in
// [-0.5, 0.5) from a counter hash (splitmix64 finaliser of seed + index), for
// the timing loops.  Their inputs must carry realistic bit patterns: on this
// hardware a stencil over all-zero buffers runs ~15% faster than over random
// ones, and a plain copy likewise (zero bits draw less power in HBM and fabric).
__device__ __forceinline__ double hash_unit(unsigned long long x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    x ^= x >> 31;
    return (double)(x >> 11) * 0x1.0p-53 - 0.5;
}
template <class S>
__global__ __launch_bounds__(kBlock) void k_fill_hash(long n, typename S::st *__restrict__ dst,
                                                      unsigned long long seed) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const unsigned long long k = seed * 0x100000001b3ull + 2ull * (unsigned long long)i;
    typedef typename S::rt T;
    S::sto(dst, i, cx<T>{(T)hash_unit(k), (T)hash_unit(k + 1)});
}
__global__ __launch_bounds__(kBlock) void k_fill_hash_real(long n, double *__restrict__ dst,
                                                           unsigned long long seed) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = hash_unit(seed * 0x100000001b3ull + (unsigned long long)i);
}

// ---- multi-shift Lanczos (rhmc/cheisdwf.f:599-745) -----------------------
struct LzState {
    double beta;        // |r| of the current Lanczos vector (beta0 of the next step)
    double alphatild;
    double phimod, resid;
    int done, itercg, ndiag, pad;
    // Shift freezing (0 = off, the reference's behaviour): once
    // |phimod rho_k| < freeze * 2 resid, shift k's vectors are no longer
    // updated (bit k of amask cleared); its scalars, and so the stopping
    // test and itercg, run on unchanged.
    double freeze;
    unsigned amask, pad2;
    double aden[kMaxShift];
    double d[kMaxShift], rho[kMaxShift], amu[kMaxShift];
};

// After |phi|^2: beta = phimod = |phi|, resid = sqrt(kferm) res.
__global__ void k_lz_init(LzState *st, const double *__restrict__ partial, int np, double resid) {
    const double t = sum_partials(partial, np);
    if (threadIdx.x == 0) {
        st->beta = sqrt(t);
        st->phimod = st->beta;
        st->resid = resid;
        st->done = 0;
        st->itercg = 0;
    }
}

// alphatild = Re <q, M^dagger M q>; then the shift recurrences of one step
// (rhmc/cheisdwf.f:666-687), which use the previous beta.
__device__ __forceinline__ void lz_alpha(LzState *st, const double *__restrict__ partial, int np, int iter) {
    const double at = sum_partials(partial, np);
    __shared__ double sat;
    if (threadIdx.x == 0) {
        st->alphatild = at;
        sat = at;
    }
    __syncthreads();
    const int k = threadIdx.x;
    if (k < st->ndiag) {
        const double alpha = sat + st->aden[k];
        const double b0 = st->beta;
        if (iter == 1) {
            st->d[k] = alpha;
            st->rho[k] = b0 / alpha;
        } else {
            const double amu = b0 / st->d[k];
            const double dm1 = st->d[k];
            const double dn = alpha - b0 * amu;
            st->amu[k] = amu;
            st->d[k] = dn;
            st->rho[k] = -amu * dm1 * st->rho[k] / dn;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_lz_alpha(LzState *st, const double *__restrict__ partial, int np,
                                                     int iter) {
    if (st->done) return;
    lz_alpha(st, partial, np, iter);
}

// new beta = |r|; convergence max_k |phimod rho_k| < 2 resid for iter >= 2
// (rhmc/cheisdwf.f:697-715).  A non-finite |r|^2 halts the solve with
// done = 2 (the host reports it; the reference has no such test).
__device__ __forceinline__ void lz_beta(LzState *st, const double *__restrict__ partial, int np, int iter) {
    const double t = sum_partials<64>(partial, np);
    if (threadIdx.x == 0) {
        st->beta = sqrt(t);
        st->itercg = iter;
        if (!isfinite(t)) {
            st->done = 2;
        } else if (iter >= 2) {
            double rhomax = 0.0;
            for (int k = 0; k < st->ndiag; k++) {
                const double a = fabs(st->phimod * st->rho[k]);
                if (k == 0 || a > rhomax) rhomax = a;
            }
            if (rhomax < 2.0 * st->resid) st->done = 1;
            if (st->freeze > 0.0) {
                unsigned m = st->amask;
                for (int k = 0; k < st->ndiag; k++)
                    if (fabs(st->phimod * st->rho[k]) < st->freeze * 2.0 * st->resid) m &= ~(1u << k);
                st->amask = m;
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_lz_beta(LzState *st, const double *__restrict__ partial, int np,
                                                    int iter) {
    if (st->done) return;
    lz_beta(st, partial, np, iter);
}

// Shift update of one element i: r = w - alphatild q - beta0 qm1 and, for
// every shift, p_k = q - mu_k p_k (iter 1: p_k = q) and
//   CMB = false: x_k += rho_k p_k (iter 1: x_k = rho_k q), x shift-major
//                (shift k at offset k*n);
//   CMB = true:  x += sum_k c_k p_k with c_k = anum[k+1] rho_k, x the single
//                running rational combination a0 phi + sum_k anum[k+1] x_k,
//                which is all qmrherm's iflag = 0 returns: the x_k are never
//                stored, which halves the shift traffic.
// Coefficients come from shared memory (lz_coefs).  Returns |r_i|^2.
struct LzCoef {
    double amu[kMaxShift], rho[kMaxShift], c[kMaxShift];
};

template <bool CMB>
__device__ __forceinline__ void lz_coefs(LzCoef &sc, const LzState *__restrict__ st, int ndiag,
                                         const double *__restrict__ anum) {
    if (threadIdx.x < ndiag) {
        sc.amu[threadIdx.x] = st->amu[threadIdx.x];
        sc.rho[threadIdx.x] = st->rho[threadIdx.x];
        if (CMB) sc.c[threadIdx.x] = anum[threadIdx.x + 1] * st->rho[threadIdx.x];
    }
    __syncthreads();
}

// Which elements of the x_k the shift update keeps (CMB = false): those whose
// offset inside their plane, i mod n5, lies outside [lo, hi).  The force
// (rhmc/cheisdwf.f:733-849) reads x_k only on the slices s = 0, 1, Ls-2,
// Ls-1 (the wall slices and their s-neighbours: derivs uses y = x_k and
// z = M x_k at s in {1, Ls} only), so it passes lo = 2 kvol, hi = (Ls-2) kvol
// and the other slices of every x_k are neither read nor written.  The
// default keeps everything.
struct XKeep {
    unsigned n5 = 1, lo = 0, hi = 0;
    __device__ __forceinline__ bool keep(long i) const {
        const unsigned e = (unsigned)i % n5;
        return e < lo || e >= hi;
    }
};

// Storage S of r, q, w, p_k (and x_k); arithmetic in FP64 on loaded values,
// rounded on store.  With CMB the running combination x is FP64 (double2).
template <class S, int NS, bool CMB>
__device__ __forceinline__ double lz_update_elem(long i, long n, const typename S::st *__restrict__ w,
                                                 const typename S::st *__restrict__ q,
                                                 const typename S::st *__restrict__ qm1,
                                                 typename S::st *__restrict__ r, typename S::st *__restrict__ p,
                                                 void *__restrict__ xp, const LzState *__restrict__ st,
                                                 const LzCoef &sc, int ndiag, int iter, unsigned amask,
                                                 XKeep xkeep = XKeep()) {
    typedef cx<typename S::rt> C;
    const double at = st->alphatild, b0 = st->beta;
    const C qs = S::ld(q, i), ws = S::ld(w, i), qms = S::ld(qm1, i);
    const double2 qv = make_double2(qs.r, qs.i), wv = make_double2(ws.r, ws.i), qm = make_double2(qms.r, qms.i);
    double2 rv;
    rv.x = wv.x - at * qv.x - b0 * qm.x;
    rv.y = wv.y - at * qv.y - b0 * qm.y;
    S::sto(r, i, C{(typename S::rt)rv.x, (typename S::rt)rv.y});
    const int ns = NS > 0 ? NS : ndiag;
    if (CMB) {
        double2 *x = (double2 *)xp;
        double2 xv = x[i];
        if (iter == 1) {
#pragma unroll 4
            for (int k = 0; k < ns; k++) {
                S::sto(p, (long)k * n + i, qs);
                xv.x += sc.c[k] * qv.x;
                xv.y += sc.c[k] * qv.y;
            }
        } else {
#pragma unroll 4
            for (int k = 0; k < ns; k++) {
                if (!((amask >> k) & 1u)) continue;
                const C ps = S::ld(p, (long)k * n + i);
                double2 pv;
                pv.x = qv.x - sc.amu[k] * (double)ps.r;
                pv.y = qv.y - sc.amu[k] * (double)ps.i;
                S::sto(p, (long)k * n + i, C{(typename S::rt)pv.x, (typename S::rt)pv.y});
                xv.x += sc.c[k] * pv.x;
                xv.y += sc.c[k] * pv.y;
            }
        }
        x[i] = xv;
    } else {
        typename S::st *x = (typename S::st *)xp;
        const bool kx = xkeep.keep(i);
        if (iter == 1) {
#pragma unroll 4
            for (int k = 0; k < ns; k++) {
                S::sto(p, (long)k * n + i, qs);
                if (kx)
                    S::sto(x, (long)k * n + i,
                           C{(typename S::rt)(sc.rho[k] * qv.x), (typename S::rt)(sc.rho[k] * qv.y)});
            }
        } else if (kx) {
#pragma unroll 4
            for (int k = 0; k < ns; k++) {
                if (!((amask >> k) & 1u)) continue;
                const C ps = S::ld(p, (long)k * n + i), xs = S::ld(x, (long)k * n + i);
                double2 pv, xv;
                pv.x = qv.x - sc.amu[k] * (double)ps.r;
                pv.y = qv.y - sc.amu[k] * (double)ps.i;
                xv.x = (double)xs.r + sc.rho[k] * pv.x;
                xv.y = (double)xs.i + sc.rho[k] * pv.y;
                S::sto(p, (long)k * n + i, C{(typename S::rt)pv.x, (typename S::rt)pv.y});
                S::sto(x, (long)k * n + i, C{(typename S::rt)xv.x, (typename S::rt)xv.y});
            }
        } else {
#pragma unroll 4
            for (int k = 0; k < ns; k++) {
                if (!((amask >> k) & 1u)) continue;
                const C ps = S::ld(p, (long)k * n + i);
                double2 pv;
                pv.x = qv.x - sc.amu[k] * (double)ps.r;
                pv.y = qv.y - sc.amu[k] * (double)ps.i;
                S::sto(p, (long)k * n + i, C{(typename S::rt)pv.x, (typename S::rt)pv.y});
            }
        }
    }
    return rv.x * rv.x + rv.y * rv.y;
}

// EPT consecutive element blocks of kBlock per workgroup (block lb covers
// element blocks lb*EPT .. lb*EPT+EPT-1), one |r|^2 partial per element block,
// so the partials do not depend on EPT.
template <class S, int NS, bool CMB, int EPT>
__device__ __forceinline__ void lz_update(long n, const typename S::st *__restrict__ w,
                                          const typename S::st *__restrict__ q,
                                          const typename S::st *__restrict__ qm1, typename S::st *__restrict__ r,
                                          typename S::st *__restrict__ p, void *__restrict__ x,
                                          const double *__restrict__ anum, const LzState *__restrict__ st, int ndiag,
                                          int iter, double *__restrict__ partial, long lb, XKeep xkeep = XKeep()) {
    __shared__ LzCoef sc;
    lz_coefs<CMB>(sc, st, ndiag, anum);
    const unsigned amask = st->amask;   // uniform: the skips are scalar branches
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const long eb = lb * EPT + e;
        const long i = eb * kBlock + threadIdx.x;
        double part = 0.0;
        if (i < n) part = lz_update_elem<S, NS, CMB>(i, n, w, q, qm1, r, p, x, st, sc, ndiag, iter, amask, xkeep);
        const double t = block_sum(part);
        if (threadIdx.x == 0) partial[eb] = t;
    }
}

template <class S, int NS, bool CMB>
__global__ __launch_bounds__(kBlock) void k_lz_update(long n, const typename S::st *__restrict__ w,
                                                      const typename S::st *__restrict__ q,
                                                      const typename S::st *__restrict__ qm1,
                                                      typename S::st *__restrict__ r, typename S::st *__restrict__ p,
                                                      void *__restrict__ x, const double *__restrict__ anum,
                                                      const LzState *__restrict__ st, int ndiag, int iter,
                                                      double *__restrict__ partial, XKeep xkeep) {
    if (st->done) return;
    lz_update<S, NS, CMB, 1>(n, w, q, qm1, r, p, x, anum, st, ndiag, iter, partial, logical_block(), xkeep);
}

// ---- propagator measurement (rhmc/cheisdwf.f:897-1058) -------------------
// Wall source: element (s = 0, every site of t-slice t0) of plane p set to val
// in a native FP64 spinor (the rest was cleared by the caller).
__global__ __launch_bounds__(kBlock) void k_wall_source(Lat lat, double2 *__restrict__ xi, int t0, int plane,
                                                        double val) {
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    const long L2 = (long)lat.L * lat.L;
    if (i < L2) xi[(long)plane * lat.ps + (long)t0 * L2 + i] = make_double2(val, 0.0);
}

// Sums over the L^2 sites of each t-slice for up to four (s, plane) pairs of
// a native FP64 spinor: out[q * Lt + t], one workgroup per (t, q), fixed order.
struct SliceReq {
    int s[4], plane[4];
};
__global__ __launch_bounds__(kBlock) void k_slice_sums(Lat lat, const double2 *__restrict__ x, SliceReq rq,
                                                       double2 *__restrict__ out) {
    const int t = blockIdx.x, q = blockIdx.y;
    const long L2 = (long)lat.L * lat.L;
    const double2 *base = x + (long)rq.plane[q] * lat.ps + (long)rq.s[q] * lat.kvol + (long)t * L2;
    double re = 0.0, im = 0.0;
    for (long i = threadIdx.x; i < L2; i += kBlock) {
        const double2 v = base[i];
        re += v.x;
        im += v.y;
    }
    re = block_sum(re);
    im = block_sum(im);
    if (threadIdx.x == 0) out[(long)q * lat.Lt + t] = make_double2(re, im);
}

// x = a0 phi + sum_k a_k x_k, summed in shift order (rhmc/cheisdwf.f:722-727)
__global__ __launch_bounds__(kBlock) void k_lz_combine(long n, const double2 *__restrict__ phi,
                                                       const double2 *__restrict__ xk, int ndiag,
                                                       const double *__restrict__ anum, double2 *__restrict__ x) {
    const long i = logical_block() * kBlock + threadIdx.x;
    if (i < n) {
        const double2 pv = phi[i];
        double2 acc = make_double2(anum[0] * pv.x, anum[0] * pv.y);
        for (int k = 0; k < ndiag; k++) {
            const double2 v = xk[(long)k * n + i];
            acc.x += anum[k + 1] * v.x;
            acc.y += anum[k + 1] * v.y;
        }
        x[i] = acc;
    }
}

// ---- pseudofermion force (rhmc/cheisdwf.f:733-849) -----------------------
// Wall bilinears of y and z = M y for one site, accumulated into F[3]
// (imass=1: Re/Im/Re of t w1, w2, w3; imass=3: the rotated pattern of
// rhmc/cheisdwf.f:804-845), t = -2 a.
__device__ __forceinline__ void derivs_site(double &F1, double &F2, double &F3, const double2 (&yl)[4][2],
                                            const double2 (&zl)[4][2], double t, int imass) {
    // yl: y at the wall that pairs with z in zl.  d=0,1: y(s=Ls) with z(s=1);
    // d=2,3: y(s=1) with z(s=Ls); the caller arranges the slices.
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const double y0r = yl[d][0].x, y0i = -yl[d][0].y, y1r = yl[d][1].x, y1i = -yl[d][1].y;
        const double z0r = zl[d][0].x, z0i = zl[d][0].y, z1r = zl[d][1].x, z1i = zl[d][1].y;
        const double Ar = y0r * z1r - y0i * z1i, Ai = y0r * z1i + y0i * z1r;
        const double Br = y1r * z0r - y1i * z0i, Bi = y1r * z0i + y1i * z0r;
        const double Cr = y0r * z0r - y0i * z0i, Ci = y0r * z0i + y0i * z0r;
        const double Dr = y1r * z1r - y1i * z1i, Di = y1r * z1i + y1i * z1r;
        if (imass == 1) {
            F1 -= t * (Ar + Br);
            F2 -= t * (Ai - Bi);
            F3 -= t * (Cr - Dr);
        } else if (imass == 3) {
            const double sg = d < 2 ? 1.0 : -1.0;
            F1 -= sg * t * (Ai + Bi);
            F2 += sg * t * (Ar - Br);
            F3 -= sg * t * (Ci - Di);
        }
    }
}

// dSdchi[a][site] += sum_k contrib[k][a][site], in shift order
__global__ __launch_bounds__(kBlock) void k_force_sum(int kvol, int ndiag, const double *__restrict__ contrib,
                                                      double *__restrict__ F) {
    const long i = logical_block() * kBlock + threadIdx.x;
    if (i >= 3L * kvol) return;
    const int a = (int)(i / kvol);
    const int site = (int)(i - (long)a * kvol);
    double acc = F[i];
    for (int k = 0; k < ndiag; k++) acc += contrib[((long)k * 3 + a) * kvol + site];
    F[i] = acc;
}

// F = 2 beta chi (rhmc/cheisdwf.f:494-498)
__global__ __launch_bounds__(kBlock) void k_force_init(long n, double beta, const double *__restrict__ chi,
                                                       double *__restrict__ F) {
    const long i = logical_block() * kBlock + threadIdx.x;
    if (i < n) F[i] = 2.0 * beta * chi[i];
}

// ---- CG state machine --------------------------------------------------------
// Scalars of CGNR (rhmc/cheisdwf.f:1064-1157) and of the reliable-update
// CG used for mixed precision.  halt = done | hold; every kernel of an
// iteration does nothing once halt is set, so the host may enqueue several
// iterations ahead and poll.
struct CgState {
    double alpha, alphan, betad, beta;
    // scaled storage (H16): p and M p are held divided by sigma, the rms of
    // the residual when p was formed; alpha_s = alpha * sigma multiplies the
    // stored M^dagger M p, and the x/p update is x += xp_a p, p = xp_r r +
    // xp_p p (xp_a = alpha sigma, xp_r = 1 / sigma', xp_p = beta sigma / sigma')
    double sigma, alpha_s, xp_a, xp_r, xp_p, nelem;
    // the x increment coefficient of the previous iteration (alpha, or alpha
    // sigma when scaled): the reliable-update CG adds two increments every
    // second iteration (k_cg_xpnext)
    double xp_a_prev;
    double resid;        // reference rule: stop when |r|^2 < resid
    double rnorm2;       // |r|^2 of the current iterated residual
    double maxr2;        // |r|^2 at the last reliable update
    double delta2;       // reliable update when |r|^2 < delta2 * maxr2
    double stop2;        // |r|^2 below which a true-residual check is forced
    double rn_hold;      // iterated |r|^2 when the host was called
    int done;            // converged (or given up)
    int hold;            // waiting for the host
    int halt;            // done | hold
    int iter;            // iterations taken
    int maxiter;
    int scaled;          // p and M p are stored scaled (H16)
    // the fused CG stencil's multiplier (alpha, or alpha sigma) as published
    // by its workgroup 0, kAlphaEmpty until then (EPI_CGW)
    unsigned long long aslot;
    int aerr;            // a wave gave up waiting for aslot
    int pad_;
};

// In-launch alpha of the reliable-update CG (EPI_CGW).  Workgroup 0 of the
// fused CG stencil sums the |M p|^2 partials the previous launch left, forms
// alpha = rnorm2 / |M p|^2 (scaled storage: alpha sigma as the multiplier)
// and publishes the multiplier in st->aslot with one agent-scope atomic
// store; every other wave, once its stencil work is done, reads st->aslot
// with agent-scope atomic loads until it is published.  Workgroup 0 is
// dispatched first and waits for nothing, so the wait ends; it is still
// bounded (kAlphaSpin polls of ~1 us, about a second), after which the wave
// sets st->aerr and uses NaN, so that every wave exits.  k_rcg_beta puts
// kAlphaEmpty back for the next iteration.  The partials are read before any
// workgroup of the launch can overwrite them: each writes its own partial
// only after it has read the published alpha.
constexpr unsigned long long kAlphaEmpty = 0x7FF4DEADBEEF0001ull;   // a signalling NaN arithmetic never makes
constexpr int kAlphaSpin = 1 << 20;

__device__ __forceinline__ void cgw_publish(CgState *st, const double *__restrict__ partial, int np) {
    const int scaled = st->scaled;
    const double rn = st->rnorm2, sig = st->sigma;
    const double t = sum_partials(partial, np);
    if (threadIdx.x != 0) return;
    double m;
    if (scaled) {
        // t = |M p|^2 / sigma^2
        const double a = rn / (sig * sig * t);
        st->alpha = a;
        st->alpha_s = a * sig;
        m = a * sig;
    } else {
        m = rn / t;
        st->alpha = m;
    }
    __hip_atomic_store(&st->aslot, __builtin_bit_cast(unsigned long long, m), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double cgw_wait(CgState *st) {
    for (int n = 0;; n++) {
        const unsigned long long v = __hip_atomic_load(&st->aslot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v != kAlphaEmpty) return __builtin_bit_cast(double, v);
        if (n >= kAlphaSpin) {
            __hip_atomic_store(&st->aerr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return __builtin_nan("");
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// reference CGNR after |r|^2 (rhmc/cheisdwf.f:1130-1148).  The reference
// updates p before its convergence test; p is not used after convergence, so
// halting before the p update leaves x and the iteration count unchanged.
__global__ void k_cg_beta_ref(CgState *st, const double *__restrict__ partial, int np, int nx) {
    if (st->halt) return;
    const double betan = sum_partials(partial, np);
    if (threadIdx.x == 0) {
        __hip_atomic_store(&st->aslot, kAlphaEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        double beta = betan / st->betad;
        st->betad = betan;
        st->alphan = betan;
        if (nx == 1) beta = 0.0;
        st->beta = beta;
        st->iter = nx;
        st->rnorm2 = betan;
        // a wave of the fused stencil that gave up waiting for alpha (aerr)
        // also ends the solve; the host reports it
        if (betan < st->resid || nx >= st->maxiter || st->aerr) {
            st->done = 1;
            st->halt = 1;
        }
    }
}

// p = r + beta p
template <class S, class SA>
__global__ __launch_bounds__(kBlock) void k_cg_pupdate(long n, typename S::st *__restrict__ p,
                                                       const typename SA::st *__restrict__ r,
                                                       const CgState *__restrict__ st) {
    if (st->halt) return;
    const long i = logical_block() * kBlock + threadIdx.x;
    typedef typename S::rt T;
    if (i < n) {
        const cx<typename SA::rt> rv = SA::ld(r, i);
        const cx<T> pv = S::ld(p, i);
        if constexpr (S::scaled) {
            const T cr = (T)st->xp_r, cp = (T)st->xp_p;
            S::sto(p, i, cx<T>{cr * (T)rv.r + cp * pv.r, cr * (T)rv.i + cp * pv.i});
        } else {
            const T b = (T)st->beta;
            S::sto(p, i, cx<T>{(T)rv.r + b * pv.r, (T)rv.i + b * pv.i});
        }
    }
}

// Reliable-update CG with two search-direction buffers used alternately
// (p_k in buffer k mod 2): iteration k forms p_{k+1} = r + beta p_k (scaled:
// r / sigma' + beta sigma / sigma' p_k) in the other buffer.  Even k defers
// its x increment (k_cg_pnext); odd k adds both, x += a_{k-1} p_{k-1} +
// a_k p_k, before p_{k+1} replaces p_{k-1} (k_cg_xpnext).  The accumulated x
// is read and written every second iteration instead of every one.
template <class S, class SA>
__device__ __forceinline__ cx<typename S::rt> cg_pnew(cx<typename SA::rt> rv, cx<typename S::rt> pv,
                                                      const CgState *__restrict__ st) {
    typedef typename S::rt T;
    if constexpr (S::scaled) {
        const T cr = (T)st->xp_r, cp = (T)st->xp_p;
        return cx<T>{cr * (T)rv.r + cp * pv.r, cr * (T)rv.i + cp * pv.i};
    } else {
        const T b = (T)st->beta;
        return cx<T>{(T)rv.r + b * pv.r, (T)rv.i + b * pv.i};
    }
}

template <class S, class SA>
__global__ __launch_bounds__(kBlock) void k_cg_pnext(long n, const typename S::st *__restrict__ pc,
                                                     typename S::st *__restrict__ po,
                                                     const typename SA::st *__restrict__ r,
                                                     const CgState *__restrict__ st) {
    if (st->halt) return;
    const long i = logical_block() * kBlock + threadIdx.x;
    if (i < n) S::sto(po, i, cg_pnew<S, SA>(SA::ld(r, i), S::ld(pc, i), st));
}

template <class S, class SA>
__global__ __launch_bounds__(kBlock) void k_cg_xpnext(long n, const typename S::st *__restrict__ pc,
                                                      typename S::st *__restrict__ po,
                                                      const typename SA::st *__restrict__ r,
                                                      typename SA::st *__restrict__ x,
                                                      const CgState *__restrict__ st) {
    if (st->halt) return;
    const long i = logical_block() * kBlock + threadIdx.x;
    typedef typename SA::rt TA;
    if (i < n) {
        const TA a0 = (TA)st->xp_a_prev, a1 = (TA)st->xp_a;
        const cx<typename S::rt> pv = S::ld(pc, i), pvo = S::ld(po, i);
        cx<TA> xv = SA::ld(x, i);
        xv = {xv.r + a0 * (TA)pvo.r, xv.i + a0 * (TA)pvo.i};
        xv = {xv.r + a1 * (TA)pv.r, xv.i + a1 * (TA)pv.i};
        SA::sto(x, i, xv);
        S::sto(po, i, cg_pnew<S, SA>(SA::ld(r, i), pv, st));
    }
}

// Re <r, p> over n complex numbers, r FP64 and p in storage S, one partial per
// workgroup (grid-stride): the component of the kept search direction along
// the new true residual at a reliable update.
template <class S>
__global__ __launch_bounds__(kBlock) void k_dot_rp(long n, const double2 *__restrict__ r,
                                                   const typename S::st *__restrict__ p,
                                                   double *__restrict__ partial) {
    const long lb = logical_block();
    const long stride = (long)gridDim.x * kBlock;
    double v = 0.0;
    for (long i = lb * kBlock + threadIdx.x; i < n; i += stride) {
        const double2 rv = r[i];
        const cx<typename S::rt> pv = S::ld(p, i);
        v += rv.x * (double)pv.r + rv.y * (double)pv.i;
    }
    const double t = block_sum(v);
    if (threadIdx.x == 0) partial[lb] = t;
}

// The direction after a reliable update: p_out = xp_r r + xp_p p_in (the host
// folds the removal of p's component along the new residual into xp_r).
template <class S, class SA>
__global__ __launch_bounds__(kBlock) void k_cg_prestart(long n, const typename S::st *__restrict__ pc,
                                                        typename S::st *__restrict__ po,
                                                        const typename SA::st *__restrict__ r,
                                                        const CgState *__restrict__ st) {
    const long i = logical_block() * kBlock + threadIdx.x;
    typedef typename S::rt T;
    if (i < n) {
        const T a = (T)st->xp_r, b = (T)st->xp_p;
        const cx<typename SA::rt> rv = SA::ld(r, i);
        const cx<T> pv = S::ld(pc, i);
        S::sto(po, i, cx<T>{a * (T)rv.r + b * pv.r, a * (T)rv.i + b * pv.i});
    }
}

// The increments still pending when the solve stops for the host after
// iteration k: a_k p_k (alpha of that step), and for odd k also a_{k-1}
// p_{k-1} (po) -- the update of that step did not run.
template <class S, class SA>
__global__ __launch_bounds__(kBlock) void k_cg_xflush(long n, const typename S::st *__restrict__ pc,
                                                      const typename S::st *__restrict__ po,
                                                      typename SA::st *__restrict__ x,
                                                      const CgState *__restrict__ st, int both) {
    const long i = logical_block() * kBlock + threadIdx.x;
    typedef typename SA::rt TA;
    if (i < n) {
        const TA a1 = (TA)(S::scaled ? st->alpha_s : st->alpha);
        cx<TA> xv = SA::ld(x, i);
        if (both) {
            const TA a0 = (TA)st->xp_a;
            const cx<typename S::rt> pvo = S::ld(po, i);
            xv = {xv.r + a0 * (TA)pvo.r, xv.i + a0 * (TA)pvo.i};
        }
        const cx<typename S::rt> pv = S::ld(pc, i);
        SA::sto(x, i, cx<TA>{xv.r + a1 * (TA)pv.r, xv.i + a1 * (TA)pv.i});
    }
}

// Reliable-update CG scalars after |r_s|^2 of the iterated residual (alpha =
// rnorm2 / |M p|^2 is formed inside the fused CG stencil, cgw_publish):
// either a plain step (beta) or a hold for the host (reliable update due,
// converged candidate, iteration cap, or a non-finite value).  The kernel
// reads the state together with the partials (one memory round trip instead
// of three: halt flag, partials, then the fields) and tests halt only before
// storing.
__global__ void k_rcg_beta(CgState *st, const double *__restrict__ partial, int np) {
    const int halt = st->halt, iter = st->iter + 1, maxiter = st->maxiter, scaled = st->scaled;
    const double rn_old = st->rnorm2, d2 = st->delta2, maxr2 = st->maxr2, stop2 = st->stop2;
    const double sig = st->sigma, alpha_s = st->alpha_s, nelem = st->nelem, xpa = st->xp_a, alpha = st->alpha;
    const double rn = sum_partials(partial, np);
    if (threadIdx.x == 0)
        __hip_atomic_store(&st->aslot, kAlphaEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x != 0 || halt) return;
    st->iter = iter;
    // hold for the host: reliable update due, converged candidate, iteration
    // cap, a non-finite value, or an iterated |r| 1000x above the last true
    // one (a diverging solve, which the host's stall test then ends)
    if (rn < d2 * maxr2 || rn < stop2 || iter >= maxiter || !(rn == rn) || rn > 1e6 * maxr2) {
        st->hold = 1;
        st->halt = 1;
        st->rn_hold = rn;
    } else {
        const double beta = rn / rn_old;
        st->beta = beta;
        st->rnorm2 = rn;
        st->xp_a_prev = xpa;
        if (scaled) {
            const double sig_new = sqrt(rn / nelem);
            st->xp_a = alpha_s;
            st->xp_r = 1.0 / sig_new;
            st->xp_p = beta * sig / sig_new;
            st->sigma = sig_new;
        } else {
            st->xp_a = alpha;
        }
    }
}

// x64 += xs (low precision accumulator), xs = 0
template <class SA>
__global__ __launch_bounds__(kBlock) void k_accum_x(long n, double2 *__restrict__ x,
                                                    typename SA::st *__restrict__ xs) {
    const long i = logical_block() * kBlock + threadIdx.x;
    if (i < n) {
        const cx<typename SA::rt> v = SA::ld(xs, i);
        double2 xv = x[i];
        xv.x += (double)v.r;
        xv.y += (double)v.i;
        x[i] = xv;
        SA::sto(xs, i, cx<typename SA::rt>{0, 0});
    }
}

// r64 = b - w (w = M^dagger M x), |r64|^2 partial, and rs = (SA) r64
template <class SA>
__global__ __launch_bounds__(kBlock) void k_true_resid(long n, const double2 *__restrict__ b,
                                                       const double2 *__restrict__ w, double2 *__restrict__ r,
                                                       typename SA::st *__restrict__ rs, double *__restrict__ partial) {
    const long lb = logical_block();
    const long i = lb * kBlock + threadIdx.x;
    double part = 0.0;
    if (i < n) {
        const double2 bv = b[i], wv = w[i];
        const double2 rv = make_double2(bv.x - wv.x, bv.y - wv.y);
        r[i] = rv;
        if (rs) SA::sto(rs, i, cx<typename SA::rt>{(typename SA::rt)rv.x, (typename SA::rt)rv.y});
        part = rv.x * rv.x + rv.y * rv.y;
    }
    const double t = block_sum(part);
    if (threadIdx.x == 0) partial[lb] = t;
}

}  // namespace chs
