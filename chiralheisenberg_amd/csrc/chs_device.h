// chs_device.h -- gfx950 device code of libchs: the domain-wall operator
// (rhmc/cheisdwf.f:2634-3501), its fused solver epilogues, the multi-shift
// Lanczos vector updates (rhmc/cheisdwf.f:655-695), CG vector updates
// (rhmc/cheisdwf.f:1098-1148), the wall-only force (rhmc/cheisdwf.f:733-849)
// and deterministic FP64 reductions.
//
// Data layout in HBM: the reference layout itself.  A spinor is 8 planes
// (Dirac d = 0..3, flavour f = 0..1, plane index d + 4 f), each plane holding
// Ls*kvol complex numbers with s fastest, then x, y, t.  That is a
// structure-of-arrays over the 8 spin-flavour components with the fifth
// dimension contiguous: a wavefront of 64 consecutive lanes (s, x) streams
// 64 consecutive complex numbers of every plane, and the x, y, t and s
// neighbours of a lane are fixed offsets (Ls, Ls*L, Ls*L*L, 1) into the same
// planes, so neighbour loads hit the lines the neighbouring lanes and
// workgroups just brought into L1/L2.
//
// Storage precisions: FP64 (double2), FP32 (float2), BF16 (re, im packed in
// one 32-bit word).  FP64 storage computes in FP64; FP32 and BF16 storage
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
    static __device__ __forceinline__ cx<double> ld(const st *p, long i) {
        double2 v = p[i];
        return {v.x, v.y};
    }
    static __device__ __forceinline__ void sto(st *p, long i, cx<double> v) { p[i] = make_double2(v.r, v.i); }
};
struct F32 {
    typedef float2 st;
    typedef float rt;
    static __device__ __forceinline__ cx<float> ld(const st *p, long i) {
        float2 v = p[i];
        return {v.x, v.y};
    }
    static __device__ __forceinline__ void sto(st *p, long i, cx<float> v) { p[i] = make_float2(v.r, v.i); }
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
    static __device__ __forceinline__ cx<float> ld(const st *p, long i) {
        uint32_t v = p[i];
        return {__uint_as_float(v << 16), __uint_as_float(v & 0xffff0000u)};
    }
    static __device__ __forceinline__ void sto(st *p, long i, cx<float> v) { p[i] = f2bf(v.r) | (f2bf(v.i) << 16); }
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

// ---- the operator at one 5D site ------------------------------------------
// Adds the hop  -1/2 bc (1 - e g_mu) psi(nb)  to o, e = +1 for the M forward
// hop; the backward hop and M^dagger flip e.  The projector keeps two
// independent combinations a, b (rhmc/cheisdwf.f:2671-2828 for M,
// :3108-3265 for M^dagger).
template <class S, int MU>
__device__ __forceinline__ void hop(cx<typename S::rt> (&o)[4][2], const typename S::st *__restrict__ in,
                                    long gn, long P, typename S::rt e, typename S::rt coef) {
    typedef typename S::rt T;
#pragma unroll
    for (int f = 0; f < 2; f++) {
        const cx<T> p0 = S::ld(in, gn + P * (0 + 4 * f));
        const cx<T> p1 = S::ld(in, gn + P * (1 + 4 * f));
        const cx<T> p2 = S::ld(in, gn + P * (2 + 4 * f));
        const cx<T> p3 = S::ld(in, gn + P * (3 + 4 * f));
        cx<T> a, b, h2, h3;
        if (MU == 0) {          // g1: a = p0 + e i p3, b = p1 + e i p2 -> (a, b, -e i b, -e i a)
            a = cadd(p0, cmuli(e, p3));
            b = cadd(p1, cmuli(e, p2));
            h2 = cmuli(-e, b);
            h3 = cmuli(-e, a);
        } else if (MU == 1) {   // g2: a = p0 + e p3, b = p1 - e p2 -> (a, b, -e b, e a)
            a = cadd(p0, cscale(e, p3));
            b = csub(p1, cscale(e, p2));
            h2 = cscale(-e, b);
            h3 = cscale(e, a);
        } else {                // g3: a = p0 + e i p2, b = p1 - e i p3 -> (a, b, -e i a, e i b)
            a = cadd(p0, cmuli(e, p2));
            b = csub(p1, cmuli(e, p3));
            h2 = cmuli(-e, a);
            h3 = cmuli(e, b);
        }
        o[0][f] = cadd(o[0][f], cscale(coef, a));
        o[1][f] = cadd(o[1][f], cscale(coef, b));
        o[2][f] = cadd(o[2][f], cscale(coef, h2));
        o[3][f] = cadd(o[3][f], cscale(coef, h3));
    }
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

// o = (M in)(s, site) (DAG = false) or (M^dagger in)(s, site); cen receives
// the centre values in[s, site] (used by fused epilogues).
template <class S, bool DAG>
__device__ __forceinline__ void site_apply(const Lat &lat, const typename S::st *__restrict__ in,
                                           const double *__restrict__ chi, double hpin, int imass, int s,
                                           int site, cx<typename S::rt> (&o)[4][2],
                                           cx<typename S::rt> (&cen)[4][2]) {
    typedef typename S::rt T;
    const int L = lat.L, Ls = lat.Ls;
    const long P = lat.n5;
    const int ix = site % L;
    const int r1 = site / L;
    const int iy = r1 % L;
    const int it = r1 / L;
    const long g = (long)site * Ls + s;
    const T dg = (T)lat.diag;
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
        for (int d = 0; d < 4; d++) {
            cen[d][f] = S::ld(in, g + P * (d + 4 * f));
            o[d][f] = cscale(dg, cen[d][f]);
        }
    const T e = DAG ? T(-1) : T(1);
    const T mh = T(-0.5);
    // x
    {
        const long gf = g + (long)Ls * ((ix + 1 == L) ? 1 - L : 1);
        const long gb = g + (long)Ls * ((ix == 0) ? L - 1 : -1);
        hop<S, 0>(o, in, gf, P, e, mh);
        hop<S, 0>(o, in, gb, P, -e, mh);
    }
    // y
    {
        const long sy = (long)Ls * L;
        const long gf = g + sy * ((iy + 1 == L) ? 1 - L : 1);
        const long gb = g + sy * ((iy == 0) ? L - 1 : -1);
        hop<S, 1>(o, in, gf, P, e, mh);
        hop<S, 1>(o, in, gb, P, -e, mh);
    }
    // t, antiperiodic across it = Lt-1 <-> 0 when apbc
    {
        const long st = (long)Ls * L * L;
        const int Lt = lat.Lt;
        const long gf = g + st * ((it + 1 == Lt) ? 1 - Lt : 1);
        const long gb = g + st * ((it == 0) ? Lt - 1 : -1);
        const T cf = (lat.apbc && it + 1 == Lt) ? T(0.5) : mh;
        const T cb = (lat.apbc && it == 0) ? T(0.5) : mh;
        hop<S, 2>(o, in, gf, P, e, cf);
        hop<S, 2>(o, in, gb, P, -e, cb);
    }
    // fifth dimension: M takes P- psi(s+1) and P+ psi(s-1); M^dagger the
    // reverse (rhmc/cheisdwf.f:2879-2908, :3313-3342).  At s = Ls-1 (s = 0)
    // the missing hop is replaced by the wall block acting on psi(0)
    // (psi(Ls-1)).
    const int dn = DAG ? 0 : 2;   // first component taking s+1
    const int dp = DAG ? 2 : 0;   // first component taking s-1
    const bool wall_ok = (imass == 1 || imass == 3);
    T c1 = 0, c2 = 0, c3 = 0;
    if ((s == 0 || s == Ls - 1) && wall_ok) {
        c1 = (T)chi[site];
        c2 = (T)chi[site + lat.kvol];
        double c3d = chi[site + 2L * lat.kvol];
        if (imass == 1 && site == 0) c3d += hpin;
        c3 = (T)c3d;
    }
    {
        const long gn = g + ((s + 1 == Ls) ? 1 - Ls : 1);
#pragma unroll
        for (int dd = 0; dd < 2; dd++) {
            const cx<T> v0 = S::ld(in, gn + P * (dn + dd));
            const cx<T> v1 = S::ld(in, gn + P * (dn + dd + 4));
            if (s + 1 < Ls) {
                o[dn + dd][0] = csub(o[dn + dd][0], v0);
                o[dn + dd][1] = csub(o[dn + dd][1], v1);
            } else if (wall_ok) {
                wall_apply<T>(o[dn + dd][0], o[dn + dd][1], v0, v1, c1, c2, c3, imass, true);
            }
        }
    }
    {
        const long gp = g + ((s == 0) ? Ls - 1 : -1);
#pragma unroll
        for (int dd = 0; dd < 2; dd++) {
            const cx<T> v0 = S::ld(in, gp + P * (dp + dd));
            const cx<T> v1 = S::ld(in, gp + P * (dp + dd + 4));
            if (s > 0) {
                o[dp + dd][0] = csub(o[dp + dd][0], v0);
                o[dp + dd][1] = csub(o[dp + dd][1], v1);
            } else if (wall_ok) {
                wall_apply<T>(o[dp + dd][0], o[dp + dd][1], v0, v1, c1, c2, c3, imass, false);
            }
        }
    }
}

// ---- fused dslash kernels ---------------------------------------------------
enum Epi {
    EPI_STORE = 0,   // out = M in
    EPI_NORM = 1,    // out = M in, partial += |out|^2
    EPI_SCALEQ = 2,  // in = r: out = (M r)/beta, aux0 = q = r/beta (Lanczos, scal[0] = beta)
    EPI_DOT = 3,     // out = M in, partial += Re <aux0, out>
    EPI_CG = 4       // no store: aux1 (x) += alpha aux0 (p), aux2 (r) -= alpha (M in),
                     // partial += |r|^2; alpha = scal[0]; flag bit0: skip x update
};

struct EpiArgs {
    double *partial;
    const double *scal;
    void *aux0, *aux1, *aux2;
    const int *skip;   // if not NULL and *skip != 0 the launch does nothing
    int flag;
};

template <class S, class SA, bool DAG, int EPI>
__global__ __launch_bounds__(kBlock) void k_dslash(Lat lat, typename S::st *__restrict__ out,
                                                   const typename S::st *__restrict__ in,
                                                   const double *__restrict__ chi, double hpin, int imass,
                                                   EpiArgs ea) {
    typedef typename S::rt T;
    if (ea.skip && *ea.skip) return;
    const long lb = logical_block();
    const long g = lb * kBlock + threadIdx.x;
    double part = 0.0;
    if (g < lat.n5) {
        const int site = (int)(g / lat.Ls);
        const int s = (int)(g - (long)site * lat.Ls);
        cx<T> o[4][2], c[4][2];
        site_apply<S, DAG>(lat, in, chi, hpin, imass, s, site, o, c);
        const long P = lat.n5;
        if (EPI == EPI_SCALEQ) {
            const T ib = (T)(1.0 / ea.scal[0]);
            typename S::st *q = (typename S::st *)ea.aux0;
#pragma unroll
            for (int f = 0; f < 2; f++)
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    S::sto(out, g + P * (d + 4 * f), cscale(ib, o[d][f]));
                    S::sto(q, g + P * (d + 4 * f), cscale(ib, c[d][f]));
                }
        } else if (EPI == EPI_CG) {
            const T alpha = (T)ea.scal[0];
            const typename S::st *p = (const typename S::st *)ea.aux0;
            typename SA::st *x = (typename SA::st *)ea.aux1;
            typename SA::st *r = (typename SA::st *)ea.aux2;
            T acc = 0;
#pragma unroll
            for (int f = 0; f < 2; f++)
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    const long i = g + P * (d + 4 * f);
                    if (!(ea.flag & 1)) {
                        cx<T> xv = SA::ld(x, i);
                        const cx<T> pv = S::ld(p, i);
                        xv = cadd(xv, cscale(alpha, pv));
                        SA::sto(x, i, xv);
                    }
                    cx<T> rv = SA::ld(r, i);
                    rv = csub(rv, cscale(alpha, o[d][f]));
                    SA::sto(r, i, rv);
                    acc += rv.r * rv.r + rv.i * rv.i;
                }
            part = (double)acc;
        } else {
            T acc = 0;
            const typename S::st *a0 = (const typename S::st *)ea.aux0;
#pragma unroll
            for (int f = 0; f < 2; f++)
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    S::sto(out, g + P * (d + 4 * f), o[d][f]);
                    if (EPI == EPI_NORM) acc += o[d][f].r * o[d][f].r + o[d][f].i * o[d][f].i;
                    if (EPI == EPI_DOT) {
                        const cx<T> av = S::ld(a0, g + P * (d + 4 * f));
                        acc += av.r * o[d][f].r + av.i * o[d][f].i;
                    }
                }
            part = (double)acc;
        }
    }
    if (EPI == EPI_NORM || EPI == EPI_DOT || EPI == EPI_CG) {
        const double t = block_sum(part);
        if (threadIdx.x == 0) ea.partial[lb] = t;
    }
}

// ---- deterministic FP64 reductions -----------------------------------------
// One workgroup sums n partials in a fixed order.
__device__ __forceinline__ double sum_partials(const double *__restrict__ part, int n) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) v += part[i];
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

// ---- multi-shift Lanczos (rhmc/cheisdwf.f:599-745) -----------------------
struct LzState {
    double beta;        // |r| of the current Lanczos vector (beta0 of the next step)
    double alphatild;
    double phimod, resid;
    int done, itercg, ndiag, pad;
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
__global__ void k_lz_alpha(LzState *st, const double *__restrict__ partial, int np, int iter) {
    if (st->done) return;
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

// new beta = |r|; convergence max_k |phimod rho_k| < 2 resid for iter >= 2
// (rhmc/cheisdwf.f:697-715).
__global__ void k_lz_beta(LzState *st, const double *__restrict__ partial, int np, int iter) {
    if (st->done) return;
    const double t = sum_partials(partial, np);
    if (threadIdx.x == 0) {
        st->beta = sqrt(t);
        st->itercg = iter;
        if (iter >= 2) {
            double rhomax = 0.0;
            for (int k = 0; k < st->ndiag; k++) {
                const double a = fabs(st->phimod * st->rho[k]);
                if (k == 0 || a > rhomax) rhomax = a;
            }
            if (rhomax < 2.0 * st->resid) st->done = 1;
        }
    }
}

// r = w - alphatild q - beta0 qm1 (|r|^2 partial) and, for every shift,
// p_k = q - mu_k p_k, x_k += rho_k p_k (iter 1: p_k = q, x_k = rho_k q).
// p and x are shift-major: shift k at offset k*n.
template <int NS>
__global__ __launch_bounds__(kBlock) void k_lz_update(long n, const double2 *__restrict__ w,
                                                      const double2 *__restrict__ q,
                                                      const double2 *__restrict__ qm1, double2 *__restrict__ r,
                                                      double2 *__restrict__ p, double2 *__restrict__ x,
                                                      const LzState *__restrict__ st, int ndiag, int iter,
                                                      double *__restrict__ partial) {
    if (st->done) return;
    __shared__ double s_amu[kMaxShift], s_rho[kMaxShift];
    if (threadIdx.x < ndiag) {
        s_amu[threadIdx.x] = st->amu[threadIdx.x];
        s_rho[threadIdx.x] = st->rho[threadIdx.x];
    }
    __syncthreads();
    const long lb = logical_block();
    const long i = lb * kBlock + threadIdx.x;
    double part = 0.0;
    if (i < n) {
        const double at = st->alphatild, b0 = st->beta;
        const double2 qv = q[i], wv = w[i];
        const double2 qm = qm1[i];
        double2 rv;
        rv.x = wv.x - at * qv.x - b0 * qm.x;
        rv.y = wv.y - at * qv.y - b0 * qm.y;
        r[i] = rv;
        part = rv.x * rv.x + rv.y * rv.y;
        const int ns = NS > 0 ? NS : ndiag;
        if (iter == 1) {
#pragma unroll 4
            for (int k = 0; k < ns; k++) {
                p[(long)k * n + i] = qv;
                x[(long)k * n + i] = make_double2(s_rho[k] * qv.x, s_rho[k] * qv.y);
            }
        } else {
#pragma unroll 4
            for (int k = 0; k < ns; k++) {
                double2 pv = p[(long)k * n + i];
                double2 xv = x[(long)k * n + i];
                pv.x = qv.x - s_amu[k] * pv.x;
                pv.y = qv.y - s_amu[k] * pv.y;
                xv.x += s_rho[k] * pv.x;
                xv.y += s_rho[k] * pv.y;
                p[(long)k * n + i] = pv;
                x[(long)k * n + i] = xv;
            }
        }
    }
    const double t = block_sum(part);
    if (threadIdx.x == 0) partial[lb] = t;
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

// One thread per (site, shift): contribution of shift k to the force at the
// site, from y = x_k and z = M x_k on the two wall slices only, into
// contrib[(k*3 + a)*kvol + site].
__global__ __launch_bounds__(kBlock) void k_force_shift(Lat lat, const double2 *__restrict__ xk, long n,
                                                        int ndiag, const double *__restrict__ anum,
                                                        const double *__restrict__ chi, double hpin, int imass,
                                                        double *__restrict__ contrib) {
    const long gid = logical_block() * kBlock + threadIdx.x;
    if (gid >= (long)lat.kvol * ndiag) return;
    const int k = (int)(gid / lat.kvol);
    const int site = (int)(gid - (long)k * lat.kvol);
    const double2 *y = xk + (long)k * n;
    cx<double> zf[4][2], zl[4][2], cf[4][2], cl[4][2];
    site_apply<F64, false>(lat, y, chi, hpin, imass, 0, site, zf, cf);
    site_apply<F64, false>(lat, y, chi, hpin, imass, lat.Ls - 1, site, zl, cl);
    double2 ya[4][2], za[4][2];
#pragma unroll
    for (int f = 0; f < 2; f++) {
#pragma unroll
        for (int d = 0; d < 2; d++) {   // y(s=Ls) with z(s=1)
            ya[d][f] = make_double2(cl[d][f].r, cl[d][f].i);
            za[d][f] = make_double2(zf[d][f].r, zf[d][f].i);
        }
#pragma unroll
        for (int d = 2; d < 4; d++) {   // y(s=1) with z(s=Ls)
            ya[d][f] = make_double2(cf[d][f].r, cf[d][f].i);
            za[d][f] = make_double2(zl[d][f].r, zl[d][f].i);
        }
    }
    double F1 = 0, F2 = 0, F3 = 0;
    derivs_site(F1, F2, F3, ya, za, -2.0 * anum[k + 1], imass);
    contrib[((long)k * 3 + 0) * lat.kvol + site] = F1;
    contrib[((long)k * 3 + 1) * lat.kvol + site] = F2;
    contrib[((long)k * 3 + 2) * lat.kvol + site] = F3;
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

// derivs on given y, z (both full spinors), one thread per site
__global__ __launch_bounds__(kBlock) void k_derivs(Lat lat, const double2 *__restrict__ y,
                                                   const double2 *__restrict__ z, double anum, int imass,
                                                   double *__restrict__ F) {
    const long site = logical_block() * kBlock + threadIdx.x;
    if (site >= lat.kvol) return;
    const long P = lat.n5;
    const long g0 = site * lat.Ls, gl = g0 + lat.Ls - 1;
    double2 ya[4][2], za[4][2];
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const long gy = d < 2 ? gl : g0, gz = d < 2 ? g0 : gl;
            ya[d][f] = y[gy + P * (d + 4 * f)];
            za[d][f] = z[gz + P * (d + 4 * f)];
        }
    double F1 = F[site], F2 = F[site + lat.kvol], F3 = F[site + 2L * lat.kvol];
    derivs_site(F1, F2, F3, ya, za, -2.0 * anum, imass);
    F[site] = F1;
    F[site + lat.kvol] = F2;
    F[site + 2L * lat.kvol] = F3;
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
    int pad;
};

// reference CGNR: alpha = alphan / |M p|^2
__global__ void k_cg_alpha(CgState *st, const double *__restrict__ partial, int np) {
    if (st->halt) return;
    const double t = sum_partials(partial, np);
    if (threadIdx.x == 0) st->alpha = st->alphan / t;
}

// reference CGNR after |r|^2 (rhmc/cheisdwf.f:1130-1148).  The reference
// updates p before its convergence test; p is not used after convergence, so
// halting before the p update leaves x and the iteration count unchanged.
__global__ void k_cg_beta_ref(CgState *st, const double *__restrict__ partial, int np, int nx) {
    if (st->halt) return;
    const double betan = sum_partials(partial, np);
    if (threadIdx.x == 0) {
        double beta = betan / st->betad;
        st->betad = betan;
        st->alphan = betan;
        if (nx == 1) beta = 0.0;
        st->beta = beta;
        st->iter = nx;
        st->rnorm2 = betan;
        if (betan < st->resid || nx >= st->maxiter) {
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
        const T b = (T)st->beta;
        const cx<typename SA::rt> rv = SA::ld(r, i);
        const cx<T> pv = S::ld(p, i);
        S::sto(p, i, cx<T>{(T)rv.r + b * pv.r, (T)rv.i + b * pv.i});
    }
}

// reliable-update CG: alpha = rnorm2 / |M p|^2
__global__ void k_rcg_alpha(CgState *st, const double *__restrict__ partial, int np) {
    if (st->halt) return;
    const double t = sum_partials(partial, np);
    if (threadIdx.x == 0) st->alpha = st->rnorm2 / t;
}

// reliable-update CG after |r_s|^2 of the iterated residual: a plain step
// (beta), or a hold for the host (reliable update due, converged candidate,
// iteration cap, or a non-finite value).
__global__ void k_rcg_beta(CgState *st, const double *__restrict__ partial, int np) {
    if (st->halt) return;
    const double rn = sum_partials(partial, np);
    if (threadIdx.x == 0) {
        st->iter += 1;
        if (rn < st->delta2 * st->maxr2 || rn < st->stop2 || st->iter >= st->maxiter || !(rn == rn)) {
            st->hold = 1;
            st->halt = 1;
            st->rn_hold = rn;
        } else {
            st->beta = rn / st->rnorm2;
            st->rnorm2 = rn;
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
