// chs_stencil.h -- the production domain-wall stencil of libchs (M and
// M^dagger of rhmc/cheisdwf.f:2634-3501) with fused solver epilogues.
//
// Device layout ("native"): a spinor is 8 spin-flavour planes (d + 4 f), each
// plane Ls three-dimensional volumes: element s*kvol + site, site = x + L y +
// L^2 t.  The reference layout keeps s fastest (site*Ls + s); the native one
// keeps s outermost inside a plane, so that a run of consecutive workgroups
// covers whole 3-D s-slices.  Workgroups go to the 8 XCDs in contiguous
// ranges (logical_block_x), so one XCD owns a few complete s-slices of one
// flavour: every x, y, t neighbour it reads is its own data, L2-resident
// within a distance of L^2 elements, and only the two s-slices at the edge of
// its range (read by the s-hop) and the walls are shared with another XCD.
// With s fastest, an XCD owned a thin t-slab instead and re-fetched the slabs
// on either side (measured: 2.0x the input bytes from beyond L2 in FP64 at
// L=24, Ls=32).  The host boundary converts layouts once per call.
//
// Threads: one per (s, site, flavour) (NF = 1) with the flavour constant per
// workgroup, so that every plane base is wave-uniform (scalar registers) and
// a lane only carries a 32-bit element offset (contexts reject lattices with
// 8*Ls*kvol >= 2^31).  Output stores are non-temporal: nothing re-reads the
// output in the same launch, and keeping it out of L2 leaves L2 to the
// neighbour reuse of the input.
#pragma once
#include "chs_device.h"

namespace chs {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// XCD-contiguous logical block along x (see logical_block)
__device__ __forceinline__ unsigned logical_block_x() {
    const unsigned b = blockIdx.x, nb = gridDim.x;
    if ((nb & 7u) == 0u) return (b & 7u) * (nb >> 3) + (b >> 3);
    return b;
}

template <class S>
__device__ __forceinline__ void store_nt(typename S::st *__restrict__ p, unsigned i, cx<typename S::rt> v) {
    typename S::st w;
    S::sto(&w, 0, v);
    if constexpr (sizeof(w) == 16) {
        u32x4 raw;
        __builtin_memcpy(&raw, &w, 16);
        __builtin_nontemporal_store(raw, reinterpret_cast<u32x4 *>(p + i));
    } else if constexpr (sizeof(w) == 8) {
        u32x2 raw;
        __builtin_memcpy(&raw, &w, 8);
        __builtin_nontemporal_store(raw, reinterpret_cast<u32x2 *>(p + i));
    } else {
        unsigned raw;
        __builtin_memcpy(&raw, &w, 4);
        __builtin_nontemporal_store(raw, reinterpret_cast<unsigned *>(p + i));
    }
}

// Plain streaming copy of n elements with the stencil's non-temporal stores:
// the reference rate for the stencil's HBM figures (chs_time_copy).
template <class S>
__global__ __launch_bounds__(kBlock) void k_copy_nt(long n, typename S::st *__restrict__ dst,
                                                    const typename S::st *__restrict__ src) {
    const long i = logical_block() * kBlock + threadIdx.x;
    if (i < n) store_nt<S>(dst, (unsigned)i, S::ld(src, i));
}

// The same copy on a persistent grid: each workgroup strides over the array
// (the streaming ceiling with a fixed number of resident workgroups).
template <class S>
__global__ __launch_bounds__(kBlock) void k_copy_nt_grid(long n, typename S::st *__restrict__ dst,
                                                         const typename S::st *__restrict__ src) {
    const long stride = (long)gridDim.x * kBlock;
    for (long i = logical_block() * kBlock + threadIdx.x; i < n; i += stride)
        store_nt<S>(dst, (unsigned)i, S::ld(src, i));
}

// o += coef * (1 - e g_mu) (p0, p1, p2, p3).  MU = 0, 1, 2 (x, y, t) with
// e = +1; MU = -1, -2, -3 the same with e = -1.
template <class T, int MU>
__device__ __forceinline__ void hop_vals(cx<T> (&o)[4], cx<T> p0, cx<T> p1, cx<T> p2, cx<T> p3, T coef) {
    constexpr int mu = MU < 0 ? -MU - 1 : MU;
    constexpr T e = MU < 0 ? T(-1) : T(1);
    cx<T> a, b, h2, h3;
    if (mu == 0) {          // g1: (a, b, -e i b, -e i a), a = p0 + e i p3, b = p1 + e i p2
        a = cadd(p0, cmuli(e, p3));
        b = cadd(p1, cmuli(e, p2));
        h2 = cmuli(-e, b);
        h3 = cmuli(-e, a);
    } else if (mu == 1) {   // g2: (a, b, -e b, e a), a = p0 + e p3, b = p1 - e p2
        a = cadd(p0, cscale(e, p3));
        b = csub(p1, cscale(e, p2));
        h2 = cscale(-e, b);
        h3 = cscale(e, a);
    } else {                // g3: (a, b, -e i a, e i b), a = p0 + e i p2, b = p1 - e i p3
        a = cadd(p0, cmuli(e, p2));
        b = csub(p1, cmuli(e, p3));
        h2 = cmuli(-e, a);
        h3 = cmuli(e, b);
    }
    o[0] = {o[0].r + coef * a.r, o[0].i + coef * a.i};
    o[1] = {o[1].r + coef * b.r, o[1].i + coef * b.i};
    o[2] = {o[2].r + coef * h2.r, o[2].i + coef * h2.i};
    o[3] = {o[3].r + coef * h3.r, o[3].i + coef * h3.i};
}

// |a|^2 and Re <a, b> formed in FP64 whatever the compute type: the epilogue
// partials of FP32 / 16-bit storage would otherwise under- or overflow for
// components beyond ~1e-19 .. 1e19 (squares outside FP32's range).
template <class U>
__device__ __forceinline__ double norm2_d(cx<U> a) {
    return (double)a.r * (double)a.r + (double)a.i * (double)a.i;
}
template <class U>
__device__ __forceinline__ double dotre_d(cx<U> a, cx<U> b) {
    return (double)a.r * (double)b.r + (double)a.i * (double)b.i;
}

// o[ff] += coef * (1 - e g_mu) psi_ff(gn); pl[ff][d] are uniform plane bases.
template <class S, int MU, int NF>
__device__ __forceinline__ void hop2(cx<typename S::rt> (&o)[NF][4], const typename S::st *const (&pl)[NF][4],
                                     unsigned gn, typename S::rt coef) {
    typedef typename S::rt T;
#pragma unroll
    for (int ff = 0; ff < NF; ff++)
        hop_vals<T, MU>(o[ff], S::ld(pl[ff][0], gn), S::ld(pl[ff][1], gn), S::ld(pl[ff][2], gn),
                        S::ld(pl[ff][3], gn), coef);
}

// the same hop for the two sites of a pair: neighbours at gn, gn + 1 (gn even)
template <class S, int MU>
__device__ __forceinline__ void hop_pair(cx<typename S::rt> (&o0)[4], cx<typename S::rt> (&o1)[4],
                                         const typename S::st *const (&pl)[4], unsigned gn, typename S::rt coef) {
    typedef typename S::rt T;
    cx<T> a[4], b[4];
#pragma unroll
    for (int d = 0; d < 4; d++) S::ld2(pl[d], gn, a[d], b[d]);
    hop_vals<T, MU>(o0, a[0], a[1], a[2], a[3], coef);
    hop_vals<T, MU>(o1, b[0], b[1], b[2], b[3], coef);
}

// Element offset of (s, site) inside a plane.
template <int LAYOUT>
__device__ __forceinline__ unsigned elem(unsigned s, unsigned site, unsigned Ls, unsigned kvol) {
    return LAYOUT ? s * kvol + site : site * Ls + s;
}

// (M in)(s, site) or (M^dagger in)(s, site) for the flavours fb..fb+NF-1.
// in: base of the whole spinor; c receives the centre values.
template <class S, bool DAG, int NF, int LAYOUT>
__device__ __forceinline__ void stencil(const Lat &lat, const typename S::st *__restrict__ in,
                                        const double *__restrict__ chi, double hpin, int imass, unsigned s,
                                        unsigned site, int fb, cx<typename S::rt> (&o)[NF][4],
                                        cx<typename S::rt> (&c)[NF][4]) {
    typedef typename S::rt T;
    typedef typename S::st st;
    const unsigned Ls = (unsigned)lat.Ls, L = (unsigned)lat.L, Lt = (unsigned)lat.Lt, kvol = (unsigned)lat.kvol;
    const size_t P = (size_t)lat.ps;
    const st *pl[NF][4];
#pragma unroll
    for (int ff = 0; ff < NF; ff++)
#pragma unroll
        for (int d = 0; d < 4; d++) pl[ff][d] = in + P * (size_t)(d + 4 * (fb + ff));
    const unsigned ix = site % L;
    const unsigned r1 = site / L;
    const unsigned iy = r1 % L;
    const unsigned it = r1 / L;
    const unsigned g = elem<LAYOUT>(s, site, Ls, kvol);
    const T dg = (T)lat.diag;
#pragma unroll
    for (int ff = 0; ff < NF; ff++)
#pragma unroll
        for (int d = 0; d < 4; d++) {
            c[ff][d] = S::ld(pl[ff][d], g);
            o[ff][d] = cscale(dg, c[ff][d]);
        }
    const T mh = T(-0.5);
    // M: forward (1 - g), backward (1 + g); M^dagger the reverse
    constexpr int XF = DAG ? -1 : 0, XB = DAG ? 0 : -1;
    constexpr int YF = DAG ? -2 : 1, YB = DAG ? 1 : -2;
    constexpr int TF = DAG ? -3 : 2, TB = DAG ? 2 : -3;
    const unsigned sx = LAYOUT ? 1u : Ls;
    const unsigned sy = sx * L, stt = sy * L;
    const unsigned ss = LAYOUT ? kvol : 1u;
    hop2<S, XF, NF>(o, pl, (ix + 1 == L) ? g - (L - 1) * sx : g + sx, mh);
    hop2<S, XB, NF>(o, pl, (ix == 0) ? g + (L - 1) * sx : g - sx, mh);
    hop2<S, YF, NF>(o, pl, (iy + 1 == L) ? g - (L - 1) * sy : g + sy, mh);
    hop2<S, YB, NF>(o, pl, (iy == 0) ? g + (L - 1) * sy : g - sy, mh);
    // antiperiodic in t across it = Lt-1 <-> 0 (rhmc/cheisdwf.f:2831-2876)
    hop2<S, TF, NF>(o, pl, (it + 1 == Lt) ? g - (Lt - 1) * stt : g + stt,
                    (lat.apbc && it + 1 == Lt) ? T(0.5) : mh);
    hop2<S, TB, NF>(o, pl, (it == 0) ? g + (Lt - 1) * stt : g - stt, (lat.apbc && it == 0) ? T(0.5) : mh);
    // fifth dimension: M takes P- psi(s+1), P+ psi(s-1); M^dagger the reverse
    // (rhmc/cheisdwf.f:2879-2908, :3313-3342).  At s = Ls-1 (s = 0) the
    // missing hop is replaced by the wall block acting on psi(0) (psi(Ls-1)),
    // both flavours (rhmc/cheisdwf.f:2912-3066, :3345-3498).
    constexpr int dn = DAG ? 0 : 2;
    constexpr int dp = DAG ? 2 : 0;
    const bool wall_ok = (imass == 1 || imass == 3);
    const bool last = (s + 1 == Ls), first = (s == 0);
    // The s-neighbour loads are unconditional (clamped to the centre at the
    // walls, where the value is then discarded by a select): inside a branch
    // each would be a memory round trip of its own after the x, y, t loads.
    // Subtracting the selected +0 leaves o bit-identical.
    {
        const unsigned gsp = last ? g : g + ss, gsm = first ? g : g - ss;
        const cx<T> zero = {T(0), T(0)};
#pragma unroll
        for (int ff = 0; ff < NF; ff++)
#pragma unroll
            for (int dd = 0; dd < 2; dd++) {
                const cx<T> up = S::ld(pl[ff][dn + dd], gsp), dw = S::ld(pl[ff][dp + dd], gsm);
                o[ff][dn + dd] = csub(o[ff][dn + dd], last ? zero : up);
                o[ff][dp + dd] = csub(o[ff][dp + dd], first ? zero : dw);
            }
    }
    if ((first || last) && wall_ok) {
        const st *q0 = in, *q1 = in + 4 * P;   // flavour 0 / 1 planes at d = 0
        const unsigned src = last ? elem<LAYOUT>(0, site, Ls, kvol) : elem<LAYOUT>(Ls - 1, site, Ls, kvol);
        // all wall operands are loaded before any is used (one round trip)
        cx<T> v[2][2];
#pragma unroll
        for (int dd = 0; dd < 2; dd++) {
            const int d = last ? dn + dd : dp + dd;
            v[dd][0] = S::ld(q0 + P * (size_t)d, src);
            v[dd][1] = S::ld(q1 + P * (size_t)d, src);
        }
        const T c1 = (T)chi[site], c2 = (T)chi[site + kvol];
        double c3d = chi[site + 2u * kvol];
        if (imass == 1 && site == 0) c3d += hpin;
        const T c3 = (T)c3d;
#pragma unroll
        for (int dd = 0; dd < 2; dd++) {
            const cx<T> v0 = v[dd][0], v1 = v[dd][1];
            cx<T> w0 = {0, 0}, w1 = {0, 0};
            wall_apply<T>(w0, w1, v0, v1, c1, c2, c3, imass, last);
#pragma unroll
            for (int ff = 0; ff < NF; ff++) {
                const cx<T> w = (fb + ff) ? w1 : w0;
                if (last) o[ff][dn + dd] = cadd(o[ff][dn + dd], w);
                else o[ff][dp + dd] = cadd(o[ff][dp + dd], w);
            }
        }
    }
}

// Workgroups per flavour for a stencil launch over n5 = Ls*kvol elements.
__host__ __device__ inline unsigned stencil_blocks_per_flavour(long n5) {
    return (unsigned)((n5 + kBlock - 1) / kBlock);
}

// out = M in / M^dagger in with the epilogues of chs::Epi.  1-D grid of
// 2 * blocks_per_flavour workgroups rounded up to a multiple of 8; logical
// block lb covers flavour lb / bpf, elements (lb % bpf) * kBlock + tid.
// Partials: one per workgroup at partial[logical block].  dsl_block is the
// work of one logical block; k_dsl launches it for one lattice and the
// grouped kernels of chs_batch.h for many.
template <class S, class SA, bool DAG, int EPI, int LAYOUT>
__device__ __forceinline__ void dsl_block(const Lat &lat, typename S::st *__restrict__ out,
                                          const typename S::st *__restrict__ in, const double *__restrict__ chi,
                                          double hpin, int imass, const EpiArgs &ea, unsigned lb) {
    typedef typename S::rt T;
    typedef typename SA::rt TA;
    const unsigned bpf = stencil_blocks_per_flavour(lat.n5);
    const int fb = (int)(lb / bpf);
    const unsigned g = (lb - (unsigned)fb * bpf) * kBlock + threadIdx.x;   // position in (s, site) order
    const size_t P = (size_t)lat.ps;
    double part = 0.0;
    if (fb < 2 && g < (unsigned)lat.n5) {
        const unsigned Ls = (unsigned)lat.Ls, kvol = (unsigned)lat.kvol;
        unsigned s, site;
        if (LAYOUT) {
            s = g / kvol;
            site = g - s * kvol;
        } else {
            site = g / Ls;
            s = g - site * Ls;
        }
        // the dot product's operand is loaded before the stencil, so that its
        // latency overlaps the neighbour loads instead of following them (the
        // CG update's r is not: read early, it slowed the kernel after it)
        cx<T> apre[4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const size_t po = P * (size_t)(d + 4 * fb);
            if (EPI == EPI_DOT) apre[d] = S::ld((const typename S::st *)ea.aux0 + po, g);
        }
        cx<T> o[1][4], c[1][4];
        stencil<S, DAG, 1, LAYOUT>(lat, in, chi, hpin, imass, s, site, fb, o, c);
        TA alpha_w = TA(0);
        if constexpr (EPI == EPI_CGW) alpha_w = (TA)cgw_wait((CgState *)ea.cgs);
        double acc = 0.0;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const size_t po = P * (size_t)(d + 4 * fb);
            if (EPI == EPI_SCALEQ) {
                const T ib = (T)(1.0 / ea.scal[0]);
                store_nt<S>(out + po, g, cscale(ib, o[0][d]));
                store_nt<S>((typename S::st *)ea.aux0 + po, g, cscale(ib, c[0][d]));
            } else if (EPI == EPI_CG || EPI == EPI_CGQ || EPI == EPI_CGW) {
                if (EPI == EPI_CGQ) store_nt<S>(out + po, g, o[0][d]);
                const TA alpha = EPI == EPI_CGW ? alpha_w : (TA)ea.scal[0];
                typename SA::st *xr = (typename SA::st *)ea.aux1 + po;
                typename SA::st *rr = (typename SA::st *)ea.aux2 + po;
                if (!(ea.flag & 1)) {
                    const cx<T> pv = S::ld((const typename S::st *)ea.aux0 + po, g);
                    cx<TA> xv = SA::ld(xr, g);
                    xv = {xv.r + alpha * (TA)pv.r, xv.i + alpha * (TA)pv.i};
                    store_nt<SA>(xr, g, xv);
                }
                cx<TA> rv = SA::ld(rr, g);
                rv = {rv.r - alpha * (TA)o[0][d].r, rv.i - alpha * (TA)o[0][d].i};
                store_nt<SA>(rr, g, rv);
                acc += norm2_d(rv);
            } else {
                store_nt<S>(out + po, g, o[0][d]);
                if (EPI == EPI_NORM) acc += norm2_d(o[0][d]);
                if (EPI == EPI_DOT) acc += dotre_d(apre[d], o[0][d]);
            }
        }
        part = acc;
    }
    if (EPI == EPI_NORM || EPI == EPI_DOT || EPI == EPI_CG || EPI == EPI_CGQ || EPI == EPI_CGW) {
        const double t = block_sum(part);
        if (threadIdx.x == 0) ea.partial[lb] = t;
    }
}

template <class S, class SA, bool DAG, int EPI, int LAYOUT = 1, int WPE = 1>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE))) void k_dsl(
    Lat lat, typename S::st *__restrict__ out, const typename S::st *__restrict__ in, const double *__restrict__ chi,
    double hpin, int imass, EpiArgs ea) {
    if (ea.skip && *ea.skip) return;
    if constexpr (EPI == EPI_CGW)
        if (blockIdx.x == 0) cgw_publish((CgState *)ea.cgs, ea.partial, (int)gridDim.x);
    dsl_block<S, SA, DAG, EPI, LAYOUT>(lat, out, in, chi, hpin, imass, ea, logical_block_x());
}

// ---- plane stencil (FP64 dslash without epilogue, L^2 <= 1024) -----------
// One workgroup per (flavour, s, t) xy-plane, one thread per site of it: the
// plane's four spin components are staged in LDS (4 L^2 elements, 64 KB at
// L = 32) and the x and y neighbours are read from there.  The t neighbours
// are loaded with the centre and their two hops summed into their own
// accumulator before the barrier, the s neighbours and the wall operands are
// loaded after it: that keeps 72 VGPRs live (no spill at the 7 waves/SIMD
// the attribute asks for), so three 576-thread workgroups share a CU at
// L = 24 instead of two.  Measured at L=24, Ls=32 with inputs from HBM:
// 28.9 us against 31.8 us with all loads before the barrier (94 VGPRs) and
// 35.1 us for the one-site kernel; forcing 8 waves/SIMD spills and takes 50
// us.  The t-hop sum joins after the x and y hops, so the result differs from
// stencil()'s term order by rounding only (<= 1 ulp of the largest term).
// The fused-epilogue stencils keep the one-site form.
template <class S, bool DAG>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(7))) void k_dsl_plane(
    Lat lat, typename S::st *__restrict__ out, const typename S::st *__restrict__ in, const double *__restrict__ chi,
    double hpin, int imass) {
    typedef typename S::rt T;
    typedef typename S::st st;
    extern __shared__ unsigned char smem[];
    st *sh = reinterpret_cast<st *>(smem);
    const unsigned L = lat.L, L2 = L * L, Lt = lat.Lt, Ls = lat.Ls, kvol = lat.kvol;
    const unsigned lb = logical_block_x();
    const unsigned per_f = Ls * Lt;
    const int fb = (int)(lb / per_f);
    const unsigned rem = lb - (unsigned)fb * per_f, s = rem / Lt, it = rem - s * Lt;
    const unsigned xy = threadIdx.x;
    const bool act = xy < L2 && fb < 2;
    // inactive threads load the plane's first site (in bounds) and only join the barrier
    const unsigned xyc = act ? xy : 0u, itc = fb < 2 ? it : 0u, sc = fb < 2 ? s : 0u;
    const int fbc = fb < 2 ? fb : 0;
    const unsigned ix = xyc % L, iy = xyc / L;
    const unsigned site = xyc + L2 * itc, g = sc * kvol + site;
    const size_t P = (size_t)lat.ps;
    const st *pl[4];
#pragma unroll
    for (int d = 0; d < 4; d++) pl[d] = in + P * (size_t)(d + 4 * fbc);
    cx<T> c[4], o[4];
    const T dg = (T)lat.diag;
#pragma unroll
    for (int d = 0; d < 4; d++) c[d] = S::ld(pl[d], g);
    const unsigned stt = L2;
    const unsigned gtf = (itc + 1 == Lt) ? g - (Lt - 1) * stt : g + stt;
    const unsigned gtb = (itc == 0) ? g + (Lt - 1) * stt : g - stt;
    cx<T> tf[4], tb[4];
#pragma unroll
    for (int d = 0; d < 4; d++) {
        tf[d] = S::ld(pl[d], gtf);
        tb[d] = S::ld(pl[d], gtb);
    }
    constexpr int dn = DAG ? 0 : 2;
    constexpr int dp = DAG ? 2 : 0;
    const bool last = (sc + 1 == Ls), first = (sc == 0);
    const unsigned gsp = last ? g : g + kvol, gsm = first ? g : g - kvol;
    if (act) {
#pragma unroll
        for (int d = 0; d < 4; d++) {
            st w;
            S::sto(&w, 0, c[d]);
            sh[d * L2 + xy] = w;
        }
    }
    const T mh = T(-0.5);
    constexpr int XF = DAG ? -1 : 0, XB = DAG ? 0 : -1;
    constexpr int YF = DAG ? -2 : 1, YB = DAG ? 1 : -2;
    constexpr int TF = DAG ? -3 : 2, TB = DAG ? 2 : -3;
    // antiperiodic in t across it = Lt-1 <-> 0 (rhmc/cheisdwf.f:2831-2876)
    cx<T> ot[4];
#pragma unroll
    for (int d = 0; d < 4; d++) ot[d] = {T(0), T(0)};
    hop_vals<T, TF>(ot, tf[0], tf[1], tf[2], tf[3], (lat.apbc && it + 1 == Lt) ? T(0.5) : mh);
    hop_vals<T, TB>(ot, tb[0], tb[1], tb[2], tb[3], (lat.apbc && it == 0) ? T(0.5) : mh);
    __syncthreads();
    if (!act) return;
    cx<T> up[2], dw[2];
#pragma unroll
    for (int dd = 0; dd < 2; dd++) {
        up[dd] = S::ld(pl[dn + dd], gsp);
        dw[dd] = S::ld(pl[dp + dd], gsm);
    }
#pragma unroll
    for (int d = 0; d < 4; d++) o[d] = cscale(dg, c[d]);
    const unsigned xf = (ix + 1 == L) ? xy - (L - 1) : xy + 1;
    const unsigned xb = (ix == 0) ? xy + (L - 1) : xy - 1;
    const unsigned yf = (iy + 1 == L) ? xy - (L - 1) * L : xy + L;
    const unsigned yb = (iy == 0) ? xy + (L - 1) * L : xy - L;
    hop_vals<T, XF>(o, S::ld(sh, xf), S::ld(sh + L2, xf), S::ld(sh + 2 * L2, xf), S::ld(sh + 3 * L2, xf), mh);
    hop_vals<T, XB>(o, S::ld(sh, xb), S::ld(sh + L2, xb), S::ld(sh + 2 * L2, xb), S::ld(sh + 3 * L2, xb), mh);
    hop_vals<T, YF>(o, S::ld(sh, yf), S::ld(sh + L2, yf), S::ld(sh + 2 * L2, yf), S::ld(sh + 3 * L2, yf), mh);
    hop_vals<T, YB>(o, S::ld(sh, yb), S::ld(sh + L2, yb), S::ld(sh + 2 * L2, yb), S::ld(sh + 3 * L2, yb), mh);
#pragma unroll
    for (int d = 0; d < 4; d++) o[d] = cadd(o[d], ot[d]);
    const cx<T> zero = {T(0), T(0)};
#pragma unroll
    for (int dd = 0; dd < 2; dd++) {
        o[dn + dd] = csub(o[dn + dd], last ? zero : up[dd]);
        o[dp + dd] = csub(o[dp + dd], first ? zero : dw[dd]);
    }
    const bool wall_ok = (imass == 1 || imass == 3);
    if ((first || last) && wall_ok) {
        const st *q0 = in, *q1 = in + 4 * P;
        const unsigned src = last ? site : (Ls - 1) * kvol + site;
        cx<T> v[2][2];
#pragma unroll
        for (int dd = 0; dd < 2; dd++) {
            const int d = last ? dn + dd : dp + dd;
            v[dd][0] = S::ld(q0 + P * (size_t)d, src);
            v[dd][1] = S::ld(q1 + P * (size_t)d, src);
        }
        const T c1 = (T)chi[site], c2 = (T)chi[site + kvol];
        double c3d = chi[site + 2u * kvol];
        if (imass == 1 && site == 0) c3d += hpin;
        const T c3 = (T)c3d;
#pragma unroll
        for (int dd = 0; dd < 2; dd++) {
            cx<T> w0 = {0, 0}, w1 = {0, 0};
            wall_apply<T>(w0, w1, v[dd][0], v[dd][1], c1, c2, c3, imass, last);
            const cx<T> w = fb ? w1 : w0;
            if (last) o[dn + dd] = cadd(o[dn + dd], w);
            else o[dp + dd] = cadd(o[dp + dd], w);
        }
    }
#pragma unroll
    for (int d = 0; d < 4; d++) store_nt<S>(out + P * (size_t)(d + 4 * fb), g, o[d]);
}

// Launch shape of k_dsl_plane: one workgroup per (flavour, s, t) plane, the
// grid rounded up to a multiple of 8 (the extra workgroups exit after the
// barrier), L^2 threads rounded up to a wavefront.
inline bool plane_stencil_fits(const Lat &lat) { return lat.L * lat.L <= 1024; }
inline unsigned plane_grid(const Lat &lat) { return (2u * (unsigned)(lat.Ls * lat.Lt) + 7u) / 8u * 8u; }
inline unsigned plane_block(const Lat &lat) { return ((unsigned)(lat.L * lat.L) + 63u) / 64u * 64u; }

// ---- paired-site stencil (FP32 / BF16 storage, even L) --------------------
// One thread computes the sites x0 and x0 + 1 (x0 even) of one (s, y, t) row
// and one flavour.  Every plane access of the centre and of the y, t and s
// neighbours is one 16-byte (FP32) or 8-byte (BF16) load of both elements,
// the inner x neighbours are the pair itself and only the two outer x
// neighbours are single loads: half the load instructions per site of the
// one-site stencil.  Each site accumulates its terms in the same order as
// stencil(), so the outputs are identical to it.
template <class S>
__device__ __forceinline__ void store_nt2(typename S::st *__restrict__ p, unsigned i, typename S::st2 v) {
    if constexpr (sizeof(v) == 16) {
        u32x4 raw;
        __builtin_memcpy(&raw, &v, 16);
        __builtin_nontemporal_store(raw, reinterpret_cast<u32x4 *>(p + i));
    } else {
        u32x2 raw;
        __builtin_memcpy(&raw, &v, 8);
        __builtin_nontemporal_store(raw, reinterpret_cast<u32x2 *>(p + i));
    }
}

template <class S, bool DAG>
__device__ __forceinline__ void stencil_pair(const Lat &lat, const typename S::st *__restrict__ in,
                                             const double *__restrict__ chi, double hpin, int imass, unsigned s,
                                             unsigned site, int fb, cx<typename S::rt> (&o0)[4],
                                             cx<typename S::rt> (&o1)[4], cx<typename S::rt> (&c0)[4],
                                             cx<typename S::rt> (&c1)[4]) {
    typedef typename S::rt T;
    typedef typename S::st st;
    const unsigned Ls = (unsigned)lat.Ls, L = (unsigned)lat.L, Lt = (unsigned)lat.Lt, kvol = (unsigned)lat.kvol;
    const size_t P = (size_t)lat.ps;
    const st *pl[4];
#pragma unroll
    for (int d = 0; d < 4; d++) pl[d] = in + P * (size_t)(d + 4 * fb);
    const unsigned ix = site % L;
    const unsigned r1 = site / L;
    const unsigned iy = r1 % L;
    const unsigned it = r1 / L;
    const unsigned g = s * kvol + site;
    const T dg = (T)lat.diag;
#pragma unroll
    for (int d = 0; d < 4; d++) {
        S::ld2(pl[d], g, c0[d], c1[d]);
        o0[d] = cscale(dg, c0[d]);
        o1[d] = cscale(dg, c1[d]);
    }
    const T mh = T(-0.5);
    constexpr int XF = DAG ? -1 : 0, XB = DAG ? 0 : -1;
    constexpr int YF = DAG ? -2 : 1, YB = DAG ? 1 : -2;
    constexpr int TF = DAG ? -3 : 2, TB = DAG ? 2 : -3;
    // x: site 0's forward and site 1's backward neighbour are the pair itself
    {
        const unsigned gf = (ix + 2 == L) ? g - (L - 2) : g + 2;
        const unsigned gb = (ix == 0) ? g + (L - 1) : g - 1;
        const cx<T> f0 = S::ld(pl[0], gf), f1 = S::ld(pl[1], gf), f2 = S::ld(pl[2], gf), f3 = S::ld(pl[3], gf);
        const cx<T> b0 = S::ld(pl[0], gb), b1 = S::ld(pl[1], gb), b2 = S::ld(pl[2], gb), b3 = S::ld(pl[3], gb);
        hop_vals<T, XF>(o0, c1[0], c1[1], c1[2], c1[3], mh);
        hop_vals<T, XB>(o0, b0, b1, b2, b3, mh);
        hop_vals<T, XF>(o1, f0, f1, f2, f3, mh);
        hop_vals<T, XB>(o1, c0[0], c0[1], c0[2], c0[3], mh);
    }
    const unsigned sy = L, stt = L * L;
    hop_pair<S, YF>(o0, o1, pl, (iy + 1 == L) ? g - (L - 1) * sy : g + sy, mh);
    hop_pair<S, YB>(o0, o1, pl, (iy == 0) ? g + (L - 1) * sy : g - sy, mh);
    hop_pair<S, TF>(o0, o1, pl, (it + 1 == Lt) ? g - (Lt - 1) * stt : g + stt,
                    (lat.apbc && it + 1 == Lt) ? T(0.5) : mh);
    hop_pair<S, TB>(o0, o1, pl, (it == 0) ? g + (Lt - 1) * stt : g - stt, (lat.apbc && it == 0) ? T(0.5) : mh);
    constexpr int dn = DAG ? 0 : 2;
    constexpr int dp = DAG ? 2 : 0;
    const bool wall_ok = (imass == 1 || imass == 3);
    const bool last = (s + 1 == Ls), first = (s == 0);
    // unconditional s-neighbour loads, clamped at the walls (as in stencil())
    {
        const unsigned gsp = last ? g : g + kvol, gsm = first ? g : g - kvol;
        const cx<T> zero = {T(0), T(0)};
#pragma unroll
        for (int dd = 0; dd < 2; dd++) {
            cx<T> a, b, e, f;
            S::ld2(pl[dn + dd], gsp, a, b);
            S::ld2(pl[dp + dd], gsm, e, f);
            o0[dn + dd] = csub(o0[dn + dd], last ? zero : a);
            o1[dn + dd] = csub(o1[dn + dd], last ? zero : b);
            o0[dp + dd] = csub(o0[dp + dd], first ? zero : e);
            o1[dp + dd] = csub(o1[dp + dd], first ? zero : f);
        }
    }
    if ((first || last) && wall_ok) {
        const st *q0 = in, *q1 = in + 4 * P;   // flavour 0 / 1 planes at d = 0
        const unsigned src = last ? site : (Ls - 1) * kvol + site;
        // all wall operands are loaded before any is used (one round trip)
        cx<T> v[2][4];
#pragma unroll
        for (int dd = 0; dd < 2; dd++) {
            const int d = last ? dn + dd : dp + dd;
            S::ld2(q0 + P * (size_t)d, src, v[dd][0], v[dd][1]);
            S::ld2(q1 + P * (size_t)d, src, v[dd][2], v[dd][3]);
        }
        const double2 ch1 = *reinterpret_cast<const double2 *>(chi + site);
        const double2 ch2 = *reinterpret_cast<const double2 *>(chi + site + kvol);
        const double2 ch3 = *reinterpret_cast<const double2 *>(chi + site + 2u * kvol);
        const T c1a = (T)ch1.x, c1b = (T)ch1.y, c2a = (T)ch2.x, c2b = (T)ch2.y;
        double c3ad = ch3.x;
        if (imass == 1 && site == 0) c3ad += hpin;
        const T c3a = (T)c3ad, c3b = (T)ch3.y;
#pragma unroll
        for (int dd = 0; dd < 2; dd++) {
            const int d = last ? dn + dd : dp + dd;
            const cx<T> v0a = v[dd][0], v0b = v[dd][1], v1a = v[dd][2], v1b = v[dd][3];
            cx<T> w0a = {0, 0}, w1a = {0, 0}, w0b = {0, 0}, w1b = {0, 0};
            wall_apply<T>(w0a, w1a, v0a, v1a, c1a, c2a, c3a, imass, last);
            wall_apply<T>(w0b, w1b, v0b, v1b, c1b, c2b, c3b, imass, last);
            o0[d] = cadd(o0[d], fb ? w1a : w0a);
            o1[d] = cadd(o1[d], fb ? w1b : w0b);
        }
    }
}

__host__ __device__ inline unsigned stencil_pair_blocks_per_flavour(long n5) {
    return (unsigned)((n5 / 2 + kBlock - 1) / kBlock);
}

template <class S, class SA, bool DAG, int EPI>
__device__ __forceinline__ void dsl_block_pair(const Lat &lat, typename S::st *__restrict__ out,
                                               const typename S::st *__restrict__ in,
                                               const double *__restrict__ chi, double hpin, int imass,
                                               const EpiArgs &ea, unsigned lb) {
    typedef typename S::rt T;
    typedef typename SA::rt TA;
    const unsigned bpf = stencil_pair_blocks_per_flavour(lat.n5);
    const int fb = (int)(lb / bpf);
    const unsigned g2 = (lb - (unsigned)fb * bpf) * kBlock + threadIdx.x;   // pair index in (s, site) order
    const size_t P = (size_t)lat.ps;
    double part = 0.0;
    if (fb < 2 && g2 < (unsigned)(lat.n5 / 2)) {
        const unsigned kvol = (unsigned)lat.kvol;
        const unsigned g = 2 * g2;
        const unsigned s = g / kvol;
        const unsigned site = g - s * kvol;
        // epilogue operands loaded ahead of the stencil (see dsl_block)
        cx<T> apa[4], apb[4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const size_t po = P * (size_t)(d + 4 * fb);
            if (EPI == EPI_DOT) S::ld2((const typename S::st *)ea.aux0 + po, g, apa[d], apb[d]);
        }
        cx<T> o0[4], o1[4], c0[4], c1[4];
        stencil_pair<S, DAG>(lat, in, chi, hpin, imass, s, site, fb, o0, o1, c0, c1);
        TA alpha_w = TA(0);
        if constexpr (EPI == EPI_CGW) alpha_w = (TA)cgw_wait((CgState *)ea.cgs);
        double acc = 0.0;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const size_t po = P * (size_t)(d + 4 * fb);
            if (EPI == EPI_SCALEQ) {
                const T ib = (T)(1.0 / ea.scal[0]);
                store_nt2<S>(out + po, g, S::pack2(cscale(ib, o0[d]), cscale(ib, o1[d])));
                store_nt2<S>((typename S::st *)ea.aux0 + po, g, S::pack2(cscale(ib, c0[d]), cscale(ib, c1[d])));
            } else if (EPI == EPI_CG || EPI == EPI_CGQ || EPI == EPI_CGW) {
                if (EPI == EPI_CGQ) store_nt2<S>(out + po, g, S::pack2(o0[d], o1[d]));
                const TA alpha = EPI == EPI_CGW ? alpha_w : (TA)ea.scal[0];
                typename SA::st *xr = (typename SA::st *)ea.aux1 + po;
                typename SA::st *rr = (typename SA::st *)ea.aux2 + po;
                if (!(ea.flag & 1)) {
                    cx<T> pa, pb;
                    S::ld2((const typename S::st *)ea.aux0 + po, g, pa, pb);
                    cx<TA> xa, xb;
                    SA::ld2(xr, g, xa, xb);
                    xa = {xa.r + alpha * (TA)pa.r, xa.i + alpha * (TA)pa.i};
                    xb = {xb.r + alpha * (TA)pb.r, xb.i + alpha * (TA)pb.i};
                    store_nt2<SA>(xr, g, SA::pack2(xa, xb));
                }
                cx<TA> ra, rb;
                SA::ld2(rr, g, ra, rb);
                ra = {ra.r - alpha * (TA)o0[d].r, ra.i - alpha * (TA)o0[d].i};
                rb = {rb.r - alpha * (TA)o1[d].r, rb.i - alpha * (TA)o1[d].i};
                store_nt2<SA>(rr, g, SA::pack2(ra, rb));
                acc += norm2_d(ra) + norm2_d(rb);
            } else {
                store_nt2<S>(out + po, g, S::pack2(o0[d], o1[d]));
                if (EPI == EPI_NORM)
                    acc += norm2_d(o0[d]) + norm2_d(o1[d]);
                if (EPI == EPI_DOT)
                    acc += dotre_d(apa[d], o0[d]) + dotre_d(apb[d], o1[d]);
            }
        }
        part = acc;
    }
    if (EPI == EPI_NORM || EPI == EPI_DOT || EPI == EPI_CG || EPI == EPI_CGQ || EPI == EPI_CGW) {
        const double t = block_sum(part);
        if (threadIdx.x == 0) ea.partial[lb] = t;
    }
}

template <class S, class SA, bool DAG, int EPI>
__global__ __launch_bounds__(kBlock) void k_dsl_pair(Lat lat, typename S::st *__restrict__ out,
                                                     const typename S::st *__restrict__ in,
                                                     const double *__restrict__ chi, double hpin, int imass,
                                                     EpiArgs ea) {
    if (ea.skip && *ea.skip) return;
    if constexpr (EPI == EPI_CGW)
        if (blockIdx.x == 0) cgw_publish((CgState *)ea.cgs, ea.partial, (int)gridDim.x);
    dsl_block_pair<S, SA, DAG, EPI>(lat, out, in, chi, hpin, imass, ea, logical_block_x());
}

// Grid size (workgroups) of a k_dsl_pair launch.
inline unsigned stencil_grid_pair(long n5) {
    unsigned nb = 2 * stencil_pair_blocks_per_flavour(n5);
    return (nb + 7u) / 8u * 8u;
}

// Grid size (workgroups) of a k_dsl launch.
inline unsigned stencil_grid(long n5) {
    unsigned nb = 2 * stencil_blocks_per_flavour(n5);
    return (nb + 7u) / 8u * 8u;
}

// ---- pseudofermion force on the two wall slices (rhmc/cheisdwf.f:733-849) --
// One thread per (site, shift): y = x_k, z = M x_k on the slices s = 0 and
// s = Ls-1 only (derivs uses no other slice), contribution of shift k at the
// site into contrib[(k*3 + a)*kvol + site].  Native layout.
// (x_k stored in S; the bilinears and their sums are FP64)
template <class S>
__global__ __launch_bounds__(kBlock) void k_force_shift(Lat lat, const typename S::st *__restrict__ xk, long n,
                                                        int ndiag, const double *__restrict__ anum,
                                                        const double *__restrict__ chi, double hpin, int imass,
                                                        double *__restrict__ contrib) {
    typedef typename S::rt T;
    const long gid = (long)blockIdx.x * kBlock + threadIdx.x;
    if (gid >= (long)lat.kvol * ndiag) return;
    const int k = (int)(gid / lat.kvol);
    const unsigned site = (unsigned)(gid - (long)k * lat.kvol);
    const typename S::st *y = xk + (size_t)k * n;
    cx<T> zf[2][4], cf[2][4], zl[2][4], cl[2][4];
    stencil<S, false, 2, 1>(lat, y, chi, hpin, imass, 0, site, 0, zf, cf);
    stencil<S, false, 2, 1>(lat, y, chi, hpin, imass, (unsigned)lat.Ls - 1, site, 0, zl, cl);
    double2 ya[4][2], za[4][2];
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
        for (int d = 0; d < 4; d++) {
            // d = 0,1: y(s=Ls) with z(s=1); d = 2,3: y(s=1) with z(s=Ls)
            const cx<T> yv = d < 2 ? cl[f][d] : cf[f][d];
            const cx<T> zv = d < 2 ? zf[f][d] : zl[f][d];
            ya[d][f] = make_double2((double)yv.r, (double)yv.i);
            za[d][f] = make_double2((double)zv.r, (double)zv.i);
        }
    double F1 = 0, F2 = 0, F3 = 0;
    derivs_site(F1, F2, F3, ya, za, -2.0 * anum[k + 1], imass);
    contrib[((long)k * 3 + 0) * lat.kvol + site] = F1;
    contrib[((long)k * 3 + 1) * lat.kvol + site] = F2;
    contrib[((long)k * 3 + 2) * lat.kvol + site] = F3;
}

// derivs on given y and z (full native spinors), one thread per site:
// F += wall bilinears (rhmc/cheisdwf.f:747-849)
__global__ __launch_bounds__(kBlock) void k_derivs(Lat lat, const double2 *__restrict__ y,
                                                   const double2 *__restrict__ z, double anum, int imass,
                                                   double *__restrict__ F) {
    const long site = (long)blockIdx.x * kBlock + threadIdx.x;
    if (site >= lat.kvol) return;
    const size_t P = (size_t)lat.ps;
    const size_t e0 = (size_t)site, el = (size_t)(lat.Ls - 1) * lat.kvol + site;
    double2 ya[4][2], za[4][2];
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const size_t gy = d < 2 ? el : e0, gz = d < 2 ? e0 : el;
            ya[d][f] = y[gy + P * (d + 4 * f)];
            za[d][f] = z[gz + P * (d + 4 * f)];
        }
    double F1 = F[site], F2 = F[site + lat.kvol], F3 = F[site + 2L * lat.kvol];
    derivs_site(F1, F2, F3, ya, za, -2.0 * anum, imass);
    F[site] = F1;
    F[site + lat.kvol] = F2;
    F[site + 2L * lat.kvol] = F3;
}

// Layout conversion between the reference order (site*Ls + s) and the native
// order (s*kvol + site) inside each of the 8 planes; dir = 0: ref -> native.
template <class S>
__global__ __launch_bounds__(kBlock) void k_relayout(Lat lat, typename S::st *__restrict__ dst,
                                                     const typename S::st *__restrict__ src, int dir) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;   // native order index over 8 planes
    const size_t n5 = (size_t)lat.n5;
    if (i >= 8 * n5) return;
    const size_t plane = i / n5;
    const unsigned e = (unsigned)(i - plane * n5);
    const unsigned s = e / (unsigned)lat.kvol, site = e - s * (unsigned)lat.kvol;
    const size_t r = plane * n5 + (size_t)site * lat.Ls + s;   // reference order index
    if (dir == 0) dst[i] = src[r];
    else dst[r] = src[i];
}

}  // namespace chs
