// chs_batch.h -- grouped kernels of the batched multi-shift solve: one launch
// per Lanczos stage covers every system of a batch (pseudofermion fields,
// lattices of different sizes), so that small lattices fill the GPU together
// and a whole step is a handful of launches that a hipGraph replays.
//
// A system is one multi-shift solve (rhmc/cheisdwf.f:599-745, iflag = 0 or 1)
// with its own geometry, source, auxiliary field, shifts and device work space;
// the spinors are stored in FP64 or FP32 (S) for the whole batch, the
// Lanczos scalars, reductions and the rational combination are FP64.
// Each grouped grid is the concatenation of the systems' own grids in system
// order; a workgroup finds its system from the table of first blocks and then
// runs exactly the per-block code of the single-system kernels (dsl_block,
// lz_alpha, lz_update, lz_beta) with the same logical block number, so every
// partial sum lands in the same slot with the same value and a batched solve
// reproduces the one-at-a-time solve bit for bit.
//
// The shift update may cover EPT element blocks per workgroup (fewer, fatter
// workgroups; one partial per element block, so no change in any value).
// The iteration counter lives on the device (itercg), and the Lanczos vector
// pair alternates with the parity of the iteration, passed as an argument,
// so a captured sequence of an even number of iterations can be replayed.
#pragma once
#include "chs_stencil.h"

#include <type_traits>

namespace chs {

struct MsSys {
    Lat lat;
    long n;                       // complex elements of a spinor (8 * Ls * kvol)
    const double *chi;
    double hpin;
    int imass, ndiag;
    int np1, np2;                 // partial counts: stencil grid, vector grid
    void *r, *q0, *q1, *v, *w, *pk;   // spinors in the batch's storage precision
    double2 *x;                   // the caller's output: a0 phi + sum_k anum[k+1] x_k (FP64)
    int pairs;                    // FP32 storage on an even L: the paired-site stencil
    int iflag;                    // 1: every shift's x_k kept (on the wall slices) for the force
    void *xk;                     // iflag = 1: the x_k, shift-major (storage precision)
    XKeep xkeep;                  // iflag = 1: the slices of x_k that are stored
    const double *anum;           // device copy of the remez numerators
    LzState *lz;
    double *partial, *partial2;
};

struct MsGroup {
    const MsSys *sys;
    const unsigned *s_dsl;        // first block of each system in the stencil grid
    const unsigned *s_vec;        // ... in the shift-update grid (np2 / EPT blocks each)
    int nsys;
    int *status;                  // [0] systems done, [1] largest itercg
};

__device__ __forceinline__ int find_sys(const unsigned *start, int nsys, unsigned b) {
    int j = 0;
    for (int k = 1; k < nsys; k++)
        if (b >= start[k]) j = k;
    return j;
}

// stage 1: v = M r / beta, q_cur = r / beta; stage 2: w = M^dagger v, <q_cur, w>
// (the stencil variant each system's one-at-a-time solve uses)
template <class S, bool DAG>
__global__ __launch_bounds__(kBlock) void k_ms_dsl(MsGroup gr, int parity) {
    typedef typename S::st st;
    const unsigned b = logical_block_x();
    const int j = find_sys(gr.s_dsl, gr.nsys, b);
    const MsSys &s = gr.sys[j];
    if (s.lz->done) return;
    st *qc = (st *)(parity ? s.q0 : s.q1);
    EpiArgs ea;
    ea.aux0 = qc;
    ea.aux1 = ea.aux2 = nullptr;
    ea.skip = nullptr;
    ea.flag = 0;
    ea.cgs = nullptr;
    const unsigned lb = b - gr.s_dsl[j];
    if (!DAG) {
        ea.partial = nullptr;
        ea.scal = &s.lz->beta;
        if constexpr (std::is_same<S, F32>::value) {
            if (s.pairs) {
                dsl_block_pair<S, S, false, EPI_SCALEQ>(s.lat, (st *)s.v, (const st *)s.r, s.chi, s.hpin, s.imass,
                                                        ea, lb);
                return;
            }
        }
        dsl_block<S, S, false, EPI_SCALEQ, 1>(s.lat, (st *)s.v, (const st *)s.r, s.chi, s.hpin, s.imass, ea, lb);
    } else {
        ea.partial = s.partial;
        ea.scal = nullptr;
        if constexpr (std::is_same<S, F32>::value) {
            if (s.pairs) {
                dsl_block_pair<S, S, true, EPI_DOT>(s.lat, (st *)s.w, (const st *)s.v, s.chi, s.hpin, s.imass, ea,
                                                    lb);
                return;
            }
        }
        dsl_block<S, S, true, EPI_DOT, 1>(s.lat, (st *)s.w, (const st *)s.v, s.chi, s.hpin, s.imass, ea, lb);
    }
}

__global__ __launch_bounds__(kBlock) void k_ms_alpha(MsGroup gr) {
    const MsSys &s = gr.sys[blockIdx.x];
    if (s.lz->done) return;
    lz_alpha(s.lz, s.partial, s.np1, s.lz->itercg + 1);
}

__global__ __launch_bounds__(kBlock) void k_ms_beta(MsGroup gr) {
    const MsSys &s = gr.sys[blockIdx.x];
    if (s.lz->done) return;
    lz_beta(s.lz, s.partial2, s.np2, s.lz->itercg + 1);
}

// Shift update with EPT element blocks per workgroup (the partials do not
// depend on EPT); the rational combination accumulates directly into x.
template <class S, int EPT>
__device__ __forceinline__ void ms_update(const MsSys &s, int parity, long lb) {
    typedef typename S::st st;
    const st *qc = (const st *)(parity ? s.q0 : s.q1);
    const st *qp = (const st *)(parity ? s.q1 : s.q0);
    const st *w = (const st *)s.w;
    st *r = (st *)s.r, *pk = (st *)s.pk;
    const int iter = s.lz->itercg + 1;
    if (s.iflag) {   // the force's solve: x_k kept on the slices derivs reads (as k_lz_update)
        if (s.ndiag == 25)
            lz_update<S, 25, false, EPT>(s.n, w, qc, qp, r, pk, s.xk, s.anum, s.lz, s.ndiag, iter, s.partial2, lb,
                                         s.xkeep);
        else if (s.ndiag == 12)
            lz_update<S, 12, false, EPT>(s.n, w, qc, qp, r, pk, s.xk, s.anum, s.lz, s.ndiag, iter, s.partial2, lb,
                                         s.xkeep);
        else
            lz_update<S, 0, false, EPT>(s.n, w, qc, qp, r, pk, s.xk, s.anum, s.lz, s.ndiag, iter, s.partial2, lb,
                                        s.xkeep);
        return;
    }
    if (s.ndiag == 25)
        lz_update<S, 25, true, EPT>(s.n, w, qc, qp, r, pk, s.x, s.anum, s.lz, s.ndiag, iter, s.partial2, lb);
    else if (s.ndiag == 12)
        lz_update<S, 12, true, EPT>(s.n, w, qc, qp, r, pk, s.x, s.anum, s.lz, s.ndiag, iter, s.partial2, lb);
    else
        lz_update<S, 0, true, EPT>(s.n, w, qc, qp, r, pk, s.x, s.anum, s.lz, s.ndiag, iter, s.partial2, lb);
}

template <class S, int EPT>
__global__ __launch_bounds__(kBlock) void k_ms_update(MsGroup gr, int parity) {
    const unsigned b = logical_block_x();
    const int j = find_sys(gr.s_vec, gr.nsys, b);
    const MsSys &s = gr.sys[j];
    if (s.lz->done) return;
    ms_update<S, EPT>(s, parity, (long)(b - gr.s_vec[j]));
}

// status[0] = systems done, status[1] = largest itercg (one workgroup)
__global__ void k_ms_status(MsGroup gr) {
    if (threadIdx.x != 0) return;
    int nd = 0, it = 0;
    for (int j = 0; j < gr.nsys; j++) {
        nd += gr.sys[j].lz->done != 0;   // 1: converged, 2: halted on a non-finite residual
        it = max(it, gr.sys[j].lz->itercg);
    }
    gr.status[0] = nd;
    gr.status[1] = it;
}

}  // namespace chs
