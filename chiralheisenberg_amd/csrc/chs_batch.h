// chs_batch.h -- grouped kernels of the batched multi-shift solve: one launch
// per Lanczos stage covers every system of a batch (pseudofermion fields,
// lattices of different sizes), so that small lattices fill the GPU together
// and a whole step is a handful of launches that a hipGraph replays.
//
// A system is one multi-shift solve (rhmc/cheisdwf.f:599-745, iflag = 0) with
// its own geometry, source, auxiliary field, shifts and device work space.
// Each grouped grid is the concatenation of the systems' own grids in system
// order; a workgroup finds its system from the table of first blocks and then
// runs exactly the per-block code of the single-system kernels (dsl_block,
// lz_alpha, lz_update, lz_beta) with the same logical block number, so every
// partial sum lands in the same slot with the same value and a batched solve
// reproduces the one-at-a-time solve bit for bit.
//
// The shift update may be split over G workgroups per block of elements
// (shift k handled by group k mod G, the residual and its norm by group 0):
// more workgroups for small lattices, no change in any value.  The
// iteration counter lives on the device (itercg), and the Lanczos vector
// pair alternates with the parity of the iteration, passed as an argument,
// so a captured sequence of an even number of iterations can be replayed.
#pragma once
#include "chs_stencil.h"

namespace chs {

struct MsSys {
    Lat lat;
    long n;                       // complex elements of a spinor (8 * Ls * kvol)
    const double *chi;
    double hpin;
    int imass, ndiag;
    int np1, np2;                 // partial counts: stencil grid, vector grid
    double2 *r, *q0, *q1, *v, *w, *pk, *xk;
    LzState *lz;
    double *partial, *partial2;
};

struct MsGroup {
    const MsSys *sys;
    const unsigned *s_dsl;        // first block of each system in the stencil grid
    const unsigned *s_vec;        // ... in the shift-update grid (G blocks per element block)
    int nsys, G;
    int *status;                  // [0] systems done, [1] largest itercg
};

__device__ __forceinline__ int find_sys(const unsigned *start, int nsys, unsigned b) {
    int j = 0;
    for (int k = 1; k < nsys; k++)
        if (b >= start[k]) j = k;
    return j;
}

// stage 1: v = M r / beta, q_cur = r / beta; stage 2: w = M^dagger v, <q_cur, w>
template <bool DAG>
__global__ __launch_bounds__(kBlock) void k_ms_dsl(MsGroup gr, int parity) {
    const unsigned b = logical_block_x();
    const int j = find_sys(gr.s_dsl, gr.nsys, b);
    const MsSys &s = gr.sys[j];
    if (s.lz->done) return;
    double2 *qc = parity ? s.q0 : s.q1;
    EpiArgs ea;
    ea.aux0 = qc;
    ea.aux1 = ea.aux2 = nullptr;
    ea.skip = nullptr;
    ea.flag = 0;
    const unsigned lb = b - gr.s_dsl[j];
    if (!DAG) {
        ea.partial = nullptr;
        ea.scal = &s.lz->beta;
        dsl_block<F64, F64, false, EPI_SCALEQ, 1>(s.lat, s.v, s.r, s.chi, s.hpin, s.imass, ea, lb);
    } else {
        ea.partial = s.partial;
        ea.scal = nullptr;
        dsl_block<F64, F64, true, EPI_DOT, 1>(s.lat, s.w, s.v, s.chi, s.hpin, s.imass, ea, lb);
    }
}

__global__ void k_ms_alpha(MsGroup gr) {
    const MsSys &s = gr.sys[blockIdx.x];
    if (s.lz->done) return;
    lz_alpha(s.lz, s.partial, s.np1, s.lz->itercg + 1);
}

__global__ void k_ms_beta(MsGroup gr) {
    const MsSys &s = gr.sys[blockIdx.x];
    if (s.lz->done) return;
    lz_beta(s.lz, s.partial2, s.np2, s.lz->itercg + 1);
}

// Shift update of one element block for shift group g of G: the same values
// as lz_update (group 0 also forms r and the |r|^2 partial).
template <int NS>
__device__ __forceinline__ void lz_update_split(const MsSys &s, const double2 *__restrict__ q,
                                                const double2 *__restrict__ qm1, int iter, long lb, int g, int G) {
    const LzState *st = s.lz;
    const int ndiag = s.ndiag;
    const long n = s.n;
    __shared__ double s_amu[kMaxShift], s_rho[kMaxShift];
    if (threadIdx.x < ndiag) {
        s_amu[threadIdx.x] = st->amu[threadIdx.x];
        s_rho[threadIdx.x] = st->rho[threadIdx.x];
    }
    __syncthreads();
    const long i = lb * kBlock + threadIdx.x;
    double part = 0.0;
    if (i < n) {
        const double2 qv = q[i];
        if (g == 0) {
            const double at = st->alphatild, b0 = st->beta;
            const double2 wv = s.w[i];
            const double2 qm = qm1[i];
            double2 rv;
            rv.x = wv.x - at * qv.x - b0 * qm.x;
            rv.y = wv.y - at * qv.y - b0 * qm.y;
            s.r[i] = rv;
            part = rv.x * rv.x + rv.y * rv.y;
        }
        const int ns = NS > 0 ? NS : ndiag;
        double2 *__restrict__ p = s.pk;
        double2 *__restrict__ x = s.xk;
        if (iter == 1) {
            for (int k = g; k < ns; k += G) {
                p[(long)k * n + i] = qv;
                x[(long)k * n + i] = make_double2(s_rho[k] * qv.x, s_rho[k] * qv.y);
            }
        } else {
            for (int k = g; k < ns; k += G) {
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
    if (g == 0) {
        const double t = block_sum(part);
        if (threadIdx.x == 0) s.partial2[lb] = t;
    }
}

__global__ __launch_bounds__(kBlock) void k_ms_update(MsGroup gr, int parity) {
    const unsigned b = logical_block_x();
    const int j = find_sys(gr.s_vec, gr.nsys, b);
    const MsSys &s = gr.sys[j];
    if (s.lz->done) return;
    const unsigned local = b - gr.s_vec[j];
    const int g = (int)(local % (unsigned)gr.G);
    const long lb = (long)(local / (unsigned)gr.G);
    const double2 *qc = parity ? s.q0 : s.q1;
    const double2 *qp = parity ? s.q1 : s.q0;
    const int iter = s.lz->itercg + 1;
    if (s.ndiag == 25) lz_update_split<25>(s, qc, qp, iter, lb, g, gr.G);
    else if (s.ndiag == 12) lz_update_split<12>(s, qc, qp, iter, lb, g, gr.G);
    else lz_update_split<0>(s, qc, qp, iter, lb, g, gr.G);
}

// status[0] = systems done, status[1] = largest itercg (one workgroup)
__global__ void k_ms_status(MsGroup gr) {
    if (threadIdx.x != 0) return;
    int nd = 0, it = 0;
    for (int j = 0; j < gr.nsys; j++) {
        nd += gr.sys[j].lz->done;
        it = max(it, gr.sys[j].lz->itercg);
    }
    gr.status[0] = nd;
    gr.status[1] = it;
}

}  // namespace chs
