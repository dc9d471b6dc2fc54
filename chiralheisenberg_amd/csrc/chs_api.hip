// chs_api.hip -- host side of libchs: context, device buffers, kernel
// launch sequences of the solvers and the extern "C" entry points declared in
// include/chs.h.
//
// The reference routines (rhmc/cheisdwf.f) operate on host arrays and COMMON
// blocks; here every field lives in HBM for the duration of a call (the
// drop-in entry points upload their inputs once and download their outputs
// once) or permanently (the _dev entry points).  Solver iterations run as
// fixed sequences of kernels whose scalars stay on the device; the host
// enqueues several iterations ahead and polls a halt flag, so no host round
// trip sits inside an iteration.
#include "chs_stencil.h"
#include "../../include/chs.h"

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

using namespace chs;

struct chs_ctx {
    int L, Lt, Ls, ibound, device;
    double am3;
    long kvol, kferm, n5;
    Lat lat;
    hipStream_t stream = nullptr;
    std::string err;
    // device work space, grown on demand and kept
    struct Buf { void *p = nullptr; size_t n = 0; };
    Buf b[20];
    Buf shifts_p, shifts_x, contrib;
    double *partial = nullptr, *partial2 = nullptr;
    int npart = 0;
    LzState *lz = nullptr;
    CgState *cg = nullptr;
    double *dscal = nullptr;      // small device scalar array
    double *dcoef = nullptr;      // remez coefficients (anum) on the device
    void *hpinned = nullptr;      // pinned host scratch for polling
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    struct NNState *nn = nullptr; // learned preconditioner (chs_learn_host.inc)
    int use_graphs = 1;           // replay solver iteration chunks from hipGraphs (CHS_GRAPHS=0: off)
    int pair_stencil = 1;         // paired-site FP32 stencil on even L (CHS_PAIRS=0: off)
    int plane_stencil = 1;        // LDS-plane plain FP64 dslash when L^2 <= 1024 (CHS_PLANES=0: off)
    int force_prec = CHS_FP64;    // storage of the force's multi-shift solves (chs_set_force_precision)
    int lz_cap = 5000;            // multi-shift iteration cap (the reference's niterc)
    double lz_freeze = 0.0;       // multi-shift shift freezing factor (chs_set_shift_freeze; 0 = off)
    void *scratch = nullptr;      // device address of the static scratch on this device
    std::string scratch_err;
    // Device mirrors of read-only host spinors, keyed by host address: the
    // copy on the device is current while its generation equals the host
    // array's (the caller bumps that with chs_mirror_touch after writing it).
    struct Mirror { const void *host; void *dev; unsigned long long gen, dev_gen; };
    std::vector<Mirror> mirrors;
    unsigned long long h2d_spinor_bytes = 0, d2h_spinor_bytes = 0;
};
static void nn_destroy(chs_ctx *c);

static thread_local std::string g_create_err;

// ---- static scratch of chs_congrad_noalloc_dev --------------------------------
// Device memory that exists for the life of the process on every GPU the
// library is loaded on, so the entry point that uses it calls no allocator.
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address comes from hipGetSymbolAddress, once per device, and only
// that address is ever passed to kernels or copies.
constexpr size_t kNoAllocPartialOff = 4096;          // after the CG state
constexpr size_t kNoAllocSpinorOff = 64 * 1024;      // after the partial sums (up to 7680 doubles)
constexpr size_t kNoAllocBytes = kNoAllocSpinorOff + 3 * 16 * (size_t)CHS_NOALLOC_MAX_KFERM;
static_assert(sizeof(chs::CgState) <= kNoAllocPartialOff, "CG state does not fit its scratch slot");
static_assert((kNoAllocSpinorOff - kNoAllocPartialOff) / sizeof(double) >=
                  (CHS_NOALLOC_MAX_KFERM + chs::kBlock - 1) / chs::kBlock + 8,
              "partial sums do not fit their scratch slot");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_dev[64];            // device address of `scratch`, per device
static std::string g_scratch_err[64];

// With `device` current: the device address of the scratch array, looked up
// on the first call for that device.
static void *scratch_address(int device, std::string &err) {
    if (device < 0 || device >= 64) {
        err = "device index outside 0..63";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device] && g_scratch_err[device].empty()) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
        else
            g_scratch_dev[device] = ptr;
    }
    err = g_scratch_err[device];
    return g_scratch_dev[device];
}

namespace {

enum BufId { B_PHI = 0, B_X, B_R, B_Q0, B_Q1, B_V, B_W, B_CHI, B_F, B_PP, B_P, B_T1, B_T2, B_XS, B_RS, B_PS, B_PS2,
             B_STAGE };

int fail(chs_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail((c), CHS_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

inline unsigned nblocks(long n) {
    long nb = (n + kBlock - 1) / kBlock;
    nb = (nb + 7) / 8 * 8;   // multiple of the XCD count for logical_block()
    return (unsigned)(nb > 0 ? nb : 8);
}

// owns an instantiated graph for the duration of a solve
struct GraphGuard {
    hipGraphExec_t exec = nullptr;
    ~GraphGuard() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

// device allocations of one call (the timing helpers' buffers), freed on every
// return, error paths included (hipFree waits for kernels still using them)
struct DevAllocs {
    std::vector<void *> p;
    hipError_t alloc(void **out, size_t bytes) {
        *out = nullptr;
        const hipError_t e = hipMalloc(out, bytes);
        if (e == hipSuccess) p.push_back(*out);
        return e;
    }
    ~DevAllocs() {
        for (void *q : p) (void)hipFree(q);
    }
};

int ensure(chs_ctx *c, chs_ctx::Buf &b, size_t bytes) {
    if (b.n >= bytes) return CHS_OK;
    if (b.p) HIPCHK(c, hipFree(b.p));
    b.p = nullptr;
    b.n = 0;
    HIPCHK(c, hipMalloc(&b.p, bytes));
    b.n = bytes;
    return CHS_OK;
}

template <typename T> T *buf(chs_ctx *c, int id) { return (T *)c->b[id].p; }

// ensure a spinor-sized buffer of `bytes_per_complex` for every listed id
int ensure_spinors(chs_ctx *c, std::initializer_list<int> ids, size_t bpc = 16) {
    for (int id : ids) {
        int rc = ensure(c, c->b[id], (size_t)c->kferm * bpc);
        if (rc) return rc;
    }
    return CHS_OK;
}

int ensure_partials(chs_ctx *c) {
    if (c->partial) return CHS_OK;
    const int np = (int)nblocks(c->kferm);   // the largest grid of any reduction
    HIPCHK(c, hipMalloc(&c->partial, sizeof(double) * np));
    HIPCHK(c, hipMalloc(&c->partial2, sizeof(double) * np));
    c->npart = np;
    return CHS_OK;
}

// Host spinors are in the reference order (s fastest inside a plane); device
// spinors in the native order (s outermost, chs_stencil.h).  Both directions
// go through a staging buffer and one relayout kernel.
int h2d_spinor(chs_ctx *c, void *dst, const void *host) {
    int rc = ensure(c, c->b[B_STAGE], 16 * (size_t)c->kferm);
    if (rc) return rc;
    c->h2d_spinor_bytes += 16ull * (unsigned long long)c->kferm;
    HIPCHK(c, hipMemcpyAsync(c->b[B_STAGE].p, host, 16 * c->kferm, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL((k_relayout<F64>), dim3((unsigned)((c->kferm + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, c->lat, (double2 *)dst, (const double2 *)c->b[B_STAGE].p, 0);
    HIPCHK(c, hipGetLastError());
    return CHS_OK;
}

int d2h_spinor(chs_ctx *c, void *host, const void *src) {
    int rc = ensure(c, c->b[B_STAGE], 16 * (size_t)c->kferm);
    if (rc) return rc;
    c->d2h_spinor_bytes += 16ull * (unsigned long long)c->kferm;
    hipLaunchKernelGGL((k_relayout<F64>), dim3((unsigned)((c->kferm + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, c->lat, (double2 *)c->b[B_STAGE].p, (const double2 *)src, 1);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(host, c->b[B_STAGE].p, 16 * c->kferm, hipMemcpyDeviceToHost, c->stream));
    return CHS_OK;
}

chs_ctx::Mirror *find_mirror(chs_ctx *c, const void *host) {
    for (auto &m : c->mirrors)
        if (m.host == host) return &m;
    return nullptr;
}

// The device copy (native layout) of a read-only host spinor argument: a
// registered mirror, uploaded only when the host array's generation has moved
// since its last upload, or else an upload into the work buffer `fallback`.
int in_spinor(chs_ctx *c, const double *host, int fallback, const double **dev) {
    if (chs_ctx::Mirror *m = find_mirror(c, host)) {
        if (m->dev_gen != m->gen) {
            int rc = h2d_spinor(c, m->dev, host);
            if (rc) return rc;
            m->dev_gen = m->gen;
        }
        *dev = (const double *)m->dev;
        return CHS_OK;
    }
    int rc = ensure(c, c->b[fallback], 16 * (size_t)c->kferm);
    if (rc) return rc;
    if ((rc = h2d_spinor(c, c->b[fallback].p, host))) return rc;
    *dev = (const double *)c->b[fallback].p;
    return CHS_OK;
}

size_t prec_bytes(int prec) { return prec == CHS_FP64 ? 16 : prec == CHS_FP32 ? 8 : 4; }
bool prec_ok(int prec) { return prec >= CHS_FP64 && prec <= CHS_FP16; }

// n complex numbers of storage precision prec at p <- hashed uniform values in
// [-0.5, 0.5) (timing inputs with realistic bit patterns; see k_fill_hash)
void fill_random(chs_ctx *c, void *p, int prec, long n, unsigned long long seed) {
    const dim3 g((unsigned)((n + kBlock - 1) / kBlock));
    if (prec == CHS_FP64) hipLaunchKernelGGL(k_fill_hash<F64>, g, dim3(kBlock), 0, c->stream, n, (double2 *)p, seed);
    else if (prec == CHS_FP32) hipLaunchKernelGGL(k_fill_hash<F32>, g, dim3(kBlock), 0, c->stream, n, (float2 *)p, seed);
    else if (prec == CHS_BF16) hipLaunchKernelGGL(k_fill_hash<B16>, g, dim3(kBlock), 0, c->stream, n, (uint32_t *)p, seed);
    else hipLaunchKernelGGL(k_fill_hash<H16>, g, dim3(kBlock), 0, c->stream, n, (uint32_t *)p, seed);
}
void fill_random_real(chs_ctx *c, double *p, long n, unsigned long long seed) {
    hipLaunchKernelGGL(k_fill_hash_real, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream, n,
                       p, seed);
}

// ---- launchers --------------------------------------------------------------
// FP32 storage on an even L uses the paired-site stencil (measured at L=24,
// Ls=32: FP32 17.5 -> 13.1 us; BF16 pairs, 8-byte loads at lower occupancy,
// were slower than its one-site stencil: 10.2 -> 11.3 us)
template <class S>
bool use_pairs(const chs_ctx *c) {
    return std::is_same<S, F32>::value && c->pair_stencil && (c->L % 2 == 0);
}

// workgroups (= partials) of the stencil launch dslash_launch<S> makes
template <class S>
unsigned dsl_grid(const chs_ctx *c) {
    return use_pairs<S>(c) ? stencil_grid_pair(c->n5) : stencil_grid(c->n5);
}

template <class S, class SA, bool DAG, int EPI>
void dslash_launch(chs_ctx *c, void *out, const void *in, const double *chi, double hpin, int imass,
                   EpiArgs ea) {
    if constexpr (std::is_same<S, F64>::value && EPI == EPI_STORE) {
        // the plain FP64 operator: xy-planes staged in LDS when a plane fits a workgroup
        if (c->plane_stencil && plane_stencil_fits(c->lat) && !(ea.skip)) {
            hipLaunchKernelGGL((k_dsl_plane<S, DAG>), dim3(plane_grid(c->lat)), dim3(plane_block(c->lat)),
                               4 * sizeof(typename S::st) * (size_t)c->L * c->L, c->stream, c->lat,
                               (typename S::st *)out, (const typename S::st *)in, chi, hpin, imass);
            return;
        }
    }
    if constexpr (std::is_same<S, F32>::value) {
        if (use_pairs<S>(c)) {
            hipLaunchKernelGGL((k_dsl_pair<S, SA, DAG, EPI>), dim3(stencil_grid_pair(c->n5)), dim3(kBlock), 0,
                               c->stream, c->lat, (typename S::st *)out, (const typename S::st *)in, chi, hpin,
                               imass, ea);
            return;
        }
    }
    hipLaunchKernelGGL((k_dsl<S, SA, DAG, EPI>), dim3(stencil_grid(c->n5)), dim3(kBlock), 0, c->stream, c->lat,
                       (typename S::st *)out, (const typename S::st *)in, chi, hpin, imass, ea);
}

EpiArgs epi(double *partial = nullptr, const double *scal = nullptr, void *a0 = nullptr, void *a1 = nullptr,
            void *a2 = nullptr, const int *skip = nullptr, int flag = 0, CgState *cgs = nullptr) {
    EpiArgs e;
    e.partial = partial;
    e.scal = scal;
    e.aux0 = a0;
    e.aux1 = a1;
    e.aux2 = a2;
    e.skip = skip;
    e.flag = flag;
    e.cgs = cgs;
    return e;
}

template <bool DAG>
void dslash_prec(chs_ctx *c, void *out, const void *in, const double *chi, double hpin, int imass, int prec) {
    if (prec == CHS_FP64) dslash_launch<F64, F64, DAG, EPI_STORE>(c, out, in, chi, hpin, imass, epi());
    else if (prec == CHS_FP32) dslash_launch<F32, F32, DAG, EPI_STORE>(c, out, in, chi, hpin, imass, epi());
    else if (prec == CHS_BF16) dslash_launch<B16, F32, DAG, EPI_STORE>(c, out, in, chi, hpin, imass, epi());
    else dslash_launch<H16, F32, DAG, EPI_STORE>(c, out, in, chi, hpin, imass, epi());
}

// |a|^2 (mode 0) / Re<a,b> (mode 1) / sum a^2 over reals (mode 2) into dst (device)
void reduce_launch(chs_ctx *c, long n, const void *a, const void *b, int mode, double *dst) {
    long nb = (n + kBlock - 1) / kBlock;
    if (nb > 2048) nb = 2048;
    nb = (nb + 7) / 8 * 8;
    hipLaunchKernelGGL(k_reduce, dim3((unsigned)nb), dim3(kBlock), 0, c->stream, n, (const double2 *)a,
                       (const double2 *)b, mode, c->partial2);
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kBlock), 0, c->stream, (const double *)c->partial2, (int)nb,
                       dst);
}

int check_imass(chs_ctx *c, int imass) {
    // the reference applies no wall term for other values (rhmc/cheisdwf.f:2912,2996); keep that
    (void)c;
    (void)imass;
    return CHS_OK;
}

int poll_int(chs_ctx *c, const int *dev, int *out) {
    HIPCHK(c, hipMemcpyAsync(c->hpinned, dev, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = *(int *)c->hpinned;
    return CHS_OK;
}

// how many solver iterations to enqueue between polls: about 2 ms of work
int chunk_for(chs_ctx *c, double bytes_per_iter) {
    double t_us = bytes_per_iter / 5.0e6;   // ~5 TB/s
    int ch = (int)(2000.0 / (t_us + 20.0));
    if (ch < 2) ch = 2;
    if (ch > 64) ch = 64;
    (void)c;
    return ch;
}

}  // namespace

// ============================================================================
extern "C" {

const char *chs_create_error(void) { return g_create_err.c_str(); }

chs_ctx *chs_ctx_create(int L, int Lt, int Ls, int ibound, double am3, int device) {
    if (L < 2 || Lt < 2 || Ls < 2) {
        g_create_err = "lattice extents must be >= 2 (L, Lt, Ls)";
        return nullptr;
    }
    long kvol = (long)L * L * Lt;
    if (8L * Ls * kvol > (1L << 31) - 1) {
        g_create_err = "lattice too large for 32-bit site indexing";
        return nullptr;
    }
    if (hipSetDevice(device) != hipSuccess) {
        g_create_err = "hipSetDevice failed (no GPU?)";
        return nullptr;
    }
    chs_ctx *c = new chs_ctx();
    c->L = L;
    c->Lt = Lt;
    c->Ls = Ls;
    c->ibound = ibound;
    c->am3 = am3;
    c->device = device;
    c->kvol = kvol;
    c->n5 = (long)Ls * kvol;
    c->kferm = 8 * c->n5;
    c->lat.L = L;
    c->lat.Lt = Lt;
    c->lat.Ls = Ls;
    c->lat.kvol = (int)kvol;
    c->lat.n5 = c->n5;
    c->lat.diag = 4.0 - am3;
    c->lat.apbc = (ibound == -1);
    c->lat.ps = c->n5;
    if (const char *g = getenv("CHS_GRAPHS")) c->use_graphs = atoi(g) != 0;
    if (const char *g = getenv("CHS_PAIRS")) c->pair_stencil = atoi(g) != 0;
    if (const char *g = getenv("CHS_PLANES")) c->plane_stencil = atoi(g) != 0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&c->lz, sizeof(LzState)) != hipSuccess || hipMalloc(&c->cg, sizeof(CgState)) != hipSuccess ||
        hipMalloc(&c->dscal, 64 * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->dcoef, 2 * (kMaxShift + 1) * sizeof(double)) != hipSuccess ||
        hipHostMalloc(&c->hpinned, 4096, hipHostMallocDefault) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        g_create_err = "HIP allocation failed in chs_ctx_create";
        chs_ctx_destroy(c);
        return nullptr;
    }
    c->scratch = scratch_address(device, c->scratch_err);
    if (ensure_partials(c) != CHS_OK) {
        g_create_err = c->err;
        chs_ctx_destroy(c);
        return nullptr;
    }
    return c;
}

void chs_ctx_destroy(chs_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto &b : c->b)
        if (b.p) hipFree(b.p);
    if (c->shifts_p.p) hipFree(c->shifts_p.p);
    if (c->shifts_x.p) hipFree(c->shifts_x.p);
    if (c->contrib.p) hipFree(c->contrib.p);
    for (auto &m : c->mirrors)
        if (m.dev) hipFree(m.dev);
    if (c->partial) hipFree(c->partial);
    if (c->partial2) hipFree(c->partial2);
    if (c->lz) hipFree(c->lz);
    if (c->cg) hipFree(c->cg);
    if (c->dscal) hipFree(c->dscal);
    if (c->dcoef) hipFree(c->dcoef);
    if (c->hpinned) hipHostFree(c->hpinned);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    nn_destroy(c);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char *chs_last_error(const chs_ctx *c) { return c ? c->err.c_str() : g_create_err.c_str(); }
long chs_kferm(const chs_ctx *c) { return c->kferm; }
long chs_kvol(const chs_ctx *c) { return c->kvol; }
void *chs_stream(chs_ctx *c) { return (void *)c->stream; }
int chs_synchronize(chs_ctx *c) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHS_OK;
}

void *chs_alloc(chs_ctx *c, size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        fail(c, CHS_ENOMEM, "hipMalloc(%zu) failed", bytes);
        return nullptr;
    }
    return p;
}
int chs_free(chs_ctx *c, void *p) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(p));
    return CHS_OK;
}
int chs_upload(chs_ctx *c, void *dst, const void *src, size_t bytes) {
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHS_OK;
}
int chs_download(chs_ctx *c, void *dst, const void *src, size_t bytes) {
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHS_OK;
}

int chs_convert_dev(chs_ctx *c, void *dst, int dp, const void *src, int sp, long n) {
    if (!prec_ok(dp) || !prec_ok(sp) || n < 0) return fail(c, CHS_EINVAL, "bad precision or size");
    dim3 g(nblocks(n)), b(kBlock);
#define CV(SD, SS)                                                                                           \
    hipLaunchKernelGGL((k_convert<SD, SS>), g, b, 0, c->stream, n, (typename SD::st *)dst,                  \
                       (const typename SS::st *)src)
    if (dp == CHS_FP64 && sp == CHS_FP64) CV(F64, F64);
    else if (dp == CHS_FP64 && sp == CHS_FP32) CV(F64, F32);
    else if (dp == CHS_FP64 && sp == CHS_BF16) CV(F64, B16);
    else if (dp == CHS_FP32 && sp == CHS_FP64) CV(F32, F64);
    else if (dp == CHS_FP32 && sp == CHS_FP32) CV(F32, F32);
    else if (dp == CHS_FP32 && sp == CHS_BF16) CV(F32, B16);
    else if (dp == CHS_BF16 && sp == CHS_FP64) CV(B16, F64);
    else if (dp == CHS_BF16 && sp == CHS_FP32) CV(B16, F32);
    else if (dp == CHS_BF16 && sp == CHS_BF16) CV(B16, B16);
    else if (dp == CHS_FP16 && sp == CHS_FP64) CV(H16, F64);
    else if (dp == CHS_FP16 && sp == CHS_FP32) CV(H16, F32);
    else if (dp == CHS_FP64 && sp == CHS_FP16) CV(F64, H16);
    else if (dp == CHS_FP32 && sp == CHS_FP16) CV(F32, H16);
    else if (dp == CHS_FP16 && sp == CHS_FP16) CV(H16, H16);
    else return fail(c, CHS_EINVAL, "no conversion between BF16 and FP16");
#undef CV
    HIPCHK(c, hipGetLastError());
    return CHS_OK;
}

int chs_layout_dev(chs_ctx *c, void *dst, const void *src, int prec, int to_native) {
    if (!prec_ok(prec)) return fail(c, CHS_EINVAL, "precision %d not in {0,1,2,3}", prec);
    if (dst == src) return fail(c, CHS_EINVAL, "layout conversion cannot work in place");
    const dim3 g((unsigned)((c->kferm + kBlock - 1) / kBlock)), b(kBlock);
    const int dir = to_native ? 0 : 1;
    if (prec == CHS_FP64)
        hipLaunchKernelGGL((k_relayout<F64>), g, b, 0, c->stream, c->lat, (double2 *)dst, (const double2 *)src, dir);
    else if (prec == CHS_FP32)
        hipLaunchKernelGGL((k_relayout<F32>), g, b, 0, c->stream, c->lat, (float2 *)dst, (const float2 *)src, dir);
    else   // BF16 and FP16: both move 32-bit words
        hipLaunchKernelGGL((k_relayout<B16>), g, b, 0, c->stream, c->lat, (uint32_t *)dst, (const uint32_t *)src,
                           dir);
    HIPCHK(c, hipGetLastError());
    return CHS_OK;
}

// ---- operators ------------------------------------------------------------------
int chs_dslash_dev(chs_ctx *c, void *out, const void *in, const double *chi, double hpin, int imass, int dagger,
                   int prec) {
    if (!prec_ok(prec)) return fail(c, CHS_EINVAL, "precision %d not in {0,1,2,3}", prec);
    if (out == in) return fail(c, CHS_EINVAL, "dslash output must not alias its input");
    if (dagger) dslash_prec<true>(c, out, in, chi, hpin, imass, prec);
    else dslash_prec<false>(c, out, in, chi, hpin, imass, prec);
    HIPCHK(c, hipGetLastError());
    return CHS_OK;
}

static int dslash_host(chs_ctx *c, double *phi, const double *r, const double *chi, double hpin, int imass,
                       int dagger) {
    int rc = ensure_spinors(c, {B_V});
    if (rc) return rc;
    rc = ensure(c, c->b[B_CHI], sizeof(double) * 3 * c->kvol);
    if (rc) return rc;
    const double *dr = nullptr;
    if ((rc = in_spinor(c, r, B_R, &dr))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->b[B_CHI].p, chi, 24 * c->kvol, hipMemcpyHostToDevice, c->stream));
    rc = chs_dslash_dev(c, c->b[B_V].p, dr, buf<double>(c, B_CHI), hpin, imass, dagger, CHS_FP64);
    if (rc) return rc;
    if ((rc = d2h_spinor(c, phi, c->b[B_V].p))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHS_OK;
}

int chs_dslash(chs_ctx *c, double *phi, const double *r, const double *chi, double am, double hpin, int imass) {
    (void)am;
    return dslash_host(c, phi, r, chi, hpin, imass, 0);
}
int chs_dslashd(chs_ctx *c, double *phi, const double *r, const double *chi, double am, double hpin, int imass) {
    (void)am;
    return dslash_host(c, phi, r, chi, hpin, imass, 1);
}

// ---- multi-shift Lanczos ------------------------------------------------------------
}  // extern "C"

namespace {

// Multi-shift Lanczos (rhmc/cheisdwf.f:599-745) with the spinors r, q, w, v,
// p_k (and x_k) stored in S.  Scalars, reductions and, for iflag = 0, the
// running rational combination x are FP64 whatever S is; FP32 storage halves
// the traffic (the force's solves; their error against FP64 is tested).
template <class S>
int qmrherm_impl(chs_ctx *c, const double *phi, double res, int *itercg, double hpin, int imass,
                 const double *anum, const double *aden, int ndiag, int iflag, const double *chi, double *x,
                 double *dsdchi) {
    typedef typename S::st st;
    if (ndiag < 1 || ndiag > kMaxShift) return fail(c, CHS_EINVAL, "ndiag=%d outside 1..%d", ndiag, kMaxShift);
    if (iflag == 1 && !dsdchi) return fail(c, CHS_EINVAL, "iflag=1 needs dsdchi");
    const long n = c->kferm;
    int rc = ensure_spinors(c, {B_R, B_Q0, B_Q1, B_V, B_W});
    if (rc) return rc;
    if ((rc = ensure(c, c->shifts_p, (size_t)ndiag * n * sizeof(st)))) return rc;
    // iflag = 0 accumulates the rational combination directly into x; only
    // the force (iflag = 1) needs every x_k; iflag = 2 (chs_qmrherm_shifts_dev)
    // writes every x_k, whole, into the caller's x
    if (iflag == 1 && (rc = ensure(c, c->shifts_x, (size_t)ndiag * n * sizeof(st)))) return rc;
    st *r = buf<st>(c, B_R), *q0 = buf<st>(c, B_Q0), *q1 = buf<st>(c, B_Q1);
    st *v = buf<st>(c, B_V), *w = buf<st>(c, B_W);
    st *pk = (st *)c->shifts_p.p;
    void *xk = iflag == 1 ? c->shifts_x.p : (void *)x;

    // scalar state: aden, ndiag; coefficients anum for the combine / force
    LzState h;
    memset(&h, 0, sizeof h);
    h.ndiag = ndiag;
    h.freeze = c->lz_freeze;
    h.amask = ndiag == 32 ? ~0u : (1u << ndiag) - 1u;
    for (int k = 0; k < ndiag; k++) h.aden[k] = aden[k];
    HIPCHK(c, hipMemcpyAsync(c->lz, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->dcoef, anum, sizeof(double) * (ndiag + 1), hipMemcpyHostToDevice, c->stream));

    // x = a0 Phi; r = Phi; q_{0} = 0; beta = phimod = |Phi|
    if (iflag != 2)
        hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, anum[0], (const double2 *)phi,
                           0.0, (double2 *)x);
    if (std::is_same<S, F64>::value)
        HIPCHK(c, hipMemcpyAsync(r, phi, 16 * n, hipMemcpyDeviceToDevice, c->stream));
    else
        hipLaunchKernelGGL((k_convert<S, F64>), dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, r,
                           (const double2 *)phi);
    HIPCHK(c, hipMemsetAsync(q1, 0, sizeof(st) * n, c->stream));
    reduce_launch(c, n, phi, nullptr, 0, c->dscal);
    hipLaunchKernelGGL(k_lz_init, dim3(1), dim3(kBlock), 0, c->stream, c->lz, (const double *)c->dscal, 1,
                       std::sqrt((double)n * res * res));

    const int niterc = c->lz_cap;
    const int np1 = (int)dsl_grid<S>(c), np2 = (int)nblocks(n);
    const int chunk = chunk_for(c, (12.0 + 2.0 * ndiag) * (double)sizeof(st) * n);
    const int *skip = &c->lz->done;
    // the force keeps its x_k only on the slices k_force_shift reads
    XKeep xkeep;
    if (iflag == 1 && c->Ls > 4) {
        xkeep.n5 = (unsigned)c->n5;
        xkeep.lo = 2u * (unsigned)c->kvol;
        xkeep.hi = (unsigned)(c->Ls - 2) * (unsigned)c->kvol;
    }
    int done = 0, iter = 1;
    while (iter <= niterc && !done) {
        const int last = std::min(niterc, iter + chunk - 1);
        for (; iter <= last; iter++) {
            st *qc = (iter & 1) ? q0 : q1, *qp = (iter & 1) ? q1 : q0;
            dslash_launch<S, S, false, EPI_SCALEQ>(c, v, r, chi, hpin, imass,
                                                  epi(nullptr, &c->lz->beta, qc, nullptr, nullptr, skip));
            dslash_launch<S, S, true, EPI_DOT>(c, w, v, chi, hpin, imass,
                                              epi(c->partial, nullptr, qc, nullptr, nullptr, skip));
            hipLaunchKernelGGL(k_lz_alpha, dim3(1), dim3(kBlock), 0, c->stream, c->lz, (const double *)c->partial,
                               np1, iter);
            const double *an = (const double *)c->dcoef;
            const LzState *lzc = c->lz;
#define LZ_UPD(NS, CMB)                                                                                         \
    hipLaunchKernelGGL((k_lz_update<S, NS, CMB>), dim3(np2), dim3(kBlock), 0, c->stream, n, (const st *)w,      \
                       (const st *)qc, (const st *)qp, r, pk, xk, an, lzc, ndiag, iter, c->partial2, xkeep)
            if (iflag == 0) {
                if (ndiag == 25) LZ_UPD(25, true);
                else if (ndiag == 12) LZ_UPD(12, true);
                else LZ_UPD(0, true);
            } else {
                if (ndiag == 25) LZ_UPD(25, false);
                else if (ndiag == 12) LZ_UPD(12, false);
                else LZ_UPD(0, false);
            }
#undef LZ_UPD
            hipLaunchKernelGGL(k_lz_beta, dim3(1), dim3(kBlock), 0, c->stream, c->lz, (const double *)c->partial2,
                               np2, iter);
        }
        HIPCHK(c, hipGetLastError());
        if ((rc = poll_int(c, &c->lz->done, &done))) return rc;
    }
    int it = 0;
    if ((rc = poll_int(c, &c->lz->itercg, &it))) return rc;
    if (itercg) *itercg = it;

    if (iflag == 1) {
        const size_t cb = (size_t)ndiag * 3 * c->kvol * sizeof(double);
        if ((rc = ensure(c, c->contrib, cb))) return rc;
        hipLaunchKernelGGL((k_force_shift<S>), dim3(nblocks((long)c->kvol * ndiag)), dim3(kBlock), 0, c->stream,
                           c->lat, (const st *)xk, n, ndiag, (const double *)c->dcoef, chi, hpin, imass,
                           (double *)c->contrib.p);
        hipLaunchKernelGGL(k_force_sum, dim3(nblocks(3 * c->kvol)), dim3(kBlock), 0, c->stream, (int)c->kvol,
                           ndiag, (const double *)c->contrib.p, dsdchi);
    }
    HIPCHK(c, hipGetLastError());
    // The reference returns silently after niterc iterations
    // (rhmc/cheisdwf.f:639) and never tests its residual for NaN; here both
    // are reported, with x and the force contributions written as the
    // reference would leave them.
    if (done == 0)
        return fail(c, CHS_ENOCONV, "multi-shift solve reached its cap of %d iterations without converging", niterc);
    if (done == 2)
        return fail(c, CHS_ENOCONV, "multi-shift solve: non-finite residual norm at iteration %d", it);
    return CHS_OK;
}

}  // namespace

extern "C" {

int chs_qmrherm_dev(chs_ctx *c, const double *phi, double res, int *itercg, double hpin, int imass,
                    const double *anum, const double *aden, int ndiag, int iflag, const double *chi, double *x,
                    double *dsdchi) {
    if (iflag != 0 && iflag != 1) return fail(c, CHS_EINVAL, "iflag=%d not 0 or 1", iflag);
    return qmrherm_impl<F64>(c, phi, res, itercg, hpin, imass, anum, aden, ndiag, iflag, chi, x, dsdchi);
}

int chs_qmrherm_prec_dev(chs_ctx *c, const double *phi, double res, int *itercg, double hpin, int imass,
                         const double *anum, const double *aden, int ndiag, int iflag, const double *chi, double *x,
                         double *dsdchi, int prec) {
    if (iflag != 0 && iflag != 1) return fail(c, CHS_EINVAL, "iflag=%d not 0 or 1", iflag);
    if (prec == CHS_FP32)
        return qmrherm_impl<F32>(c, phi, res, itercg, hpin, imass, anum, aden, ndiag, iflag, chi, x, dsdchi);
    if (prec == CHS_FP64)
        return qmrherm_impl<F64>(c, phi, res, itercg, hpin, imass, anum, aden, ndiag, iflag, chi, x, dsdchi);
    return fail(c, CHS_EINVAL, "multi-shift precision must be FP64 or FP32");
}

int chs_qmrherm_shifts_dev(chs_ctx *c, const double *phi, double res, int *itercg, double hpin, int imass,
                           const double *aden, int ndiag, const double *chi, double *xk) {
    if (ndiag < 1 || ndiag > kMaxShift) return fail(c, CHS_EINVAL, "ndiag=%d outside 1..%d", ndiag, kMaxShift);
    if (!xk) return fail(c, CHS_EINVAL, "xk is NULL");
    const std::vector<double> anum(ndiag + 1, 0.0);   // no combination is formed
    return qmrherm_impl<F64>(c, phi, res, itercg, hpin, imass, anum.data(), aden, ndiag, 2, chi, xk, nullptr);
}

int chs_set_shift_freeze(chs_ctx *c, double f) {
    if (!(f >= 0.0 && f <= 1.0)) return fail(c, CHS_EINVAL, "shift freezing factor %g outside [0, 1]", f);
    c->lz_freeze = f;
    return CHS_OK;
}

int chs_set_lanczos_cap(chs_ctx *c, int niterc) {
    if (niterc < 0) return fail(c, CHS_EINVAL, "multi-shift iteration cap %d < 0", niterc);
    c->lz_cap = niterc ? niterc : 5000;
    return CHS_OK;
}

int chs_mirror_register(chs_ctx *c, const void *host) {
    if (!host) return fail(c, CHS_EINVAL, "mirror of a NULL host array");
    if (find_mirror(c, host)) return CHS_OK;
    chs_ctx::Mirror m{host, nullptr, 1, 0};
    HIPCHK(c, hipMalloc(&m.dev, 16 * (size_t)c->kferm));
    c->mirrors.push_back(m);
    return CHS_OK;
}

int chs_mirror_touch(chs_ctx *c, const void *host) {
    chs_ctx::Mirror *m = find_mirror(c, host);
    if (!m) return fail(c, CHS_EINVAL, "no mirror registered for host array %p", host);
    m->gen++;
    return CHS_OK;
}

int chs_mirror_release(chs_ctx *c, const void *host) {
    for (size_t i = 0; i < c->mirrors.size(); i++) {
        if (c->mirrors[i].host != host) continue;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipFree(c->mirrors[i].dev));
        c->mirrors.erase(c->mirrors.begin() + (long)i);
        return CHS_OK;
    }
    return fail(c, CHS_EINVAL, "no mirror registered for host array %p", host);
}

int chs_transfer_bytes(const chs_ctx *c, unsigned long long *h2d, unsigned long long *d2h) {
    if (h2d) *h2d = c->h2d_spinor_bytes;
    if (d2h) *d2h = c->d2h_spinor_bytes;
    return CHS_OK;
}

int chs_set_force_precision(chs_ctx *c, int prec) {
    if (prec != CHS_FP64 && prec != CHS_FP32 && prec != CHS_FP16)
        return fail(c, CHS_EINVAL, "force precision must be FP64, FP32 or FP16");
    c->force_prec = prec;
    return CHS_OK;
}

int chs_qmrherm(chs_ctx *c, const double *phi, double res, int *itercg, double am, double hpin, int imass,
                const double *anum, const double *aden, int ndiag, int iflag, const double *chi, double *x,
                double *dsdchi) {
    (void)am;
    int rc = ensure_spinors(c, {B_X});
    if (rc) return rc;
    if ((rc = ensure(c, c->b[B_CHI], 24 * c->kvol))) return rc;
    if ((rc = ensure(c, c->b[B_F], 24 * c->kvol))) return rc;
    if (iflag == 1 && !dsdchi) return fail(c, CHS_EINVAL, "iflag=1 needs dsdchi");
    const double *dphi = nullptr;
    if ((rc = in_spinor(c, phi, B_PHI, &dphi))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->b[B_CHI].p, chi, 24 * c->kvol, hipMemcpyHostToDevice, c->stream));
    if (iflag == 1) HIPCHK(c, hipMemcpyAsync(c->b[B_F].p, dsdchi, 24 * c->kvol, hipMemcpyHostToDevice, c->stream));
    const int rs = chs_qmrherm_dev(c, dphi, res, itercg, hpin, imass, anum, aden, ndiag, iflag,
                                   buf<double>(c, B_CHI), buf<double>(c, B_X), buf<double>(c, B_F));
    if (rs != CHS_OK && rs != CHS_ENOCONV) return rs;
    // results are written also when the solve stopped at its cap
    const std::string msg = c->err;
    if (x && (rc = d2h_spinor(c, x, c->b[B_X].p))) return rc;
    if (iflag == 1) HIPCHK(c, hipMemcpyAsync(dsdchi, c->b[B_F].p, 24 * c->kvol, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->err = msg;
    return rs;
}

// ---- CGNR with the reference recurrences --------------------------------------
}  // extern "C"

namespace {

// CGNR (rhmc/cheisdwf.f:1064-1157) with its work spinors p, r, x1 = M p, the
// stencils' partial sums and the CG state in the buffers the caller passes:
// the context's (chs_congrad_dev) or the library's static scratch
// (chs_congrad_noalloc_dev).  partial holds at least nblocks(kferm) doubles.
int congrad_impl(chs_ctx *c, const double *phi, double res, int *itercg, double hpin, int imass, const double *chi,
                 double *x, double2 *p, double2 *r, double2 *x1, double *partial, CgState *cg) {
    const long n = c->kferm;
    int rc = CHS_OK;
    CgState h;
    memset(&h, 0, sizeof h);
    h.alpha = 1.0;
    h.betad = 1.0;
    h.resid = (double)n * res * res;
    h.maxiter = (int)(c->Ls * c->kvol);
    h.aslot = kAlphaEmpty;
    HIPCHK(c, hipMemcpyAsync(cg, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(p, x, 16 * n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(r, phi, 16 * n, hipMemcpyDeviceToDevice, c->stream));
    const int np = (int)stencil_grid(c->n5);
    const int *skip = &cg->halt;
    const int chunk = chunk_for(c, 11.0 * 16.0 * n);
    int halt = 0, nx = 1;
    while (nx <= h.maxiter && !halt) {
        const int last = std::min(h.maxiter, nx + chunk - 1);
        for (; nx <= last; nx++) {
            if (nx == 1) {
                dslash_launch<F64, F64, false, EPI_STORE>(c, x1, p, chi, hpin, imass,
                                                         epi(nullptr, nullptr, nullptr, nullptr, nullptr, skip));
                dslash_launch<F64, F64, true, EPI_CG>(c, nullptr, x1, chi, hpin, imass,
                                                     epi(partial, &cg->alpha, p, x, r, skip, 1));
            } else {
                // alpha = alphan / |M p|^2 (alphan = |r|^2 = rnorm2) inside the
                // fused stencil, from the M p launch's partials
                dslash_launch<F64, F64, false, EPI_NORM>(c, x1, p, chi, hpin, imass,
                                                        epi(partial, nullptr, nullptr, nullptr, nullptr, skip));
                dslash_launch<F64, F64, true, EPI_CGW>(c, nullptr, x1, chi, hpin, imass,
                                                      epi(partial, nullptr, p, x, r, skip, 0, cg));
            }
            hipLaunchKernelGGL(k_cg_beta_ref, dim3(1), dim3(kBlock), 0, c->stream, cg, (const double *)partial,
                               np, nx);
            hipLaunchKernelGGL((k_cg_pupdate<F64, F64>), dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, p,
                               (const double2 *)r, (const CgState *)cg);
        }
        HIPCHK(c, hipGetLastError());
        if ((rc = poll_int(c, &cg->halt, &halt))) return rc;
    }
    CgState fin;
    HIPCHK(c, hipMemcpyAsync(&fin, cg, sizeof fin, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (itercg) *itercg = fin.iter;
    if (fin.aerr) return fail(c, CHS_EHIP, "congrad: alpha not published within the wait bound (iteration %d)",
                              fin.iter);
    if (!(fin.rnorm2 < fin.resid))
        return fail(c, CHS_ENOCONV, "# iterations of congrad exceeds niterc (%d)", fin.iter);
    return CHS_OK;
}

}  // namespace

extern "C" {

int chs_congrad_dev(chs_ctx *c, const double *phi, double res, int *itercg, double hpin, int imass,
                    const double *chi, double *x) {
    int rc = ensure_spinors(c, {B_P, B_R, B_T1});
    if (rc) return rc;
    return congrad_impl(c, phi, res, itercg, hpin, imass, chi, x, buf<double2>(c, B_P), buf<double2>(c, B_R),
                        buf<double2>(c, B_T1), c->partial, c->cg);
}

// The same CGNR with every buffer it writes besides x in the static scratch
// array (layout: CgState at 0, the partial sums at kNoAllocPartialOff, then
// p, r and M p, 16*kferm bytes each).  Nothing is allocated and nothing is
// freed; the scratch keeps the last call's contents, and every field the
// solve reads is written by it first (the CG state by one host-to-device
// copy, p and r from x and phi, the partials by the launch before the one
// that reads them).
int chs_congrad_noalloc_dev(chs_ctx *c, const double *phi, double res, int *itercg, double hpin, int imass,
                            const double *chi, double *x) {
    if (c->kferm > CHS_NOALLOC_MAX_KFERM)
        return fail(c, CHS_EINVAL,
                    "congrad without allocation: kferm = %ld exceeds the static scratch's %ld (%zu bytes)", c->kferm,
                    (long)CHS_NOALLOC_MAX_KFERM, kNoAllocBytes);
    if (!c->scratch) return fail(c, CHS_EHIP, "congrad without allocation: %s", c->scratch_err.c_str());
    unsigned char *base = (unsigned char *)c->scratch;
    double2 *p = (double2 *)(base + kNoAllocSpinorOff);
    double2 *r = p + c->kferm, *x1 = r + c->kferm;
    return congrad_impl(c, phi, res, itercg, hpin, imass, chi, x, p, r, x1, (double *)(base + kNoAllocPartialOff),
                        (CgState *)base);
}


// fermion (rhmc/cheisdwf.f:897-1058): five wall sources at t-slices
// mod(ittt0 + k (Lt/5), Lt) + 1, Dirac 3 and 4, flavour 1, 1/L^2 at s = 1;
// x = (M^dagger M)^{-1} M^dagger source from the initial guess M^dagger
// source (congrad, or the FP32 / scaled-FP16 reliable-update solve to the same
// residual);
// correlators from the t-slice sums of x at s = Ls (cferm1, Dirac = source)
// and s = 1 (cferm2, gamma_3 partner), with the antiperiodic sign for slices
// behind the source.
// appends (phi, x, chi) to the learned preconditioner's pair store while it
// has room (chs_learn_host.inc)
static int nn_store_pair(chs_ctx *c, const double2 *phi, const double2 *x, const double *chi);

int chs_fermion(chs_ctx *c, const double *chi, double res, double hpin, int imass, int ittt0, int prec,
                double *cferm1, double *cferm2, int *iterations) {
    if (ittt0 < 1 || ittt0 > c->Lt) return fail(c, CHS_EINVAL, "ittt0=%d outside 1..%d", ittt0, c->Lt);
    if (prec != CHS_FP64 && prec != CHS_FP32 && prec != CHS_FP16 && prec != CHS_LEARNED)
        return fail(c, CHS_EINVAL, "fermion solver form must be FP64, FP32, FP16 or LEARNED");
    if (prec == CHS_LEARNED && !c->nn) return fail(c, CHS_EINVAL, "learned form needs chs_nn_init first");
    const int nsource = 5, Lt = c->Lt, Ls = c->Ls;
    const long n = c->kferm, L2 = (long)c->L * c->L;
    int rc = ensure_spinors(c, {B_PHI, B_X, B_Q0});
    if (rc) return rc;
    if ((rc = ensure(c, c->b[B_CHI], 24 * c->kvol))) return rc;
    if ((rc = ensure(c, c->b[B_PP], std::max<size_t>(24 * c->kvol, 64 * (size_t)Lt)))) return rc;
    double2 *xi = buf<double2>(c, B_Q0), *phi = buf<double2>(c, B_PHI), *x = buf<double2>(c, B_X);
    double2 *sums = buf<double2>(c, B_PP);
    double *dchi = buf<double>(c, B_CHI);
    HIPCHK(c, hipMemcpyAsync(dchi, chi, 24 * c->kvol, hipMemcpyHostToDevice, c->stream));
    std::vector<double> c1(4 * Lt, 0.0), c2(4 * Lt, 0.0);   // [f][t] (re, im)
    std::vector<double2> S[2];
    int iter = 0, ncap = 0;
    for (int ks = 1; ks <= nsource; ks++) {
        const int ittt = (ittt0 + ks * (Lt / nsource)) % Lt + 1;
        for (int ids = 0; ids < 2; ids++) {
            HIPCHK(c, hipMemsetAsync(xi, 0, 16 * n, c->stream));
            hipLaunchKernelGGL(k_wall_source, dim3((unsigned)((L2 + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                               c->stream, c->lat, xi, ittt - 1, 2 + ids, 1.0 / (double)L2);
            dslash_prec<true>(c, phi, xi, dchi, hpin, imass, CHS_FP64);
            HIPCHK(c, hipMemcpyAsync(x, phi, 16 * n, hipMemcpyDeviceToDevice, c->stream));
            int it = 0;
            if (prec == CHS_FP64) {
                rc = chs_congrad_dev(c, (const double *)phi, res, &it, hpin, imass, dchi, (double *)x);
            } else {
                // the reference's stopping rule |r|^2 < kferm res^2 as a relative tolerance
                reduce_launch(c, n, phi, nullptr, 0, c->dscal);
                double b2 = 0.0;
                HIPCHK(c, hipMemcpyAsync(&b2, c->dscal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
                HIPCHK(c, hipStreamSynchronize(c->stream));
                chs_solve_stats st;
                const double tol = std::sqrt((double)n * res * res / b2);
                if (prec == CHS_LEARNED)
                    rc = chs_solve_mdagm_fcg_dev(c, (const double *)phi, (double *)x, dchi, hpin, imass, tol, 0.0,
                                                 0, 1, &st);
                else
                    rc = chs_solve_mdagm_dev(c, (const double *)phi, (double *)x, dchi, hpin, imass, tol, prec, 0.0,
                                             0, &st);
                it = st.iterations;
            }
            if (rc == CHS_ENOCONV) ncap++;   // the reference carries on with the last iterate
            else if (rc) return rc;
            if (prec != CHS_LEARNED && (rc = nn_store_pair(c, phi, x, dchi))) return rc;
            iter += it;
            SliceReq rq;
            rq.s[0] = Ls - 1, rq.plane[0] = 2 + ids;       // cferm1, flavour 1
            rq.s[1] = Ls - 1, rq.plane[1] = 2 + ids + 4;   // cferm1, flavour 2
            rq.s[2] = 0, rq.plane[2] = ids;                 // cferm2 (gamin(3, 3 + ids) = 1 + ids)
            rq.s[3] = 0, rq.plane[3] = ids + 4;
            hipLaunchKernelGGL(k_slice_sums, dim3(Lt, 4), dim3(kBlock), 0, c->stream, c->lat, (const double2 *)x,
                               rq, sums);
            S[ids].resize(4 * Lt);
            HIPCHK(c, hipMemcpyAsync(S[ids].data(), sums, 16 * 4 * Lt, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        // cferm1 += sign akappa prop0L(idd, idd); cferm2 += sign gamval(3, idd)
        // prop00(idd, gamin(3, idd)), gamval(3, 3) = i/2, gamval(3, 4) = -i/2
        for (int idd = 0; idd < 2; idd++) {
            const double gi = idd == 0 ? 0.5 : -0.5;
            for (int it = 0; it < Lt; it++) {
                const int itt = (ittt + it - 1) % Lt + 1;
                const double sgn = itt >= ittt ? 1.0 : (double)c->ibound;
                for (int f = 0; f < 2; f++) {
                    const double2 a = S[idd][(size_t)f * Lt + itt - 1], b = S[idd][(size_t)(2 + f) * Lt + itt - 1];
                    double *p1 = &c1[2 * (it + (size_t)Lt * f)], *p2 = &c2[2 * (it + (size_t)Lt * f)];
                    p1[0] += sgn * 0.5 * a.x;
                    p1[1] += sgn * 0.5 * a.y;
                    p2[0] += sgn * (-gi * b.y);
                    p2[1] += sgn * (gi * b.x);
                }
            }
        }
    }
    for (int k = 0; k < 4 * Lt; k++) {
        cferm1[k] = c1[k] / nsource;
        cferm2[k] = c2[k] / nsource;
    }
    if (iterations) *iterations = iter;
    // the message begins with the number of solves at the cap: the reference
    // writes its congrad message once for each (rhmc/cheisdwf.f:1151-1153)
    if (ncap) return fail(c, CHS_ENOCONV, "%d of the %d propagator solves exceeded the iteration cap", ncap, 2 * nsource);
    return CHS_OK;
}

int chs_congrad(chs_ctx *c, const double *phi, double res, int *itercg, double am, double hpin, int imass,
                const double *chi, double *x) {
    (void)am;
    int rc = ensure_spinors(c, {B_X});
    if (rc) return rc;
    if ((rc = ensure(c, c->b[B_CHI], 24 * c->kvol))) return rc;
    const double *dphi = nullptr;
    if ((rc = in_spinor(c, phi, B_PHI, &dphi))) return rc;
    if ((rc = h2d_spinor(c, c->b[B_X].p, x))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->b[B_CHI].p, chi, 24 * c->kvol, hipMemcpyHostToDevice, c->stream));
    int rc2 = chs_congrad_dev(c, dphi, res, itercg, hpin, imass, buf<double>(c, B_CHI), buf<double>(c, B_X));
    if (rc2 != CHS_OK && rc2 != CHS_ENOCONV) return rc2;
    const std::string msg = c->err;
    if ((rc = d2h_spinor(c, x, c->b[B_X].p))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (rc2) c->err = msg;
    return rc2;
}

// ---- force, derivs, hamilton ------------------------------------------------
int chs_force(chs_ctx *c, const double *phi, double res1, double am, double hpin, int imass, const double *chi,
              double beta, const double *anum, const double *aden, const double *bnum, const double *bden, int ndiag,
              double *dsdchi, double *x, int *its) {
    (void)am;
    int rc = ensure_spinors(c, {B_X, B_T2});
    if (rc) return rc;
    if ((rc = ensure(c, c->b[B_CHI], 24 * c->kvol))) return rc;
    if ((rc = ensure(c, c->b[B_F], 24 * c->kvol))) return rc;
    double *dx = buf<double>(c, B_X), *dx2 = buf<double>(c, B_T2);
    double *dchi = buf<double>(c, B_CHI), *dF = buf<double>(c, B_F);
    const double *dphi = nullptr;
    if ((rc = in_spinor(c, phi, B_PHI, &dphi))) return rc;
    HIPCHK(c, hipMemcpyAsync(dchi, chi, 24 * c->kvol, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_force_init, dim3(nblocks(3 * c->kvol)), dim3(kBlock), 0, c->stream, 3 * c->kvol, beta,
                       (const double *)dchi, dF);
    int i1 = 0, i2 = 0;
    // X = r_b(M^dagger M) Phi with the remez2g b-set, then the force from the a-set
    // (rhmc/cheisdwf.f:500-513)
    // (solver storage per the context's force precision policy)
    // (CHS_FP16 runs the multi-shift solves in FP32: shifted systems have no
    // reliable updates, and 16-bit storage does not converge there)
    auto ms = c->force_prec != CHS_FP64 ? qmrherm_impl<F32> : qmrherm_impl<F64>;
    // a solve that stops at its cap (or on a non-finite residual) does not
    // stop the force: the reference carries on, and the status is returned
    // with the results written
    std::string msg;
    int rs = ms(c, dphi, res1, &i1, hpin, imass, bnum, bden, ndiag, 0, dchi, dx, nullptr);
    if (rs != CHS_OK && rs != CHS_ENOCONV) return rs;
    if (rs) msg = c->err;
    HIPCHK(c, hipMemcpyAsync(dx2, dx, 16 * c->kferm, hipMemcpyDeviceToDevice, c->stream));
    rc = ms(c, dx2, res1, &i2, hpin, imass, anum, aden, ndiag, 1, dchi, dx, dF);
    if (rc != CHS_OK && rc != CHS_ENOCONV) return rc;
    if (rc && !rs) {
        rs = rc;
        msg = c->err;
    }
    HIPCHK(c, hipMemcpyAsync(dsdchi, dF, 24 * c->kvol, hipMemcpyDeviceToHost, c->stream));
    if (x && (rc = d2h_spinor(c, x, dx))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (its) {
        its[0] = i1;
        its[1] = i2;
    }
    if (rs) c->err = msg;
    return rs;
}

int chs_derivs(chs_ctx *c, const double *r, const double *x2, double anum, int imass, double *dsdchi) {
    int rc = ensure(c, c->b[B_F], 24 * c->kvol);
    if (rc) return rc;
    const double *dr = nullptr, *dx2 = nullptr;
    if ((rc = in_spinor(c, r, B_T1, &dr))) return rc;
    if ((rc = in_spinor(c, x2, B_T2, &dx2))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->b[B_F].p, dsdchi, 24 * c->kvol, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_derivs, dim3(nblocks(c->kvol)), dim3(kBlock), 0, c->stream, c->lat, (const double2 *)dr,
                       (const double2 *)dx2, anum, imass, buf<double>(c, B_F));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(dsdchi, c->b[B_F].p, 24 * c->kvol, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHS_OK;
}

// HMC sibling (mh/cheisdwf_hmc.f:393-557): X1 = (M^dagger M)^{-1} Phi by
// congrad from the guess x0 (hpin = 0: the sibling's operator has no pinning
// term), dSdchi = 2 (2 beta chi) + derivs(X1, M X1) with a = -1, and the
// linear chronological guess x0 <- 2 X1 - x0 for the next call.
int chs_hmc_force(chs_ctx *c, const double *phi, double *x0, double res, double am, int imass, const double *chi,
                  double beta, double *dsdchi, int *itercg) {
    (void)am;
    int rc = ensure_spinors(c, {B_X, B_Q0, B_T2});
    if (rc) return rc;
    if ((rc = ensure(c, c->b[B_CHI], 24 * c->kvol))) return rc;
    if ((rc = ensure(c, c->b[B_F], 24 * c->kvol))) return rc;
    const long n = c->kferm;
    double *dx1 = buf<double>(c, B_X), *dx0 = buf<double>(c, B_Q0);
    double *dx2 = buf<double>(c, B_T2), *dchi = buf<double>(c, B_CHI), *dF = buf<double>(c, B_F);
    const double *dphi = nullptr;
    if ((rc = in_spinor(c, phi, B_PHI, &dphi))) return rc;
    if ((rc = h2d_spinor(c, dx0, x0))) return rc;
    HIPCHK(c, hipMemcpyAsync(dchi, chi, 24 * c->kvol, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dx1, dx0, 16 * n, hipMemcpyDeviceToDevice, c->stream));
    int it = 0;
    if (c->force_prec == CHS_FP64) {
        rc = chs_congrad_dev(c, dphi, res, &it, 0.0, imass, dchi, dx1);
    } else {
        // the reliable-update CG (FP32, or binary16 search directions) from
        // the same guess, to the congrad rule |r|^2 < kferm res^2 taken as an
        // FP64 true relative residual
        reduce_launch(c, n, (const double2 *)dphi, nullptr, 0, c->dscal);
        double b2 = 0.0;
        HIPCHK(c, hipMemcpyAsync(&b2, c->dscal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const double tol = b2 > 0.0 ? std::sqrt((double)n * res * res / b2) : 1.0;   // Phi = 0: x = 0
        chs_solve_stats st;
        rc = chs_solve_mdagm_dev(c, dphi, dx1, dchi, 0.0, imass, tol, c->force_prec, 0.0, 0, &st);
        it = st.iterations;
    }
    if (rc && rc != CHS_ENOCONV) return rc;   // the sibling, too, carries on at its iteration cap
    const bool capped = rc == CHS_ENOCONV;
    // x0 = 2 X1 - x0 (mh/cheisdwf_hmc.f:434-438)
    hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, 2.0, (const double2 *)dx1, -1.0,
                       (double2 *)dx0);
    dslash_prec<false>(c, dx2, dx1, dchi, 0.0, imass, CHS_FP64);
    hipLaunchKernelGGL(k_force_init, dim3(nblocks(3 * c->kvol)), dim3(kBlock), 0, c->stream, 3 * c->kvol, 2.0 * beta,
                       (const double *)dchi, dF);
    hipLaunchKernelGGL(k_derivs, dim3(nblocks(c->kvol)), dim3(kBlock), 0, c->stream, c->lat, (const double2 *)dx1,
                       (const double2 *)dx2, -1.0, imass, dF);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(dsdchi, dF, 24 * c->kvol, hipMemcpyDeviceToHost, c->stream));
    if ((rc = d2h_spinor(c, x0, dx0))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (itercg) *itercg = it;
    // the force and x0 are written; the sibling writes its congrad message
    // and carries on (mh/cheisdwf_hmc.f:705-707)
    if (capped) return fail(c, CHS_ENOCONV, "1 of the 1 force solves exceeded the iteration cap");
    return CHS_OK;
}

// HMC sibling hamilton (mh/cheisdwf_hmc.f:561-620): hg = 2 beta sum chi^2,
// hp = sum p^2 / 2, hf = Re Phi^dagger X1 with X1 by congrad from x0, which
// receives X1.  h4 = (h, hg, hp, s = hg + hf).
int chs_hmc_hamilton(chs_ctx *c, const double *phi, double *x0, double res, double am, int imass, const double *chi,
                     const double *pp, double beta, double *h4, int *itercg) {
    (void)am;
    int rc = ensure_spinors(c, {B_X});
    if (rc) return rc;
    if ((rc = ensure(c, c->b[B_CHI], 24 * c->kvol))) return rc;
    if ((rc = ensure(c, c->b[B_PP], 24 * c->kvol))) return rc;
    double *dx = buf<double>(c, B_X), *dchi = buf<double>(c, B_CHI);
    double *dpp = buf<double>(c, B_PP);
    const double *dphi = nullptr;
    if ((rc = in_spinor(c, phi, B_PHI, &dphi))) return rc;
    if ((rc = h2d_spinor(c, dx, x0))) return rc;
    HIPCHK(c, hipMemcpyAsync(dchi, chi, 24 * c->kvol, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dpp, pp, 24 * c->kvol, hipMemcpyHostToDevice, c->stream));
    reduce_launch(c, 3 * c->kvol, dpp, nullptr, 2, c->dscal + 8);
    reduce_launch(c, 3 * c->kvol, dchi, nullptr, 2, c->dscal + 9);
    int it = 0;
    rc = chs_congrad_dev(c, dphi, res, &it, 0.0, imass, dchi, dx);
    if (rc && rc != CHS_ENOCONV) return rc;
    const bool capped = rc == CHS_ENOCONV;
    reduce_launch(c, c->kferm, dphi, dx, 1, c->dscal + 10);
    double sums[3];
    HIPCHK(c, hipMemcpyAsync(sums, c->dscal + 8, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if ((rc = d2h_spinor(c, x0, dx))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double hp = 0.5 * sums[0], hg = 2.0 * beta * sums[1], hf = sums[2];
    h4[0] = hg + hp + hf;
    h4[1] = hg;
    h4[2] = hp;
    h4[3] = hg + hf;
    if (itercg) *itercg = it;
    if (capped) return fail(c, CHS_ENOCONV, "1 of the 1 Hamiltonian solves exceeded the iteration cap");
    return CHS_OK;
}

int chs_hamilton(chs_ctx *c, const double *phi, double *h, double *hg, double *hp, double *s, double res2,
                 int isweep, double am, double hpin, int imass, const double *chi, const double *pp, double beta,
                 const double *bnum, const double *bden, int ndiag, double *x, int *itercg) {
    (void)am;
    (void)isweep;
    int rc = ensure_spinors(c, {B_X});
    if (rc) return rc;
    if ((rc = ensure(c, c->b[B_CHI], 24 * c->kvol))) return rc;
    if ((rc = ensure(c, c->b[B_PP], 24 * c->kvol))) return rc;
    double *dx = buf<double>(c, B_X), *dchi = buf<double>(c, B_CHI);
    double *dpp = buf<double>(c, B_PP);
    const double *dphi = nullptr;
    if ((rc = in_spinor(c, phi, B_PHI, &dphi))) return rc;
    HIPCHK(c, hipMemcpyAsync(dchi, chi, 24 * c->kvol, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dpp, pp, 24 * c->kvol, hipMemcpyHostToDevice, c->stream));
    // sums of p^2 and chi^2 (rhmc/cheisdwf.f:546-559), FP64 on the device
    reduce_launch(c, 3 * c->kvol, dpp, nullptr, 2, c->dscal + 8);
    reduce_launch(c, 3 * c->kvol, dchi, nullptr, 2, c->dscal + 9);
    int it = 0;
    const int rs = chs_qmrherm_dev(c, dphi, res2, &it, hpin, imass, bnum, bden, ndiag, 0, dchi, dx, nullptr);
    if (rs != CHS_OK && rs != CHS_ENOCONV) return rs;
    const std::string msg = c->err;
    reduce_launch(c, c->kferm, dphi, dx, 1, c->dscal + 10);   // Re Phi^dagger X (rhmc/cheisdwf.f:578-580)
    double sums[3];
    HIPCHK(c, hipMemcpyAsync(sums, c->dscal + 8, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (x && (rc = d2h_spinor(c, x, dx))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double vhp = 0.5 * sums[0], vhg = beta * sums[1], vhf = sums[2];
    *hp = vhp;
    *hg = vhg;
    *h = vhg + vhp + vhf;
    *s = vhg + vhf;
    if (itercg) *itercg = it;
    if (rs) c->err = msg;
    return rs;
}

}  // extern "C"

// ---- M^dagger M solves to an FP64 true residual ----------------------------------
namespace {

// w = M^dagger M x in FP64 using t as scratch
void mdagm64(chs_ctx *c, double2 *w, const double2 *x, double2 *t, const double *chi, double hpin, int imass) {
    dslash_launch<F64, F64, false, EPI_STORE>(c, t, x, chi, hpin, imass, epi());
    dslash_launch<F64, F64, true, EPI_STORE>(c, w, t, chi, hpin, imass, epi());
}

template <class S, class SA>
int rcg_solve(chs_ctx *c, const double *bvec, double *xvec, const double *chi, double hpin, int imass, double tol,
              double delta, int maxiter, chs_solve_stats *stats) {
    const long n = c->kferm;
    const size_t bs = sizeof(typename S::st), ba = sizeof(typename SA::st);
    int rc;
    if ((rc = ensure_spinors(c, {B_T1, B_T2, B_R}))) return rc;   // FP64 scratch + true residual
    if ((rc = ensure(c, c->b[B_PS], n * bs))) return rc;         // p of even iterations (dslash input)
    if ((rc = ensure(c, c->b[B_PS2], n * bs))) return rc;        // p of odd iterations
    if ((rc = ensure(c, c->b[B_V], n * bs))) return rc;          // M p
    if ((rc = ensure(c, c->b[B_XS], n * ba))) return rc;         // low-precision x accumulator
    if ((rc = ensure(c, c->b[B_RS], n * ba))) return rc;         // iterated residual
    double2 *t1 = buf<double2>(c, B_T1), *t2 = buf<double2>(c, B_T2), *r64 = buf<double2>(c, B_R);
    void *ps = c->b[B_PS].p, *mp = c->b[B_V].p, *xs = c->b[B_XS].p, *rs = c->b[B_RS].p;
    void *const pbuf[2] = {ps, c->b[B_PS2].p};   // p_k lives in pbuf[k % 2]
    const double2 *b = (const double2 *)bvec;
    double2 *x = (double2 *)xvec;
    const unsigned gn = nblocks(n);
    const int np = (int)dsl_grid<S>(c);

    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    reduce_launch(c, n, b, nullptr, 0, c->dscal);
    // true residual r64 = b - M^dagger M x, rs = r64
    mdagm64(c, t2, x, t1, chi, hpin, imass);
    hipLaunchKernelGGL((k_true_resid<SA>), dim3(gn), dim3(kBlock), 0, c->stream, n, b, (const double2 *)t2, r64,
                       (typename SA::st *)rs, c->partial2);
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kBlock), 0, c->stream, (const double *)c->partial2, (int)gn,
                       c->dscal + 1);
    double s2[2];
    HIPCHK(c, hipMemcpyAsync(s2, c->dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double bnorm2 = s2[0];
    double rn2 = s2[1];
    chs_solve_stats st;
    memset(&st, 0, sizeof st);
    const double target2 = tol * tol * bnorm2;
    if (bnorm2 == 0.0) {
        HIPCHK(c, hipMemsetAsync(x, 0, 16 * n, c->stream));
        rn2 = 0.0;
    }
    int status = CHS_OK;
    if (rn2 > target2) {
        CgState h;
        memset(&h, 0, sizeof h);
        h.rnorm2 = rn2;
        h.maxr2 = rn2;
        h.delta2 = delta * delta;
        double stop_fac = 0.5;
        h.stop2 = stop_fac * target2;
        h.maxiter = maxiter;
        h.aslot = kAlphaEmpty;
        HIPCHK(c, hipMemsetAsync(xs, 0, n * ba, c->stream));
        if constexpr (S::scaled) {
            // p = rs / sigma, sigma = rms of the residual (p starts at zero so
            // that the update's p term adds nothing)
            h.scaled = 1;
            h.nelem = (double)n;
            h.sigma = std::sqrt(rn2 / (double)n);
            h.xp_r = 1.0 / h.sigma;
            h.xp_p = 0.0;
            HIPCHK(c, hipMemsetAsync(ps, 0, n * bs, c->stream));
            HIPCHK(c, hipMemcpyAsync(c->cg, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
            hipLaunchKernelGGL((k_cg_pupdate<S, SA>), dim3(gn), dim3(kBlock), 0, c->stream, n,
                               (typename S::st *)ps, (const typename SA::st *)rs, (const CgState *)c->cg);
        } else {
            // p = rs
            hipLaunchKernelGGL((k_convert<S, SA>), dim3(gn), dim3(kBlock), 0, c->stream, n, (typename S::st *)ps,
                               (const typename SA::st *)rs);
            HIPCHK(c, hipMemcpyAsync(c->cg, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
        }
        const int *skip = &c->cg->halt;
        // an even number of iterations per chunk: a chunk starts at an even
        // iteration, so its p buffers are the same in every replay
        const int chunk = std::max(2, chunk_for(c, 10.0 * bs * n) & ~1);
        typedef typename S::st sst;
        typedef typename SA::st ast;
        // iteration k (parity par = k % 2): four launches.  The fused CG
        // stencil forms alpha itself (its workgroup 0, from the M p launch's
        // partials; the stored M^dagger M p is multiplied by alpha, or alpha
        // sigma for scaled storage) and applies r -= alpha M^dagger M p; the
        // x increments are in the p update
        auto enqueue_iter = [&](int par) {
            void *pc = pbuf[par], *po = pbuf[par ^ 1];
            dslash_launch<S, SA, false, EPI_NORM>(c, mp, pc, chi, hpin, imass,
                                                 epi(c->partial, nullptr, nullptr, nullptr, nullptr, skip));
            dslash_launch<S, SA, true, EPI_CGW>(c, nullptr, mp, chi, hpin, imass,
                                               epi(c->partial, nullptr, pc, xs, rs, skip, 1, c->cg));
            hipLaunchKernelGGL(k_rcg_beta, dim3(1), dim3(kBlock), 0, c->stream, c->cg, (const double *)c->partial,
                               np);
            if (par == 0)
                hipLaunchKernelGGL((k_cg_pnext<S, SA>), dim3(gn), dim3(kBlock), 0, c->stream, n, (const sst *)pc,
                                   (sst *)po, (const ast *)rs, (const CgState *)c->cg);
            else
                hipLaunchKernelGGL((k_cg_xpnext<S, SA>), dim3(gn), dim3(kBlock), 0, c->stream, n, (const sst *)pc,
                                   (sst *)po, (const ast *)rs, (ast *)xs, (const CgState *)c->cg);
        };
        // one chunk, replayed from a hipGraph (the pointers are fixed)
        auto enqueue_chunk = [&]() {
            for (int k = 0; k < chunk; k += 2) {
                enqueue_iter(0);
                enqueue_iter(1);
            }
        };
        // a solve whose true residual has not improved on its best in kStall
        // reliable updates in a row, restarts included, is diverging: it stops
        // with CHS_ENOCONV instead of running to maxiter.  Converging forms
        // gain ~1/delta^2 in |r|^2 per update.
        constexpr int kStall = 8;
        double best_rn2 = rn2;
        int stall = 0;
        GraphGuard gg;
        if (c->use_graphs) {
            hipGraph_t graph;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            enqueue_chunk();
            HIPCHK(c, hipStreamEndCapture(c->stream, &graph));
            const hipError_t e = hipGraphInstantiate(&gg.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIPCHK(c, e);
        }
        for (;;) {
            // after a reliable update at an even iteration the next one is
            // odd: run it alone so that the chunk starts even again
            if (h.iter & 1) enqueue_iter(1);
            if (gg.exec) HIPCHK(c, hipGraphLaunch(gg.exec, c->stream));
            else enqueue_chunk();
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipMemcpyAsync(&h, c->cg, sizeof h, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if (h.aerr) {
                status = fail(c, CHS_EHIP, "CG stencil: alpha not published within the wait bound (iteration %d)",
                              h.iter);
                break;
            }
            if (!h.hold) continue;
            // iteration k raised the hold; its p update did not run.  Add the
            // pending increments (a_k p_k, and for odd k a_{k-1} p_{k-1}), then
            // the reliable update: x += xs; r64 = b - M^dagger M x; rs = r64
            const int kh = h.iter - 1, par = kh & 1;
            hipLaunchKernelGGL((k_cg_xflush<S, SA>), dim3(gn), dim3(kBlock), 0, c->stream, n,
                               (const sst *)pbuf[par], (const sst *)pbuf[par ^ 1], (ast *)xs,
                               (const CgState *)c->cg, par);
            hipLaunchKernelGGL((k_accum_x<SA>), dim3(gn), dim3(kBlock), 0, c->stream, n, x, (ast *)xs);
            mdagm64(c, t2, x, t1, chi, hpin, imass);
            hipLaunchKernelGGL((k_true_resid<SA>), dim3(gn), dim3(kBlock), 0, c->stream, n, b, (const double2 *)t2,
                               r64, (ast *)rs, c->partial2);
            hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kBlock), 0, c->stream, (const double *)c->partial2,
                               (int)gn, c->dscal + 1);
            // Re <r, p_k>: the kept direction's component along the new residual
            {
                long nbd = std::min<long>(2048, (n + kBlock - 1) / kBlock);
                nbd = (nbd + 7) / 8 * 8;
                hipLaunchKernelGGL((k_dot_rp<S>), dim3((unsigned)nbd), dim3(kBlock), 0, c->stream, n,
                                   (const double2 *)r64, (const sst *)pbuf[par], c->partial2);
                hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kBlock), 0, c->stream, (const double *)c->partial2,
                                   (int)nbd, c->dscal + 2);
            }
            double rd[2];
            HIPCHK(c, hipMemcpyAsync(rd, c->dscal + 1, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            rn2 = rd[0];
            const double rp = rd[1];
            st.reliable_updates++;
            if (rn2 <= target2) break;
            if (!(rn2 == rn2)) {
                status = fail(c, CHS_ENOCONV, "non-finite residual after %d iterations", h.iter);
                break;
            }
            if (h.iter >= maxiter) {
                status = fail(c, CHS_ENOCONV, "no convergence to %g in %d iterations (true resid %g)", tol, h.iter,
                              std::sqrt(rn2 / bnorm2));
                break;
            }
            bool restart = false;
            if (rn2 < best_rn2) {
                best_rn2 = rn2;
                stall = 0;
            } else if (++stall >= kStall) {
                status = fail(c, CHS_ENOCONV,
                              "true residual diverging: %g after %d iterations, best %g, no improvement in %d "
                              "reliable updates",
                              std::sqrt(rn2 / bnorm2), h.iter, std::sqrt(best_rn2 / bnorm2), kStall);
                break;
            } else {
                // the true residual did not improve since the last update: the
                // kept direction carries more rounding than Krylov information
                // (BF16 storage at L=16, Ls=32, beta=0.45 reaches this), so CG
                // restarts from p = r
                restart = true;
                st.restarts++;
            }
            // the iterated residual had reached the stopping region but the
            // true one had not: tighten the trigger
            if (h.rn_hold < h.stop2) {
                stop_fac *= 0.25;
                h.stop2 = stop_fac * target2;
            }
            // continue CG from the true residual with the search direction kept,
            // first made orthogonal to the new residual (p -= (<r,p>/|r|^2) r,
            // as the reliable-update CG of QUDA does): p_{k+1} = r + beta p_k'
            // from p_k into the other buffer; nothing is pending
            h.beta = restart ? 0.0 : rn2 / h.rnorm2;
            const double cr = rp / rn2;   // in p's stored units
            if constexpr (S::scaled) {
                const double sig_new = std::sqrt(rn2 / (double)n);
                h.xp_p = h.beta * h.sigma / sig_new;
                h.xp_r = 1.0 / sig_new - h.xp_p * cr;
                h.sigma = sig_new;
            } else {
                h.xp_p = h.beta;
                h.xp_r = 1.0 - h.beta * cr;
            }
            h.xp_a = 0.0;
            h.rnorm2 = rn2;
            h.maxr2 = rn2;
            h.hold = 0;
            h.halt = 0;
            h.aslot = kAlphaEmpty;
            HIPCHK(c, hipMemcpyAsync(c->cg, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
            hipLaunchKernelGGL((k_cg_prestart<S, SA>), dim3(gn), dim3(kBlock), 0, c->stream, n,
                               (const sst *)pbuf[par], (sst *)pbuf[par ^ 1], (const ast *)rs, (const CgState *)c->cg);
        }
        st.iterations = h.iter;
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    st.seconds = ms * 1e-3;
    st.true_resid = bnorm2 > 0 ? std::sqrt(rn2 / bnorm2) : 0.0;
    if (stats) *stats = st;
    return status;
}

}  // namespace

extern "C" {

int chs_solve_mdagm_dev(chs_ctx *c, const double *b, double *x, const double *chi, double hpin, int imass,
                        double tol, int prec, double delta, int maxiter, chs_solve_stats *stats) {
    if (!(tol > 0)) return fail(c, CHS_EINVAL, "tol must be > 0");
    if (maxiter <= 0) maxiter = 100000;
    if (prec == CHS_FP64) {
        if (delta <= 0) delta = 1e-30;   // plain CG: true residual only at the end
        return rcg_solve<F64, F64>(c, b, x, chi, hpin, imass, tol, delta, maxiter, stats);
    }
    if (prec == CHS_FP32) {
        if (delta <= 0) delta = 0.1;
        return rcg_solve<F32, F32>(c, b, x, chi, hpin, imass, tol, delta, maxiter, stats);
    }
    if (prec == CHS_BF16) {
        if (delta <= 0) delta = 0.1;
        return rcg_solve<B16, F32>(c, b, x, chi, hpin, imass, tol, delta, maxiter, stats);
    }
    if (prec == CHS_FP16) {
        if (delta <= 0) delta = 0.1;
        return rcg_solve<H16, F32>(c, b, x, chi, hpin, imass, tol, delta, maxiter, stats);
    }
    return fail(c, CHS_EINVAL, "precision %d not in {0,1,2,3}", prec);
}

// ---- timing helper for the roofline figure ------------------------------------
int chs_time_dslash(chs_ctx *c, int prec, int dagger, int nbuf, int iters, double hpin, int imass, double *avg_ms) {
    if (!prec_ok(prec) || nbuf < 1 || iters < 1) return fail(c, CHS_EINVAL, "bad arguments");
    const size_t bytes = (size_t)c->kferm * prec_bytes(prec);
    std::vector<void *> ins(nbuf), outs(nbuf);
    DevAllocs mem;
    double *chi = nullptr;
    HIPCHK(c, mem.alloc((void **)&chi, 24 * c->kvol));
    fill_random_real(c, chi, 3L * c->kvol, 1);
    for (int i = 0; i < nbuf; i++) {
        HIPCHK(c, mem.alloc(&ins[i], bytes));
        HIPCHK(c, mem.alloc(&outs[i], bytes));
        fill_random(c, ins[i], prec, c->kferm, 2 + i);   // random bits: zeros run ~15% faster
    }
    HIPCHK(c, hipGetLastError());
    for (int i = 0; i < nbuf; i++) chs_dslash_dev(c, outs[i], ins[i], chi, hpin, imass, dagger, prec);   // warm-up
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < iters; i++) chs_dslash_dev(c, outs[i % nbuf], ins[i % nbuf], chi, hpin, imass, dagger, prec);
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = ms / iters;
    return CHS_OK;
}

// A plain copy of one spinor (2V bytes, non-temporal stores) over nbuf
// rotating buffer pairs of random contents: the streaming rate the stencil's
// HBM figure is set beside.
static int time_copy_impl(chs_ctx *c, int prec, int nbuf, int iters, long pgrid, double *avg_ms) {
    if (!prec_ok(prec) || nbuf < 1 || iters < 1) return fail(c, CHS_EINVAL, "bad arguments");
    if (c->kferm >= (1L << 32)) return fail(c, CHS_EINVAL, "spinor too large for the copy kernel");
    const size_t bytes = (size_t)c->kferm * prec_bytes(prec);
    std::vector<void *> ins(nbuf), outs(nbuf);
    DevAllocs mem;
    for (int i = 0; i < nbuf; i++) {
        HIPCHK(c, mem.alloc(&ins[i], bytes));
        HIPCHK(c, mem.alloc(&outs[i], bytes));
        fill_random(c, ins[i], prec, c->kferm, 2 + i);
    }
    const long n = c->kferm;
    const dim3 g(nblocks(n));
    auto launch = [&](int i) {
        void *o = outs[i % nbuf], *in = ins[i % nbuf];
        if (prec == CHS_FP64 && pgrid > 0)
            hipLaunchKernelGGL(k_copy_nt_grid<F64>, dim3((unsigned)pgrid), dim3(kBlock), 0, c->stream, n, (double2 *)o,
                               (const double2 *)in);
        else if (prec == CHS_FP64)
            hipLaunchKernelGGL(k_copy_nt<F64>, g, dim3(kBlock), 0, c->stream, n, (double2 *)o, (const double2 *)in);
        else if (prec == CHS_FP32)
            hipLaunchKernelGGL(k_copy_nt<F32>, g, dim3(kBlock), 0, c->stream, n, (float2 *)o, (const float2 *)in);
        else   // BF16 and FP16: 32-bit words
            hipLaunchKernelGGL(k_copy_nt<B16>, g, dim3(kBlock), 0, c->stream, n, (uint32_t *)o, (const uint32_t *)in);
    };
    for (int i = 0; i < nbuf; i++) launch(i);   // warm-up
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < iters; i++) launch(i);
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = ms / iters;
    return CHS_OK;
}

int chs_time_copy(chs_ctx *c, int prec, int nbuf, int iters, double *avg_ms) {
    return time_copy_impl(c, prec, nbuf, iters, 0, avg_ms);
}

int chs_time_copy_grid(chs_ctx *c, int nbuf, int iters, int grid, double *avg_ms) {
    return time_copy_impl(c, CHS_FP64, nbuf, iters, grid, avg_ms);
}

// The fused CG stencil of the reliable-update solve (M^dagger applied to M p
// with r -= alpha M^dagger M p and |r|^2 partials), cycling nbuf sets of its
// arrays; alpha = 0, so the traffic is that of a solve (r is rewritten with
// its own random values).
int chs_time_cg_stencil(chs_ctx *c, int prec, int nbuf, int iters, double hpin, int imass, double *avg_ms) {
    if (!prec_ok(prec) || nbuf < 1 || iters < 1) return fail(c, CHS_EINVAL, "bad arguments");
    // M p and p in the storage precision; x and r in the accumulation one
    // (FP64 for FP64, FP32 otherwise), as in chs_solve_mdagm_dev
    const int pacc = prec == CHS_FP64 ? CHS_FP64 : CHS_FP32;
    std::vector<void *> a(4 * nbuf, nullptr);
    DevAllocs mem;
    double *chi = nullptr;
    HIPCHK(c, mem.alloc((void **)&chi, 24 * c->kvol));
    fill_random_real(c, chi, 3L * c->kvol, 1);
    HIPCHK(c, hipMemsetAsync(c->dscal, 0, sizeof(double), c->stream));   // alpha
    unsigned long long seed = 2;
    for (size_t k = 0; k < a.size(); k++) {
        const int pk = (k % 4) < 2 ? prec : pacc;
        HIPCHK(c, mem.alloc(&a[k], (size_t)c->kferm * prec_bytes(pk)));
        fill_random(c, a[k], pk, c->kferm, seed++);   // random bits: zeros run ~15% faster
    }
    HIPCHK(c, hipGetLastError());
    auto launch = [&](int i) {
        void **s = &a[4 * (i % nbuf)];   // M p, p, x, r
        EpiArgs e = epi(c->partial, c->dscal, s[1], s[2], s[3], nullptr, 1);
        if (prec == CHS_FP64) dslash_launch<F64, F64, true, EPI_CG>(c, nullptr, s[0], chi, hpin, imass, e);
        else if (prec == CHS_FP32) dslash_launch<F32, F32, true, EPI_CG>(c, nullptr, s[0], chi, hpin, imass, e);
        else if (prec == CHS_BF16) dslash_launch<B16, F32, true, EPI_CG>(c, nullptr, s[0], chi, hpin, imass, e);
        else dslash_launch<H16, F32, true, EPI_CG>(c, nullptr, s[0], chi, hpin, imass, e);
    };
    for (int i = 0; i < nbuf; i++) launch(i);   // warm-up
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < iters; i++) launch(i);
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = ms / iters;
    return CHS_OK;
}

}  // extern "C"

// learned preconditioner and flexible CG (same translation unit)
#include "chs_learn_host.inc"
#include "chs_batch_host.inc"
