/*
 * chs.h -- C-ABI of libchs, the MI355X (gfx950) implementation of the
 * fermion hot path of the 2+1D chiral Heisenberg RHMC (reference:
 * rhmc/cheisdwf.f of j-ostmeyer/chiralHeisenberg).
 *
 * Conventions
 *  - Plain pointers and sizes only.  Every field is in the reference layout:
 *    spinors are complex*16 Phi(Ls, L, L, Lt, 4, 2) column-major with
 *    interleaved (re, im), i.e. element offset s + Ls*(site + kvol*(d + 4 f)),
 *    site = ix + L*iy + L*L*it (rhmc/cheisdwf.f:50,54,2644); auxiliary field
 *    chi(kvol, 3) real*8; momenta and forces (kvol, 3) real*8.
 *  - Entry points whose name has no _dev suffix take HOST arrays exactly as
 *    the reference routines do (the drop-in boundary); _dev entry points take
 *    DEVICE pointers into memory that stays resident in HBM.
 *  - Numerics follow the reference compiled with -fdefault-real-8
 *    (rhmc/Makefile:21): every real is FP64.  To reproduce the shipped
 *    REAL*4 rounding of chi inside dslash (rhmc/cheisdwf.f:2917), pass chi
 *    already rounded to FP32.
 *  - The reference routines return no status; here every entry point returns
 *    0 on success and a nonzero CHS_E* code otherwise, and chs_last_error()
 *    gives the message.  The Fortran shim keeps the reference signatures and
 *    checks the status itself.
 *  - One context = one lattice geometry + one HIP stream (default priority).
 *    Entry points are not reentrant per context; distinct contexts may be
 *    used from distinct host threads.
 *  - am (the bare mass) is accepted for signature parity and has no effect,
 *    as in the reference (rhmc/cheisdwf.f:26).
 */
#ifndef CHS_H
#define CHS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chs_ctx chs_ctx;

enum {
    CHS_OK = 0,
    CHS_EINVAL = 1,     /* bad argument (size, imass, precision, ndiag, ...) */
    CHS_EHIP = 2,       /* HIP runtime error */
    CHS_ENOCONV = 3,    /* solver hit its iteration cap (result still written) */
    CHS_ENOMEM = 4
};

/* storage precision of a device spinor */
enum { CHS_FP64 = 0, CHS_FP32 = 1, CHS_BF16 = 2, CHS_FP16 = 3 };
/* solver form of chs_fermion's propagator solves beyond the storage
 * precisions: flexible CG with the learned preconditioner (chs_nn_*) */
enum { CHS_LEARNED = 4 };

/* Upper bound on the number of rational shifts in one solve. */
#define CHS_MAX_SHIFTS 32

/* ---- context ------------------------------------------------------------ */
/* Replaces the compile-time parameter(ksize, ksizet, kthird) and the COMMON
 * /para/ am3, ibound (rhmc/cheisdwf.f:31,40,85).  Selects device `device`. */
chs_ctx *chs_ctx_create(int L, int Lt, int Ls, int ibound, double am3, int device);
void chs_ctx_destroy(chs_ctx *ctx);
const char *chs_last_error(const chs_ctx *ctx);
/* message for a failed chs_ctx_create (no context exists then) */
const char *chs_create_error(void);
/* kvol = L*L*Lt, kferm = 8*Ls*kvol */
long chs_kferm(const chs_ctx *ctx);
long chs_kvol(const chs_ctx *ctx);
/* the context's HIP stream, as an opaque handle (hipStream_t) */
void *chs_stream(chs_ctx *ctx);
int chs_synchronize(chs_ctx *ctx);

/* ---- operators (host arrays) -------------------------------------------- */
/* Phi = M R.  Replaces dslash(Phi,R,chi,am,hpin,imass), rhmc/cheisdwf.f:2634. */
int chs_dslash(chs_ctx *ctx, double *phi, const double *r, const double *chi,
               double am, double hpin, int imass);
/* Phi = M^dagger R.  Replaces dslashd, rhmc/cheisdwf.f:3071. */
int chs_dslashd(chs_ctx *ctx, double *phi, const double *r, const double *chi,
                double am, double hpin, int imass);

/* ---- solvers (host arrays) ---------------------------------------------- */
/* Multi-shift Lanczos on (M^dagger M + aden_k) x_k = Phi.
 * Replaces qmrherm(Phi,res,itercg,am,hpin,imass,anum,aden,ndiag,iflag),
 * rhmc/cheisdwf.f:599.  The COMMON inputs become arguments: chi = /trial/
 * chit; x = /vector/ x (output: a0 Phi + sum a_k x_k for iflag=0, a0 Phi for
 * iflag=1); dsdchi = /sforce/ dSdchi (accumulated for iflag=1, may be NULL
 * for iflag=0).  iflag must be 0 or 1 (CHS_EINVAL otherwise; the reference
 * takes any nonzero value as 1, and its program passes only 0 and 1).  anum
 * has ndiag+1 entries, aden ndiag.  Returns CHS_ENOCONV,
 * with x and dsdchi written as the reference leaves them, when the Lanczos
 * loop reaches its cap (niterc = 5000, rhmc/cheisdwf.f:603,639, where the
 * reference returns silently) or |r|^2 turns non-finite; chs_force and
 * chs_hamilton pass that status on, their results written too. */
int chs_qmrherm(chs_ctx *ctx, const double *phi, double res, int *itercg,
                double am, double hpin, int imass, const double *anum,
                const double *aden, int ndiag, int iflag, const double *chi,
                double *x, double *dsdchi);

/* CGNR on M^dagger M x = Phi with x the initial guess on entry and the
 * solution on exit (/vector/).  Replaces congrad(Phi,res,itercg,am,hpin,
 * imass), rhmc/cheisdwf.f:1064.  Stops when |r|^2 < kferm res^2 or after
 * Ls*kvol iterations (then returns CHS_ENOCONV). */
int chs_congrad(chs_ctx *ctx, const double *phi, double res, int *itercg,
                double am, double hpin, int imass, const double *chi,
                double *x);

/* dSdchi = 2 beta chi + pseudofermion force with the remez2g set.
 * Replaces force(Phi,res1,am,hpin,imass), rhmc/cheisdwf.f:470.  beta =
 * /para/ beta; anum..bden = /remez2g/; dsdchi = /sforce/ (output); x, if not
 * NULL, receives /vector/ as the reference leaves it; its, if not NULL,
 * receives the guidance and force iteration counts. */
int chs_force(chs_ctx *ctx, const double *phi, double res1, double am,
              double hpin, int imass, const double *chi, double beta,
              const double *anum, const double *aden, const double *bnum,
              const double *bden, int ndiag, double *dsdchi, double *x,
              int *its);

/* chs_qmrherm_dev with the spinors stored in prec (CHS_FP64 or CHS_FP32;
 * Lanczos scalars, reductions, the rational combination x and the derivs
 * sums stay FP64).  No counterpart in the reference. */
int chs_qmrherm_prec_dev(chs_ctx *ctx, const double *phi, double res, int *itercg, double hpin, int imass,
                         const double *anum, const double *aden, int ndiag, int iflag, const double *chi,
                         double *x, double *dsdchi, int prec);

/* Storage precision of the two multi-shift solves inside chs_force
 * (CHS_FP64, the default, or CHS_FP32: spinors in FP32, Lanczos scalars,
 * reductions and the derivs bilinears and sums in FP64; CHS_FP16 runs them
 * in FP32 too, since shifted systems have no reliable updates and 16-bit
 * storage does not converge there).  For the HMC sibling's chs_hmc_force
 * the setting picks the CG: the reference CGNR (CHS_FP64) or the
 * reliable-update CG of chs_solve_mdagm_dev in CHS_FP32 or CHS_FP16 form
 * from the same guess, to the congrad rule taken as an FP64 true relative
 * residual.  No counterpart in the reference (which computes in FP64
 * throughout); hamilton, and with it the Metropolis test, stays FP64. */
int chs_set_force_precision(chs_ctx *ctx, int prec);

/* Iteration cap of every multi-shift solve of the context (the reference's
 * compile-time niterc = 5000, rhmc/cheisdwf.f:603; 0 restores it).  No
 * counterpart as a call: for tests of the cap and for callers that want an
 * earlier bound. */
int chs_set_lanczos_cap(chs_ctx *ctx, int niterc);

/* Every shifted solution of the multi-shift solve: x_k = (M^dagger M +
 * aden[k])^{-1} Phi for k < ndiag, FP64 native device spinors, shift-major
 * (x_k at xk + 2*kferm*k doubles), to the reference's stopping rule
 * (max_k |phimod rho_k| < 2 sqrt(kferm) res).  No counterpart as a call: the
 * reference keeps them in qmrherm's local x1(ndiag, kferm)
 * (rhmc/cheisdwf.f:613, :672-687) and returns only their rational
 * combination (iflag = 0) or force (iflag = 1); for callers that combine
 * the shifts themselves (other coefficient sets over the same poles) and
 * for checking each shift's true residual. */
int chs_qmrherm_shifts_dev(chs_ctx *ctx, const double *phi, double res, int *itercg, double hpin, int imass,
                           const double *aden, int ndiag, const double *chi, double *xk);

/* Shift freezing in every multi-shift solve of the context (no counterpart
 * in the reference, which updates every shift's vectors until the slowest
 * has converged, rhmc/cheisdwf.f:688-715): with 0 < f <= 1, shift k's p_k and
 * x_k are no longer updated from the iteration after |phimod rho_k| falls
 * below f * 2 resid -- the reference's stopping test applied to that shift
 * alone, tightened by f.  The scalar recurrences of every shift continue, so
 * the stopping test and itercg are unchanged; each frozen x_k lacks only
 * updates smaller than f times what the reference's own stop leaves out.
 * f = 0 (the default) updates every shift as the reference does.  For
 * callers that want the force's solves faster (their largest shifts converge
 * long before the smallest); the batched solve follows each system's
 * context. */
int chs_set_shift_freeze(chs_ctx *ctx, double f);

/* ---- device mirrors of host spinors (boundary convention) ----------------
 * The host-array entry points upload each spinor argument once per call.
 * A read-only spinor argument that the caller passes again and again
 * unchanged (Phi: every force and hamilton of a trajectory; R of repeated
 * dslash calls) can be registered: the library then keeps a device copy,
 * keyed by the host address, and uploads it only when the array's generation
 * has moved.  The caller bumps the generation with chs_mirror_touch after
 * every write to the array (in the reference's driver: after the
 * pseudofermion heatbath sets Phi, rhmc/cheisdwf.f:238-245).  Arrays written
 * by a call (x, /vector/, dsdchi) are always transferred.  No counterpart in
 * the reference. */
int chs_mirror_register(chs_ctx *ctx, const void *host);   /* kferm complex*16 */
int chs_mirror_touch(chs_ctx *ctx, const void *host);
int chs_mirror_release(chs_ctx *ctx, const void *host);
/* spinor bytes moved host -> device and device -> host by the host-array
 * entry points since the context was created */
int chs_transfer_bytes(const chs_ctx *ctx, unsigned long long *h2d, unsigned long long *d2h);

/* H = beta sum chi^2 + 1/2 sum p^2 + Re Phi^dagger (M^dagger M)^{-1/2} Phi.
 * Replaces hamilton(Phi,h,hg,hp,s,res2,isweep,am,hpin,imass),
 * rhmc/cheisdwf.f:522, with /trial/ chit, pp and /remez2/ b-set explicit.
 * All global sums are FP64 on the device.  x, if not NULL, receives
 * /vector/; itercg, if not NULL, the iteration count. */
int chs_hamilton(chs_ctx *ctx, const double *phi, double *h, double *hg,
                 double *hp, double *s, double res2, int isweep, double am,
                 double hpin, int imass, const double *chi, const double *pp,
                 double beta, const double *bnum, const double *bden,
                 int ndiag, double *x, int *itercg);

/* Wall-source fermion correlators on the domain walls.  Replaces
 * fermion(res,aviter,am,hpin,imass), rhmc/cheisdwf.f:897-1058, with /trial/
 * chi explicit and ittt0 = int(Lt*rano(yran,idum)) + 1 (the one random draw
 * of the routine, 1..Lt) passed in.  Five wall sources, Dirac 3 and 4,
 * congrad each (prec = CHS_FP64: the reference CGNR and stopping rule;
 * CHS_FP32 / CHS_FP16: the FP32 or scaled-binary16 reliable-update solve to
 * the same residual; CHS_LEARNED: flexible CG with the learned
 * preconditioner, chs_nn_init first, to the same residual).  While a pair
 * store is open (chs_nn_collect), every solve of a CHS_FP64 / CHS_FP32 /
 * CHS_FP16 call appends its (source, solution) pair to it.
 * cferm1 = fort.5001/5002, cferm2 = fort.5011/5012: complex (re, im)
 * [flavour][t], t fastest, 2*Lt complex each.  iterations = total solver
 * iterations (the reference's aviter times 10).  A solve at its iteration
 * cap keeps its last iterate, as the reference's congrad does; the results
 * are written and CHS_ENOCONV returned, chs_last_error() beginning with the
 * number of such solves (the reference writes its message once for each). */
int chs_fermion(chs_ctx *ctx, const double *chi, double res, double hpin, int imass, int ittt0, int prec,
                double *cferm1, double *cferm2, int *iterations);

/* ---- HMC sibling mh/cheisdwf_hmc.f (its operator is M with hpin = 0) ---- */
/* Both calls, like the sibling's congrad, keep the last iterate at the
 * iteration cap: the results are written and CHS_ENOCONV is returned,
 * chs_last_error() beginning with the number of capped solves (1). */
/* Replaces force(Phi,X0,dSdchi,iflag,res1,am,imass), mh/cheisdwf_hmc.f:393:
 * X1 = (M^dagger M)^{-1} Phi by congrad from the guess x0 (/vector/ in the
 * reference), dSdchi = 4 beta chi - 2 (wall bilinears of X1 and M X1), and
 * x0 <- 2 X1 - x0.  /trial/ chi and /para/ beta explicit; host arrays. */
int chs_hmc_force(chs_ctx *ctx, const double *phi, double *x0, double res, double am, int imass, const double *chi,
                  double beta, double *dsdchi, int *itercg);
/* Replaces hamilton(Phi,X0,h,hg,hp,s,res2,isweep,am,imass),
 * mh/cheisdwf_hmc.f:561: h4 = (h, hg = 2 beta sum chi^2, hp = sum p^2 / 2,
 * s = hg + Re Phi^dagger X1), X1 by congrad from x0, which receives X1. */
int chs_hmc_hamilton(chs_ctx *ctx, const double *phi, double *x0, double res, double am, int imass,
                     const double *chi, const double *pp, double beta, double *h4, int *itercg);

/* dSdchi += 2 a wall bilinears of r and x2 = M r.
 * Replaces derivs(R,X2,anum,imass), rhmc/cheisdwf.f:747. */
int chs_derivs(chs_ctx *ctx, const double *r, const double *x2, double anum,
               int imass, double *dsdchi);

/* ---- device-resident interface ------------------------------------------ */
/* Device spinors of the _dev entry points are in the library's NATIVE
 * layout: the same 8 spin-flavour planes (d + 4 f), but inside a plane the
 * element of (s, site) sits at s*kvol + site (s outermost) instead of the
 * reference's site*Ls + s.  The stencil needs this to keep each XCD's
 * neighbour reads inside its own L2.  chs_layout_dev converts between the two
 * orders on the device (to_native = 1: reference -> native). */
int chs_layout_dev(chs_ctx *ctx, void *dst, const void *src, int prec, int to_native);

/* Device buffers allocated by the library (hipMalloc), freed with chs_free. */
void *chs_alloc(chs_ctx *ctx, size_t bytes);
int chs_free(chs_ctx *ctx, void *p);
int chs_upload(chs_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
int chs_download(chs_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);
/* Convert a device FP64 spinor to another storage precision and back. */
int chs_convert_dev(chs_ctx *ctx, void *dst, int dst_prec, const void *src,
                    int src_prec, long ncomplex);

/* out = M in (dagger=0) or M^dagger in (dagger=1) on device spinors of
 * storage precision prec (FP64: double2, FP32: float2, BF16 / FP16: bf16 or
 * IEEE binary16 re,im pair).  chi is a device real*8 (kvol,3) array. */
int chs_dslash_dev(chs_ctx *ctx, void *out, const void *in, const double *chi,
                   double hpin, int imass, int dagger, int prec);

/* Multi-shift Lanczos on device arrays (same recurrences and stopping rule
 * as chs_qmrherm).  phi, x: device FP64 spinors; dsdchi: device (kvol,3)
 * FP64, accumulated for iflag=1. */
int chs_qmrherm_dev(chs_ctx *ctx, const double *phi, double res, int *itercg,
                    double hpin, int imass, const double *anum,
                    const double *aden, int ndiag, int iflag,
                    const double *chi, double *x, double *dsdchi);

/* CGNR with the reference's recurrences on device arrays. */
int chs_congrad_dev(chs_ctx *ctx, const double *phi, double res, int *itercg,
                    double hpin, int imass, const double *chi, double *x);

/* chs_congrad_dev that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code that must not
 * call hipMalloc, such as tight serving loops or harnesses that forbid it).
 * The work spinors p, r and M p, the partial sums and the CG state live in a
 * static device array of the library, allocated with the code object on
 * every GPU and held until the process ends: kferm <= CHS_NOALLOC_MAX_KFERM
 * (2^20 complex numbers, e.g. L = 16, Ls = 32; 48 MiB + 64 KiB of scratch),
 * larger lattices are rejected with CHS_EINVAL.  Same iterates, results and
 * status as chs_congrad_dev (bit for bit).  The array is shared by every
 * call on one GPU: calls that overlap on one GPU (two host threads, two
 * contexts' streams) race and must be serialised by the caller. */
#define CHS_NOALLOC_MAX_KFERM (1L << 20)
int chs_congrad_noalloc_dev(chs_ctx *ctx, const double *phi, double res, int *itercg,
                            double hpin, int imass, const double *chi, double *x);

/* Statistics of one M^dagger M solve. */
typedef struct {
    int iterations;        /* total inner (Krylov) iterations */
    int reliable_updates;  /* FP64 true-residual recomputations */
    int restarts;          /* fallbacks / restarts taken */
    double true_resid;     /* final |b - M^dagger M x| / |b| in FP64 */
    double seconds;        /* wall time of the solve on the device stream */
} chs_solve_stats;

/* Solve M^dagger M x = b on device arrays to the FP64 true relative
 * residual tol.  prec selects the storage precision of the inner CG
 * (CHS_FP64: plain FP64 CG; CHS_FP32 / CHS_BF16 / CHS_FP16: reliable-update
 * CG with FP64 true residuals every time the inner residual has dropped by
 * delta, delta <= 0 picks a default; in CHS_FP16 the search direction and its
 * image are stored in binary16 divided by the residual's rms, the residual
 * and the x increment in FP32).  x holds the initial guess on entry.
 * A reliable update whose FP64 true residual is no better than the best so
 * far restarts CG from p = r (stats->restarts; BF16 storage at L=16, Ls=32,
 * beta=0.45 converges this way).  Returns CHS_ENOCONV after maxiter
 * iterations (maxiter <= 0: 100000), on a non-finite residual, or when the
 * true residual has not improved in 8 reliable updates in a row (a diverging
 * solve).  The CG step length alpha is formed
 * inside the fused CG stencil launch, whose waves wait (bounded, about a
 * second) for it: CHS_EHIP if a wave gave up. */
int chs_solve_mdagm_dev(chs_ctx *ctx, const double *b, double *x,
                        const double *chi, double hpin, int imass, double tol,
                        int prec, double delta, int maxiter,
                        chs_solve_stats *stats);

/* ---- learned preconditioner (no counterpart in the reference) ------------
 * A linear, symmetric positive definite 3-D convolutional network on the
 * Ls-stacked spinor (16*Ls channels: s, spin-flavour plane, re/im):
 *   P(r) = c u + C1^T C2^T C2 C1 u,  u = r / rms(r),  c = exp(theta),
 * C1 a 1x1x1 convolution 16*Ls -> 16 channels, C2 a 3x3x3 convolution over
 * (x, y, t), periodic, 16 -> 16; BF16 operands on gfx950 MFMA
 * (v_mfma_f32_16x16x32_bf16), FP32 accumulation.  P is SPD for any weights,
 * so preconditioned CG keeps its conjugacy.  `hidden` must be 16. */
int chs_nn_init(chs_ctx *ctx, int hidden, unsigned long long seed);
long chs_nn_num_weights(chs_ctx *ctx);
/* all parameters, FP32: W1[n][k] (16 x 16*Ls), W2[tap][n][k] (27 x 16 x 16,
 * tap = (dx+1) + 3 (dy+1) + 9 (dt+1)), theta */
int chs_nn_set_weights(chs_ctx *ctx, const float *w, long n);
int chs_nn_get_weights(chs_ctx *ctx, float *w, long n);
/* z = P(r): FP32 native-layout device spinors */
int chs_nn_apply_dev(chs_ctx *ctx, const float *r, float *z);
/* loss and gradient (host array of chs_nn_num_weights floats, may be NULL)
 * for one pair: with u = src / rms(src), w = P(u) - sol / rms(src), the
 * relative A-norm <w, A w> / <u, sol / rms(src)>, A = M^dagger M of the
 * operator given by chi, hpin, imass (device arrays) */
int chs_nn_loss_grad_dev(chs_ctx *ctx, const float *src, const float *sol, const double *chi, double hpin, int imass,
                         double *loss, float *grad);
/* Adam training over npairs (source, solution) pairs stored contiguously as
 * FP32 native spinors; epoch_loss (epochs entries) receives the mean loss of
 * each epoch. */
int chs_nn_train_dev(chs_ctx *ctx, const float *src, const float *sol, int npairs, int epochs, float lr,
                     const double *chi, double hpin, int imass, double *epoch_loss);
/* Training pairs from the propagator solves: chs_nn_collect opens a device
 * store for max_pairs (source, solution) pairs -- FP32 native spinors and
 * the chi each was solved with -- emptying any earlier one; max_pairs = 0
 * frees it.  While it has room, every propagator solve of chs_fermion in a
 * plain form (CHS_FP64 / CHS_FP32 / CHS_FP16) appends its pair (source
 * M^dagger xi, solution x of M^dagger M x = M^dagger xi).  chs_nn_init
 * (which recreates the network) frees the store. */
int chs_nn_collect(chs_ctx *ctx, int max_pairs);
/* number of pairs in the store (0 without a network or a store) */
int chs_nn_num_collected(chs_ctx *ctx);
/* device addresses of stored pair i: src and sol FP32 native spinors, chi
 * 3*kvol doubles; valid until the next chs_nn_collect or chs_nn_init */
int chs_nn_pair_dev(chs_ctx *ctx, int i, const float **src, const float **sol, const double **chi);
/* chs_nn_train_dev over the stored pairs, each with the chi it was solved
 * with; the same updates, in the same order, as chs_nn_train_dev on the
 * same pairs when they share one chi */
int chs_nn_train_collected(chs_ctx *ctx, int epochs, float lr, double hpin, int imass, double *epoch_loss);
/* Flexible CG (FP32 iteration, FP64 reliable updates every delta of
 * residual reduction, answer accepted on the FP64 true relative residual) with
 * z = P(r) (use_nn = 1) or z = r (use_nn = 0, plain CG).  Plain-CG fallback:
 * when with the network |r|^2 has not fallen by a factor 8 within 64
 * iterations the solve continues from its current x as plain CG
 * (stats->restarts counts it). */
int chs_solve_mdagm_fcg_dev(chs_ctx *ctx, const double *b, double *x, const double *chi, double hpin, int imass,
                            double tol, double delta, int maxiter, int use_nn, chs_solve_stats *stats);

/* ---- batched multi-shift solves (no counterpart in the reference) ----------
 * Many qmrherm solves at once -- pseudofermion fields, lattices of different
 * sizes, the heatbath / Hamiltonian kind (iflag = 0) and the force kind
 * (iflag = 1) mixed -- iterated in lock step with one grouped launch per
 * Lanczos stage, a chunk of steps replayed from a hipGraph.  Every system's
 * result (x, and for iflag = 1 the dsdchi contribution) and itercg equal
 * those of chs_qmrherm_dev on the same inputs bit for bit.  Systems may share
 * a context (its geometry is all that is used); all must be on the batch's
 * device.  Device arrays in the native layout. */
typedef struct chs_batch chs_batch;
typedef struct {
    chs_ctx *ctx;           /* lattice geometry */
    const double *phi;      /* source (device, FP64) */
    const double *chi;      /* auxiliary field (device) */
    double *x;              /* out: anum[0] phi + sum_k anum[k+1] x_k (iflag = 0), anum[0] phi (iflag = 1) */
    double hpin;
    int imass;
    int ndiag;
    const double *anum;     /* host, ndiag + 1 */
    const double *aden;     /* host, ndiag */
    double res;             /* qmrherm's res: stop when max_k |rho_k| phimod < 2 sqrt(kferm) res */
    int itercg;             /* out */
    int iflag;              /* 0, or 1: the force's solve (rhmc/cheisdwf.f:505-513, :733-740) */
    double *dsdchi;         /* iflag = 1: (3, kvol) FP64 device array, the shifts' derivs accumulated */
} chs_ms_system;
/* launch parameters: ept = element blocks (of 256) per workgroup of the
 * shift update (1, 2 or 4), chunk = iterations per poll (even), graph =
 * replay chunks from a hipGraph; us_per_iter = the tuner's measurement */
typedef struct {
    int ept, chunk, graph;
    double us_per_iter;
} chs_batch_tune;
typedef struct {
    int iterations;         /* lock-step iterations run (>= every itercg) */
    int graph_replays;
    int launches;           /* kernel launches enqueued one by one */
    double seconds;
    chs_batch_tune tune;    /* the parameters used after tuning */
} chs_batch_stats;
chs_batch *chs_batch_create(int device);   /* NULL on failure: chs_create_error() */
void chs_batch_destroy(chs_batch *b);
const char *chs_batch_last_error(chs_batch *b);
/* prec: CHS_FP64 or CHS_FP32 storage of every system's spinors (Lanczos
 * scalars, reductions and x stay FP64; FP32 equals chs_qmrherm_prec_dev with
 * CHS_FP32 bit for bit).  tune = NULL: autotune on the first chunks (cached
 * per batch shape). */
int chs_qmrherm_batch_dev(chs_batch *b, chs_ms_system *sys, int nsys, int prec, const chs_batch_tune *tune,
                          chs_batch_stats *stats);

/* Average duration in milliseconds of `iters` back-to-back dslash launches
 * cycling through nbuf (in, out) device buffer pairs, measured with HIP
 * events on the context stream.  Used for the HBM roofline figure.  Inputs
 * and phi hold hashed uniform random values: all-zero buffers would time
 * ~15% fast on this hardware. */
int chs_time_dslash(chs_ctx *ctx, int prec, int dagger, int nbuf, int iters,
                    double hpin, int imass, double *avg_ms);

/* Average duration in milliseconds of a plain copy of one spinor of
 * precision prec (2V bytes; non-temporal stores as in the stencil) cycling
 * nbuf buffer pairs of random contents, HIP events on the context stream:
 * the streaming rate the stencil's HBM figure is compared with.  A measuring
 * helper with no counterpart in the reference. */
int chs_time_copy(chs_ctx *ctx, int prec, int nbuf, int iters, double *avg_ms);

/* The same FP64 copy with `grid` workgroups of 256 threads striding over the
 * spinor (grid <= 0: one element per thread, as chs_time_copy): the
 * streaming ceiling of a persistent grid.  A measuring helper with no
 * counterpart in the reference. */
int chs_time_copy_grid(chs_ctx *ctx, int nbuf, int iters, int grid, double *avg_ms);

/* Average duration in milliseconds of the fused CG stencil of
 * chs_solve_mdagm_dev (M^dagger applied to M p, fused with
 * r -= alpha M^dagger M p and the |r|^2 partials; any storage precision, with
 * x and r in FP64 for CHS_FP64 and in FP32 otherwise),
 * cycling nbuf sets of its arrays (random contents), HIP events on the
 * context stream.  The dominant kernel of the mixed-precision solve. */
int chs_time_cg_stencil(chs_ctx *ctx, int prec, int nbuf, int iters, double hpin, int imass, double *avg_ms);

/* ---- random numbers (rhmc/cheisdwf.f:1678-1801) -------------------------- */
/* Host-side generator state: 48-bit LCG + 97-entry shuffle shared by the
 * yran and yrun pointer streams, FP64 uniforms. */
typedef struct {
    uint64_t s;
    double v[97];
    int idum;
    double yran, yrun;
} chs_rng;
void chs_rng_set(chs_rng *r, double seed);      /* rranset */
double chs_rng_get(const chs_rng *r);            /* rranget */
double chs_rano(chs_rng *r, int stream);         /* 0: yran, 1: yrun */
/* gauss0 (E|z|^2 = 1) and gaussp (N(0,1)): ps(kvol, 2) */
void chs_gauss0(chs_rng *r, double *ps, int kvol);
void chs_gaussp(chs_rng *r, double *ps, int kvol);

#ifdef __cplusplus
}
#endif
#endif
