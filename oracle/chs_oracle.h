/*
 * chs_oracle.h -- CPU oracle for the fermion hot path of the 2+1D chiral
 * Heisenberg RHMC with domain-wall fermions (reference: rhmc/cheisdwf.f).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libchs, chiralheisenberg_amd/) never links or calls it.
 *
 * Numerics follow the reference built with -fdefault-real-8 (its own
 * commented Makefile flag, rhmc/Makefile:21): every real is FP64.  Arrays use
 * the reference layout Phi(Ls, L, L, Lt, 4, 2) column-major with interleaved
 * complex (re, im), i.e. offset = s + Ls*(site + kvol*(d + 4*f)), site =
 * ix + L*iy + L*L*it, all 0-based.  chi is chi(kvol, 3): offset site + kvol*a.
 */
#ifndef CHS_ORACLE_H
#define CHS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int L, Lt, Ls;   /* spatial extent, time extent, fifth dimension */
    int ibound;      /* -1: antiperiodic in t (the reference's only setting) */
    double am3;      /* domain-wall height M5 */
} orc_lat;

/* Phi = M R (dagger=0) or Phi = M^dagger R (dagger=1).
 * rhmc/cheisdwf.f:2634-3069 (dslash), :3071-3501 (dslashd). */
void orc_dslash(const orc_lat *g, double *phi, const double *r,
                const double *chi, double hpin, int imass, int dagger);

/* dSdchi += wall bilinears of y and z = M y (rhmc/cheisdwf.f:747-849). */
void orc_derivs(const orc_lat *g, double *dSdchi, const double *y,
                const double *z, double anum, int imass);

/* Multi-shift Lanczos solve (rhmc/cheisdwf.f:599-745).  x (kferm complex)
 * receives a0*Phi + sum_k a_k x_k for iflag=0 and a0*Phi for iflag=1, in
 * which case dSdchi (3*kvol) is accumulated.  Returns the iteration count. */
int orc_qmrherm(const orc_lat *g, const double *phi, double res,
                const double *anum, const double *aden, int ndiag, int iflag,
                const double *chi, double hpin, int imass,
                double *x, double *dSdchi);

/* Same solve, but also stores every shifted solution x_k (ndiag*kferm
 * complex, shift-major) in xk when xk is not NULL. */
void orc_set_lanczos_cap(int niterc);   /* 0: the reference's 5000 */
int orc_qmrherm_shifts(const orc_lat *g, const double *phi, double res,
                       const double *anum, const double *aden, int ndiag,
                       int iflag, const double *chi, double hpin, int imass,
                       double *x, double *dSdchi, double *xk);

/* CGNR on M^dagger M x = Phi, x holds the initial guess on entry
 * (rhmc/cheisdwf.f:1064-1157).  Returns the iteration count. */
int orc_congrad(const orc_lat *g, const double *phi, double res,
                const double *chi, double hpin, int imass, double *x);

/* Exactly niter CGNR iterations from x = 0 (timing sample). */
void orc_cg_iterations(const orc_lat *g, const double *phi, const double *chi,
                       double hpin, int imass, int niter, double *x);

/* force (rhmc/cheisdwf.f:470-518): dSdchi = 2 beta chi + fermion force.
 * x receives the reference's /vector/ contents on exit.  its[0..1] receive
 * the guidance and force iteration counts. */
void orc_force(const orc_lat *g, const double *phi, double res1,
               const double *chi, double beta,
               const double *anum, const double *aden,
               const double *bnum, const double *bden, int ndiag,
               double hpin, int imass, double *dSdchi, double *x, int *its);

/* hamilton (rhmc/cheisdwf.f:522-589).  out[0..3] = h, hg, hp, s. */
int orc_hamilton(const orc_lat *g, const double *phi, double res2,
                 const double *chi, const double *pp, double beta,
                 const double *bnum, const double *bden, int ndiag,
                 double hpin, int imass, double *out, double *x);

/* fermion (rhmc/cheisdwf.f:897-1058): wall-source propagators on the domain
 * walls.  nsource = 5 wall sources at t-slices (1-based)
 * mod(ittt0 + k*(Lt/5), Lt) + 1, k = 1..5, ittt0 in 1..Lt being the caller's
 * draw int(Lt*rano) + 1; Dirac source 3 and 4, flavour 1, value 1/L^2 at
 * s = 1; x = (M^dagger M)^{-1} M^dagger source by congrad from the initial
 * guess M^dagger source.  cferm1, cferm2: complex (re, im) [f][t], t fastest,
 * f = 0, 1 (the reference's fort.5001/5002 and fort.5011/5012).  Returns the
 * total congrad iterations (the reference's aviter is that / 10). */
int orc_fermion(const orc_lat *g, const double *chi, double res, double hpin,
                int imass, int ittt0, double *cferm1, double *cferm2);

/* ---- HMC sibling mh/cheisdwf_hmc.f (its operator is M at hpin = 0) --- */
/* force (mh/cheisdwf_hmc.f:393-557): X1 = (M^dagger M)^{-1} Phi by congrad
 * from the initial guess x0; dSdchi = 4 beta chi - 2 (wall bilinears of X1,
 * M X1), i.e. 2 (2 beta chi) plus derivs with a = -1; on exit
 * x0 = 2 X1 - x0 (the linear chronological guess).  Returns the iterations. */
int orc_hmc_force(const orc_lat *g, const double *phi, double *x0, double res,
                  const double *chi, double beta, int imass, double *dSdchi);

/* hamilton (mh/cheisdwf_hmc.f:561-620): hg = 2 beta sum chi^2,
 * hp = sum p^2 / 2, hf = Re Phi^dagger X1 with X1 = (M^dagger M)^{-1} Phi by
 * congrad from x0; x0 = X1 on exit; out = h, hg, hp, s = hg + hf.  Returns
 * the iterations. */
int orc_hmc_hamilton(const orc_lat *g, const double *phi, double *x0, double res,
                     const double *chi, const double *pp, double beta, int imass,
                     double *out);

/* ---- random numbers (rhmc/cheisdwf.f:1678-1801) ---------------------- */
typedef struct {
    uint64_t s;       /* 48-bit LCG state */
    double v[97];     /* shuffle table shared by both streams */
    int idum;         /* <0: table not yet initialised */
    double yran;      /* shuffle pointer of the main stream */
    double yrun;      /* shuffle pointer of the MD/Metropolis stream */
} orc_rng;

void orc_rng_set(orc_rng *r, double seed);    /* rranset */
double orc_rng_get(const orc_rng *r);          /* rranget */
double orc_rranf(orc_rng *r);                  /* full-precision draw */
double orc_rano(orc_rng *r, int stream);       /* 0: yran, 1: yrun */
void orc_gauss0(orc_rng *r, double *ps, int kvol);  /* ps(kvol,2), E|z|^2=1 */
void orc_gaussp(orc_rng *r, double *ps, int kvol);  /* ps(kvol,2), N(0,1) */

/* ---- RHMC driver (rhmc/cheisdwf.f:207-413) --------------------------- */
typedef struct {
    double dt, beta, hpin;
    int imass, iterl, itermax, iprint;
    double rescga, rescgg;
    /* remez sets: coefficient arrays, anum has n+1 entries */
    const double *anum2, *aden2, *bnum2, *bden2; int n2;
    const double *anum4, *aden4, *bnum4, *bden4; int n4;
    const double *anum2g, *aden2g, *bnum2g, *bden2g; int n2g;
} orc_rhmc_par;

typedef struct {
    double H0, H1, S0, S1, dH;
    int accepted;
    int md_steps;
    int it_pf, it_h0, it_h1, it_force_total;
} orc_traj_out;

/* Runs ntraj trajectories starting at trajectory number isweep0 (1-based);
 * chi (3*kvol) is updated in place on acceptance. */
void orc_rhmc(const orc_lat *g, const orc_rhmc_par *p, orc_rng *rng,
              double *chi, int isweep0, int ntraj, orc_traj_out *out);

/* helpers for tests / cpu baseline */
double orc_norm2(long n, const double *a);
double orc_redot(long n, const double *a, const double *b);
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
