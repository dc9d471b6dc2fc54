/*
 * chs_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, never shipped as the
 * product path) for the fermion hot path of rhmc/cheisdwf.f.
 *
 * A clean-room restatement of the mathematics, not a transliteration: the
 * operator is written in the generic gamma-matrix form
 *
 *   (M psi)(s,x) = (4-M5) psi(s,x)
 *                - 1/2 sum_mu [ (1-g_mu) psi(s,x+mu) + (1+g_mu) psi(s,x-mu) ]
 *                - [s<Ls] P- psi(s+1,x) - [s>1] P+ psi(s-1,x)
 *                + [s=Ls] P- Y psi(1,x) + [s=1] P+ Y' psi(Ls,x)
 *
 * with antiperiodic time boundary (ibound=-1), P+ = diag(1,1,0,0),
 * P- = diag(0,0,1,1), gamma tables of rhmc/cheisdwf.f:1523-1557 and the wall
 * coupling of rhmc/cheisdwf.f:2912-3066; M^dagger flips the hop signs, swaps
 * the projectors of the s-hop and uses the adjoint wall blocks
 * (rhmc/cheisdwf.f:3313-3498).  The reference itself evaluates the same
 * operator as a sequence of projected array passes (rhmc/cheisdwf.f:2652-3066).
 *
 * Every reduction is deterministic: fixed 4096-element chunks summed in
 * chunk order, whatever the thread count.
 */
#include "chs_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CHUNK 4096L

static long kvol_of(const orc_lat *g) { return (long)g->L * g->L * g->Lt; }
static long kferm_of(const orc_lat *g) { return 8L * g->Ls * kvol_of(g); }

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- deterministic reductions over interleaved complex arrays --------- */
double orc_norm2(long n, const double *a) {
    long nch = (n + CHUNK - 1) / CHUNK;
    double *part = (double *)malloc(sizeof(double) * (nch ? nch : 1));
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nch; c++) {
        long lo = c * CHUNK, hi = lo + CHUNK < n ? lo + CHUNK : n;
        double t = 0.0;
        for (long i = lo; i < hi; i++) t += a[2 * i] * a[2 * i] + a[2 * i + 1] * a[2 * i + 1];
        part[c] = t;
    }
    double s = 0.0;
    for (long c = 0; c < nch; c++) s += part[c];
    free(part);
    return s;
}

/* Re sum conj(a) b */
double orc_redot(long n, const double *a, const double *b) {
    long nch = (n + CHUNK - 1) / CHUNK;
    double *part = (double *)malloc(sizeof(double) * (nch ? nch : 1));
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nch; c++) {
        long lo = c * CHUNK, hi = lo + CHUNK < n ? lo + CHUNK : n;
        double t = 0.0;
        for (long i = lo; i < hi; i++) t += a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1];
        part[c] = t;
    }
    double s = 0.0;
    for (long c = 0; c < nch; c++) s += part[c];
    free(part);
    return s;
}

/* ---- the operator ------------------------------------------------------ */
/* gamma_mu psi: component d = gval[mu][d] * psi[gcol[mu][d]]
 * (rhmc/cheisdwf.f:1523-1557, mu = x, y, t). */
static const int gcol[3][4] = {{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}};
static const double gre[3][4] = {{0, 0, 0, 0}, {-1, 1, 1, -1}, {0, 0, 0, 0}};
static const double gim[3][4] = {{-1, -1, 1, 1}, {0, 0, 0, 0}, {-1, 1, 1, -1}};

/* 2x2 flavour matrix acting at the walls, complex entries y[f][g][re/im] */
typedef struct { double m[2][2][2]; } flav2;

/* Y = phi.sigma (+ hpin sigma3 at the pinned site) for imass=1,
 * -i phi.sigma for imass=3 (rhmc/cheisdwf.f:2912-3066).  which=1 gives Y'. */
static flav2 wall_matrix(const double *chi, long kvol, long site, double hpin,
                         int imass, int which) {
    flav2 y;
    memset(&y, 0, sizeof y);
    double c1 = chi[site], c2 = chi[site + kvol], c3 = chi[site + 2 * kvol];
    if (imass == 1) {
        double h = (site == 0) ? hpin : 0.0;
        y.m[0][0][0] = c3 + h;
        y.m[1][1][0] = -(c3 + h);
        y.m[0][1][0] = c1; y.m[0][1][1] = -c2;
        y.m[1][0][0] = c1; y.m[1][0][1] = c2;
    } else if (imass == 3) {
        /* s * (-i) * phi.sigma, s = +1 for Y, -1 for Y' */
        double sg = which ? -1.0 : 1.0;
        /* -i * c3 */
        y.m[0][0][1] = -sg * c3;
        y.m[1][1][1] = sg * c3;
        /* -i*(c1 - i c2) = -c2 - i c1 */
        y.m[0][1][0] = -sg * c2; y.m[0][1][1] = -sg * c1;
        /* -i*(c1 + i c2) = c2 - i c1 */
        y.m[1][0][0] = sg * c2; y.m[1][0][1] = -sg * c1;
    }
    return y;
}

static flav2 adjoint2(flav2 a) {
    flav2 b;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            b.m[i][j][0] = a.m[j][i][0];
            b.m[i][j][1] = -a.m[j][i][1];
        }
    return b;
}

void orc_dslash(const orc_lat *g, double *phi, const double *r,
                const double *chi, double hpin, int imass, int dagger) {
    const long L = g->L, Lt = g->Lt, Ls = g->Ls;
    const long kvol = kvol_of(g);
    const long plane = Ls * kvol;               /* complex per (d,f) plane */
    const double diag = 4.0 - g->am3;
    const double sig = dagger ? -1.0 : 1.0;      /* M: (1 - g) forward */

#pragma omp parallel for schedule(static)
    for (long site = 0; site < kvol; site++) {
        long ix = site % L, iy = (site / L) % L, it = site / (L * L);
        long fw[3], bw[3];
        double bf[3] = {1, 1, 1}, bb[3] = {1, 1, 1};
        fw[0] = ((ix + 1) % L) + L * iy + L * L * it;
        bw[0] = ((ix + L - 1) % L) + L * iy + L * L * it;
        fw[1] = ix + L * ((iy + 1) % L) + L * L * it;
        bw[1] = ix + L * ((iy + L - 1) % L) + L * L * it;
        fw[2] = ix + L * iy + L * L * ((it + 1) % Lt);
        bw[2] = ix + L * iy + L * L * ((it + Lt - 1) % Lt);
        if (g->ibound == -1) {
            if (it == Lt - 1) bf[2] = -1.0;
            if (it == 0) bb[2] = -1.0;
        }
        flav2 Y = wall_matrix(chi, kvol, site, hpin, imass, 0);
        flav2 Yp = wall_matrix(chi, kvol, site, hpin, imass, 1);
        flav2 Wl, Wf; /* wall blocks at s=Ls and at s=1 */
        if (!dagger) { Wl = Y; Wf = Yp; }
        else { Wl = adjoint2(Yp); Wf = adjoint2(Y); }

        for (long s = 0; s < Ls; s++) {
            for (int f = 0; f < 2; f++) {
                for (int d = 0; d < 4; d++) {
                    long o = s + Ls * site + plane * (d + 4 * f);
                    double are = diag * r[2 * o], aim = diag * r[2 * o + 1];
                    for (int mu = 0; mu < 3; mu++) {
                        for (int dir = 0; dir < 2; dir++) {
                            long nb = dir == 0 ? fw[mu] : bw[mu];
                            double bc = dir == 0 ? bf[mu] : bb[mu];
                            /* forward: (1 - sig g), backward: (1 + sig g) */
                            double sgn = dir == 0 ? -sig : sig;
                            long od = s + Ls * nb + plane * (d + 4 * f);
                            long oc = s + Ls * nb + plane * (gcol[mu][d] + 4 * f);
                            double pr = r[2 * od], pi = r[2 * od + 1];
                            double qr = r[2 * oc], qi = r[2 * oc + 1];
                            double gr = gre[mu][d], gi = gim[mu][d];
                            double hr = pr + sgn * (gr * qr - gi * qi);
                            double hi = pi + sgn * (gr * qi + gi * qr);
                            are -= 0.5 * bc * hr;
                            aim -= 0.5 * bc * hi;
                        }
                    }
                    /* fifth-dimension hop */
                    int upper = d < 2; /* P+ components */
                    int take_next = dagger ? upper : !upper;
                    if (take_next && s < Ls - 1) {
                        long on = (s + 1) + Ls * site + plane * (d + 4 * f);
                        are -= r[2 * on]; aim -= r[2 * on + 1];
                    }
                    if (!take_next && s > 0) {
                        long on = (s - 1) + Ls * site + plane * (d + 4 * f);
                        are -= r[2 * on]; aim -= r[2 * on + 1];
                    }
                    /* wall coupling */
                    const flav2 *W = NULL; long src_s = -1;
                    if (s == Ls - 1 && take_next) { W = &Wl; src_s = 0; }
                    if (s == 0 && !take_next) { W = &Wf; src_s = Ls - 1; }
                    if (W && (imass == 1 || imass == 3)) {
                        for (int h = 0; h < 2; h++) {
                            long ow = src_s + Ls * site + plane * (d + 4 * h);
                            double wr = W->m[f][h][0], wi = W->m[f][h][1];
                            are += wr * r[2 * ow] - wi * r[2 * ow + 1];
                            aim += wr * r[2 * ow + 1] + wi * r[2 * ow];
                        }
                    }
                    phi[2 * o] = are;
                    phi[2 * o + 1] = aim;
                }
            }
        }
    }
}

/* ---- force bilinears ---------------------------------------------------- */
void orc_derivs(const orc_lat *g, double *F, const double *y, const double *z,
                double anum, int imass) {
    const long kvol = kvol_of(g), Ls = g->Ls, plane = Ls * kvol;
    const double t = -2.0 * anum;
    if (imass != 1 && imass != 3) return;
#pragma omp parallel for schedule(static)
    for (long site = 0; site < kvol; site++) {
        for (int pair = 0; pair < 2; pair++) {
            /* d=0,1: conj(y at s=Ls) z at s=1; d=2,3: conj(y at s=1) z at s=Ls */
            long sy = pair == 0 ? Ls - 1 : 0, sz = pair == 0 ? 0 : Ls - 1;
            double sgn = pair == 0 ? 1.0 : -1.0; /* imass=3 sign flip */
            for (int dd = 0; dd < 2; dd++) {
                int d = 2 * pair + dd;
                long oy0 = sy + Ls * site + plane * (d + 0), oy1 = sy + Ls * site + plane * (d + 4);
                long oz0 = sz + Ls * site + plane * (d + 0), oz1 = sz + Ls * site + plane * (d + 4);
                double y0r = y[2 * oy0], y0i = -y[2 * oy0 + 1];
                double y1r = y[2 * oy1], y1i = -y[2 * oy1 + 1];
                double z0r = z[2 * oz0], z0i = z[2 * oz0 + 1];
                double z1r = z[2 * oz1], z1i = z[2 * oz1 + 1];
                /* A = conj(y_f1) z_f2, B = conj(y_f2) z_f1, C = conj(y_f1) z_f1, D = conj(y_f2) z_f2 */
                double Ar = y0r * z1r - y0i * z1i, Ai = y0r * z1i + y0i * z1r;
                double Br = y1r * z0r - y1i * z0i, Bi = y1r * z0i + y1i * z0r;
                double Cr = y0r * z0r - y0i * z0i, Ci = y0r * z0i + y0i * z0r;
                double Dr = y1r * z1r - y1i * z1i, Di = y1r * z1i + y1i * z1r;
                if (imass == 1) {
                    F[site] -= t * (Ar + Br);
                    F[site + kvol] -= t * (Ai - Bi);
                    F[site + 2 * kvol] -= t * (Cr - Dr);
                } else {
                    F[site] -= sgn * t * (Ai + Bi);
                    F[site + kvol] += sgn * t * (Ar - Br);
                    F[site + 2 * kvol] -= sgn * t * (Ci - Di);
                }
            }
        }
    }
}

/* ---- multi-shift Lanczos ------------------------------------------------ */
/* niterc of rhmc/cheisdwf.f:603 (5000); settable so that tests can compare
 * solves stopped at a small cap */
static int g_lz_cap = 5000;
void orc_set_lanczos_cap(int niterc) { g_lz_cap = niterc > 0 ? niterc : 5000; }

int orc_qmrherm_shifts(const orc_lat *g, const double *phi, double res,
                       const double *anum, const double *aden, int ndiag,
                       int iflag, const double *chi, double hpin, int imass,
                       double *x, double *dSdchi, double *xk) {
    const long n = kferm_of(g);
    const int niterc = g_lz_cap;
    double resid = sqrt((double)n * res * res);
    double *r = (double *)malloc(sizeof(double) * 2 * n);
    double *q = (double *)malloc(sizeof(double) * 2 * n);
    double *qm1 = (double *)calloc(2 * n, sizeof(double));
    double *v = (double *)malloc(sizeof(double) * 2 * n);
    double *w = (double *)malloc(sizeof(double) * 2 * n);
    double *p = (double *)malloc(sizeof(double) * 2 * n * ndiag);
    double *x1 = (double *)malloc(sizeof(double) * 2 * n * ndiag);
    double *alpha = (double *)malloc(sizeof(double) * ndiag);
    double *d = (double *)malloc(sizeof(double) * ndiag);
    double *rho = (double *)malloc(sizeof(double) * ndiag);
    double *amu = (double *)malloc(sizeof(double) * ndiag);

    memcpy(r, phi, sizeof(double) * 2 * n);
    for (long i = 0; i < 2 * n; i++) x[i] = anum[0] * phi[i];
    double beta = sqrt(orc_norm2(n, r));
    const double phimod = beta;
    int itercg = 0;
    for (int iter = 1; iter <= niterc; iter++) {
        itercg++;
        const double ib = 1.0 / beta;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < 2 * n; i++) q[i] = r[i] * ib;
        orc_dslash(g, v, q, chi, hpin, imass, 0);
        orc_dslash(g, w, v, chi, hpin, imass, 1);
        double at = orc_redot(n, q, w);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < 2 * n; i++) {
            r[i] = w[i] - at * q[i] - beta * qm1[i];
            qm1[i] = q[i];
        }
        double beta0 = beta;
        beta = sqrt(orc_norm2(n, r));
        for (int k = 0; k < ndiag; k++) alpha[k] = at + aden[k];
        if (iter == 1) {
            for (int k = 0; k < ndiag; k++) { d[k] = alpha[k]; rho[k] = beta0 / alpha[k]; }
#pragma omp parallel for schedule(static)
            for (long i = 0; i < 2 * n; i++)
                for (int k = 0; k < ndiag; k++) {
                    p[2 * n * k + i] = q[i];
                    x1[2 * n * k + i] = rho[k] * q[i];
                }
        } else {
            for (int k = 0; k < ndiag; k++) {
                amu[k] = beta0 / d[k];
                double dm1 = d[k];
                d[k] = alpha[k] - beta0 * amu[k];
                rho[k] = -amu[k] * dm1 * rho[k] / d[k];
            }
#pragma omp parallel for schedule(static)
            for (long i = 0; i < 2 * n; i++)
                for (int k = 0; k < ndiag; k++) {
                    double pk = q[i] - amu[k] * p[2 * n * k + i];
                    p[2 * n * k + i] = pk;
                    x1[2 * n * k + i] += rho[k] * pk;
                }
            double rhomax = 0.0;
            for (int k = 0; k < ndiag; k++) {
                double a = fabs(phimod * rho[k]);
                if (k == 0 || a > rhomax) rhomax = a;
            }
            if (rhomax < 2.0 * resid) break;
        }
    }
    if (iflag == 0) {
        for (int k = 0; k < ndiag; k++) {
            const double a = anum[k + 1];
#pragma omp parallel for schedule(static)
            for (long i = 0; i < 2 * n; i++) x[i] += a * x1[2 * n * k + i];
        }
    } else {
        for (int k = 0; k < ndiag; k++) {
            orc_dslash(g, v, x1 + 2 * n * k, chi, hpin, imass, 0);
            orc_derivs(g, dSdchi, x1 + 2 * n * k, v, anum[k + 1], imass);
        }
    }
    if (xk) memcpy(xk, x1, sizeof(double) * 2 * n * ndiag);
    free(r); free(q); free(qm1); free(v); free(w); free(p); free(x1);
    free(alpha); free(d); free(rho); free(amu);
    return itercg;
}

int orc_qmrherm(const orc_lat *g, const double *phi, double res,
                const double *anum, const double *aden, int ndiag, int iflag,
                const double *chi, double hpin, int imass,
                double *x, double *dSdchi) {
    return orc_qmrherm_shifts(g, phi, res, anum, aden, ndiag, iflag, chi, hpin,
                              imass, x, dSdchi, NULL);
}

/* ---- CGNR --------------------------------------------------------------- */
int orc_congrad(const orc_lat *g, const double *phi, double res,
                const double *chi, double hpin, int imass, double *x) {
    const long n = kferm_of(g);
    const long niterc = (long)g->Ls * kvol_of(g);
    const double resid = (double)n * res * res;
    double *x1 = (double *)malloc(sizeof(double) * 2 * n);
    double *x2 = (double *)malloc(sizeof(double) * 2 * n);
    double *p = (double *)malloc(sizeof(double) * 2 * n);
    double *r = (double *)malloc(sizeof(double) * 2 * n);
    double alpha = 1.0, alphan = 0.0, betad = 1.0;
    int itercg = 0;
    for (long nx = 1; nx <= niterc; nx++) {
        itercg++;
        if (nx == 1) {
            memcpy(p, x, sizeof(double) * 2 * n);
            memcpy(r, phi, sizeof(double) * 2 * n);
            betad = 1.0;
            alpha = 1.0;
        }
        orc_dslash(g, x1, p, chi, hpin, imass, 0);
        if (nx != 1) {
            double alphad = orc_norm2(n, x1);
            alpha = alphan / alphad;
#pragma omp parallel for schedule(static)
            for (long i = 0; i < 2 * n; i++) x[i] += alpha * p[i];
        }
        orc_dslash(g, x2, x1, chi, hpin, imass, 1);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < 2 * n; i++) r[i] -= alpha * x2[i];
        double betan = orc_norm2(n, r);
        double beta = betan / betad;
        betad = betan;
        alphan = betan;
        if (nx == 1) beta = 0.0;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < 2 * n; i++) p[i] = r[i] + beta * p[i];
        if (betan < resid) break;
    }
    free(x1); free(x2); free(p); free(r);
    return itercg;
}

/* Exactly niter CGNR iterations of the same recurrences from x = 0 (used to
 * time a bounded sample of a solve). */
void orc_cg_iterations(const orc_lat *g, const double *phi, const double *chi,
                       double hpin, int imass, int niter, double *x) {
    const long n = kferm_of(g);
    double *x1 = (double *)malloc(sizeof(double) * 2 * n);
    double *x2 = (double *)malloc(sizeof(double) * 2 * n);
    double *p = (double *)malloc(sizeof(double) * 2 * n);
    double *r = (double *)malloc(sizeof(double) * 2 * n);
    memset(x, 0, sizeof(double) * 2 * n);
    memcpy(r, phi, sizeof(double) * 2 * n);
    memcpy(p, phi, sizeof(double) * 2 * n);
    double rn = orc_norm2(n, r);
    for (int it = 0; it < niter; it++) {
        orc_dslash(g, x1, p, chi, hpin, imass, 0);
        double alpha = rn / orc_norm2(n, x1);
        orc_dslash(g, x2, x1, chi, hpin, imass, 1);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < 2 * n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * x2[i]; }
        double rn2 = orc_norm2(n, r);
        double beta = rn2 / rn;
        rn = rn2;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < 2 * n; i++) p[i] = r[i] + beta * p[i];
    }
    free(x1); free(x2); free(p); free(r);
}

/* ---- force and Hamiltonian ---------------------------------------------- */
void orc_force(const orc_lat *g, const double *phi, double res1,
               const double *chi, double beta,
               const double *anum, const double *aden,
               const double *bnum, const double *bden, int ndiag,
               double hpin, int imass, double *dSdchi, double *x, int *its) {
    const long kvol = kvol_of(g), n = kferm_of(g);
    for (long i = 0; i < 3 * kvol; i++) dSdchi[i] = 2.0 * beta * chi[i];
    double *x2 = (double *)malloc(sizeof(double) * 2 * n);
    int i1 = orc_qmrherm(g, phi, res1, bnum, bden, ndiag, 0, chi, hpin, imass, x, dSdchi);
    memcpy(x2, x, sizeof(double) * 2 * n);
    int i2 = orc_qmrherm(g, x2, res1, anum, aden, ndiag, 1, chi, hpin, imass, x, dSdchi);
    if (its) { its[0] = i1; its[1] = i2; }
    free(x2);
}

/* rhmc/cheisdwf.f:897-1058.  Spinor element (s, site, d, f) of the reference
 * layout: s + Ls*(site + kvol*(d + 4 f)). */
int orc_fermion(const orc_lat *g, const double *chi, double res, double hpin,
                int imass, int ittt0, double *cferm1, double *cferm2) {
    const int nsource = 5;
    const long kvol = kvol_of(g), n = kferm_of(g), L2 = (long)g->L * g->L;
    const int Lt = g->Lt, Ls = g->Ls;
    double *xi = (double *)calloc(2 * n, sizeof(double));
    double *phi = (double *)malloc(sizeof(double) * 2 * n);
    double *x[2];   /* solutions for Dirac source 3 and 4 */
    x[0] = (double *)malloc(sizeof(double) * 2 * n);
    x[1] = (double *)malloc(sizeof(double) * 2 * n);
    memset(cferm1, 0, sizeof(double) * 2 * 2 * Lt);
    memset(cferm2, 0, sizeof(double) * 2 * 2 * Lt);
    int iter = 0;
#define EL(s, site, d, f) ((long)(s) + Ls * ((long)(site) + kvol * ((long)(d) + 4L * (f))))
    for (int ks = 1; ks <= nsource; ks++) {
        const int ittt = (ittt0 + ks * (Lt / nsource)) % Lt + 1;   /* 1-based */
        for (int ids = 0; ids < 2; ids++) {                       /* Dirac source 3, 4 */
            memset(xi, 0, sizeof(double) * 2 * n);
            for (long i = 0; i < L2; i++) xi[2 * EL(0, (ittt - 1) * L2 + i, 2 + ids, 0)] = 1.0 / (double)L2;
            orc_dslash(g, phi, xi, chi, hpin, imass, 1);
            memcpy(x[ids], phi, sizeof(double) * 2 * n);
            iter += orc_congrad(g, phi, res, chi, hpin, imass, x[ids]);
        }
        /* cferm1 += sign 0.5 prop0L(idd, idd); cferm2 += sign gamval(3,idd)
         * prop00(idd, gamin(3,idd)) with gamval(3,3) = 0.5 i, gamin(3,3) = 1,
         * gamval(3,4) = -0.5 i, gamin(3,4) = 2 (rhmc/cheisdwf.f:1017-1036,
         * :1555-1563, scaled by akappa = 0.5 at :1591-1595) */
        for (int idd = 0; idd < 2; idd++) {
            const double gi = idd == 0 ? 0.5 : -0.5;   /* gamval = gi * i */
            const int dsink2 = idd;                     /* gamin - 1 */
            for (int it = 0; it < Lt; it++) {
                const int itt = (ittt + it - 1) % Lt + 1;
                const double sgn = itt >= ittt ? 1.0 : (double)g->ibound;
                const long ioff = (long)(itt - 1) * L2;
                for (int f = 0; f < 2; f++) {
                    double *c1 = cferm1 + 2 * (it + (long)Lt * f), *c2 = cferm2 + 2 * (it + (long)Lt * f);
                    for (long i = 0; i < L2; i++) {
                        const double *pL = x[idd] + 2 * EL(Ls - 1, ioff + i, 2 + idd, f);
                        c1[0] += sgn * 0.5 * pL[0];
                        c1[1] += sgn * 0.5 * pL[1];
                        const double *p0 = x[idd] + 2 * EL(0, ioff + i, dsink2, f);
                        /* (gi i) (a + i b) = -gi b + i gi a */
                        c2[0] += sgn * (-gi * p0[1]);
                        c2[1] += sgn * (gi * p0[0]);
                    }
                }
            }
        }
    }
#undef EL
    for (int k = 0; k < 2 * 2 * Lt; k++) {
        cferm1[k] /= nsource;
        cferm2[k] /= nsource;
    }
    free(xi); free(phi); free(x[0]); free(x[1]);
    return iter;
}

int orc_hmc_force(const orc_lat *g, const double *phi, double *x0, double res,
                  const double *chi, double beta, int imass, double *dSdchi) {
    const long n = kferm_of(g), kvol = kvol_of(g);
    double *x1 = (double *)malloc(sizeof(double) * 2 * n);
    double *x2 = (double *)malloc(sizeof(double) * 2 * n);
    memcpy(x1, x0, sizeof(double) * 2 * n);
    const int it = orc_congrad(g, phi, res, chi, 0.0, imass, x1);
    for (long i = 0; i < 2 * n; i++) x0[i] = 2.0 * x1[i] - x0[i];
    orc_dslash(g, x2, x1, chi, 0.0, imass, 0);
    for (long i = 0; i < 3 * kvol; i++) dSdchi[i] = 2.0 * beta * chi[i];
    /* the reference subtracts the bilinears and doubles everything
     * (mh/cheisdwf_hmc.f:440-552): 2 (2 beta chi) + derivs with a = -1 */
    for (long i = 0; i < 3 * kvol; i++) dSdchi[i] *= 2.0;
    orc_derivs(g, dSdchi, x1, x2, -1.0, imass);
    free(x1); free(x2);
    return it;
}

int orc_hmc_hamilton(const orc_lat *g, const double *phi, double *x0, double res,
                     const double *chi, const double *pp, double beta, int imass,
                     double *out) {
    const long n = kferm_of(g), kvol = kvol_of(g);
    double hp = 0.0, hg = 0.0;
    for (long i = 0; i < 3 * kvol; i++) hp += pp[i] * pp[i];
    hp *= 0.5;
    for (long i = 0; i < 3 * kvol; i++) hg += chi[i] * chi[i];
    hg *= 2.0 * beta;
    const int it = orc_congrad(g, phi, res, chi, 0.0, imass, x0);
    const double hf = orc_redot(n, phi, x0);   /* n complex numbers */
    out[0] = hg + hp + hf;
    out[1] = hg;
    out[2] = hp;
    out[3] = hg + hf;
    return it;
}

int orc_hamilton(const orc_lat *g, const double *phi, double res2,
                 const double *chi, const double *pp, double beta,
                 const double *bnum, const double *bden, int ndiag,
                 double hpin, int imass, double *out, double *x) {
    const long kvol = kvol_of(g), n = kferm_of(g);
    double hp = 0.0, hg = 0.0, hf = 0.0;
    for (long i = 0; i < 3 * kvol; i++) hp += pp[i] * pp[i];
    hp *= 0.5;
    for (long i = 0; i < 3 * kvol; i++) hg += chi[i] * chi[i];
    hg *= beta;
    int it = orc_qmrherm(g, phi, res2, bnum, bden, ndiag, 0, chi, hpin, imass, x, NULL);
    hf = orc_redot(n, phi, x);
    out[0] = hg + hp + hf;
    out[1] = hg;
    out[2] = hp;
    out[3] = hg + hf;
    return it;
}

/* ---- random numbers ------------------------------------------------------ */
/* 48-bit multiplicative LCG, multiplier 2651554*2^24 + 15184245
 * (rhmc/cheisdwf.f:1765-1801). */
static const uint64_t RMUL = 2651554ULL * 16777216ULL + 15184245ULL;
static const uint64_t RMASK = (1ULL << 48) - 1;

void orc_rng_set(orc_rng *r, double seed) {
    r->s = ((uint64_t)seed) & RMASK;
    r->idum = -1;
    r->yran = 0.0;
    r->yrun = 0.0;
    memset(r->v, 0, sizeof r->v);
}
double orc_rng_get(const orc_rng *r) { return (double)r->s; }

double orc_rranf(orc_rng *r) {
    /* (s * RMUL) mod 2^48; unsigned 64-bit wrap keeps the low 48 bits exact */
    r->s = (r->s * RMUL) & RMASK;
    return (double)r->s * (1.0 / 281474976710656.0);
}

/* Bays-Durham shuffle (rhmc/cheisdwf.f:1722-1751); the two pointer
 * streams yran and yrun share one table. */
double orc_rano(orc_rng *r, int stream) {
    double *y = stream ? &r->yrun : &r->yran;
    if (r->idum < 0) {
        r->idum = 1;
        for (int j = 0; j < 97; j++) (void)orc_rranf(r);
        for (int j = 0; j < 97; j++) r->v[j] = orc_rranf(r);
        *y = orc_rranf(r);
    }
    int j = 1 + (int)(97.0 * *y);
    if (j > 97) j = 97;
    if (j < 1) j = 1;
    *y = r->v[j - 1];
    double out = *y;
    r->v[j - 1] = orc_rranf(r);
    return out;
}

static const double TWO_PI = 6.283185307179586476925286766559;

/* all radii first, then all angles (rhmc/cheisdwf.f:1703-1716) */
void orc_gauss0(orc_rng *r, double *ps, int kvol) {
    for (int i = 0; i < kvol; i++) ps[i + kvol] = sqrt(-log(orc_rano(r, 0)));
    for (int i = 0; i < kvol; i++) {
        double th = TWO_PI * orc_rano(r, 0);
        ps[i] = ps[i + kvol] * sin(th);
        ps[i + kvol] = ps[i + kvol] * cos(th);
    }
}

void orc_gaussp(orc_rng *r, double *ps, int kvol) {
    for (int i = 0; i < kvol; i++) ps[i + kvol] = sqrt(-2.0 * log(orc_rano(r, 0)));
    for (int i = 0; i < kvol; i++) {
        double th = TWO_PI * orc_rano(r, 0);
        ps[i] = ps[i + kvol] * sin(th);
        ps[i + kvol] = ps[i + kvol] * cos(th);
    }
}

/* ---- RHMC trajectories (rhmc/cheisdwf.f:207-413) ------------------------ */
void orc_rhmc(const orc_lat *g, const orc_rhmc_par *P, orc_rng *rng,
              double *chi, int isweep0, int ntraj, orc_traj_out *out) {
    const long kvol = kvol_of(g), n = kferm_of(g), Ls = g->Ls;
    double *chit = (double *)malloc(sizeof(double) * 3 * kvol);
    double *pp = (double *)malloc(sizeof(double) * 3 * kvol);
    double *F = (double *)malloc(sizeof(double) * 3 * kvol);
    double *ps = (double *)malloc(sizeof(double) * 2 * kvol);
    double *R = (double *)malloc(sizeof(double) * 2 * n);
    double *Phi = (double *)malloc(sizeof(double) * 2 * n);
    double *X = (double *)malloc(sizeof(double) * 2 * n);
    const double proby = 1.0 / (double)P->iterl;
    const long plane = Ls * kvol;
    for (int t = 0; t < ntraj; t++) {
        int isweep = isweep0 + t;
        orc_traj_out *o = &out[t];
        memset(o, 0, sizeof *o);
        memcpy(chit, chi, sizeof(double) * 3 * kvol);
        /* pseudofermion heatbath: flavour, Dirac, s order (:221-230) */
        for (int f = 0; f < 2; f++)
            for (int d = 0; d < 4; d++)
                for (long s = 0; s < Ls; s++) {
                    orc_gauss0(rng, ps, (int)kvol);
                    for (long i = 0; i < kvol; i++) {
                        long off = s + Ls * i + plane * (d + 4 * f);
                        R[2 * off] = ps[i];
                        R[2 * off + 1] = ps[i + kvol];
                    }
                }
        o->it_pf = orc_qmrherm(g, R, P->rescga, P->anum4, P->aden4, P->n4, 0,
                               chit, P->hpin, P->imass, Phi, NULL);
        orc_gaussp(rng, ps, (int)kvol);
        for (long i = 0; i < kvol; i++) { pp[i] = ps[i]; pp[i + kvol] = ps[i + kvol]; }
        orc_gaussp(rng, ps, (int)kvol);
        for (long i = 0; i < kvol; i++) pp[i + 2 * kvol] = ps[i];
        double hv[4];
        o->it_h0 = orc_hamilton(g, Phi, P->rescga, chit, pp, P->beta, P->bnum2,
                                P->bden2, P->n2, P->hpin, P->imass, hv, X);
        o->H0 = hv[0]; o->S0 = hv[3];
        int its[2];
        orc_force(g, Phi, P->rescgg, chit, P->beta, P->anum2g, P->aden2g, P->bnum2g,
                  P->bden2g, P->n2g, P->hpin, P->imass, F, X, its);
        o->it_force_total += its[0] + its[1];
        for (long i = 0; i < 3 * kvol; i++) pp[i] -= 0.5 * P->dt * F[i];
        for (int iter = 1; iter <= P->itermax; iter++) {
            for (long i = 0; i < 3 * kvol; i++) chit[i] += P->dt * pp[i];
            orc_force(g, Phi, P->rescgg, chit, P->beta, P->anum2g, P->aden2g, P->bnum2g,
                      P->bden2g, P->n2g, P->hpin, P->imass, F, X, its);
            o->it_force_total += its[0] + its[1];
            o->md_steps = iter;
            double ytest = orc_rano(rng, 1);
            if (ytest < proby) {
                for (long i = 0; i < 3 * kvol; i++) pp[i] -= 0.5 * P->dt * F[i];
                break;
            }
            for (long i = 0; i < 3 * kvol; i++) pp[i] -= P->dt * F[i];
        }
        o->it_h1 = orc_hamilton(g, Phi, P->rescga, chit, pp, P->beta, P->bnum2,
                                P->bden2, P->n2, P->hpin, P->imass, hv, X);
        o->H1 = hv[0]; o->S1 = hv[3];
        o->dH = o->H0 - o->H1;
        double y = exp(o->dH);
        double xr = orc_rano(rng, 1);
        o->accepted = xr < y;
        if (o->accepted) memcpy(chi, chit, sizeof(double) * 3 * kvol);
        /* the measurement routine draws one number from the main stream every
         * iprint trajectories (rhmc/cheisdwf.f:359,929) */
        if (P->iprint > 0 && isweep % P->iprint == 0) (void)orc_rano(rng, 0);
    }
    free(chit); free(pp); free(F); free(ps); free(R); free(Phi); free(X);
}
