/*
 * oracle.c -- FP64 CPU restatement of the many-sphere Stokes mobility operator.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * CPU-baseline leg of the benchmark.  Nothing in the product links this file.
 *
 * Each function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference tree).  The spherical-harmonic numerics of
 * the un-vendored SCTL library are re-derived from the mathematics (see the
 * header comment of oracle.h); the STKFMM far field is replaced by the exact
 * direct sum; Belos GMRES by a restarted GMRES with two-pass modified
 * Gram-Schmidt ("IMGS", mobilitySolver.yaml:15) and Givens rotations.
 *
 * Pinning (what checks each function; tests/test_oracle_*.py):
 *   orc_gl                  closed-form moments, exactness on polynomials
 *   orc_solid_harmonics     finite differences of the ladder identities,
 *                           explicit low-order closed forms, Laplace equation
 *   orc_sh_analysis/synth   idempotence, exactness on band-limited fields,
 *                           rotation covariance, orthonormality by quadrature
 *   orc_sh_self_spectral    translation/rotation closed forms (S[c] = 2R/3 c,
 *                           S[c x n] = R/3 c x n, K'_ext[rho0] = -rho0), jump
 *                           relation, agreement with the exterior expansion at r=1
 *   orc_sh_eval_exterior    brute-force quadrature of the Stokeslet/stresslet over
 *                           an upsampled grid (the Shexp_test.cpp:109-524 method)
 *   orc_mobility_solve      single sphere V=F/(6 pi mu R), Omega=T/(8 pi mu R^3);
 *                           Stimson-Jeffery two-sphere drag; rigid-motion
 *                           covariance; symmetric mobility matrix
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.14159265358979323846264338327950288
#define MAXP 64

typedef double complex cplx;

static inline int tri(int n, int m) { return n * (n + 1) / 2 + m; }
static inline int tri_size(int P) { return (P + 1) * (P + 2) / 2; }

int orc_num_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------- */
/* Gauss-Legendre rule (SimToolbox Gauss_Legendre_Nodes_and_Weights, called at
 * Sphere/Shexp.cpp:221).  Newton iteration on P_n from the Tricomi guess. */
void orc_gl(int n, double *x, double *w) {
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = cos(PI * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) {
                p0 = 1.0;
                p1 = z;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        /* recompute derivative at the converged node */
        {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
        }
        x[n - 1 - i] = z;
        x[i] = -z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}

/* ------------------------------------------------------------------------- */
/* Irregular solid harmonics D_n^m(x) = (dx + i dy)^m dz^(n-m) (1/r), computed by
 *   D_m^m     = -(2m-1) (x+iy)/r^2 D_{m-1}^{m-1}
 *   D_{n+1}^m = -[(2n+1) z D_n^m + (n^2-m^2) D_{n-1}^m] / r^2              */
void orc_solid_harmonics(int P, const double *xp, double *Dout) {
    cplx *D = (cplx *)Dout;
    const double x = xp[0], y = xp[1], z = xp[2];
    const double r2 = x * x + y * y + z * z;
    const double ir2 = 1.0 / r2;
    const cplx xy = x + I * y;
    D[0] = 1.0 / sqrt(r2);
    for (int m = 1; m <= P; m++) D[tri(m, m)] = -(2.0 * m - 1.0) * xy * ir2 * D[tri(m - 1, m - 1)];
    for (int m = 0; m < P; m++) {
        D[tri(m + 1, m)] = -((2.0 * m + 1.0) * z * D[tri(m, m)]) * ir2;
        for (int n = m + 1; n < P; n++)
            D[tri(n + 1, m)] =
                -((2.0 * n + 1.0) * z * D[tri(n, m)] + ((double)n * n - (double)m * m) * D[tri(n - 1, m)]) * ir2;
    }
}

/* Squared norm of y_nm = D_n^m on the unit sphere: 4 pi (n-m)! (n+m)! / (2n+1). */
static double norm_nm(int n, int m) {
    double f = 4.0 * PI / (2.0 * n + 1.0);
    for (int k = 2; k <= n - m; k++) f *= k;
    for (int k = 2; k <= n + m; k++) f *= k;
    return f;
}

/* Unit grid of Sphere/Shexp.cpp:231-246 with the poles dropped
 * (BIE/SphereSTKSHOperator.cpp:112-114): cos(theta_j) = x^GL_{p-j},
 * phi_k = 2 pi k/(2p+2), index j(2p+2)+k, weight 2 pi/(2p+2) w^GL_{p-j}. */
void orc_unit_grid(int p, double *eta, double *w) {
    double xg[MAXP + 2], wg[MAXP + 2];
    orc_gl(p + 1, xg, wg);
    const int np = 2 * p + 2;
    for (int j = 0; j <= p; j++) {
        const double ct = xg[p - j];
        const double st = sqrt(1.0 - ct * ct);
        for (int k = 0; k < np; k++) {
            const double ph = 2.0 * PI * k / np;
            const int q = j * np + k;
            eta[3 * q + 0] = st * cos(ph);
            eta[3 * q + 1] = st * sin(ph);
            eta[3 * q + 2] = ct;
            w[q] = 2.0 * PI / np * wg[p - j];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Per-order tables of y, grad_s y and r x grad_s y at the unit grid points.    */
typedef struct {
    int p, npts, Q;
    double *eta, *w;
    cplx *Y;  /* npts*Q */
    cplx *GS; /* npts*Q*3 */
    cplx *XT; /* npts*Q*3 */
    double *N;
} sh_table;

static sh_table *g_tables[MAXP + 1];

/* Cartesian gradient of D_n^m from D_{n+1}^*:  dz D_n^m = D_{n+1}^m,
 *   m>=1: dx = (D_{n+1}^{m+1} - D_{n+1}^{m-1})/2, dy = (D_{n+1}^{m+1} + D_{n+1}^{m-1})/(2i)
 *   m=0:  dx = Re D_{n+1}^1, dy = Im D_{n+1}^1                                  */
static void grad_D(const cplx *D, int n, int m, cplx g[3]) {
    g[2] = D[tri(n + 1, m)];
    if (m >= 1) {
        g[0] = 0.5 * (D[tri(n + 1, m + 1)] - D[tri(n + 1, m - 1)]);
        g[1] = -0.5 * I * (D[tri(n + 1, m + 1)] + D[tri(n + 1, m - 1)]);
    } else {
        g[0] = creal(D[tri(n + 1, 1)]);
        g[1] = cimag(D[tri(n + 1, 1)]);
    }
}

static sh_table *get_table(int p) {
    sh_table *t = g_tables[p];
    if (t) return t;
#pragma omp critical(orc_tables)
    {
        if (!g_tables[p]) {
            t = (sh_table *)calloc(1, sizeof(sh_table));
            t->p = p;
            t->npts = (p + 1) * (2 * p + 2);
            t->Q = tri_size(p);
            t->eta = (double *)malloc(sizeof(double) * 3 * t->npts);
            t->w = (double *)malloc(sizeof(double) * t->npts);
            t->Y = (cplx *)malloc(sizeof(cplx) * t->npts * t->Q);
            t->GS = (cplx *)malloc(sizeof(cplx) * t->npts * t->Q * 3);
            t->XT = (cplx *)malloc(sizeof(cplx) * t->npts * t->Q * 3);
            t->N = (double *)malloc(sizeof(double) * t->Q);
            orc_unit_grid(p, t->eta, t->w);
            for (int n = 0; n <= p; n++)
                for (int m = 0; m <= n; m++) t->N[tri(n, m)] = norm_nm(n, m);
            cplx *D = (cplx *)malloc(sizeof(cplx) * tri_size(p + 1));
            for (int q = 0; q < t->npts; q++) {
                const double *e = t->eta + 3 * q;
                orc_solid_harmonics(p + 1, e, (double *)D);
                for (int n = 0; n <= p; n++)
                    for (int m = 0; m <= n; m++) {
                        const int k = tri(n, m);
                        cplx g[3];
                        grad_D(D, n, m, g);
                        const cplx y = D[k];
                        cplx *gs = t->GS + ((long)q * t->Q + k) * 3;
                        cplx *xt = t->XT + ((long)q * t->Q + k) * 3;
                        t->Y[(long)q * t->Q + k] = y;
                        for (int c = 0; c < 3; c++) gs[c] = g[c] + (n + 1.0) * y * e[c];
                        xt[0] = e[1] * g[2] - e[2] * g[1];
                        xt[1] = e[2] * g[0] - e[0] * g[2];
                        xt[2] = e[0] * g[1] - e[1] * g[0];
                    }
            }
            free(D);
            g_tables[p] = t;
        }
        t = g_tables[p];
    }
    return t;
}

/* Grid -> coefficients (the role of Shexp::calcSpectralCoeff, Shexp.cpp:408-447,
 * body frame): orthogonal projection in the quadrature inner product, exact for
 * fields of vector-SH degree <= p. */
void orc_sh_analysis(int p, const double *vals, double *coef) {
    sh_table *t = get_table(p);
    const int Q = t->Q;
    cplx *Fr = (cplx *)coef, *Fs = Fr + Q, *Ft = Fs + Q;
    for (int k = 0; k < 3 * Q; k++) Fr[k] = 0;
    for (int q = 0; q < t->npts; q++) {
        const double *v = vals + 3 * q;
        const double *e = t->eta + 3 * q;
        const double w = t->w[q];
        const double vr = v[0] * e[0] + v[1] * e[1] + v[2] * e[2];
        for (int k = 0; k < Q; k++) {
            const cplx *gs = t->GS + ((long)q * Q + k) * 3;
            const cplx *xt = t->XT + ((long)q * Q + k) * 3;
            Fr[k] += w * vr * conj(t->Y[(long)q * Q + k]);
            Fs[k] += w * (v[0] * conj(gs[0]) + v[1] * conj(gs[1]) + v[2] * conj(gs[2]));
            Ft[k] += w * (v[0] * conj(xt[0]) + v[1] * conj(xt[1]) + v[2] * conj(xt[2]));
        }
    }
    for (int n = 0; n <= p; n++)
        for (int m = 0; m <= n; m++) {
            const int k = tri(n, m);
            Fr[k] /= t->N[k];
            if (n == 0) {
                Fs[k] = Ft[k] = 0;
            } else {
                Fs[k] /= (n * (n + 1.0)) * t->N[k];
                Ft[k] /= (n * (n + 1.0)) * t->N[k];
            }
            if (m == 0) {
                Fr[k] = creal(Fr[k]);
                Fs[k] = creal(Fs[k]);
                Ft[k] = creal(Ft[k]);
            }
        }
}

/* Coefficients -> grid (the role of Shexp::calcGridValue, Shexp.cpp:342-379). */
void orc_sh_synthesis(int p, const double *coef, double *vals) {
    sh_table *t = get_table(p);
    const int Q = t->Q;
    const cplx *Fr = (const cplx *)coef, *Fs = Fr + Q, *Ft = Fs + Q;
    for (int q = 0; q < t->npts; q++) {
        const double *e = t->eta + 3 * q;
        double v[3] = {0, 0, 0};
        for (int n = 0; n <= p; n++)
            for (int m = 0; m <= n; m++) {
                const int k = tri(n, m);
                const double eps = m ? 2.0 : 1.0;
                const cplx *gs = t->GS + ((long)q * Q + k) * 3;
                const cplx *xt = t->XT + ((long)q * Q + k) * 3;
                const cplx ry = Fr[k] * t->Y[(long)q * Q + k];
                for (int c = 0; c < 3; c++) v[c] += eps * creal(ry * e[c] + Fs[k] * gs[c] + Ft[k] * xt[c]);
            }
        vals[3 * q + 0] = v[0];
        vals[3 * q + 1] = v[1];
        vals[3 * q + 2] = v[2];
    }
}

/* Per-degree spectra of the unit-sphere single layer (mu=1, kernel (1/8pi)G) and
 * of the traction of that flow (the reference's SCTL StokesEvalSL at r=1 and
 * StokesEvalKSelf ext/int, Shexp.cpp:531, :572), from Lamb's solution:
 *   Dn = (2n-1)(2n+1)(2n+3)
 *   S:      u_r = n(n+1)(2F_r+3F_s)/Dn, u_s = (3F_r+(2n^2+2n+3)F_s)/Dn, u_t = F_t/(2n+1)
 *   K_ext:  t_r = -n((4n^2+6n-1)F_r+3(n+1)F_s)/Dn,
 *           t_s = -(3F_r+(4n^3+6n^2-n+3)F_s)/Dn, t_t = -(n+2)F_t/(2n+1)
 *   K_int = K_ext + I,  K_PV = K_ext + I/2;  degree 0 (radial): S = 0, K_ext = 0. */
void orc_sh_self_spectral(int p, int which, const double *cin, double *cout) {
    const int Q = tri_size(p);
    const cplx *Fr = (const cplx *)cin, *Fs = Fr + Q, *Ft = Fs + Q;
    cplx *Gr = (cplx *)cout, *Gs = Gr + Q, *Gt = Gs + Q;
    for (int n = 0; n <= p; n++) {
        const double nn = n;
        const double Dn = (2 * nn - 1) * (2 * nn + 1) * (2 * nn + 3);
        for (int m = 0; m <= n; m++) {
            const int k = tri(n, m);
            cplx r = 0, s = 0, tt = 0;
            if (n >= 1) {
                if (which == 0) {
                    r = nn * (nn + 1) * (2 * Fr[k] + 3 * Fs[k]) / Dn;
                    s = (3 * Fr[k] + (2 * nn * nn + 2 * nn + 3) * Fs[k]) / Dn;
                    tt = Ft[k] / (2 * nn + 1);
                } else {
                    r = -nn * ((4 * nn * nn + 6 * nn - 1) * Fr[k] + 3 * (nn + 1) * Fs[k]) / Dn;
                    s = -(3 * Fr[k] + (4 * nn * nn * nn + 6 * nn * nn - nn + 3) * Fs[k]) / Dn;
                    tt = -(nn + 2) * Ft[k] / (2 * nn + 1);
                }
            }
            if (which == 2) {
                r += Fr[k];
                s += Fs[k];
                tt += Ft[k];
            } else if (which == 3) {
                r += 0.5 * Fr[k];
                s += 0.5 * Fs[k];
                tt += 0.5 * Ft[k];
            }
            Gr[k] = r;
            Gs[k] = s;
            Gt[k] = tt;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Exterior expansion.  A real harmonic function is stored as H = sum_{n,m>=0}
 * eps_m Re(h_nm D_n^m) (eps_0=1, eps_m=2).  Cartesian derivatives act on the
 * coefficients (degree +1):
 *   z: h'_{n+1,k} = h_{n,k}
 *   x: h'_{n+1,0} = -h_{n,1};      h'_{n+1,k} = (h_{n,k-1} - h_{n,k+1})/2
 *   y: h'_{n+1,0} = -i h_{n,1};    h'_{n+1,k} = -i (h_{n,k-1} + h_{n,k+1})/2
 * valid when the m=0 coefficients are real (enforced before each use). */
static inline cplx hget(const cplx *h, int P, int n, int k) {
    if (k < 0 || k > n || n > P) return 0;
    return h[tri(n, k)];
}

static void hderiv(const cplx *h, int P, int dir, cplx *out) {
    for (int k = 0; k < tri_size(P + 1); k++) out[k] = 0;
    for (int n = 0; n <= P; n++) {
        for (int k = 0; k <= n + 1; k++) {
            cplx v;
            if (dir == 2) {
                v = hget(h, P, n, k);
            } else if (dir == 0) {
                v = (k == 0) ? -hget(h, P, n, 1) : 0.5 * (hget(h, P, n, k - 1) - hget(h, P, n, k + 1));
            } else {
                v = (k == 0) ? -I * hget(h, P, n, 1) : -0.5 * I * (hget(h, P, n, k - 1) + hget(h, P, n, k + 1));
            }
            out[tri(n + 1, k)] = v;
        }
    }
    for (int n = 0; n <= P + 1; n++) out[tri(n, 0)] = creal(out[tri(n, 0)]);
}

static double heval(const cplx *h, int P, const cplx *D) {
    double s = 0;
    for (int n = 0; n <= P; n++) {
        s += creal(h[tri(n, 0)] * D[tri(n, 0)]);
        for (int m = 1; m <= n; m++) s += 2.0 * creal(h[tri(n, m)] * D[tri(n, m)]);
    }
    return s;
}

/* Exterior single-layer flow of a unit-sphere density (the role of SCTL
 * StokesEvalSL / StokesEvalKL called at Shexp.cpp:531, :613):
 *   u = grad H1 + r^2 grad H2 + x H3 + x cross grad H4,   p = P
 * with, per degree n>=1 and A_n=(2-n)/(2n(2n-1)), B_n=(n+1)/(n(2n-1)):
 *   a = n(F_r+(n+3)F_s)/(2(2n+1)(2n+3)),  b = n(F_r+(n+1)F_s)/(2n+1),  alpha = F_t/(2n+1)
 *   H1 = a, H2 = A_n b, H3 = B_n b, H4 = alpha, P = b  (coefficients on D_n^m). */
void orc_sh_eval_exterior(int p, const double *coef, int ntrg, const double *xi, const double *nu, double *uout,
                          double *tout) {
    const int Q = tri_size(p);
    const int Q1 = tri_size(p + 1), Q2 = tri_size(p + 2);
    const cplx *Fr = (const cplx *)coef, *Fs = Fr + Q, *Ft = Fs + Q;
    cplx *H = (cplx *)calloc(5 * Q, sizeof(cplx));
    cplx *H1 = H, *H2 = H + Q, *H3 = H + 2 * Q, *H4 = H + 3 * Q, *HP = H + 4 * Q;
    for (int n = 1; n <= p; n++) {
        const double nn = n;
        const double An = (2 - nn) / (2 * nn * (2 * nn - 1)), Bn = (nn + 1) / (nn * (2 * nn - 1));
        for (int m = 0; m <= n; m++) {
            const int k = tri(n, m);
            const cplx a = nn * (Fr[k] + (nn + 3) * Fs[k]) / (2 * (2 * nn + 1) * (2 * nn + 3));
            const cplx b = nn * (Fr[k] + (nn + 1) * Fs[k]) / (2 * nn + 1);
            H1[k] = a;
            H2[k] = An * b;
            H3[k] = Bn * b;
            H4[k] = Ft[k] / (2 * nn + 1);
            HP[k] = b;
        }
    }
    /* gradients (degree p+1) and Hessians (degree p+2) */
    cplx *G = (cplx *)calloc(12 * Q1, sizeof(cplx)); /* [H1,H2,H3,H4] x [x,y,z] */
    cplx *Hs = (cplx *)calloc(18 * Q2, sizeof(cplx)); /* [H1,H2,H4] x [xx,xy,xz,yy,yz,zz] */
    const cplx *base[4] = {H1, H2, H3, H4};
    for (int f = 0; f < 4; f++)
        for (int d = 0; d < 3; d++) hderiv(base[f], p, d, G + (f * 3 + d) * Q1);
    const int hf[3] = {0, 1, 3};
    const int hd[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int f = 0; f < 3; f++)
        for (int e = 0; e < 6; e++)
            hderiv(G + (hf[f] * 3 + hd[e][0]) * Q1, p + 1, hd[e][1], Hs + (f * 6 + e) * Q2);

#pragma omp parallel if (ntrg > 64)
    {
        cplx *D = (cplx *)malloc(sizeof(cplx) * Q2);
#pragma omp for
        for (int j = 0; j < ntrg; j++) {
            const double *x = xi + 3 * j;
            orc_solid_harmonics(p + 2, x, (double *)D);
            double g[4][3];
            for (int f = 0; f < 4; f++)
                for (int d = 0; d < 3; d++) g[f][d] = heval(G + (f * 3 + d) * Q1, p + 1, D);
            const double h3 = heval(H3, p, D);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (uout) {
                double *u = uout + 3 * j;
                const double cr[3] = {x[1] * g[3][2] - x[2] * g[3][1], x[2] * g[3][0] - x[0] * g[3][2],
                                      x[0] * g[3][1] - x[1] * g[3][0]};
                for (int c = 0; c < 3; c++) u[c] = g[0][c] + r2 * g[1][c] + x[c] * h3 + cr[c];
            }
            if (tout) {
                const double *v = nu + 3 * j;
                double Hm[3][3][3];
                for (int f = 0; f < 3; f++)
                    for (int e = 0; e < 6; e++) {
                        const double val = heval(Hs + (f * 6 + e) * Q2, p + 2, D);
                        Hm[f][hd[e][0]][hd[e][1]] = val;
                        Hm[f][hd[e][1]][hd[e][0]] = val;
                    }
                const double pr = heval(HP, p, D);
                double H1n[3], H2n[3], H4n[3], H4vx[3];
                const double vx[3] = {v[1] * x[2] - v[2] * x[1], v[2] * x[0] - v[0] * x[2], v[0] * x[1] - v[1] * x[0]};
                for (int a = 0; a < 3; a++) {
                    H1n[a] = H2n[a] = H4n[a] = H4vx[a] = 0;
                    for (int b = 0; b < 3; b++) {
                        H1n[a] += Hm[0][a][b] * v[b];
                        H2n[a] += Hm[1][a][b] * v[b];
                        H4n[a] += Hm[2][a][b] * v[b];
                        H4vx[a] += Hm[2][a][b] * vx[b];
                    }
                }
                const double xv = x[0] * v[0] + x[1] * v[1] + x[2] * v[2];
                const double vg2 = v[0] * g[1][0] + v[1] * g[1][1] + v[2] * g[1][2];
                const double vg3 = v[0] * g[2][0] + v[1] * g[2][1] + v[2] * g[2][2];
                const double xh4[3] = {x[1] * H4n[2] - x[2] * H4n[1], x[2] * H4n[0] - x[0] * H4n[2],
                                       x[0] * H4n[1] - x[1] * H4n[0]};
                double *t = tout + 3 * j;
                for (int c = 0; c < 3; c++)
                    t[c] = -pr * v[c] + 2 * H1n[c] + 2 * xv * g[1][c] + 2 * x[c] * vg2 + 2 * r2 * H2n[c] +
                           2 * h3 * v[c] + x[c] * vg3 + xv * g[2][c] + xh4[c] + H4vx[c];
            }
        }
        free(D);
    }
    free(H);
    free(G);
    free(Hs);
}

/* ------------------------------------------------------------------------- */
/* Quaternion (w,x,y,z) rotation; inv applies the conjugate. */
static inline void qrot(const double *q, const double *v, double *o, int inv) {
    const double w = q[0];
    const double s = inv ? -1.0 : 1.0;
    const double ux = s * q[1], uy = s * q[2], uz = s * q[3];
    const double tx = 2 * (uy * v[2] - uz * v[1]);
    const double ty = 2 * (uz * v[0] - ux * v[2]);
    const double tz = 2 * (ux * v[1] - uy * v[0]);
    o[0] = v[0] + w * tx + (uy * tz - uz * ty);
    o[1] = v[1] + w * ty + (uz * tx - ux * tz);
    o[2] = v[2] + w * tz + (ux * ty - uy * tx);
}

struct orc_op {
    int nsph;
    double *pos, *rad, *quat;
    int *order;
    double cId, cSL, cTrac, cLOP;
    long npts;
    long *off;
    double *coords, *norms, *w;
    /* SoA copies for the far-field loops */
    double *px, *py, *pz;
    int *sph;
    long *near_off;
    int *near_src;
    long nnear;
    long max_coef;
    long *coef_off;
};

/* Operator set-up: SphereSTKSHOperator::setupDOF / setupFMM / setupNearEval
 * (BIE/SphereSTKSHOperator.cpp:37-144, :539-573). */
orc_op *orc_create(int nsph, const double *pos, const double *radius, const double *quat, const int *order, double cId,
                   double cSL, double cTrac, double cLOP) {
    orc_op *op = (orc_op *)calloc(1, sizeof(orc_op));
    op->nsph = nsph;
    op->cId = cId;
    op->cSL = cSL;
    op->cTrac = cTrac;
    op->cLOP = cLOP;
    op->pos = (double *)malloc(sizeof(double) * 3 * nsph);
    op->rad = (double *)malloc(sizeof(double) * nsph);
    op->quat = (double *)malloc(sizeof(double) * 4 * nsph);
    op->order = (int *)malloc(sizeof(int) * nsph);
    memcpy(op->pos, pos, sizeof(double) * 3 * nsph);
    memcpy(op->rad, radius, sizeof(double) * nsph);
    memcpy(op->order, order, sizeof(int) * nsph);
    for (int i = 0; i < nsph; i++) {
        const double *q = quat + 4 * i;
        const double nq = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (int c = 0; c < 4; c++) op->quat[4 * i + c] = q[c] / nq;
    }
    op->off = (long *)malloc(sizeof(long) * (nsph + 1));
    op->coef_off = (long *)malloc(sizeof(long) * (nsph + 1));
    op->off[0] = 0;
    op->coef_off[0] = 0;
    for (int i = 0; i < nsph; i++) {
        const int p = order[i];
        op->off[i + 1] = op->off[i] + (long)(p + 1) * (2 * p + 2);
        op->coef_off[i + 1] = op->coef_off[i] + 3L * (p + 1) * (p + 2);
        get_table(p);
    }
    const long n = op->npts = op->off[nsph];
    op->coords = (double *)malloc(sizeof(double) * 3 * n);
    op->norms = (double *)malloc(sizeof(double) * 3 * n);
    op->w = (double *)malloc(sizeof(double) * n);
    op->px = (double *)malloc(sizeof(double) * n);
    op->py = (double *)malloc(sizeof(double) * n);
    op->pz = (double *)malloc(sizeof(double) * n);
    op->sph = (int *)malloc(sizeof(int) * n);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < nsph; i++) {
        sh_table *t = get_table(op->order[i]);
        const double R = op->rad[i];
        for (int q = 0; q < t->npts; q++) {
            const long g = op->off[i] + q;
            double nv[3];
            qrot(op->quat + 4 * i, t->eta + 3 * q, nv, 0);
            for (int c = 0; c < 3; c++) {
                op->norms[3 * g + c] = nv[c];
                op->coords[3 * g + c] = op->pos[3 * i + c] + R * nv[c];
            }
            op->w[g] = R * R * t->w[q];
            op->px[g] = op->coords[3 * g];
            op->py[g] = op->coords[3 * g + 1];
            op->pz[g] = op->coords[3 * g + 2];
            op->sph[g] = i;
        }
    }
    /* near list: 0 < |c_s-c_t|^2 < (1.5 R_s + 1.5 R_t)^2, grouped by target, sorted */
    long *cnt = (long *)calloc(nsph + 1, sizeof(long));
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < nsph; t++) {
        long c = 0;
        for (int s = 0; s < nsph; s++) {
            if (s == t) continue;
            const double dx = pos[3 * s] - pos[3 * t], dy = pos[3 * s + 1] - pos[3 * t + 1],
                         dz = pos[3 * s + 2] - pos[3 * t + 2];
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double rr = 1.5 * radius[s] + 1.5 * radius[t];
            if (d2 > 0 && d2 < rr * rr) c++;
        }
        cnt[t + 1] = c;
    }
    op->near_off = (long *)malloc(sizeof(long) * (nsph + 1));
    op->near_off[0] = 0;
    for (int t = 0; t < nsph; t++) op->near_off[t + 1] = op->near_off[t] + cnt[t + 1];
    op->nnear = op->near_off[nsph];
    op->near_src = (int *)malloc(sizeof(int) * (op->nnear + 1));
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < nsph; t++) {
        long c = op->near_off[t];
        for (int s = 0; s < nsph; s++) {
            if (s == t) continue;
            const double dx = pos[3 * s] - pos[3 * t], dy = pos[3 * s + 1] - pos[3 * t + 1],
                         dz = pos[3 * s + 2] - pos[3 * t + 2];
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double rr = 1.5 * radius[s] + 1.5 * radius[t];
            if (d2 > 0 && d2 < rr * rr) op->near_src[c++] = s;
        }
    }
    free(cnt);
    return op;
}

void orc_destroy(orc_op *op) {
    if (!op) return;
    free(op->pos);
    free(op->rad);
    free(op->quat);
    free(op->order);
    free(op->off);
    free(op->coef_off);
    free(op->coords);
    free(op->norms);
    free(op->w);
    free(op->px);
    free(op->py);
    free(op->pz);
    free(op->sph);
    free(op->near_off);
    free(op->near_src);
    free(op);
}

long orc_npts(const orc_op *op) { return op->npts; }

void orc_get_grid(const orc_op *op, double *coords, double *norms, double *weights, long *offsets) {
    if (coords) memcpy(coords, op->coords, sizeof(double) * 3 * op->npts);
    if (norms) memcpy(norms, op->norms, sizeof(double) * 3 * op->npts);
    if (weights) memcpy(weights, op->w, sizeof(double) * op->npts);
    if (offsets) memcpy(offsets, op->off, sizeof(long) * (op->nsph + 1));
}

long orc_num_near_pairs(const orc_op *op) { return op->nnear; }

void orc_get_near_pairs(const orc_op *op, int *trg, int *src) {
    for (int t = 0; t < op->nsph; t++)
        for (long k = op->near_off[t]; k < op->near_off[t + 1]; k++) {
            trg[k] = t;
            src[k] = op->near_src[k];
        }
}

static int is_near(const orc_op *op, int t, int s) {
    for (long k = op->near_off[t]; k < op->near_off[t + 1]; k++)
        if (op->near_src[k] == s) return 1;
    return 0;
}

/* Body-frame coefficients of every sphere (cacheSHCoeff, SphereSTKSHOperator.cpp:523-537). */
static void all_coefficients(const orc_op *op, const double *in, double *coef) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < op->nsph; i++) {
        const int p = op->order[i];
        const long n = op->off[i + 1] - op->off[i];
        double *body = (double *)malloc(sizeof(double) * 3 * n);
        for (long q = 0; q < n; q++) qrot(op->quat + 4 * i, in + 3 * (op->off[i] + q), body + 3 * q, 1);
        orc_sh_analysis(p, body, coef + op->coef_off[i]);
        free(body);
    }
}

/* Far-field direct sums: the exact limit of the STKFMM evaluation
 * (SphereSTKSHOperator.cpp:287-303 and :365-446) restricted to pairs that are
 * neither on the same sphere nor near (the reference subtracts those afterwards,
 * :332-348, :408-433, :641-690).
 *   SL:       u_t += cSL/(8 pi) sum [f/r + r (r.f)/r^3]
 *   traction: t_t += -cTrac 3/(4 pi) sum r (r.f)(r.n_t)/r^5,  r = x_t - y_s, f = w sigma */
void orc_far_sums(const orc_op *op, const double *in, double *out, double cSL, double cTrac) {
    orc_far_sums_range(op, in, out, cSL, cTrac, 0, op->npts);
}

/* Far sums for the target points [t0, t1) only (all sources). */
void orc_far_sums_range(const orc_op *op, const double *in, double *out, double cSL, double cTrac, long t0, long t1) {
    const long n = op->npts;
    double *fx = (double *)malloc(sizeof(double) * n), *fy = (double *)malloc(sizeof(double) * n),
           *fz = (double *)malloc(sizeof(double) * n);
    for (long i = 0; i < n; i++) {
        fx[i] = in[3 * i] * op->w[i];
        fy[i] = in[3 * i + 1] * op->w[i];
        fz[i] = in[3 * i + 2] * op->w[i];
    }
    const int doSL = fabs(cSL) > 1e-9, doT = fabs(cTrac) > 1e-9;
    const double cs = cSL / (8 * PI), ct = -cTrac * 3 / (4 * PI);
    const double *px = op->px, *py = op->py, *pz = op->pz;
#pragma omp parallel for schedule(dynamic, 64)
    for (long it = t0; it < t1; it++) {
        const int T = op->sph[it];
        const double x = px[it], y = py[it], z = pz[it];
        const double nx = op->norms[3 * it], ny = op->norms[3 * it + 1], nz = op->norms[3 * it + 2];
        double su0 = 0, su1 = 0, su2 = 0, st0 = 0, st1 = 0, st2 = 0;
        for (int S = 0; S < op->nsph; S++) {
            if (S == T || is_near(op, T, S)) continue;
            const long b = op->off[S], e = op->off[S + 1];
            if (doT) {
                double a0 = 0, a1 = 0, a2 = 0;
#pragma omp simd reduction(+ : a0, a1, a2)
                for (long js = b; js < e; js++) {
                    const double rx = x - px[js], ry = y - py[js], rz = z - pz[js];
                    const double r2 = rx * rx + ry * ry + rz * rz;
                    const double ri = 1.0 / sqrt(r2);
                    const double ri2 = ri * ri;
                    const double s = (rx * fx[js] + ry * fy[js] + rz * fz[js]) * (rx * nx + ry * ny + rz * nz) *
                                     ri2 * ri2 * ri;
                    a0 += s * rx;
                    a1 += s * ry;
                    a2 += s * rz;
                }
                st0 += a0;
                st1 += a1;
                st2 += a2;
            }
            if (doSL) {
                double a0 = 0, a1 = 0, a2 = 0;
#pragma omp simd reduction(+ : a0, a1, a2)
                for (long js = b; js < e; js++) {
                    const double rx = x - px[js], ry = y - py[js], rz = z - pz[js];
                    const double r2 = rx * rx + ry * ry + rz * rz;
                    const double ri = 1.0 / sqrt(r2);
                    const double s = (rx * fx[js] + ry * fy[js] + rz * fz[js]) * ri * ri * ri;
                    a0 += fx[js] * ri + s * rx;
                    a1 += fy[js] * ri + s * ry;
                    a2 += fz[js] * ri + s * rz;
                }
                su0 += a0;
                su1 += a1;
                su2 += a2;
            }
        }
        out[3 * it + 0] += cs * su0 + ct * st0;
        out[3 * it + 1] += cs * su1 + ct * st1;
        out[3 * it + 2] += cs * su2 + ct * st2;
    }
    free(fx);
    free(fy);
    free(fz);
}

/* applyP2POP (SphereSTKSHOperator.cpp:262-485): out += cId in + cSL S[in] + cTrac K'[in]
 * where the self terms are cSL R S_self(P in) + cTrac (K_ext+K_int)/2 (P in)
 * (:340-348, :417-419, :452) and the near terms are the exterior SH evaluations
 * of the near sources' projected densities (:575-695). */
void orc_apply_p2p(const orc_op *op, const double *in, double *out, double cId, double cSL, double cTrac) {
    orc_apply_p2p_spheres(op, in, out, cId, cSL, cTrac, 0, op->nsph);
}

/* applyP2POP on the points of target spheres [T0, T1) only: the same terms as
 * orc_apply_p2p (all spheres act as sources), written into those spheres'
 * entries of out; the other entries are left untouched. */
void orc_apply_p2p_spheres(const orc_op *op, const double *in, double *out, double cId, double cSL, double cTrac,
                           int T0, int T1) {
    const int doSL = fabs(cSL) > 1e-9, doT = fabs(cTrac) > 1e-9;
    if (T0 < 0) T0 = 0;
    if (T1 > op->nsph) T1 = op->nsph;
    if (T1 <= T0) return;
    if (doSL || doT) {
        double *coef = (double *)malloc(sizeof(double) * op->coef_off[op->nsph]);
        all_coefficients(op, in, coef);
        orc_far_sums_range(op, in, out, cSL, cTrac, op->off[T0], op->off[T1]);
        /* self */
#pragma omp parallel for schedule(dynamic, 4)
        for (int i = T0; i < T1; i++) {
            const int p = op->order[i];
            const long n = op->off[i + 1] - op->off[i];
            const long nc = 3L * (p + 1) * (p + 2);
            double *g = (double *)malloc(sizeof(double) * nc);
            double *v = (double *)malloc(sizeof(double) * 3 * n);
            double *acc = (double *)calloc(3 * n, sizeof(double));
            if (doSL) {
                orc_sh_self_spectral(p, 0, coef + op->coef_off[i], g);
                orc_sh_synthesis(p, g, v);
                for (long k = 0; k < 3 * n; k++) acc[k] += cSL * op->rad[i] * v[k];
            }
            if (doT) {
                orc_sh_self_spectral(p, 3, coef + op->coef_off[i], g);
                orc_sh_synthesis(p, g, v);
                for (long k = 0; k < 3 * n; k++) acc[k] += cTrac * v[k];
            }
            for (long q = 0; q < n; q++) {
                double o[3];
                qrot(op->quat + 4 * i, acc + 3 * q, o, 0);
                for (int c = 0; c < 3; c++) out[3 * (op->off[i] + q) + c] += o[c];
            }
            free(g);
            free(v);
            free(acc);
        }
        /* near */
#pragma omp parallel for schedule(dynamic, 1)
        for (int T = T0; T < T1; T++) {
            const long n = op->off[T + 1] - op->off[T];
            double *xi = (double *)malloc(sizeof(double) * 3 * n), *nu = (double *)malloc(sizeof(double) * 3 * n);
            double *u = (double *)malloc(sizeof(double) * 3 * n), *t = (double *)malloc(sizeof(double) * 3 * n);
            for (long k = op->near_off[T]; k < op->near_off[T + 1]; k++) {
                const int S = op->near_src[k];
                const double RS = op->rad[S];
                for (long q = 0; q < n; q++) {
                    const long g = op->off[T] + q;
                    double d[3];
                    for (int c = 0; c < 3; c++) d[c] = (op->coords[3 * g + c] - op->pos[3 * S + c]) / RS;
                    qrot(op->quat + 4 * S, d, xi + 3 * q, 1);
                    qrot(op->quat + 4 * S, op->norms + 3 * g, nu + 3 * q, 1);
                }
                orc_sh_eval_exterior(op->order[S], coef + op->coef_off[S], (int)n, xi, nu, doSL ? u : NULL,
                                     doT ? t : NULL);
                for (long q = 0; q < n; q++) {
                    double o[3];
                    const long g = op->off[T] + q;
                    if (doSL) {
                        qrot(op->quat + 4 * S, u + 3 * q, o, 0);
                        for (int c = 0; c < 3; c++) out[3 * g + c] += cSL * RS * o[c];
                    }
                    if (doT) {
                        qrot(op->quat + 4 * S, t + 3 * q, o, 0);
                        for (int c = 0; c < 3; c++) out[3 * g + c] += cTrac * o[c];
                    }
                }
            }
            free(xi);
            free(nu);
            free(u);
            free(t);
        }
        free(coef);
    }
    if (fabs(cId) > 1e-9) {
        for (long k = 3 * op->off[T0]; k < 3 * op->off[T1]; k++) out[k] += cId * in[k];
    }
}

/* applyLOP (SphereSTKSHOperator.cpp:487-521). */
void orc_apply_lop(const orc_op *op, const double *in, double *out, double cLOP) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < op->nsph; i++) {
        const double R = op->rad[i];
        double A[3] = {0, 0, 0}, B[3] = {0, 0, 0};
        for (long g = op->off[i]; g < op->off[i + 1]; g++) {
            const double *v = in + 3 * g, *nv = op->norms + 3 * g;
            const double w = op->w[g];
            for (int c = 0; c < 3; c++) A[c] += w * v[c];
            B[0] += w * R * (nv[1] * v[2] - nv[2] * v[1]);
            B[1] += w * R * (nv[2] * v[0] - nv[0] * v[2]);
            B[2] += w * R * (nv[0] * v[1] - nv[1] * v[0]);
        }
        for (int c = 0; c < 3; c++) {
            A[c] *= 1.0 / (4 * PI * R * R);
            B[c] *= 3.0 / (8 * PI * R * R * R * R);
        }
        for (long g = op->off[i]; g < op->off[i + 1]; g++) {
            const double rn[3] = {R * op->norms[3 * g], R * op->norms[3 * g + 1], R * op->norms[3 * g + 2]};
            out[3 * g + 0] += cLOP * (A[0] + B[1] * rn[2] - B[2] * rn[1]);
            out[3 * g + 1] += cLOP * (A[1] + B[2] * rn[0] - B[0] * rn[2]);
            out[3 * g + 2] += cLOP * (A[2] + B[0] * rn[1] - B[1] * rn[0]);
        }
    }
}

/* apply (SphereSTKSHOperator.cpp:153-236), one column. */
void orc_apply(const orc_op *op, const double *X, double *Y, double alpha, double beta) {
    const long n3 = 3 * op->npts;
    double *tmp = (double *)calloc(n3, sizeof(double));
    orc_apply_p2p(op, X, tmp, op->cId, op->cSL, op->cTrac);
    orc_apply_lop(op, X, tmp, op->cLOP);
    for (long k = 0; k < n3; k++) Y[k] = (beta == 0.0 ? 0.0 : beta * Y[k]) + alpha * tmp[k];
    free(tmp);
}

/* projectNullSpace (SphereSTKSHOperator.cpp:239-260). */
void orc_project(const orc_op *op, double *inout) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < op->nsph; i++) {
        const int p = op->order[i];
        const long n = op->off[i + 1] - op->off[i];
        double *body = (double *)malloc(sizeof(double) * 3 * n);
        double *coef = (double *)malloc(sizeof(double) * 3 * (p + 1) * (p + 2));
        for (long q = 0; q < n; q++) qrot(op->quat + 4 * i, inout + 3 * (op->off[i] + q), body + 3 * q, 1);
        orc_sh_analysis(p, body, coef);
        orc_sh_synthesis(p, coef, body);
        for (long q = 0; q < n; q++) qrot(op->quat + 4 * i, body + 3 * q, inout + 3 * (op->off[i] + q), 0);
        free(body);
        free(coef);
    }
}

/* ------------------------------------------------------------------------- */
static double dot(const double *a, const double *b, long n) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) if (n > 100000)
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* Restarted GMRES (Belos "GMRES" with Num Blocks = restart, Orthogonalization
 * IMGS, residuals scaled by ||b||: mobilitySolver.yaml:3-16).  Convergence is
 * declared when the implicit (Givens) residual and, if requested, the explicit
 * residual ||b - A x||/||b|| are both <= tol. */
int orc_gmres(const orc_op *A, const double *b, double *x, int x_is_zero, orc_gmres_params prm, orc_solve_info *info) {
    const long n = 3 * A->npts;
    const int m = prm.restart > 0 ? prm.restart : 30;
    memset(info, 0, sizeof(*info));
    double *V = (double *)malloc(sizeof(double) * n * (m + 1));
    double *r = (double *)malloc(sizeof(double) * n);
    double *H = (double *)calloc((m + 1) * m, sizeof(double));
    double *cs = (double *)malloc(sizeof(double) * m), *sn = (double *)malloc(sizeof(double) * m);
    double *g = (double *)malloc(sizeof(double) * (m + 1)), *y = (double *)malloc(sizeof(double) * m);
    const double bnorm = sqrt(dot(b, b, n));
    if (bnorm == 0.0) {
        memset(x, 0, sizeof(double) * n);
        info->converged = 1;
        free(V), free(r), free(H), free(cs), free(sn), free(g), free(y);
        return 0;
    }
    if (x_is_zero) {
        memset(x, 0, sizeof(double) * n);
        memcpy(r, b, sizeof(double) * n);
    } else {
        orc_apply(A, x, r, -1.0, 0.0);
        info->applies++;
        for (long k = 0; k < n; k++) r[k] += b[k];
    }
    double beta = sqrt(dot(r, r, n));
    info->relres_implicit = info->relres_explicit = beta / bnorm;
    if (beta / bnorm <= prm.tol) info->converged = 1;
    while (!info->converged && info->iters < prm.max_iters) {
        for (long k = 0; k < n; k++) V[k] = r[k] / beta;
        memset(H, 0, sizeof(double) * (m + 1) * m);
        for (int k = 0; k <= m; k++) g[k] = 0;
        g[0] = beta;
        int j, kdim = 0;
        for (j = 0; j < m && info->iters < prm.max_iters; j++) {
            double *w = V + (long)(j + 1) * n;
            orc_apply(A, V + (long)j * n, w, 1.0, 0.0);
            info->applies++;
            for (int pass = 0; pass < 2; pass++)
                for (int i = 0; i <= j; i++) {
                    const double h = dot(w, V + (long)i * n, n);
                    H[i * m + j] += h;
                    for (long k = 0; k < n; k++) w[k] -= h * V[(long)i * n + k];
                }
            const double hn = sqrt(dot(w, w, n));
            H[(j + 1) * m + j] = hn;
            if (hn > 0)
                for (long k = 0; k < n; k++) w[k] /= hn;
            for (int i = 0; i < j; i++) {
                const double t0 = cs[i] * H[i * m + j] + sn[i] * H[(i + 1) * m + j];
                const double t1 = -sn[i] * H[i * m + j] + cs[i] * H[(i + 1) * m + j];
                H[i * m + j] = t0;
                H[(i + 1) * m + j] = t1;
            }
            const double a = H[j * m + j], bb = H[(j + 1) * m + j];
            const double den = sqrt(a * a + bb * bb);
            cs[j] = den > 0 ? a / den : 1.0;
            sn[j] = den > 0 ? bb / den : 0.0;
            H[j * m + j] = den;
            H[(j + 1) * m + j] = 0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            info->iters++;
            kdim = j + 1;
            info->relres_implicit = fabs(g[j + 1]) / bnorm;
            if (info->relres_implicit <= prm.tol || hn == 0.0) break;
        }
        for (int i = kdim - 1; i >= 0; i--) {
            double s = g[i];
            for (int k = i + 1; k < kdim; k++) s -= H[i * m + k] * y[k];
            y[i] = s / H[i * m + i];
        }
        for (int i = 0; i < kdim; i++)
            for (long k = 0; k < n; k++) x[k] += y[i] * V[(long)i * n + k];
        const int imp_ok = info->relres_implicit <= prm.tol;
        if (imp_ok && !prm.explicit_check) {
            info->converged = 1;
            break;
        }
        orc_apply(A, x, r, -1.0, 0.0);
        info->applies++;
        for (long k = 0; k < n; k++) r[k] += b[k];
        beta = sqrt(dot(r, r, n));
        info->relres_explicit = beta / bnorm;
        if (imp_ok && info->relres_explicit <= prm.tol) info->converged = 1;
        if (!info->converged) info->restarts++;
        if (beta == 0.0) {
            info->converged = 1;
            break;
        }
    }
    free(V), free(r), free(H), free(cs), free(sn), free(g), free(y);
    return 0;
}

/* SphereSTKMobMat::solveMob (BIE/SphereSTKMobMat.cpp:146-286). */
int orc_mobility_solve(const orc_op *A, double viscosity, const double *force, double *vel, double *x_inout,
                       orc_gmres_params prm, orc_solve_info *info) {
    const long n3 = 3 * A->npts;
    double *rho = (double *)malloc(sizeof(double) * n3);
    double *b = (double *)calloc(n3, sizeof(double));
    double *x = (double *)calloc(n3, sizeof(double));
    /* rho0 = F/(4 pi R^2) + 3/(8 pi R^4) T x (R n)   (:174-192) */
    for (int i = 0; i < A->nsph; i++) {
        const double R = A->rad[i];
        const double *f = force + 6 * i, *t = force + 6 * i + 3;
        const double a = 1 / (4 * PI * R * R), c = 3 / (8 * PI * R * R * R * R);
        for (long g = A->off[i]; g < A->off[i + 1]; g++) {
            const double rn[3] = {R * A->norms[3 * g], R * A->norms[3 * g + 1], R * A->norms[3 * g + 2]};
            rho[3 * g + 0] = a * f[0] + c * (t[1] * rn[2] - t[2] * rn[1]);
            rho[3 * g + 1] = a * f[1] + c * (t[2] * rn[0] - t[0] * rn[2]);
            rho[3 * g + 2] = a * f[2] + c * (t[0] * rn[1] - t[1] * rn[0]);
        }
    }
    /* b = -(1/2 I + K') rho0   (:196) */
    orc_apply_p2p(A, rho, b, -0.5, 0.0, -1.0);
    const double bnorm = sqrt(dot(b, b, n3));
    memset(info, 0, sizeof(*info));
    if (bnorm < 1e-5) { /* (:216-219) */
        info->shortcut = 1;
        info->converged = 1;
    } else {
        int zero = 1;
        if (x_inout) {
            memcpy(x, x_inout, sizeof(double) * n3);
            for (long k = 0; k < n3; k++)
                if (x[k] != 0.0) {
                    zero = 0;
                    break;
                }
        }
        orc_gmres(A, b, x, zero, prm, info);
    }
    if (x_inout) memcpy(x_inout, x, sizeof(double) * n3);
    /* rho = rho0 + x; u = S[rho]  (:253-259) */
    for (long k = 0; k < n3; k++) {
        rho[k] += x[k];
        b[k] = 0;
    }
    orc_apply_p2p(A, rho, b, 0.0, 1.0, 0.0);
    info->applies += 2;
    /* rigid-body averages (:261-285) */
    const double iv = 1.0 / viscosity;
    for (int i = 0; i < A->nsph; i++) {
        const double R = A->rad[i];
        double V[3] = {0, 0, 0}, W[3] = {0, 0, 0};
        for (long g = A->off[i]; g < A->off[i + 1]; g++) {
            const double *u = b + 3 * g, *nv = A->norms + 3 * g;
            const double w = A->w[g];
            for (int c = 0; c < 3; c++) V[c] += w * u[c];
            W[0] += R * w * (nv[1] * u[2] - nv[2] * u[1]);
            W[1] += R * w * (nv[2] * u[0] - nv[0] * u[2]);
            W[2] += R * w * (nv[0] * u[1] - nv[1] * u[0]);
        }
        for (int c = 0; c < 3; c++) {
            vel[6 * i + c] = V[c] / (4 * PI * R * R) * iv;
            vel[6 * i + 3 + c] = W[c] * 3 / (8 * PI * R * R * R * R) * iv;
        }
    }
    free(rho);
    free(b);
    free(x);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* The steps beside the solve: Brownian velocities and collision pairs.      */

/* Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with
 * multipliers 0xD2511F53 and 0xCD9E8D57, the key bumped by the Weyl constants
 * 0x9E3779B9 and 0xBB67AE85 between rounds. */
void orc_philox4x32_10(const unsigned *ctr, const unsigned *key, unsigned *out) {
    unsigned c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const unsigned long long p0 = (unsigned long long)0xD2511F53u * c0, p1 = (unsigned long long)0xCD9E8D57u * c2;
        const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0, hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* six standard normals of sphere i at `step`: three Philox blocks with counter
 * (i, c, step_lo, step_hi), key (seed_lo, seed_hi); each block gives two
 * 53-bit uniforms in (0,1) and one Box-Muller pair */
static void brown_normals(unsigned long long seed, unsigned long long step, int i, double z[6]) {
    const unsigned key[2] = {(unsigned)seed, (unsigned)(seed >> 32)};
    for (int c = 0; c < 3; c++) {
        const unsigned ctr[4] = {(unsigned)i, (unsigned)c, (unsigned)step, (unsigned)(step >> 32)};
        unsigned r[4];
        orc_philox4x32_10(ctr, key, r);
        const double ua = ((double)(((unsigned long long)r[0] << 21) | (r[1] >> 11)) + 0.5) * 0x1p-53;
        const double ub = ((double)(((unsigned long long)r[2] << 21) | (r[3] >> 11)) + 0.5) * 0x1p-53;
        const double m = sqrt(-2.0 * log(ua)), th = 2.0 * PI * ub;
        z[2 * c] = m * cos(th);
        z[2 * c + 1] = m * sin(th);
    }
}

/* SphereSystem::getVelocityBrown (SRC/SphereSystem.cpp:323-360): per sphere
 * V = sqrt(1/(6 pi R mu)) sqrt(2 kBT/dt) scale N(0,1)^3 and
 * W = sqrt(1/(8 pi R^3 mu)) sqrt(2 kBT/dt) scale N(0,1)^3.  The reference draws
 * from a per-thread TRNG pool; the normals here come from brown_normals. */
void orc_brownian_velocity(int n, const double *radius, double mu, double kBT, double dt, double scale,
                           unsigned long long seed, unsigned long long step, double *vel) {
    const double fack = sqrt(2.0 * kBT / dt) * scale;
    for (int i = 0; i < n; i++) {
        const double R = radius[i];
        const double dx = sqrt(1.0 / (6.0 * PI * R * mu)) * fack, dr = sqrt(1.0 / (8.0 * PI * R * R * R * mu)) * fack;
        double z[6];
        brown_normals(seed, step, i, z);
        for (int k = 0; k < 3; k++) {
            vel[6 * i + k] = dx * z[k];
            vel[6 * i + 3 + k] = dr * z[3 + k];
        }
    }
}

/* Collision pairs (CollisionSphere::collide, Sphere/CollisionSphere.hpp:51-90):
 * sphere a is the I side of the pair with b when gid[a] < gid[b] and the
 * separation sep = |x_b - x_a| - Rc_a - Rc_b is below colbuf * Rc_a; the block
 * holds phi0 = sep and normI = -(x_b - x_a)/|x_b - x_a|.  Brute force over all
 * pairs, in the order (a ascending, then b ascending).  Returns the number of
 * pairs; arrays (may be NULL) are written up to `cap` entries. */
long orc_collision_pairs(int n, const double *pos, const double *rcol, const int *gid, double colbuf, long cap,
                         int *idx_i, int *idx_j, double *phi0, double *norm_i) {
    long np = 0;
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) {
            const int ga = gid ? gid[a] : a, gb = gid ? gid[b] : b;
            if (ga >= gb) continue;
            const double dx = pos[3 * b] - pos[3 * a], dy = pos[3 * b + 1] - pos[3 * a + 1],
                         dz = pos[3 * b + 2] - pos[3 * a + 2];
            const double r = sqrt(fma(dx, dx, fma(dy, dy, dz * dz)));
            const double sep = (r - rcol[a]) - rcol[b];
            if (!(sep < colbuf * rcol[a])) continue;
            if (np < cap) {
                const double s = -1.0 / r;
                if (idx_i) idx_i[np] = a;
                if (idx_j) idx_j[np] = b;
                if (phi0) phi0[np] = sep;
                if (norm_i) {
                    norm_i[3 * np] = dx * s;
                    norm_i[3 * np + 1] = dy * s;
                    norm_i[3 * np + 2] = dz * s;
                }
            }
            np++;
        }
    return np;
}
