/*
 * oracle.h -- FP64 CPU restatement of the many-sphere Stokes boundary-integral
 * mobility operator (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the checker for the HIP product path.  Only tests/, the smoke
 * entry point and the CPU-baseline leg of the benchmark may load it.  The
 * product library never links or calls it.
 *
 * Reference behaviour restated here (files relative to the reference tree):
 *   grid                 Sphere/Shexp.cpp:206-271 (getGridWithPole, poles dropped
 *                        as in BIE/SphereSTKSHOperator.cpp:103-121)
 *   SH projection        Sphere/Shexp.cpp:342-379, :408-447 (calcGridValue /
 *                        calcSpectralCoeff, body frame via q^-1 :437, :374)
 *   self/near SH evals   Sphere/Shexp.cpp:503-624 (calcSDLNF / calcKSelf / calcKNF)
 *   operator             BIE/SphereSTKSHOperator.cpp:153-236 (apply),
 *                        :262-485 (applyP2POP), :487-521 (applyLOP)
 *   near criterion       MPI/NearInteraction.txx:430-441 with radiusNear = 1.5 R
 *                        (BIE/SphereSTKSHOperator.cpp:547)
 *   mobility solve       BIE/SphereSTKMobMat.cpp:146-286, GMRES parameters
 *                        mobilitySolver.yaml:3-16
 *   Brownian velocities  SRC/SphereSystem.cpp:323-360 (normals from Philox4x32-10
 *                        + Box-Muller in place of the reference's TRNG pool)
 *   collision pairs      Sphere/CollisionSphere.hpp:51-90 (brute force)
 *
 * The spherical-harmonic arithmetic the reference takes from SCTL (an un-vendored,
 * unpinned dependency) is re-derived: vector SH of degree <= p in the basis
 * (y r, grad_s y, r x grad_s y) with y_nm = D_n^m restricted to the unit sphere,
 * D_n^m = (d/dx + i d/dy)^m (d/dz)^(n-m) (1/r).  Self operators use the exact
 * per-degree spectra of the single layer and of the traction (adjoint double
 * layer) operator on a sphere, derived from Lamb's general solution; near
 * evaluations use the exterior solid-harmonic expansion of the same flow.
 * The FMM of the reference is replaced by its exact limit, the direct sum.
 */
#ifndef SPHSTK_ORACLE_H
#define SPHSTK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_op orc_op;

/* Gauss-Legendre nodes (ascending) and weights on [-1,1]. */
void orc_gl(int n, double *x, double *w);

/* D_n^m(x) for 0<=m<=n<=P, complex interleaved, index n(n+1)/2+m. */
void orc_solid_harmonics(int P, const double *x, double *D);

/* Unit-sphere body-frame grid of order p (no poles): n=(p+1)(2p+2) points. */
void orc_unit_grid(int p, double *eta /*3n*/, double *w /*n*/);

/* Vector SH analysis / synthesis on the order-p unit grid (body frame).
 * coef: 3*(p+1)(p+2) doubles = F_r, F_s, F_t, each (p+1)(p+2)/2 complex. */
void orc_sh_analysis(int p, const double *vals, double *coef);
void orc_sh_synthesis(int p, const double *coef, double *vals);

/* On-surface self operators of the unit sphere applied to a coefficient set:
 * which=0 single layer, 1 traction exterior limit, 2 traction interior limit,
 * 3 traction principal value. */
void orc_sh_self_spectral(int p, int which, const double *coef_in, double *coef_out);

/* Exterior single-layer velocity u and traction t (normal nu) of the unit-sphere
 * density given by coef, at body-frame points xi (|xi|>=1). u or t may be NULL. */
void orc_sh_eval_exterior(int p, const double *coef, int ntrg, const double *xi, const double *nu, double *u,
                          double *t);

/* Operator construction.  quat is (w,x,y,z) per sphere. */
orc_op *orc_create(int nsph, const double *pos, const double *radius, const double *quat, const int *order, double cId,
                   double cSL, double cTrac, double cLOP);
void orc_destroy(orc_op *op);
long orc_npts(const orc_op *op);
void orc_get_grid(const orc_op *op, double *coords, double *norms, double *weights, long *offsets);
long orc_num_near_pairs(const orc_op *op);
void orc_get_near_pairs(const orc_op *op, int *trg, int *src);

/* out += cId*in + cSL*S[in] + cTrac*K'[in]   (SphereSTKSHOperator.cpp:262-485) */
void orc_apply_p2p(const orc_op *op, const double *in, double *out, double cId, double cSL, double cTrac);
/* The same on the points of target spheres [T0, T1) only (all spheres are sources). */
void orc_apply_p2p_spheres(const orc_op *op, const double *in, double *out, double cId, double cSL, double cTrac,
                           int T0, int T1);
/* out += cLOP * L[in]                        (SphereSTKSHOperator.cpp:487-521) */
void orc_apply_lop(const orc_op *op, const double *in, double *out, double cLOP);
/* Y = beta*Y + alpha*A*X with the stored coefficients (:153-236) */
void orc_apply(const orc_op *op, const double *X, double *Y, double alpha, double beta);
/* in-place projection onto band-limited vector fields (:239-260) */
void orc_project(const orc_op *op, double *inout);

/* Far-field direct pair sums only (no self/near terms): out += sum over far pairs. */
void orc_far_sums(const orc_op *op, const double *in, double *out, double cSL, double cTrac);
void orc_far_sums_range(const orc_op *op, const double *in, double *out, double cSL, double cTrac, long t0, long t1);

typedef struct {
    double tol;
    int restart;
    int max_iters;
    int explicit_check;
} orc_gmres_params;

typedef struct {
    int iters;
    int restarts;
    int converged;
    int shortcut;
    double relres_implicit;
    double relres_explicit;
    int applies;
} orc_solve_info;

/* GMRES(restart) with two-pass modified Gram-Schmidt.  Solves A x = b. */
int orc_gmres(const orc_op *A, const double *b, double *x, int x_is_zero, orc_gmres_params prm, orc_solve_info *info);

/* One mobility solve (SphereSTKMobMat::solveMob, MobMat.cpp:146-286). The operator
 * must have been created with coefficients (0.5, 0, 1, 1).  x_inout is the
 * warm start on entry (NULL = zero) and the solution x on exit. */
int orc_mobility_solve(const orc_op *A, double viscosity, const double *force, double *vel, double *x_inout,
                       orc_gmres_params prm, orc_solve_info *info);

int orc_num_threads(void);

/* Philox4x32-10 block: out = philox(ctr[4], key[2]). */
void orc_philox4x32_10(const unsigned *ctr, const unsigned *key, unsigned *out);
/* Brownian velocities, 6 per sphere (SRC/SphereSystem.cpp:323-360). */
void orc_brownian_velocity(int n, const double *radius, double mu, double kBT, double dt, double scale,
                           unsigned long long seed, unsigned long long step, double *vel);
/* Collision pairs (Sphere/CollisionSphere.hpp:51-90), brute force; returns the count. */
long orc_collision_pairs(int n, const double *pos, const double *rcol, const int *gid, double colbuf, long cap,
                         int *idx_i, int *idx_j, double *phi0, double *norm_i);

#ifdef __cplusplus
}
#endif
#endif
