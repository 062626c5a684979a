/*
 * sphstk.h -- C ABI of the MI355X-native many-sphere Stokes boundary-integral
 * mobility operator (libsphstk.so).
 *
 * Drop-in boundary for the reference's BIE layer (paths relative to the
 * reference tree):
 *   SphereSTKSHOperator  BIE/SphereSTKSHOperator.hpp:112-231
 *   SphereSTKMobMat      BIE/SphereSTKMobMat.hpp:26-79
 * Tpetra vectors are replaced by plain double arrays (host or device); the
 * Sphere vector by flat arrays (pos 3N, radius N, quaternion w,x,y,z 4N,
 * SH order N).  The DOF layout matches the reference: per quadrature point the
 * interleaved (x,y,z) value, spheres in input order, points in
 * Shexp::getGridWithPole order with the poles removed (Sphere/Shexp.cpp:206-271,
 * BIE/SphereSTKSHOperator.cpp:103-121).
 *
 * Conventions
 *  - Every function returns SPHSTK_OK (0) or a negative error code; the message
 *    of the most recent failure on the calling thread is sphstk_last_error().
 *    Errors that the reference raises as TEUCHOS_ASSERT / std::logic_error
 *    (size or mode mismatch, cpp:155-185, MobMat.cpp:100-127) are returned as
 *    SPHSTK_ERR_INVALID; no C++ exception crosses this boundary.
 *  - One host thread per handle.  A handle owns one HIP stream on the device
 *    that was current when it was created; calls on one handle are serialised
 *    on that stream and are not re-entrant (the reference's scratch members are
 *    `mutable`, hpp:123-157).
 *  - Functions whose names end in _device take device pointers that must be
 *    valid on the handle's device; the others take host pointers.
 *  - No function falls back to a CPU path: without a usable HIP device,
 *    creation fails with SPHSTK_ERR_DEVICE.
 */
#ifndef SPHSTK_H
#define SPHSTK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPHSTK_OK 0
#define SPHSTK_ERR_INVALID (-1)
#define SPHSTK_ERR_DEVICE (-2)
#define SPHSTK_ERR_NOMEM (-3)
#define SPHSTK_ERR_INTERNAL (-4)

/* precision modes of the mobility solve */
#define SPHSTK_PREC_FP64 0  /* plain form: every operator apply in FP64 */
#define SPHSTK_PREC_MIXED 1 /* FP32 inner GMRES inside FP64 iterative refinement */

/* creation flags */
#define SPHSTK_FLAG_NONE 0
#define SPHSTK_FLAG_TIMING 1      /* record HIP-event timings of the far-field kernels */
#define SPHSTK_FLAG_FP32_DIRECT 2 /* FP32 inner operator: direct far field (no multi-level coarse grids) */
#define SPHSTK_FLAG_AUTOTUNE 4    /* time candidate far-field tile shapes at set-up and keep the fastest */
#define SPHSTK_FLAG_GRAPH 8       /* replay the FP32 inner operator's apply from a captured HIP graph */
#define SPHSTK_FLAG_INNER_PRECOND 16 /* mixed form: right-precondition the inner solve by the isolated-sphere
                                        inverse (sphstk_operator_apply_self_inverse) */
#define SPHSTK_FLAG_NO_CELLS 32   /* FP32 inner operator: no source-cell aggregation on its two coarsest levels */
#define SPHSTK_FLAG_EXPLICIT_REFINE 64 /* mixed form: an explicit FP64 residual at every refinement step (default:
                                          once twice the inner operator's measured relative error times the
                                          correction's size is below tol / 2, a step takes the inner solve's
                                          residual plus that bound) */
#define SPHSTK_FLAG_NO_COARSE 128 /* mixed form: plain GMRES on the FP32 operator for the inner solves (default:
                                     refinement with a coarse FP32 operator, FP32 residuals) */

typedef struct sphstk_operator sphstk_operator;
typedef struct sphstk_mobility sphstk_mobility;

/* Belos parameters of mobilitySolver.yaml:3-16 that the solve honours. */
typedef struct {
    double tol;         /* Convergence Tolerance, residual scaled by ||b|| (yaml :4, :12-13) */
    int restart;        /* Num Blocks (yaml :7); 30 */
    int max_iters;      /* Maximum Iterations (yaml :6) */
    int explicit_check; /* 1: confirm convergence with ||b - A x|| (Belos explicit residual test) */
    double inner_tol;   /* mixed precision only: inner FP32 GMRES tolerance (3e-5) */
    int max_refine;     /* mixed precision only: maximum outer refinement steps */
} sphstk_solver_params;

typedef struct {
    int iters;              /* Krylov iterations (operator applies inside GMRES) summed over the solve */
    int outer;              /* refinement steps (mixed precision), 1 for FP64 */
    int converged;          /* the reference ignores Belos' status (MobMat.cpp:236); reported here */
    int shortcut;           /* ||b|| < 1e-5 shortcut taken (MobMat.cpp:216-219) */
    double relres;          /* final ||b - A x||/||b|| (explicit when computed, else implicit).  FP64
                               form: A is the FP64 operator with the direct far-field sum.  Mixed form:
                               A is the multi-level FP64 operator of sphstk_operator_apply_outer (about
                               1e-9 from the direct-sum operator), b its right-hand side with distant
                               spheres' rho0 in closed form; after a refinement step without an FP64
                               apply (see SPHSTK_FLAG_EXPLICIT_REFINE) it is the bound
                               (|inner residual| + measured inner-operator error) / |b| */
    int applies_fp64;       /* FP64 operator applies, including the RHS and final single-layer applies */
    int applies_fp32;       /* FP32 operator applies (mixed precision) */
} sphstk_solve_info;

/* Last error message of the calling thread ("" if none). */
const char *sphstk_last_error(void);
/* Library version string. */
const char *sphstk_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int sphstk_device_count(void);

/* ---------------------------------------------------------------------------
 * Operator  Y = beta Y + alpha (cId I + cSL S + cTrac K' + cLOP L) X
 * Replaces SphereSTKSHOperator(const std::vector<Sphere>*, layerName,
 *   shared_ptr<STKFMM>&, cId, cSL, cTrac, cLOP)  (SphereSTKSHOperator.hpp:182-184).
 * Inputs are copied; the grid, near-pair list and tables are built on device. */
int sphstk_operator_create(int n_spheres, const double *pos, const double *radius, const double *quat_wxyz,
                           const int *order, double cId, double cSL, double cTrac, double cLOP, int flags,
                           sphstk_operator **out);
/* Batched form: n_susp independent suspensions, sphere_counts[b] spheres each,
 * concatenated in the arrays.  Sphere pairs in different suspensions do not
 * interact. */
int sphstk_operator_create_batched(int n_susp, const int *sphere_counts, const double *pos, const double *radius,
                                   const double *quat_wxyz, const int *order, double cId, double cSL, double cTrac,
                                   double cLOP, int flags, sphstk_operator **out);
int sphstk_operator_destroy(sphstk_operator *op);

/* Sizes: spheres, quadrature points, DOF (= 3 * points), near pairs (ordered). */
int sphstk_operator_sizes(const sphstk_operator *op, int64_t *n_spheres, int64_t *n_points, int64_t *n_dof,
                          int64_t *n_near_pairs);

/* apply: Y := beta Y + alpha A X for ncol columns of length n_dof, column-major
 * (SphereSTKSHOperator::apply, cpp:153-236; hpp:195-198). */
int sphstk_operator_apply(sphstk_operator *op, const double *X, double *Y, int ncol, double alpha, double beta);
int sphstk_operator_apply_device(sphstk_operator *op, const double *dX, double *dY, int ncol, double alpha,
                                 double beta);

/* The FP32 inner operator of the mixed-precision solve, Y := beta Y + alpha A32 X
 * (one column).  No counterpart in the reference: A32 has the far field in FP32
 * and, unless SPHSTK_FLAG_FP32_DIRECT was given at creation, multi-level (sphere
 * pairs farther apart than 3 (R_s + R_t) interact through coarser grids of order
 * p/2+1, max(5, p/3+1), 4 and 3 from 3, 6, 9 and 16 (R_s + R_t) on; orders
 * below 10 keep their own grid at the first level); self and near terms stay FP64;
 * relative error about 4e-6 against A64. */
int sphstk_operator_apply_inner(sphstk_operator *op, const double *X, double *Y, double alpha, double beta);

/* The FP64 outer operator of the mixed-precision solve (its right-hand side and
 * refinement residuals), Y := beta Y + alpha A64ml X (one column).  No
 * counterpart in the reference (whose far field is an FMM to a set tolerance):
 * FP64 pair sums with sphere pairs farther apart than 7 (R_s + R_t)
 * interacting through coarser grids of order min(p, 8), min(p, 6), min(p, 5)
 * and min(p, 4) from 7, 14, 28 and 56 (R_s + R_t) on; self and near terms as
 * A64; relative error about 1e-9 against A64. */
int sphstk_operator_apply_outer(sphstk_operator *op, const double *X, double *Y, double alpha, double beta);

/* The coarse FP32 operator of the mixed-precision solve's inner solves (their
 * refinement steps run GMRES on it), Y := beta Y + alpha Ac X (one column).  No
 * counterpart in the reference: as the FP32 inner operator, with coarse grids of
 * order 5, 4, 3 and 2 (capped at p) from 2, 3, 5 and 8 (R_s + R_t) on
 * (SPHSTK_COARSE_ALPHA); relative error about 1e-2 against A64; unless
 * SPHSTK_FLAG_FP32_DIRECT was given at creation, in which case it is A32. */
int sphstk_operator_apply_coarse(sphstk_operator *op, const double *X, double *Y, double alpha, double beta);

/* applyP2POP: out += cId in + cSL S[in] + cTrac K'[in]  (cpp:262-485; hpp:204-206). */
int sphstk_operator_apply_p2p(sphstk_operator *op, const double *in, double *out, double cId, double cSL,
                              double cTrac);
int sphstk_operator_apply_p2p_device(sphstk_operator *op, const double *din, double *dout, double cId, double cSL,
                                     double cTrac);

/* applyLOP: out += cLOP L[in]  (cpp:487-521; hpp:208-210). */
int sphstk_operator_apply_lop(sphstk_operator *op, const double *in, double *out, double cLOP);

/* Values of the spherical-harmonic expansion of the grid field `in` (ndof) at
 * the north and south poles of every sphere's SH frame, lab frame: poles[6 S +
 * 0..2] north, [6 S + 3..5] south (Shexp::calcPoleValue, Shexp.cpp:382-405,
 * which feeds the grid-with-poles output of Shexp::writeVTU, :96-204). */
int sphstk_operator_pole_values(sphstk_operator *op, const double *in, double *poles);

/* out = M^-1 in, M the block-diagonal part of the mobility operator
 * 1/2 I + K' + L (each sphere's self operator, the far and near terms
 * dropped), inverted exactly per sphere and SH degree.  The operator must have
 * the coefficients (0.5, any, 1, 1).  The mixed form's inner solve uses it as a
 * right preconditioner.  No counterpart in the reference (Belos runs
 * unpreconditioned there). */
int sphstk_operator_apply_self_inverse(sphstk_operator *op, const double *in, double *out);
/* projectNullSpace: in-place projection onto vector SH of degree <= p
 * (cpp:239-260; hpp:200-202). */
int sphstk_operator_project_nullspace(sphstk_operator *op, double *inout);

/* Geometry getters (getGridWeights / getGridCoords / getGridNorms /
 * getGridNumberIndex / getGridNumberLength, hpp:218-223).  Any pointer may be
 * NULL.  index/length are per sphere (n_spheres entries). */
int sphstk_operator_get_grid(const sphstk_operator *op, double *weights, double *coords, double *norms,
                             int64_t *index, int64_t *length);

/* Near pairs (target, source) in the order used internally (grouped by target). */
int sphstk_operator_get_near_pairs(const sphstk_operator *op, int32_t *trg, int32_t *src);

/* Far-field work-list shapes in use (chosen by SPHSTK_FLAG_AUTOTUNE, else the
 * defaults): targets per lane and source slices of the FP64 direct lists, and
 * targets per lane of the FP32 inner operator's lists.  No counterpart in the
 * reference. */
int sphstk_operator_far_shape(const sphstk_operator *op, int *tpl, int *split, int *inner_tpl);

/* Far-field kernel timing collected with SPHSTK_FLAG_TIMING: total milliseconds,
 * launches, and the algorithmic pair count per launch, for the traction (kind 0)
 * or single-layer (kind 1) pair sums of the FP64 operator (prec 0), the FP32
 * inner operator (prec 1), the FP64 multi-level outer operator (prec 2) or the
 * coarse FP32 operator of the mixed form's inner solves (prec 3). */
int sphstk_operator_kernel_stats(const sphstk_operator *op, int kind, int prec, double *total_ms, int64_t *launches,
                                 double *pairs_per_launch);
int sphstk_operator_reset_stats(sphstk_operator *op);

/* ---------------------------------------------------------------------------
 * Mobility  force/torque (6N) -> velocity/angular velocity (6N)
 * Replaces SphereSTKMobMat(std::vector<Sphere>*, name, shared_ptr<STKFMM>&,
 *   viscosity) (SphereSTKMobMat.hpp:57-58); the operator it builds is
 * A = 0.5 I + K' + L (MobMat.cpp:17). */
int sphstk_mobility_create(int n_spheres, const double *pos, const double *radius, const double *quat_wxyz,
                           const int *order, double viscosity, const sphstk_solver_params *params, int precision,
                           int flags, sphstk_mobility **out);
int sphstk_mobility_create_batched(int n_susp, const int *sphere_counts, const double *pos, const double *radius,
                                   const double *quat_wxyz, const int *order, double viscosity,
                                   const sphstk_solver_params *params, int precision, int flags,
                                   sphstk_mobility **out);
int sphstk_mobility_destroy(sphstk_mobility *mob);
/* The operator A owned by the mobility handle (do not destroy). */
sphstk_operator *sphstk_mobility_operator(sphstk_mobility *mob);
int sphstk_mobility_set_params(sphstk_mobility *mob, const sphstk_solver_params *params, int precision);

/* solveMob (MobMat.cpp:146-286).  force = [fx,fy,fz,tx,ty,tz] per sphere,
 * vel = [vx,vy,vz,wx,wy,wz].  The GMRES initial guess is the stored density
 * unknown x (xLast, MobMat.cpp:222, :245); info may be NULL.  In batched form
 * every suspension is solved; info aggregates (iters = max over suspensions). */
int sphstk_mobility_solve(sphstk_mobility *mob, const double *force, double *vel, sphstk_solve_info *info);
/* Several right-hand sides on the handle's geometry in one lock-step solve
 * (the reference's apply asserts one column, MobMat.cpp:127): force and vel
 * hold ncol column blocks of 6N.  FP64 form: the traction far field is
 * computed once per point pair for blocks of up to 8 columns (d, r^2, r^-5 and
 * d.n shared).  Mixed form: the columns are solved as a batch of independent
 * copies (no shared pair sums).  Starts from x0 = 0 and
 * leaves the handle's own warm start untouched; the first call for a given
 * ncol builds the copies (set-up).  No counterpart in the reference. */
int sphstk_mobility_solve_multi(sphstk_mobility *mob, int ncol, const double *force, double *vel,
                                sphstk_solve_info *info);
/* apply: Y := beta Y + alpha M X, one mobility solve per column (MobMat.cpp:98-144). */
int sphstk_mobility_apply(sphstk_mobility *mob, const double *X, double *Y, int ncol, double alpha, double beta);

/* Density access: getDensitySolution (rho = rho0 + x of the last solve, hpp:76),
 * writeBackDensitySolution semantics (hpp:78) and the warm start x (xLast). */
int sphstk_mobility_get_density(const sphstk_mobility *mob, double *rho);
int sphstk_mobility_get_unknown(const sphstk_mobility *mob, double *x);
int sphstk_mobility_set_unknown(sphstk_mobility *mob, const double *x);
/* The constructor's warm start from the density stored on the spheres' layers
 * (MobMat.cpp:67-93): x = projectNullSpace(rho).  The reference subtracts its
 * member rho there, which is still zero, so rho0 stays in the guess; pass the
 * density a previous handle returned by get_density (the values
 * writeBackDensitySolution stores, :288-307) to reproduce its next step.
 * set_unknown with the previous x is the closer guess. */
int sphstk_mobility_set_density(sphstk_mobility *mob, const double *rho);

/* ---- The steps beside the solve (host arrays; the work runs on the GPU) ---- */

/* Brownian velocities, replacing SphereSystem::getVelocityBrown
 * (SRC/SphereSystem.cpp:323-360): per sphere i of radius R,
 *   vel[6i..6i+2] = sqrt(1/(6 pi R mu)) sqrt(2 kBT/dt) scale z_0..2,
 *   vel[6i+3..6i+5] = sqrt(1/(8 pi R^3 mu)) sqrt(2 kBT/dt) scale z_3..5,
 * with standard normals z from Philox4x32-10 blocks (counter (i, c, step_lo,
 * step_hi), c = 0..2, key (seed_lo, seed_hi)) through Box-Muller: the same
 * (seed, step) gives the same numbers on any device and launch.  The reference
 * draws from a per-thread TRNG pool instead, so its streams are not reproduced. */
int sphstk_brownian_velocity(int n, const double *radius, double viscosity, double kBT, double dt, double scale,
                             uint64_t seed, uint64_t step, double *vel);

/* Collision pairs, replacing the near search of SphereSystem::resolveCollision
 * (SRC/SphereSystem.cpp:414-431) and CollisionSphere::collide
 * (Sphere/CollisionSphere.hpp:51-90): every pair (I, J) with gid[I] < gid[J]
 * (gid NULL: the index) and sep = |x_J - x_I| - Rc_I - Rc_J < colbuf * Rc_I,
 * ordered by I, then J.  Per pair: idx_i, idx_j (sphere indices), phi0 = sep,
 * norm_i = -(x_J - x_I)/|x_J - x_I| (3 doubles; normJ = -normI, the initial
 * gamma max(-sep, 0) and the stress |r| normI normI^T follow from these).
 * *npairs receives the count; the arrays (any may be NULL) are written only
 * when it is <= capacity, so capacity 0 asks for the count alone.  colbuf is
 * the collision buffer of SimToolbox's GeoCommon.h, an argument here. */
int sphstk_collision_pairs(int n, const double *pos, const double *radius_col, const int *gid, double colbuf,
                           int64_t capacity, int64_t *npairs, int *idx_i, int *idx_j, double *phi0, double *norm_i);

/* sphstk_collision_pairs on DEVICE arrays, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device memory
 * (code that a HIP graph capture surrounds, tight serving loops, harnesses
 * that forbid hipMalloc).  pos, radius_col, gid (may be NULL: the index) and
 * the outputs are device pointers on the current device; the four outputs are
 * all given or all NULL (count only).  Every intermediate array (cell keys,
 * sort buffers, cell bounds, counts, offsets, the sort workspace) lives in one
 * static device array of about 40 MB per device, which holds searches of at
 * most SPHSTK_NOALLOC_MAX_SPHERES spheres; larger n is rejected with
 * SPHSTK_ERR_INVALID.  The work runs on `stream` (a hipStream_t; NULL: the
 * default stream) and the call returns when it is done (*npairs is a host
 * value).  Calls that overlap on one device share that array and race. */
#define SPHSTK_NOALLOC_MAX_SPHERES 262144
int sphstk_collision_pairs_noalloc(int n, const double *pos, const double *radius_col, const int *gid, double colbuf,
                                   int64_t capacity, int64_t *npairs, int *idx_i, int *idx_j, double *phi0,
                                   double *norm_i, void *stream);

/* The collision LCP around the mobility solve, replacing SimToolbox's
 * CollisionSolver as called by SphereSystem::resolveCollision
 * (SRC/SphereSystem.cpp:445-474; the solver is not in the reference tree):
 *   gamma >= 0, phi = phi0 + buffer + dt D^T (vel_known + M D gamma) >= 0, gamma . phi = 0,
 * with D gamma adding gamma norm_i to sphere idx_i's force and -gamma norm_i to
 * idx_j's (contact on the line of centres, no torque), solved by
 * Barzilai-Borwein projected gradient descent to max_k |min(gamma_k, phi_k)| <= tol
 * (runConfig "LCP tol", 2000 iterations in the reference).  M is one full
 * mobility solve of `mob` per iteration, warm-started from the previous one
 * (SphereSTKMobMat::apply inside the LCP), or, with mob NULL, the free-drain
 * diagonal 1/(6 pi mu R), 1/(8 pi mu R^3) of SphereSystem::getMobOperator
 * without hydrodynamics (SRC/SphereSystem.cpp:237-268; radius and viscosity
 * are read only then).  gamma: initial guess in (NULL entries not allowed;
 * negative values are clipped), solution out.  force_col = D gamma and
 * vel_col = M D gamma (6N each; the reference's getForceCol / getVelocityCol,
 * vel = vel_col + vel_known at SRC/SphereSystem.cpp:472-474). */
typedef struct {
    int iters;           /* BBPGD iterations */
    int mobility_solves; /* applications of M */
    int converged;       /* residual <= tol */
    double residual;     /* max_k |min(gamma_k, phi_k)| at return */
} sphstk_lcp_info;
int sphstk_collision_lcp(sphstk_mobility *mob, int n, const double *radius, double viscosity, int64_t npairs,
                         const int *idx_i, const int *idx_j, const double *phi0, const double *norm_i,
                         const double *vel_known, double dt, double buffer, double tol, int max_iters, double *gamma,
                         double *force_col, double *vel_col, sphstk_lcp_info *info);

#ifdef __cplusplus
}
#endif
#endif
