// extern "C" boundary of libsphstk.so (declared in include/sphstk.h).  Every
// entry point catches library errors and returns a status code; the message is
// kept per calling thread for sphstk_last_error().
#include "../../include/sphstk.h"
#include "engine.hpp"
#include "lcp.hpp"
#include "steps.hpp"

#include <cstring>
#include <new>

using namespace sphstk;

static_assert(SPHSTK_NOALLOC_MAX_SPHERES == kNoAllocMaxSpheres, "header and library disagree on the scratch limit");

struct sphstk_operator {
    std::unique_ptr<Operator> own;
    Operator *op = nullptr;
};

struct sphstk_mobility {
    std::unique_ptr<Mobility> mob;
    sphstk_operator view;
    std::vector<SegInfo> last;
    // the creation inputs, kept for the several-right-hand-side solver
    std::vector<int> counts, order;
    std::vector<double> pos, radius, quat;
    double viscosity = 1.0;
    int flags = 0;
    std::unique_ptr<Mobility> multi; // copies of the suspensions, one per right-hand side
    int multi_ncol = 0, multi_prec = 0;
};

namespace {

thread_local std::string g_err;

// solve info aggregated over the suspensions (iters = max)
void fill_info(sphstk_solve_info *info, const std::vector<SegInfo> &seg, int outer, int a64, int a32) {
    if (!info) return;
    std::memset(info, 0, sizeof(*info));
    info->converged = 1;
    info->shortcut = 1;
    for (const auto &s : seg) {
        info->iters = std::max(info->iters, s.iters);
        info->converged = info->converged && s.converged;
        info->shortcut = info->shortcut && s.shortcut;
        info->relres = std::max(info->relres, s.relres);
    }
    info->outer = outer;
    info->applies_fp64 = a64;
    info->applies_fp32 = a32;
}

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

template <typename F>
int guard(F &&f) {
    try {
        g_err.clear();
        f();
        return SPHSTK_OK;
    } catch (const Error &e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(SPHSTK_ERR_NOMEM, "out of host memory");
    } catch (const std::exception &e) {
        return fail(SPHSTK_ERR_INTERNAL, e.what());
    } catch (...) {
        return fail(SPHSTK_ERR_INTERNAL, "unknown error");
    }
}

SolverParams to_params(const sphstk_solver_params *p) {
    SolverParams s;
    if (p) {
        s.tol = p->tol;
        s.restart = p->restart;
        s.max_iters = p->max_iters;
        s.explicit_check = p->explicit_check;
        s.inner_tol = p->inner_tol;
        s.max_refine = p->max_refine;
    }
    SPHSTK_REQUIRE(s.tol > 0 && s.tol < 1, "tolerance must be in (0, 1)");
    SPHSTK_REQUIRE(s.restart >= 1 && s.restart <= 500, "restart must be in [1, 500]");
    SPHSTK_REQUIRE(s.max_iters >= 0, "max_iters must be >= 0");
    SPHSTK_REQUIRE(s.inner_tol > 0 && s.inner_tol < 1, "inner tolerance must be in (0, 1)");
    SPHSTK_REQUIRE(s.max_refine >= 1, "max_refine must be >= 1");
    return s;
}

// one column on device: Y = beta Y + alpha A X with host or device buffers
void op_apply(Operator &A, const double *X, double *Y, int ncol, double alpha, double beta, bool device) {
    SPHSTK_REQUIRE(ncol >= 1, "ncol must be >= 1");
    SPHSTK_REQUIRE(X && Y, "X and Y must not be NULL");
    const int64_t n = A.ndof();
    if (device) {
        for (int c = 0; c < ncol; c++) A.apply(X + c * n, Y + c * n, alpha, beta, 0);
        SPHSTK_HIP(hipStreamSynchronize(A.stream()));
        return;
    }
    DevBuf<double> dx, dy;
    dx.alloc(std::max<int64_t>(n, 1));
    dy.alloc(std::max<int64_t>(n, 1));
    for (int c = 0; c < ncol; c++) {
        SPHSTK_HIP(hipMemcpyAsync(dx.p, X + c * n, sizeof(double) * n, hipMemcpyHostToDevice, A.stream()));
        if (beta != 0.0)
            SPHSTK_HIP(hipMemcpyAsync(dy.p, Y + c * n, sizeof(double) * n, hipMemcpyHostToDevice, A.stream()));
        A.apply(dx.p, dy.p, alpha, beta, 0);
        SPHSTK_HIP(hipMemcpyAsync(Y + c * n, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost, A.stream()));
    }
    SPHSTK_HIP(hipStreamSynchronize(A.stream()));
}

void op_p2p(Operator &A, const double *in, double *out, double cId, double cSL, double cTrac, bool device) {
    SPHSTK_REQUIRE(in && out, "in and out must not be NULL");
    const int64_t n = A.ndof();
    if (device) {
        A.p2p(in, out, cId, cSL, cTrac, 0, nullptr);
        SPHSTK_HIP(hipStreamSynchronize(A.stream()));
        return;
    }
    DevBuf<double> di, dout;
    di.alloc(std::max<int64_t>(n, 1));
    dout.alloc(std::max<int64_t>(n, 1));
    SPHSTK_HIP(hipMemcpyAsync(di.p, in, sizeof(double) * n, hipMemcpyHostToDevice, A.stream()));
    SPHSTK_HIP(hipMemcpyAsync(dout.p, out, sizeof(double) * n, hipMemcpyHostToDevice, A.stream()));
    A.p2p(di.p, dout.p, cId, cSL, cTrac, 0, nullptr);
    SPHSTK_HIP(hipMemcpyAsync(out, dout.p, sizeof(double) * n, hipMemcpyDeviceToHost, A.stream()));
    SPHSTK_HIP(hipStreamSynchronize(A.stream()));
}

} // namespace

extern "C" {

const char *sphstk_last_error(void) { return g_err.c_str(); }

const char *sphstk_version(void) { return "sphstk 0.1.0 (gfx950)"; }

int sphstk_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int sphstk_operator_create(int n_spheres, const double *pos, const double *radius, const double *quat_wxyz,
                           const int *order, double cId, double cSL, double cTrac, double cLOP, int flags,
                           sphstk_operator **out) {
    return guard([&] {
        SPHSTK_REQUIRE(out, "out must not be NULL");
        SPHSTK_REQUIRE(n_spheres >= 0, "n_spheres must be >= 0");
        auto h = std::make_unique<sphstk_operator>();
        h->own = std::make_unique<Operator>(1, &n_spheres, pos, radius, quat_wxyz, order, cId, cSL, cTrac, cLOP,
                                            flags);
        h->op = h->own.get();
        *out = h.release();
    });
}

int sphstk_operator_create_batched(int n_susp, const int *sphere_counts, const double *pos, const double *radius,
                                   const double *quat_wxyz, const int *order, double cId, double cSL, double cTrac,
                                   double cLOP, int flags, sphstk_operator **out) {
    return guard([&] {
        SPHSTK_REQUIRE(out, "out must not be NULL");
        auto h = std::make_unique<sphstk_operator>();
        h->own = std::make_unique<Operator>(n_susp, sphere_counts, pos, radius, quat_wxyz, order, cId, cSL, cTrac,
                                            cLOP, flags);
        h->op = h->own.get();
        *out = h.release();
    });
}

int sphstk_operator_destroy(sphstk_operator *op) {
    return guard([&] {
        if (op && op->own) delete op; // views owned by a mobility handle are not deleted
    });
}

int sphstk_operator_sizes(const sphstk_operator *op, int64_t *n_spheres, int64_t *n_points, int64_t *n_dof,
                          int64_t *n_near_pairs) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        if (n_spheres) *n_spheres = op->op->nsph();
        if (n_points) *n_points = op->op->npts();
        if (n_dof) *n_dof = op->op->ndof();
        if (n_near_pairs) *n_near_pairs = op->op->nnear();
    });
}

int sphstk_operator_apply(sphstk_operator *op, const double *X, double *Y, int ncol, double alpha, double beta) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        op_apply(*op->op, X, Y, ncol, alpha, beta, false);
    });
}

int sphstk_operator_apply_device(sphstk_operator *op, const double *dX, double *dY, int ncol, double alpha,
                                 double beta) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        op_apply(*op->op, dX, dY, ncol, alpha, beta, true);
    });
}

static int apply_prec(sphstk_operator *op, const double *X, double *Y, double alpha, double beta, int prec) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        SPHSTK_REQUIRE(X && Y, "X and Y must not be NULL");
        Operator &A = *op->op;
        const int64_t n = A.ndof();
        DevBuf<double> dx, dy;
        dx.alloc(std::max<int64_t>(n, 1));
        dy.alloc(std::max<int64_t>(n, 1));
        SPHSTK_HIP(hipMemcpyAsync(dx.p, X, sizeof(double) * n, hipMemcpyHostToDevice, A.stream()));
        if (beta != 0.0) SPHSTK_HIP(hipMemcpyAsync(dy.p, Y, sizeof(double) * n, hipMemcpyHostToDevice, A.stream()));
        A.apply(dx.p, dy.p, alpha, beta, prec);
        SPHSTK_HIP(hipMemcpyAsync(Y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost, A.stream()));
        SPHSTK_HIP(hipStreamSynchronize(A.stream()));
    });
}

int sphstk_operator_apply_inner(sphstk_operator *op, const double *X, double *Y, double alpha, double beta) {
    return apply_prec(op, X, Y, alpha, beta, 1);
}

int sphstk_operator_apply_outer(sphstk_operator *op, const double *X, double *Y, double alpha, double beta) {
    return apply_prec(op, X, Y, alpha, beta, 2);
}

int sphstk_operator_apply_coarse(sphstk_operator *op, const double *X, double *Y, double alpha, double beta) {
    return apply_prec(op, X, Y, alpha, beta, 3);
}

int sphstk_operator_apply_p2p(sphstk_operator *op, const double *in, double *out, double cId, double cSL,
                              double cTrac) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        op_p2p(*op->op, in, out, cId, cSL, cTrac, false);
    });
}

int sphstk_operator_apply_p2p_device(sphstk_operator *op, const double *din, double *dout, double cId, double cSL,
                                     double cTrac) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        op_p2p(*op->op, din, dout, cId, cSL, cTrac, true);
    });
}

int sphstk_operator_apply_lop(sphstk_operator *op, const double *in, double *out, double cLOP) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        SPHSTK_REQUIRE(in && out, "in and out must not be NULL");
        Operator &A = *op->op;
        const int64_t n = A.ndof();
        DevBuf<double> di, dout;
        di.alloc(std::max<int64_t>(n, 1));
        dout.alloc(std::max<int64_t>(n, 1));
        SPHSTK_HIP(hipMemcpyAsync(di.p, in, sizeof(double) * n, hipMemcpyHostToDevice, A.stream()));
        SPHSTK_HIP(hipMemcpyAsync(dout.p, out, sizeof(double) * n, hipMemcpyHostToDevice, A.stream()));
        A.lop(di.p, dout.p, cLOP, nullptr);
        SPHSTK_HIP(hipMemcpyAsync(out, dout.p, sizeof(double) * n, hipMemcpyDeviceToHost, A.stream()));
        SPHSTK_HIP(hipStreamSynchronize(A.stream()));
    });
}

int sphstk_operator_pole_values(sphstk_operator *op, const double *in, double *poles) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        SPHSTK_REQUIRE(in && poles, "in and poles must not be NULL");
        Operator &A = *op->op;
        const int64_t n = A.ndof();
        DevBuf<double> d;
        d.alloc(std::max<int64_t>(n, 1));
        SPHSTK_HIP(hipMemcpyAsync(d.p, in, sizeof(double) * n, hipMemcpyHostToDevice, A.stream()));
        A.pole_values(d.p, poles);
    });
}

int sphstk_operator_apply_self_inverse(sphstk_operator *op, const double *in, double *out) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        SPHSTK_REQUIRE(in && out, "in and out must not be NULL");
        Operator &A = *op->op;
        const int64_t n = A.ndof();
        DevBuf<double> a, b;
        a.alloc(std::max<int64_t>(n, 1));
        b.alloc(std::max<int64_t>(n, 1));
        SPHSTK_HIP(hipMemcpyAsync(a.p, in, sizeof(double) * n, hipMemcpyHostToDevice, A.stream()));
        A.precondition(a.p, b.p, nullptr);
        SPHSTK_HIP(hipMemcpyAsync(out, b.p, sizeof(double) * n, hipMemcpyDeviceToHost, A.stream()));
        SPHSTK_HIP(hipStreamSynchronize(A.stream()));
    });
}

int sphstk_operator_project_nullspace(sphstk_operator *op, double *inout) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        SPHSTK_REQUIRE(inout, "inout must not be NULL");
        Operator &A = *op->op;
        const int64_t n = A.ndof();
        DevBuf<double> d;
        d.alloc(std::max<int64_t>(n, 1));
        SPHSTK_HIP(hipMemcpyAsync(d.p, inout, sizeof(double) * n, hipMemcpyHostToDevice, A.stream()));
        A.project(d.p);
        SPHSTK_HIP(hipMemcpyAsync(inout, d.p, sizeof(double) * n, hipMemcpyDeviceToHost, A.stream()));
        SPHSTK_HIP(hipStreamSynchronize(A.stream()));
    });
}

int sphstk_operator_get_grid(const sphstk_operator *op, double *weights, double *coords, double *norms,
                             int64_t *index, int64_t *length) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        op->op->grid(weights, coords, norms, index, length);
    });
}

int sphstk_operator_get_near_pairs(const sphstk_operator *op, int32_t *trg, int32_t *src) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        op->op->near_pairs(trg, src);
    });
}

int sphstk_operator_far_shape(const sphstk_operator *op, int *tpl, int *split, int *inner_tpl) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        if (tpl) *tpl = op->op->far_tpl();
        if (split) *split = op->op->far_split();
        if (inner_tpl) *inner_tpl = op->op->inner_tpl();
    });
}

int sphstk_operator_kernel_stats(const sphstk_operator *op, int kind, int prec, double *total_ms, int64_t *launches,
                                 double *pairs_per_launch) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        op->op->kernel_stats(kind, prec, total_ms, launches, pairs_per_launch);
    });
}

int sphstk_operator_reset_stats(sphstk_operator *op) {
    return guard([&] {
        SPHSTK_REQUIRE(op && op->op, "invalid operator handle");
        op->op->reset_stats();
    });
}

int sphstk_mobility_create(int n_spheres, const double *pos, const double *radius, const double *quat_wxyz,
                           const int *order, double viscosity, const sphstk_solver_params *params, int precision,
                           int flags, sphstk_mobility **out) {
    return sphstk_mobility_create_batched(1, &n_spheres, pos, radius, quat_wxyz, order, viscosity, params, precision,
                                          flags, out);
}

int sphstk_mobility_create_batched(int n_susp, const int *sphere_counts, const double *pos, const double *radius,
                                   const double *quat_wxyz, const int *order, double viscosity,
                                   const sphstk_solver_params *params, int precision, int flags,
                                   sphstk_mobility **out) {
    return guard([&] {
        SPHSTK_REQUIRE(out, "out must not be NULL");
        const SolverParams prm = to_params(params);
        auto h = std::make_unique<sphstk_mobility>();
        h->mob = std::make_unique<Mobility>(n_susp, sphere_counts, pos, radius, quat_wxyz, order, viscosity, prm,
                                            precision, flags);
        h->view.op = &h->mob->op();
        const int n = h->mob->op().nsph();
        h->counts.assign(sphere_counts, sphere_counts + n_susp);
        h->order.assign(order, order + n);
        h->pos.assign(pos, pos + 3 * n);
        h->radius.assign(radius, radius + n);
        h->quat.assign(quat_wxyz, quat_wxyz + 4 * n);
        h->viscosity = viscosity;
        h->flags = flags;
        *out = h.release();
    });
}

int sphstk_mobility_destroy(sphstk_mobility *mob) {
    return guard([&] { delete mob; });
}

sphstk_operator *sphstk_mobility_operator(sphstk_mobility *mob) { return mob ? &mob->view : nullptr; }

int sphstk_mobility_set_params(sphstk_mobility *mob, const sphstk_solver_params *params, int precision) {
    return guard([&] {
        SPHSTK_REQUIRE(mob, "invalid mobility handle");
        SPHSTK_REQUIRE(precision == 0 || precision == 1, "precision must be 0 (FP64) or 1 (mixed)");
        mob->mob->set_params(to_params(params), precision);
    });
}

int sphstk_mobility_solve(sphstk_mobility *mob, const double *force, double *vel, sphstk_solve_info *info) {
    return guard([&] {
        SPHSTK_REQUIRE(mob, "invalid mobility handle");
        SPHSTK_REQUIRE(mob->mob->op().nsph() == 0 || (force && vel), "force and vel must not be NULL");
        int outer = 0, a64 = 0, a32 = 0;
        mob->mob->solve(force, vel, mob->last, outer, a64, a32);
        fill_info(info, mob->last, outer, a64, a32);
    });
}

int sphstk_mobility_apply(sphstk_mobility *mob, const double *X, double *Y, int ncol, double alpha, double beta) {
    return guard([&] {
        SPHSTK_REQUIRE(mob, "invalid mobility handle");
        SPHSTK_REQUIRE(ncol >= 1 && X && Y, "invalid arguments");
        const int64_t n6 = 6 * (int64_t)mob->mob->op().nsph();
        std::vector<double> v(n6);
        int outer, a64, a32;
        for (int c = 0; c < ncol; c++) {
            mob->mob->solve(X + c * n6, v.data(), mob->last, outer, a64, a32);
            for (int64_t i = 0; i < n6; i++) Y[c * n6 + i] = (beta == 0.0 ? 0.0 : beta * Y[c * n6 + i]) + alpha * v[i];
        }
    });
}

int sphstk_mobility_solve_multi(sphstk_mobility *mob, int ncol, const double *force, double *vel,
                                sphstk_solve_info *info) {
    return guard([&] {
        SPHSTK_REQUIRE(mob && force && vel, "invalid arguments");
        SPHSTK_REQUIRE(ncol >= 1 && ncol <= 64, "ncol must lie in [1, 64]");
        const int prec = mob->mob->precision();
        const int B = (int)mob->counts.size();
        if (!mob->multi || mob->multi_ncol != ncol || mob->multi_prec != prec) {
            mob->multi.reset();
            std::vector<int> counts, order;
            std::vector<double> pos, radius, quat;
            for (int c = 0; c < ncol; c++) {
                counts.insert(counts.end(), mob->counts.begin(), mob->counts.end());
                order.insert(order.end(), mob->order.begin(), mob->order.end());
                pos.insert(pos.end(), mob->pos.begin(), mob->pos.end());
                radius.insert(radius.end(), mob->radius.begin(), mob->radius.end());
                quat.insert(quat.end(), mob->quat.begin(), mob->quat.end());
            }
            // FP64: copies that share the traction pair sums; mixed: the copies
            // as independent suspensions of one batched solve
            mob->multi = std::make_unique<Mobility>(ncol * B, counts.data(), pos.data(), radius.data(), quat.data(),
                                                    order.data(), mob->viscosity, mob->mob->params(), prec,
                                                    mob->flags & ~4, prec == 0 ? ncol : 1);
            mob->multi_ncol = ncol;
            mob->multi_prec = prec;
        }
        mob->multi->set_params(mob->mob->params(), prec);
        std::vector<SegInfo> seg;
        int outer = 0, a64 = 0, a32 = 0;
        mob->multi->set_unknown_zero();
        mob->multi->solve(force, vel, seg, outer, a64, a32);
        fill_info(info, seg, outer, a64, a32);
    });
}

int sphstk_mobility_get_density(const sphstk_mobility *mob, double *rho) {
    return guard([&] {
        SPHSTK_REQUIRE(mob && rho, "invalid arguments");
        mob->mob->get_density(rho);
    });
}

int sphstk_mobility_get_unknown(const sphstk_mobility *mob, double *x) {
    return guard([&] {
        SPHSTK_REQUIRE(mob && x, "invalid arguments");
        mob->mob->get_unknown(x);
    });
}

int sphstk_mobility_set_unknown(sphstk_mobility *mob, const double *x) {
    return guard([&] {
        SPHSTK_REQUIRE(mob && x, "invalid arguments");
        mob->mob->set_unknown(x);
    });
}

int sphstk_mobility_set_density(sphstk_mobility *mob, const double *rho) {
    return guard([&] {
        SPHSTK_REQUIRE(mob && rho, "invalid arguments");
        mob->mob->set_density(rho);
    });
}

int sphstk_brownian_velocity(int n, const double *radius, double viscosity, double kBT, double dt, double scale,
                             uint64_t seed, uint64_t step, double *vel) {
    return guard([&] { brownian_velocity(n, radius, viscosity, kBT, dt, scale, seed, step, vel); });
}

int sphstk_collision_pairs(int n, const double *pos, const double *radius_col, const int *gid, double colbuf,
                           int64_t capacity, int64_t *npairs, int *idx_i, int *idx_j, double *phi0, double *norm_i) {
    return guard([&] {
        SPHSTK_REQUIRE(npairs, "npairs must not be NULL");
        *npairs = collision_pairs(n, pos, radius_col, gid, colbuf, capacity, idx_i, idx_j, phi0, norm_i);
    });
}

int sphstk_collision_pairs_noalloc(int n, const double *pos, const double *radius_col, const int *gid, double colbuf,
                                   int64_t capacity, int64_t *npairs, int *idx_i, int *idx_j, double *phi0,
                                   double *norm_i, void *stream) {
    return guard([&] {
        SPHSTK_REQUIRE(npairs, "npairs must not be NULL");
        *npairs = collision_pairs_noalloc(n, pos, radius_col, gid, colbuf, capacity, idx_i, idx_j, phi0, norm_i,
                                          (hipStream_t)stream);
    });
}

int sphstk_collision_lcp(sphstk_mobility *mob, int n, const double *radius, double viscosity, int64_t npairs,
                         const int *idx_i, const int *idx_j, const double *phi0, const double *norm_i,
                         const double *vel_known, double dt, double buffer, double tol, int max_iters, double *gamma,
                         double *force_col, double *vel_col, sphstk_lcp_info *info) {
    return guard([&] {
        MobilityFn M;
        if (mob) {
            SPHSTK_REQUIRE(n == mob->mob->op().nsph(), "n must equal the handle's number of spheres");
            M = [mob](const double *F, double *V) {
                int outer = 0, a64 = 0, a32 = 0;
                mob->mob->solve(F, V, mob->last, outer, a64, a32);
            };
        } else {
            SPHSTK_REQUIRE(n == 0 || radius, "radius must not be NULL without a mobility handle");
            SPHSTK_REQUIRE(std::isfinite(viscosity) && viscosity > 0, "viscosity must be > 0");
            for (int i = 0; i < n; i++) SPHSTK_REQUIRE(std::isfinite(radius[i]) && radius[i] > 0, "radii must be > 0");
            M = [n, radius, viscosity](const double *F, double *V) { free_drain_mobility(n, radius, viscosity, F, V); };
        }
        const LcpResult r = collision_lcp(M, n, npairs, idx_i, idx_j, phi0, norm_i, vel_known, dt, buffer, tol,
                                          max_iters, gamma, force_col, vel_col);
        if (info) {
            info->iters = r.iters;
            info->mobility_solves = r.mobility_solves;
            info->converged = r.converged;
            info->residual = r.residual;
        }
    });
}

} // extern "C"
