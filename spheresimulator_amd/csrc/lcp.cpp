// The collision LCP around the mobility solve: the caller of
// SphereSTKMobMat::apply in SphereSystem::resolveCollision
// (SRC/SphereSystem.cpp:445-474), whose solver (SimToolbox's CollisionSolver)
// is not part of the reference tree.  Restated from the published method:
// Barzilai-Borwein projected gradient descent (BBPGD; Barzilai and Borwein,
// IMA J. Numer. Anal. 8 (1988); used for this LCP by Yan et al., J. Comput.
// Phys. 2019) on
//
//   find gamma >= 0 with phi(gamma) = phi0 + buffer + dt D^T (v_known + M D gamma) >= 0,
//   gamma^T phi = 0,
//
// where D maps the pair forces gamma to sphere forces (F_I += gamma normI,
// F_J -= gamma normI: contact on the line of centres, no torque) and M is the
// mobility: one full mobility solve (warm-started from the previous one) per
// iteration, or the free-drain diagonal 1/(6 pi mu R), 1/(8 pi mu R^3) of
// SphereSystem::getMobOperator without hydrodynamics (SRC/SphereSystem.cpp:237-268).
#include "lcp.hpp"

#include <algorithm>
#include <cmath>

namespace sphstk {

namespace {

struct Pairs {
    int64_t m;
    const int *ii, *jj;
    const double *nrm;
    // F = D gamma (6 per sphere; torques stay zero)
    void force(const std::vector<double> &g, std::vector<double> &F) const {
        std::fill(F.begin(), F.end(), 0.0);
        for (int64_t k = 0; k < m; k++)
            for (int c = 0; c < 3; c++) {
                F[6 * (int64_t)ii[k] + c] += g[k] * nrm[3 * k + c];
                F[6 * (int64_t)jj[k] + c] -= g[k] * nrm[3 * k + c];
            }
    }
    // q = D^T v: the rate of change of each gap
    void rate(const double *v, std::vector<double> &q) const {
        for (int64_t k = 0; k < m; k++) {
            double s = 0.0;
            for (int c = 0; c < 3; c++) s += nrm[3 * k + c] * (v[6 * (int64_t)ii[k] + c] - v[6 * (int64_t)jj[k] + c]);
            q[k] = s;
        }
    }
};

double dot(const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
}

// max_k |min(gamma_k, phi_k)|: zero exactly at a solution of the LCP
double minmap(const std::vector<double> &g, const std::vector<double> &phi) {
    double r = 0.0;
    for (size_t k = 0; k < g.size(); k++) r = std::max(r, std::fabs(std::min(g[k], phi[k])));
    return r;
}

} // namespace

LcpResult collision_lcp(const MobilityFn &mobility, int n, int64_t m, const int *ii, const int *jj,
                        const double *phi0, const double *nrm, const double *vel_known, double dt, double buffer,
                        double tol, int max_iters, double *gamma, double *force_col, double *vel_col) {
    SPHSTK_REQUIRE(n >= 0 && m >= 0, "sizes must be >= 0");
    SPHSTK_REQUIRE(std::isfinite(dt) && dt > 0, "dt must be > 0");
    SPHSTK_REQUIRE(std::isfinite(tol) && tol > 0, "tol must be > 0");
    SPHSTK_REQUIRE(max_iters >= 1, "max_iters must be >= 1");
    SPHSTK_REQUIRE(std::isfinite(buffer), "buffer must be finite");
    SPHSTK_REQUIRE(m == 0 || (ii && jj && phi0 && nrm && gamma), "pair arrays must not be NULL");
    SPHSTK_REQUIRE(vel_known && force_col && vel_col, "velocity and force arrays must not be NULL");
    for (int64_t k = 0; k < m; k++) {
        SPHSTK_REQUIRE(ii[k] >= 0 && ii[k] < n && jj[k] >= 0 && jj[k] < n && ii[k] != jj[k],
                       "pair indices must name two different spheres");
        SPHSTK_REQUIRE(std::isfinite(phi0[k]), "phi0 must be finite");
    }
    LcpResult res;
    const size_t n6 = 6 * (size_t)n;
    std::fill(force_col, force_col + n6, 0.0);
    std::fill(vel_col, vel_col + n6, 0.0);
    if (m == 0) {
        res.converged = 1;
        return res;
    }
    const Pairs P{m, ii, jj, nrm};
    // b = phi0 + buffer + dt D^T v_known
    std::vector<double> b(m), q(m);
    P.rate(vel_known, q);
    for (int64_t k = 0; k < m; k++) b[k] = phi0[k] + buffer + dt * q[k];
    std::vector<double> F(n6), U(n6);
    // A x = dt D^T M D x (one mobility application); U keeps M D x
    auto applyA = [&](const std::vector<double> &x, std::vector<double> &Ax, std::vector<double> &Ux) {
        P.force(x, F);
        mobility(F.data(), Ux.data());
        res.mobility_solves++;
        P.rate(Ux.data(), Ax);
        for (int64_t k = 0; k < m; k++) Ax[k] *= dt;
    };
    std::vector<double> g(m), gn(m), phi(m), phin(m), s(m), y(m), Ag(m), Ug(n6), Un(n6);
    for (int64_t k = 0; k < m; k++) g[k] = std::max(gamma[k], 0.0);
    applyA(g, phi, U);
    for (int64_t k = 0; k < m; k++) phi[k] += b[k];
    res.residual = minmap(g, phi);
    if (res.residual > tol) {
        // first step: exact line search along the gradient, alpha = g^T g / g^T A g
        applyA(phi, Ag, Ug);
        const double gAg = dot(phi, Ag);
        double alpha = gAg > 0 ? dot(phi, phi) / gAg : 1.0;
        for (int it = 1; it <= max_iters; it++) {
            for (int64_t k = 0; k < m; k++) gn[k] = std::max(g[k] - alpha * phi[k], 0.0);
            applyA(gn, phin, Un);
            for (int64_t k = 0; k < m; k++) phin[k] += b[k];
            res.iters = it;
            res.residual = minmap(gn, phin);
            for (int64_t k = 0; k < m; k++) {
                s[k] = gn[k] - g[k];
                y[k] = phin[k] - phi[k];
            }
            g.swap(gn);
            phi.swap(phin);
            U.swap(Un);
            if (res.residual <= tol) break;
            // alternating Barzilai-Borwein steps
            const double ss = dot(s, s), sy = dot(s, y), yy = dot(y, y);
            if (ss == 0.0) break; // no movement: the projection holds the iterate
            const double a = (it % 2) ? ss / sy : sy / yy;
            if (std::isfinite(a) && a > 0) alpha = a;
        }
    }
    res.converged = res.residual <= tol;
    for (int64_t k = 0; k < m; k++) gamma[k] = g[k];
    P.force(g, F);
    std::copy(F.begin(), F.end(), force_col);
    std::copy(U.begin(), U.end(), vel_col);
    return res;
}

void free_drain_mobility(int n, const double *radius, double mu, const double *F, double *V) {
    for (int i = 0; i < n; i++) {
        const double R = radius[i];
        const double t = 1.0 / (6.0 * kPi * R * mu), r = 1.0 / (8.0 * kPi * R * R * R * mu);
        for (int c = 0; c < 3; c++) {
            V[6 * i + c] = t * F[6 * i + c];
            V[6 * i + 3 + c] = r * F[6 * i + 3 + c];
        }
    }
}

} // namespace sphstk
