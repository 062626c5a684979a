// Collision LCP (BBPGD) around the mobility: host-side iteration, one
// mobility application per step.
#pragma once
#include "common.hpp"

#include <cstdint>
#include <functional>

namespace sphstk {

// V = M F for 6 values per sphere (force/torque -> velocity/angular velocity)
using MobilityFn = std::function<void(const double *F, double *V)>;

struct LcpResult {
    int iters = 0;
    int mobility_solves = 0;
    int converged = 0;
    double residual = 0.0; // max_k |min(gamma_k, phi_k)|
};

// gamma: initial guess in (negative entries are clipped), solution out;
// force_col = D gamma, vel_col = M D gamma (6 per sphere)
LcpResult collision_lcp(const MobilityFn &mobility, int n, int64_t m, const int *ii, const int *jj,
                        const double *phi0, const double *nrm, const double *vel_known, double dt, double buffer,
                        double tol, int max_iters, double *gamma, double *force_col, double *vel_col);

// the free-drain (no hydrodynamic interaction) mobility of SRC/SphereSystem.cpp:237-268
void free_drain_mobility(int n, const double *radius, double mu, const double *F, double *V);

} // namespace sphstk
