// The steps beside the mobility solve: Brownian velocities and collision pairs
// (host arrays in and out; the work runs on the current HIP device).
#pragma once
#include "common.hpp"

#include <cstdint>

namespace sphstk {

// vel[6i..6i+5] = (sqrt(1/(6 pi R mu)) N^3, sqrt(1/(8 pi R^3 mu)) N^3) * sqrt(2 kBT/dt) * scale
// with standard normals from Philox4x32-10 (counter (i, c, step), key seed) and Box-Muller
void brownian_velocity(int n, const double *radius, double mu, double kBT, double dt, double scale,
                       unsigned long long seed, unsigned long long step, double *vel);

// pairs (I, J) with gid[I] < gid[J] and |x_J - x_I| - Rc_I - Rc_J < colbuf Rc_I, ordered by I then J;
// returns the count; the arrays (any may be NULL) are written only when the count <= capacity
int64_t collision_pairs(int n, const double *pos, const double *rcol, const int *gid, double colbuf,
                        int64_t capacity, int *ii, int *jj, double *phi0, double *norm);

// Near pairs of the boundary-integral operator on the device: (t, s) in one
// suspension (susp per sphere) with 0 < |c_s - c_t|^2 < (1.5 R_s + 1.5 R_t)^2,
// grouped by t with s ascending, returned as CSR on the host (pos/rad: host
// copies of the device arrays dpos/drad, for the cell grid).  Throws when two
// spheres of a suspension overlap (centre distance < R_s + R_t).
void near_pairs(int n, const double *pos, const double *rad, const double *dpos, const double *drad, const int *dsusp,
                hipStream_t s, std::vector<int64_t> &near_off, std::vector<int> &near_src);

// The same search on device arrays with no device allocation: every
// intermediate array lives in a static device scratch sized for
// kNoAllocMaxSpheres spheres; gid may be NULL (the index); the four outputs
// are all given or all NULL.  Work runs on stream s, which is synchronised
// before returning.  Calls that overlap on one device share the scratch and race.
constexpr int kNoAllocMaxSpheres = 1 << 18;
int64_t collision_pairs_noalloc(int n, const double *pos, const double *rcol, const int *gid, double colbuf,
                                int64_t capacity, int *ii, int *jj, double *phi0, double *norm, hipStream_t s);

} // namespace sphstk
