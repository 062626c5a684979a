// Internal declarations shared by the host engine and the HIP kernels of
// libsphstk.so.  Nothing here is part of the C ABI (include/sphstk.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace sphstk {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Error type thrown inside the library and turned into a status code at the C ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define SPHSTK_HIP(call)                                                                                          \
    do {                                                                                                           \
        hipError_t e_ = (call);                                                                                    \
        if (e_ != hipSuccess)                                                                                      \
            throw ::sphstk::Error(-2, std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + ":" + \
                                          std::to_string(__LINE__) + " (" #call ")");                              \
    } while (0)

#define SPHSTK_REQUIRE(cond, msg)                                                                                  \
    do {                                                                                                           \
        if (!(cond)) throw ::sphstk::Error(-1, std::string(msg));                                                  \
    } while (0)

inline int tri(int n, int m) { return n * (n + 1) / 2 + m; }
inline int tri_size(int P) { return (P + 1) * (P + 2) / 2; }
inline int grid_points(int p) { return (p + 1) * (2 * p + 2); }

// Far-field source interval: sources [begin, end) interact with the lanes whose
// bit is set in mask[t] (word t covers targets t*64 .. t*64+63 of the group).
constexpr int kMaxTPL = 4;
struct FarInterval {
    int64_t begin, end;
    uint64_t mask[kMaxTPL];
};

// One far-field work item: up to 64*TPL consecutive targets of one suspension
// against one slice (`slot`) of their source intervals.
struct FarGroup {
    int64_t tstart;
    int64_t iv_begin, iv_end;
    int32_t tcount;
    int32_t susp;
    int32_t slot;
    int32_t pad;
};

// Number of complex arrays in the per-sphere expansion records.
constexpr int kRecSL = 10;   // grad H1, grad H2, H3, grad H4
constexpr int kRecTrac = 33; // 26 traction arrays + 7 self-synthesis arrays
constexpr int kRecTracNear = 26;

// Host-side spherical-harmonic tables for one order p.
struct ShTables {
    int p = 0, npts = 0, Q = 0;
    std::vector<double> eta; // 3*npts unit body-frame grid
    std::vector<double> w;   // npts unit-sphere weights
    // analysis matrix, column-major: A[c*rows + r], rows = 6Q, cols = 3*npts
    std::vector<double> anal;
    // synthesis matrix, column-major: rows = 3*npts, cols = 6Q
    std::vector<double> synth;
};

const ShTables &sh_tables(int p);

// The synthesis rows at the body-frame poles +z and -z: 6 rows (north x, y, z,
// south x, y, z) of 6Q coefficients each, row-major.
std::vector<double> pole_rows(int p);

// Per-order resampling matrices between the order-p grid and a coarser grid of
// order pc <= p, both acting on body-frame values (3 per point), column-major:
//   down (3 nc x 3 n): coarse values of the degree-pc truncation of the projection
//   up   (3 n x 3 nc): fine values of the degree-pc projection of coarse values
// (pc == p: identity in both directions).
struct Resample {
    int p = 0, pc = 0, n = 0, nc = 0;
    std::vector<double> down, up;
};
const Resample &resample_tables(int p, int pc);
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w);

// Equivalent-source template of the far-field cell aggregation (equiv.cpp):
// K equivalent points and Kc check points on the unit sphere (3 each), to be
// placed at radii re*rho and rc*rho around a cell of radius rho, and
// P (3K x 3Kc, row-major) with P G = I, G the Stokeslet matrix from the
// equivalent to the check points of the unit cell: a cell's point forces are
// represented outside its check sphere by rho P u_check.
struct EquivTables {
    int K = 0, Kc = 0;
    double re = 0.0, rc = 0.0;
    std::vector<double> eq, chk, P;
};
constexpr int kEquivK = 200; // equivalent point forces per cell of the FP32 operator
const EquivTables &equiv_tables(int K = kEquivK);

} // namespace sphstk
