// The steps beside the mobility solve in the reference's time loop, on the GPU:
// Brownian velocities (SphereSystem::getVelocityBrown, SRC/SphereSystem.cpp:323-360)
// and collision pairs (CollisionSphere::collide, Sphere/CollisionSphere.hpp:51-90,
// found by SphereSystem::resolveCollision, SRC/SphereSystem.cpp:414-431).
//
// Both are integer/geometry work bound by memory latency, not arithmetic:
// * Brownian: one thread per sphere, three Philox4x32-10 blocks and three
//   Box-Muller pairs, 8 B in and 48 B out per sphere.
// * Collision pairs: a uniform cell grid of edge >= the largest interaction
//   distance, spheres sorted by cell (radix sort), then one thread per sphere
//   scans the 27 neighbouring cells twice: once to count its pairs, once (after
//   an exclusive scan of the counts) to write them, sorted by partner, into its
//   own segment.  No atomics; the output order (I ascending, then J) does not
//   depend on the launch.
#include "steps.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

namespace sphstk {

namespace {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC'11)

__device__ inline void philox4x32_10(unsigned c[4], unsigned k0, unsigned k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const unsigned a = c[0], b = c[2];
        const unsigned hi0 = __umulhi(0xD2511F53u, a), lo0 = 0xD2511F53u * a;
        const unsigned hi1 = __umulhi(0xCD9E8D57u, b), lo1 = 0xCD9E8D57u * b;
        const unsigned n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

// 53-bit uniform in (0, 1) from two 32-bit words
__device__ inline double u53(unsigned hi, unsigned lo) {
    return ((double)(((unsigned long long)hi << 21) | (lo >> 11)) + 0.5) * 0x1p-53;
}

__global__ void k_brownian(int n, const double *__restrict__ radius, double mu, double fack, unsigned long long seed,
                           unsigned long long step, double *__restrict__ vel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double R = radius[i];
    const double dx = sqrt(1.0 / (6.0 * kPi * R * mu)) * fack, dr = sqrt(1.0 / (8.0 * kPi * R * R * R * mu)) * fack;
    double z[6];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        unsigned b[4] = {(unsigned)i, (unsigned)c, (unsigned)step, (unsigned)(step >> 32)};
        philox4x32_10(b, (unsigned)seed, (unsigned)(seed >> 32));
        const double ua = u53(b[0], b[1]), ub = u53(b[2], b[3]);
        const double m = sqrt(-2.0 * log(ua)), th = 2.0 * kPi * ub;
        z[2 * c] = m * cos(th);
        z[2 * c + 1] = m * sin(th);
    }
    double *o = vel + 6 * (int64_t)i;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        o[k] = dx * z[k];
        o[3 + k] = dr * z[3 + k];
    }
}

// ---------------------------------------------------------------------------
// Collision pairs on a uniform cell grid

struct CellGrid {
    double lo[3];
    double inv_h; // 1 / cell edge
    int dim[3];
};

__device__ inline int cell_coord(double x, double lo, double inv_h, int dim) {
    int c = (int)floor((x - lo) * inv_h);
    return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
}

__global__ void k_cell_keys(int n, const double *__restrict__ pos, CellGrid g, unsigned *__restrict__ key,
                            int *__restrict__ idx) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int cx = cell_coord(pos[3 * a], g.lo[0], g.inv_h, g.dim[0]);
    const int cy = cell_coord(pos[3 * a + 1], g.lo[1], g.inv_h, g.dim[1]);
    const int cz = cell_coord(pos[3 * a + 2], g.lo[2], g.inv_h, g.dim[2]);
    key[a] = (unsigned)((cx * g.dim[1] + cy) * g.dim[2] + cz);
    idx[a] = a;
}

__global__ void k_fill_i32(int64_t n, int v, int *__restrict__ y) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

// begin/end of each occupied cell in the sorted order
__global__ void k_cell_bounds(int n, const unsigned *__restrict__ key, int *__restrict__ begin,
                              int *__restrict__ end) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const unsigned c = key[k];
    if (k == 0 || key[k - 1] != c) begin[c] = k;
    if (k == n - 1 || key[k + 1] != c) end[c] = k + 1;
}

// CollisionSphere::collide's test (Sphere/CollisionSphere.hpp:55-70) for the
// ordered pair (a, b); r is |x_b - x_a|
__device__ inline bool collides(const double *__restrict__ pos, const double *__restrict__ rcol,
                                const int *__restrict__ gid, double colbuf, int a, int b, double &dx, double &dy,
                                double &dz, double &r, double &sep) {
    if (gid[a] >= gid[b]) return false;
    dx = pos[3 * b] - pos[3 * a];
    dy = pos[3 * b + 1] - pos[3 * a + 1];
    dz = pos[3 * b + 2] - pos[3 * a + 2];
    r = sqrt(fma(dx, dx, fma(dy, dy, dz * dz)));
    sep = (r - rcol[a]) - rcol[b];
    return sep < colbuf * rcol[a];
}

// FILL = false: count the pairs of sphere a (I side) into count[a].
// FILL = true: write them at off[a], sorted by partner index.
template <bool FILL>
__global__ void k_pairs(int n, const double *__restrict__ pos, const double *__restrict__ rcol,
                        const int *__restrict__ gid, double colbuf, CellGrid g, const int *__restrict__ sorted,
                        const int *__restrict__ begin, const int *__restrict__ end, int64_t *__restrict__ count,
                        const int64_t *__restrict__ off, int *__restrict__ ii, int *__restrict__ jj,
                        double *__restrict__ phi0, double *__restrict__ norm) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int cx = cell_coord(pos[3 * a], g.lo[0], g.inv_h, g.dim[0]);
    const int cy = cell_coord(pos[3 * a + 1], g.lo[1], g.inv_h, g.dim[1]);
    const int cz = cell_coord(pos[3 * a + 2], g.lo[2], g.inv_h, g.dim[2]);
    int m = 0;
    const int64_t o = FILL ? off[a] : 0;
    for (int x = max(cx - 1, 0); x <= min(cx + 1, g.dim[0] - 1); x++)
        for (int y = max(cy - 1, 0); y <= min(cy + 1, g.dim[1] - 1); y++)
            for (int z = max(cz - 1, 0); z <= min(cz + 1, g.dim[2] - 1); z++) {
                const int c = (x * g.dim[1] + y) * g.dim[2] + z;
                for (int k = begin[c]; k < end[c]; k++) {
                    const int b = sorted[k];
                    double dx, dy, dz, r, sep;
                    if (!collides(pos, rcol, gid, colbuf, a, b, dx, dy, dz, r, sep)) continue;
                    if (FILL) {
                        // insertion into the sorted segment jj[o, o + m)
                        int t = m;
                        while (t > 0 && jj[o + t - 1] > b) {
                            jj[o + t] = jj[o + t - 1];
                            t--;
                        }
                        jj[o + t] = b;
                    }
                    m++;
                }
            }
    if (!FILL) {
        count[a] = m;
        return;
    }
    for (int t = 0; t < m; t++) {
        const int b = jj[o + t];
        double dx, dy, dz, r, sep;
        collides(pos, rcol, gid, colbuf, a, b, dx, dy, dz, r, sep);
        const double s = -1.0 / r;
        ii[o + t] = a;
        phi0[o + t] = sep;
        norm[3 * (o + t)] = dx * s;
        norm[3 * (o + t) + 1] = dy * s;
        norm[3 * (o + t) + 2] = dz * s;
    }
}

template <typename T>
struct Dev {
    T *p = nullptr;
    explicit Dev(size_t n) {
        if (n) SPHSTK_HIP(hipMalloc(&p, sizeof(T) * n));
    }
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { SPHSTK_HIP(hipStreamCreate(&s)); }
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
};

unsigned grid_of(int n) { return (unsigned)((n + 255) / 256); }
// the pair passes run one wave per block, so that even 16k spheres spread over
// all 256 CUs (one thread per sphere, latency-bound cell scans)
unsigned grid_of64(int n) { return (unsigned)((n + 63) / 64); }

} // namespace

void brownian_velocity(int n, const double *radius, double mu, double kBT, double dt, double scale,
                       unsigned long long seed, unsigned long long step, double *vel) {
    SPHSTK_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return;
    SPHSTK_REQUIRE(radius && vel, "radius and vel must not be NULL");
    SPHSTK_REQUIRE(std::isfinite(mu) && mu > 0, "viscosity must be > 0");
    SPHSTK_REQUIRE(std::isfinite(kBT) && kBT >= 0, "kBT must be >= 0");
    SPHSTK_REQUIRE(std::isfinite(dt) && dt > 0, "dt must be > 0");
    SPHSTK_REQUIRE(std::isfinite(scale) && scale >= 0, "scale must be >= 0");
    for (int i = 0; i < n; i++)
        SPHSTK_REQUIRE(std::isfinite(radius[i]) && radius[i] > 0, "radii must be > 0");
    Stream st;
    Dev<double> r(n), v(6 * (size_t)n);
    SPHSTK_HIP(hipMemcpyAsync(r.p, radius, sizeof(double) * n, hipMemcpyHostToDevice, st.s));
    const double fack = std::sqrt(2.0 * kBT / dt) * scale;
    k_brownian<<<grid_of(n), 256, 0, st.s>>>(n, r.p, mu, fack, seed, step, v.p);
    SPHSTK_HIP(hipGetLastError());
    SPHSTK_HIP(hipMemcpyAsync(vel, v.p, sizeof(double) * 6 * n, hipMemcpyDeviceToHost, st.s));
    SPHSTK_HIP(hipStreamSynchronize(st.s));
}

int64_t collision_pairs(int n, const double *pos, const double *rcol, const int *gid_in, double colbuf,
                        int64_t capacity, int *ii, int *jj, double *phi0, double *norm) {
    SPHSTK_REQUIRE(n >= 0, "n must be >= 0");
    if (n < 2) return 0;
    SPHSTK_REQUIRE(pos && rcol, "pos and radius_col must not be NULL");
    SPHSTK_REQUIRE(std::isfinite(colbuf) && colbuf > -1.0, "colbuf must be > -1");
    SPHSTK_REQUIRE(capacity >= 0, "capacity must be >= 0");
    double lo[3] = {pos[0], pos[1], pos[2]}, hi[3] = {pos[0], pos[1], pos[2]}, rmax = 0.0;
    for (int a = 0; a < n; a++) {
        for (int d = 0; d < 3; d++) {
            SPHSTK_REQUIRE(std::isfinite(pos[3 * a + d]), "positions must be finite");
            lo[d] = std::min(lo[d], pos[3 * a + d]);
            hi[d] = std::max(hi[d], pos[3 * a + d]);
        }
        SPHSTK_REQUIRE(std::isfinite(rcol[a]) && rcol[a] > 0, "collision radii must be > 0");
        rmax = std::max(rmax, rcol[a]);
    }
    std::vector<int> gid(n);
    for (int a = 0; a < n; a++) gid[a] = gid_in ? gid_in[a] : a;
    // a colliding pair has |x_b - x_a| < R_a + R_b + colbuf R_a <= (2 + max(colbuf, 0)) R_max:
    // such pairs lie in the same or adjacent cells of edge h >= that distance
    double h = (2.0 + std::max(colbuf, 0.0)) * rmax * (1.0 + 1e-9);
    CellGrid g{};
    const double max_cells = std::max(64.0, 8.0 * n);
    for (;;) {
        double cells = 1.0;
        for (int d = 0; d < 3; d++) cells *= std::floor((hi[d] - lo[d]) / h) + 1.0;
        if (cells <= max_cells) break;
        h *= std::max(1.01, std::cbrt(cells / max_cells));
    }
    for (int d = 0; d < 3; d++) {
        g.lo[d] = lo[d];
        g.dim[d] = (int)(std::floor((hi[d] - lo[d]) / h) + 1.0);
    }
    g.inv_h = 1.0 / h;
    const int64_t ncells = (int64_t)g.dim[0] * g.dim[1] * g.dim[2];
    SPHSTK_REQUIRE(ncells < (int64_t(1) << 31), "too many spheres for the cell grid");
    int key_bits = 1;
    while (key_bits < 32 && (1u << key_bits) < (unsigned)ncells) key_bits++;

    Stream st;
    Dev<double> dpos(3 * (size_t)n), drc(n);
    Dev<int> dgid(n), idx(n), sorted(n), begin(ncells), end(ncells);
    Dev<int64_t> count(n + 1);
    Dev<unsigned> key(n), skey(n);
    Dev<int64_t> off(n + 1);
    SPHSTK_HIP(hipMemcpyAsync(dpos.p, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st.s));
    SPHSTK_HIP(hipMemcpyAsync(drc.p, rcol, sizeof(double) * n, hipMemcpyHostToDevice, st.s));
    SPHSTK_HIP(hipMemcpyAsync(dgid.p, gid.data(), sizeof(int) * n, hipMemcpyHostToDevice, st.s));
    k_cell_keys<<<grid_of(n), 256, 0, st.s>>>(n, dpos.p, g, key.p, idx.p);
    size_t tmp_sort = 0, tmp_scan = 0;
    SPHSTK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_sort, key.p, skey.p, idx.p, sorted.p, n, 0, key_bits,
                                                  st.s));
    SPHSTK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_scan, count.p, off.p, n + 1, st.s));
    Dev<unsigned char> tmp(std::max(tmp_sort, tmp_scan));
    SPHSTK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_sort, key.p, skey.p, idx.p, sorted.p, n, 0, key_bits,
                                                  st.s));
    const unsigned cg = (unsigned)((ncells + 255) / 256);
    k_fill_i32<<<cg, 256, 0, st.s>>>(ncells, 0, begin.p);
    k_fill_i32<<<cg, 256, 0, st.s>>>(ncells, 0, end.p);
    SPHSTK_HIP(hipMemsetAsync(count.p + n, 0, sizeof(int64_t), st.s));
    k_cell_bounds<<<grid_of(n), 256, 0, st.s>>>(n, skey.p, begin.p, end.p);
    k_pairs<false><<<grid_of64(n), 64, 0, st.s>>>(n, dpos.p, drc.p, dgid.p, colbuf, g, sorted.p, begin.p, end.p,
                                                 count.p, nullptr, nullptr, nullptr, nullptr, nullptr);
    SPHSTK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_scan, count.p, off.p, n + 1, st.s));
    SPHSTK_HIP(hipGetLastError());
    int64_t total = 0;
    SPHSTK_HIP(hipMemcpyAsync(&total, off.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, st.s));
    SPHSTK_HIP(hipStreamSynchronize(st.s));
    const bool want = ii || jj || phi0 || norm;
    if (!want || total > capacity || total == 0) return total;
    Dev<int> dii(total), djj(total);
    Dev<double> dphi(total), dnorm(3 * (size_t)total);
    k_pairs<true><<<grid_of64(n), 64, 0, st.s>>>(n, dpos.p, drc.p, dgid.p, colbuf, g, sorted.p, begin.p, end.p,
                                                nullptr, off.p, dii.p, djj.p, dphi.p, dnorm.p);
    SPHSTK_HIP(hipGetLastError());
    if (ii) SPHSTK_HIP(hipMemcpyAsync(ii, dii.p, sizeof(int) * total, hipMemcpyDeviceToHost, st.s));
    if (jj) SPHSTK_HIP(hipMemcpyAsync(jj, djj.p, sizeof(int) * total, hipMemcpyDeviceToHost, st.s));
    if (phi0) SPHSTK_HIP(hipMemcpyAsync(phi0, dphi.p, sizeof(double) * total, hipMemcpyDeviceToHost, st.s));
    if (norm) SPHSTK_HIP(hipMemcpyAsync(norm, dnorm.p, sizeof(double) * 3 * total, hipMemcpyDeviceToHost, st.s));
    SPHSTK_HIP(hipStreamSynchronize(st.s));
    return total;
}

} // namespace sphstk
