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
#include <mutex>
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
// (gid NULL: the sphere index is the id)
__device__ inline bool collides(const double *__restrict__ pos, const double *__restrict__ rcol,
                                const int *__restrict__ gid, double colbuf, int a, int b, double &dx, double &dy,
                                double &dz, double &r, double &sep) {
    if ((gid ? gid[a] : a) >= (gid ? gid[b] : b)) return false;
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

// Near pairs of the boundary-integral operator: (t, s) of one suspension with
// 0 < |c_s - c_t|^2 < (1.5 R_s + 1.5 R_t)^2 (MPI/NearInteraction.txx:430-441
// with radiusNear = 1.5 R, BIE/SphereSTKSHOperator.cpp:547).  Products and
// sums unfused, in the order of the host's scalar code.  COUNT pass: count[t]
// and the smallest s of the suspension closer than R_s + R_t (overlapping,
// -1 if none) in overlap[t]; FILL pass: the sources of t at off[t], ascending.
template <bool FILL>
__global__ void k_near(int n, const double *__restrict__ pos, const double *__restrict__ rad,
                       const int *__restrict__ susp, CellGrid g, const int *__restrict__ sorted,
                       const int *__restrict__ begin, const int *__restrict__ end, int64_t *__restrict__ count,
                       int *__restrict__ overlap, const int64_t *__restrict__ off, int *__restrict__ src) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double xt = pos[3 * t], yt = pos[3 * t + 1], zt = pos[3 * t + 2], rt = rad[t];
    const int bt = susp[t];
    const int cx = cell_coord(xt, g.lo[0], g.inv_h, g.dim[0]);
    const int cy = cell_coord(yt, g.lo[1], g.inv_h, g.dim[1]);
    const int cz = cell_coord(zt, g.lo[2], g.inv_h, g.dim[2]);
    int m = 0, ov = -1;
    const int64_t o = FILL ? off[t] : 0;
    for (int x = max(cx - 1, 0); x <= min(cx + 1, g.dim[0] - 1); x++)
        for (int y = max(cy - 1, 0); y <= min(cy + 1, g.dim[1] - 1); y++)
            for (int z = max(cz - 1, 0); z <= min(cz + 1, g.dim[2] - 1); z++) {
                const int c = (x * g.dim[1] + y) * g.dim[2] + z;
                for (int k = begin[c]; k < end[c]; k++) {
                    const int s = sorted[k];
                    if (s == t || susp[s] != bt) continue;
                    const double dx = pos[3 * s] - xt, dy = pos[3 * s + 1] - yt, dz = pos[3 * s + 2] - zt;
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    const double rs = rad[s] + rt;
                    if (d2 < rs * rs && (ov < 0 || s < ov)) ov = s;
                    const double rr = 1.5 * rad[s] + 1.5 * rt;
                    if (!(d2 > 0 && d2 < rr * rr)) continue;
                    if (FILL) {
                        int q = m;
                        while (q > 0 && src[o + q - 1] > s) {
                            src[o + q] = src[o + q - 1];
                            q--;
                        }
                        src[o + q] = s;
                    }
                    m++;
                }
            }
    if (!FILL) {
        count[t] = m;
        overlap[t] = ov;
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

// The cell grid of the collision search over spheres in the box [lo, hi] with
// largest collision radius rmax: a colliding pair has |x_b - x_a| < R_a + R_b +
// colbuf R_a <= (2 + max(colbuf, 0)) R_max, so such pairs lie in the same or
// adjacent cells of edge h >= that distance; h grows while the grid would have
// more than max(64, 8n) cells.  key_bits: bits of the largest cell key.
CellGrid cell_grid(int n, const double lo[3], const double hi[3], double rmax, double colbuf, int &key_bits) {
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
    key_bits = 1;
    while (key_bits < 32 && (1u << key_bits) < (unsigned)ncells) key_bits++;
    return g;
}

// Bounding box, largest collision radius and input check of n spheres on the
// device, by one workgroup: out[0..2] = lo, out[3..5] = hi, out[6] = max Rc,
// out[7] = 1 if a position is not finite or a radius is not > 0.
constexpr int kBoundsThreads = 256;
__global__ __launch_bounds__(kBoundsThreads) void k_bounds(int n, const double *__restrict__ pos,
                                                           const double *__restrict__ rcol, double *__restrict__ out) {
    __shared__ double sh[8][kBoundsThreads];
    double v[8] = {HUGE_VAL, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL, 0.0, 0.0};
    for (int a = threadIdx.x; a < n; a += blockDim.x) {
        for (int d = 0; d < 3; d++) {
            const double x = pos[3 * a + d];
            if (!isfinite(x)) v[7] = 1.0;
            v[d] = fmin(v[d], x);
            v[3 + d] = fmax(v[3 + d], x);
        }
        const double r = rcol[a];
        if (!(isfinite(r) && r > 0.0)) v[7] = 1.0;
        v[6] = fmax(v[6], r);
    }
    for (int k = 0; k < 8; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 8; k++) {
                const double a = sh[k][threadIdx.x], b = sh[k][threadIdx.x + o];
                sh[k][threadIdx.x] = k < 3 ? fmin(a, b) : fmax(a, b);
            }
        __syncthreads();
    }
    if (threadIdx.x < 8) out[threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_fill_i64(int64_t n, int64_t v, int64_t *__restrict__ y) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

// ---------------------------------------------------------------------------
// Scratch of the collision search that allocates nothing: one file-scope
// device array, reached from the host through the address that
// hipGetSymbolAddress returns for the current device (each device holds its
// own copy).  Layout for n <= kNoAllocMaxSpheres spheres: bounds (8 doubles),
// keys, sorted keys, indices, sorted indices (n each), cell begin / end (up to
// 8n cells each), pair counts and offsets (n + 1 each), then the radix sort and
// scan workspace.
constexpr int64_t kNA = kNoAllocMaxSpheres;
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kOffBounds = 0;
constexpr size_t kOffKey = align256(kOffBounds + 8 * sizeof(double));
constexpr size_t kOffSKey = align256(kOffKey + kNA * sizeof(unsigned));
constexpr size_t kOffIdx = align256(kOffSKey + kNA * sizeof(unsigned));
constexpr size_t kOffSorted = align256(kOffIdx + kNA * sizeof(int));
constexpr size_t kOffBegin = align256(kOffSorted + kNA * sizeof(int));
constexpr size_t kOffEnd = align256(kOffBegin + 8 * kNA * sizeof(int));
constexpr size_t kOffCount = align256(kOffEnd + 8 * kNA * sizeof(int));
constexpr size_t kOffOff = align256(kOffCount + (kNA + 1) * sizeof(int64_t));
constexpr size_t kOffTmp = align256(kOffOff + (kNA + 1) * sizeof(int64_t));
constexpr size_t kTmpBytes = size_t(16) << 20;
constexpr size_t kScratchBytes = kOffTmp + kTmpBytes;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the current device, looked up once per device.
unsigned char *scratch_on_current_device() {
    static std::mutex mu;
    static std::vector<unsigned char *> ptr;
    int dev = 0;
    SPHSTK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if ((int)ptr.size() <= dev) ptr.resize(dev + 1, nullptr);
    if (!ptr[dev]) {
        void *p = nullptr;
        SPHSTK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SPHSTK_REQUIRE(p != nullptr, "no device address for the collision-search scratch");
        ptr[dev] = (unsigned char *)p;
    }
    return ptr[dev];
}
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
    int key_bits = 0;
    const CellGrid g = cell_grid(n, lo, hi, rmax, colbuf, key_bits);
    const int64_t ncells = (int64_t)g.dim[0] * g.dim[1] * g.dim[2];

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

void near_pairs(int n, const double *pos, const double *rad, const double *dpos, const double *drad, const int *dsusp,
                hipStream_t s, std::vector<int64_t> &near_off, std::vector<int> &near_src) {
    near_off.assign(n + 1, 0);
    near_src.clear();
    if (n < 2) return;
    double lo[3] = {pos[0], pos[1], pos[2]}, hi[3] = {pos[0], pos[1], pos[2]}, rmax = 0.0;
    for (int a = 0; a < n; a++) {
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], pos[3 * a + d]);
            hi[d] = std::max(hi[d], pos[3 * a + d]);
        }
        rmax = std::max(rmax, rad[a]);
    }
    // the near distance 1.5 (R_s + R_t) and the overlap distance R_s + R_t are
    // at most 3 R_max: cells of edge (2 + 1) R_max
    int key_bits = 0;
    const CellGrid g = cell_grid(n, lo, hi, rmax, 1.0, key_bits);
    const int64_t ncells = (int64_t)g.dim[0] * g.dim[1] * g.dim[2];
    Dev<int> idx(n), sorted(n), begin(ncells), end(ncells), overlap(n);
    Dev<unsigned> key(n), skey(n);
    Dev<int64_t> count(n + 1), off(n + 1);
    k_cell_keys<<<grid_of(n), 256, 0, s>>>(n, dpos, g, key.p, idx.p);
    size_t tmp_sort = 0, tmp_scan = 0;
    SPHSTK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_sort, key.p, skey.p, idx.p, sorted.p, n, 0, key_bits,
                                                  s));
    SPHSTK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_scan, count.p, off.p, n + 1, s));
    Dev<unsigned char> tmp(std::max(tmp_sort, tmp_scan));
    SPHSTK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_sort, key.p, skey.p, idx.p, sorted.p, n, 0, key_bits, s));
    const unsigned cg = (unsigned)((ncells + 255) / 256);
    k_fill_i32<<<cg, 256, 0, s>>>(ncells, 0, begin.p);
    k_fill_i32<<<cg, 256, 0, s>>>(ncells, 0, end.p);
    k_fill_i64<<<1, 64, 0, s>>>(1, 0, count.p + n);
    k_cell_bounds<<<grid_of(n), 256, 0, s>>>(n, skey.p, begin.p, end.p);
    k_near<false><<<grid_of64(n), 64, 0, s>>>(n, dpos, drad, dsusp, g, sorted.p, begin.p, end.p, count.p, overlap.p,
                                              nullptr, nullptr);
    SPHSTK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_scan, count.p, off.p, n + 1, s));
    SPHSTK_HIP(hipGetLastError());
    std::vector<int> ov(n);
    SPHSTK_HIP(hipMemcpyAsync(near_off.data(), off.p, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, s));
    SPHSTK_HIP(hipMemcpyAsync(ov.data(), overlap.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    SPHSTK_HIP(hipStreamSynchronize(s));
    for (int t = 0; t < n; t++)
        SPHSTK_REQUIRE(ov[t] < 0, "overlapping spheres " + std::to_string(t) + " and " + std::to_string(ov[t]) +
                                      " (centre distance < R_s + R_t)");
    const int64_t total = near_off[n];
    near_src.resize(total);
    if (total == 0) return;
    Dev<int> dsrc(total);
    k_near<true><<<grid_of64(n), 64, 0, s>>>(n, dpos, drad, dsusp, g, sorted.p, begin.p, end.p, nullptr, nullptr, off.p,
                                             dsrc.p);
    SPHSTK_HIP(hipGetLastError());
    SPHSTK_HIP(hipMemcpyAsync(near_src.data(), dsrc.p, sizeof(int) * total, hipMemcpyDeviceToHost, s));
    SPHSTK_HIP(hipStreamSynchronize(s));
}

int64_t collision_pairs_noalloc(int n, const double *dpos, const double *drc, const int *dgid, double colbuf,
                                int64_t capacity, int *dii, int *djj, double *dphi, double *dnorm,
                                hipStream_t s) {
    SPHSTK_REQUIRE(n >= 0, "n must be >= 0");
    SPHSTK_REQUIRE(n <= kNoAllocMaxSpheres, "n = " + std::to_string(n) + " exceeds the " +
                                                std::to_string(kNoAllocMaxSpheres) +
                                                " spheres of the collision-search scratch");
    if (n < 2) return 0;
    SPHSTK_REQUIRE(dpos && drc, "pos and radius_col must not be NULL");
    SPHSTK_REQUIRE(std::isfinite(colbuf) && colbuf > -1.0, "colbuf must be > -1");
    SPHSTK_REQUIRE(capacity >= 0, "capacity must be >= 0");
    const bool want = dii || djj || dphi || dnorm;
    SPHSTK_REQUIRE(!want || (dii && djj && dphi && dnorm), "the four output arrays must all be given, or none");
    unsigned char *base = scratch_on_current_device();
    double *bounds = (double *)(base + kOffBounds);
    unsigned *key = (unsigned *)(base + kOffKey), *skey = (unsigned *)(base + kOffSKey);
    int *idx = (int *)(base + kOffIdx), *sorted = (int *)(base + kOffSorted);
    int *begin = (int *)(base + kOffBegin), *end = (int *)(base + kOffEnd);
    int64_t *count = (int64_t *)(base + kOffCount), *off = (int64_t *)(base + kOffOff);
    void *tmp = base + kOffTmp;

    k_bounds<<<1, kBoundsThreads, 0, s>>>(n, dpos, drc, bounds);
    SPHSTK_HIP(hipGetLastError());
    double hb[8];
    SPHSTK_HIP(hipMemcpyAsync(hb, bounds, sizeof hb, hipMemcpyDeviceToHost, s));
    SPHSTK_HIP(hipStreamSynchronize(s));
    SPHSTK_REQUIRE(hb[7] == 0.0, "positions must be finite and collision radii > 0");
    int key_bits = 0;
    const CellGrid g = cell_grid(n, hb, hb + 3, hb[6], colbuf, key_bits);
    const int64_t ncells = (int64_t)g.dim[0] * g.dim[1] * g.dim[2];
    SPHSTK_REQUIRE(ncells <= 8 * kNA, "internal: cell grid larger than the scratch");

    size_t tmp_sort = 0, tmp_scan = 0;
    SPHSTK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_sort, key, skey, idx, sorted, n, 0, key_bits, s));
    SPHSTK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_scan, count, off, n + 1, s));
    SPHSTK_REQUIRE(std::max(tmp_sort, tmp_scan) <= kTmpBytes, "internal: sort workspace larger than the scratch");
    // everything read before it is written starts from a fill kernel (the
    // scratch keeps the previous call's contents)
    k_cell_keys<<<grid_of(n), 256, 0, s>>>(n, dpos, g, key, idx);
    SPHSTK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_sort, key, skey, idx, sorted, n, 0, key_bits, s));
    const unsigned cg = (unsigned)((ncells + 255) / 256);
    k_fill_i32<<<cg, 256, 0, s>>>(ncells, 0, begin);
    k_fill_i32<<<cg, 256, 0, s>>>(ncells, 0, end);
    k_fill_i64<<<1, 64, 0, s>>>(1, 0, count + n);
    k_cell_bounds<<<grid_of(n), 256, 0, s>>>(n, skey, begin, end);
    k_pairs<false><<<grid_of64(n), 64, 0, s>>>(n, dpos, drc, dgid, colbuf, g, sorted, begin, end, count, nullptr,
                                              nullptr, nullptr, nullptr, nullptr);
    SPHSTK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_scan, count, off, n + 1, s));
    SPHSTK_HIP(hipGetLastError());
    int64_t total = 0;
    SPHSTK_HIP(hipMemcpyAsync(&total, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SPHSTK_HIP(hipStreamSynchronize(s));
    if (!want || total > capacity || total == 0) return total;
    k_pairs<true><<<grid_of64(n), 64, 0, s>>>(n, dpos, drc, dgid, colbuf, g, sorted, begin, end, nullptr, off, dii,
                                             djj, dphi, dnorm);
    SPHSTK_HIP(hipGetLastError());
    SPHSTK_HIP(hipStreamSynchronize(s));
    return total;
}

} // namespace sphstk
