// HIP kernels of the many-sphere Stokes boundary-integral operator (gfx950).
//
// Far field   k_far_f64/_lvl direct Stokeslet / stresslet pair sums over the
//             k_far_f32      far pairs of every target (replaces the STKFMM
//                            evaluation, BIE/SphereSTKSHOperator.cpp:287-303,
//                            :365-446), over the fine grid or a level's grids.
//                            One 64-lane wave per group of 64*TPL targets;
//                            sources are wave-uniform and reach the VALU
//                            through scalar loads (no LDS staging).
// Centre sums k_far_centre   Faxen sphere-centre sums of the final velocities;
//             k_far_rigid    closed-form traction of distant rigid-body densities.
// SH matrices k_tiled_apply  grid values -> vector-SH coefficients per sphere
//                            (Shexp::calcSpectralCoeff, Sphere/Shexp.cpp:408-447)
//                            and resampling between grids, 16 spheres per tile.
// Expansion   k_expand       coefficients -> exterior solid-harmonic expansion
//                            records (the flow of SCTL StokesEvalSL/KL/KSelf).
// Near/self   k_sheval_src,  self (SphereSTKSHOperator.cpp:305-351, :375-456)
//             k_sheval       and near-pair (:575-695) SH evaluations.
// LOP         k_lop          rigid-body projector (:487-521).
// BLAS-1      k_seg_*        segmented (one segment per suspension) GMRES
//                            vector kernels.
#include "device_math.hpp"
#include "kernels.hpp"

namespace sphstk {

// ---------------------------------------------------------------------------
// helpers

__device__ __forceinline__ void qrot(const double *q, const double v[3], double o[3], bool inv) {
    const double w = q[0];
    const double s = inv ? -1.0 : 1.0;
    const double ux = s * q[1], uy = s * q[2], uz = s * q[3];
    const double tx = 2.0 * (uy * v[2] - uz * v[1]);
    const double ty = 2.0 * (uz * v[0] - ux * v[2]);
    const double tz = 2.0 * (ux * v[1] - uy * v[0]);
    o[0] = v[0] + w * tx + (uy * tz - uz * ty);
    o[1] = v[1] + w * ty + (uz * tx - ux * tz);
    o[2] = v[2] + w * tz + (ux * ty - uy * tx);
}

// FP64 reciprocal square root: hardware estimate refined by one Newton step,
// y1 = y0 + y0 (1 - r2 y0^2)/2.
__device__ __forceinline__ double rsqrt_f64(double r2) {
    const double y0 = __builtin_amdgcn_rsq(r2);
    const double h = r2 * y0;
    const double e = __builtin_fma(-h, y0, 1.0);
    return __builtin_fma(0.5 * y0, e, y0);
}

__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

__device__ __forceinline__ int dtri(int n, int m) { return n * (n + 1) / 2 + m; }

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    T r = 0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; i++) r += sh[i];
    return r;
}

// ---------------------------------------------------------------------------
// source / target packing

__global__ void k_pack_src_f64(int64_t n, const double *__restrict__ coords, const double *__restrict__ w,
                               const double *__restrict__ in, double *__restrict__ src) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double wi = w[i];
    double *o = src + 8 * i;
    o[0] = coords[3 * i];
    o[1] = coords[3 * i + 1];
    o[2] = coords[3 * i + 2];
    o[3] = wi * in[3 * i];
    o[4] = wi * in[3 * i + 1];
    o[5] = wi * in[3 * i + 2];
    o[6] = 0.0;
    o[7] = 0.0;
}

__global__ void k_pack_src_f32(int64_t n, const double *__restrict__ coords, const double *__restrict__ w,
                               const double *__restrict__ in, float *__restrict__ src) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double wi = w[i];
    float *o = src + 8 * i;
    o[0] = (float)coords[3 * i];
    o[1] = (float)coords[3 * i + 1];
    o[2] = (float)coords[3 * i + 2];
    o[3] = (float)(wi * in[3 * i]);
    o[4] = (float)(wi * in[3 * i + 1]);
    o[5] = (float)(wi * in[3 * i + 2]);
    o[6] = 0.f;
    o[7] = 0.f;
}

__global__ void k_pack_trg_f32(int64_t n, const double *__restrict__ coords, const double *__restrict__ norms,
                               float *__restrict__ trg) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *o = trg + 8 * i;
    for (int c = 0; c < 3; c++) {
        o[c] = (float)coords[3 * i + c];
        o[3 + c] = (float)norms[3 * i + c];
    }
    o[6] = 0.f;
    o[7] = 0.f;
}

void launch_pack_src_f64(int64_t n, const double *coords, const double *w, const double *in, double *src,
                         hipStream_t s) {
    if (n > 0) k_pack_src_f64<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(n, coords, w, in, src);
}
void launch_pack_src_f32(int64_t n, const double *coords, const double *w, const double *in, float *src,
                         hipStream_t s) {
    if (n > 0) k_pack_src_f32<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(n, coords, w, in, src);
}
void launch_pack_trg_f32(int64_t n, const double *coords, const double *norms, float *trg, hipStream_t s) {
    if (n > 0) k_pack_trg_f32<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(n, coords, norms, trg);
}

// ---------------------------------------------------------------------------
// far field, FP64.  KIND 0: traction  t = -(3/4pi) sum r (r.f)(r.n_t)/r^5
//                   KIND 1: single layer u = (1/8pi) sum [f/r + r (r.f)/r^3]
// (the -(3/4pi) and 1/(8pi) prefactors are folded into `coef`).

// One source against TPL targets.  MASKED: per-lane inclusion flags select the
// reciprocal distance (a select, so an excluded coincident point cannot turn
// into inf*0).
template <int KIND, int TPL, bool MASKED, typename R>
__device__ __forceinline__ void far_pair(const R sx, const R sy, const R sz, const R fx, const R fy, const R fz,
                                         const R *x, const R *y, const R *z, const R *nx, const R *ny, const R *nz,
                                         const bool *incl, R *ax, R *ay, R *az) {
#pragma unroll
    for (int t = 0; t < TPL; t++) {
        const R dx = x[t] - sx, dy = y[t] - sy, dz = z[t] - sz;
        const R r2 = fmad(dx, dx, fmad(dy, dy, dz * dz));
        const R df = fmad(dx, fx, fmad(dy, fy, dz * fz));
        if constexpr (KIND == 0 && sizeof(R) == 8) {
            R w = trac_w_f64(r2);
            if (MASKED) w = incl[t] ? w : R(0);
            const R dn = fmad(dx, nx[t], fmad(dy, ny[t], dz * nz[t]));
            const R q = (df * dn) * w;
            ax[t] = fmad(q, dx, ax[t]);
            ay[t] = fmad(q, dy, ay[t]);
            az[t] = fmad(q, dz, az[t]);
            continue;
        }
        R ri;
        if constexpr (sizeof(R) == 8)
            ri = rsqrt_f64(r2);
        else
            ri = __builtin_amdgcn_rsqf(r2);
        if (MASKED) ri = incl[t] ? ri : R(0);
        if (KIND == 0) {
            const R dn = fmad(dx, nx[t], fmad(dy, ny[t], dz * nz[t]));
            const R ri2 = ri * ri;
            const R q = (df * dn) * (ri2 * ri2 * ri);
            ax[t] = fmad(q, dx, ax[t]);
            ay[t] = fmad(q, dy, ay[t]);
            az[t] = fmad(q, dz, az[t]);
        } else {
            const R q = df * (ri * ri * ri);
            ax[t] = fmad(q, dx, fmad(fx, ri, ax[t]));
            ay[t] = fmad(q, dy, fmad(fy, ri, ay[t]));
            az[t] = fmad(q, dz, fmad(fz, ri, az[t]));
        }
    }
}

// Sources [b, e) of an 8-word record array, four at a time so that the
// wave-uniform (scalar) loads of a batch are issued before its arithmetic.
// FP32 records: the loads of the next batch are issued before the arithmetic
// of the current one, so their latency (the sources stream from beyond the
// L2) hides behind it (16384 spheres p=12: 12% faster).  FP64 records keep
// one batch in flight: two would not fit the wave's scalar registers.
template <int KIND, int TPL, bool MASKED, typename R>
__device__ __forceinline__ void far_range(const R *__restrict__ src, int64_t b, int64_t e, const R *x, const R *y,
                                          const R *z, const R *nx, const R *ny, const R *nz, const bool *incl, R *ax,
                                          R *ay, R *az) {
    int64_t s = b;
    const int64_t e4 = b + ((e - b) & ~int64_t(3));
    if constexpr (sizeof(R) == 4) {
        if (s != e4) {
            R v[4][6];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int c = 0; c < 6; c++) v[u][c] = src[8 * (s + u) + c];
            for (; s + 4 != e4; s += 4) {
                R w[4][6];
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int c = 0; c < 6; c++) w[u][c] = src[8 * (s + 4 + u) + c];
#pragma unroll
                for (int u = 0; u < 4; u++)
                    far_pair<KIND, TPL, MASKED>(v[u][0], v[u][1], v[u][2], v[u][3], v[u][4], v[u][5], x, y, z, nx,
                                                ny, nz, incl, ax, ay, az);
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int c = 0; c < 6; c++) v[u][c] = w[u][c];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                far_pair<KIND, TPL, MASKED>(v[u][0], v[u][1], v[u][2], v[u][3], v[u][4], v[u][5], x, y, z, nx, ny,
                                            nz, incl, ax, ay, az);
            s = e4;
        }
    }
    for (; s != e4; s += 4) {
        R v[4][6];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int c = 0; c < 6; c++) v[u][c] = src[8 * (s + u) + c];
#pragma unroll
        for (int u = 0; u < 4; u++)
            far_pair<KIND, TPL, MASKED>(v[u][0], v[u][1], v[u][2], v[u][3], v[u][4], v[u][5], x, y, z, nx, ny, nz,
                                        incl, ax, ay, az);
    }
    for (; s != e; s++) {
        const R *sp = src + 8 * s;
        far_pair<KIND, TPL, MASKED>(sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], x, y, z, nx, ny, nz, incl, ax, ay, az);
    }
}

// accumulate != 0: out += coef*acc.  accumulate == 0: out[slot*slot_stride + ...] = coef*acc
// (one partial-sum slot per source slice, reduced afterwards in slot order).
template <int TPL>
__device__ __forceinline__ void store_far(const FarGroup &G, int lane, const double *ax, const double *ay,
                                          const double *az, double *__restrict__ out, double coef, int accumulate,
                                          int64_t slot_stride) {
    if (!accumulate) out += G.slot * slot_stride;
#pragma unroll
    for (int t = 0; t < TPL; t++) {
        const int idx = t * 64 + lane;
        if (idx < G.tcount) {
            double *o = out + 3 * (G.tstart + idx);
            if (accumulate) {
                o[0] += coef * ax[t];
                o[1] += coef * ay[t];
                o[2] += coef * az[t];
            } else {
                o[0] = coef * ax[t];
                o[1] = coef * ay[t];
                o[2] = coef * az[t];
            }
        }
    }
}

template <int KIND, int TPL>
__device__ __forceinline__ void far_f64_body(const double *__restrict__ src, const double *__restrict__ coords,
                                             const double *__restrict__ norms, const int *__restrict__ gorder,
                                             const FarGroup *__restrict__ groups, const FarInterval *__restrict__ iv,
                                             const int *__restrict__ active, double *__restrict__ out, double coef,
                                             int accumulate, int64_t slot_stride) {
    const int gi = gorder[blockIdx.x];
    const FarGroup G = groups[gi];
    if (active && !active[G.susp]) return;
    const int lane = threadIdx.x;
    double x[TPL], y[TPL], z[TPL], nx[TPL], ny[TPL], nz[TPL], ax[TPL], ay[TPL], az[TPL];
#pragma unroll
    for (int t = 0; t < TPL; t++) {
        const int idx = t * 64 + lane;
        const int64_t tid = G.tstart + (idx < G.tcount ? idx : 0);
        x[t] = coords[3 * tid];
        y[t] = coords[3 * tid + 1];
        z[t] = coords[3 * tid + 2];
        nx[t] = norms[3 * tid];
        ny[t] = norms[3 * tid + 1];
        nz[t] = norms[3 * tid + 2];
        ax[t] = ay[t] = az[t] = 0.0;
    }
    for (int64_t k = G.iv_begin; k < G.iv_end; k++) {
        const int64_t b = iv[k].begin, e = iv[k].end;
        bool full = true, incl[TPL];
#pragma unroll
        for (int t = 0; t < TPL; t++) {
            const uint64_t m = iv[k].mask[t];
            full = full && (m == ~0ull);
            incl[t] = ((m >> lane) & 1ull) != 0;
        }
        if (full)
            far_range<KIND, TPL, false>(src, b, e, x, y, z, nx, ny, nz, incl, ax, ay, az);
        else
            far_range<KIND, TPL, true>(src, b, e, x, y, z, nx, ny, nz, incl, ax, ay, az);
    }
    store_far<TPL>(G, lane, ax, ay, az, out, KIND == 0 ? kTracW * coef : coef, accumulate, slot_stride);
}

// The direct FP64 far field (the whole fine grid) and the same sum over the
// work lists of a multi-level far field, under two kernel names so that
// profiles keep them apart.
template <int KIND, int TPL>
__global__ __launch_bounds__(64) void k_far_f64(const double *__restrict__ src, const double *__restrict__ coords,
                                                const double *__restrict__ norms, const int *__restrict__ gorder,
                                                const FarGroup *__restrict__ groups,
                                                const FarInterval *__restrict__ iv, const int *__restrict__ active,
                                                double *__restrict__ out, double coef, int accumulate,
                                                int64_t slot_stride) {
    far_f64_body<KIND, TPL>(src, coords, norms, gorder, groups, iv, active, out, coef, accumulate, slot_stride);
}
template <int KIND, int TPL>
__global__ __launch_bounds__(64) void k_far_f64_lvl(const double *__restrict__ src, const double *__restrict__ coords,
                                                    const double *__restrict__ norms, const int *__restrict__ gorder,
                                                    const FarGroup *__restrict__ groups,
                                                    const FarInterval *__restrict__ iv, const int *__restrict__ active,
                                                    double *__restrict__ out, double coef, int accumulate,
                                                    int64_t slot_stride) {
    far_f64_body<KIND, TPL>(src, coords, norms, gorder, groups, iv, active, out, coef, accumulate, slot_stride);
}

// far field, FP32 pair arithmetic with two-level accumulation: FP32 partial sums
// over blocks of at most kF32Block sources, flushed into FP64 accumulators.
constexpr int kF32Block = 256;

template <int KIND, int TPL>
__global__ __launch_bounds__(64) void k_far_f32(const float *__restrict__ src, const float *__restrict__ trg,
                                                const int *__restrict__ gorder, const FarGroup *__restrict__ groups,
                                                const FarInterval *__restrict__ iv, const int *__restrict__ active,
                                                double *__restrict__ out, double coef, int accumulate,
                                                int64_t slot_stride) {
    const int gi = gorder[blockIdx.x];
    const FarGroup G = groups[gi];
    if (active && !active[G.susp]) return;
    const int lane = threadIdx.x;
    float x[TPL], y[TPL], z[TPL], nx[TPL], ny[TPL], nz[TPL];
    double ax[TPL], ay[TPL], az[TPL];
#pragma unroll
    for (int t = 0; t < TPL; t++) {
        const int idx = t * 64 + lane;
        const int64_t tid = G.tstart + (idx < G.tcount ? idx : 0);
        const float *tp = trg + 8 * tid;
        x[t] = tp[0];
        y[t] = tp[1];
        z[t] = tp[2];
        nx[t] = tp[3];
        ny[t] = tp[4];
        nz[t] = tp[5];
        ax[t] = ay[t] = az[t] = 0.0;
    }
    for (int64_t k = G.iv_begin; k < G.iv_end; k++) {
        const int64_t b = iv[k].begin, e = iv[k].end;
        bool full = true, incl[TPL];
#pragma unroll
        for (int t = 0; t < TPL; t++) {
            const uint64_t m = iv[k].mask[t];
            full = full && (m == ~0ull);
            incl[t] = ((m >> lane) & 1ull) != 0;
        }
        for (int64_t s0 = b; s0 < e; s0 += kF32Block) {
            const int64_t s1 = (s0 + kF32Block < e) ? s0 + kF32Block : e;
            float px[TPL], py[TPL], pz[TPL];
#pragma unroll
            for (int t = 0; t < TPL; t++) px[t] = py[t] = pz[t] = 0.f;
            if (full)
                far_range<KIND, TPL, false>(src, s0, s1, x, y, z, nx, ny, nz, incl, px, py, pz);
            else
                far_range<KIND, TPL, true>(src, s0, s1, x, y, z, nx, ny, nz, incl, px, py, pz);
#pragma unroll
            for (int t = 0; t < TPL; t++) {
                ax[t] += (double)px[t];
                ay[t] += (double)py[t];
                az[t] += (double)pz[t];
            }
        }
    }
    store_far<TPL>(G, lane, ax, ay, az, out, coef, accumulate, slot_stride);
}

struct FarLaunch {
    const void *src;
    const double *tcoords, *tnorms;
    const float *trg32;
    int64_t ntrg;
    bool levels;
};

template <int KIND, int TPL>
static void far_dispatch_prec(int prec, const FarLaunch &a, int ngroups, const int *go, const FarGroup *gr,
                              const FarInterval *iv, const int *active, double *out, double coef, int acc,
                              hipStream_t s) {
    const int64_t stride = 3 * a.ntrg;
    if (prec == 0 && a.levels)
        k_far_f64_lvl<KIND, TPL><<<dim3(ngroups), dim3(64), 0, s>>>((const double *)a.src, a.tcoords, a.tnorms, go,
                                                                   gr, iv, active, out, coef, acc, stride);
    else if (prec == 0)
        k_far_f64<KIND, TPL><<<dim3(ngroups), dim3(64), 0, s>>>((const double *)a.src, a.tcoords, a.tnorms, go, gr,
                                                               iv, active, out, coef, acc, stride);
    else
        k_far_f32<KIND, TPL><<<dim3(ngroups), dim3(64), 0, s>>>((const float *)a.src, a.trg32, go, gr, iv, active,
                                                               out, coef, acc, stride);
}

template <int KIND>
static void far_dispatch_tpl(int tpl, int prec, const FarLaunch &a, int ngroups, const int *go, const FarGroup *gr,
                             const FarInterval *iv, const int *active, double *out, double coef, int acc,
                             hipStream_t s) {
    switch (tpl) {
    case 1: far_dispatch_prec<KIND, 1>(prec, a, ngroups, go, gr, iv, active, out, coef, acc, s); break;
    case 2: far_dispatch_prec<KIND, 2>(prec, a, ngroups, go, gr, iv, active, out, coef, acc, s); break;
    case 4: far_dispatch_prec<KIND, 4>(prec, a, ngroups, go, gr, iv, active, out, coef, acc, s); break;
    default: throw Error(-1, "unsupported targets-per-lane");
    }
}

void launch_far(int kind, int prec, int tpl, const void *src, const double *tcoords, const double *tnorms,
                const float *trg32, int64_t ntrg, int ngroups, const int *go, const FarGroup *gr,
                const FarInterval *iv, const int *active, double *out, double coef, int acc, hipStream_t s,
                bool levels) {
    if (ngroups <= 0) return;
    const FarLaunch a{src, tcoords, tnorms, trg32, ntrg, levels};
    if (kind == 0)
        far_dispatch_tpl<0>(tpl, prec, a, ngroups, go, gr, iv, active, out, coef, acc, s);
    else
        far_dispatch_tpl<1>(tpl, prec, a, ngroups, go, gr, iv, active, out, coef, acc, s);
}

// ---------------------------------------------------------------------------
// Equivalent point forces of the far-field cells: g_c = rho_c P u_c for every
// cell c (P: 3K x 3Kc, column-major; u_c: the cell's check velocities), written
// as the force words (3..5) of the cell's FP32 source records, whose positions
// were set at set-up.  kEqCells cells per workgroup share each load of P; their
// check velocities are staged in LDS; thread i owns rows i, i + 256, ....
constexpr int kEqCells = 4;
constexpr int kEqThreads = 256;
__global__ __launch_bounds__(kEqThreads) void k_equiv_forces(int ncell, int K, int Kc, const double *__restrict__ P,
                                                             const double *__restrict__ u,
                                                             const double *__restrict__ rho,
                                                             const int *__restrict__ susp,
                                                             const int *__restrict__ active, float *__restrict__ eq) {
    extern __shared__ double su[]; // kEqCells x 3Kc
    const int c0 = blockIdx.x * kEqCells;
    const int nc = min(kEqCells, ncell - c0);
    const int m = 3 * Kc, n = 3 * K;
    for (int k = 0; k < nc; k++)
        for (int j = threadIdx.x; j < m; j += blockDim.x) su[k * m + j] = u[(int64_t)(c0 + k) * m + j];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double acc[kEqCells] = {};
        for (int j = 0; j < m; j++) {
            const double pij = P[(int64_t)j * n + i];
#pragma unroll
            for (int k = 0; k < kEqCells; k++) acc[k] = fmad(pij, su[k * m + j], acc[k]);
        }
        for (int k = 0; k < nc; k++) {
            const int c = c0 + k;
            if (active && !active[susp[c]]) continue;
            eq[8 * ((int64_t)c * K + i / 3) + 3 + i % 3] = (float)(rho[c] * acc[k]);
        }
    }
}

void launch_equiv_forces(int ncell, int K, int Kc, const double *P, const double *u, const double *rho,
                         const int *susp, const int *active, float *eq, hipStream_t s) {
    if (ncell <= 0) return;
    const size_t lds = sizeof(double) * kEqCells * 3 * Kc;
    k_equiv_forces<<<dim3((ncell + kEqCells - 1) / kEqCells), dim3(kEqThreads), lds, s>>>(ncell, K, Kc, P, u, rho,
                                                                                          susp, active, eq);
}

// ---------------------------------------------------------------------------
// Per-sphere matrices applied to body-frame values, one tile of up to kTileSph
// = 16 spheres of the same order per workgroup, as a GEMM on the matrix cores
// (the tables are up to 12.7 MB at p = 16; each element is read once per tile):
//   analysis:  coef[S] = A_p (q_S^-1 in_S)                (A_p: 6Q x 3n)
//   resample:  out_S (+)= q_S M_p (q_S^-1 in_S)            (M_p: 3 nout x 3 nin)
// A workgroup owns kRowBlock = 192 matrix rows (64 points): each of its four
// waves three blocks of 16 rows, accumulated by v_mfma_f64_16x16x4f64 over the
// columns (16 rows x 4 columns of the matrix, read from global memory, times
// 4 columns x 16 spheres of the rotated input).  The input is staged through
// LDS in chunks of kChunkPts points, stored [column][sphere].  Operand layout
// of the instruction: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// D[i][j] in lane j + 16 (i % 4), element i / 4.
// The FP32 inner operator's level resampling runs the same GEMM with FP32
// tables, FP32 staging and v_mfma_f32_16x16x4f32 (same A and B layout, D[i][j]
// in lane j + 16 (i / 4), element i % 4; FP32 accumulation: about 1e-7
// relative against the FP64 resampling, far below that operator's own 4e-6).
constexpr int kRowBlock = 192;
constexpr int kWaveRowBlocks = 3;
constexpr int kChunkPts = 64;
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <typename T>
struct MatTabsT {
    const T *M[kMaxOrder + 1] = {};
    int rows[kMaxOrder + 1] = {};
    int nin[kMaxOrder + 1] = {};
};
using MatTabs = MatTabsT<double>;

__device__ __forceinline__ dbl4 mfma16x16x4(double a, double b, dbl4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ flt4 mfma16x16x4(float a, float b, flt4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <bool RESAMPLE, typename T>
__global__ __launch_bounds__(256) void k_tiled_apply(DevGeom g, MatTabsT<T> tab, const int64_t *__restrict__ in_off,
                                                     const int64_t *__restrict__ out_off,
                                                     const double *__restrict__ in, double *__restrict__ out,
                                                     int accumulate, const int *__restrict__ active) {
    typedef T T4 __attribute__((ext_vector_type(4)));
    __shared__ T V[3 * kChunkPts * kTileSph];
    __shared__ T P[kRowBlock * kTileSph];
    const SphTile Ti = g.tiles[blockIdx.x];
    const int p = Ti.p, rows = tab.rows[p], nin = tab.nin[p];
    const int r0 = blockIdx.y * kRowBlock;
    if (r0 >= rows) return;
    const T *__restrict__ M = tab.M[p];
    const int *sph = g.tile_sph + Ti.first;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int li = lane & 15, kq = lane >> 4;
    const int rw = r0 + wv * 16 * kWaveRowBlocks; // this wave's first row
    T4 acc[kWaveRowBlocks];
#pragma unroll
    for (int rb = 0; rb < kWaveRowBlocks; rb++) acc[rb] = T4{T(0), T(0), T(0), T(0)};
    for (int k0 = 0; k0 < nin; k0 += kChunkPts) {
        const int kc = min(kChunkPts, nin - k0);
        __syncthreads();
        for (int e = tid; e < kChunkPts * kTileSph; e += blockDim.x) {
            const int t = e / kChunkPts, i = e % kChunkPts;
            double o[3] = {0.0, 0.0, 0.0};
            if (t < Ti.count && i < kc) {
                const int S = sph[t];
                const double *v = in + 3 * (in_off[S] + k0 + i);
                const double vv[3] = {v[0], v[1], v[2]};
                qrot(g.quat + 4 * S, vv, o, true);
            }
            V[(3 * i) * kTileSph + t] = (T)o[0];
            V[(3 * i + 1) * kTileSph + t] = (T)o[1];
            V[(3 * i + 2) * kTileSph + t] = (T)o[2];
        }
        __syncthreads();
        const int nc = 3 * kc;
        const T *Mc = M + (int64_t)(3 * k0) * rows;
        for (int c = 0; c < nc; c += 4) {
            const int col = c + kq;
            const bool cin = col < nc;
            const T b = cin ? V[col * kTileSph + li] : T(0);
#pragma unroll
            for (int rb = 0; rb < kWaveRowBlocks; rb++) {
                const int r = rw + 16 * rb + li;
                const T a = (cin && r < rows) ? Mc[(int64_t)col * rows + r] : T(0);
                acc[rb] = mfma16x16x4(a, b, acc[rb]);
            }
        }
    }
    // row of element q of this lane's accumulator within its 16-row block:
    // FP32 instruction 4 (lane / 16) + q, FP64 instruction (lane / 16) + 4 q
#pragma unroll
    for (int rb = 0; rb < kWaveRowBlocks; rb++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = sizeof(T) == 8 ? kq + 4 * q : 4 * kq + q;
            P[(wv * 16 * kWaveRowBlocks + 16 * rb + i) * kTileSph + li] = acc[rb][q];
        }
    __syncthreads();
    if (!RESAMPLE) {
        const int nr = min(rows, r0 + kRowBlock) - r0;
        for (int e = tid; e < nr * Ti.count; e += blockDim.x) {
            const int t = e / nr, i = e % nr;
            const int S = sph[t];
            if (active && !active[g.sph_susp[S]]) continue;
            out[g.coef_off[S] + r0 + i] = P[i * kTileSph + t];
        }
        return;
    }
    // resample: rotate each output point back to the lab frame
    const int npt = (min(rows, r0 + kRowBlock) - r0) / 3;
    for (int e = tid; e < npt * Ti.count; e += blockDim.x) {
        const int t = e / npt, i = e % npt;
        const int S = sph[t];
        if (active && !active[g.sph_susp[S]]) continue;
        const double vb[3] = {(double)P[(3 * i) * kTileSph + t], (double)P[(3 * i + 1) * kTileSph + t],
                              (double)P[(3 * i + 2) * kTileSph + t]};
        double o[3];
        qrot(g.quat + 4 * S, vb, o, false);
        double *dst = out + 3 * (out_off[S] + r0 / 3 + i);
        if (accumulate) {
            dst[0] += o[0];
            dst[1] += o[1];
            dst[2] += o[2];
        } else {
            dst[0] = o[0];
            dst[1] = o[1];
            dst[2] = o[2];
        }
    }
}

void launch_resample(const DevGeom &g, const ResampleTabs &tab, const int64_t *in_off, const int64_t *out_off,
                     const double *in, double *out, int accumulate, const int *active, hipStream_t s,
                     int max_points) {
    if (!g.nsph || !g.ntiles) return;
    const int rb = (3 * max_points + kRowBlock - 1) / kRowBlock;
    if (tab.fp32) {
        MatTabsT<float> mt;
        for (int p = 0; p <= kMaxOrder; p++) {
            mt.M[p] = tab.M32[p];
            mt.rows[p] = 3 * tab.nout[p];
            mt.nin[p] = tab.nin[p];
        }
        k_tiled_apply<true, float><<<dim3(g.ntiles, rb), dim3(256), 0, s>>>(g, mt, in_off, out_off, in, out,
                                                                           accumulate, active);
        return;
    }
    MatTabs mt;
    for (int p = 0; p <= kMaxOrder; p++) {
        mt.M[p] = tab.M[p];
        mt.rows[p] = 3 * tab.nout[p];
        mt.nin[p] = tab.nin[p];
    }
    k_tiled_apply<true, double><<<dim3(g.ntiles, rb), dim3(256), 0, s>>>(g, mt, in_off, out_off, in, out, accumulate,
                                                                        active);
}
void launch_analysis(const DevGeom &g, const double *in, double *coef, const int *active, hipStream_t s,
                     int max_order) {
    if (!g.nsph || !g.ntiles) return;
    MatTabs mt;
    for (int p = 0; p <= kMaxOrder; p++) {
        mt.M[p] = g.anal[p];
        mt.rows[p] = 3 * (p + 1) * (p + 2);
        mt.nin[p] = (p + 1) * (2 * p + 2);
    }
    const int rb = (3 * (max_order + 1) * (max_order + 2) + kRowBlock - 1) / kRowBlock;
    k_tiled_apply<false, double><<<dim3(g.ntiles, rb), dim3(256), 0, s>>>(g, mt, g.pt_off, nullptr, in, coef, 0,
                                                                         active);
}

// ---------------------------------------------------------------------------
// Expansion records.  Harmonic functions H = sum_{n,m>=0} eps_m Re(h_nm D_n^m).
// Cartesian derivatives raise the degree by one:
//   z: h'_{n+1,k} = h_{n,k}
//   x: h'_{n+1,0} = -h_{n,1},    h'_{n+1,k} = (h_{n,k-1} - h_{n,k+1})/2
//   y: h'_{n+1,0} = -i h_{n,1},  h'_{n+1,k} = -i (h_{n,k-1} + h_{n,k+1})/2
// (m=0 coefficients are kept real).  Exterior single-layer flow of a unit-sphere
// density with vector-SH coefficients (F_r, F_s, F_t), per degree n>=1:
//   a = n(F_r+(n+3)F_s)/(2(2n+1)(2n+3)), b = n(F_r+(n+1)F_s)/(2n+1), alpha = F_t/(2n+1)
//   H1 = a, H2 = A_n b, H3 = B_n b, H4 = alpha, P = b,
//   A_n = (2-n)/(2n(2n-1)), B_n = (n+1)/(n(2n-1))
//   u = grad H1 + r^2 grad H2 + x H3 + x cross grad H4,  pressure P.
// Self synthesis of the projected density at r=1:
//   P sigma = Qr x + grad Qs + x cross grad Qt,  Qr = F_r + (n+1) F_s, Qs = F_s, Qt = F_t.
// Records store, per (n,m) with n <= p+2, the K arrays' coefficients interleaved
// as (eps_m Re c, -eps_m Im c) so that eps_m Re(c D) = c0 D.re + c1 D.im.

__device__ __forceinline__ double2 hget(const double2 *h, int P, int n, int k) {
    if (k < 0 || k > n || n > P || n < 0) return make_double2(0.0, 0.0);
    return h[dtri(n, k)];
}

// coefficient (n1, k) of the derivative along dir of the array h (degree <= P)
__device__ __forceinline__ double2 dcoef(const double2 *h, int P, int n1, int k, int dir) {
    if (n1 < 1) return make_double2(0.0, 0.0);
    const int n = n1 - 1;
    double2 r;
    if (dir == 2) {
        r = hget(h, P, n, k);
    } else if (dir == 0) {
        if (k == 0) {
            const double2 a = hget(h, P, n, 1);
            r = make_double2(-a.x, -a.y);
        } else {
            const double2 a = hget(h, P, n, k - 1), b = hget(h, P, n, k + 1);
            r = make_double2(0.5 * (a.x - b.x), 0.5 * (a.y - b.y));
        }
    } else {
        if (k == 0) {
            const double2 a = hget(h, P, n, 1); // -i a
            r = make_double2(a.y, -a.x);
        } else {
            const double2 a = hget(h, P, n, k - 1), b = hget(h, P, n, k + 1);
            const double2 s = make_double2(a.x + b.x, a.y + b.y); // -i s / 2
            r = make_double2(0.5 * s.y, -0.5 * s.x);
        }
    }
    if (k == 0) r.y = 0.0;
    return r;
}

__global__ __launch_bounds__(256) void k_expand(DevGeom g, int mode, const double *__restrict__ coef,
                                                double *__restrict__ rec, const int *__restrict__ active) {
    extern __shared__ double2 lds[];
    const int S = blockIdx.x;
    if (active && !active[g.sph_susp[S]]) return;
    const int p = g.order[S];
    const int Q = (p + 1) * (p + 2) / 2, P1 = p + 1, Q1 = (P1 + 1) * (P1 + 2) / 2, P2 = p + 2,
              Q2 = (P2 + 1) * (P2 + 2) / 2;
    const int nb = (mode == 0) ? 4 : 8;
    double2 *base = lds;           // nb arrays of Q
    double2 *grad = lds + nb * Q;  // (mode 0: H1,H2,H4; mode 1: H1,H2,H3,H4,Qs,Qt) x 3 arrays of Q1
    const double2 *F = reinterpret_cast<const double2 *>(coef + g.coef_off[S]);
    // base arrays
    for (int k = threadIdx.x; k < Q; k += blockDim.x) {
        int n = 0;
        while ((n + 1) * (n + 2) / 2 <= k) n++;
        const double2 fr = F[k], fs = F[Q + k], ft = F[2 * Q + k];
        double2 h1 = make_double2(0, 0), h2 = h1, h3 = h1, h4 = h1, hp = h1;
        if (n >= 1) {
            const double nn = n;
            const double An = (2.0 - nn) / (2.0 * nn * (2.0 * nn - 1.0));
            const double Bn = (nn + 1.0) / (nn * (2.0 * nn - 1.0));
            const double ca = nn / (2.0 * (2.0 * nn + 1.0) * (2.0 * nn + 3.0));
            const double cb = nn / (2.0 * nn + 1.0);
            const double2 a = make_double2(ca * (fr.x + (nn + 3.0) * fs.x), ca * (fr.y + (nn + 3.0) * fs.y));
            const double2 b = make_double2(cb * (fr.x + (nn + 1.0) * fs.x), cb * (fr.y + (nn + 1.0) * fs.y));
            h1 = a;
            h2 = make_double2(An * b.x, An * b.y);
            h3 = make_double2(Bn * b.x, Bn * b.y);
            h4 = make_double2(ft.x / (2.0 * nn + 1.0), ft.y / (2.0 * nn + 1.0));
            hp = b;
        }
        base[k] = h1;
        base[Q + k] = h2;
        base[2 * Q + k] = h3;
        base[3 * Q + k] = h4;
        if (mode == 1) {
            base[4 * Q + k] = hp;
            base[5 * Q + k] = make_double2(fr.x + (n + 1.0) * fs.x, fr.y + (n + 1.0) * fs.y);
            base[6 * Q + k] = fs;
            base[7 * Q + k] = ft;
        }
    }
    __syncthreads();
    // gradient arrays
    const int ng = (mode == 0) ? 3 : 6;
    const int gsrc0[3] = {0, 1, 3};
    const int gsrc1[6] = {0, 1, 2, 3, 6, 7};
    for (int idx = threadIdx.x; idx < ng * 3 * Q1; idx += blockDim.x) {
        const int a = idx / (3 * Q1), rem = idx % (3 * Q1), dir = rem / Q1, k1 = rem % Q1;
        int n1 = 0;
        while ((n1 + 1) * (n1 + 2) / 2 <= k1) n1++;
        const int kk = k1 - n1 * (n1 + 1) / 2;
        const int f = (mode == 0) ? gsrc0[a] : gsrc1[a];
        grad[(a * 3 + dir) * Q1 + k1] = dcoef(base + f * Q, p, n1, kk, dir);
    }
    __syncthreads();
    // record
    const int K = (mode == 0) ? kRecSL : kRecTrac;
    double *R = rec + g.rec_off[S];
    const int hd0[6] = {0, 0, 0, 1, 1, 2}, hd1[6] = {0, 1, 2, 1, 2, 2};
    for (int idx = threadIdx.x; idx < K * Q2; idx += blockDim.x) {
        const int k2 = idx / K, a = idx % K;
        int n2 = 0;
        while ((n2 + 1) * (n2 + 2) / 2 <= k2) n2++;
        const int m = k2 - n2 * (n2 + 1) / 2;
        double2 c = make_double2(0, 0);
        if (mode == 0) {
            // 0-2 grad H1, 3-5 grad H2, 6 H3, 7-9 grad H4
            if (a < 6) {
                const int gi = a / 3, d = a % 3;
                c = hget(grad + (gi * 3 + d) * Q1, P1, n2, m);
            } else if (a == 6) {
                c = hget(base + 2 * Q, p, n2, m);
            } else {
                c = hget(grad + (2 * 3 + (a - 7)) * Q1, P1, n2, m);
            }
        } else {
            // 0-5 Hess H1, 6-8 grad H2, 9-14 Hess H2, 15 H3, 16-18 grad H3, 19-24 Hess H4, 25 P,
            // 26 Qr, 27-29 grad Qs, 30-32 grad Qt
            if (a < 6) {
                c = dcoef(grad + (0 * 3 + hd0[a]) * Q1, P1, n2, m, hd1[a]);
            } else if (a < 9) {
                c = hget(grad + (1 * 3 + (a - 6)) * Q1, P1, n2, m);
            } else if (a < 15) {
                c = dcoef(grad + (1 * 3 + hd0[a - 9]) * Q1, P1, n2, m, hd1[a - 9]);
            } else if (a == 15) {
                c = hget(base + 2 * Q, p, n2, m);
            } else if (a < 19) {
                c = hget(grad + (2 * 3 + (a - 16)) * Q1, P1, n2, m);
            } else if (a < 25) {
                c = dcoef(grad + (3 * 3 + hd0[a - 19]) * Q1, P1, n2, m, hd1[a - 19]);
            } else if (a == 25) {
                c = hget(base + 4 * Q, p, n2, m);
            } else if (a == 26) {
                c = hget(base + 5 * Q, p, n2, m);
            } else if (a < 30) {
                c = hget(grad + (4 * 3 + (a - 27)) * Q1, P1, n2, m);
            } else {
                c = hget(grad + (5 * 3 + (a - 30)) * Q1, P1, n2, m);
            }
        }
        const double eps = m ? 2.0 : 1.0;
        R[2 * (k2 * K + a)] = eps * c.x;
        R[2 * (k2 * K + a) + 1] = -eps * c.y;
    }
}

// Per-degree blocks on the SH coefficients of each sphere, in place:
// (F_r, F_s) <- M_n (F_r, F_s) (real 2x2, tab[5n..5n+3] row-major),
// F_t <- tab[5n+4] F_t.  Used for the inverse of the single-sphere operator.
__global__ __launch_bounds__(256) void k_degree_blocks(DevGeom g, double *__restrict__ coef,
                                                       const double *__restrict__ tab,
                                                       const int *__restrict__ active) {
    const int S = blockIdx.x;
    if (active && !active[g.sph_susp[S]]) return;
    const int p = g.order[S];
    const int Q = (p + 1) * (p + 2) / 2;
    double2 *F = reinterpret_cast<double2 *>(coef + g.coef_off[S]);
    for (int k = threadIdx.x; k < Q; k += blockDim.x) {
        int n = 0;
        while ((n + 1) * (n + 2) / 2 <= k) n++;
        const double *t = tab + 5 * n;
        const double2 r = F[k], s = F[Q + k], u = F[2 * Q + k];
        F[k] = make_double2(t[0] * r.x + t[1] * s.x, t[0] * r.y + t[1] * s.y);
        F[Q + k] = make_double2(t[2] * r.x + t[3] * s.x, t[2] * r.y + t[3] * s.y);
        F[2 * Q + k] = make_double2(t[4] * u.x, t[4] * u.y);
    }
}

void launch_degree_blocks(const DevGeom &g, double *coef, const double *tab, const int *active, hipStream_t s) {
    if (g.nsph) k_degree_blocks<<<dim3(g.nsph), dim3(256), 0, s>>>(g, coef, tab, active);
}

void launch_expand(const DevGeom &g, int mode, const double *coef, double *rec, const int *active, hipStream_t s,
                   int max_order) {
    if (!g.nsph) return;
    const int Q = tri_size(max_order), Q1 = tri_size(max_order + 1);
    const size_t lds = sizeof(double) * 2 * ((mode == 0 ? 4 : 8) * Q + (mode == 0 ? 9 : 18) * Q1);
    static size_t lds_set = 0;
    if (lds > 65536 && lds > lds_set) {
        if (hipFuncSetAttribute((const void *)k_expand, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
            hipSuccess)
            throw Error(-2, "cannot raise the LDS limit of the expansion kernel to " + std::to_string(lds));
        lds_set = lds;
    }
    k_expand<<<dim3(g.nsph), dim3(256), lds, s>>>(g, mode, coef, rec, active);
}

// ---------------------------------------------------------------------------
// Expansion evaluation: acc[a] = sum_{n<=P, m<=n} (c0 D.re + c1 D.im) for arrays
// a in [KBEG, KBEG+KCNT) of a record with KTOT arrays, D_n^m by recurrence.

template <int KTOT, int KBEG, int KCNT>
__device__ __forceinline__ void eval_record(const double *__restrict__ rec, int P, double x, double y, double z,
                                            double *acc) {
#pragma unroll
    for (int a = 0; a < KCNT; a++) acc[a] = 0.0;
    const double r2 = x * x + y * y + z * z;
    const double ir2 = 1.0 / r2;
    double dmr = rsqrt_f64(r2), dmi = 0.0; // D_m^m
    for (int m = 0; m <= P; m++) {
        if (m > 0) {
            const double f = -(2.0 * m - 1.0) * ir2;
            const double nr = f * (dmr * x - dmi * y), ni = f * (dmr * y + dmi * x);
            dmr = nr;
            dmi = ni;
        }
        double pr = 0.0, pi = 0.0, cr = dmr, ci = dmi;
        const double mm = (double)m * m;
        for (int n = m; n <= P; n++) {
            const double *c = rec + 2 * (KTOT * dtri(n, m) + KBEG);
#pragma unroll
            for (int a = 0; a < KCNT; a++) acc[a] = __builtin_fma(c[2 * a], cr, __builtin_fma(c[2 * a + 1], ci, acc[a]));
            const double f1 = -(2.0 * n + 1.0) * z * ir2, f2 = -((double)n * n - mm) * ir2;
            const double nr = __builtin_fma(f1, cr, f2 * pr), ni = __builtin_fma(f1, ci, f2 * pi);
            pr = cr;
            pi = ci;
            cr = nr;
            ci = ni;
        }
    }
}

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// symmetric 3x3 from (xx,xy,xz,yy,yz,zz) times v
__device__ __forceinline__ void symv(const double *h, const double v[3], double o[3]) {
    o[0] = h[0] * v[0] + h[1] * v[1] + h[2] * v[2];
    o[1] = h[1] * v[0] + h[3] * v[1] + h[4] * v[2];
    o[2] = h[2] * v[0] + h[4] * v[1] + h[5] * v[2];
}

// traction of the exterior flow from the 26 traction arrays (layout of k_expand)
__device__ __forceinline__ void traction_from(const double *A, const double x[3], const double v[3], double t[3]) {
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    const double xv = x[0] * v[0] + x[1] * v[1] + x[2] * v[2];
    const double *g2 = A + 6, *g3 = A + 16;
    const double h3 = A[15], pr = A[25];
    const double vg2 = v[0] * g2[0] + v[1] * g2[1] + v[2] * g2[2];
    const double vg3 = v[0] * g3[0] + v[1] * g3[1] + v[2] * g3[2];
    double H1v[3], H2v[3], H4v[3], vx[3], H4vx[3], xH4v[3];
    symv(A + 0, v, H1v);
    symv(A + 9, v, H2v);
    symv(A + 19, v, H4v);
    cross3(v, x, vx);
    symv(A + 19, vx, H4vx);
    cross3(x, H4v, xH4v);
    for (int c = 0; c < 3; c++)
        t[c] = -pr * v[c] + 2.0 * H1v[c] + 2.0 * xv * g2[c] + 2.0 * x[c] * vg2 + 2.0 * r2 * H2v[c] +
               2.0 * h3 * v[c] + x[c] * vg3 + xv * g3[c] + xH4v[c] + H4vx[c];
}

// One work item = one 64-lane wave: targets [tstart, tstart+tcount) of sphere T
// against the expansion of sphere S (S == T: self).  MODE 0: single layer,
// 2: synthesis of the projected density (self items only); the traction runs
// in k_sheval_src.  MODE 0 writes the item's contribution (lab frame,
// coefficients applied) to scratch[slot + 3*lane]; MODE 2 writes out directly.
// Workgroups go to the 8 XCDs round-robin, so workgroup b takes item
// (b % 8) * ceil(nitems / 8) + b / 8: each XCD runs a contiguous range of the
// items, whose chunks and neighbouring targets read the same source records
// (63 KB at p = 12) from its own L2.  The grid is 8 ceil(nitems / 8).
template <int MODE>
__global__ __launch_bounds__(64) void k_sheval(DevGeom g, const double *__restrict__ rec,
                                               const EvalItem *__restrict__ items, int nitems,
                                               double *__restrict__ scratch, double *__restrict__ out, double cSL,
                                               double cTrac, const int *__restrict__ active) {
    const int chunk = (int)(gridDim.x >> 3);
    const int pos = (int)(blockIdx.x & 7) * chunk + (int)(blockIdx.x >> 3);
    if (pos >= nitems) return;
    const EvalItem it = items[pos];
    const int T = it.trg, S = it.src;
    if (active && !active[g.sph_susp[T]]) return;
    const int lane = threadIdx.x;
    if (lane >= it.tcount) return;
    const int64_t gp = g.pt_off[T] + it.tstart + lane;
    const double X[3] = {g.coords[3 * gp], g.coords[3 * gp + 1], g.coords[3 * gp + 2]};
    const double *q = g.quat + 4 * S;
    const double R = g.rad[S];
    const double iR = 1.0 / R;
    const double d[3] = {(X[0] - g.pos[3 * S]) * iR, (X[1] - g.pos[3 * S + 1]) * iR, (X[2] - g.pos[3 * S + 2]) * iR};
    double xi[3], res[3], lab[3];
    qrot(q, d, xi, true);
    const int P = g.order[S] + 2;
    const double *recS = rec + g.rec_off[S];
    if (MODE == 0) {
        double A[kRecSL];
        eval_record<kRecSL, 0, kRecSL>(recS, P, xi[0], xi[1], xi[2], A);
        const double r2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
        double xg4[3];
        cross3(xi, A + 7, xg4);
        for (int c = 0; c < 3; c++) res[c] = A[c] + r2 * A[3 + c] + xi[c] * A[6] + xg4[c];
        qrot(q, res, lab, false);
        for (int c = 0; c < 3; c++) scratch[it.slot + 3 * lane + c] = cSL * R * lab[c];
    } else {
        double A[7];
        eval_record<kRecTrac, 26, 7>(recS, P, xi[0], xi[1], xi[2], A);
        double xq[3];
        cross3(xi, A + 4, xq);
        for (int c = 0; c < 3; c++) res[c] = A[0] * xi[c] + A[1 + c] + xq[c];
        qrot(q, res, lab, false);
        for (int c = 0; c < 3; c++) out[3 * gp + c] = lab[c];
    }
}

// Traction evaluation of the self and near items with the source's record in
// LDS: one workgroup of kSrcWaves waves per source sphere S runs all items
// whose source is S (its self item and its near targets' chunks), the waves
// taking the items in turn.  The record (kRecTrac complex arrays per term,
// (p+3)(p+4)/2 terms: 63 KB at p = 12) is read from global memory once per
// source instead of once per item through the scalar cache, whose loads each
// term of the item loop waited on; with up to 256 VGPRs per lane (2 waves per
// SIMD, which the LDS allows anyway) each term's LDS reads issue together.
// eval_lds is eval_record with the coefficients read from LDS: the same
// operations in the same order.
// One launch per sphere order over that order's spheres (in k-d order), LDS
// sized to its record; workgroup b takes list entry (b % 8) ceil(n / 8) + b / 8
// (XCD-contiguous, as k_sheval).
constexpr int kSrcWaves = 4;

template <int KCNT>
__device__ __forceinline__ void eval_lds(const double2 *__restrict__ srec, int P, double x, double y, double z,
                                         double *acc) {
#pragma unroll
    for (int a = 0; a < KCNT; a++) acc[a] = 0.0;
    const double r2 = x * x + y * y + z * z;
    const double ir2 = 1.0 / r2;
    double dmr = rsqrt_f64(r2), dmi = 0.0; // D_m^m
    for (int m = 0; m <= P; m++) {
        if (m > 0) {
            const double f = -(2.0 * m - 1.0) * ir2;
            const double nr = f * (dmr * x - dmi * y), ni = f * (dmr * y + dmi * x);
            dmr = nr;
            dmi = ni;
        }
        double pr = 0.0, pi = 0.0, cr = dmr, ci = dmi;
        const double mm = (double)m * m;
        for (int n = m; n <= P; n++) {
            const double2 *c = srec + kRecTrac * dtri(n, m);
#pragma unroll
            for (int a = 0; a < KCNT; a++) {
                const double2 v = c[a];
                acc[a] = __builtin_fma(v.x, cr, __builtin_fma(v.y, ci, acc[a]));
            }
            const double f1 = -(2.0 * n + 1.0) * z * ir2, f2 = -((double)n * n - mm) * ir2;
            const double nr = __builtin_fma(f1, cr, f2 * pr), ni = __builtin_fma(f1, ci, f2 * pi);
            pr = cr;
            pi = ci;
            cr = nr;
            ci = ni;
        }
    }
}

__global__ __launch_bounds__(64 * kSrcWaves) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_sheval_src(DevGeom g, const double *__restrict__ rec,
                                                              const EvalItem *__restrict__ items,
                                                              const int64_t *__restrict__ src_off,
                                                              const int *__restrict__ srcs, int nsrc,
                                                              double *__restrict__ scratch, double cTrac,
                                                              const int *__restrict__ active) {
    extern __shared__ double2 srec[];
    const int chunk = (int)(gridDim.x >> 3);
    const int pos = (int)(blockIdx.x & 7) * chunk + (int)(blockIdx.x >> 3);
    if (pos >= nsrc) return;
    const int S = srcs[pos];
    // near pairs lie within one suspension: the test is uniform over the workgroup
    if (active && !active[g.sph_susp[S]]) return;
    const int64_t i0 = src_off[S], i1 = src_off[S + 1];
    if (i0 == i1) return;
    const int P = g.order[S] + 2;
    const int nrec = kRecTrac * ((P + 1) * (P + 2) / 2);
    const double2 *recS = reinterpret_cast<const double2 *>(rec + g.rec_off[S]);
    for (int i = threadIdx.x; i < nrec; i += blockDim.x) srec[i] = recS[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const double *q = g.quat + 4 * S;
    const double R = g.rad[S];
    const double iR = 1.0 / R;
    const double cx = g.pos[3 * S], cy = g.pos[3 * S + 1], cz = g.pos[3 * S + 2];
    for (int64_t k = i0 + (threadIdx.x >> 6); k < i1; k += kSrcWaves) {
        const EvalItem it = items[k];
        if (lane >= it.tcount) continue;
        const int T = it.trg;
        const int64_t gp = g.pt_off[T] + it.tstart + lane;
        const double d[3] = {(g.coords[3 * gp] - cx) * iR, (g.coords[3 * gp + 1] - cy) * iR,
                             (g.coords[3 * gp + 2] - cz) * iR};
        const double Nl[3] = {g.norms[3 * gp], g.norms[3 * gp + 1], g.norms[3 * gp + 2]};
        double xi[3], vb[3], res[3], lab[3];
        qrot(q, d, xi, true);
        qrot(q, Nl, vb, true);
        if (T == S) {
            double A[kRecTrac];
            eval_lds<kRecTrac>(srec, P, xi[0], xi[1], xi[2], A);
            traction_from(A, xi, vb, res);
            double xq[3];
            cross3(xi, A + 30, xq);
            for (int c = 0; c < 3; c++) res[c] += 0.5 * (A[26] * xi[c] + A[27 + c] + xq[c]);
        } else {
            double A[kRecTracNear];
            eval_lds<kRecTracNear>(srec, P, xi[0], xi[1], xi[2], A);
            traction_from(A, xi, vb, res);
        }
        qrot(q, res, lab, false);
        for (int c = 0; c < 3; c++) scratch[it.slot + 3 * lane + c] = cTrac * lab[c];
    }
}

// out[T's points] += sum over T's items (self first, then near pairs in list
// order) of their scratch contributions + cId in
__global__ __launch_bounds__(256) void k_sheval_gather(DevGeom g, const int64_t *__restrict__ trg_slot_off,
                                                       const int *__restrict__ trg_nsrc, const double *__restrict__ scratch,
                                                       const double *__restrict__ in, double *__restrict__ out, double cId,
                                                       const int *__restrict__ active) {
    const int T = blockIdx.x;
    if (active && !active[g.sph_susp[T]]) return;
    const int64_t off = g.pt_off[T];
    const int nT = (int)(g.pt_off[T + 1] - off);
    const int64_t base = trg_slot_off[T];
    const int ns = trg_nsrc[T];
    for (int i = threadIdx.x; i < 3 * nT; i += blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < ns; k++) s += scratch[base + (int64_t)k * 3 * nT + i];
        out[3 * off + i] += s + cId * in[3 * off + i];
    }
}

void launch_sheval_items(const DevGeom &g, int mode, const double *rec, const EvalItem *items, int nitems,
                         double *scratch, double cSL, double cTrac, const int *active, hipStream_t s,
                         const SrcItems *by_src) {
    if (!g.nsph || nitems <= 0) return;
    if (mode == 1) {
        if (!by_src) throw Error(-2, "internal: the traction evaluation needs the items grouped by source");
        for (int p = 0; p <= kMaxOrder; p++) {
            const int a = by_src->order_off[p], n = by_src->order_off[p + 1] - a;
            if (n <= 0) continue;
            const int P = p + 2;
            const size_t lds = sizeof(double2) * kRecTrac * ((P + 1) * (P + 2) / 2);
            static size_t lds_set = 0;
            if (lds > 65536 && lds > lds_set) {
                if (hipFuncSetAttribute((const void *)k_sheval_src, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds) != hipSuccess)
                    throw Error(-2, "cannot raise the LDS limit of the self/near evaluation to " +
                                        std::to_string(lds));
                lds_set = lds;
            }
            k_sheval_src<<<dim3(8 * ((n + 7) / 8)), dim3(64 * kSrcWaves), lds, s>>>(
                g, rec, by_src->items, by_src->src_off, by_src->srcs + a, n, scratch, cTrac, active);
        }
        return;
    }
    k_sheval<0><<<dim3(8 * ((nitems + 7) / 8)), dim3(64), 0, s>>>(g, rec, items, nitems, scratch, nullptr, cSL,
                                                                   cTrac, active);
}

void launch_sheval_gather(const DevGeom &g, int nitems, const int64_t *trg_slot_off, const int *trg_nsrc,
                          const double *scratch, const double *in, double *out, double cId, const int *active,
                          hipStream_t s) {
    if (!g.nsph || nitems <= 0) return;
    k_sheval_gather<<<dim3(g.nsph), dim3(256), 0, s>>>(g, trg_slot_off, trg_nsrc, scratch, in, out, cId, active);
}

void launch_sheval(const DevGeom &g, int mode, const double *rec, const EvalItem *items, int nitems,
                   const int64_t *trg_slot_off, const int *trg_nsrc, double *scratch, const double *in, double *out,
                   double cSL, double cTrac, double cId, const int *active, hipStream_t s, const SrcItems *by_src) {
    launch_sheval_items(g, mode, rec, items, nitems, scratch, cSL, cTrac, active, s, by_src);
    launch_sheval_gather(g, nitems, trg_slot_off, trg_nsrc, scratch, in, out, cId, active, s);
}

void launch_synthesis(const DevGeom &g, const double *rec, const EvalItem *self_items, int nitems, double *out,
                      hipStream_t s, const int *active) {
    if (!g.nsph || nitems <= 0) return;
    k_sheval<2><<<dim3(8 * ((nitems + 7) / 8)), dim3(64), 0, s>>>(g, rec, self_items, nitems, nullptr, out, 0.0, 0.0,
                                                                  active);
}

__global__ void k_identity(DevGeom g, const double *__restrict__ in, double *__restrict__ out, double cId,
                           const int *__restrict__ active) {
    const int S = blockIdx.x;
    if (active && !active[g.sph_susp[S]]) return;
    for (int64_t i = 3 * g.pt_off[S] + threadIdx.x; i < 3 * g.pt_off[S + 1]; i += blockDim.x) out[i] += cId * in[i];
}

void launch_identity(const DevGeom &g, const double *in, double *out, double cId, const int *active,
                     hipStream_t s) {
    if (!g.nsph) return;
    k_identity<<<dim3(g.nsph), dim3(256), 0, s>>>(g, in, out, cId, active);
}

// ---------------------------------------------------------------------------
// rigid-body projector L (SphereSTKSHOperator.cpp:487-521):
//   A = sum w sigma / (4 pi R^2),  B = 3/(8 pi R^4) sum w R n x sigma,
//   out += cLOP (A + B x R n)

__global__ __launch_bounds__(256) void k_lop(DevGeom g, const double *__restrict__ in, double *__restrict__ out,
                                             double cLOP, const int *__restrict__ active) {
    __shared__ double sh[8];
    const int S = blockIdx.x;
    if (active && !active[g.sph_susp[S]]) return;
    const int64_t b = g.pt_off[S], e = g.pt_off[S + 1];
    const double R = g.rad[S];
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const double w = g.w[i];
        const double v[3] = {in[3 * i], in[3 * i + 1], in[3 * i + 2]};
        const double nv[3] = {g.norms[3 * i], g.norms[3 * i + 1], g.norms[3 * i + 2]};
        a[0] += w * v[0];
        a[1] += w * v[1];
        a[2] += w * v[2];
        a[3] += w * R * (nv[1] * v[2] - nv[2] * v[1]);
        a[4] += w * R * (nv[2] * v[0] - nv[0] * v[2]);
        a[5] += w * R * (nv[0] * v[1] - nv[1] * v[0]);
    }
    double tot[6];
    for (int c = 0; c < 6; c++) tot[c] = block_sum(a[c], sh);
    const double ca = 1.0 / (4.0 * kPi * R * R), cb = 3.0 / (8.0 * kPi * R * R * R * R);
    const double A[3] = {tot[0] * ca, tot[1] * ca, tot[2] * ca};
    const double B[3] = {tot[3] * cb, tot[4] * cb, tot[5] * cb};
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const double rn[3] = {R * g.norms[3 * i], R * g.norms[3 * i + 1], R * g.norms[3 * i + 2]};
        out[3 * i] += cLOP * (A[0] + B[1] * rn[2] - B[2] * rn[1]);
        out[3 * i + 1] += cLOP * (A[1] + B[2] * rn[0] - B[0] * rn[2]);
        out[3 * i + 2] += cLOP * (A[2] + B[0] * rn[1] - B[1] * rn[0]);
    }
}

void launch_lop(const DevGeom &g, const double *in, double *out, double cLOP, const int *active, hipStream_t s) {
    if (!g.nsph) return;
    k_lop<<<dim3(g.nsph), dim3(256), 0, s>>>(g, in, out, cLOP, active);
}

// ---------------------------------------------------------------------------
// segmented BLAS-1 (deterministic: per-chunk partials reduced in chunk order)

__global__ __launch_bounds__(256) void k_seg_dot(const double *__restrict__ a, const double *__restrict__ b,
                                                 const Chunk *__restrict__ ch, double *__restrict__ partial,
                                                 const int *__restrict__ active) {
    __shared__ double sh[4];
    const Chunk c = ch[blockIdx.x];
    double s = 0.0;
    if (!active || active[c.seg])
        for (int64_t i = c.begin + threadIdx.x; i < c.end; i += blockDim.x) s = __builtin_fma(a[i], b[i], s);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_seg_reduce(const double *__restrict__ partial, const int *__restrict__ off, int nseg,
                             double *__restrict__ out, int op, const int *__restrict__ active) {
    const int sgi = blockIdx.x * blockDim.x + threadIdx.x;
    if (sgi >= nseg) return;
    if (active && !active[sgi]) return;
    double s = 0.0;
    for (int k = off[sgi]; k < off[sgi + 1]; k++) s += partial[k];
    if (op == 1) s = (s > 0.0) ? 1.0 / sqrt(s) : 0.0;
    if (op == 2) s = -s;
    out[sgi] = s;
}

__global__ __launch_bounds__(256) void k_seg_axpy(double *__restrict__ y, const double *__restrict__ x,
                                                  const double *__restrict__ alpha, double sign,
                                                  const Chunk *__restrict__ ch, const int *__restrict__ active) {
    const Chunk c = ch[blockIdx.x];
    if (active && !active[c.seg]) return;
    const double a = sign * alpha[c.seg];
    for (int64_t i = c.begin + threadIdx.x; i < c.end; i += blockDim.x) y[i] = __builtin_fma(a, x[i], y[i]);
}

__global__ __launch_bounds__(256) void k_seg_scale(double *__restrict__ y, const double *__restrict__ x,
                                                   const double *__restrict__ alpha, const Chunk *__restrict__ ch,
                                                   const int *__restrict__ active) {
    const Chunk c = ch[blockIdx.x];
    if (active && !active[c.seg]) return;
    const double a = alpha[c.seg];
    for (int64_t i = c.begin + threadIdx.x; i < c.end; i += blockDim.x) y[i] = a * x[i];
}

// x (+)= sum_j coef[j nseg + seg] V_j over the active segments (SET: x = the
// sum).  Zero coefficients skip their vector: a suspension that left a GMRES
// cycle early has basis vectors beyond its dimension that this cycle never
// wrote.
template <bool SET>
__global__ __launch_bounds__(256) void k_seg_multi_axpy(double *__restrict__ x, const double *__restrict__ V,
                                                        int64_t ldv, const double *__restrict__ coef, int k,
                                                        int nseg, const Chunk *__restrict__ ch,
                                                        const int *__restrict__ active) {
    const Chunk c = ch[blockIdx.x];
    if (active && !active[c.seg]) return;
    for (int64_t i = c.begin + threadIdx.x; i < c.end; i += blockDim.x) {
        double s = SET ? 0.0 : x[i];
        for (int j = 0; j < k; j++) {
            const double cf = coef[(int64_t)j * nseg + c.seg];
            if (cf != 0.0) s = __builtin_fma(cf, V[(int64_t)j * ldv + i], s);
        }
        x[i] = s;
    }
}

// Block-partial dot products of w with k basis vectors V_j (stride ldv) over
// one chunk: partial[j * nchunks + chunk], j < k <= kMaxDots, in one pass over w.
constexpr int kMaxDots = 64;
__global__ __launch_bounds__(256) void k_seg_multi_dot(const double *__restrict__ w, const double *__restrict__ V,
                                                       int64_t ldv, int k, const Chunk *__restrict__ ch,
                                                       double *__restrict__ partial, const int *__restrict__ active) {
    __shared__ double sh[4][kMaxDots];
    const Chunk c = ch[blockIdx.x];
    const bool on = !active || active[c.seg];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j0 = 0; j0 < k; j0 += 16) {
        double s[16];
#pragma unroll
        for (int u = 0; u < 16; u++) s[u] = 0.0;
        if (on)
            for (int64_t i = c.begin + threadIdx.x; i < c.end; i += blockDim.x) {
                const double wi = w[i];
#pragma unroll
                for (int u = 0; u < 16; u++)
                    if (j0 + u < k) s[u] = __builtin_fma(wi, V[(int64_t)(j0 + u) * ldv + i], s[u]);
            }
#pragma unroll
        for (int u = 0; u < 16; u++) {
            double v = s[u];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) sh[wid][j0 + u < kMaxDots ? j0 + u : 0] = v;
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < k; j += blockDim.x)
        partial[(int64_t)j * gridDim.x + blockIdx.x] = (sh[0][j] + sh[1][j]) + (sh[2][j] + sh[3][j]);
}

// out[j * nseg + seg] = (op 2: -) sum of the chunk partials of seg, for j < k
__global__ void k_seg_multi_reduce(const double *__restrict__ partial, int nchunks, const int *__restrict__ off,
                                   int nseg, int k, double *__restrict__ out, int op, const int *__restrict__ active) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nseg * k) return;
    const int sgi = idx % nseg, j = idx / nseg;
    if (active && !active[sgi]) return;
    double s = 0.0;
    for (int q = off[sgi]; q < off[sgi + 1]; q++) s += partial[(int64_t)j * nchunks + q];
    out[(int64_t)j * nseg + sgi] = op == 2 ? -s : s;
}

// out[i] += sum_k part[k*stride + i], k = 0..nslot-1, over the DOF chunks of active segments
// ---------------------------------------------------------------------------
// GMRES least-squares bookkeeping on the device (Belos GMRES, one Krylov space
// per suspension): with it a restart cycle needs no host synchronisation.
//
// k_gmres_init: start of a cycle: g_s = beta_s e_1, active = the cycle's set.
__global__ void k_gmres_init(int B, int m, const double *__restrict__ beta, const int *__restrict__ cyc,
                             double *__restrict__ g, int *__restrict__ cur, int *__restrict__ kdim,
                             int *__restrict__ impok) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    g[(size_t)s * (m + 1)] = beta[s];
    cur[s] = cyc[s];
    kdim[s] = 0;
    impok[s] = 0;
}

// k_gmres_givens: column j of every active suspension's Hessenberg matrix from
// the orthogonalisation products hh (two passes, then |w|^2; hsign -1 when the
// passes stored -h), the previous Givens rotations applied to it, the new
// rotation, the rotated right-hand side g and the implicit residual
// |g_{j+1}| / |b|; a suspension whose residual reaches tol (or whose Krylov
// space is exhausted) leaves the active set.  nact = suspensions still active.
// One workgroup loops over the suspensions (no atomics).
__global__ __launch_bounds__(256) void k_gmres_givens(int B, int j, int m, const double *__restrict__ hh,
                                                      double hsign, double *__restrict__ H, double *__restrict__ cs,
                                                      double *__restrict__ sn, double *__restrict__ g,
                                                      const double *__restrict__ bn, double tol,
                                                      int *__restrict__ cur, int *__restrict__ kdim,
                                                      int *__restrict__ impok, double *__restrict__ relres,
                                                      int *__restrict__ nact) {
    __shared__ int sh[256];
    int cnt = 0;
    for (int s = threadIdx.x; s < B; s += blockDim.x) {
        if (!cur[s]) continue;
        double *Hs = H + (size_t)s * (m + 1) * m;
        double *c = cs + (size_t)s * m, *sv = sn + (size_t)s * m, *gs = g + (size_t)s * (m + 1);
        for (int i = 0; i <= j; i++) Hs[i * m + j] = hsign * (hh[(size_t)i * B + s] + hh[(size_t)((m + 1) + i) * B + s]);
        const double hn = sqrt(fmax(hh[(size_t)(2 * (m + 1)) * B + s], 0.0));
        Hs[(j + 1) * m + j] = hn;
        for (int i = 0; i < j; i++) {
            const double t0 = c[i] * Hs[i * m + j] + sv[i] * Hs[(i + 1) * m + j];
            const double t1 = -sv[i] * Hs[i * m + j] + c[i] * Hs[(i + 1) * m + j];
            Hs[i * m + j] = t0;
            Hs[(i + 1) * m + j] = t1;
        }
        const double a = Hs[j * m + j], bb = Hs[(j + 1) * m + j];
        const double den = sqrt(a * a + bb * bb);
        c[j] = den > 0 ? a / den : 1.0;
        sv[j] = den > 0 ? bb / den : 0.0;
        Hs[j * m + j] = den;
        Hs[(j + 1) * m + j] = 0.0;
        gs[j + 1] = -sv[j] * gs[j];
        gs[j] = c[j] * gs[j];
        kdim[s] = j + 1;
        const double rr = fabs(gs[j + 1]) / bn[s];
        relres[s] = rr;
        if (rr <= tol) impok[s] = 1;
        if (rr <= tol || hn == 0.0)
            cur[s] = 0;
        else
            cnt++;
    }
    sh[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *nact = sh[0];
}

// k_gmres_backsub: y_s = R_s^-1 g_s (back substitution of the rotated
// Hessenberg system, kdim_s columns) as coef[i B + s], zero beyond kdim_s and
// for suspensions outside the cycle.
__global__ void k_gmres_backsub(int B, int m, const double *__restrict__ H, const double *__restrict__ g,
                                const int *__restrict__ kdim, const int *__restrict__ cyc, double *__restrict__ coef) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    for (int i = 0; i < m; i++) coef[(size_t)i * B + s] = 0.0;
    if (!cyc[s]) return;
    const double *Hs = H + (size_t)s * (m + 1) * m;
    const double *gs = g + (size_t)s * (m + 1);
    const int k = kdim[s];
    for (int i = k - 1; i >= 0; i--) {
        double acc = gs[i];
        for (int q = i + 1; q < k; q++) acc -= Hs[i * m + q] * coef[(size_t)q * B + s];
        coef[(size_t)i * B + s] = Hs[i * m + i] != 0.0 ? acc / Hs[i * m + i] : 0.0;
    }
}

// k_gmres_resid_coef: coefficients of the residual at the end of a cycle in
// the cycle's basis.  With Omega = G_{k-1} ... G_0 the product of the Givens
// rotations (Omega Hbar = [R; 0], Omega beta e_1 = g) and y = R^-1 g[0..k),
//   r = V_{k+1} (beta e_1 - Hbar y) = V_{k+1} Omega^T (0, ..., 0, g_k)^T,
// so coef = g_k G_0^T ... G_{k-1}^T e_k (k + 1 entries, as coef[i B + s]);
// zero for suspensions outside the cycle.
__global__ void k_gmres_resid_coef(int B, int m, const double *__restrict__ g, const double *__restrict__ cs,
                                   const double *__restrict__ sn, const int *__restrict__ kdim,
                                   const int *__restrict__ cyc, double *__restrict__ coef) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    for (int i = 0; i <= m; i++) coef[(size_t)i * B + s] = 0.0;
    const int k = kdim[s];
    if (!cyc[s] || k <= 0) return;
    const double *c = cs + (size_t)s * m, *sv = sn + (size_t)s * m;
    coef[(size_t)k * B + s] = g[(size_t)s * (m + 1) + k];
    for (int j = k - 1; j >= 0; j--) {
        const double a = coef[(size_t)j * B + s], b = coef[(size_t)(j + 1) * B + s];
        coef[(size_t)j * B + s] = c[j] * a - sv[j] * b;
        coef[(size_t)(j + 1) * B + s] = sv[j] * a + c[j] * b;
    }
}

void launch_gmres_resid_coef(int B, int m, const double *g, const double *cs, const double *sn, const int *kdim,
                             const int *cyc, double *coef, hipStream_t s) {
    if (B > 0) k_gmres_resid_coef<<<dim3((B + 255) / 256), dim3(256), 0, s>>>(B, m, g, cs, sn, kdim, cyc, coef);
}

void launch_gmres_init(int B, int m, const double *beta, const int *cyc, double *g, int *cur, int *kdim, int *impok,
                       hipStream_t s) {
    if (B > 0) k_gmres_init<<<dim3((B + 255) / 256), dim3(256), 0, s>>>(B, m, beta, cyc, g, cur, kdim, impok);
}
void launch_gmres_givens(int B, int j, int m, const double *hh, double hsign, double *H, double *cs, double *sn,
                         double *g, const double *bn, double tol, int *cur, int *kdim, int *impok, double *relres,
                         int *nact, hipStream_t s) {
    k_gmres_givens<<<dim3(1), dim3(256), 0, s>>>(B, j, m, hh, hsign, H, cs, sn, g, bn, tol, cur, kdim, impok, relres,
                                                 nact);
}
void launch_gmres_backsub(int B, int m, const double *H, const double *g, const int *kdim, const int *cyc,
                          double *coef, hipStream_t s) {
    if (B > 0) k_gmres_backsub<<<dim3((B + 255) / 256), dim3(256), 0, s>>>(B, m, H, g, kdim, cyc, coef);
}

__global__ __launch_bounds__(256) void k_slot_reduce(double *__restrict__ out, const double *__restrict__ part,
                                                     int nslot, int64_t stride, const Chunk *__restrict__ ch,
                                                     const int *__restrict__ active) {
    const Chunk c = ch[blockIdx.x];
    if (active && !active[c.seg]) return;
    for (int64_t i = c.begin + threadIdx.x; i < c.end; i += blockDim.x) {
        double s = out[i];
        for (int k = 0; k < nslot; k++) s += part[k * stride + i];
        out[i] = s;
    }
}

void launch_slot_reduce(double *out, const double *part, int nslot, int64_t stride, const Chunk *chunks,
                        int nchunks, const int *active, hipStream_t s) {
    if (nchunks > 0) k_slot_reduce<<<dim3(nchunks), dim3(256), 0, s>>>(out, part, nslot, stride, chunks, active);
}

__global__ void k_axpby(int64_t n, double a, const double *__restrict__ x, double b, double *__restrict__ y) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = (b == 0.0) ? a * x[i] : __builtin_fma(a, x[i], b * y[i]);
}

__global__ void k_fill(int64_t n, double v, double *__restrict__ y) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

void launch_seg_dot(const double *a, const double *b, const Chunk *chunks, int nchunks, double *partial,
                    const int *active, hipStream_t s) {
    if (nchunks > 0) k_seg_dot<<<dim3(nchunks), dim3(256), 0, s>>>(a, b, chunks, partial, active);
}
void launch_seg_reduce(const double *partial, const int *off, int nseg, double *out, int op, const int *active,
                       hipStream_t s) {
    if (nseg > 0) k_seg_reduce<<<dim3((nseg + 63) / 64), dim3(64), 0, s>>>(partial, off, nseg, out, op, active);
}
void launch_seg_axpy(double *y, const double *x, const double *alpha, double sign, const Chunk *chunks, int nchunks,
                     const int *active, hipStream_t s) {
    if (nchunks > 0) k_seg_axpy<<<dim3(nchunks), dim3(256), 0, s>>>(y, x, alpha, sign, chunks, active);
}
void launch_seg_scale(double *y, const double *x, const double *alpha, const Chunk *chunks, int nchunks,
                      const int *active, hipStream_t s) {
    if (nchunks > 0) k_seg_scale<<<dim3(nchunks), dim3(256), 0, s>>>(y, x, alpha, chunks, active);
}
void launch_seg_multi_axpy(double *x, const double *V, int64_t ldv, const double *coef, int k, int nseg,
                           const Chunk *chunks, int nchunks, const int *active, hipStream_t s) {
    if (nchunks > 0 && k > 0)
        k_seg_multi_axpy<false><<<dim3(nchunks), dim3(256), 0, s>>>(x, V, ldv, coef, k, nseg, chunks, active);
}
void launch_seg_multi_comb(double *x, const double *V, int64_t ldv, const double *coef, int k, int nseg,
                           const Chunk *chunks, int nchunks, const int *active, hipStream_t s) {
    if (nchunks > 0 && k > 0)
        k_seg_multi_axpy<true><<<dim3(nchunks), dim3(256), 0, s>>>(x, V, ldv, coef, k, nseg, chunks, active);
}
void launch_seg_multi_dot(const double *w, const double *V, int64_t ldv, int k, const Chunk *chunks, int nchunks,
                          double *partial, const int *active, hipStream_t s) {
    if (k > kMaxDots) throw Error(-1, "internal: too many dot products in one pass");
    if (nchunks > 0 && k > 0)
        k_seg_multi_dot<<<dim3(nchunks), dim3(256), 0, s>>>(w, V, ldv, k, chunks, partial, active);
}
void launch_seg_multi_reduce(const double *partial, int nchunks, const int *off, int nseg, int k, double *out, int op,
                             const int *active, hipStream_t s) {
    const int n = nseg * k;
    if (n > 0)
        k_seg_multi_reduce<<<dim3((n + 63) / 64), dim3(64), 0, s>>>(partial, nchunks, off, nseg, k, out, op, active);
}
void launch_axpby(int64_t n, double a, const double *x, double b, double *y, hipStream_t s) {
    if (n > 0) k_axpby<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(n, a, x, b, y);
}
void launch_fill(int64_t n, double v, double *y, hipStream_t s) {
    if (n > 0) k_fill<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(n, v, y);
}

// ---------------------------------------------------------------------------
// Sphere-centre far field for the rigid-body averages of a single-layer flow
// (Faxen's laws, exact for flows whose singularities lie outside the sphere):
//   mean of u over the sphere = u(c) + R^2/6 lap u(c),   rotation = curl u(c) / 2.
// For a Stokeslet of strength f at y (r = c - y), times 1/(8 pi):
//   u = f/r + r (r.f)/r^3,  lap u = 2 f/r^3 - 6 r (r.f)/r^5,  curl u / 2 = (f x r)/r^3.
// Targets are sphere centres; out[9*T + (0..2: u, 3..5: lap u, 6..8: curl/2)] or,
// with slots, part[slot*9*ntrg + ...].
__global__ __launch_bounds__(64) void k_far_centre_f64(const double *__restrict__ src,
                                                       const double *__restrict__ centres,
                                                       const int *__restrict__ gorder,
                                                       const FarGroup *__restrict__ groups,
                                                       const FarInterval *__restrict__ iv, double *__restrict__ part,
                                                       int64_t ntrg) {
    const FarGroup G = groups[gorder[blockIdx.x]];
    const int lane = threadIdx.x;
    const int64_t tid = G.tstart + (lane < G.tcount ? lane : 0);
    const double x = centres[3 * tid], y = centres[3 * tid + 1], z = centres[3 * tid + 2];
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t k = G.iv_begin; k < G.iv_end; k++) {
        const bool incl = ((iv[k].mask[0] >> lane) & 1ull) != 0;
        for (int64_t s = iv[k].begin; s != iv[k].end; s++) {
            const double *sp = src + 8 * s;
            const double dx = x - sp[0], dy = y - sp[1], dz = z - sp[2];
            const double fx = sp[3], fy = sp[4], fz = sp[5];
            const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
            double ri = rsqrt_f64(r2);
            ri = incl ? ri : 0.0;
            const double ri2 = ri * ri, ri3 = ri2 * ri, ri5 = ri3 * ri2;
            const double df = __builtin_fma(dx, fx, __builtin_fma(dy, fy, dz * fz));
            const double q3 = df * ri3, q5 = -6.0 * df * ri5, t3 = 2.0 * ri3;
            a[0] = __builtin_fma(fx, ri, __builtin_fma(q3, dx, a[0]));
            a[1] = __builtin_fma(fy, ri, __builtin_fma(q3, dy, a[1]));
            a[2] = __builtin_fma(fz, ri, __builtin_fma(q3, dz, a[2]));
            a[3] = __builtin_fma(fx, t3, __builtin_fma(q5, dx, a[3]));
            a[4] = __builtin_fma(fy, t3, __builtin_fma(q5, dy, a[4]));
            a[5] = __builtin_fma(fz, t3, __builtin_fma(q5, dz, a[5]));
            a[6] = __builtin_fma(fy * dz - fz * dy, ri3, a[6]);
            a[7] = __builtin_fma(fz * dx - fx * dz, ri3, a[7]);
            a[8] = __builtin_fma(fx * dy - fy * dx, ri3, a[8]);
        }
    }
    if (lane < G.tcount) {
        double *o = part + G.slot * 9 * ntrg + 9 * (G.tstart + lane);
        for (int c = 0; c < 9; c++) o[c] = a[c];
    }
}

// vel[6T..] = (near/self rigid averages) + far part from the centre sums, / mu
__global__ void k_centre_finish(int nsph, int nslot, const double *__restrict__ part, const double *__restrict__ rad,
                                const double *__restrict__ vel_ns, double inv_mu, double *__restrict__ vel) {
    const int T = blockIdx.x * blockDim.x + threadIdx.x;
    if (T >= nsph) return;
    double a[9];
    for (int c = 0; c < 9; c++) {
        double s = 0.0;
        for (int k = 0; k < nslot; k++) s += part[(int64_t)k * 9 * nsph + 9 * T + c];
        a[c] = s;
    }
    const double R = rad[T], f = 1.0 / (8.0 * kPi);
    for (int c = 0; c < 3; c++) {
        vel[6 * T + c] = vel_ns[6 * T + c] + f * (a[c] + R * R / 6.0 * a[3 + c]) * inv_mu;
        vel[6 * T + 3 + c] = vel_ns[6 * T + 3 + c] + f * a[6 + c] * inv_mu;
    }
}

void launch_far_centre(const double *src, const double *centres, int ngroups, const int *go, const FarGroup *gr,
                       const FarInterval *iv, double *part, int64_t ntrg, hipStream_t s) {
    if (ngroups > 0) k_far_centre_f64<<<dim3(ngroups), dim3(64), 0, s>>>(src, centres, go, gr, iv, part, ntrg);
}
void launch_centre_finish(int nsph, int nslot, const double *part, const double *rad, const double *vel_ns,
                          double inv_mu, double *vel, hipStream_t s) {
    if (nsph > 0)
        k_centre_finish<<<dim3((nsph + 127) / 128), dim3(128), 0, s>>>(nsph, nslot, part, rad, vel_ns, inv_mu, vel);
}

// ---------------------------------------------------------------------------
// FP64 traction far field of NC (1..8) densities on one geometry (NC
// right-hand sides solved in lock step: copies of one set of suspensions).
// The work lists are those of the first copy; per source the record holds
// (x, y, z, 0, w f_0, 0, w f_1, 0, ...), 4 (NC + 1) doubles, so that d, r^2,
// r^-5 and d.n are computed once per pair for all NC densities:
// 15 VALU instructions + one v_rsq_f64 shared, 7 per density.
template <int NC>
__global__ void k_pack_src_multi(int64_t n0, const double *__restrict__ coords, const double *__restrict__ w,
                                 const double *__restrict__ in, int64_t col_stride, double *__restrict__ src) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n0) return;
    double *o = src + 4 * (NC + 1) * i;
    const double wi = w[i];
    o[0] = coords[3 * i];
    o[1] = coords[3 * i + 1];
    o[2] = coords[3 * i + 2];
    o[3] = 0.0;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const double *v = in + c * col_stride + 3 * i;
        o[4 * (c + 1)] = wi * v[0];
        o[4 * (c + 1) + 1] = wi * v[1];
        o[4 * (c + 1) + 2] = wi * v[2];
        o[4 * (c + 1) + 3] = 0.0;
    }
}

template <int NC>
__global__ __launch_bounds__(64) void k_far_trac_multi_f64(const double *__restrict__ src,
                                                           const double *__restrict__ coords,
                                                           const double *__restrict__ norms,
                                                           const int *__restrict__ gorder,
                                                           const FarGroup *__restrict__ groups,
                                                           const FarInterval *__restrict__ iv,
                                                           const int *__restrict__ active, int nsusp0,
                                                           double *__restrict__ out, double coef, int accumulate,
                                                           int64_t slot_stride, int64_t col_stride) {
    const FarGroup G = groups[gorder[blockIdx.x]];
    bool any = !active;
#pragma unroll
    for (int c = 0; c < NC; c++) any = any || active[c * nsusp0 + G.susp];
    if (!any) return;
    const int lane = threadIdx.x;
    const int64_t t = G.tstart + (lane < G.tcount ? lane : 0);
    const double x = coords[3 * t], y = coords[3 * t + 1], z = coords[3 * t + 2];
    const double nx = norms[3 * t], ny = norms[3 * t + 1], nz = norms[3 * t + 2];
    double ax[NC], ay[NC], az[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) ax[c] = ay[c] = az[c] = 0.0;
    constexpr int RS = 4 * (NC + 1);
    auto pair = [&](const double *v, bool incl) {
        const double dx = x - v[0], dy = y - v[1], dz = z - v[2];
        const double r2 = fmad(dx, dx, fmad(dy, dy, dz * dz));
        double wq = trac_w_f64(r2);
        wq = incl ? wq : 0.0;
        const double dnw = fmad(dx, nx, fmad(dy, ny, dz * nz)) * wq;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const double *f = v + 4 * (c + 1);
            const double q = fmad(dx, f[0], fmad(dy, f[1], dz * f[2])) * dnw;
            ax[c] = fmad(q, dx, ax[c]);
            ay[c] = fmad(q, dy, ay[c]);
            az[c] = fmad(q, dz, az[c]);
        }
    };
    // up to 4 columns: two sources per step, their wave-uniform (scalar) loads
    // issued before the arithmetic of either; more columns: one (a record of
    // 4 (NC + 1) doubles fills the scalar registers)
    constexpr int U = NC <= 4 ? 2 : 1;
    for (int64_t k = G.iv_begin; k < G.iv_end; k++) {
        const int64_t b = iv[k].begin, e = iv[k].end;
        const bool incl = ((iv[k].mask[0] >> lane) & 1ull) != 0;
        int64_t s = b;
        for (; s + U <= e; s += U) {
            double v[U][RS];
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int j = 0; j < RS; j++) v[u][j] = src[RS * (s + u) + j];
#pragma unroll
            for (int u = 0; u < U; u++) pair(v[u], incl);
        }
        if (s < e) {
            double v[RS];
#pragma unroll
            for (int j = 0; j < RS; j++) v[j] = src[RS * s + j];
            pair(v, incl);
        }
    }
    if (lane >= G.tcount) return;
    const double cf = kTracW * coef;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        if (active && !active[c * nsusp0 + G.susp]) continue;
        double *o = out + (accumulate ? 0 : G.slot * slot_stride) + c * col_stride + 3 * (G.tstart + lane);
        if (accumulate) {
            o[0] += cf * ax[c];
            o[1] += cf * ay[c];
            o[2] += cf * az[c];
        } else {
            o[0] = cf * ax[c];
            o[1] = cf * ay[c];
            o[2] = cf * az[c];
        }
    }
}

template <int NC>
static void far_trac_multi(int64_t n0, const double *coords, const double *norms, const double *w, const double *in,
                           int64_t col_stride, double *src, int ngroups, const int *go, const FarGroup *gr,
                           const FarInterval *iv, const int *active, int nsusp0, double *out, double coef, int acc,
                           int64_t slot_stride, hipStream_t s) {
    if (n0 > 0)
        k_pack_src_multi<NC><<<dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, s>>>(n0, coords, w, in, col_stride,
                                                                                     src);
    if (ngroups > 0)
        k_far_trac_multi_f64<NC><<<dim3(ngroups), dim3(64), 0, s>>>(src, coords, norms, go, gr, iv, active, nsusp0, out,
                                                                   coef, acc, slot_stride, col_stride);
}

void launch_far_trac_multi(int ncol, int64_t n0, const double *coords, const double *norms, const double *w,
                           const double *in, int64_t col_stride, double *src, int ngroups, const int *go,
                           const FarGroup *gr, const FarInterval *iv, const int *active, int nsusp0, double *out,
                           double coef, int acc, int64_t slot_stride, hipStream_t s) {
    switch (ncol) {
#define SPHSTK_MULTI_CASE(NC)                                                                                      \
    case NC:                                                                                                       \
        far_trac_multi<NC>(n0, coords, norms, w, in, col_stride, src, ngroups, go, gr, iv, active, nsusp0, out, coef, \
                           acc, slot_stride, s);                                                                   \
        break;
        SPHSTK_MULTI_CASE(1)
        SPHSTK_MULTI_CASE(2)
        SPHSTK_MULTI_CASE(3)
        SPHSTK_MULTI_CASE(4)
        SPHSTK_MULTI_CASE(5)
        SPHSTK_MULTI_CASE(6)
        SPHSTK_MULTI_CASE(7)
        SPHSTK_MULTI_CASE(8)
#undef SPHSTK_MULTI_CASE
    default:
        throw Error(-1, "shared far field supports 1 to 8 right-hand sides per launch");
    }
}

// ---------------------------------------------------------------------------
// Traction far field of the rigid-body densities rho0 = F/(4 pi R^2) +
// 3/(8 pi R^4) T x (R n) (BIE/SphereSTKMobMat.cpp:174-192) of distant spheres,
// in closed form at the sphere centres.  Outside its sphere, rho0 of sphere S
// generates the flow of a Stokeslet F, its degenerate quadrupole (R^2/6) lap G F
// and a rotlet T (all with the 1/(8 pi) normalisation of the single layer);
// K'[rho0](x) is that flow's traction on a surface of normal n at x.  With
// r = x - c:
//   Stokeslet:   -(3/4pi) r (r.F)(r.n)/r^5
//   quadrupole:  (R^2/12pi) [-3 (F (r.n) + r (F.n) + n (r.F))/r^5 + 15 (r.F)(r.n) r/r^7]
//   rotlet:      -(3/8pi) [(T x r)(r.n) + r ((T x r).n)]/r^5
// The order-p grid sum of the traction kernel equals this to ~(R/d)^(2p+2).
// Targets are grid points (one per lane); intervals run over sphere indices.
__global__ __launch_bounds__(64) void k_far_rigid_f64(const double *__restrict__ coords,
                                                      const double *__restrict__ norms,
                                                      const double *__restrict__ pos, const double *__restrict__ force,
                                                      const double *__restrict__ rad, const int *__restrict__ gorder,
                                                      const FarGroup *__restrict__ groups,
                                                      const FarInterval *__restrict__ iv, double *__restrict__ out,
                                                      double coef, int accumulate, int64_t slot_stride) {
    const FarGroup G = groups[gorder[blockIdx.x]];
    const int lane = threadIdx.x;
    const int64_t tid = G.tstart + (lane < G.tcount ? lane : 0);
    const double x = coords[3 * tid], y = coords[3 * tid + 1], z = coords[3 * tid + 2];
    const double nx = norms[3 * tid], ny = norms[3 * tid + 1], nz = norms[3 * tid + 2];
    double ax = 0.0, ay = 0.0, az = 0.0;
    constexpr double cS = -3.0 / (4.0 * kPi), cD = 1.0 / (12.0 * kPi), cR = -3.0 / (8.0 * kPi);
    for (int64_t k = G.iv_begin; k < G.iv_end; k++) {
        const bool incl = ((iv[k].mask[0] >> lane) & 1ull) != 0;
        for (int64_t s = iv[k].begin; s < iv[k].end; s++) {
            const double rx = x - pos[3 * s], ry = y - pos[3 * s + 1], rz = z - pos[3 * s + 2];
            const double Fx = force[6 * s], Fy = force[6 * s + 1], Fz = force[6 * s + 2];
            const double Tx = force[6 * s + 3], Ty = force[6 * s + 4], Tz = force[6 * s + 5];
            const double R2 = rad[s] * rad[s];
            double ri = rsqrt_f64(fmad(rx, rx, fmad(ry, ry, rz * rz)));
            ri = incl ? ri : 0.0;
            const double ri2 = ri * ri, ri5 = ri2 * ri2 * ri;
            const double rn = fmad(rx, nx, fmad(ry, ny, rz * nz));
            const double rF = fmad(rx, Fx, fmad(ry, Fy, rz * Fz));
            const double Fn = fmad(Fx, nx, fmad(Fy, ny, Fz * nz));
            const double qx = Ty * rz - Tz * ry, qy = Tz * rx - Tx * rz, qz = Tx * ry - Ty * rx; // T x r
            const double qn = fmad(qx, nx, fmad(qy, ny, qz * nz));
            // coefficients of r, F, n and T x r
            const double a_r = (cS * rF * rn + cD * R2 * 15.0 * rF * rn * ri2 - cD * R2 * 3.0 * Fn + cR * qn) * ri5;
            const double a_F = -cD * R2 * 3.0 * rn * ri5;
            const double a_n = -cD * R2 * 3.0 * rF * ri5;
            const double a_q = cR * rn * ri5;
            ax += a_r * rx + a_F * Fx + a_n * nx + a_q * qx;
            ay += a_r * ry + a_F * Fy + a_n * ny + a_q * qy;
            az += a_r * rz + a_F * Fz + a_n * nz + a_q * qz;
        }
    }
    if (lane >= G.tcount) return;
    if (!accumulate) out += G.slot * slot_stride;
    double *o = out + 3 * (G.tstart + lane);
    if (accumulate) {
        o[0] += coef * ax;
        o[1] += coef * ay;
        o[2] += coef * az;
    } else {
        o[0] = coef * ax;
        o[1] = coef * ay;
        o[2] = coef * az;
    }
}

void launch_far_rigid(const double *coords, const double *norms, const double *pos, const double *force,
                      const double *rad, int64_t ntrg, int ngroups, const int *go, const FarGroup *gr,
                      const FarInterval *iv, double *out, double coef, int acc, hipStream_t s) {
    if (ngroups > 0)
        k_far_rigid_f64<<<dim3(ngroups), dim3(64), 0, s>>>(coords, norms, pos, force, rad, go, gr, iv, out, coef, acc,
                                                          3 * ntrg);
}

// ---------------------------------------------------------------------------
// mobility helpers (SphereSTKMobMat::solveMob, BIE/SphereSTKMobMat.cpp:174-192, :261-285)

// rho0 = F/(4 pi R^2) + 3/(8 pi R^4) T x (R n) on every point of sphere S
__global__ void k_rho0(DevGeom g, const double *__restrict__ force, double *__restrict__ rho) {
    const int S = blockIdx.x;
    const double R = g.rad[S];
    const double a = 1.0 / (4.0 * kPi * R * R), c = 3.0 / (8.0 * kPi * R * R * R * R);
    const double *f = force + 6 * S, *t = force + 6 * S + 3;
    for (int64_t i = g.pt_off[S] + threadIdx.x; i < g.pt_off[S + 1]; i += blockDim.x) {
        const double rn[3] = {R * g.norms[3 * i], R * g.norms[3 * i + 1], R * g.norms[3 * i + 2]};
        rho[3 * i] = a * f[0] + c * (t[1] * rn[2] - t[2] * rn[1]);
        rho[3 * i + 1] = a * f[1] + c * (t[2] * rn[0] - t[0] * rn[2]);
        rho[3 * i + 2] = a * f[2] + c * (t[0] * rn[1] - t[1] * rn[0]);
    }
}

// V = sum w u/(4 pi R^2 mu), Omega = 3 sum w (R n x u)/(8 pi R^4 mu)
__global__ __launch_bounds__(256) void k_rigid_avg(DevGeom g, const double *__restrict__ u, double inv_mu,
                                                   double *__restrict__ vel) {
    __shared__ double sh[8];
    const int S = blockIdx.x;
    const double R = g.rad[S];
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = g.pt_off[S] + threadIdx.x; i < g.pt_off[S + 1]; i += blockDim.x) {
        const double w = g.w[i];
        const double v[3] = {u[3 * i], u[3 * i + 1], u[3 * i + 2]};
        const double nv[3] = {g.norms[3 * i], g.norms[3 * i + 1], g.norms[3 * i + 2]};
        a[0] += w * v[0];
        a[1] += w * v[1];
        a[2] += w * v[2];
        a[3] += R * w * (nv[1] * v[2] - nv[2] * v[1]);
        a[4] += R * w * (nv[2] * v[0] - nv[0] * v[2]);
        a[5] += R * w * (nv[0] * v[1] - nv[1] * v[0]);
    }
    double tot[6];
    for (int c = 0; c < 6; c++) tot[c] = block_sum(a[c], sh);
    if (threadIdx.x == 0) {
        const double cv = inv_mu / (4.0 * kPi * R * R), cw = 3.0 * inv_mu / (8.0 * kPi * R * R * R * R);
        for (int c = 0; c < 3; c++) {
            vel[6 * S + c] = tot[c] * cv;
            vel[6 * S + 3 + c] = tot[3 + c] * cw;
        }
    }
}

void launch_rho0(const DevGeom &g, const double *force, double *rho, hipStream_t s) {
    if (g.nsph) k_rho0<<<dim3(g.nsph), dim3(128), 0, s>>>(g, force, rho);
}
void launch_rigid_avg(const DevGeom &g, const double *u, double inv_mu, double *vel, hipStream_t s) {
    if (g.nsph) k_rigid_avg<<<dim3(g.nsph), dim3(256), 0, s>>>(g, u, inv_mu, vel);
}

} // namespace sphstk
