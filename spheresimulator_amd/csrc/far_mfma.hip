// Far-field traction sum with the three per-pair dot products on the matrix
// cores (v_mfma_f64_16x16x4f64) and the rest on the vector ALUs.
//
// Pair term (A.2; BIE/SphereSTKSHOperator.cpp:365-369, :435-446):
//   t_t = sum_s q_ts (x_t - y_s),  q_ts = (d.f_s)(d.n_t) / |d|^5,  d = x_t - y_s.
// With coordinates taken relative to the centre c of the target sphere
// (x' = x - c, y' = y - c; |x'| = R, so the expansions below lose no accuracy
// for the closest far pairs) the three dot products are 4-term contractions:
//   |d|^2 = |x'|^2 + [y', |y'|^2] . [-2x', 1]
//   d.n   = x'.n   + [y', |y'|^2] . [-n, 0]
//   d.f   =          [f, -y'.f]  . [x', 1]
// and the accumulation splits as t = x' sum_s q - sum_s q y'.
//
// One wave per work item: up to 64 targets of ONE sphere (column tiles ct of 16
// targets) against its source intervals, 16 sources per tile (matrix rows).
// Operand layouts of the 16x16x4 f64 instruction (lane l, j = l & 15,
// k = l >> 4): A[j][k], B[k][j]; the result item v of lane l is row 4k + v,
// column j.  The source coordinates are moved into the result layout by three
// more products with one-hot B columns, so no LDS or cross-lane traffic is
// needed inside the loop.  Each lane accumulates its four rows; the four lane
// groups are reduced once at the end in a fixed order.
#include "device_math.hpp"
#include "kernels.hpp"

namespace sphstk {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <int NCT>
__device__ __forceinline__ void tile_terms(const double ag, const double af, const double *br2, const double *bdn,
                                           const double *bdf, const double *cr2, const double *cdn, const d4 sx,
                                           const d4 sy, const d4 sz, double *Q, double *ax, double *ay,
                                           double *az) {
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
        const d4 r2 = mma(ag, br2[ct], d4{cr2[ct], cr2[ct], cr2[ct], cr2[ct]});
        const d4 dn = mma(ag, bdn[ct], d4{cdn[ct], cdn[ct], cdn[ct], cdn[ct]});
        const d4 df = mma(af, bdf[ct], d4{0.0, 0.0, 0.0, 0.0});
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double q = (df[v] * dn[v]) * trac_w_f64(r2[v]);
            Q[ct] += q;
            ax[ct] = __builtin_fma(q, sx[v], ax[ct]);
            ay[ct] = __builtin_fma(q, sy[v], ay[ct]);
            az[ct] = __builtin_fma(q, sz[v], az[ct]);
        }
    }
}

template <int NCT>
__global__ __launch_bounds__(64) void k_far_trac_mfma_f64(const double *__restrict__ src,
                                                          const double *__restrict__ coords,
                                                          const double *__restrict__ norms,
                                                          const double *__restrict__ gcen,
                                                          const int *__restrict__ gorder,
                                                          const FarGroup *__restrict__ groups,
                                                          const FarInterval *__restrict__ iv,
                                                          const int *__restrict__ active, double *__restrict__ out,
                                                          double coef, int accumulate, int64_t slot_stride) {
    const int gi = gorder[blockIdx.x];
    const FarGroup G = groups[gi];
    if (active && !active[G.susp]) return;
    const int l = threadIdx.x, j = l & 15, k = l >> 4;
    const double cx = gcen[4 * gi], cy = gcen[4 * gi + 1], cz = gcen[4 * gi + 2];
    // one-hot weights of this lane's operand component k
    const double e0 = k == 0 ? 1.0 : 0.0, e1 = k == 1 ? 1.0 : 0.0, e2 = k == 2 ? 1.0 : 0.0,
                 e3 = k == 3 ? 1.0 : 0.0;
    double br2[NCT], bdn[NCT], bdf[NCT], cr2[NCT], cdn[NCT], Q[NCT], ax[NCT], ay[NCT], az[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
        const int idx = 16 * ct + j;
        const int64_t t = G.tstart + (idx < G.tcount ? idx : 0);
        const double x = coords[3 * t] - cx, y = coords[3 * t + 1] - cy, z = coords[3 * t + 2] - cz;
        const double nx = norms[3 * t], ny = norms[3 * t + 1], nz = norms[3 * t + 2];
        br2[ct] = __builtin_fma(-2.0 * e0, x, __builtin_fma(-2.0 * e1, y, __builtin_fma(-2.0 * e2, z, e3)));
        bdn[ct] = -__builtin_fma(e0, nx, __builtin_fma(e1, ny, e2 * nz));
        bdf[ct] = __builtin_fma(e0, x, __builtin_fma(e1, y, __builtin_fma(e2, z, e3)));
        cr2[ct] = __builtin_fma(x, x, __builtin_fma(y, y, z * z));
        cdn[ct] = __builtin_fma(x, nx, __builtin_fma(y, ny, z * nz));
        Q[ct] = ax[ct] = ay[ct] = az[ct] = 0.0;
    }
    const d4 zero{0.0, 0.0, 0.0, 0.0};
    for (int64_t q = G.iv_begin; q < G.iv_end; q++) {
        const int64_t b = iv[q].begin, e = iv[q].end;
        for (int64_t s0 = b; s0 < e; s0 += 16) {
            // matrix row j of this tile is source s0 + j; rows past the interval
            // end repeat its last source with zero density (q = 0 for them)
            const int64_t si = s0 + j;
            const bool valid = si < e;
            const double *sp = src + 8 * (valid ? si : e - 1);
            const double sx = sp[0] - cx, sy = sp[1] - cy, sz = sp[2] - cz;
            const double fx = valid ? sp[3] : 0.0, fy = valid ? sp[4] : 0.0, fz = valid ? sp[5] : 0.0;
            const double ss = __builtin_fma(sx, sx, __builtin_fma(sy, sy, sz * sz));
            const double sf = __builtin_fma(sx, fx, __builtin_fma(sy, fy, sz * fz));
            const double ag = __builtin_fma(e0, sx, __builtin_fma(e1, sy, __builtin_fma(e2, sz, e3 * ss)));
            const double af = __builtin_fma(e0, fx, __builtin_fma(e1, fy, __builtin_fma(e2, fz, -e3 * sf)));
            const d4 px = mma(ag, e0, zero), py = mma(ag, e1, zero), pz = mma(ag, e2, zero);
            tile_terms<NCT>(ag, af, br2, bdn, bdf, cr2, cdn, px, py, pz, Q, ax, ay, az);
        }
    }
    // reduce the four lane groups (rows 4k..4k+3) in a fixed order
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
        Q[ct] += __shfl_xor(Q[ct], 16);
        ax[ct] += __shfl_xor(ax[ct], 16);
        ay[ct] += __shfl_xor(ay[ct], 16);
        az[ct] += __shfl_xor(az[ct], 16);
        Q[ct] += __shfl_xor(Q[ct], 32);
        ax[ct] += __shfl_xor(ax[ct], 32);
        ay[ct] += __shfl_xor(ay[ct], 32);
        az[ct] += __shfl_xor(az[ct], 32);
    }
    if (k != 0) return;
    if (!accumulate) out += G.slot * slot_stride;
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
        const int idx = 16 * ct + j;
        if (idx < G.tcount) {
            const int64_t t = G.tstart + idx;
            const double x = coords[3 * t] - cx, y = coords[3 * t + 1] - cy, z = coords[3 * t + 2] - cz;
            const double c = kTracW * coef;
            const double tx = __builtin_fma(x, Q[ct], -ax[ct]);
            const double ty = __builtin_fma(y, Q[ct], -ay[ct]);
            const double tz = __builtin_fma(z, Q[ct], -az[ct]);
            double *o = out + 3 * t;
            if (accumulate) {
                o[0] += c * tx;
                o[1] += c * ty;
                o[2] += c * tz;
            } else {
                o[0] = c * tx;
                o[1] = c * ty;
                o[2] = c * tz;
            }
        }
    }
}

void launch_far_trac_mfma_f64(const double *src, const double *tcoords, const double *tnorms, const double *gcen,
                              int64_t ntrg, int ngroups, const int *go, const FarGroup *gr, const FarInterval *iv,
                              const int *active, double *out, double coef, int acc, hipStream_t s) {
    if (ngroups <= 0) return;
    k_far_trac_mfma_f64<4><<<dim3(ngroups), dim3(64), 0, s>>>(src, tcoords, tnorms, gcen, go, gr, iv, active, out,
                                                               coef, acc, 3 * ntrg);
}

} // namespace sphstk
