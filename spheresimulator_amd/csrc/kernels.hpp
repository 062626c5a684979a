// Device views and launch wrappers of the HIP kernels in kernels.hip.
#pragma once

#include "common.hpp"

namespace sphstk {

constexpr int kMaxOrder = 32;

// Read-only device view of one (possibly batched) operator's geometry.
// Spheres of one order, at most kTileSph of them (indices tile_sph[first ..
// first + count)), processed together by the tiled per-sphere matrix kernels.
constexpr int kTileSph = 16;
struct SphTile {
    int32_t p, first, count, pad;
};

struct DevGeom {
    int nsph = 0, nsusp = 0;
    int ntiles = 0;
    const SphTile *tiles = nullptr; // spheres grouped by order
    const int *tile_sph = nullptr;
    int64_t npts = 0;
    const double *pos = nullptr;   // 3N
    const double *rad = nullptr;   // N
    const double *quat = nullptr;  // 4N (w,x,y,z), unit
    const int *order = nullptr;    // N
    const int *sph_susp = nullptr; // N
    const int64_t *pt_off = nullptr;   // N+1
    const int64_t *coef_off = nullptr; // N+1 (6Q doubles per sphere)
    const int64_t *rec_off = nullptr;  // N+1 (2*kRecTrac*tri(p+2) doubles per sphere)
    const double *coords = nullptr;    // 3P lab
    const double *norms = nullptr;     // 3P lab
    const double *w = nullptr;         // P
    const int64_t *near_off = nullptr; // N+1
    const int *near_src = nullptr;
    const double *anal[kMaxOrder + 1] = {}; // analysis matrix per order (column-major)
};

// Chunks of the DOF vector used by the segmented BLAS-1 kernels.
struct Chunk {
    int64_t begin, end;
    int32_t seg, pad;
};

// Per-order resampling matrices on device (see Resample in common.hpp).
// per-order resampling matrices (column-major, 3 nout x 3 nin); M32: the same
// in FP32 (set for the FP32 inner operator's levels, whose resampling then runs
// on FP32 matrix cores with FP32 accumulation), else NULL
struct ResampleTabs {
    bool fp32 = false;
    const double *M[kMaxOrder + 1] = {};
    const float *M32[kMaxOrder + 1] = {};
    int nin[kMaxOrder + 1] = {};
    int nout[kMaxOrder + 1] = {};
};

// source records (x, y, z, w*in, 0, 0) and FP32 target records (x, n, 0, 0) of n points
void launch_pack_src_f64(int64_t n, const double *coords, const double *w, const double *in, double *src,
                         hipStream_t s);
void launch_pack_src_f32(int64_t n, const double *coords, const double *w, const double *in, float *src,
                         hipStream_t s);
void launch_pack_trg_f32(int64_t n, const double *coords, const double *norms, float *trg, hipStream_t s);

// far field over ntrg targets (FP64: tcoords/tnorms, FP32: trg32 records):
// accumulate=1: out += coef * sum over the listed pairs;
// accumulate=0: slot k of out (stride 3*ntrg) = coef * the sum over source slice k
void launch_far(int kind, int prec, int tpl, const void *src, const double *tcoords, const double *tnorms,
                const float *trg32, int64_t ntrg, int ngroups, const int *group_order, const FarGroup *groups,
                const FarInterval *iv, const int *active, double *out, double coef, int accumulate, hipStream_t s,
                bool levels = false);

// Equivalent point forces of the far-field cells: the force words of eq
// (FP32 source records, K per cell) = rho_c P u_c (P 3K x 3Kc column-major, u
// 3Kc check velocities per cell); cells of inactive suspensions are skipped.
void launch_equiv_forces(int ncell, int K, int Kc, const double *P, const double *u, const double *rho,
                         const int *susp, const int *active, float *eq, hipStream_t s);

// out_S (+)= R_S M_p R_S^-1 in_S between two point sets with per-sphere offsets
void launch_resample(const DevGeom &g, const ResampleTabs &tab, const int64_t *in_off, const int64_t *out_off,
                     const double *in, double *out, int accumulate, const int *active, hipStream_t s,
                     int max_points);

void launch_analysis(const DevGeom &g, const double *in, double *coef, const int *active, hipStream_t s,
                     int max_order);
// per sphere and degree n: (F_r, F_s) <- 2x2 block tab[5n..5n+3], F_t <- tab[5n+4] F_t (in place)
void launch_degree_blocks(const DevGeom &g, double *coef, const double *tab, const int *active, hipStream_t s);
// mode 0: single-layer record (kRecSL arrays), 1: traction record (kRecTrac arrays)
void launch_expand(const DevGeom &g, int mode, const double *coef, double *rec, const int *active, hipStream_t s,
                   int max_order);
// One self/near evaluation work item: targets [tstart, tstart+tcount) (<= 64) of
// sphere trg against the expansion of sphere src; its contribution goes to
// scratch[slot ...] (3 doubles per target).
struct EvalItem {
    int32_t trg, src;
    int32_t tstart, tcount;
    int64_t slot;
};

// The same items grouped by source sphere: source S owns items
// [src_off[S], src_off[S + 1]).  The traction evaluation (mode 1) runs one
// workgroup per source with the source's record staged in LDS, one launch per
// sphere order (LDS sized to that order's record): srcs lists the spheres by
// order, those of order p at [order_off[p], order_off[p + 1]).
struct SrcItems {
    const EvalItem *items = nullptr;
    const int64_t *src_off = nullptr;
    const int *srcs = nullptr;
    int order_off[kMaxOrder + 2] = {};
};

// out += cSL*S_self/near + cTrac*K_self/near + cId*in   (mode 0: single layer, 1: traction)
void launch_sheval(const DevGeom &g, int mode, const double *rec, const EvalItem *items, int nitems,
                   const int64_t *trg_slot_off, const int *trg_nsrc, double *scratch, const double *in, double *out,
                   double cSL, double cTrac, double cId, const int *active, hipStream_t s,
                   const SrcItems *by_src = nullptr);
// the two halves of launch_sheval: the per-item evaluations into scratch, and
// the per-target gather (out += sum of the target's items + cId in)
void launch_sheval_items(const DevGeom &g, int mode, const double *rec, const EvalItem *items, int nitems,
                         double *scratch, double cSL, double cTrac, const int *active, hipStream_t s,
                         const SrcItems *by_src = nullptr);
void launch_sheval_gather(const DevGeom &g, int nitems, const int64_t *trg_slot_off, const int *trg_nsrc,
                          const double *scratch, const double *in, double *out, double cId, const int *active,
                          hipStream_t s);
void launch_identity(const DevGeom &g, const double *in, double *out, double cId, const int *active,
                     hipStream_t s);
void launch_lop(const DevGeom &g, const double *in, double *out, double cLOP, const int *active, hipStream_t s);
// out = P in (grid values of the band-limited projection) from a traction record (self items only)
void launch_synthesis(const DevGeom &g, const double *rec, const EvalItem *self_items, int nitems, double *out,
                      hipStream_t s, const int *active = nullptr);

// segmented BLAS-1 over chunks (one segment per suspension)
void launch_seg_dot(const double *a, const double *b, const Chunk *chunks, int nchunks, double *partial,
                    const int *active, hipStream_t s);
// out[seg] = sum of partials of seg; op 0: sum, 1: 1/sqrt(sum) (0 if sum==0), 2: -sum
void launch_seg_reduce(const double *partial, const int *seg_chunk_off, int nseg, double *out, int op,
                       const int *active, hipStream_t s);
// y += sign * alpha[seg] * x
void launch_seg_axpy(double *y, const double *x, const double *alpha, double sign, const Chunk *chunks, int nchunks,
                     const int *active, hipStream_t s);
// y = alpha[seg] * x
void launch_seg_scale(double *y, const double *x, const double *alpha, const Chunk *chunks, int nchunks,
                      const int *active, hipStream_t s);
// x += sum_i coef[i*nseg + seg] * V[i]   (i < k)
// partial[j * nchunks + chunk] = w . V_j over the chunk, j < k <= 64 (one pass over w)
void launch_seg_multi_dot(const double *w, const double *V, int64_t ldv, int k, const Chunk *chunks, int nchunks,
                          double *partial, const int *active, hipStream_t s);
// out[j * nseg + seg] = sum of seg's chunk partials (op 2: negated), j < k
void launch_seg_multi_reduce(const double *partial, int nchunks, const int *off, int nseg, int k, double *out, int op,
                             const int *active, hipStream_t s);
void launch_seg_multi_axpy(double *x, const double *V, int64_t ldv, const double *coef, int k, int nseg,
                           const Chunk *chunks, int nchunks, const int *active, hipStream_t s);
// x = sum_j coef[j nseg + seg] V_j over the active segments (x overwritten there)
void launch_seg_multi_comb(double *x, const double *V, int64_t ldv, const double *coef, int k, int nseg,
                           const Chunk *chunks, int nchunks, const int *active, hipStream_t s);
// GMRES least-squares bookkeeping on the device, one Krylov space per
// suspension (arrays: H B x (m+1) x m, cs/sn B x m, g B x (m+1), per-suspension
// ints and doubles of length B): cycle start (g = beta e_1, cur = cyc), the
// Givens update of column j from the orthogonalisation products hh (active set
// cur shrinks as suspensions converge; *nact = those still active), and the
// back substitution into coef[i B + s].
// residual coefficients of a finished cycle in its basis (m + 1 per suspension, coef[i B + s])
void launch_gmres_resid_coef(int B, int m, const double *g, const double *cs, const double *sn, const int *kdim,
                             const int *cyc, double *coef, hipStream_t s);
void launch_gmres_init(int B, int m, const double *beta, const int *cyc, double *g, int *cur, int *kdim, int *impok,
                       hipStream_t s);
void launch_gmres_givens(int B, int j, int m, const double *hh, double hsign, double *H, double *cs, double *sn,
                         double *g, const double *bn, double tol, int *cur, int *kdim, int *impok, double *relres,
                         int *nact, hipStream_t s);
void launch_gmres_backsub(int B, int m, const double *H, const double *g, const int *kdim, const int *cyc,
                          double *coef, hipStream_t s);
// out += sum of nslot partial-sum slots (far-field source slices)
void launch_slot_reduce(double *out, const double *part, int nslot, int64_t stride, const Chunk *chunks,
                        int nchunks, const int *active, hipStream_t s);
// y = a*x + b*y (plain, whole vector)
void launch_axpby(int64_t n, double a, const double *x, double b, double *y, hipStream_t s);
void launch_fill(int64_t n, double v, double *y, hipStream_t s);

// FP64 traction far field of ncol (1..8) densities on one geometry: `in` holds
// ncol DOF vectors col_stride apart, of which the first n0 points carry the
// geometry and work lists; src is scratch of 4 (ncol + 1) doubles per point;
// out (or its split slots, slot_stride apart) has the same column layout.
void launch_far_trac_multi(int ncol, int64_t n0, const double *coords, const double *norms, const double *w,
                           const double *in, int64_t col_stride, double *src, int ngroups, const int *go,
                           const FarGroup *gr, const FarInterval *iv, const int *active, int nsusp0, double *out,
                           double coef, int acc, int64_t slot_stride, hipStream_t s);
// sphere-centre far field (Faxen) of packed FP64 sources, 9 sums per centre and slot
// out (+)= coef K'[rho0] at grid points from distant spheres' rigid-body densities, in closed form (Stokeslet,
// degenerate quadrupole and rotlet at the centres); force holds (F, T) per sphere
void launch_far_rigid(const double *coords, const double *norms, const double *pos, const double *force,
                      const double *rad, int64_t ntrg, int ngroups, const int *go, const FarGroup *gr,
                      const FarInterval *iv, double *out, double coef, int acc, hipStream_t s);
void launch_far_centre(const double *src, const double *centres, int ngroups, const int *go, const FarGroup *gr,
                       const FarInterval *iv, double *part, int64_t ntrg, hipStream_t s);
void launch_centre_finish(int nsph, int nslot, const double *part, const double *rad, const double *vel_ns,
                          double inv_mu, double *vel, hipStream_t s);
void launch_rho0(const DevGeom &g, const double *force, double *rho, hipStream_t s);
void launch_rigid_avg(const DevGeom &g, const double *u, double inv_mu, double *vel, hipStream_t s);

} // namespace sphstk
