// Host engine: operator set-up (grid, near list, far-field work lists and the
// level sets of the multi-level far fields, the tile-shape autotuner), the
// operator apply pipeline (and its captured HIP graph), batched restarted GMRES
// and the mobility solve in FP64 and in mixed precision (FP32 multi-level inner
// operator inside FP64 iterative refinement with multi-level FP64 residuals).
#include "engine.hpp"
#include "steps.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <functional>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <unordered_map>

namespace sphstk {

namespace {

constexpr int kMaxSupportedOrder = 20; // LDS of the expansion kernel bounds the order
constexpr int64_t kChunk = 4096;  // DOF per block of the segmented vector kernels
constexpr int64_t kWantItems = 40000; // far-field work items wanted when a problem is small
constexpr int64_t kLookaheadDof = 300000; // GMRES below this many unknowns launches one iteration ahead
constexpr double kTuneBudget = 3.0;    // seconds per autotuning pass (SPHSTK_FLAG_AUTOTUNE)

// SPHSTK_VERBOSE=1: set-up phase times on stderr
static bool setup_verbose() {
    static const bool v = [] {
        const char *e = std::getenv("SPHSTK_VERBOSE");
        return e && std::atoi(e) != 0;
    }();
    return v;
}
struct PhaseTimer {
    const char *what;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    explicit PhaseTimer(const char *w) : what(w) {}
    ~PhaseTimer() {
        if (setup_verbose())
            std::fprintf(stderr, "[sphstk] setup %s: %.3f s\n", what,
                         std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
};

inline void qrot_host(const double *q, const double *v, double *o) {
    const double w = q[0], ux = q[1], uy = q[2], uz = q[3];
    const double tx = 2 * (uy * v[2] - uz * v[1]), ty = 2 * (uz * v[0] - ux * v[2]), tz = 2 * (ux * v[1] - uy * v[0]);
    o[0] = v[0] + w * tx + (uy * tz - uz * ty);
    o[1] = v[1] + w * ty + (uz * tx - ux * tz);
    o[2] = v[2] + w * tz + (ux * ty - uy * tx);
}

using Mask = std::array<uint64_t, kMaxTPL>;

} // namespace

static std::vector<Chunk> make_chunks(int nsusp, const std::vector<int64_t> &susp_sph_off,
                                      const std::vector<int64_t> &off, std::vector<int> *seg_off);

// ---------------------------------------------------------------------------
// Operator set-up: SphereSTKSHOperator::setupDOF / setupFMM / setupNearEval
// (BIE/SphereSTKSHOperator.cpp:37-144, :539-573).

Operator::Operator(int nsusp, const int *counts, const double *pos, const double *radius, const double *quat,
                   const int *order, double cId, double cSL, double cTrac, double cLOP, int flags, int ncopy)
    : nsph_(0), nsusp_(nsusp), npts_(0), ncopy_(ncopy), flags_(flags) {
    PhaseTimer pt("operator");
    c_[0] = cId;
    c_[1] = cSL;
    c_[2] = cTrac;
    c_[3] = cLOP;
    SPHSTK_REQUIRE((flags & ~255) == 0, "unknown flags (known: TIMING 1, FP32_DIRECT 2, AUTOTUNE 4, GRAPH 8, "
                                        "INNER_PRECOND 16, NO_CELLS 32, EXPLICIT_REFINE 64, NO_COARSE 128)");
    SPHSTK_REQUIRE(nsusp >= 1, "number of suspensions must be >= 1");
    SPHSTK_REQUIRE(ncopy >= 1 && ncopy <= 64 && nsusp % ncopy == 0, "copies must divide the suspensions");
    SPHSTK_REQUIRE(counts != nullptr, "sphere counts must not be NULL");
    susp_sph_off_.assign(nsusp + 1, 0);
    for (int b = 0; b < nsusp; b++) {
        SPHSTK_REQUIRE(counts[b] >= 0, "sphere count must be >= 0");
        susp_sph_off_[b + 1] = susp_sph_off_[b] + counts[b];
    }
    nsph_ = (int)susp_sph_off_[nsusp];
    SPHSTK_REQUIRE(nsph_ == 0 || (pos && radius && quat && order), "input arrays must not be NULL");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw Error(-2, "no HIP device available");
    SPHSTK_HIP(hipGetDevice(&device_));

    pos_.assign(pos, pos + 3 * (size_t)nsph_);
    rad_.assign(radius, radius + nsph_);
    order_.assign(order, order + nsph_);
    quat_.resize(4 * (size_t)nsph_);
    sph_susp_.resize(nsph_);
    for (int b = 0; b < nsusp; b++)
        for (int64_t i = susp_sph_off_[b]; i < susp_sph_off_[b + 1]; i++) sph_susp_[i] = b;
    for (int i = 0; i < nsph_; i++) {
        SPHSTK_REQUIRE(std::isfinite(pos_[3 * i]) && std::isfinite(pos_[3 * i + 1]) && std::isfinite(pos_[3 * i + 2]),
                       "sphere position is not finite (sphere " + std::to_string(i) + ")");
        SPHSTK_REQUIRE(std::isfinite(rad_[i]) && rad_[i] > 0, "sphere radius must be > 0 (sphere " +
                                                                   std::to_string(i) + ")");
        SPHSTK_REQUIRE(order_[i] >= 1 && order_[i] <= kMaxSupportedOrder,
                       "SH order must be in [1, " + std::to_string(kMaxSupportedOrder) + "] (sphere " +
                           std::to_string(i) + ")");
        const double *q = quat + 4 * i;
        const double nq = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        SPHSTK_REQUIRE(std::isfinite(nq) && nq > 0, "quaternion must be non-zero (sphere " + std::to_string(i) + ")");
        for (int c = 0; c < 4; c++) quat_[4 * i + c] = q[c] / nq;
        max_order_ = std::max(max_order_, order_[i]);
    }

    // DOF layout (setupDOF, cpp:51-80): points per sphere (p+1)(2p+2), spectral 3(p+1)(p+2)
    pt_off_.assign(nsph_ + 1, 0);
    coef_off_.assign(nsph_ + 1, 0);
    rec_off_.assign(nsph_ + 1, 0);
    for (int i = 0; i < nsph_; i++) {
        const int p = order_[i];
        pt_off_[i + 1] = pt_off_[i] + grid_points(p);
        coef_off_[i + 1] = coef_off_[i] + 3 * (p + 1) * (p + 2);
        rec_off_[i + 1] = rec_off_[i] + 2 * kRecTrac * tri_size(p + 2);
    }
    npts_ = pt_off_[nsph_];
    coords_.resize(3 * npts_);
    norms_.resize(3 * npts_);
    w_.resize(npts_);
    pt_sph_.resize(npts_);
    for (int i = 0; i < nsph_; i++) {
        const ShTables &t = sh_tables(order_[i]);
        const double R = rad_[i];
        for (int q = 0; q < t.npts; q++) {
            const int64_t g = pt_off_[i] + q;
            double nv[3];
            qrot_host(&quat_[4 * i], &t.eta[3 * q], nv);
            for (int c = 0; c < 3; c++) {
                norms_[3 * g + c] = nv[c];
                coords_[3 * g + c] = pos_[3 * i + c] + R * nv[c];
            }
            w_[g] = R * R * t.w[q];
            pt_sph_[g] = i;
        }
    }

    // near pairs: 0 < |c_s - c_t|^2 < (1.5 R_s + 1.5 R_t)^2 within each suspension
    // (MPI/NearInteraction.txx:430-441, radiusNear = 1.5 R at cpp:547), grouped
    // by target and sorted (:527), found on the device over a cell grid (which
    // replaces the Morton octree); overlapping spheres are rejected there
    stream_ = stream_owner_.create();
    aux_ = aux_owner_.create();
    SPHSTK_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    SPHSTK_HIP(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
    d_pos_.upload(pos_);
    d_rad_.upload(rad_);
    d_sph_susp_.upload(sph_susp_);
    {
        PhaseTimer pn("near pairs");
        ::sphstk::near_pairs(nsph_, pos_.data(), rad_.data(), d_pos_.p, d_rad_.p, d_sph_susp_.p, stream_, near_off_,
                             near_src_);
    }

    // self/near evaluation work items: per target sphere its own expansion first,
    // then its near sources in list order, each cut into chunks of 64 targets;
    // the same items grouped by source for the traction evaluation
    {
        std::vector<EvalItem> items, selfi;
        std::vector<std::vector<EvalItem>> by_src(nsph_);
        std::vector<int64_t> tso(nsph_ + 1, 0);
        std::vector<int> tns(nsph_, 0);
        int64_t slot = 0;
        for (int t = 0; t < nsph_; t++) {
            const int nT = (int)(pt_off_[t + 1] - pt_off_[t]);
            tso[t] = slot;
            tns[t] = 1 + (int)(near_off_[t + 1] - near_off_[t]);
            for (int k = 0; k < tns[t]; k++) {
                const int s = (k == 0) ? t : near_src_[near_off_[t] + k - 1];
                for (int j = 0; j < nT; j += 64) {
                    EvalItem it{t, s, j, std::min(64, nT - j), slot + 3 * (int64_t)j};
                    items.push_back(it);
                    by_src[s].push_back(it);
                    if (k == 0) selfi.push_back(it);
                }
                slot += 3 * (int64_t)nT;
            }
        }
        tso[nsph_] = slot;
        items_.upload(items);
        self_items_.upload(selfi);
        std::vector<EvalItem> si;
        std::vector<int64_t> soff(nsph_ + 1, 0);
        si.reserve(items.size());
        for (int q = 0; q < nsph_; q++) {
            si.insert(si.end(), by_src[q].begin(), by_src[q].end());
            soff[q + 1] = (int64_t)si.size();
        }
        std::vector<int> sl;
        sl.reserve(nsph_);
        for (int p = 0; p <= kMaxOrder; p++) {
            src_items_.order_off[p] = (int)sl.size();
            for (int q = 0; q < nsph_; q++)
                if (order_[q] == p) sl.push_back(q);
        }
        src_items_.order_off[kMaxOrder + 1] = (int)sl.size();
        d_src_items_.upload(si);
        d_src_off_.upload(soff);
        d_src_list_.upload(sl);
        src_items_.items = d_src_items_.p;
        src_items_.src_off = d_src_off_.p;
        src_items_.srcs = d_src_list_.p;
        d_trg_slot_off_.upload(tso);
        d_trg_nsrc_.upload(tns);
        evscratch_.alloc(std::max<int64_t>(slot, 1));
    }

    // device state
    d_quat_.upload(quat_);
    d_order_.upload(order_);
    d_pt_off_.upload(pt_off_);
    d_coef_off_.upload(coef_off_);
    d_rec_off_.upload(rec_off_);
    d_coords_.upload(coords_);
    d_norms_.upload(norms_);
    d_w_.upload(w_);
    d_near_off_.upload(near_off_);
    d_near_src_.upload(near_src_);
    geom_.nsph = nsph_;
    geom_.nsusp = nsusp_;
    geom_.npts = npts_;
    geom_.pos = d_pos_.p;
    geom_.rad = d_rad_.p;
    geom_.quat = d_quat_.p;
    geom_.order = d_order_.p;
    geom_.sph_susp = d_sph_susp_.p;
    geom_.pt_off = d_pt_off_.p;
    geom_.coef_off = d_coef_off_.p;
    geom_.rec_off = d_rec_off_.p;
    geom_.coords = d_coords_.p;
    geom_.norms = d_norms_.p;
    geom_.w = d_w_.p;
    geom_.near_off = d_near_off_.p;
    geom_.near_src = d_near_src_.p;
    {
        // spheres grouped by order, in index order within an order, tiles of kTileSph
        std::vector<int> ts(nsph_);
        std::iota(ts.begin(), ts.end(), 0);
        std::stable_sort(ts.begin(), ts.end(), [&](int a, int b) { return order_[a] < order_[b]; });
        std::vector<SphTile> tl;
        for (int i = 0; i < nsph_;) {
            int j = i;
            while (j < nsph_ && j - i < kTileSph && order_[ts[j]] == order_[ts[i]]) j++;
            tl.push_back(SphTile{order_[ts[i]], i, j - i, 0});
            i = j;
        }
        d_tile_sph_.upload(ts);
        d_tiles_.upload(tl);
        geom_.ntiles = (int)tl.size();
        geom_.tiles = d_tiles_.p;
        geom_.tile_sph = d_tile_sph_.p;
    }
    std::vector<bool> used(kMaxOrder + 1, false);
    for (int p : order_) used[p] = true;
    for (int p = 0; p <= kMaxOrder; p++)
        if (used[p]) {
            d_anal_[p].upload(sh_tables(p).anal);
            geom_.anal[p] = d_anal_[p].p;
        }

    // segmented-vector chunks (one segment per suspension)
    std::vector<int> sco;
    const std::vector<Chunk> ch = make_chunks(nsusp_, susp_sph_off_, pt_off_, &sco);
    chunks_.upload(ch);
    seg_chunk_off_.upload(sco);
    seg_chunk_off_host_ = sco;
    partial_.alloc(std::max<size_t>(ch.size(), 1));

    int tpl = 1, split = 0;
    // FP32 inner operator: orders from level_order().  FP64 outer applies: fine
    // below s' = 4.8, then orders 10, 9, 8, 7, 6, 5 from s' = 4.8, 5.8, 7.35,
    // 10.05, 15.1, 26 (the radius-aware separation of ensure_levels()), where a
    // level's worst truncation ratio (2 s' - 1)^-(q + 1) is 5.6e-11 to 5.9e-11.
    ls32_.alo[0] = 3.0, ls32_.alo[1] = 6.0, ls32_.alo[2] = 9.0, ls32_.alo[3] = 16.0;
    ls32_.order[0] = ls32_.order[1] = ls32_.order[2] = ls32_.order[3] = -1;
    ls32_.fp64 = false;
    ls32_.cells = (flags & 32) == 0;
    ls64_.nlev = 6;
    {
        const double a[6] = {4.8, 5.8, 7.35, 10.05, 15.1, 26.0};
        const int q[6] = {10, 9, 8, 7, 6, 5};
        std::copy(a, a + 6, ls64_.alo);
        std::copy(q, q + 6, ls64_.order);
    }
    ls64_.fp64 = true;
    // coarse FP32 operator of the mixed form's preconditioning solves: orders
    // 5, 4, 3, 2 from 2, 3, 5, 8 (R_s + R_t) (capped at p), source cells of 32
    // equivalent point forces from 2.5 cell radii (on the check sphere)
    lsc_.alo[0] = 2.0, lsc_.alo[1] = 3.0, lsc_.alo[2] = 5.0, lsc_.alo[3] = 8.0;
    lsc_.order[0] = 5, lsc_.order[1] = 4, lsc_.order[2] = 3, lsc_.order[3] = 2;
    lsc_.fp64 = false;
    lsc_.cells = ls32_.cells;
    lsc_.cell_theta = 2.5;
    lsc_.equiv_k = 32;
    if (const char *v = std::getenv("SPHSTK_COARSE_ALPHA")) {
        const int got = std::sscanf(v, "%lf,%lf,%lf,%lf:%d,%d,%d,%d", &lsc_.alo[0], &lsc_.alo[1], &lsc_.alo[2],
                                    &lsc_.alo[3], &lsc_.order[0], &lsc_.order[1], &lsc_.order[2], &lsc_.order[3]);
        if (got != 4 && got != 8) throw Error(-1, "SPHSTK_COARSE_ALPHA: a0,a1,a2,a3[:p0,p1,p2,p3]");
    }
    // separations, optionally followed by orders (-1: the order-dependent default)
    if (const char *v = std::getenv("SPHSTK_LEVEL_ALPHA")) {
        const int got = std::sscanf(v, "%lf,%lf,%lf,%lf:%d,%d,%d,%d", &ls32_.alo[0], &ls32_.alo[1], &ls32_.alo[2],
                                    &ls32_.alo[3], &ls32_.order[0], &ls32_.order[1], &ls32_.order[2],
                                    &ls32_.order[3]);
        if (got != 4 && got != 8) throw Error(-1, "SPHSTK_LEVEL_ALPHA: a0,a1,a2,a3[:p0,p1,p2,p3]");
    }
    for (const LevelSet *S : {&ls32_, &ls64_, &lsc_}) {
        SPHSTK_REQUIRE(S->alo[0] >= 1.5, "level separations must start at >= 1.5");
        SPHSTK_REQUIRE(S->nlev >= 1 && S->nlev <= kMaxLevels && (!S->cells || S->nlev == kNumLevels),
                       "internal: level count");
        for (int l = 1; l < S->nlev; l++)
            SPHSTK_REQUIRE(S->alo[l] > S->alo[l - 1], "level separations must increase");
        for (int l = 0; l < S->nlev; l++)
            SPHSTK_REQUIRE(S->order[l] == -1 || (S->order[l] >= 1 && S->order[l] <= kMaxOrder),
                           "level orders must lie in [1, max order]");
    }
    multi_level_ = (flags_ & 2) == 0;
    {
        PhaseTimer pt2("fine far-field lists");
        std::vector<std::vector<int>> excl(nsph_);
        for (int t = 0; t < nsph_; t++) {
            excl[t].push_back(t);
            for (int64_t k = near_off_[t]; k < near_off_[t + 1]; k++) excl[t].push_back(near_src_[k]);
            std::sort(excl[t].begin(), excl[t].end());
        }
        runs_fine_.assign(nsph_, Runs{});
        for (int t = 0; t < nsph_; t++) {
            const int b = sph_susp_[t];
            runs_fine_[t] = runs_of(excl[t], (int)susp_sph_off_[b], (int)susp_sph_off_[b + 1], false);
        }
        if (ncopy_ > 1) {
            // work lists of the first copy only; its geometry serves all copies
            const int b0 = nsusp_ / ncopy_, s0 = (int)susp_sph_off_[b0];
            for (int c = 1; c < ncopy_; c++)
                for (int i = 0; i < s0; i++) {
                    const int j = c * s0 + i;
                    SPHSTK_REQUIRE(order_[j] == order_[i] && rad_[j] == rad_[i] && pos_[3 * j] == pos_[3 * i] &&
                                       pos_[3 * j + 1] == pos_[3 * i + 1] && pos_[3 * j + 2] == pos_[3 * i + 2] &&
                                       std::equal(&quat_[4 * j], &quat_[4 * j + 4], &quat_[4 * i]),
                                   "copies of the suspensions must be identical");
                }
            const std::vector<int64_t> so0(susp_sph_off_.begin(), susp_sph_off_.begin() + b0 + 1),
                po0(pt_off_.begin(), pt_off_.begin() + s0 + 1);
            const std::vector<Runs> r0(runs_fine_.begin(), runs_fine_.begin() + s0);
            build_far_lists(lists_fine_, b0, so0, po0, po0, r0, 1, split, kWantItems);
            if (lists_fine_.split > 1) lists_fine_.part.alloc(size_t(lists_fine_.split) * 3 * npts_);
            srcm_.alloc(std::max<int64_t>(4 * (std::min(ncopy_, kMultiCols) + 1) * pt_off_[s0], 8));
        } else {
            build_far_lists(lists_fine_, nsusp_, susp_sph_off_, pt_off_, pt_off_, runs_fine_, tpl, split, kWantItems);
        }
    }
    ls32_.tpl = ls64_.tpl = lsc_.tpl = tpl;

    src64_.alloc(std::max<int64_t>(8 * npts_, 8));
    coef_.alloc(std::max<int64_t>(coef_off_[nsph_], 1));
    rec_.alloc(std::max<int64_t>(rec_off_[nsph_], 1));
    tmp_.alloc(std::max<int64_t>(3 * npts_, 1));
    if ((flags_ & 4) && ncopy_ == 1) {
        double budget = kTuneBudget;
        autotune_fine(budget);
    }
}

// ---------------------------------------------------------------------------
// Tile-shape autotuner (SPHSTK_FLAG_AUTOTUNE).  The far-field kernels take two
// shape parameters: targets per lane (work items of 64, 128 or 256 targets)
// and the number of source slices per item (more, smaller items balance
// better; each slice adds a partial-sum pass).  Candidates are timed on this
// problem's own work lists with a unit density, two launches each (the first
// warms caches and code), within a time budget so that very large problems
// keep the default shapes instead of spending minutes tuning.
double Operator::time_far_pass(FarLists &L, int prec, const void *src, const double *tcoords, const double *tnorms,
                               const float *trg32, const Chunk *chunks, int nchunks, double *out) {
    hipEvent_t a, b;
    SPHSTK_HIP(hipEventCreate(&a));
    SPHSTK_HIP(hipEventCreate(&b));
    far_run(L, 0, prec, src, tcoords, tnorms, trg32, chunks, nchunks, 1.0, out, nullptr);
    SPHSTK_HIP(hipEventRecord(a, stream_));
    far_run(L, 0, prec, src, tcoords, tnorms, trg32, chunks, nchunks, 1.0, out, nullptr);
    SPHSTK_HIP(hipEventRecord(b, stream_));
    SPHSTK_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    SPHSTK_HIP(hipEventElapsedTime(&ms, a, b));
    SPHSTK_HIP(hipEventDestroy(a));
    SPHSTK_HIP(hipEventDestroy(b));
    return ms;
}

void Operator::autotune_fine(double &budget_s) {
    if (nsph_ == 0) return;
    launch_fill(3 * npts_, 1.0, tmp_.p, stream_);
    launch_pack_src_f64(npts_, d_coords_.p, d_w_.p, tmp_.p, src64_.p, stream_);
    DevBuf<double> out;
    out.alloc(std::max<int64_t>(3 * npts_, 1));
    auto timed = [&](FarLists &L) {
        const double ms = time_far_pass(L, 0, src64_.p, d_coords_.p, d_norms_.p, nullptr, chunks_.p,
                                        (int)chunks_.n, out.p);
        budget_s -= 2e-3 * ms;
        return ms;
    };
    double best = timed(lists_fine_);
    char buf[160];
    std::snprintf(buf, sizeof buf, "fine FP64: tpl %d split %d %.3f ms", lists_fine_.tpl, lists_fine_.split, best);
    tune_log_ += buf;
    const int s0 = lists_fine_.split;
    for (int tpl : {1, 2, 4})
        for (int f : {1, 2, 4}) {
            const int split = std::max(1, std::min(64, f == 1 ? s0 : f == 2 ? s0 / 2 : 2 * s0));
            if (tpl == lists_fine_.tpl && split == lists_fine_.split) continue;
            if (budget_s < 2e-3 * best * 1.5) break;
            FarLists C;
            build_far_lists(C, nsusp_, susp_sph_off_, pt_off_, pt_off_, runs_fine_, tpl, split, kWantItems);
            const double ms = timed(C);
            std::snprintf(buf, sizeof buf, "; tpl %d split %d %.3f ms", tpl, split, ms);
            tune_log_ += buf;
            if (ms < best) {
                best = ms;
                lists_fine_ = std::move(C);
            }
        }
    ls64_.tpl = lists_fine_.tpl;
    std::snprintf(buf, sizeof buf, " -> tpl %d split %d\n", lists_fine_.tpl, lists_fine_.split);
    tune_log_ += buf;
    if (setup_verbose()) std::fprintf(stderr, "[sphstk] autotune %s", tune_log_.c_str());
}

// FP32 inner operator: targets per lane of all its level lists, timed as one
// multi-level pass.
void Operator::far_levels_pass(LevelSet &S, const double *in, double *out) {
    launch_pack_src_f32(npts_, d_coords_.p, d_w_.p, in, src32_.p, stream_);
    far_levels(S, 0, 1.0, in, out, nullptr, nullptr, nullptr);
}

void Operator::autotune_inner(double &budget_s) {
    if (nsph_ == 0 || !multi_level_) return;
    if (!trg32_.p) {
        trg32_.alloc(8 * std::max<int64_t>(npts_, 1));
        src32_.alloc(8 * std::max<int64_t>(npts_, 1));
        launch_pack_trg_f32(npts_, d_coords_.p, d_norms_.p, trg32_.p, stream_);
    }
    launch_fill(3 * npts_, 1.0, tmp_.p, stream_);
    DevBuf<double> out;
    out.alloc(std::max<int64_t>(3 * npts_, 1));
    auto timed = [&]() {
        auto t0 = std::chrono::steady_clock::now();
        far_levels_pass(ls32_, tmp_.p, out.p);
        hipEvent_t a, b;
        SPHSTK_HIP(hipEventCreate(&a));
        SPHSTK_HIP(hipEventCreate(&b));
        SPHSTK_HIP(hipEventRecord(a, stream_));
        far_levels_pass(ls32_, tmp_.p, out.p);
        SPHSTK_HIP(hipEventRecord(b, stream_));
        SPHSTK_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        SPHSTK_HIP(hipEventElapsedTime(&ms, a, b));
        SPHSTK_HIP(hipEventDestroy(a));
        SPHSTK_HIP(hipEventDestroy(b));
        budget_s -= std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return (double)ms;
    };
    ensure_levels(ls32_);
    const int t_def = ls32_.tpl;
    double best = timed();
    int best_tpl = t_def;
    char buf[160];
    std::snprintf(buf, sizeof buf, "FP32 levels: tpl %d %.3f ms", t_def, best);
    std::string log = buf;
    for (int tpl : {1, 2, 4}) {
        if (tpl == t_def) continue;
        if (budget_s < 0) break;
        LevelSet keep = std::move(ls32_);
        ls32_.tpl = tpl;
        ls32_.nlev = keep.nlev;
        std::copy(keep.alo, keep.alo + kMaxLevels, ls32_.alo);
        std::copy(keep.order, keep.order + kMaxLevels, ls32_.order);
        ls32_.fp64 = false;
        ls32_.cells = keep.cells;
        ls32_.ready = false;
        ls32_.levels.clear();
        auto tb = std::chrono::steady_clock::now();
        ensure_levels(ls32_);
        budget_s -= std::chrono::duration<double>(std::chrono::steady_clock::now() - tb).count();
        const double ms = timed();
        std::snprintf(buf, sizeof buf, "; tpl %d %.3f ms", tpl, ms);
        log += buf;
        if (ms < best) {
            best = ms;
            best_tpl = tpl;
        } else {
            ls32_ = std::move(keep);
        }
    }
    std::snprintf(buf, sizeof buf, " -> tpl %d\n", best_tpl);
    log += buf;
    tune_log_ += log;
    if (setup_verbose()) std::fprintf(stderr, "[sphstk] autotune %s", log.c_str());
}

Operator::~Operator() {
    if (ls64_build_.valid()) ls64_build_.wait();
    if (ls32_build_.valid()) ls32_build_.wait();
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (aux_) (void)hipStreamSynchronize(aux_);
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (gx_inner_) (void)hipGraphExecDestroy(gx_inner_);
    if (g_inner_) (void)hipGraphDestroy(g_inner_);
    for (auto &e : pending_) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (auto &e : free_events_) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
}

// Runs of the sorted sphere list `lst` within [s0, s1) (include) or of its
// complement there (exclude).
Runs runs_of(const std::vector<int> &lst, int s0, int s1, bool include) {
    Runs r;
    if (include) {
        for (int s : lst)
            if (s >= s0 && s < s1) run_append(r, s);
        return r;
    }
    int cur = s0;
    for (int s : lst) {
        if (s < cur || s >= s1) continue;
        if (s > cur) r.push_back({cur, s});
        cur = s + 1;
    }
    if (cur < s1) r.push_back({cur, s1});
    return r;
}

// Runs a parallel loop over [0, n) on the host's cores (at most 16 threads),
// handing out `chunk` indices at a time (0: automatic); serial below `serial_below`.
static void parallel_for(int n, const std::function<void(int, int)> &body, int chunk = 0, int serial_below = 256) {
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < serial_below || nt == 1) {
        body(0, n);
        return;
    }
    if (chunk <= 0) chunk = std::max(16, n / (8 * nt));
    std::vector<std::thread> th;
    std::atomic<int> next{0};
    for (int k = 0; k < std::min(nt, (n + chunk - 1) / chunk); k++)
        th.emplace_back([&] {
            for (;;) {
                const int a = next.fetch_add(chunk);
                if (a >= n) break;
                body(a, std::min(n, a + chunk));
            }
        });
    for (auto &t : th) t.join();
}

// Far-field work lists.  Targets (a point set with per-sphere offsets `toff`)
// are cut into groups of 64*tpl consecutive points of one suspension.  Each
// target sphere T interacts with the source spheres of `runs[T]` (ranges of
// sphere indices of the source point set `soff`, sorted, disjoint, within T's
// suspension): for the direct far field all spheres of the suspension minus T
// itself and its near spheres (the pairs the reference removes again after the
// FMM, cpp:332-348, :408-433, :641-690).  A group's source intervals are the
// point ranges of its spheres' runs, with a per-lane mask where the group
// spans spheres.  Each group's list is cut into `split` slices of nearly equal
// length, one work item each, launched in order of decreasing work.  With
// `sph_pos` (sphere centres) the groups are aligned to spheres instead (at most
// 64 points of one sphere each, so every interval applies to all lanes) and
// each group records its sphere's centre.
void build_far_lists(FarLists &L, int nsusp, const std::vector<int64_t> &susp_sph_off,
                     const std::vector<int64_t> &toff, const std::vector<int64_t> &soff, const std::vector<Runs> &runs,
                     int tpl, int split, int64_t want_items, const double *sph_pos,
                     const std::vector<int64_t> *src_susp_off) {
    const std::vector<int64_t> &sso = src_susp_off ? *src_susp_off : susp_sph_off;
    SPHSTK_REQUIRE(tpl == 1 || tpl == 2 || tpl == 4, "targets per lane must be 1, 2 or 4");
    SPHSTK_REQUIRE(split >= 0 && split <= 64, "source split must be in [0, 64] (0 = automatic)");
    const bool aligned = sph_pos != nullptr;
    SPHSTK_REQUIRE(!aligned || tpl == 1, "sphere-aligned far-field lists use one target per lane");
    const int nsph = (int)toff.size() - 1;
    SPHSTK_REQUIRE((int)runs.size() == nsph, "internal: one run list per target sphere");
    const int64_t ntrg = toff[nsph];
    std::vector<int32_t> tsph(ntrg);
    for (int i = 0; i < nsph; i++)
        for (int64_t k = toff[i]; k < toff[i + 1]; k++) tsph[k] = i;
    const int G = 64 * tpl;
    int64_t nbase = 0;
    for (int b = 0; b < nsusp; b++) {
        if (aligned) {
            for (int64_t T = susp_sph_off[b]; T < susp_sph_off[b + 1]; T++) nbase += (toff[T + 1] - toff[T] + G - 1) / G;
        } else {
            const int64_t np = toff[susp_sph_off[b + 1]] - toff[susp_sph_off[b]];
            nbase += (np + G - 1) / G;
        }
    }
    if (split == 0) {
        const int64_t nb = std::max<int64_t>(nbase, 1);
        split = (int)std::max<int64_t>(1, std::min<int64_t>(8, (want_items + nb - 1) / nb));
    }
    L.tpl = tpl;
    L.split = split;
    L.ntrg = ntrg;
    // the groups (first target, count, suspension), then their interval lists
    // built in parallel over chunks of consecutive groups and concatenated
    struct GSpec {
        int64_t t0;
        int cnt, b;
    };
    std::vector<GSpec> gs;
    for (int b = 0; b < nsusp; b++) {
        const int64_t p0 = toff[susp_sph_off[b]], p1 = toff[susp_sph_off[b + 1]];
        for (int64_t t0 = p0; t0 < p1;) {
            const int64_t tend = aligned ? toff[tsph[t0] + 1] : p1;
            const int cnt = (int)std::min<int64_t>(G, tend - t0);
            gs.push_back(GSpec{t0, cnt, b});
            t0 += cnt;
        }
    }
    struct Part {
        std::vector<FarGroup> groups;
        std::vector<FarInterval> ivs;
        std::vector<int64_t> work;
        std::vector<double> cen;
        double pairs = 0.0;
    };
    const int nparts = (int)std::min<size_t>(64, std::max<size_t>(1, gs.size() / 512));
    std::vector<Part> parts(nparts);
    Mask full;
    full.fill(~0ull);
    auto build_part = [&](int pi) {
        Part &P = parts[pi];
        const size_t g0 = gs.size() * pi / nparts, g1 = gs.size() * (pi + 1) / nparts;
        std::vector<FarInterval> giv;
        std::vector<std::pair<int, Mask>> ev;
        // groups inside one sphere all get that sphere's runs with full masks, so
        // consecutive groups of the same sphere share one copy of the intervals
        int last_T = -1;
        int64_t last_work = 0;
        std::vector<std::array<int64_t, 3>> last_slots; // iv_begin, iv_end, quota per slot
        for (size_t gi = g0; gi < g1; gi++) {
            const int64_t t0 = gs[gi].t0;
            const int cnt = gs[gi].cnt, b = gs[gi].b;
            if (tsph[t0] == tsph[t0 + cnt - 1] && tsph[t0] == last_T) {
                for (int sl = 0; sl < split; sl++) {
                    P.groups.push_back(FarGroup{t0, last_slots[sl][0], last_slots[sl][1], cnt, b, sl, 0});
                    P.work.push_back(last_slots[sl][2]);
                    if (aligned) {
                        for (int c = 0; c < 3; c++) P.cen.push_back(sph_pos[3 * last_T + c]);
                        P.cen.push_back(0.0);
                    }
                }
                P.pairs += double(last_work) * cnt;
                continue;
            }
            Mask act{};
            for (int i = 0; i < cnt; i++) act[i / 64] |= 1ull << (i % 64);
            giv.clear();
            int64_t gwork = 0;
            auto emit = [&](int64_t bb, int64_t ee, const Mask &m) {
                if (bb >= ee) return;
                int lanes = 0;
                for (int w = 0; w < tpl; w++) lanes += __builtin_popcountll(m[w] & act[w]);
                if (lanes == 0) return;
                if (!giv.empty() && giv.back().end == bb &&
                    std::memcmp(giv.back().mask, m.data(), sizeof(uint64_t) * kMaxTPL) == 0) {
                    giv.back().end = ee;
                } else {
                    FarInterval iv;
                    iv.begin = bb;
                    iv.end = ee;
                    std::memcpy(iv.mask, m.data(), sizeof(uint64_t) * kMaxTPL);
                    giv.push_back(iv);
                }
                P.pairs += double(ee - bb) * lanes;
                gwork += ee - bb;
            };
            const int Ta = tsph[t0], Tb = tsph[t0 + cnt - 1];
            if (Ta == Tb) {
                for (const auto &r : runs[Ta]) emit(soff[r.first], soff[r.second], full);
            } else {
                // sweep over the run boundaries of the group's spheres; the lane
                // sets of different spheres are disjoint, so toggling a sphere's
                // lanes at its run boundaries gives the covering mask
                ev.clear();
                for (int T = Ta; T <= Tb; T++) {
                    const int64_t lo = std::max(toff[T], t0) - t0, hi = std::min(toff[T + 1], t0 + cnt) - t0;
                    Mask lanes{};
                    for (int64_t i = lo; i < hi; i++) lanes[i / 64] |= 1ull << (i % 64);
                    for (const auto &r : runs[T]) {
                        ev.push_back({r.first, lanes});
                        ev.push_back({r.second, lanes});
                    }
                }
                std::sort(ev.begin(), ev.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
                Mask cur{};
                for (size_t k = 0; k < ev.size();) {
                    const int pos = ev[k].first;
                    while (k < ev.size() && ev[k].first == pos) {
                        for (int w = 0; w < kMaxTPL; w++) cur[w] ^= ev[k].second[w];
                        k++;
                    }
                    if (k < ev.size()) {
                        bool any = false;
                        for (int w = 0; w < kMaxTPL; w++) any = any || cur[w] != 0;
                        if (any) {
                            Mask m = cur;
                            bool all = true;
                            for (int w = 0; w < tpl; w++) all = all && (m[w] & act[w]) == act[w];
                            emit(soff[pos], soff[ev[k].first], all ? full : m);
                        }
                    }
                }
            }
            const bool single = Ta == Tb;
            last_T = single ? Ta : -1;
            last_work = gwork;
            last_slots.assign(split, {0, 0, 0});
            size_t iv_i = 0;
            int64_t iv_pos = giv.empty() ? 0 : giv[0].begin;
            for (int sl = 0; sl < split; sl++) {
                const int64_t quota = gwork * (sl + 1) / split - gwork * sl / split;
                FarGroup g{t0, (int64_t)P.ivs.size(), 0, cnt, b, sl, 0};
                int64_t left = quota;
                while (left > 0 && iv_i < giv.size()) {
                    const int64_t take = std::min(giv[iv_i].end - iv_pos, left);
                    FarInterval piece = giv[iv_i];
                    piece.begin = iv_pos;
                    piece.end = iv_pos + take;
                    P.ivs.push_back(piece);
                    left -= take;
                    iv_pos += take;
                    if (iv_pos == giv[iv_i].end) {
                        iv_i++;
                        if (iv_i < giv.size()) iv_pos = giv[iv_i].begin;
                    }
                }
                g.iv_end = (int64_t)P.ivs.size();
                P.groups.push_back(g);
                P.work.push_back(quota);
                last_slots[sl] = {g.iv_begin, g.iv_end, quota};
                if (aligned) {
                    for (int c = 0; c < 3; c++) P.cen.push_back(sph_pos[3 * Ta + c]);
                    P.cen.push_back(0.0);
                }
            }
        }
    };
    parallel_for(
        nparts,
        [&](int a, int e) {
            for (int pi = a; pi < e; pi++) build_part(pi);
        },
        1, 2);
    std::vector<FarGroup> groups;
    std::vector<FarInterval> ivs;
    std::vector<int64_t> work;
    std::vector<double> cen;
    double pairs = 0.0;
    {
        size_t ng = 0, ni = 0;
        for (const Part &P : parts) {
            ng += P.groups.size();
            ni += P.ivs.size();
        }
        groups.reserve(ng);
        ivs.reserve(ni);
        work.reserve(ng);
        for (Part &P : parts) {
            const int64_t base = (int64_t)ivs.size();
            for (FarGroup g : P.groups) {
                g.iv_begin += base;
                g.iv_end += base;
                groups.push_back(g);
            }
            ivs.insert(ivs.end(), P.ivs.begin(), P.ivs.end());
            work.insert(work.end(), P.work.begin(), P.work.end());
            cen.insert(cen.end(), P.cen.begin(), P.cen.end());
            pairs += P.pairs;
            Part().groups.swap(P.groups);
            std::vector<FarInterval>().swap(P.ivs);
        }
    }
    for (const FarGroup &g : groups) {
        const int b = g.susp;
        const int64_t tlo = toff[susp_sph_off[b]], thi = toff[susp_sph_off[b + 1]];
        const int64_t slo = soff[sso[b]], shi = soff[sso[b + 1]];
        SPHSTK_REQUIRE(g.tstart >= tlo && g.tstart + g.tcount <= thi && g.tcount > 0 && g.tcount <= G &&
                           g.slot >= 0 && g.slot < split,
                       "internal: far-field group out of range");
        for (int64_t k = g.iv_begin; k < g.iv_end; k++) {
            SPHSTK_REQUIRE(ivs[k].begin >= slo && ivs[k].end <= shi && ivs[k].begin < ivs[k].end,
                           "internal: far-field interval out of range");
            SPHSTK_REQUIRE(!aligned || ivs[k].mask[0] == ~0ull, "internal: masked interval in an aligned list");
        }
        SPHSTK_REQUIRE(!aligned || tsph[g.tstart] == tsph[g.tstart + g.tcount - 1],
                       "internal: aligned far-field group spans spheres");
    }
    std::vector<int> ord(groups.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return work[a] > work[b]; });
    L.ngroups = (int)groups.size();
    L.pairs = pairs;
    if (setup_verbose())
        std::fprintf(stderr, "[sphstk]     %d work items, %zu intervals, %.4g pairs\n", L.ngroups, ivs.size(), pairs);
    L.groups.upload(groups);
    L.iv.upload(ivs);
    L.order.upload(ord);
    L.aligned = aligned;
    if (aligned) L.cen.upload(cen);
    if (split > 1) L.part.alloc(size_t(split) * 3 * std::max<int64_t>(ntrg, 1));
}

static std::vector<Chunk> make_chunks(int nsusp, const std::vector<int64_t> &susp_sph_off,
                                      const std::vector<int64_t> &off, std::vector<int> *seg_off) {
    std::vector<Chunk> ch;
    if (seg_off) seg_off->assign(nsusp + 1, 0);
    for (int b = 0; b < nsusp; b++) {
        const int64_t d0 = 3 * off[susp_sph_off[b]], d1 = 3 * off[susp_sph_off[b + 1]];
        for (int64_t a = d0; a < d1; a += kChunk) ch.push_back(Chunk{a, std::min(a + kChunk, d1), b, 0});
        if (seg_off) (*seg_off)[b + 1] = (int)ch.size();
    }
    return ch;
}

// Multi-level far field of the FP32 inner operator.  With s = |c_S - c_T| /
// (R_S + R_T), a far pair (s >= 1.5) interacts
//   s < 3        point to point on the fine grids ("mid"),
//   3 <= s < 6   between grids of order p/2+1 (p < 10: p) (level 0),
//   6 <= s < 11  between grids of order max(5, p/3+1)    (level 1),
//   11 <= s < 18 between grids of order 4                (level 2),
//   s >= 18      between grids of order 3                (level 3),
// (orders capped by p) through the degree-truncated density of the source
// (Resample::down) with the target-grid field interpolated back to the fine grid
// (Resample::up).  The field of degree-n density decays like (1/(2s))^(n+2) and
// its angular content on the target like (1/(2s-1))^(n+1); these thresholds put
// the operator's error near 4e-6 relative (1024 spheres, p=12), well inside what
// two refinement steps against the FP64 residual need.  Level point sets are
// stored in k-d order of the sphere centres (per suspension) so that the
// spatially compact exclusion sets of the last level are few index ranges.
namespace {
int level_order(int lvl, int p) {
    int pc = p;
    if (lvl == 0) pc = p < 10 ? p : p / 2 + 1;
    if (lvl == 1) pc = std::max(5, p / 3 + 1);
    if (lvl == 2) pc = 4;
    if (lvl == 3) pc = 3;
    return std::min(pc, p);
}
// Balanced k-d order of the spheres idx[b, e): split at a multiple of `leaf`
// along the longest extent of the centres, the lower half first, down to
// ranges of at most `leaf` spheres.  Every leaf starts at b + a multiple of
// leaf, so the leaves are the consecutive runs of `leaf` ranks; they are
// compacter than runs of a Morton order (the 16384-sphere packing at vf 0.2:
// median cell radius 6.6 R against 8.2 R).
void kd_order(const double *pos, int *idx, int b, int e, int leaf) {
    const int n = e - b;
    if (n <= leaf) return;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = b; i < e; i++)
        for (int c = 0; c < 3; c++) {
            lo[c] = std::min(lo[c], pos[3 * idx[i] + c]);
            hi[c] = std::max(hi[c], pos[3 * idx[i] + c]);
        }
    int ax = 0;
    for (int c = 1; c < 3; c++)
        if (hi[c] - lo[c] > hi[ax] - lo[ax]) ax = c;
    const int h = n < 2 * leaf ? leaf : (n / leaf / 2) * leaf;
    std::nth_element(idx + b, idx + b + h, idx + e, [&](int u, int v) {
        const double pu = pos[3 * u + ax], pv = pos[3 * v + ax];
        return pu < pv || (pu == pv && u < v);
    });
    kd_order(pos, idx, b, b + h, leaf);
    kd_order(pos, idx, b + h, e, leaf);
}
} // namespace

void Operator::ensure_levels(LevelSet &S) {
    if (S.ready) return;
    PhaseTimer pt(S.fp64 ? "FP64 outer level set" : "FP32 inner level set");
    const double *a = S.alo;
    auto lorder = [&](int l, int p) { return S.order[l] < 0 ? level_order(l, p) : std::min(p, S.order[l]); };
    // k-d order of the spheres within each suspension
    std::vector<int> perm(nsph_), rank(nsph_);
    for (int i = 0; i < nsph_; i++) perm[i] = i;
    for (int b = 0; b < nsusp_; b++)
        kd_order(pos_.data(), perm.data(), (int)susp_sph_off_[b], (int)susp_sph_off_[b + 1], kCellSize);
    for (int k = 0; k < nsph_; k++) rank[perm[k]] = k;
    // Pair classes by one scan per target sphere over its suspension's spheres
    // in k-d order: near (< 1.5 (R_s + R_t), the operator's near-pair
    // predicate) and self are skipped, mid (< a0) is summed on the fine grid,
    // level l covers [a_l, a_l+1).  Level runs are in k-d rank space, the
    // mid runs in sphere index space.
    auto tn = std::make_unique<PhaseTimer>("  level pair classes");
    std::vector<Runs> mid(nsph_);
    const int nlev = S.nlev;
    std::vector<std::vector<Runs>> lvl(nlev, std::vector<Runs>(nsph_));
    std::vector<int> sph_s0(nsph_), sph_s1(nsph_);
    for (int b = 0; b < nsusp_; b++)
        for (int64_t t = susp_sph_off_[b]; t < susp_sph_off_[b + 1]; t++) {
            sph_s0[t] = (int)susp_sph_off_[b];
            sph_s1[t] = (int)susp_sph_off_[b + 1];
        }
    // cells of the FP32 set: the k-d leaves (kCellSize consecutive ranks) of one
    // suspension, with centre (mean of the member centres) and radius (largest
    // member distance + member radius)
    std::vector<int> cell_r0, cell_r1;
    std::vector<double> ccen, crho;
    std::vector<int64_t> susp_cell_off(nsusp_ + 1, 0);
    if (S.cells) {
        for (int b = 0; b < nsusp_; b++) {
            const int s0 = (int)susp_sph_off_[b], s1 = (int)susp_sph_off_[b + 1];
            for (int r = s0; r < s1; r += kCellSize) {
                const int r1 = std::min(r + kCellSize, s1);
                double c[3] = {0.0, 0.0, 0.0};
                for (int q = r; q < r1; q++)
                    for (int k = 0; k < 3; k++) c[k] += pos_[3 * perm[q] + k] / (r1 - r);
                double rho = 0.0;
                for (int q = r; q < r1; q++) {
                    const int sq = perm[q];
                    const double e = std::sqrt((pos_[3 * sq] - c[0]) * (pos_[3 * sq] - c[0]) +
                                               (pos_[3 * sq + 1] - c[1]) * (pos_[3 * sq + 1] - c[1]) +
                                               (pos_[3 * sq + 2] - c[2]) * (pos_[3 * sq + 2] - c[2]));
                    rho = std::max(rho, e + rad_[sq]);
                }
                cell_r0.push_back(r);
                cell_r1.push_back(r1);
                ccen.insert(ccen.end(), c, c + 3);
                crho.push_back(rho);
            }
            susp_cell_off[b + 1] = (int64_t)cell_r0.size();
        }
    }
    const int ncell = (int)cell_r0.size();
    // cell runs of each target (by k-d rank) on the level-2 and level-3 grids,
    // and per target the pair evaluations its cell interactions save
    std::vector<std::array<Runs, 2>> crun(S.cells ? nsph_ : 0);
    std::vector<double> saved(S.cells ? nsph_ : 0, 0.0);
    const int Keq = equiv_tables(S.equiv_k).K;
    auto scan = [&](int ta, int tb) {
        std::vector<int> ml;
        // pair class of (t, s): -3 self, -2 near, -1 mid, else the level.  The
        // FP32 set uses the separation s = d / (R_s + R_t).  The FP64 set, whose
        // accuracy the refinement's residual rests on, uses the equal-radii
        // separation with the same worst truncation ratio, s' = (1/q + 1)/2 with
        // q = max(R_s / (d - R_t), R_t / (d - R_s)) (the source density's
        // truncated degrees decay like R_s / (d - R_t) at the target, the
        // target grid's interpolation error like R_t / (d - R_s)): s' = s for
        // equal radii, smaller (a finer level) for unequal ones.
        auto klass = [&](int t, int s) {
            if (s == t) return -3;
            const double ddx = pos_[3 * s] - pos_[3 * t], ddy = pos_[3 * s + 1] - pos_[3 * t + 1],
                         ddz = pos_[3 * s + 2] - pos_[3 * t + 2];
            const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
            const double rr = 1.5 * rad_[s] + 1.5 * rad_[t], rs = rad_[s] + rad_[t];
            if (d2 < rr * rr) return -2;
            double sep;
            if (S.fp64) {
                const double d = std::sqrt(d2);
                const double q = std::max(rad_[s] / (d - rad_[t]), rad_[t] / (d - rad_[s]));
                sep = 0.5 * (1.0 / q + 1.0);
            } else {
                sep = std::sqrt(d2) / rs;
            }
            if (sep < a[0]) return -1;
            int l = 0;
            while (l + 1 < nlev && sep >= a[l + 1]) l++;
            return l;
        };
        auto member = [&](int t, int r, int rt) {
            const int s = perm[r];
            const int l = klass(t, s);
            if (l == -1) ml.push_back(s);
            if (l >= 0) run_append(lvl[l][rt], r);
        };
        for (int t = ta; t < tb; t++) {
            const int s0 = sph_s0[t], s1 = sph_s1[t], rt = rank[t];
            ml.clear();
            if (!S.cells) {
                for (int r = s0; r < s1; r++) member(t, r, rt);
            } else {
                const int b = sph_susp_[t];
                for (int64_t c = susp_cell_off[b]; c < susp_cell_off[b + 1]; c++) {
                    const double ex = ccen[3 * c] - pos_[3 * t], ey = ccen[3 * c + 1] - pos_[3 * t + 1],
                                 ez = ccen[3 * c + 2] - pos_[3 * t + 2];
                    if (std::sqrt(ex * ex + ey * ey + ez * ez) >= S.cell_theta * crho[c] + rad_[t]) {
                        int lmin = nlev;
                        double member_pairs = 0.0;
                        for (int r = cell_r0[c]; r < cell_r1[c] && lmin >= 2; r++) {
                            const int sm = perm[r], lm = klass(t, sm);
                            lmin = std::min(lmin, lm);
                            if (lm >= 0)
                                member_pairs += double(grid_points(lorder(lm, order_[sm]))) *
                                                grid_points(lorder(lm, order_[t]));
                        }
                        if (lmin >= 2) {
                            run_append(crun[rt][lmin - 2], (int)c);
                            saved[t] += member_pairs - double(Keq) * grid_points(lorder(lmin, order_[t]));
                            continue;
                        }
                    }
                    for (int r = cell_r0[c]; r < cell_r1[c]; r++) member(t, r, rt);
                }
            }
            std::sort(ml.begin(), ml.end());
            mid[t] = runs_of(ml, s0, s1, true);
        }
    };
    parallel_for(nsph_, scan);
    if (S.cells) {
        // keep the cells only where they pay: the pair evaluations they save
        // against the check velocities of the cells in use (Kc check points x
        // the members' fine grids, plus the fit, ~K Kc / 5 pair evaluations)
        double gain = 0.0, cost = 0.0;
        for (double v : saved) gain += v;
        std::vector<char> used(ncell, 0);
        for (const auto &cr : crun)
            for (const Runs &rr : cr)
                for (const auto &r : rr)
                    for (int c = r.first; c < r.second; c++) used[c] = 1;
        const EquivTables &E = equiv_tables(S.equiv_k);
        for (int c = 0; c < ncell; c++)
            if (used[c]) {
                double pts = 0.0;
                for (int r = cell_r0[c]; r < cell_r1[c]; r++) pts += grid_points(order_[perm[r]]);
                cost += E.Kc * pts + E.K * E.Kc / 5.0;
            }
        if (gain < 3.0 * cost) {
            S.cells = false;
            crun.clear();
            for (auto &lv : lvl)
                for (Runs &r : lv) r.clear();
            parallel_for(nsph_, scan);
        }
    }
    tn.reset();
    {
        PhaseTimer pm("  mid-range lists");
        build_far_lists(S.mid, nsusp_, susp_sph_off_, pt_off_, pt_off_, mid, S.tpl, 0, kWantItems);
    }
    S.levels.clear();
    std::vector<std::vector<int64_t>> level_off(nlev);
    for (int l = 0; l < nlev; l++) {
        auto L = std::make_unique<FarLevel>();
        std::vector<int64_t> &off = level_off[l];
        off.assign(nsph_ + 1, 0);
        std::vector<int64_t> off_by_sph(nsph_);
        int maxpts = 0;
        for (int r = 0; r < nsph_; r++) {
            const int s = perm[r];
            const int p = order_[s], pc = lorder(l, p);
            off[r + 1] = off[r] + grid_points(pc);
            off_by_sph[s] = off[r];
            maxpts = std::max(maxpts, std::max(grid_points(p), grid_points(pc)));
        }
        const int64_t nc = off[nsph_];
        std::vector<double> cc(3 * nc), cn(3 * nc), cw(nc);
        for (int s = 0; s < nsph_; s++) {
            const ShTables &t = sh_tables(lorder(l, order_[s]));
            for (int q = 0; q < t.npts; q++) {
                const int64_t g = off_by_sph[s] + q;
                double nv[3];
                qrot_host(&quat_[4 * s], &t.eta[3 * q], nv);
                for (int c = 0; c < 3; c++) {
                    cn[3 * g + c] = nv[c];
                    cc[3 * g + c] = pos_[3 * s + c] + rad_[s] * nv[c];
                }
                cw[g] = rad_[s] * rad_[s] * t.w[q];
            }
        }
        L->npts = nc;
        L->max_pts = maxpts;
        L->d_off_by_sph.upload(off_by_sph);
        L->coords.upload(cc);
        L->norms.upload(cn);
        L->w.upload(cw);
        std::vector<bool> used(kMaxOrder + 1, false);
        for (int p : order_) used[p] = true;
        for (int p = 0; p <= kMaxOrder; p++)
            if (used[p]) {
                const Resample &rs = resample_tables(p, lorder(l, p));
                L->d_down[p].upload(rs.down);
                L->d_up[p].upload(rs.up);
                L->down.M[p] = L->d_down[p].p;
                L->down.nin[p] = rs.n;
                L->down.nout[p] = rs.nc;
                L->up.M[p] = L->d_up[p].p;
                L->up.nin[p] = rs.nc;
                L->up.nout[p] = rs.n;
                if (!S.fp64) {
                    L->d_down32[p].upload(std::vector<float>(rs.down.begin(), rs.down.end()));
                    L->d_up32[p].upload(std::vector<float>(rs.up.begin(), rs.up.end()));
                    L->down.M32[p] = L->d_down32[p].p;
                    L->up.M32[p] = L->d_up32[p].p;
                }
            }
        L->down.fp32 = L->up.fp32 = !S.fp64;
        PhaseTimer pl("  level lists");
        build_far_lists(L->lists, nsusp_, susp_sph_off_, off, off, lvl[l], S.tpl, 0, kWantItems);
        std::vector<int> dummy;
        L->chunks.upload(make_chunks(nsusp_, susp_sph_off_, off, &dummy));
        L->sv.alloc(std::max<int64_t>(3 * nc, 1));
        L->tv.alloc(std::max<int64_t>(3 * nc, 1));
        if (S.fp64) {
            L->src64.alloc(std::max<int64_t>(8 * nc, 8));
        } else {
            L->src32.alloc(std::max<int64_t>(8 * nc, 8));
            L->trg32.alloc(std::max<int64_t>(8 * nc, 8));
            launch_pack_trg_f32(nc, L->coords.p, L->norms.p, L->trg32.p, stream_);
        }
        S.levels.push_back(std::move(L));
    }
    S.agg.reset();
    // keep the cells some target uses, renumbered in order (runs stay sorted)
    std::vector<int> cell_new(ncell, -1);
    int nused = 0;
    if (S.cells) {
        std::vector<char> used(ncell, 0);
        for (const auto &cr : crun)
            for (const Runs &rr : cr)
                for (const auto &r : rr)
                    for (int c = r.first; c < r.second; c++) used[c] = 1;
        std::vector<int64_t> sco(nsusp_ + 1, 0);
        for (int b = 0; b < nsusp_; b++) {
            for (int64_t c = susp_cell_off[b]; c < susp_cell_off[b + 1]; c++)
                if (used[c]) cell_new[c] = nused++;
            sco[b + 1] = nused;
        }
        std::vector<int> r0, r1;
        std::vector<double> cc, cr;
        for (int c = 0; c < ncell; c++)
            if (cell_new[c] >= 0) {
                r0.push_back(cell_r0[c]);
                r1.push_back(cell_r1[c]);
                cc.insert(cc.end(), &ccen[3 * c], &ccen[3 * c] + 3);
                cr.push_back(crho[c]);
            }
        for (auto &cv : crun)
            for (Runs &rr : cv) {
                Runs nr;
                for (const auto &r : rr)
                    for (int c = r.first; c < r.second; c++) run_append(nr, cell_new[c]);
                rr.swap(nr);
            }
        cell_r0.swap(r0);
        cell_r1.swap(r1);
        ccen.swap(cc);
        crho.swap(cr);
        susp_cell_off.swap(sco);
    }
    if (S.cells && nused > 0) {
        const int ncell = nused;
        PhaseTimer pc("  cells");
        auto A = std::make_unique<CellAgg>();
        const EquivTables &E = equiv_tables(S.equiv_k);
        const int K = E.K, Kc = E.Kc, m = 3 * Kc, n = 3 * K;
        A->ncell = ncell;
        A->E = &E;
        A->susp_cell_off = susp_cell_off;
        std::vector<float> chk(8 * (size_t)ncell * Kc, 0.f), eq(8 * (size_t)ncell * K, 0.f);
        std::vector<int> csusp(ncell);
        for (int c = 0; c < ncell; c++) {
            for (int i = 0; i < Kc; i++)
                for (int k = 0; k < 3; k++)
                    chk[8 * ((size_t)c * Kc + i) + k] = (float)(ccen[3 * c + k] + E.rc * crho[c] * E.chk[3 * i + k]);
            for (int i = 0; i < K; i++)
                for (int k = 0; k < 3; k++)
                    eq[8 * ((size_t)c * K + i) + k] = (float)(ccen[3 * c + k] + E.re * crho[c] * E.eq[3 * i + k]);
            csusp[c] = sph_susp_[perm[cell_r0[c]]];
        }
        std::vector<double> Pcm((size_t)n * m);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < m; j++) Pcm[(size_t)j * n + i] = E.P[(size_t)i * m + j];
        A->chk32.upload(chk);
        A->eq32.upload(eq);
        A->P.upload(Pcm);
        A->rho.upload(crho);
        A->susp.upload(csusp);
        A->chk_u.alloc((size_t)ncell * m);
        // check points of cell c x the fine grids of its members (sphere index space)
        std::vector<int64_t> chk_off(ncell + 1), eq_off(ncell + 1);
        std::vector<Runs> chk_runs(ncell);
        for (int c = 0; c <= ncell; c++) {
            chk_off[c] = (int64_t)c * Kc;
            eq_off[c] = (int64_t)c * K;
        }
        for (int c = 0; c < ncell; c++) {
            std::vector<int> mem(perm.begin() + cell_r0[c], perm.begin() + cell_r1[c]);
            std::sort(mem.begin(), mem.end());
            for (int sm : mem) run_append(chk_runs[c], sm);
        }
        build_far_lists(A->chk_lists, nsusp_, susp_cell_off, chk_off, pt_off_, chk_runs, 1, 0, kWantItems, nullptr,
                        &susp_sph_off_);
        A->chk_chunks.upload(make_chunks(nsusp_, susp_cell_off, chk_off, nullptr));
        A->check_pairs = A->chk_lists.pairs;
        for (int i = 0; i < 2; i++) {
            std::vector<Runs> runs(nsph_);
            for (int r = 0; r < nsph_; r++) runs[r] = crun[r][i];
            build_far_lists(A->lists[i], nsusp_, susp_sph_off_, level_off[2 + i], eq_off, runs, S.tpl, 0, kWantItems,
                            nullptr, &susp_cell_off);
            A->pairs += A->lists[i].pairs;
        }
        S.agg = std::move(A);
    }
    S.ready = true;
}

void Operator::harvest_events() {
    for (auto &e : pending_) {
        SPHSTK_HIP(hipEventSynchronize(e.b));
        float ms = 0.f;
        SPHSTK_HIP(hipEventElapsedTime(&ms, e.a, e.b));
        stats_[e.kind][e.prec].ms += ms;
        stats_[e.kind][e.prec].launches++;
        free_events_.push_back(e);
    }
    pending_.clear();
}

void Operator::far_run(FarLists &L, int kind, int prec, const void *src, const double *tcoords,
                       const double *tnorms, const float *trg32, const Chunk *chunks, int nchunks, double coef,
                       double *out, const int *active) {
    if (L.pairs == 0.0) return; // e.g. a level that no sphere pair reaches
    double *dst = L.split > 1 ? L.part.p : out;
    const int acc = L.split > 1 ? 0 : 1;
    launch_far(kind, prec, L.tpl, src, tcoords, tnorms, trg32, L.ntrg, L.ngroups, L.order.p, L.groups.p, L.iv.p,
               active, dst, coef, acc, stream_, &L != &lists_fine_);
    if (L.split > 1) launch_slot_reduce(out, L.part.p, L.split, 3 * L.ntrg, chunks, nchunks, active, stream_);
}

// Multi-level far field (the packed fine-grid sources of `in` are already in
// src64_ or src32_): each level resamples `in` to its grid, sums its shell of
// sphere pairs there and interpolates the result back onto the fine grid; the
// pairs closer than the first level are summed on the fine grid.  ev_a and
// ev_b bracket the pair sums only.
void Operator::far_levels(LevelSet &S, int kind, double coef, const double *in, double *out, const int *active,
                          hipEvent_t ev_a, hipEvent_t ev_b) {
    ensure_levels(S);
    const int prec = S.fp64 ? 0 : 1;
    for (auto &L : S.levels) {
        launch_resample(geom_, L->down, d_pt_off_.p, L->d_off_by_sph.p, in, L->sv.p, 0, active, stream_, L->max_pts);
        if (S.fp64)
            launch_pack_src_f64(L->npts, L->coords.p, L->w.p, L->sv.p, L->src64.p, stream_);
        else
            launch_pack_src_f32(L->npts, L->coords.p, L->w.p, L->sv.p, L->src32.p, stream_);
        launch_fill(3 * L->npts, 0.0, L->tv.p, stream_);
    }
    CellAgg *A = S.agg.get();
    if (A) {
        // check velocities of each cell's raw fine-grid sources (single layer,
        // FP32 pairs, FP64 sums), then the equivalent point forces
        launch_fill((int64_t)A->ncell * 3 * A->E->Kc, 0.0, A->chk_u.p, stream_);
        far_run(A->chk_lists, 1, 1, src32_.p, nullptr, nullptr, A->chk32.p, A->chk_chunks.p, (int)A->chk_chunks.n,
                1.0 / (8.0 * kPi), A->chk_u.p, active);
        launch_equiv_forces(A->ncell, A->E->K, A->E->Kc, A->P.p, A->chk_u.p, A->rho.p, A->susp.p,
                            active, A->eq32.p, stream_);
    }
    if (ev_a) SPHSTK_HIP(hipEventRecord(ev_a, stream_));
    if (A)
        for (int i = 0; i < 2; i++) {
            FarLevel &L = *S.levels[2 + i];
            far_run(A->lists[i], kind, 1, A->eq32.p, nullptr, nullptr, L.trg32.p, L.chunks.p, (int)L.chunks.n, coef,
                    L.tv.p, active);
        }
    for (auto &L : S.levels) {
        if (S.fp64)
            far_run(L->lists, kind, 0, L->src64.p, L->coords.p, L->norms.p, nullptr, L->chunks.p, (int)L->chunks.n,
                    coef, L->tv.p, active);
        else
            far_run(L->lists, kind, 1, L->src32.p, nullptr, nullptr, L->trg32.p, L->chunks.p, (int)L->chunks.n, coef,
                    L->tv.p, active);
    }
    if (S.fp64)
        far_run(S.mid, kind, prec, src64_.p, d_coords_.p, d_norms_.p, nullptr, chunks_.p, (int)chunks_.n, coef, out,
                active);
    else
        far_run(S.mid, kind, prec, src32_.p, nullptr, nullptr, trg32_.p, chunks_.p, (int)chunks_.n, coef, out,
                active);
    if (ev_b) SPHSTK_HIP(hipEventRecord(ev_b, stream_));
    for (auto &L : S.levels)
        launch_resample(geom_, L->up, L->d_off_by_sph.p, d_pt_off_.p, L->tv.p, out, 1, active, stream_, L->max_pts);
}

// FP64 far field of ncopy identical copies of a set of suspensions (several
// right-hand sides on one geometry), with the first copy's work lists.  The
// traction sum is computed once per pair for blocks of up to 8 copies
// (k_far_trac_multi_f64); the single layer one copy at a time on the first
// copy's geometry.
void Operator::far_copies(int kind, double coef, const double *in, double *out, const int *active, hipEvent_t ev_a) {
    FarLists &L = lists_fine_;
    const int b0 = nsusp_ / ncopy_;
    const int64_t n0 = npts_ / ncopy_, col = 3 * n0;
    const bool split = L.split > 1;
    if (ev_a) SPHSTK_HIP(hipEventRecord(ev_a, stream_));
    if (L.pairs == 0.0) return;
    if (kind == 0) {
        for (int c0 = 0; c0 < ncopy_; c0 += kMultiCols) {
            const int nc = std::min(kMultiCols, ncopy_ - c0);
            launch_far_trac_multi(nc, n0, d_coords_.p, d_norms_.p, d_w_.p, in + c0 * col, col, srcm_.p, L.ngroups,
                                  L.order.p, L.groups.p, L.iv.p, active ? active + c0 * b0 : nullptr, b0,
                                  (split ? L.part.p : out) + c0 * col, coef, split ? 0 : 1, 3 * npts_, stream_);
        }
        if (split) launch_slot_reduce(out, L.part.p, L.split, 3 * npts_, chunks_.p, (int)chunks_.n, active, stream_);
        return;
    }
    const int nch0 = seg_chunk_off_host_[b0];
    for (int c = 0; c < ncopy_; c++) {
        const int *act = active ? active + c * b0 : nullptr;
        launch_pack_src_f64(n0, d_coords_.p, d_w_.p, in + c * col, src64_.p, stream_);
        launch_far(kind, 0, L.tpl, src64_.p, d_coords_.p, d_norms_.p, nullptr, L.ntrg, L.ngroups, L.order.p,
                   L.groups.p, L.iv.p, act, split ? L.part.p : out + c * col, coef, split ? 0 : 1, stream_);
        if (split) launch_slot_reduce(out + c * col, L.part.p, L.split, 3 * L.ntrg, chunks_.p, nch0, act, stream_);
    }
}

// The FP64 multi-level far field is used when its pair evaluations (fine mid
// range plus all levels) are at most 3/4 of the direct sum's; below that the
// resampling and the extra launches eat the saving (many small suspensions).
void Operator::wait_outer_levels() {
    if (ls64_build_.valid()) ls64_build_.get();
}
void Operator::wait_inner_levels() {
    if (ls32_build_.valid()) ls32_build_.get();
}

bool Operator::outer_levels_pay() {
    wait_outer_levels();
    if (outer_pay_ < 0) {
        ensure_levels(ls64_);
        double ml = ls64_.mid.pairs;
        for (auto &L : ls64_.levels) ml += L->lists.pairs;
        outer_pay_ = ml <= 0.75 * lists_fine_.pairs ? 1 : 0;
    }
    return outer_pay_ == 1;
}

// Far field of `in` into out (accumulating).  FP64: fine grid, direct sum over
// all far pairs.  FP32: multi-level (fine for mid-range pairs, coarser grids
// with distance) unless disabled, in which case the FP32 direct sum.
void Operator::far(int kind, int prec, double coef, const double *in, double *out, const int *active) {
    const bool timing = (flags_ & 1) != 0 && !capturing_;
    EvPair ev{};
    if (timing) {
        if (free_events_.empty()) {
            SPHSTK_HIP(hipEventCreate(&ev.a));
            SPHSTK_HIP(hipEventCreate(&ev.b));
        } else {
            ev = free_events_.back();
            free_events_.pop_back();
        }
        ev.kind = kind;
        ev.prec = prec;
    }
    if (ncopy_ > 1) {
        SPHSTK_REQUIRE(prec == 0, "several right-hand sides on one geometry are solved in FP64 only");
        far_copies(kind, coef, in, out, active, timing ? ev.a : nullptr);
    } else if (prec == 0) {
        launch_pack_src_f64(npts_, d_coords_.p, d_w_.p, in, src64_.p, stream_);
        if (timing) SPHSTK_HIP(hipEventRecord(ev.a, stream_));
        far_run(lists_fine_, kind, 0, src64_.p, d_coords_.p, d_norms_.p, nullptr, chunks_.p, (int)chunks_.n, coef,
                out, active);
    } else if (prec == 2 && !outer_levels_pay()) {
        // the coarse levels would save too few pair evaluations: direct FP64 sum
        launch_pack_src_f64(npts_, d_coords_.p, d_w_.p, in, src64_.p, stream_);
        if (timing) SPHSTK_HIP(hipEventRecord(ev.a, stream_));
        far_run(lists_fine_, kind, 0, src64_.p, d_coords_.p, d_norms_.p, nullptr, chunks_.p, (int)chunks_.n, coef,
                out, active);
    } else if (prec == 2) {
        launch_pack_src_f64(npts_, d_coords_.p, d_w_.p, in, src64_.p, stream_);
        far_levels(ls64_, kind, coef, in, out, active, timing ? ev.a : nullptr, timing ? ev.b : nullptr);
        if (timing) {
            pending_.push_back(ev);
            timing_done_ = true;
        }
    } else {
        if (prec == 1) wait_inner_levels();
        if (!trg32_.p) {
            trg32_.alloc(8 * std::max<int64_t>(npts_, 1));
            src32_.alloc(8 * std::max<int64_t>(npts_, 1));
            launch_pack_trg_f32(npts_, d_coords_.p, d_norms_.p, trg32_.p, stream_);
        }
        launch_pack_src_f32(npts_, d_coords_.p, d_w_.p, in, src32_.p, stream_);
        if (multi_level_) {
            far_levels(prec == 3 ? lsc_ : ls32_, kind, coef, in, out, active, timing ? ev.a : nullptr,
                       timing ? ev.b : nullptr);
            if (timing) {
                pending_.push_back(ev);
                timing_done_ = true;
            }
        } else {
            if (timing) SPHSTK_HIP(hipEventRecord(ev.a, stream_));
            far_run(lists_fine_, kind, 1, src32_.p, nullptr, nullptr, trg32_.p, chunks_.p, (int)chunks_.n, coef, out,
                    active);
        }
    }
    SPHSTK_HIP(hipGetLastError());
    if (timing && !timing_done_) {
        SPHSTK_HIP(hipEventRecord(ev.b, stream_));
        pending_.push_back(ev);
    }
    timing_done_ = false;
    if (timing && pending_.size() > 64) harvest_events();
}

// applyP2POP (cpp:262-485), accumulating into out.
void Operator::p2p(const double *in, double *out, double cId, double cSL, double cTrac, int prec, const int *active) {
    const bool doSL = std::fabs(cSL) > 1e-9, doT = std::fabs(cTrac) > 1e-9;
    double cid = std::fabs(cId) > 1e-9 ? cId : 0.0;
    if (!doSL && !doT) {
        if (cid != 0.0) launch_identity(geom_, in, out, cid, active, stream_);
        return;
    }
    if (nsph_ == 0) return;
    if (doT && !capturing_) {
        // the self and near terms (SH analysis, expansion, evaluation into
        // scratch) on the second stream while the far field runs on the first;
        // the gather into out waits for both
        SPHSTK_HIP(hipEventRecord(ev_fork_, stream_));
        SPHSTK_HIP(hipStreamWaitEvent(aux_, ev_fork_, 0));
        launch_analysis(geom_, in, coef_.p, active, aux_, max_order_);
        launch_expand(geom_, 1, coef_.p, rec_.p, active, aux_, max_order_);
        launch_sheval_items(geom_, 1, rec_.p, items_.p, (int)items_.n, evscratch_.p, 0.0, cTrac, active, aux_,
                            &src_items_);
        SPHSTK_HIP(hipEventRecord(ev_join_, aux_));
        far(0, prec, -3.0 / (4.0 * kPi) * cTrac, in, out, active);
        SPHSTK_HIP(hipStreamWaitEvent(stream_, ev_join_, 0));
        launch_sheval_gather(geom_, (int)items_.n, d_trg_slot_off_.p, d_trg_nsrc_.p, evscratch_.p, in, out, cid,
                             active, stream_);
        cid = 0.0;
    } else {
        launch_analysis(geom_, in, coef_.p, active, stream_, max_order_);
    }
    if (doT && capturing_) {
        far(0, prec, -3.0 / (4.0 * kPi) * cTrac, in, out, active);
        launch_expand(geom_, 1, coef_.p, rec_.p, active, stream_, max_order_);
        launch_sheval(geom_, 1, rec_.p, items_.p, (int)items_.n, d_trg_slot_off_.p, d_trg_nsrc_.p, evscratch_.p, in,
                      out, 0.0, cTrac, cid, active, stream_, &src_items_);
        cid = 0.0;
    }
    if (doSL) {
        far(1, prec, cSL / (8.0 * kPi), in, out, active);
        launch_expand(geom_, 0, coef_.p, rec_.p, active, stream_, max_order_);
        launch_sheval(geom_, 0, rec_.p, items_.p, (int)items_.n, d_trg_slot_off_.p, d_trg_nsrc_.p, evscratch_.p, in,
                      out, cSL, 0.0, cid, active, stream_);
    }
    SPHSTK_HIP(hipGetLastError());
}

// Other spheres of the same suspension with centre distance below
// amax (R_s + R_t), per sphere, sorted.
std::vector<std::vector<int>> Operator::neighbours_within(double amax) const {
    std::vector<std::vector<int>> nb(nsph_);
    for (int b = 0; b < nsusp_; b++) {
        const int s0 = (int)susp_sph_off_[b], s1 = (int)susp_sph_off_[b + 1];
        if (s1 <= s0) continue;
        double rmax = 0, lo[3] = {1e300, 1e300, 1e300};
        for (int i = s0; i < s1; i++) {
            rmax = std::max(rmax, rad_[i]);
            for (int c = 0; c < 3; c++) lo[c] = std::min(lo[c], pos_[3 * i + c]);
        }
        const double h = 2.0 * amax * rmax;
        std::unordered_map<int64_t, std::vector<int>> cells;
        auto key = [](int64_t a, int64_t bb, int64_t c) { return (a * 1000003LL + bb) * 1000033LL + c; };
        std::vector<std::array<int64_t, 3>> cid(s1 - s0);
        for (int i = s0; i < s1; i++) {
            for (int c = 0; c < 3; c++) cid[i - s0][c] = (int64_t)std::floor((pos_[3 * i + c] - lo[c]) / h);
            cells[key(cid[i - s0][0], cid[i - s0][1], cid[i - s0][2])].push_back(i);
        }
        for (int t = s0; t < s1; t++) {
            const auto &ct = cid[t - s0];
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dz = -1; dz <= 1; dz++) {
                        auto it = cells.find(key(ct[0] + dx, ct[1] + dy, ct[2] + dz));
                        if (it == cells.end()) continue;
                        for (int s : it->second) {
                            const double ex = pos_[3 * s] - pos_[3 * t], ey = pos_[3 * s + 1] - pos_[3 * t + 1],
                                         ez = pos_[3 * s + 2] - pos_[3 * t + 2];
                            if (s != t && std::sqrt(ex * ex + ey * ey + ez * ez) < amax * (rad_[s] + rad_[t]))
                                nb[t].push_back(s);
                        }
                    }
            std::sort(nb[t].begin(), nb[t].end());
        }
    }
    return nb;
}

// Rigid-body averages of u = S[rho] (V = mean u / mu, Omega = 3/(8 pi R^4 mu) sum w R n x u):
// self and near terms by quadrature of the SH evaluations on the grid (as the
// reference), far pairs closer than kFaxenSep (R_s + R_t) by the direct FP64
// sum on the grid and the same quadrature, farther pairs through the
// sphere-centre sums (Faxen's laws: the exact surface averages, which the grid
// quadrature of order p reproduces to ~(R_t / d)^(2p+2), below 1e-10 for p >= 6
// beyond 4 (R_s + R_t) when R_t / (R_s + R_t) <= 0.75).
constexpr double kFaxenSep = 4.0;
void Operator::ensure_centre() {
    if (!centre_ready_) {
        PhaseTimer pt("final-velocity lists");
        std::vector<int64_t> toff(nsph_ + 1);
        for (int i = 0; i <= nsph_; i++) toff[i] = i;
        // excluded from the centre sums: the sphere itself and every sphere
        // closer than kFaxenSep; of those, the non-near ones form the grid list
        std::vector<std::vector<int>> excl = neighbours_within(kFaxenSep), grid_mid(nsph_);
        for (int t = 0; t < nsph_; t++) {
            for (int s : excl[t]) {
                const double ex = pos_[3 * s] - pos_[3 * t], ey = pos_[3 * s + 1] - pos_[3 * t + 1],
                             ez = pos_[3 * s + 2] - pos_[3 * t + 2];
                if (std::sqrt(ex * ex + ey * ey + ez * ez) >= 1.5 * (rad_[s] + rad_[t])) grid_mid[t].push_back(s);
            }
            excl[t].push_back(t);
            std::sort(excl[t].begin(), excl[t].end());
        }
        std::vector<Runs> rmid(nsph_), rcen(nsph_);
        for (int t = 0; t < nsph_; t++) {
            const int b = sph_susp_[t], s0 = (int)susp_sph_off_[b], s1 = (int)susp_sph_off_[b + 1];
            rmid[t] = runs_of(grid_mid[t], s0, s1, true);
            rcen[t] = runs_of(excl[t], s0, s1, false);
        }
        build_far_lists(lists_rv_mid_, nsusp_, susp_sph_off_, pt_off_, pt_off_, rmid, lists_fine_.tpl, 0, kWantItems);
        int64_t nb = 0;
        for (int b = 0; b < nsusp_; b++) nb += (susp_sph_off_[b + 1] - susp_sph_off_[b] + 63) / 64;
        const int split = (int)std::max<int64_t>(1, std::min<int64_t>(64, (kWantItems + nb - 1) / std::max<int64_t>(nb, 1)));
        build_far_lists(lists_centre_, nsusp_, susp_sph_off_, toff, pt_off_, rcen, 1, split, kWantItems);
        build_far_lists(lists_rigid_, nsusp_, susp_sph_off_, pt_off_, toff, rcen, 1, 1, kWantItems);
        centre_part_.alloc(size_t(lists_centre_.split) * 9 * nsph_);
        vel_ns_.alloc(6 * (size_t)nsph_);
        u_ns_.alloc(std::max<int64_t>(3 * npts_, 1));
        centre_ready_ = true;
    }
}

// Builds everything the mixed-precision solve uses (otherwise built on first
// use): the FP32 and FP64 level sets and the final-velocity lists.  The FP64
// set is first needed by the first refinement residual, after a whole inner
// solve, so it is built on a host thread meanwhile: its construction is host
// work, allocations and synchronous copies (no launches on the operator's
// non-blocking stream), and nothing else reads ls64_ until
// outer_levels_pay() has joined the thread (16384 spheres p=12: 3.1 s of the
// 4.5 s set-up).
void Operator::prepare_mixed(bool outer_levels, bool coarse) {
    if (nsph_ == 0) return;
    if (outer_levels && !ls64_.ready && !ls64_build_.valid()) {
        const int dev = device_;
        ls64_build_ = std::async(std::launch::async, [this, dev] {
            SPHSTK_HIP(hipSetDevice(dev));
            ensure_levels(ls64_);
        });
    }
    if (multi_level_) {
        if (flags_ & 4) {
            double budget = kTuneBudget;
            autotune_inner(budget);
        }
        if (coarse && !ls32_.ready && !ls32_build_.valid()) {
            // with the coarse operator first, A32 is first needed after the
            // first coarse GMRES: its set is built beside it the same way (its
            // one launch, the target packing, writes only ls32_'s buffers)
            const int dev = device_;
            ls32_build_ = std::async(std::launch::async, [this, dev] {
                SPHSTK_HIP(hipSetDevice(dev));
                ensure_levels(ls32_);
            });
        } else {
            ensure_levels(ls32_);
        }
        if (coarse) ensure_levels(lsc_);
    }
    ensure_centre();
    SPHSTK_HIP(hipStreamSynchronize(stream_));
}

void Operator::rigid_p2p(const double *force, const double *rho0, double *out, double cId, double cTrac) {
    if (nsph_ == 0) return;
    ensure_centre();
    launch_analysis(geom_, rho0, coef_.p, nullptr, stream_, max_order_);
    launch_pack_src_f64(npts_, d_coords_.p, d_w_.p, rho0, src64_.p, stream_);
    far_run(lists_rv_mid_, 0, 0, src64_.p, d_coords_.p, d_norms_.p, nullptr, chunks_.p, (int)chunks_.n,
            -3.0 / (4.0 * kPi) * cTrac, out, nullptr);
    launch_far_rigid(d_coords_.p, d_norms_.p, d_pos_.p, force, d_rad_.p, npts_, lists_rigid_.ngroups,
                     lists_rigid_.order.p, lists_rigid_.groups.p, lists_rigid_.iv.p, out, cTrac, 1, stream_);
    launch_expand(geom_, 1, coef_.p, rec_.p, nullptr, stream_, max_order_);
    launch_sheval(geom_, 1, rec_.p, items_.p, (int)items_.n, d_trg_slot_off_.p, d_trg_nsrc_.p, evscratch_.p, rho0, out,
                  0.0, cTrac, cId, nullptr, stream_, &src_items_);
    SPHSTK_HIP(hipGetLastError());
}

void Operator::rigid_velocity(const double *rho, double inv_mu, double *vel) {
    if (nsph_ == 0) return;
    ensure_centre();
    launch_analysis(geom_, rho, coef_.p, nullptr, stream_, max_order_);
    launch_expand(geom_, 0, coef_.p, rec_.p, nullptr, stream_, max_order_);
    launch_fill(3 * npts_, 0.0, u_ns_.p, stream_);
    launch_sheval(geom_, 0, rec_.p, items_.p, (int)items_.n, d_trg_slot_off_.p, d_trg_nsrc_.p, evscratch_.p, rho,
                  u_ns_.p, 1.0, 0.0, 0.0, nullptr, stream_);
    launch_pack_src_f64(npts_, d_coords_.p, d_w_.p, rho, src64_.p, stream_);
    far_run(lists_rv_mid_, 1, 0, src64_.p, d_coords_.p, d_norms_.p, nullptr, chunks_.p, (int)chunks_.n,
            1.0 / (8.0 * kPi), u_ns_.p, nullptr);
    launch_rigid_avg(geom_, u_ns_.p, inv_mu, vel_ns_.p, stream_);
    launch_far_centre(src64_.p, d_pos_.p, lists_centre_.ngroups, lists_centre_.order.p, lists_centre_.groups.p,
                      lists_centre_.iv.p, centre_part_.p, nsph_, stream_);
    launch_centre_finish(nsph_, lists_centre_.split, centre_part_.p, d_rad_.p, vel_ns_.p, inv_mu, vel, stream_);
    SPHSTK_HIP(hipGetLastError());
}

void Operator::lop(const double *in, double *out, double cLOP, const int *active) {
    if (std::fabs(cLOP) > 1e-9) launch_lop(geom_, in, out, cLOP, active, stream_);
}

void Operator::apply_into(const double *in, double *out, int prec, const int *active) {
    launch_fill(3 * npts_, 0.0, out, stream_);
    p2p(in, out, c_[0], c_[1], c_[2], prec, active);
    lop(in, out, c_[3], active);
}

// The inner operator's apply as one HIP graph: captured on the first call (or
// when the `active` array changes), replayed afterwards between two copies to
// and from fixed buffers.  The captured work is kernels only (the fills are
// kernels, not memsets); timing events are not recorded while capturing.
void Operator::apply_inner_into(const double *in, double *out, const int *active) {
    if (!(flags_ & 8)) {
        apply_into(in, out, 1, active);
        return;
    }
    const int64_t n = 3 * npts_;
    if (gx_inner_ && g_act_ == active) {
        SPHSTK_HIP(hipMemcpyAsync(g_in_.p, in, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
        SPHSTK_HIP(hipGraphLaunch(gx_inner_, stream_));
        SPHSTK_HIP(hipMemcpyAsync(out, g_out_.p, sizeof(double) * n, hipMemcpyDeviceToDevice, stream_));
        return;
    }
    // this call runs eagerly, which also completes every lazy allocation and
    // attribute setting before the capture
    apply_into(in, out, 1, active);
    if (gx_inner_) {
        SPHSTK_HIP(hipGraphExecDestroy(gx_inner_));
        SPHSTK_HIP(hipGraphDestroy(g_inner_));
        gx_inner_ = nullptr;
        g_inner_ = nullptr;
    }
    g_in_.alloc(std::max<int64_t>(n, 1));
    g_out_.alloc(std::max<int64_t>(n, 1));
    SPHSTK_HIP(hipStreamSynchronize(stream_));
    SPHSTK_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    capturing_ = true;
    try {
        apply_into(g_in_.p, g_out_.p, 1, active);
    } catch (...) {
        capturing_ = false;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    capturing_ = false;
    SPHSTK_HIP(hipStreamEndCapture(stream_, &g_inner_));
    SPHSTK_HIP(hipGraphInstantiate(&gx_inner_, g_inner_, nullptr, nullptr, 0));
    g_act_ = active;
}

// apply (cpp:153-236): Y := beta Y + alpha A X
void Operator::apply(const double *X, double *Y, double alpha, double beta, int prec) {
    if (prec == 1)
        apply_inner_into(X, tmp_.p, nullptr);
    else
        apply_into(X, tmp_.p, prec, nullptr);
    launch_axpby(3 * npts_, alpha, tmp_.p, beta, Y, stream_);
}

// projectNullSpace (cpp:239-260)
// Values of the SH expansion of `in` (device DOF vector) at every sphere's two
// poles, lab frame: out[6 S + (0..2 north, 3..5 south)] (Shexp::calcPoleValue,
// Shexp.cpp:382-405; used by the grid-with-poles output).  The coefficients
// come from the device analysis; the 6 synthesis rows per order are evaluated
// on the host.
void Operator::pole_values(const double *in, double *out) {
    if (nsph_ == 0) return;
    launch_analysis(geom_, in, coef_.p, nullptr, stream_, max_order_);
    std::vector<double> coef(std::max<int64_t>(coef_off_[nsph_], 1));
    SPHSTK_HIP(hipMemcpyAsync(coef.data(), coef_.p, sizeof(double) * coef_off_[nsph_], hipMemcpyDeviceToHost,
                              stream_));
    SPHSTK_HIP(hipStreamSynchronize(stream_));
    std::vector<std::vector<double>> rows(kMaxOrder + 1);
    for (int S = 0; S < nsph_; S++) {
        const int p = order_[S], ncoef = 6 * tri_size(p);
        if (rows[p].empty()) rows[p] = pole_rows(p);
        const double *c = coef.data() + coef_off_[S];
        for (int pole = 0; pole < 2; pole++) {
            double body[3];
            for (int k = 0; k < 3; k++) {
                const double *r = rows[p].data() + size_t(3 * pole + k) * ncoef;
                double s = 0.0;
                for (int j = 0; j < ncoef; j++) s += r[j] * c[j];
                body[k] = s;
            }
            qrot_host(&quat_[4 * S], body, out + 6 * S + 3 * pole);
        }
    }
}

void Operator::project(double *inout) {
    if (nsph_ == 0) return;
    launch_analysis(geom_, inout, coef_.p, nullptr, stream_, max_order_);
    launch_expand(geom_, 1, coef_.p, rec_.p, nullptr, stream_, max_order_);
    launch_synthesis(geom_, rec_.p, self_items_.p, (int)self_items_.n, inout, stream_);
    SPHSTK_HIP(hipGetLastError());
}

// Blocks (B_n^-1 - 2 I) per SH degree n of the single-sphere operator
// B = 1/2 I + K'_PV P + L restricted to degree-n coefficients (F_r, F_s, F_t):
// B_n = I + K_ext(n) + L(n) with the exterior traction spectra of a sphere
//   K_ext: (F_r, F_s) -> -(n (4n^2+6n-1) F_r + 3n(n+1) F_s, 3 F_r + (4n^3+6n^2-n+3) F_s) / D_n,
//          F_t -> -(n+2)/(2n+1) F_t,  D_n = (2n-1)(2n+1)(2n+3),
// and the rigid-body projector L on degree 1 (translations F_r = F_s, rotations
// F_t): (F_r, F_s) -> (F_r + 2 F_s)/3 (1, 1), F_t -> F_t.  Degree 0 (the normal
// field) and the part of a density beyond degree p see 1/2 I, hence the "- 2 I":
// M^-1 x = 2 x + synthesis((B_n^-1 - 2 I) analysis(x)).  Scale-free in R.
static std::vector<double> self_inverse_blocks() {
    std::vector<double> t(5 * (kMaxOrder + 1));
    t[0] = -1.0;
    t[3] = -1.0;
    t[4] = -1.0;
    for (int n = 1; n <= kMaxOrder; n++) {
        const double nn = n, D = (2 * nn - 1) * (2 * nn + 1) * (2 * nn + 3);
        double a = 1.0 - nn * (4 * nn * nn + 6 * nn - 1) / D, b = -3.0 * nn * (nn + 1) / D, c = -3.0 / D,
               e = 1.0 - (4 * nn * nn * nn + 6 * nn * nn - nn + 3) / D, bt = 1.0 - (nn + 2) / (2 * nn + 1);
        if (n == 1) {
            a += 1.0 / 3.0;
            b += 2.0 / 3.0;
            c += 1.0 / 3.0;
            e += 2.0 / 3.0;
            bt += 1.0;
        }
        const double det = a * e - b * c;
        t[5 * n + 0] = e / det - 2.0;
        t[5 * n + 1] = -b / det;
        t[5 * n + 2] = -c / det;
        t[5 * n + 3] = a / det - 2.0;
        t[5 * n + 4] = 1.0 / bt - 2.0;
    }
    return t;
}

// out = M^-1 in, M the block-diagonal (one block per sphere) part of the
// operator: the exact inverse of each sphere's self operator (isolated-sphere
// preconditioner).  Suspensions not active keep out = 2 in.
void Operator::precondition(const double *in, double *out, const int *active) {
    if (nsph_ == 0) return;
    if (!self_inv_.p) self_inv_.upload(self_inverse_blocks());
    launch_analysis(geom_, in, coef_.p, active, stream_, max_order_);
    launch_degree_blocks(geom_, coef_.p, self_inv_.p, active, stream_);
    launch_expand(geom_, 1, coef_.p, rec_.p, active, stream_, max_order_);
    launch_fill(3 * npts_, 0.0, out, stream_);
    launch_synthesis(geom_, rec_.p, self_items_.p, (int)self_items_.n, out, stream_, active);
    launch_axpby(3 * npts_, 2.0, in, 1.0, out, stream_);
}

void Operator::grid(double *weights, double *coords, double *norms, int64_t *index, int64_t *length) const {
    if (weights) std::copy(w_.begin(), w_.end(), weights);
    if (coords) std::copy(coords_.begin(), coords_.end(), coords);
    if (norms) std::copy(norms_.begin(), norms_.end(), norms);
    for (int i = 0; i < nsph_; i++) {
        if (index) index[i] = pt_off_[i];
        if (length) length[i] = pt_off_[i + 1] - pt_off_[i];
    }
}

void Operator::near_pairs(int32_t *trg, int32_t *src) const {
    for (int t = 0; t < nsph_; t++)
        for (int64_t k = near_off_[t]; k < near_off_[t + 1]; k++) {
            if (trg) trg[k] = t;
            if (src) src[k] = near_src_[k];
        }
}

void Operator::kernel_stats(int kind, int prec, double *ms, int64_t *launches, double *pairs) {
    SPHSTK_REQUIRE(kind == 0 || kind == 1, "kind must be 0 (traction) or 1 (single layer)");
    SPHSTK_REQUIRE(prec >= 0 && prec < kNumPrec,
                   "prec must be 0 (FP64), 1 (FP32), 2 (FP64 multi-level) or 3 (coarse FP32)");
    if (prec == 1) wait_inner_levels();
    harvest_events();
    if (ms) *ms = stats_[kind][prec].ms;
    if (launches) *launches = stats_[kind][prec].launches;
    if (pairs) {
        const LevelSet *S = (prec == 2 && outer_pay_ == 1) ? &ls64_
                            : (prec == 1 && multi_level_) ? &ls32_
                            : (prec == 3 && multi_level_) ? &lsc_
                                                          : nullptr;
        if (S && S->ready) {
            *pairs = S->mid.pairs + (S->agg ? S->agg->pairs : 0.0);
            for (auto &L : S->levels) *pairs += L->lists.pairs;
        } else {
            *pairs = lists_fine_.pairs * (prec == 0 ? ncopy_ : 1); // copies: per-copy lists, all copies summed
        }
    }
}

void Operator::reset_stats() {
    harvest_events();
    for (auto &a : stats_)
        for (auto &b : a) b = KernelStats{};
}

// ---------------------------------------------------------------------------
// Mobility: SphereSTKMobMat (BIE/SphereSTKMobMat.cpp:6-307) with A = 0.5 I + K' + L (:17)

Mobility::Mobility(int nsusp, const int *counts, const double *pos, const double *radius, const double *quat,
                   const int *order, double viscosity, const SolverParams &prm, int precision, int flags, int ncopy)
    : mu_(viscosity), prm_(prm), prec_(precision) {
    inner_restart_ = std::max(prm.restart, 60);
    inner_precond_ = (flags & 16) != 0;
    explicit_refine_ = (flags & 64) != 0;
    SPHSTK_REQUIRE(std::isfinite(viscosity) && viscosity > 0, "viscosity must be > 0");
    SPHSTK_REQUIRE(precision == 0 || precision == 1, "precision must be 0 (FP64) or 1 (mixed)");
    if (const char *v = std::getenv("SPHSTK_VERBOSE")) verbose_ = std::atoi(v) != 0;
    SPHSTK_REQUIRE(ncopy == 1 || precision == 0, "several right-hand sides on one geometry: FP64 form only");
    A_ = std::make_unique<Operator>(nsusp, counts, pos, radius, quat, order, 0.5, 0.0, 1.0, 1.0, flags, ncopy);
    const int64_t n = std::max<int64_t>(A_->ndof(), 1);
    rho_.alloc(n);
    b_.alloc(n);
    x_.alloc(n);
    r_.alloc(n);
    d_.alloc(n);
    launch_fill(n, 0.0, x_.p, A_->stream());
    x_zero_ = true;
    force_.alloc(std::max(6 * A_->nsph(), 1));
    vel_.alloc(std::max(6 * A_->nsph(), 1));
    act_.alloc(A_->nsusp());
    scal_.alloc(4 * (size_t)A_->nsusp());
    d_nact_.alloc(1);
    coarse_inner_ = prec_ == 1 && (flags & 128) == 0 && A_->has_coarse();
    if (prec_ == 1) A_->prepare_mixed(outer_prec_ == 2, coarse_inner_);
    SPHSTK_HIP(hipStreamSynchronize(A_->stream()));
    // last, so that nothing after them can throw
    for (auto &e : nact_ev_) SPHSTK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    SPHSTK_HIP(hipHostMalloc((void **)&nact_host_, 2 * sizeof(int), hipHostMallocDefault));
    nact_host_[0] = nact_host_[1] = 1;
}

Mobility::~Mobility() {
    if (A_) (void)hipStreamSynchronize(A_->stream());
    for (auto &e : nact_ev_)
        if (e) (void)hipEventDestroy(e);
    if (nact_host_) (void)hipHostFree(nact_host_);
}

void Mobility::seg_norms(const double *v, std::vector<double> &out) {
    Operator &A = *A_;
    const int B = A.nsusp();
    launch_seg_dot(v, v, A.chunks(), A.nchunks(), A.scratch_partial(), nullptr, A.stream());
    launch_seg_reduce(A.scratch_partial(), A.seg_chunk_off(), B, scal_.p, 0, nullptr, A.stream());
    out.resize(B);
    SPHSTK_HIP(hipMemcpyAsync(out.data(), scal_.p, sizeof(double) * B, hipMemcpyDeviceToHost, A.stream()));
    SPHSTK_HIP(hipStreamSynchronize(A.stream()));
    for (auto &x : out) x = std::sqrt(std::max(x, 0.0));
}

// Restarted GMRES over all suspensions in lock step (Belos GMRES with Num Blocks =
// restart, IMGS orthogonalisation = two passes of modified Gram-Schmidt, residual
// scaled by ||b||; mobilitySolver.yaml:3-16).  Each suspension has its own
// Krylov space, Hessenberg matrix and Givens rotations; converged suspensions
// drop out of the operator applies through the device `active` mask.
void Mobility::gmres(const double *b, double *x, bool x_zero, double tol, int prec, std::vector<int> &act,
                     std::vector<SegInfo> &info, int &applies, double *res) {
    Operator &A = *A_;
    hipStream_t st = A.stream();
    const int B = A.nsusp();
    const int m = std::max(1, prec != 0 ? inner_restart_ : prm_.restart);
    const int64_t n = A.ndof();
    if (V_.n < size_t(m + 1) * n) V_.alloc(size_t(m + 1) * n);
    h_.alloc(size_t(2 * (m + 1) + 1) * B);
    if (prec != 0 && m + 1 <= 64) mdot_.alloc(size_t(m + 1) * std::max(A.nchunks(), 1));
    coefv_.alloc(size_t(m + 1) * B);
    const Chunk *ch = A.chunks();
    const int nch = A.nchunks();
    const int *sco = A.seg_chunk_off();
    double *part = A.scratch_partial();
    // inner solve: right preconditioning by the isolated-sphere inverse,
    // A M^-1 u = r, d = M^-1 u (the residual stays the unpreconditioned one)
    const bool rprec = (prec == 1 || prec == 3) && inner_precond_;
    if (rprec) {
        z_.alloc(std::max<int64_t>(n, 1));
        zt_.alloc(std::max<int64_t>(n, 1));
    }

    std::vector<double> bn, beta;
    seg_norms(b, bn);
    for (int s = 0; s < B; s++)
        if (bn[s] == 0.0 && act[s]) {
            act[s] = 0;
            info[s].converged = 1;
            info[s].relres = 0.0;
        }
    auto upload_act = [&](const std::vector<int> &a) {
        SPHSTK_HIP(hipMemcpyAsync(act_.p, a.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
    };
    // r = b - A x
    if (x_zero) {
        launch_axpby(n, 1.0, b, 0.0, r_.p, st);
    } else {
        upload_act(act);
        A.apply_into(x, r_.p, prec, act_.p);
        applies++;
        launch_axpby(n, 1.0, b, -1.0, r_.p, st);
    }
    if (res) launch_axpby(n, 1.0, r_.p, 0.0, res, st);
    seg_norms(r_.p, beta);
    for (int s = 0; s < B; s++)
        if (act[s]) {
            info[s].relres = beta[s] / bn[s];
            if (info[s].relres <= tol) {
                act[s] = 0;
                info[s].converged = 1;
            }
        }
    // The least-squares bookkeeping (Hessenberg columns, Givens rotations,
    // implicit residuals, the shrinking active set) runs on the device, so a
    // restart cycle needs no host synchronisation: the host reads the count of
    // suspensions still active one iteration late (a pinned copy behind an
    // event) to stop launching, and reads the per-suspension state once per
    // cycle.  With SPHSTK_VERBOSE it also waits for every iteration's residual.
    d_H_.alloc(size_t(B) * (m + 1) * m);
    d_cs_.alloc(size_t(B) * m);
    d_sn_.alloc(size_t(B) * m);
    d_g_.alloc(size_t(B) * (m + 1));
    d_bn_.alloc(B);
    d_beta_.alloc(B);
    d_relres_.alloc(B);
    d_cyc_.alloc(B);
    d_kdim_.alloc(B);
    d_impok_.alloc(B);
    SPHSTK_HIP(hipMemcpyAsync(d_bn_.p, bn.data(), sizeof(double) * B, hipMemcpyHostToDevice, st));
    std::vector<double> inv(B), rel(B);
    std::vector<int> kdim(B), impok(B);
    int total = 0;
    auto any = [&](const std::vector<int> &a) { return std::any_of(a.begin(), a.end(), [](int v) { return v != 0; }); };
    while (any(act) && total < prm_.max_iters) {
        const std::vector<int> cyc = act;
        SPHSTK_HIP(hipMemcpyAsync(d_cyc_.p, cyc.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
        SPHSTK_HIP(hipMemcpyAsync(d_beta_.p, beta.data(), sizeof(double) * B, hipMemcpyHostToDevice, st));
        launch_gmres_init(B, m, d_beta_.p, d_cyc_.p, d_g_.p, act_.p, d_kdim_.p, d_impok_.p, st);
        for (int s = 0; s < B; s++) inv[s] = cyc[s] ? 1.0 / beta[s] : 0.0;
        SPHSTK_HIP(hipMemcpyAsync(scal_.p, inv.data(), sizeof(double) * B, hipMemcpyHostToDevice, st));
        launch_seg_scale(V_.p, r_.p, scal_.p, ch, nch, act_.p, st);
        int kmax = 0;
        for (int j = 0; j < m; j++) {
            double *W = V_.p + size_t(j + 1) * n;
            const double *Vj = V_.p + size_t(j) * n;
            if (rprec) {
                A.precondition(Vj, z_.p, act_.p);
                if (prec == 3)
                    A.apply_coarse_into(z_.p, W, act_.p);
                else
                    A.apply_inner_into(z_.p, W, act_.p);
            } else if (prec == 1)
                A.apply_inner_into(Vj, W, act_.p);
            else if (prec == 3)
                A.apply_coarse_into(Vj, W, act_.p);
            else
                A.apply_into(Vj, W, prec, act_.p);
            applies++;
            // h = V^T w with two orthogonalisation passes.  FP64 solve: IMGS, the
            // reference's Belos setting (mobilitySolver.yaml:3-16), one basis
            // vector at a time.  Inner FP32-operator solve: CGS2 (classical
            // Gram-Schmidt twice: all j+1 products of a pass in one kernel),
            // which stores -h.
            const bool cgs2 = prec != 0 && m + 1 <= 64;
            for (int pass = 0; pass < 2; pass++) {
                if (cgs2) {
                    double *hv = h_.p + size_t(pass * (m + 1)) * B;
                    launch_seg_multi_dot(W, V_.p, n, j + 1, ch, nch, mdot_.p, act_.p, st);
                    launch_seg_multi_reduce(mdot_.p, nch, sco, B, j + 1, hv, 2, act_.p, st);
                    launch_seg_multi_axpy(W, V_.p, n, hv, j + 1, B, ch, nch, act_.p, st);
                    continue;
                }
                for (int i = 0; i <= j; i++) {
                    double *hv = h_.p + size_t(pass * (m + 1) + i) * B;
                    launch_seg_dot(W, V_.p + size_t(i) * n, ch, nch, part, act_.p, st);
                    launch_seg_reduce(part, sco, B, hv, 0, act_.p, st);
                    launch_seg_axpy(W, V_.p + size_t(i) * n, hv, -1.0, ch, nch, act_.p, st);
                }
            }
            double *nv = h_.p + size_t(2 * (m + 1)) * B;
            launch_seg_dot(W, W, ch, nch, part, act_.p, st);
            launch_seg_reduce(part, sco, B, nv, 0, act_.p, st);
            launch_seg_reduce(part, sco, B, scal_.p, 1, act_.p, st);
            launch_seg_scale(W, W, scal_.p, ch, nch, act_.p, st);
            launch_gmres_givens(B, j, m, h_.p, cgs2 ? -1.0 : 1.0, d_H_.p, d_cs_.p, d_sn_.p, d_g_.p, d_bn_.p, tol,
                                act_.p, d_kdim_.p, d_impok_.p, d_relres_.p, d_nact_.p, st);
            SPHSTK_HIP(hipMemcpyAsync(nact_host_ + (j & 1), d_nact_.p, sizeof(int), hipMemcpyDeviceToHost, st));
            SPHSTK_HIP(hipEventRecord(nact_ev_[j & 1], st));
            total++;
            kmax = j + 1;
            if (verbose_) {
                SPHSTK_HIP(hipMemcpyAsync(rel.data(), d_relres_.p, sizeof(double) * B, hipMemcpyDeviceToHost, st));
                SPHSTK_HIP(hipStreamSynchronize(st));
                double worst = 0.0;
                for (int s = 0; s < B; s++)
                    if (cyc[s]) worst = std::max(worst, rel[s]);
                std::fprintf(stderr, "[sphstk] gmres(%s) iter %d: max implicit relres %.3e, %d/%d suspensions active\n",
                             prec == 3 ? "fp32 coarse" : prec ? "fp32 inner" : "fp64", total, worst, nact_host_[j & 1],
                             B);
                std::fflush(stderr);
                if (nact_host_[j & 1] == 0) break;
            }
            if (total >= prm_.max_iters) break;
            // Large problems wait for this iteration's count before launching the
            // next (an iteration is milliseconds; the wait costs tens of
            // microseconds).  Small ones keep one iteration of lookahead, so the
            // host's turnaround never idles the GPU, at the price of one
            // iteration launched after the last one that was needed.
            if (n >= kLookaheadDof) {
                SPHSTK_HIP(hipEventSynchronize(nact_ev_[j & 1]));
                if (nact_host_[j & 1] == 0) break;
            } else if (j >= 1) {
                SPHSTK_HIP(hipEventSynchronize(nact_ev_[(j - 1) & 1]));
                if (nact_host_[(j - 1) & 1] == 0) break;
            }
        }
        // x += V y (back substitution of the rotated Hessenberg systems), over
        // the suspensions of this cycle
        launch_gmres_backsub(B, m, d_H_.p, d_g_.p, d_kdim_.p, d_cyc_.p, coefv_.p, st);
        SPHSTK_HIP(hipMemcpyAsync(act_.p, d_cyc_.p, sizeof(int) * B, hipMemcpyDeviceToDevice, st));
        if (rprec) {
            // x += M^-1 (V y); zero outside the suspensions of this cycle
            launch_fill(n, 0.0, zt_.p, st);
            launch_seg_multi_axpy(zt_.p, V_.p, n, coefv_.p, kmax, B, ch, nch, act_.p, st);
            A.precondition(zt_.p, z_.p, act_.p);
            launch_axpby(n, 1.0, z_.p, 1.0, x, st);
        } else {
            launch_seg_multi_axpy(x, V_.p, n, coefv_.p, kmax, B, ch, nch, act_.p, st);
        }
        if (res) {
            // res = V_{k+1} Omega^T g_k e_k, the cycle's final residual (for the
            // suspensions of this cycle; the others keep theirs)
            launch_gmres_resid_coef(B, m, d_g_.p, d_cs_.p, d_sn_.p, d_kdim_.p, d_cyc_.p, coefv_.p, st);
            launch_seg_multi_comb(res, V_.p, n, coefv_.p, kmax + 1, B, ch, nch, act_.p, st);
        }
        SPHSTK_HIP(hipMemcpyAsync(kdim.data(), d_kdim_.p, sizeof(int) * B, hipMemcpyDeviceToHost, st));
        SPHSTK_HIP(hipMemcpyAsync(impok.data(), d_impok_.p, sizeof(int) * B, hipMemcpyDeviceToHost, st));
        SPHSTK_HIP(hipMemcpyAsync(rel.data(), d_relres_.p, sizeof(double) * B, hipMemcpyDeviceToHost, st));
        SPHSTK_HIP(hipStreamSynchronize(st));
        for (int s = 0; s < B; s++)
            if (cyc[s]) {
                info[s].iters += kdim[s];
                if (kdim[s] > 0) info[s].relres = rel[s];
            }
        // Belos' explicit residual test (yaml :12-13) for the FP64 solve.  The
        // FP32-operator solves of the mixed form skip it once the implicit
        // residual has converged: the refinement computes the residual of the
        // updated iterate right afterwards (or takes the solve's own).
        const bool explicit_final = prm_.explicit_check != 0 && prec == 0;
        bool need_res = explicit_final;
        for (int s = 0; s < B; s++)
            if (cyc[s] && !impok[s]) need_res = true;
        if (need_res) {
            A.apply_into(x, r_.p, prec, act_.p);
            applies++;
            launch_axpby(n, 1.0, b, -1.0, r_.p, st);
            seg_norms(r_.p, beta);
        }
        for (int s = 0; s < B; s++) {
            if (!cyc[s]) continue;
            bool conv = impok[s] != 0;
            if (need_res) {
                info[s].relres = beta[s] / bn[s];
                if (explicit_final) conv = conv && info[s].relres <= tol;
            }
            if (conv) {
                info[s].converged = 1;
                act[s] = 0;
            }
        }
    }
}

// The mixed form's inner solve A32 d = rhs (to itol, relative to |rhs|) by
// refinement with the coarse FP32 operator Ac (the FP32 far field on the coarse
// level set: about a third of A32's pair sums, relative error ~1e-2): d = 0,
// r = rhs; repeat { GMRES on Ac e = r to max(kCoarseTol, tol needed); d += e;
// r = rhs - A32 d (one A32 apply) } until |r| <= itol |rhs|.  Each step gains
// about Ac's error (1e-2), the steps' Krylov spaces are those of the cheap
// operator, and A32 enters only through the residuals.  If a step gains less
// than 2x on some suspension, the rest is solved with A32 itself.  On return
// res = rhs - A32 d, exact up to A32's rounding.  Without the coarse operator
// (its flag, or an FP32 direct far field) this is GMRES on A32.
constexpr double kCoarseTol = 1e-2;
constexpr int kMaxCoarseSteps = 8;
void Mobility::inner_solve(const double *rhs, double *d, double itol, std::vector<int> &act,
                           std::vector<SegInfo> &info, int &app32, double *res) {
    if (!coarse_inner_) {
        gmres(rhs, d, true, itol, 1, act, info, app32, res);
        return;
    }
    Operator &A = *A_;
    hipStream_t st = A.stream();
    const int B = A.nsusp();
    const int64_t n = A.ndof();
    ce_.alloc(std::max<int64_t>(n, 1));
    ct_.alloc(std::max<int64_t>(n, 1));
    std::vector<double> bn, rn;
    seg_norms(rhs, bn);
    launch_fill(n, 0.0, d, st);
    launch_axpby(n, 1.0, rhs, 0.0, res, st);
    rn = bn;
    std::vector<int> a(B, 0), stalled(B, 0);
    std::vector<double> ones(B, 1.0);
    SPHSTK_HIP(hipMemcpyAsync(scal_.p + 3 * B, ones.data(), sizeof(double) * B, hipMemcpyHostToDevice, st));
    auto live = [&]() {
        bool any = false;
        for (int s = 0; s < B; s++) {
            a[s] = act[s] && bn[s] > 0.0 && rn[s] > itol * bn[s] && !stalled[s];
            any = any || a[s];
        }
        return any;
    };
    for (int step = 0; step < kMaxCoarseSteps && live(); step++) {
        double worst = 0.0;
        for (int s = 0; s < B; s++)
            if (a[s]) worst = std::max(worst, rn[s] / bn[s]);
        const double ctol = std::max(kCoarseTol, 0.5 * itol / worst);
        std::vector<SegInfo> sub(B);
        std::vector<int> sa = a;
        launch_fill(n, 0.0, ce_.p, st);
        gmres(res, ce_.p, true, ctol, 3, sa, sub, app32, nullptr);
        // d += e (e is zero outside this step's suspensions); ct = A32 d there;
        // res = rhs - ct on those suspensions only
        launch_axpby(n, 1.0, ce_.p, 1.0, d, st);
        SPHSTK_HIP(hipMemcpyAsync(act_.p, a.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
        A.apply_inner_into(d, ct_.p, act_.p);
        app32++;
        launch_seg_scale(res, rhs, scal_.p + 3 * B, A.chunks(), A.nchunks(), act_.p, st);
        launch_seg_axpy(res, ct_.p, scal_.p + 3 * B, -1.0, A.chunks(), A.nchunks(), act_.p, st);
        std::vector<double> prev = rn;
        seg_norms(res, rn);
        for (int s = 0; s < B; s++)
            if (a[s]) {
                info[s].iters += sub[s].iters;
                if (rn[s] > 0.5 * prev[s]) stalled[s] = 1;
            }
        if (verbose_) {
            double w = 0.0;
            for (int s = 0; s < B; s++)
                if (a[s]) w = std::max(w, rn[s] / bn[s]);
            std::fprintf(stderr, "[sphstk] coarse step %d: tolerance %.2e, FP32 relres %.3e\n", step + 1, ctol, w);
        }
    }
    // what is left (a stalled or unfinished suspension): GMRES on A32 for the
    // rest of the correction, from a copy of the residual
    bool rest = false;
    for (int s = 0; s < B; s++) {
        a[s] = act[s] && bn[s] > 0.0 && rn[s] > itol * bn[s];
        rest = rest || a[s];
    }
    if (rest) {
        double tmin = 1.0;
        for (int s = 0; s < B; s++)
            if (a[s]) tmin = std::min(tmin, itol * bn[s] / rn[s]);
        std::vector<SegInfo> sub(B);
        launch_axpby(n, 1.0, res, 0.0, ct_.p, st);
        launch_fill(n, 0.0, ce_.p, st);
        gmres(ct_.p, ce_.p, true, tmin, 1, a, sub, app32, res);
        launch_axpby(n, 1.0, ce_.p, 1.0, d, st);
        for (int s = 0; s < B; s++) info[s].iters += sub[s].iters;
        seg_norms(res, rn);
        if (verbose_) std::fprintf(stderr, "[sphstk] inner solve finished with the FP32 operator\n");
    }
    for (int s = 0; s < B; s++)
        if (act[s] && bn[s] > 0.0) info[s].relres = rn[s] / bn[s];
}

// solveMob (BIE/SphereSTKMobMat.cpp:146-286)
void Mobility::solve(const double *force, double *vel, std::vector<SegInfo> &info, int &outer, int &app64,
                     int &app32) {
    Operator &A = *A_;
    hipStream_t st = A.stream();
    const int B = A.nsusp();
    const int N = A.nsph();
    const int64_t n = A.ndof();
    info.assign(B, SegInfo{});
    outer = 0;
    app64 = app32 = 0;
    if (N == 0) return;
    SPHSTK_HIP(hipMemcpyAsync(force_.p, force, sizeof(double) * 6 * N, hipMemcpyHostToDevice, st));
    launch_rho0(A.geom(), force_.p, rho_.p, st);
    // b = -(1/2 I + K') rho0   (:196)
    launch_fill(n, 0.0, b_.p, st);
    const int p64 = prec_ == 0 ? 0 : outer_prec_; // FP64 direct, or FP64 multi-level in the mixed form
    if (prec_ == 0)
        A.p2p(rho_.p, b_.p, -0.5, 0.0, -1.0, 0, nullptr);
    else // distant spheres' rho0 in closed form (exact to the grid quadrature's ~(R/d)^(2p+2))
        A.rigid_p2p(force_.p, rho_.p, b_.p, -0.5, -1.0);
    app64++;
    std::vector<double> bn;
    seg_norms(b_.p, bn);
    std::vector<int> act(B, 0);
    std::vector<double> keep(B, 1.0);
    const std::vector<int64_t> &so = A.susp_sph_off();
    for (int s = 0; s < B; s++) {
        if (so[s + 1] == so[s]) {
            info[s].converged = 1;
            continue;
        }
        if (bn[s] < 1e-5) { // (:216-219) forcing ~ zero: x = 0
            info[s].shortcut = 1;
            info[s].converged = 1;
            keep[s] = 0.0;
        } else {
            act[s] = 1;
        }
    }
    if (!x_zero_) {
        SPHSTK_HIP(hipMemcpyAsync(scal_.p + B, keep.data(), sizeof(double) * B, hipMemcpyHostToDevice, st));
        launch_seg_scale(x_.p, x_.p, scal_.p + B, A.chunks(), A.nchunks(), nullptr, st);
    }
    if (prec_ == 0) {
        gmres(b_.p, x_.p, x_zero_, prm_.tol, 0, act, info, app64);
        outer = 1;
    } else {
        // iterative refinement: r = b - A64 x; inner FP32-far-field GMRES on A d = r; x += d.
        // A step's new residual is r - A64 d.  The inner solve returns its own
        // residual rin = r - A32 d (from its basis, no apply), so
        //   r - A64 d = rin + (A32 - A64) d,
        // and |(A32 - A64) d| ~ eps |A32 d| with eps the inner operator's
        // relative error, measured per suspension at every explicit step as
        // |r_explicit - rin| / |A32 d|.  Once the corrections are small, that
        // term is below the tolerance: a step whose bound, with eps doubled
        // for a correction whose content differs from the measured one, stays
        // under tol / 2 takes rin as the residual, adds the bound to a running
        // error bound, and skips the FP64 apply.  relres is then
        // (|r| + bound) / |b|.
        std::vector<double> rn, eps(B, -1.0), errb(B, 0.0), rinn, a32n, dn;
        if (x_zero_) {
            launch_axpby(n, 1.0, b_.p, 0.0, r_.p, st);
        } else {
            A.apply_into(x_.p, r_.p, p64, nullptr);
            app64++;
            launch_axpby(n, 1.0, b_.p, -1.0, r_.p, st);
        }
        seg_norms(r_.p, rn);
        for (int s = 0; s < B; s++)
            if (act[s]) info[s].relres = rn[s] / bn[s];
        rin_.alloc(n);
        while (outer < prm_.max_refine) {
            std::vector<int> a2(B, 0);
            bool anyact = false;
            for (int s = 0; s < B; s++) {
                if (act[s] && info[s].relres <= prm_.tol) {
                    info[s].converged = 1;
                    act[s] = 0;
                }
                a2[s] = act[s];
                anyact = anyact || act[s];
            }
            if (!anyact) break;
            std::vector<SegInfo> inner(B);
            launch_fill(n, 0.0, d_.p, st);
            std::vector<int> ia = a2;
            // inexact inner solve: a refinement step only has to take the
            // residual below the target, so its tolerance is the larger of
            // inner_tol and a tenth of tol / (current relative residual)
            double worst = 0.0;
            for (int s = 0; s < B; s++)
                if (a2[s]) worst = std::max(worst, info[s].relres);
            const double itol = std::max({prm_.inner_tol, prm_.tol, worst > 0 ? 0.1 * prm_.tol / worst : 0.0});
            // the inner solve overwrites r_, so its right-hand side is a copy of the outer residual
            gmres_inner_rhs_.alloc(n);
            launch_axpby(n, 1.0, r_.p, 0.0, gmres_inner_rhs_.p, st);
            inner_solve(gmres_inner_rhs_.p, d_.p, itol, ia, inner, app32, rin_.p);
            SPHSTK_HIP(hipMemcpyAsync(act_.p, a2.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
            std::vector<double> ones(B, 1.0);
            SPHSTK_HIP(hipMemcpyAsync(scal_.p + 2 * B, ones.data(), sizeof(double) * B, hipMemcpyHostToDevice, st));
            launch_seg_axpy(x_.p, d_.p, scal_.p + 2 * B, 1.0, A.chunks(), A.nchunks(), act_.p, st);
            // A32 d = r - rin (into the right-hand side's copy) and the norms
            launch_axpby(n, -1.0, rin_.p, 1.0, gmres_inner_rhs_.p, st);
            seg_norms(rin_.p, rinn);
            seg_norms(gmres_inner_rhs_.p, a32n);
            // per suspension: the updated residual where its bound allows,
            // else an FP64 apply (masked to those suspensions)
            std::vector<int> upd(B, 0), exp_(B, 0);
            int nupd = 0, nexp = 0;
            for (int s = 0; s < B; s++) {
                if (!a2[s]) continue;
                const bool ok = !explicit_refine_ && eps[s] >= 0.0 &&
                                errb[s] + 2.0 * eps[s] * a32n[s] <= 0.5 * prm_.tol * bn[s];
                (ok ? upd : exp_)[s] = 1;
                (ok ? nupd : nexp)++;
            }
            if (verbose_) {
                double we = -1.0, wb = 0.0;
                for (int s = 0; s < B; s++)
                    if (a2[s]) {
                        we = std::max(we, eps[s]);
                        wb = std::max(wb, (errb[s] + 2.0 * eps[s] * a32n[s]) / bn[s]);
                    }
                std::fprintf(stderr, "[sphstk] refinement step %d: max inner-operator error %.3e, residual bound %.3e: "
                             "%d updated residuals, %d FP64 residuals\n", outer + 1, we, wb, nupd, nexp);
            }
            if (nupd && nexp) launch_axpby(n, 1.0, rin_.p, 0.0, d_.p, st); // the correction is in x already
            if (nexp) {
                SPHSTK_HIP(hipMemcpyAsync(act_.p, exp_.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
                A.apply_into(x_.p, r_.p, p64, act_.p);
                app64++;
                launch_axpby(n, 1.0, b_.p, -1.0, r_.p, st);
                seg_norms(r_.p, rn);
                // the inner operator's error on this step's correction
                launch_axpby(n, 1.0, r_.p, -1.0, rin_.p, st);
                seg_norms(rin_.p, dn);
                for (int s = 0; s < B; s++)
                    if (exp_[s]) {
                        info[s].relres = rn[s] / bn[s];
                        eps[s] = a32n[s] > 0.0 ? dn[s] / a32n[s] : -1.0;
                        errb[s] = 0.0;
                    }
                if (verbose_) {
                    double we = 0.0, wr = 0.0;
                    for (int s = 0; s < B; s++)
                        if (exp_[s]) {
                            we = std::max(we, eps[s]);
                            wr = std::max(wr, info[s].relres);
                        }
                    std::fprintf(stderr, "[sphstk] refinement step %d: FP64 relres %.3e, measured inner-operator error %.3e\n",
                                 outer + 1, wr, we);
                }
            }
            if (nupd) {
                const double *src = nexp ? d_.p : rin_.p;
                if (nexp) {
                    // r = rin on the updated suspensions only (ones at scal_ + 2B)
                    SPHSTK_HIP(hipMemcpyAsync(act_.p, upd.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
                    launch_seg_scale(r_.p, src, scal_.p + 2 * B, A.chunks(), A.nchunks(), act_.p, st);
                } else {
                    launch_axpby(n, 1.0, src, 0.0, r_.p, st);
                }
                for (int s = 0; s < B; s++)
                    if (upd[s]) {
                        errb[s] += 2.0 * eps[s] * a32n[s];
                        info[s].relres = (rinn[s] + errb[s]) / bn[s];
                    }
            }
            for (int s = 0; s < B; s++)
                if (a2[s]) info[s].iters += inner[s].iters;
            outer++;
        }
        for (int s = 0; s < B; s++)
            if (act[s] && info[s].relres <= prm_.tol) info[s].converged = 1;
    }
    x_zero_ = false;
    // rho = rho0 + x; u = S[rho]; rigid-body averages / mu   (:253-285)
    launch_axpby(n, 1.0, x_.p, 1.0, rho_.p, st);
    if (prec_ == 0) {
        launch_fill(n, 0.0, b_.p, st);
        A.p2p(rho_.p, b_.p, 0.0, 1.0, 0.0, 0, nullptr);
        app64++;
        launch_rigid_avg(A.geom(), b_.p, 1.0 / mu_, vel_.p, st);
    } else {
        A.rigid_velocity(rho_.p, 1.0 / mu_, vel_.p);
    }
    SPHSTK_HIP(hipMemcpyAsync(vel, vel_.p, sizeof(double) * 6 * N, hipMemcpyDeviceToHost, st));
    SPHSTK_HIP(hipStreamSynchronize(st));
}

void Mobility::get_density(double *rho) const {
    SPHSTK_HIP(hipStreamSynchronize(A_->stream()));
    SPHSTK_HIP(hipMemcpy(rho, rho_.p, sizeof(double) * A_->ndof(), hipMemcpyDeviceToHost));
}
void Mobility::get_unknown(double *x) const {
    SPHSTK_HIP(hipStreamSynchronize(A_->stream()));
    SPHSTK_HIP(hipMemcpy(x, x_.p, sizeof(double) * A_->ndof(), hipMemcpyDeviceToHost));
}
void Mobility::set_unknown_zero() {
    launch_fill(A_->ndof(), 0.0, x_.p, A_->stream());
    x_zero_ = true;
}

void Mobility::set_unknown(const double *x) {
    SPHSTK_HIP(hipStreamSynchronize(A_->stream()));
    x_zero_ = std::all_of(x, x + A_->ndof(), [](double v) { return v == 0.0; });
    SPHSTK_HIP(hipMemcpy(x_.p, x, sizeof(double) * A_->ndof(), hipMemcpyHostToDevice));
}

// The constructor's warm start from the density stored on the spheres
// (BIE/SphereSTKMobMat.cpp:67-93): x = P(density), P the null-space projection.
void Mobility::set_density(const double *rho) {
    set_unknown(rho);
    if (!x_zero_) A_->project(x_.p);
}

} // namespace sphstk
