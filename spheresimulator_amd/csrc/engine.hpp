// Host engine of libsphstk.so: the MI355X-native counterparts of
// SphereSTKSHOperator (BIE/SphereSTKSHOperator.hpp:112-231) and SphereSTKMobMat
// (BIE/SphereSTKMobMat.hpp:26-79), over a batch of independent suspensions
// (a single suspension is a batch of one).
#pragma once

#include "kernels.hpp"

#include <future>
#include <memory>
#include <string>

namespace sphstk {

// Owning device buffer.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        n = count;
        if (count) SPHSTK_HIP(hipMalloc(&p, sizeof(T) * count));
    }
    void upload(const std::vector<T> &v) {
        alloc(v.size());
        if (!v.empty()) SPHSTK_HIP(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    }
};

// Owning HIP stream (default priority, non-blocking).  A member of this type
// releases the stream also when a constructor throws after creating it.
struct StreamOwner {
    hipStream_t s = nullptr;
    StreamOwner() = default;
    StreamOwner(const StreamOwner &) = delete;
    StreamOwner &operator=(const StreamOwner &) = delete;
    ~StreamOwner() {
        if (s) (void)hipStreamDestroy(s);
    }
    hipStream_t create() {
        SPHSTK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        return s;
    }
};

// Far-field work lists over one target point set (see build_far_lists).
struct FarLists {
    DevBuf<FarGroup> groups;
    DevBuf<FarInterval> iv;
    DevBuf<int> order;
    DevBuf<double> part; // split partial-sum slots of 3*ntrg doubles
    DevBuf<double> cen;  // sphere-aligned lists: centre of each group's target sphere (4 per group)
    bool aligned = false;
    int ngroups = 0, tpl = 1, split = 1;
    double pairs = 0.0; // algorithmic pairs per launch
    int64_t ntrg = 0;
};

// Half-open ranges [a, b) of sphere indices, sorted and disjoint.
using Runs = std::vector<std::pair<int32_t, int32_t>>;
inline void run_append(Runs &r, int s) {
    if (!r.empty() && r.back().second == s)
        r.back().second = s + 1;
    else
        r.push_back({s, s + 1});
}
Runs runs_of(const std::vector<int> &lst, int s0, int s1, bool include);

// src_susp_off: first source "sphere" (entry of soff) of each suspension when
// the source point set is not the target's sphere set (NULL: susp_sph_off).
void build_far_lists(FarLists &L, int nsusp, const std::vector<int64_t> &susp_sph_off,
                     const std::vector<int64_t> &toff, const std::vector<int64_t> &soff, const std::vector<Runs> &runs,
                     int tpl, int split, int64_t want_items, const double *sph_pos = nullptr,
                     const std::vector<int64_t> *src_susp_off = nullptr);

// One level of the multi-level FP32 far field: per-sphere grids of a lower
// order, their geometry, resampling tables and work lists.
struct FarLevel {
    int64_t npts = 0;
    int max_pts = 0;
    DevBuf<int64_t> d_off_by_sph; // first level point of each sphere
    DevBuf<double> coords, norms, w, sv, tv;
    DevBuf<float> src32, trg32;
    DevBuf<double> src64;
    DevBuf<Chunk> chunks;
    DevBuf<double> d_down[kMaxOrder + 1], d_up[kMaxOrder + 1];
    DevBuf<float> d_down32[kMaxOrder + 1], d_up32[kMaxOrder + 1]; // FP32 copies (FP32 level sets)
    ResampleTabs down, up;
    FarLists lists;
};

// Source-cell aggregation of the two coarsest FP32 levels (the treecode part
// of the far field): cells of kCellSize spheres (the leaves of a balanced k-d
// split of each suspension's centres), each
// represented beyond cell_theta cell radii (kCellTheta on the FP32 operator)
// by the equiv_k equivalent point forces of equiv_tables(equiv_k); a (target sphere, cell) pair whose members all lie in
// levels 2-3 of the target interacts through them on the target's grid of the
// finer of its members' levels.
constexpr int kCellSize = 16;
constexpr double kCellTheta = 4.0;
struct CellAgg {
    int ncell = 0;
    const EquivTables *E = nullptr;   // the set's equivalent-source template
    double pairs = 0.0;               // pair evaluations of the two cell lists (per apply)
    double check_pairs = 0.0;         // pair evaluations of the check velocities (per apply)
    std::vector<int64_t> susp_cell_off; // first cell of each suspension
    DevBuf<float> chk32;              // check points as FP32 target records (zero normals)
    DevBuf<float> eq32;               // equivalent points as FP32 source records (forces per apply)
    DevBuf<double> chk_u;             // check velocities, 3 Kc per cell
    DevBuf<double> P;                 // template fit, column-major (3K x 3Kc)
    DevBuf<double> rho;               // cell radius
    DevBuf<int> susp;                 // cell suspension
    DevBuf<Chunk> chk_chunks;
    FarLists chk_lists;               // check points of a cell x the fine grids of its members
    FarLists lists[2];                // level-2 / level-3 target grids x equivalent points
};

// A multi-level far field: fine-grid sums for the pairs closer than the first
// level's separation (mid), then one coarser grid per separation shell.
constexpr int kNumLevels = 4; // levels of the FP32 sets (source cells on their last two)
constexpr int kMaxLevels = 6;
struct LevelSet {
    int nlev = kNumLevels;   // levels in use
    double alo[kMaxLevels];  // separations (in units of R_s + R_t) where levels 0..nlev-1 start
    int order[kMaxLevels];   // grid order of each level (capped by the sphere's order)
    bool fp64 = false;       // FP64 pair sums (else FP32)
    int tpl = 1;             // targets per lane of its work lists
    std::vector<std::unique_ptr<FarLevel>> levels;
    FarLists mid;
    bool cells = false;               // aggregate sources over cells on levels 2-3
    double cell_theta = kCellTheta;   // cells serve targets beyond cell_theta cell radii
    int equiv_k = kEquivK;            // equivalent point forces per cell
    std::unique_ptr<CellAgg> agg;
    bool ready = false;
};

struct KernelStats {
    double ms = 0.0;
    int64_t launches = 0;
};

struct SolverParams {
    double tol = 1e-8;
    int restart = 30;
    int max_iters = 1000;
    int explicit_check = 1;
    double inner_tol = 3e-5;
    int max_refine = 10;
};

struct SegInfo {
    int iters = 0, converged = 0, shortcut = 0;
    double relres = 0.0;
};

class Operator {
  public:
    // ncopy > 1: the nsusp suspensions are ncopy identical copies of the first
    // nsusp / ncopy (several right-hand sides on one geometry, FP64 form): the
    // far field is computed once per pair for all copies.
    Operator(int nsusp, const int *counts, const double *pos, const double *radius, const double *quat,
             const int *order, double cId, double cSL, double cTrac, double cLOP, int flags, int ncopy = 1);
    ~Operator();

    // sizes
    int nsph() const { return nsph_; }
    int nsusp() const { return nsusp_; }
    int64_t npts() const { return npts_; }
    int64_t ndof() const { return 3 * npts_; }
    int64_t nnear() const { return (int64_t)near_src_.size(); }
    double coeff(int i) const { return c_[i]; }
    hipStream_t stream() const { return stream_; }
    const DevGeom &geom() const { return geom_; }

    // device-pointer operations (all on stream()); `active` is a device array of
    // one int per suspension or nullptr for all.
    // out += cId in + cSL S[in] + cTrac K'[in].  Far field: prec 0 FP64 direct
    // sum, 1 FP32 (multi-level unless disabled), 2 FP64 multi-level.
    static constexpr int kNumPrec = 4;
    void p2p(const double *in, double *out, double cId, double cSL, double cTrac, int prec, const int *active);
    // out += cLOP L[in]
    void lop(const double *in, double *out, double cLOP, const int *active);
    // out = A in with the stored coefficients (overwrites out)
    void apply_into(const double *in, double *out, int prec, const int *active);
    // Y = beta Y + alpha A X
    void apply(const double *X, double *Y, double alpha, double beta, int prec);
    // out = A32 in (the FP32 inner operator); with SPHSTK_FLAG_GRAPH replayed
    // from a HIP graph captured on the first call for this `active` array
    void apply_inner_into(const double *in, double *out, const int *active);
    // the coarse FP32 operator Ac (prec 3): the FP32 operator on the coarse level set lsc_
    void apply_coarse_into(const double *in, double *out, const int *active) { apply_into(in, out, 3, active); }
    bool has_coarse() const { return multi_level_; }
    void project(double *inout);
    // out = M^-1 in with M the per-sphere self operator (exact inverse per sphere)
    void precondition(const double *in, double *out, const int *active);
    // SH expansion of in (device) at the two poles of every sphere, lab frame, 6 per sphere (host out)
    void pole_values(const double *in, double *out);
    // rigid-body velocities (6 per sphere) of the single layer of rho, distant pairs by Faxen's laws
    void rigid_velocity(const double *rho, double inv_mu, double *vel);
    // builds the level sets and final-velocity lists of the mixed-precision solve now
    void prepare_mixed(bool outer_levels, bool coarse);
    // out += cId rho0 + cTrac K'[rho0] for the rigid-body density rho0 of `force`
    // (6 per sphere): self/near and pairs closer than 4 (R_s + R_t) as p2p, distant
    // pairs in closed form at the sphere centres
    void rigid_p2p(const double *force, const double *rho0, double *out, double cId, double cTrac);

    // host accessors
    void grid(double *weights, double *coords, double *norms, int64_t *index, int64_t *length) const;
    void near_pairs(int32_t *trg, int32_t *src) const;
    void kernel_stats(int kind, int prec, double *ms, int64_t *launches, double *pairs);
    void reset_stats();
    int far_tpl() const { return lists_fine_.tpl; }
    int inner_tpl() const { return ls32_.tpl; }
    const std::string &tune_log() const { return tune_log_; }
    int far_split() const { return lists_fine_.split; }

    // segmented-vector helpers (one segment per suspension)
    const Chunk *chunks() const { return chunks_.p; }
    int nchunks() const { return (int)chunks_.n; }
    const int *seg_chunk_off() const { return seg_chunk_off_.p; }
    const std::vector<int64_t> &susp_sph_off() const { return susp_sph_off_; }
    const std::vector<int64_t> &pt_off() const { return pt_off_; }
    double *scratch_partial() { return partial_.p; }

  private:
    void far(int kind, int prec, double coef, const double *in, double *out, const int *active);
    void far_run(FarLists &L, int kind, int prec, const void *src, const double *tcoords, const double *tnorms,
                 const float *trg32, const Chunk *chunks, int nchunks, double coef, double *out, const int *active);
    void ensure_levels(LevelSet &S);
    void far_copies(int kind, double coef, const double *in, double *out, const int *active, hipEvent_t ev_a);
    // tile-shape autotuner: times candidate work-list shapes on this problem
    void autotune_fine(double &budget_s);
    void autotune_inner(double &budget_s);
    double time_far_pass(FarLists &L, int prec, const void *src, const double *tcoords, const double *tnorms,
                         const float *trg32, const Chunk *chunks, int nchunks, double *out);
    void far_levels_pass(LevelSet &S, const double *in, double *out);
    std::vector<std::vector<int>> neighbours_within(double amax) const;
    bool outer_levels_pay();
    void ensure_centre();
    // multi-level far field of `in` into out; ev_a and ev_b (if set) bracket
    // the pair-sum launches
    void far_levels(LevelSet &S, int kind, double coef, const double *in, double *out, const int *active,
                    hipEvent_t ev_a, hipEvent_t ev_b);

    StreamOwner stream_owner_; // declared before every buffer, so destroyed after them
    StreamOwner aux_owner_;    // second stream: self/near terms beside the far field
    int nsph_, nsusp_;
    int64_t npts_;
    static constexpr int kMultiCols = 8; // right-hand sides per shared far-field launch
    int ncopy_ = 1;         // identical copies of one set of suspensions (several right-hand sides)
    DevBuf<double> srcm_;   // source records of the shared far field of the copies
    DevBuf<double> self_inv_; // per-degree blocks of the isolated-sphere inverse
    double c_[4];
    int flags_;
    int max_order_ = 0;
    hipStream_t stream_ = nullptr, aux_ = nullptr;
    hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
    int device_ = 0;

    std::vector<double> pos_, rad_, quat_;
    std::vector<int> order_, sph_susp_;
    std::vector<int64_t> susp_sph_off_, pt_off_, coef_off_, rec_off_, near_off_;
    std::vector<int> near_src_;
    std::vector<double> coords_, norms_, w_;
    std::vector<int32_t> pt_sph_;

    DevBuf<double> d_pos_, d_rad_, d_quat_, d_coords_, d_norms_, d_w_;
    DevBuf<int> d_order_, d_sph_susp_, d_near_src_, d_tile_sph_;
    DevBuf<SphTile> d_tiles_;
    DevBuf<int64_t> d_pt_off_, d_coef_off_, d_rec_off_, d_near_off_;
    DevBuf<double> d_anal_[kMaxOrder + 1];
    DevGeom geom_;

    // far-field work lists: fine grid (FP64 and direct FP32); the multi-level
    // far fields of the FP32 inner operator and of the FP64 outer applies
    FarLists lists_fine_;
    std::vector<Runs> runs_fine_; // source runs of the direct far field per target sphere
    std::string tune_log_;         // autotuner decisions
    LevelSet ls32_, ls64_;
    LevelSet lsc_; // coarse FP32 operator (prec 3): the preconditioning solves of the mixed form
    bool multi_level_ = true, timing_done_ = false;
    int outer_pay_ = -1; // FP64 multi-level outer far field in use: -1 undecided, 0 no, 1 yes
    std::future<void> ls64_build_; // ls64_ being built on a host thread (prepare_mixed)
    void wait_outer_levels();      // joins that build, rethrowing its error
    std::future<void> ls32_build_; // ls32_ likewise, when the coarse operator runs first
    void wait_inner_levels();

    // captured inner-operator apply (SPHSTK_FLAG_GRAPH): g_in_ -> g_out_
    hipGraph_t g_inner_ = nullptr;
    hipGraphExec_t gx_inner_ = nullptr;
    const int *g_act_ = nullptr;
    DevBuf<double> g_in_, g_out_;
    bool capturing_ = false;

    // rigid_velocity: fine-grid sums of the pairs closer than kFaxenSep, sphere-centre sums of the rest
    FarLists lists_rv_mid_, lists_centre_;
    FarLists lists_rigid_; // grid targets x distant sphere centres (closed-form rigid-density traction)
    bool centre_ready_ = false;
    DevBuf<double> centre_part_, vel_ns_, u_ns_;

    // self/near evaluation work items and their scratch
    DevBuf<EvalItem> items_, self_items_;
    DevBuf<EvalItem> d_src_items_;    // items_ grouped by source sphere
    DevBuf<int64_t> d_src_off_;
    DevBuf<int> d_src_list_;          // spheres by order
    SrcItems src_items_;
    DevBuf<int64_t> d_trg_slot_off_;
    DevBuf<int> d_trg_nsrc_;
    DevBuf<double> evscratch_;

    // per-apply scratch
    DevBuf<double> src64_, coef_, rec_, tmp_;
    DevBuf<float> src32_, trg32_;
    DevBuf<Chunk> chunks_;
    DevBuf<int> seg_chunk_off_;
    std::vector<int> seg_chunk_off_host_;
    DevBuf<double> partial_;

    // timing
    struct EvPair {
        hipEvent_t a, b;
        int kind, prec;
    };
    std::vector<EvPair> pending_;
    std::vector<EvPair> free_events_;
    KernelStats stats_[2][kNumPrec];
    void harvest_events();
};

class Mobility {
  public:
    Mobility(int nsusp, const int *counts, const double *pos, const double *radius, const double *quat,
             const int *order, double viscosity, const SolverParams &prm, int precision, int flags, int ncopy = 1);
    ~Mobility();
    Mobility(const Mobility &) = delete;
    Mobility &operator=(const Mobility &) = delete;
    Operator &op() { return *A_; }
    int precision() const { return prec_; }
    const SolverParams &params() const { return prm_; }
    // the next solve starts from x0 = 0
    void set_unknown_zero();
    void set_params(const SolverParams &prm, int precision) {
        prm_ = prm;
        prec_ = precision;
    }
    // host arrays: force/vel 6N
    void solve(const double *force, double *vel, std::vector<SegInfo> &info, int &outer, int &app64, int &app32);
    void get_density(double *rho) const;
    void get_unknown(double *x) const;
    void set_unknown(const double *x);
    // warm start from a stored density: x = P(rho)
    void set_density(const double *rho);

  private:
    // GMRES on the segments flagged in `act` (host, one per suspension); x and b device
    // res (optional): the final residual b - A x of every suspension, from the
    // last cycle's basis (no operator apply)
    void gmres(const double *b, double *x, bool x_zero, double tol, int prec, std::vector<int> &act,
               std::vector<SegInfo> &info, int &applies, double *res = nullptr);
    void seg_norms(const double *v, std::vector<double> &out);

    std::unique_ptr<Operator> A_;
    double mu_;
    SolverParams prm_;
    int prec_;
    int outer_prec_ = 2; // far field of the mixed form's FP64 applies: 2 multi-level, 0 direct
    int inner_restart_ = 30; // GMRES restart of the inner FP32-operator solve
    bool coarse_inner_ = false; // inner solves by refinement with the coarse operator Ac
    DevBuf<double> ce_, ct_;    // coarse correction and A32 d of those solves
    void inner_solve(const double *rhs, double *d, double itol, std::vector<int> &act, std::vector<SegInfo> &info,
                     int &app32, double *res);
    bool inner_precond_ = false; // right-precondition the inner solve by the isolated-sphere inverse
    DevBuf<double> z_, zt_;      // preconditioned basis vector / update
    bool x_zero_ = true; // the stored unknown is known to be zero (cold start)
    bool verbose_ = false; // SPHSTK_VERBOSE=1: residual history on stderr (Belos Verbosity 63)
    DevBuf<double> rho_, b_, x_, r_, d_, V_, h_, scal_, coefv_, force_, vel_, gmres_inner_rhs_, mdot_, rin_;
    bool explicit_refine_ = false; // mixed form: an explicit FP64 residual at every refinement step
    DevBuf<int> act_;
    // device-side GMRES least-squares state (per suspension) and the pinned
    // count of active suspensions read one iteration late behind an event
    DevBuf<double> d_H_, d_cs_, d_sn_, d_g_, d_bn_, d_beta_, d_relres_;
    DevBuf<int> d_cyc_, d_kdim_, d_impok_, d_nact_;
    int *nact_host_ = nullptr;
    hipEvent_t nact_ev_[2] = {nullptr, nullptr};
};

} // namespace sphstk
