// solver.hpp -- a batch of independent DQG v2RDM SDPs solved together on one
// GPU by the boundary-point method.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "dbuf.hpp"
#include "constraints.hpp"
#include "layout.hpp"
#include "orbopt.hpp"
#include "projector.hpp"
#include "sell.hpp"

namespace v2h {


// Gather tables and maps of one problem shape, resident in HBM.
struct Structure {
    explicit Structure(const ProblemDesc& d);
    Layout L;
    int64_t n_primal, n_dual, nsrc;
    SellHost hA, hAT;
    std::vector<double> b;
    std::vector<int4> unitsA, unitsAT;   // per-instance workgroup unit lists
    DBuf<uint32_t> a_ent, a_slice, a_lstart, a_lent, at_ent, at_slice, at_lstart, at_lent;
    DBuf<int32_t> a_lrow, at_lrow, cost1, cost2, adj_ptr, adj_ent;
    DBuf<double> a_coef, at_coef, b_dev;
    StructDev dev() const;
};

struct InstStatus {
    double mu, energy_primal, energy_dual, eps_p, eps_d;
    int64_t oiter, iiter;
    int converged, active, status, cg_last;
};

using Monitor = std::function<void(int inst, const InstStatus&)>;

// Inputs of Batch::write_fcidump: the density-fitted integrals of the full
// orbital space and the placement of the instance's orbitals in it.
struct FcidumpArgs {
    int nfull = 0;
    const int* full_irreps = nullptr;  // [nfull], nullptr = C1
    const double* h_full = nullptr;    // [nfull][nfull]
    int nq = 0;
    const double* qmo = nullptr;       // [nq][nfull (nfull + 1) / 2]
    int ncore = 0;
    const int* core = nullptr;         // [ncore] doubly occupied
    const int* active = nullptr;       // [nmo] full index of each orbital of the instance
    double enuc = 0.0;
    const char* int_path = nullptr;
    const char* rdm_path = nullptr;
};

// Resident density-fitted integrals of one instance (set_df_integrals): the
// full-space slices, h, the accumulated orbital rotation and the scratch of
// the intake and of the orbital step.
struct DfState {
    DfState() = default;
    DfState(const DfState&) = delete;
    DfState& operator=(const DfState&) = delete;
    ~DfState() { if (pin) (void)hipHostFree(pin); }
    int nfull = 0, nq = 0, n = 0, ncore = 0, no = 0;
    std::vector<int> core, active;
    double efzc = 0.0;
    int64_t m_rows = 0;
    // pinned read-back slots of the orbital step: |g|^2 and the model change,
    // ||K||_1, E_fzc, c.x (copied asynchronously, read after one sync)
    double* pin = nullptr;
    DBuf<double> qmo, h, bd, tmp, uacc, uscr, hscr;
    DBuf<int64_t> idx;
    DBuf<int> act, cor;
    DBuf<double> rows, dvec, gu, sq, e;
    // orbital step, set up by the first optimize_orbitals
    std::vector<int> pair_key;
    int nrot = -1;
    OoTables tabs{};
    DBuf<int2> pairs;
    DBuf<int> occ, pos;
    DBuf<int4> ptab;
    DBuf<int64_t> d1tab;
    DBuf<double> D, d, BO, BFO, Bdiag, gO, Gw, Z, Zw, F, Fw, Jw, Jd, Kd, g, H, kref, kcur, ksave, K, U, T1, red;
};

// Options and results of one orbital-optimisation call (the ORBOPT_* options
// of hilbert and OrbitalOptimizer::optimize_orbitals' outputs).
struct OrbOptArgs {
    const int* full_irreps = nullptr;  // [nfull], nullptr = C1
    double g_convergence = 1.0e-4, e_convergence = 1.0e-8;
    int maxiter = 10;
    int active_active = 0;
};
struct OrbOptResult {
    int iterations = 0, converged = 0, nrot = 0, evaluations = 0, rotations = 0;
    double gnorm = 0.0, dE = 0.0, energy = 0.0, efzc = 0.0;
};

class Batch {
  public:
    Batch(const std::vector<ProblemDesc>& descs, const SolverOptions& opt, int device);
    ~Batch();

    int count() const { return (int)inst_sid_.size(); }
    const Layout& layout(int i) const { return structs_[inst_sid_[i]]->L; }
    int64_t n_primal(int i) const { return structs_[inst_sid_[i]]->n_primal; }
    int64_t n_dual(int i) const { return structs_[inst_sid_[i]]->n_dual; }
    const SolverOptions& options() const { return opt_; }
    // New options; tighter E_CONVERGENCE or R_CONVERGENCE re-open the
    // instances that converged under the old ones, so that the next solve
    // continues them from their state to the new thresholds.
    void set_options(const SolverOptions& o);
    // Factor f of the CG stopping rule: relative residual below f x the
    // larger of the last primal and dual errors (capped at 1e-2, at least
    // CG_CONVERGENCE); default kCgTolFactor.
    void set_cg_tolerance_factor(double f);
    double cg_tolerance_factor() const { return cg_factor_; }

    // Hamiltonian in the user's orbital order: h[n*n], g[n^4] chemists' (pq|rs).
    void set_integrals(int i, const double* h, const double* g);
    void set_cost(int i, const double* c);        // primal-layout cost vector
    // Density-fitted Hamiltonian (qmo[nq][nfull (nfull + 1) / 2], h_full
    // [nfull][nfull]) with core[ncore] doubly occupied; active[nmo] are the
    // full-space indices of the caller's active orbitals.  Returns E_fzc.
    double set_df_integrals(int i, int nfull, const double* h_full, int nq, const double* qmo, int ncore,
                            const int* core, const int* active);
    // The integral and RDM files of hilbert's FCIDUMP option for instance i
    // (its current primal vector); record counts out (may be null).
    void write_fcidump(int i, const FcidumpArgs& a, int64_t* n_int, int64_t* n_rdm);
    // One orbital-optimisation call on the instance's resident DF integrals
    // with its current primal vector (orbopt_driver.cpp); rotates the integrals,
    // refreshes the cost vector and E_fzc.
    void optimize_orbitals(int i, const OrbOptArgs& a, OrbOptResult* r);
    // Rotation pairs, gradient and exact diagonal Hessian at the current
    // orbitals and primal vector, nothing rotated; returns the pair count
    // (the arrays are written when cap >= that count).
    int orbital_gradient(int i, const OrbOptArgs& a, int* pairs, double* g, double* H, int cap);
    // Accumulated rotation U (new orbitals = orbitals given to
    // set_df_integrals times U) and the current h_full and packed qmo; any
    // pointer may be null.
    void get_df_orbitals(int i, double* U, double* h_full, double* qmo);
    void set_guess(int i, const double* x);       // nullptr: hilbert's random guess
    void set_guess_hf(int i);                     // hilbert's TPDM_GUESS HF
    void reset_state(int i);                      // y = z = 0, mu = mu0, counters cleared

    // Run up to maxiter outer iterations for every unconverged instance.
    // Returns the number of converged instances.
    int solve(int maxiter, const Monitor& mon = nullptr);

    InstStatus status(int i);
    void get_primal(int i, double* x);
    void get_dual(int i, double* y);
    void get_slack(int i, double* z);
    void get_cost(int i, double* c);
    void get_b(int i, double* b) const;
    void set_dual(int i, const double* y);
    void set_slack(int i, const double* z);
    double get_mu(int i);
    void set_mu(int i, double mu);

    // Maps on host vectors (copy in, apply on the device, copy out).
    void apply_A(int i, const double* u, double* out);
    // out_dev[n_dual] = A in_dev[n_primal] on the batch stream (device
    // pointers; out_dev is cleared first); no allocation, no synchronisation.
    void apply_A_device(int i, const double* in_dev, double* out_dev);
    int device() const { return device_; }
    void apply_AT(int i, const double* y, double* out);
    // dE/dsrc for src = [h; g] (user order) from a primal vector (default x).
    void cost_adjoint(int i, const double* xbar, double* dh, double* dg);

    hipStream_t stream() const { return stream_; }

    // HIP-event timing of every kernel launch of solve(), per kernel kind:
    // 0 pre, 1 cg_at, 2 cg_a, 3 post, 4..10 projection classes: Jacobi d <= 16,
    // tridiagonal d <= 64 / 128 / 192 / 224 / 256, Jacobi d > 256 (matrix in
    // global memory); 11: the device part of one set_df_integrals (its
    // contraction kernels and the cost packing); 12: the pair Gram matrix of
    // write_fcidump; 13: the device work of optimize_orbitals outside the
    // intake (densities, Fock and Hessian products, exponential, transform);
    // 14: the warm-started refinement of the 65..256 blocks (all its launches
    // of one outer iteration); 15: the eigenvector formation of the blocks
    // that fell back to the tridiagonal path.  Classes 2..5 then time only
    // their fallback launches.
    static constexpr int kProjClasses = Projector::kClasses;
    static constexpr int kKindDf = 4 + kProjClasses;
    static constexpr int kKindGram = kKindDf + 1;
    static constexpr int kKindOrb = kKindGram + 1;
    static constexpr int kKindRefine = kKindOrb + 1;
    static constexpr int kKindSeed = kKindRefine + 1;
    static constexpr int kKinds = kKindSeed + 1;
    // The projection classes run concurrently on stream_ (lane 0) and three
    // side streams: with the runtime's default of four hardware queues per
    // process, every stream of a batch keeps a queue of its own, so classes
    // on different lanes do not serialise behind each other.
    static constexpr int kSideStreams = 3;
    // Launch configuration, none of which changes a result bit (the same
    // kernels on the same data, ordered by the same events; the extra CG
    // steps of a larger first chunk are skipped by the instances that have
    // stopped): what the autotuner (hilbert_amd/autotune.py) chooses.
    struct Tuning {
        // lane (0 = stream_, 1..3 side streams) of each size class (Jacobi
        // <= 16, tri 64 / 128 / 192 / 224 / 256, Jacobi > 256) when it is not
        // refined, chosen to balance the per-iteration class times of the
        // (16e,16o) workloads
        int class_lane[kProjClasses] = {0, 0, 2, 3, 1, 1, 2};
        // lanes of the tridiagonal fallback of the refined classes (2..5)
        int fallback_lane[kProjClasses] = {0, 0, 2, 3, 3, 2, 2};
        int refine_lane = 1;  // the warm-started refinement and the eigenvector formation (1..3)
        int cg_margin = 2;    // first CG chunk: the largest previous CG step count + cg_margin
        // gather workgroups dealt to the XCDs in contiguous ranges (1) or in
        // launch order (0, default: measured faster on the (16e,16o) batches)
        int xcd_remap = 0;
        int graphs = 0;       // kGraphOff / kGraphCg / kGraphAll (set_graphs)
        // Pipelined solve: while at least this many instances run, they form
        // two groups that take turns on stream_ (pre, CG, post), the
        // projection of one group running on the side streams beside the CG
        // of the other; below it, and with 0 or with graphs, the whole batch
        // iterates in lock-step.
        int pipeline = 0;
    };
    void set_tuning(const Tuning& t);
    const Tuning& tuning() const { return tuning_; }
    // Refined projection of the 65..256 blocks on (default) or off (every
    // block takes the tridiagonal eigensolver).
    void set_refine(bool on);
    bool refine() const { return proj_.refine_on(); }
    // [block projections by the refinement, by the fallback, rejections for
    // divergence, correction size, coupled clusters]
    void projection_stats(int64_t* out) const { proj_.refine_stats(out); }
    void set_refine_params(const RefParams& p) {
        proj_.set_refine_params(p);  // the parameters are kernel arguments: re-capture
        drop_graphs();
    }
    void set_profiling(bool on) { profile_ = on; }
    // hipGraph replay of the fixed launch sequences of an outer iteration:
    // kGraphCg, the first CG chunk (its k_cg_at / k_cg_a / k_cg_reduce
    // launches); kGraphAll, also k_post with the whole projection (every size
    // class on its lane, the refinement and its fallback passes, joined into
    // stream_), the eigenvector formation then running as a branch of the
    // next CG graph.  Each graph is captured once per launch plan (the set of
    // running instances) and step count, then replayed with one
    // hipGraphLaunch; with the kernel timers on, the launches go one by one.
    // kGraphOff (default): the same launches one by one.  Both graph modes
    // give the eager bits, also over solves split into short calls with the
    // mode switched between them.
    static constexpr int kGraphOff = 0, kGraphCg = 1, kGraphAll = 2;
    void set_graphs(int mode);
    int graphs() const { return tuning_.graphs; }
    // ms/counts: all launches; full_ms/full_counts: launches made while every
    // instance of the batch was active (and, for CG kernels, still iterating)
    // at the last host poll, whose work is therefore the whole batch's.
    void kernel_times(double* ms, int64_t* counts, double* full_ms, int64_t* full_counts);
    void reset_kernel_times();
    void launch_stats(int64_t* st) const {
        const int64_t nblk = proj_.nblocks();
        st[0] = launches_;
        st[1] = nwgA_;
        st[2] = nwgAT_;
        st[3] = nblk;
    }
    double last_solve_seconds() const { return last_seconds_; }
    int64_t kernel_launches() const { return launches_; }

  private:
    void build_units();
    void build_projection();
    void sync_state_to_host();
    void sync_state_to_device();
    void refresh_norms(int i);
    // g < 0: the lock-step projection, joined into stream_; g = 0, 1: the
    // projection of group g on the side streams only, its completion recorded
    // in ev_pdone_[g] (the next pre of the group waits for it).
    void launch_projection(const BatchDev& B, bool seed, int g = -1);
    void wait_projection(int g);    // stream_ waits for group g's last projection
    void set_groups(const std::vector<int>& grp);  // upload the groups (projection idle)
    void launch_seed(const BatchDev& B);
    void join_seed();
    DfState& df_state(int i);
    double run_df_intake(int i, bool norms = true);  // resident DF integrals -> c, v1, E_fzc (and ||c||)
    void launch_df_intake_only(int i);    // the same launches, no read-back
    double instance_energy(int i);        // c.x + E_fzc
    void oo_setup(int i, const OrbOptArgs& a);
    void oo_gradient(int i);
    double oo_apply(int i, const double* kappa, double sign);  // rotate, re-run the intake, energy
    std::vector<std::unique_ptr<DfState>> df_;
    BatchDev dev() const;

    SolverOptions opt_;
    double cg_factor_ = kCgTolFactor;
    int device_ = 0;
    hipStream_t stream_ = nullptr;
    std::vector<std::shared_ptr<Structure>> structs_;
    std::vector<int> inst_sid_;
    std::vector<InstDev> inst_;
    std::vector<InstState> state_;      // host mirror
    int64_t np_total_ = 0, nd_total_ = 0, ns_total_ = 0;

    DBuf<StructDev> d_structs_;
    DBuf<InstDev> d_inst_;
    DBuf<InstState> d_state_;
    DBuf<double> x_, z_, c_, v1_, u_, aty_, q_, av1_, ax_, aaty_;
    DBuf<CgRow> cg_;
    DBuf<double> src_, partials_, scratch_;
    DBuf<int4> unitsA_, unitsAT_;        // work-unit codes (per-instance maps)
    DBuf<UnitRec> recA_, recAT_;          // the same units, resolved (batched gathers)
    DBuf<int> wgA_inst_, wgAT_inst_, wgA_begin_, wgAT_begin_;
    int nwgA_ = 0, nwgAT_ = 0;
    std::vector<int> hwgA_begin_, hwgAT_begin_;

    Projector proj_;
    // side streams of the projection, joined back into stream_ by events
    hipStream_t pstream_[kSideStreams] = {};
    // the fork / join events of the projection: one set for eager launches and
    // one used only inside graph captures (an event is never recorded both in
    // a capture and outside one)
    struct ProjEvents {
        hipEvent_t post = nullptr, state = nullptr, ref = nullptr, proj[kSideStreams] = {};
    };
    ProjEvents ev_eager_, ev_capture_;
    // pipelined solve: group of every instance (host and device copies), the
    // completion of each group's projection per side stream
    std::vector<int> grp_;
    DBuf<int> d_grp_;
    hipEvent_t ev_pdone_[2][kSideStreams] = {};
    bool pdone_used_[2][kSideStreams] = {};
    hipEvent_t ev_seed_ = nullptr;
    bool seed_pending_ = false;

    struct KTimer {
        std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;  // every event pair of this kind
        std::vector<std::pair<hipEvent_t, hipEvent_t>> free;  // pairs not in flight
        std::pair<hipEvent_t, hipEvent_t> open{};            // between tstart and tstop
        struct Pending {
            hipEvent_t a, b;
            char full;
        };
        std::vector<Pending> pending;  // recorded, not yet read (may still be in flight)
        double ms = 0.0, full_ms = 0.0;
        int64_t count = 0, full_count = 0;
    };
    bool all_running_ = false, all_iterating_ = false;
    KTimer timers_[kKinds];
    bool profile_ = false;
    void tstart(int k, hipStream_t s = nullptr);
    void tstop(int k, hipStream_t s = nullptr);
    void tflush();
    char timer_full(int k) const { return (k == 1 || k == 2) ? all_iterating_ : all_running_; }

    // Captured launch sequences (set_graphs).
    struct GraphEntry {
        std::string key;                 // plan, step count and refinement state
        hipGraphExec_t exec = nullptr;
        int64_t launches = 0;            // kernel launches per replay
    };
    template <class Body>
    void run_graph(GraphEntry& e, const std::string& key, Body&& body);
    GraphEntry& cg_graph(int steps, bool seed);
    std::string graph_key(const std::string& tag) const;
    static constexpr size_t kCgGraphs = 8;  // step counts kept per plan
    void drop_graph(GraphEntry& e);
    void drop_graphs();
    bool seed_owed_ = false;  // a projection graph ran; its eigenvector formation is not launched yet
    hipEvent_t ev_gfork_ = nullptr, ev_gjoin_ = nullptr;  // the formation's branch in a CG graph
    hipEvent_t ev_joined_ = nullptr;  // stream_ after the projection's join, for the formation
    Tuning tuning_;
    GraphEntry* capture_ = nullptr;      // the entry being captured
    std::vector<GraphEntry> g_cg_;       // first CG chunk, one entry per step count
    GraphEntry g_post_;                  // k_post and the projection

    // Launch plans of the gather kernels (UnitsDev::plan): the instances whose
    // workgroups are launched.  Host copy in pinned memory, device copy
    // [A plan | A^T plan], each 2 count + 1 ints.
    void set_plan(const std::vector<char>& sel);
    UnitsDev units_a() const;
    UnitsDev units_at() const;
    std::vector<char> plan_sel_;
    int* plan_pinned_ = nullptr;
    DBuf<int> plan_dev_;
    int planA_m_ = 0, planAT_m_ = 0, gridA_ = 0, gridAT_ = 0;
    bool plan_all_ = true, plan_copy_pending_ = false;

    InstState* pinned_ = nullptr;
    bool constructed_ = false;
    double last_seconds_ = 0.0;
    int64_t launches_ = 0;
};

}  // namespace v2h
