// solver.cpp -- host driver of the batched boundary-point solve.
//
// Per outer iteration (all instances of the batch at once, converged ones
// skipped inside the kernels):
//   k_pre      A [c - z, x, A^T y]; finish the previous iteration
//   CG steps   (k_cg_at, k_cg_a) until every instance's CG has stopped
//   k_post     A^T y, U = mu x + A^T y - c
//   k_project  one workgroup per PSD block (four size classes)
// The host only reads the small per-instance state array between CG chunks.
#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "df.hpp"
#include "kernels.hpp"

namespace v2h {

namespace {
constexpr int kLongThreshold = 32;

std::vector<int4> unit_list(const SellHost& S) {
    std::vector<int> u;
    for (int s = 0; s < S.nslices; s++) u.push_back(s);
    for (int l = 0; l < S.nlong(); l++) u.push_back(kLongFlag | l);
    while (u.size() % 4) u.push_back(-1);
    std::vector<int4> out(u.size() / 4);
    for (size_t g = 0; g < out.size(); g++) out[g] = make_int4(u[4 * g], u[4 * g + 1], u[4 * g + 2], u[4 * g + 3]);
    return out;
}

}  // namespace

// ------------------------------------------------------------------ Structure

Structure::Structure(const ProblemDesc& d) : L(d) {
    Csr A = build_constraints(L);
    n_primal = L.n_primal;
    n_dual = A.nrows;
    if (n_dual >= ((int64_t)1 << 24)) throw std::invalid_argument("problem too large: more than 2^24 constraints");
    b = A.rhs;
    Csr AT = A.transpose();
    hA = build_sell(A, kLongThreshold);
    hAT = build_sell(AT, kLongThreshold);
    unitsA = unit_list(hA);
    unitsAT = unit_list(hAT);

    CostMap cm = build_cost_map(L);
    const int64_t n = L.n;
    nsrc = n * n + n * n * n * n;
    std::vector<int32_t> ptr(nsrc + 1, 0), ent;
    for (int64_t e = 0; e < n_primal; e++) {
        if (cm.i1[e] >= 0) ptr[cm.i1[e] + 1]++;
        if (cm.i2[e] >= 0) ptr[cm.i2[e] + 1]++;
    }
    for (int64_t t = 0; t < nsrc; t++) ptr[t + 1] += ptr[t];
    ent.resize(ptr[nsrc]);
    std::vector<int32_t> cur(ptr.begin(), ptr.end() - 1);
    for (int64_t e = 0; e < n_primal; e++) {
        if (cm.i1[e] >= 0) ent[cur[cm.i1[e]]++] = (int32_t)e;
        if (cm.i2[e] >= 0) ent[cur[cm.i2[e]]++] = (int32_t)((uint32_t)e | 0x80000000u);
    }

    a_ent.upload(hA.ent); a_slice.upload(hA.slice_start); a_lstart.upload(hA.long_start);
    a_lent.upload(hA.long_ent); a_lrow.upload(hA.long_row); a_coef.upload(hA.coef);
    at_ent.upload(hAT.ent); at_slice.upload(hAT.slice_start); at_lstart.upload(hAT.long_start);
    at_lent.upload(hAT.long_ent); at_lrow.upload(hAT.long_row); at_coef.upload(hAT.coef);
    b_dev.upload(b);
    cost1.upload(cm.i1);
    cost2.upload(cm.i2);
    adj_ptr.upload(ptr);
    adj_ent.upload(ent);
}

StructDev Structure::dev() const {
    StructDev s{};
    s.A = SellDev{a_ent.p, a_slice.p, a_lstart.p, a_lent.p, a_lrow.p, a_coef.p, hA.nrows, hA.nslices, hA.nlong()};
    s.AT = SellDev{at_ent.p, at_slice.p, at_lstart.p, at_lent.p, at_lrow.p, at_coef.p, hAT.nrows, hAT.nslices,
                   hAT.nlong()};
    s.b = b_dev.p;
    s.cost1 = cost1.p;
    s.cost2 = cost2.p;
    s.adj_ptr = adj_ptr.p;
    s.adj_ent = adj_ent.p;
    s.n_primal = n_primal;
    s.n_dual = n_dual;
    s.nsrc = nsrc;
    return s;
}

// ---------------------------------------------------------------------- Batch

Batch::Batch(const std::vector<ProblemDesc>& descs, const SolverOptions& opt, int device) : opt_(opt), device_(device) {
    if (descs.empty()) throw std::invalid_argument("empty batch");
    V2H_HIP(hipSetDevice(device_));
    V2H_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (int k = 0; k < kSideStreams; k++) V2H_HIP(hipStreamCreateWithFlags(&pstream_[k], hipStreamNonBlocking));
    for (ProjEvents* E : {&ev_eager_, &ev_capture_}) {
        for (int k = 0; k < kSideStreams; k++) V2H_HIP(hipEventCreateWithFlags(&E->proj[k], hipEventDisableTiming));
        V2H_HIP(hipEventCreateWithFlags(&E->post, hipEventDisableTiming));
        V2H_HIP(hipEventCreateWithFlags(&E->ref, hipEventDisableTiming));
        V2H_HIP(hipEventCreateWithFlags(&E->state, hipEventDisableTiming));
    }
    V2H_HIP(hipEventCreateWithFlags(&ev_seed_, hipEventDisableTiming));
    V2H_HIP(hipEventCreateWithFlags(&ev_gfork_, hipEventDisableTiming));
    V2H_HIP(hipEventCreateWithFlags(&ev_joined_, hipEventDisableTiming));
    V2H_HIP(hipEventCreateWithFlags(&ev_gjoin_, hipEventDisableTiming));
    for (auto& row : ev_pdone_)
        for (auto& e : row) V2H_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));

    std::map<std::string, int> sid_of;
    for (const auto& d : descs) {
        std::string k = d.key();
        auto it = sid_of.find(k);
        int sid;
        if (it == sid_of.end()) {
            sid = (int)structs_.size();
            structs_.push_back(std::make_shared<Structure>(d));
            sid_of[k] = sid;
        } else {
            sid = it->second;
        }
        // identical shape but a different orbital order: own structure (the
        // cost map follows the user's order)
        if (structs_[sid]->L.to_user != Layout(d).to_user) {
            sid = (int)structs_.size();
            structs_.push_back(std::make_shared<Structure>(d));
        }
        inst_sid_.push_back(sid);
    }
    const int N = count();
    inst_.resize(N);
    for (int i = 0; i < N; i++) {
        const Structure& S = *structs_[inst_sid_[i]];
        inst_[i].poff = np_total_;
        inst_[i].doff = nd_total_;
        inst_[i].soff = ns_total_;
        inst_[i].sid = inst_sid_[i];
        np_total_ += S.n_primal;
        nd_total_ += S.n_dual;
        ns_total_ += S.nsrc;
    }
    build_units();
    build_projection();

    for (auto* b : {&x_, &z_, &c_, &v1_, &u_, &aty_, &q_}) b->alloc(np_total_);
    for (auto* b : {&av1_, &ax_, &aaty_}) b->alloc(nd_total_);
    cg_.alloc(nd_total_);
    src_.alloc(ns_total_);
    partials_.alloc(4 * (size_t)std::max(nwgA_, nwgAT_) + 4);
    int64_t smax = 0;
    for (const auto& s : structs_) smax = std::max({smax, s->n_primal, s->n_dual, s->nsrc});
    scratch_.alloc(std::max<int64_t>(2 * smax, 257));  // launch_dot1 needs 1 + 256
    for (auto* b : {&x_, &z_, &c_, &v1_, &u_, &aty_, &q_}) launch_fill(b->p, np_total_, 0.0, stream_);
    for (auto* b : {&av1_, &ax_, &aaty_}) launch_fill(b->p, nd_total_, 0.0, stream_);
    launch_fill(reinterpret_cast<double*>(cg_.p), 4 * nd_total_, 0.0, stream_);
    launch_fill(src_.p, ns_total_, 0.0, stream_);

    std::vector<StructDev> sd;
    for (const auto& s : structs_) sd.push_back(s->dev());
    d_structs_.upload(sd);
    d_inst_.upload(inst_);
    grp_.assign(N, 0);
    d_grp_.upload(grp_);
    state_.assign(N, InstState{});
    for (int i = 0; i < N; i++) {
        const auto& b = structs_[inst_sid_[i]]->b;
        double s = 0.0;
        for (double v : b) s += v * v;
        state_[i].bnorm = s > 0 ? std::sqrt(s) : 1.0;
        state_[i].cnorm = 1.0;
    }
    d_state_.alloc(N);
    V2H_HIP(hipHostMalloc((void**)&pinned_, sizeof(InstState) * N, hipHostMallocDefault));
    V2H_HIP(hipHostMalloc((void**)&plan_pinned_, sizeof(int) * 2 * (2 * (size_t)N + 1), hipHostMallocDefault));
    plan_dev_.alloc(2 * (2 * (size_t)N + 1));
    plan_sel_.assign(N, 1);
    gridA_ = nwgA_;
    gridAT_ = nwgAT_;
    for (int i = 0; i < N; i++) {
        set_guess(i, nullptr);
        reset_state(i);
    }
    sync_state_to_device();
    V2H_HIP(hipStreamSynchronize(stream_));
    constructed_ = true;
}

Batch::~Batch() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph(g_post_);
    for (auto& e : g_cg_) drop_graph(e);
    for (auto& t : timers_)
        for (auto& e : t.pool) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    if (pinned_) (void)hipHostFree(pinned_);
    if (plan_pinned_) (void)hipHostFree(plan_pinned_);
    for (int k = 0; k < kSideStreams; k++)
        if (pstream_[k]) (void)hipStreamDestroy(pstream_[k]);
    for (ProjEvents* E : {&ev_eager_, &ev_capture_}) {
        for (int k = 0; k < kSideStreams; k++)
            if (E->proj[k]) (void)hipEventDestroy(E->proj[k]);
        for (hipEvent_t e : {E->post, E->ref, E->state})
            if (e) (void)hipEventDestroy(e);
    }
    if (ev_seed_) (void)hipEventDestroy(ev_seed_);
    if (ev_gfork_) (void)hipEventDestroy(ev_gfork_);
    if (ev_joined_) (void)hipEventDestroy(ev_joined_);
    if (ev_gjoin_) (void)hipEventDestroy(ev_gjoin_);
    for (auto& row : ev_pdone_)
        for (auto& e : row)
            if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

namespace {
// The UnitRec of one work-unit code (slice s, kLongFlag | long row l, or -1).
UnitRec unit_rec(const SellHost& h, const uint32_t* ent, const uint32_t* lent, const double* coef, int code) {
    UnitRec r{ent, coef, 0u, -1};
    if (code < 0) return r;
    if (code & kLongFlag) {
        const int l = code & ~kLongFlag;
        r.ent = lent + h.long_start[l];
        r.meta = (h.long_start[l + 1] - h.long_start[l]) | kRecLong;
        r.row = h.long_row[l];
        return r;
    }
    const uint32_t w = h.slice_start[code + 1] - h.slice_start[code];
    const int64_t valid = std::min<int64_t>(64, h.nrows - (int64_t)code * 64);
    if (w > 0xFFFFu || valid <= 0) throw std::logic_error("SELL slice out of range");
    r.ent = ent + (size_t)h.slice_start[code] * 64;
    r.meta = w | ((uint32_t)valid << 16);
    r.row = code * 64;
    return r;
}
}  // namespace

void Batch::build_units() {
    std::vector<int4> uA, uAT;
    std::vector<int> iA, iAT;
    std::vector<UnitRec> rA, rAT;
    hwgA_begin_.assign(1, 0);
    hwgAT_begin_.assign(1, 0);
    for (int i = 0; i < count(); i++) {
        const Structure& S = *structs_[inst_sid_[i]];
        uA.insert(uA.end(), S.unitsA.begin(), S.unitsA.end());
        uAT.insert(uAT.end(), S.unitsAT.begin(), S.unitsAT.end());
        for (const int4& g : S.unitsA)
            for (int code : {g.x, g.y, g.z, g.w}) rA.push_back(unit_rec(S.hA, S.a_ent.p, S.a_lent.p, S.a_coef.p, code));
        for (const int4& g : S.unitsAT)
            for (int code : {g.x, g.y, g.z, g.w})
                rAT.push_back(unit_rec(S.hAT, S.at_ent.p, S.at_lent.p, S.at_coef.p, code));
        iA.insert(iA.end(), S.unitsA.size(), i);
        iAT.insert(iAT.end(), S.unitsAT.size(), i);
        hwgA_begin_.push_back((int)uA.size());
        hwgAT_begin_.push_back((int)uAT.size());
    }
    nwgA_ = (int)uA.size();
    nwgAT_ = (int)uAT.size();
    unitsA_.upload(uA);
    unitsAT_.upload(uAT);
    recA_.upload(rA);
    recAT_.upload(rAT);
    wgA_inst_.upload(iA);
    wgAT_inst_.upload(iAT);
    wgA_begin_.upload(hwgA_begin_);
    wgAT_begin_.upload(hwgAT_begin_);
}

void Batch::build_projection() {
    std::vector<std::vector<int>> dims(count());
    for (int i = 0; i < count(); i++) {
        int64_t off = 0;
        for (const Block& b : structs_[inst_sid_[i]]->L.blocks) {
            if (b.off != off) throw std::logic_error("PSD blocks are not contiguous in the primal vector");
            off += (int64_t)b.dim * b.dim;
            dims[i].push_back(b.dim);
        }
    }
    proj_.build(dims, inst_, stream_);
}

BatchDev Batch::dev() const {
    BatchDev B{};
    B.structs = d_structs_.p;
    B.inst = d_inst_.p;
    B.st = d_state_.p;
    B.x = x_.p; B.z = z_.p; B.c = c_.p; B.v1 = v1_.p; B.u = u_.p; B.aty = aty_.p; B.q = q_.p;
    B.cg = cg_.p; B.av1 = av1_.p; B.ax = ax_.p; B.aaty = aaty_.p;
    B.partials = partials_.p;
    B.proj_part = proj_.proj_part();
    B.opt = opt_;
    B.opt.cg_tol_factor = cg_factor_;
    B.grp = d_grp_.p;
    B.pgroup = -1;
    return B;
}

UnitsDev Batch::units_a() const {
    const int N = count();
    return UnitsDev{recA_.p, wgA_inst_.p, wgA_begin_.p, nwgA_, N, plan_all_ ? nullptr : plan_dev_.p, planA_m_,
                    gridA_, tuning_.xcd_remap};
}

UnitsDev Batch::units_at() const {
    const int N = count();
    return UnitsDev{recAT_.p, wgAT_inst_.p, wgAT_begin_.p, nwgAT_, N,
                    plan_all_ ? nullptr : plan_dev_.p + 2 * N + 1, planAT_m_, gridAT_, tuning_.xcd_remap};
}

// Launch only the workgroups of the selected instances.  The selection must
// contain every instance that the kernels would not skip (the host's view of
// the state is never behind the device's in that sense: instances only stop).
void Batch::set_plan(const std::vector<char>& sel) {
    if (sel == plan_sel_) return;
    const int N = count();
    if (plan_copy_pending_) V2H_HIP(hipStreamSynchronize(stream_));  // the pinned copy is still being read
    plan_sel_ = sel;
    plan_all_ = std::all_of(sel.begin(), sel.end(), [](char c) { return c != 0; });
    if (plan_all_) {
        gridA_ = nwgA_;
        gridAT_ = nwgAT_;
        return;
    }
    auto build = [&](const std::vector<int>& wgb, int* out, int& m) {
        // runs of consecutive selected instances
        std::vector<int> pre{0}, first;
        for (int i = 0; i < N;) {
            if (!sel[i] || wgb[i + 1] == wgb[i]) {
                i++;
                continue;
            }
            int j = i;
            while (j < N && sel[j]) j++;
            first.push_back(wgb[i]);
            pre.push_back(pre.back() + wgb[j] - wgb[i]);
            i = j;
        }
        m = (int)first.size();
        std::copy(pre.begin(), pre.end(), out);
        std::copy(first.begin(), first.end(), out + m + 1);
        return pre.back();
    };
    gridA_ = build(hwgA_begin_, plan_pinned_, planA_m_);
    gridAT_ = build(hwgAT_begin_, plan_pinned_ + 2 * N + 1, planAT_m_);
    V2H_HIP(hipMemcpyAsync(plan_dev_.p, plan_pinned_, sizeof(int) * 2 * (2 * (size_t)N + 1), hipMemcpyHostToDevice,
                           stream_));
    plan_copy_pending_ = true;
}

void Batch::sync_state_to_host() {
    V2H_HIP(hipMemcpyAsync(pinned_, d_state_.p, sizeof(InstState) * count(), hipMemcpyDeviceToHost, stream_));
    V2H_HIP(hipStreamSynchronize(stream_));
    plan_copy_pending_ = false;
    std::memcpy(state_.data(), pinned_, sizeof(InstState) * count());
    tflush();
}

void Batch::tstart(int k, hipStream_t s) {
    if (!profile_) return;
    if (!s) s = stream_;
    KTimer& t = timers_[k];
    if (t.free.empty()) {
        hipEvent_t a, b;
        V2H_HIP(hipEventCreate(&a));
        V2H_HIP(hipEventCreate(&b));
        t.pool.push_back({a, b});
        t.free.push_back({a, b});
    }
    t.open = t.free.back();
    t.free.pop_back();
    V2H_HIP(hipEventRecord(t.open.first, s));
}

void Batch::tstop(int k, hipStream_t s) {
    if (!profile_) return;
    if (!s) s = stream_;
    KTimer& t = timers_[k];
    V2H_HIP(hipEventRecord(t.open.second, s));
    t.pending.push_back({t.open.first, t.open.second, timer_full(k)});
}

// Reads the timers whose launches have finished; the others (on side
// streams that stream_ has not joined yet: the eigenvector formation, the
// projection of a pipelined group) stay pending until a later call.
void Batch::tflush() {
    for (auto& t : timers_) {
        size_t keep = 0;
        for (size_t j = 0; j < t.pending.size(); j++) {
            const auto r = t.pending[j];
            const hipError_t q = hipEventQuery(r.b);
            if (q == hipErrorNotReady) {
                t.pending[keep++] = r;
                continue;
            }
            V2H_HIP(q);
            float ms = 0.f;
            V2H_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            t.ms += ms;
            t.count++;
            if (r.full) {
                t.full_ms += ms;
                t.full_count++;
            }
            t.free.push_back({r.a, r.b});
        }
        t.pending.resize(keep);
    }
}

// Replay the launches of `body` on stream_ as one graph: captured on first use
// and whenever `key` changes, then launched with hipGraphLaunch.  `body` may
// fork onto the side streams by events but must join them back into stream_,
// and may not synchronise, copy or allocate.  With the kernel timers on,
// `body` runs eagerly instead: its timers are events on the streams.
template <class Body>
void Batch::run_graph(GraphEntry& e, const std::string& key, Body&& body) {
    if (profile_) {  // with the kernel timers on, the launches go one by one
        body();
        return;
    }
    if (!e.exec || e.key != key) {
        drop_graph(e);
        const int64_t l0 = launches_;
        V2H_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        capture_ = &e;
        hipGraph_t g = nullptr;
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            capture_ = nullptr;
            drop_graph(e);
            throw;
        }
        capture_ = nullptr;
        V2H_HIP(hipStreamEndCapture(stream_, &g));
        const hipError_t rc = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (rc != hipSuccess) {
            e.exec = nullptr;
            drop_graph(e);
            V2H_HIP(rc);
        }
        e.key = key;
        e.launches = launches_ - l0;
        launches_ = l0;
    }
    V2H_HIP(hipGraphLaunch(e.exec, stream_));
    launches_ += e.launches;
}

void Batch::drop_graph(GraphEntry& e) {
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    e.exec = nullptr;
    e.key.clear();
    e.launches = 0;
}

void Batch::drop_graphs() {
    drop_graph(g_post_);
    for (auto& e : g_cg_) drop_graph(e);
    g_cg_.clear();
}

void Batch::set_tuning(const Tuning& t) {
    for (int k = 0; k < kProjClasses; k++)
        if (t.class_lane[k] < 0 || t.class_lane[k] > kSideStreams || t.fallback_lane[k] < 0 ||
            t.fallback_lane[k] > kSideStreams)
            throw std::invalid_argument("tuning: lanes are 0.." + std::to_string(kSideStreams));
    if (t.refine_lane < 1 || t.refine_lane > kSideStreams)
        throw std::invalid_argument("tuning: the refinement lane is a side stream (1.." +
                                    std::to_string(kSideStreams) + ")");
    if (t.cg_margin < 0 || t.cg_margin > 64) throw std::invalid_argument("tuning: cg_margin is 0..64");
    if (t.xcd_remap != 0 && t.xcd_remap != 1) throw std::invalid_argument("tuning: xcd_remap is 0 or 1");
    if (t.pipeline < 0) throw std::invalid_argument("tuning: pipeline is 0 (lock-step) or an instance count");
    if (t.graphs < kGraphOff || t.graphs > kGraphAll)
        throw std::invalid_argument("tuning: graphs is 0 (off), 1 (CG chunk) or 2 (CG chunk, post and projection)");
    V2H_HIP(hipSetDevice(device_));
    join_seed();  // the eigenvector formation in flight ran on the old refinement lane
    V2H_HIP(hipStreamSynchronize(stream_));
    tuning_ = t;
    drop_graphs();  // lanes and the remap flag are baked into the captures
}

void Batch::set_graphs(int mode) {
    Tuning t = tuning_;
    t.graphs = mode;
    set_tuning(t);
}

void Batch::kernel_times(double* ms, int64_t* counts, double* full_ms, int64_t* full_counts) {
    V2H_HIP(hipStreamSynchronize(stream_));
    for (hipStream_t p : pstream_) V2H_HIP(hipStreamSynchronize(p));  // timers of unjoined side work
    tflush();
    for (int k = 0; k < kKinds; k++) {
        if (ms) ms[k] = timers_[k].ms;
        if (counts) counts[k] = timers_[k].count;
        if (full_ms) full_ms[k] = timers_[k].full_ms;
        if (full_counts) full_counts[k] = timers_[k].full_count;
    }
}

void Batch::reset_kernel_times() {
    V2H_HIP(hipStreamSynchronize(stream_));
    for (hipStream_t p : pstream_) V2H_HIP(hipStreamSynchronize(p));  // timers of unjoined side work
    tflush();
    for (auto& t : timers_) {
        t.ms = t.full_ms = 0.0;
        t.count = t.full_count = 0;
    }
}

void Batch::sync_state_to_device() {
    std::memcpy(pinned_, state_.data(), sizeof(InstState) * count());
    V2H_HIP(hipMemcpyAsync(d_state_.p, pinned_, sizeof(InstState) * count(), hipMemcpyHostToDevice, stream_));
    V2H_HIP(hipStreamSynchronize(stream_));
}

void Batch::refresh_norms(int i) {
    const InstDev& I = inst_[i];
    const Structure& S = *structs_[inst_sid_[i]];
    launch_dot1(c_.p + I.poff, c_.p + I.poff, S.n_primal, scratch_.p, stream_);
    double cc = 0.0;
    V2H_HIP(hipMemcpyAsync(&cc, scratch_.p, sizeof(double), hipMemcpyDeviceToHost, stream_));
    V2H_HIP(hipStreamSynchronize(stream_));
    sync_state_to_host();
    state_[i].cnorm = cc > 0 ? std::sqrt(cc) : 1.0;
    // A non-finite cost (NaN or infinite integrals) fails this instance
    // before any kernel iterates it; the rest of the batch is unaffected.
    if (!std::isfinite(cc)) {
        state_[i].status |= kStatusNonFinite;
        state_[i].active = 0;
    }
    sync_state_to_device();
}

void Batch::set_integrals(int i, const double* h, const double* g) {
    const InstDev& I = inst_[i];
    const Structure& S = *structs_[inst_sid_[i]];
    const int64_t n = S.L.n;
    V2H_HIP(hipMemcpyAsync(src_.p + I.soff, h, sizeof(double) * n * n, hipMemcpyHostToDevice, stream_));
    V2H_HIP(hipMemcpyAsync(src_.p + I.soff + n * n, g, sizeof(double) * n * n * n * n, hipMemcpyHostToDevice, stream_));
    launch_pack_cost(S.cost1.p, S.cost2.p, src_.p + I.soff, c_.p + I.poff, S.n_primal, stream_);
    // v1 = c - z
    launch_copy(v1_.p + I.poff, c_.p + I.poff, S.n_primal, stream_);
    refresh_norms(i);
}

// The density-fitted intake (df.hip): h' and (pq|rs) of the active space are
// formed on the device in the source array, then packed like host integrals.
double Batch::set_df_integrals(int i, int nfull, const double* h_full, int nq, const double* qmo, int ncore,
                               const int* core, const int* active) {
    const int n = layout(i).n;
    if (nfull < 1 || nq < 1 || ncore < 0 || ncore + n > nfull)
        throw std::invalid_argument("set_df_integrals: need nfull >= ncore + nmo, nq >= 1");
    std::vector<char> seen(nfull, 0);
    auto take = [&](int k) {
        if (k < 0 || k >= nfull) throw std::invalid_argument("set_df_integrals: orbital index out of range");
        if (seen[k]) throw std::invalid_argument("set_df_integrals: orbital listed twice (core and active are disjoint)");
        seen[k] = 1;
    };
    for (int c = 0; c < ncore; c++) take(core[c]);
    for (int p = 0; p < n; p++) take(active[p]);
    auto tri = [](int64_t a, int64_t b) { return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; };
    std::vector<int64_t> idx;
    const int64_t nu = (int64_t)n * (n + 1) / 2;
    idx.reserve(nu + ncore + (size_t)n * ncore + (size_t)ncore * ncore);
    for (int p = 0; p < n; p++)
        for (int q = 0; q <= p; q++) idx.push_back(tri(active[p], active[q]));
    for (int c = 0; c < ncore; c++) idx.push_back(tri(core[c], core[c]));
    for (int p = 0; p < n; p++)
        for (int c = 0; c < ncore; c++) idx.push_back(tri(active[p], core[c]));
    for (int c = 0; c < ncore; c++)
        for (int c2 = 0; c2 < ncore; c2++) idx.push_back(tri(core[c], core[c2]));
    const int64_t ntri = (int64_t)nfull * (nfull + 1) / 2;
    if (df_.size() != (size_t)count()) df_.resize(count());
    // same shape as the resident state: keep its buffers (the orbital step's
    // workspace included); the orbital lists may differ, so the rotation
    // pairs and density tables are rebuilt by the next orbital call
    const bool reuse = df_[i] && df_[i]->nfull == nfull && df_[i]->nq == nq && df_[i]->n == n &&
                       df_[i]->ncore == ncore;
    if (!reuse) df_[i].reset(new DfState());
    DfState& D = *df_[i];
    D.nrot = -1;
    D.pair_key.clear();
    D.nfull = nfull;
    D.nq = nq;
    D.n = n;
    D.ncore = ncore;
    D.no = ncore + n;
    D.core.assign(core, core + ncore);
    D.active.assign(active, active + n);
    if (!reuse) D.qmo.alloc((size_t)nq * ntri);
    V2H_HIP(hipMemcpyAsync(D.qmo.p, qmo, sizeof(double) * nq * ntri, hipMemcpyHostToDevice, stream_));
    if (!reuse) D.h.alloc((size_t)nfull * nfull);
    V2H_HIP(hipMemcpyAsync(D.h.p, h_full, sizeof(double) * nfull * nfull, hipMemcpyHostToDevice, stream_));
    D.idx.upload(idx);
    D.act.upload(std::vector<int>(active, active + n));
    D.cor.upload(ncore > 0 ? std::vector<int>(core, core + ncore) : std::vector<int>{0});
    D.m_rows = (int64_t)idx.size();
    if (!reuse) {
        D.rows.alloc(idx.size() * (size_t)nq);
        D.dvec.alloc(nq);
        D.gu.alloc((size_t)kDfMaxSplits * nu * nu);
        D.sq.alloc(std::max<size_t>(1, (size_t)ncore * ncore));
        D.e.alloc(1);
        D.uacc.alloc((size_t)nfull * nfull);
    }
    launch_oo_identity(D.uacc.p, nfull, stream_);
    return run_df_intake(i);
}

DfState& Batch::df_state(int i) {
    if ((size_t)i >= df_.size() || !df_[i])
        throw std::invalid_argument("instance " + std::to_string(i) + " has no density-fitted integrals (set_df_integrals)");
    return *df_[i];
}

double Batch::run_df_intake(int i, bool norms) {
    launch_df_intake_only(i);
    DfState& D = df_state(i);
    V2H_HIP(hipMemcpyAsync(&D.efzc, D.e.p, sizeof(double), hipMemcpyDeviceToHost, stream_));
    V2H_HIP(hipStreamSynchronize(stream_));
    if (norms) refresh_norms(i);
    return D.efzc;
}

void Batch::launch_df_intake_only(int i) {
    const InstDev& I = inst_[i];
    const Structure& S = *structs_[inst_sid_[i]];
    DfState& D = df_state(i);
    DfArgs a{};
    a.qmo = D.qmo.p;
    a.h_full = D.h.p;
    a.idx = D.idx.p;
    a.active = D.act.p;
    a.core = D.cor.p;
    a.ntri = (int64_t)D.nfull * (D.nfull + 1) / 2;
    a.m_rows = D.m_rows;
    a.nfull = D.nfull;
    a.nq = D.nq;
    a.n = D.n;
    a.ncore = D.ncore;
    a.rows = D.rows.p;
    a.dvec = D.dvec.p;
    a.gu = D.gu.p;
    a.sq = D.sq.p;
    a.h_out = src_.p + I.soff;
    a.g_out = src_.p + I.soff + (int64_t)D.n * D.n;
    a.efzc = D.e.p;
    tstart(kKindDf);
    launch_df_intake(a, stream_);
    launch_pack_cost(S.cost1.p, S.cost2.p, src_.p + I.soff, c_.p + I.poff, S.n_primal, stream_);
    launch_oo_sub(c_.p + I.poff, z_.p + I.poff, v1_.p + I.poff, S.n_primal, stream_);  // v1 = c - z
    tstop(kKindDf);
    V2H_HIP(hipGetLastError());
}

double Batch::instance_energy(int i) {
    const InstDev& I = inst_[i];
    launch_dot1(c_.p + I.poff, x_.p + I.poff, n_primal(i), scratch_.p, stream_);
    double cx = 0.0;
    V2H_HIP(hipMemcpyAsync(&cx, scratch_.p, sizeof(double), hipMemcpyDeviceToHost, stream_));
    V2H_HIP(hipStreamSynchronize(stream_));
    return cx + df_state(i).efzc;
}

void Batch::set_cost(int i, const double* c) {
    const InstDev& I = inst_[i];
    const Structure& S = *structs_[inst_sid_[i]];
    V2H_HIP(hipMemcpyAsync(c_.p + I.poff, c, sizeof(double) * S.n_primal, hipMemcpyHostToDevice, stream_));
    launch_copy(v1_.p + I.poff, c_.p + I.poff, S.n_primal, stream_);
    refresh_norms(i);
}

void Batch::set_guess(int i, const double* x) {
    const InstDev& I = inst_[i];
    const int64_t np = structs_[inst_sid_[i]]->n_primal;
    std::vector<double> g;
    if (!x) {  // v2rdm_solver.cc:1799-1806
        g.resize(np);
        srand(0);
        for (int64_t k = 0; k < np; k++) g[k] = ((double)rand() / RAND_MAX - 1.0) * 2.0;
        x = g.data();
    }
    V2H_HIP(hipMemcpyAsync(x_.p + I.poff, x, sizeof(double) * np, hipMemcpyHostToDevice, stream_));
    V2H_HIP(hipStreamSynchronize(stream_));
}

void Batch::set_guess_hf(int i) {
    const Layout& L = layout(i);
    // the external-Hamiltonian path sets doccpi = nalpha, soccpi = 0
    // (v2rdm_solver.cc:1004-1007), so Guess() occupies nalpha beta orbitals too
    std::vector<double> x = hf_guess(L, build_constraints(L), L.na, L.na);
    set_guess(i, x.data());
}

void Batch::reset_state(int i) {
    const InstDev& I = inst_[i];
    const Structure& S = *structs_[inst_sid_[i]];
    launch_fill(reinterpret_cast<double*>(cg_.p + I.doff), 4 * S.n_dual, 0.0, stream_);
    launch_fill(z_.p + I.poff, S.n_primal, 0.0, stream_);
    launch_fill(aty_.p + I.poff, S.n_primal, 0.0, stream_);
    launch_copy(v1_.p + I.poff, c_.p + I.poff, S.n_primal, stream_);
    proj_.forget(i, stream_);  // the eigenvalues that warm-start the tridiagonal projection
    if (constructed_) sync_state_to_host();
    InstState& st = state_[i];
    double bn = st.bnorm, cn = st.cnorm;
    st = InstState{};
    st.bnorm = bn;
    st.cnorm = cn;
    st.mu = opt_.mu0 > 0 ? opt_.mu0 : 1.0;
    st.eps_p = st.eps_d = 1.0;
    st.active = 1;
    if (constructed_) sync_state_to_device();
    V2H_HIP(hipStreamSynchronize(stream_));
}

// The outer iterations of one call.  Lock-step: every running instance goes
// through pre, its CG (chunks of steps between host polls), post and the
// projection together.  Pipelined (tuning_.pipeline): the running instances
// form two groups that take turns on stream_; while one group runs pre, its
// CG and post, the other group's projection runs on the side streams, so the
// latency-bound projection hides behind the bandwidth-bound CG.  Every
// instance goes through the same kernels on the same data in the same order
// either way, so both schedules give the same bits.
int Batch::solve(int maxiter, const Monitor& mon) {
    auto t0 = std::chrono::steady_clock::now();
    V2H_HIP(hipSetDevice(device_));
    sync_state_to_host();
    for (auto& st : state_) {
        bool runnable = !st.converged && st.status == 0 && st.oiter < opt_.maxiter;
        st.active = runnable ? 1 : 0;
        st.maxiter_left = maxiter;
    }
    sync_state_to_device();
    const BatchDev B = dev();
    BatchDev Bg[2] = {B, B};
    Bg[0].pgroup = 0;
    Bg[1].pgroup = 1;
    std::vector<char> sel(count());
    // running instances of group g (g < 0: of the batch), as far as the host knows
    auto in_group = [&](int i, int g) { return state_[i].active != 0 && (g < 0 || grp_[i] == g); };
    auto running = [&](int g) {
        int n = 0;
        for (int i = 0; i < count(); i++) n += in_group(i, g) ? 1 : 0;
        return n;
    };
    auto select = [&](int g, bool iterating) {
        for (int i = 0; i < count(); i++)
            sel[i] = in_group(i, g) && (!iterating || !state_[i].cg_done) ? 1 : 0;
        set_plan(sel);
    };
    std::vector<int64_t> seen(count());
    for (int i = 0; i < count(); i++) seen[i] = state_[i].oiter;

    int cg_est[3] = {4, 4, 4};  // first CG chunk of group 0, group 1, the whole batch
    for (const auto& st : state_) cg_est[2] = std::max(cg_est[2], st.cg_last + tuning_.cg_margin);
    cg_est[0] = cg_est[1] = cg_est[2];
    // timer attribution: a launch covers the whole batch only in lock-step
    auto refresh_flags = [&](int g) {
        all_running_ = all_iterating_ = g < 0;
        for (const auto& st : state_) {
            all_running_ &= st.active != 0;
            all_iterating_ &= st.active && !st.cg_done;
        }
    };
    refresh_flags(-1);
    all_iterating_ = all_running_;

    // One outer iteration of group g (g < 0: the whole batch) on stream_:
    // pre, the CG in chunks with host polls, post and the projection.
    // Returns false when none of its instances runs any more.
    auto iterate = [&](int g) -> bool {
        const BatchDev& Bx = g < 0 ? B : Bg[g];
        int& est = cg_est[g < 0 ? 2 : g];
        refresh_flags(g);
        all_iterating_ = all_running_;
        select(g, false);  // the instances running, as far as the host knows
        tstart(0);
        launch_pre(Bx, units_a(), stream_);
        tstop(0);
        launches_++;
        if (mon) {
            sync_state_to_host();
            for (int i = 0; i < count(); i++)
                if (state_[i].oiter != seen[i]) {
                    seen[i] = state_[i].oiter;
                    mon(i, status(i));
                }
        }
        int steps = 0, chunk = est;
        bool any_active = true;
        const bool graphs = g < 0 && tuning_.graphs >= kGraphCg;
        for (;;) {
            const UnitsDev UA = units_a(), UAT = units_at();
            // the eigenvector formation owed by a projection graph runs as a
            // branch beside the first chunk, joined at its end
            const bool with_seed = graphs && steps == 0 && seed_owed_;
            auto cg_chunk = [&, chunk, with_seed] {
                hipStream_t rs = pstream_[tuning_.refine_lane - 1];
                if (with_seed) {
                    V2H_HIP(hipEventRecord(ev_gfork_, stream_));
                    V2H_HIP(hipStreamWaitEvent(rs, ev_gfork_, 0));
                    tstart(kKindSeed, rs);
                    proj_.refine_seed(Bx, rs);
                    tstop(kKindSeed, rs);
                    launches_++;
                    V2H_HIP(hipEventRecord(ev_gjoin_, rs));
                }
                for (int k = 0; k < chunk; k++) {
                    tstart(1);
                    launch_cg_at(Bx, UAT, stream_);
                    tstop(1);
                    tstart(2);
                    launch_cg_a(Bx, UA, stream_);
                    tstop(2);
                }
                launches_ += 2 * chunk;
                if (with_seed) V2H_HIP(hipStreamWaitEvent(stream_, ev_gjoin_, 0));
            };
            if (graphs && steps == 0) {  // the first chunk covers every running instance
                run_graph(cg_graph(chunk, with_seed), graph_key("cg" + std::to_string(chunk) + (with_seed ? "s" : "")),
                          cg_chunk);
                seed_owed_ = false;
            } else {
                cg_chunk();
            }
            steps += chunk;
            sync_state_to_host();
            refresh_flags(g);
            bool pending = false;
            any_active = false;
            for (int i = 0; i < count(); i++) {
                any_active |= in_group(i, g);
                pending |= in_group(i, g) && !state_[i].cg_done;
            }
            if (!pending || steps > opt_.cg_maxiter) break;
            select(g, true);  // the instances whose CG goes on
            chunk = 4;
        }
        if (!any_active) return false;
        est = 4;
        for (int i = 0; i < count(); i++)
            if (in_group(i, g)) est = std::max(est, state_[i].cg_last + tuning_.cg_margin);
        all_iterating_ = all_running_;  // the next CG solve starts for every running instance
        select(g, false);
        auto post_projection = [&](bool seed) {
            tstart(3);
            launch_post(Bx, units_at(), stream_);
            tstop(3);
            launches_++;
            // the size classes are independent: largest first, on the streams
            // of tuning_.class_lane, joined back into stream_ (lock-step)
            launch_projection(Bx, seed, g);
        };
        if (g < 0 && tuning_.graphs == kGraphAll) {
            // the graph joins every lane it forks, so the eigenvector formation
            // (off the critical path) goes into the next CG graph
            run_graph(g_post_, graph_key(proj_.refine_on() ? "post_r" : "post"), [&] { post_projection(false); });
            seed_owed_ = proj_.refine_on();
        } else {
            post_projection(true);
        }
        V2H_HIP(hipGetLastError());
        return true;
    };

    // two halves of the running instances, in batch order (no projection may
    // be in flight: set_groups waits for every stream)
    auto split = [&] {
        const int n = running(-1);
        std::vector<int> grp(count(), 0);
        int k = 0;
        for (int i = 0; i < count(); i++)
            if (state_[i].active) grp[i] = 2 * k++ >= n ? 1 : 0;
        set_groups(grp);
    };
    const bool pipe = tuning_.pipeline > 0 && tuning_.graphs == kGraphOff;
    bool piped = pipe && running(-1) >= tuning_.pipeline;
    if (piped) split();
    int left[2] = {maxiter, maxiter};  // outer iterations left to each group, or (left[0]) to the batch
    for (int turn = 0;;) {
        if (!piped) {
            if (left[0] <= 0 || !iterate(-1)) break;
            left[0]--;
            continue;
        }
        const int n0 = running(0), n1 = running(1);
        const bool live0 = n0 > 0 && left[0] > 0, live1 = n1 > 0 && left[1] > 0;
        if (!live0 && !live1) break;
        if (n0 + n1 < tuning_.pipeline) {  // too few instances to overlap: lock-step from here
            wait_projection(0);
            wait_projection(1);
            piped = false;
            left[0] = std::max(left[0], left[1]);
            continue;
        }
        if (std::max(n0, n1) - std::min(n0, n1) > 1 && 3 * std::min(n0, n1) < 2 * std::max(n0, n1)) {
            // unbalanced: split the running instances again
            wait_projection(0);
            wait_projection(1);
            split();
            left[0] = left[1] = std::max(left[0], left[1]);
            continue;
        }
        const int g = turn;
        turn ^= 1;
        if (!(g == 0 ? live0 : live1)) continue;
        wait_projection(g);  // the group's last projection
        if (iterate(g)) left[g]--;
    }
    wait_projection(0);
    wait_projection(1);
    select(-1, false);
    tstart(0);
    launch_pre(B, units_a(), stream_);  // book-keeping of the last iteration
    tstop(0);
    launches_++;
    if (seed_owed_) {  // the last projection graph's eigenvector formation
        launch_seed(B);
        seed_owed_ = false;
    }
    join_seed();
    V2H_HIP(hipGetLastError());
    sync_state_to_host();
    if (mon)
        for (int i = 0; i < count(); i++)
            if (state_[i].oiter != seen[i]) mon(i, status(i));
    last_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int nconv = 0;
    for (const auto& st : state_) nconv += st.converged;
    return nconv;
}

// The PSD projection of one outer iteration, after k_post on stream_.
//
// Without the refinement every size class starts right after k_post on its
// lane (tuning_.class_lane) and stream_ joins the lanes.  With it, the lane
// tuning_.refine_lane runs the warm-started refinement of the 65..256
// classes: its first kernel sets each block's route, after which the
// tridiagonal launches of the blocks known to need the fallback start on
// their lanes (tuning_.fallback_lane) beside the refinement; after the
// refinement the same lanes take the blocks it rejected.  stream_ joins every
// lane that writes x and z; the eigenvectors of the fallback blocks (needed
// only by the next iteration's refinement) are formed on the refinement lane
// after the fallback lanes, off the critical path.
void Batch::launch_projection(const BatchDev& B, bool seed, int g) {
    const int rl = tuning_.refine_lane;
    const ProjEvents& E = capture_ ? ev_capture_ : ev_eager_;
    V2H_HIP(hipEventRecord(E.post, stream_));
    bool used[kSideStreams + 1] = {};
    const bool refine = proj_.refine_on();
    hipStream_t rs = pstream_[rl - 1];
    if (refine) {
        V2H_HIP(hipStreamWaitEvent(rs, E.post, 0));
        used[rl] = true;
        tstart(kKindRefine, rs);
        proj_.refine_begin(B, rs);
        V2H_HIP(hipEventRecord(E.state, rs));
        proj_.refine_run(B, rs);
        tstop(kKindRefine, rs);
        launches_ += proj_.refine_launches();
        V2H_HIP(hipEventRecord(E.ref, rs));
    }
    // a pipelined group's projection stays off stream_, which runs the other
    // group's CG: the classes of lane 0 go to a side stream other than the
    // refinement's
    const int lane0 = g < 0 ? 0 : (rl == 2 ? 3 : 2);
    auto lane_of = [&](int ln) { return ln == 0 ? lane0 : ln; };
    auto lane_stream = [&](int ln) { return ln == 0 ? stream_ : pstream_[ln - 1]; };
    for (int k = kProjClasses - 1; k >= 0; k--) {
        if (!proj_.has(k) || (refine && Projector::refined_class(k))) continue;
        const int ln = lane_of(tuning_.class_lane[k]);
        hipStream_t ps = lane_stream(ln);
        if (ln > 0 && !used[ln]) V2H_HIP(hipStreamWaitEvent(ps, E.post, 0));
        used[ln] = true;
        tstart(4 + k, ps);
        proj_.launch(k, B, ps);
        tstop(4 + k, ps);
        launches_++;
    }
    if (refine) {
        for (int pass = 0; pass < 2; pass++) {  // fallback known at the start, then the rejections
            bool waited[kSideStreams + 1] = {};
            for (int k = kProjClasses - 1; k >= 0; k--) {
                if (!proj_.has(k) || !Projector::refined_class(k)) continue;
                const int ln = lane_of(tuning_.fallback_lane[k]);
                hipStream_t ps = lane_stream(ln);
                if (!waited[ln]) V2H_HIP(hipStreamWaitEvent(ps, pass == 0 ? E.state : E.ref, 0));
                waited[ln] = used[ln] = true;
                tstart(4 + k, ps);
                proj_.launch(k, B, ps, pass == 0 ? kRouteFallback : kRouteRejected);
                tstop(4 + k, ps);
                launches_++;
            }
        }
    }
    if (g >= 0) {
        // the group's next pre waits for these events (wait_projection); the
        // eigenvector formation, after every fallback launch of the group
        for (int ln = 1; ln <= kSideStreams; ln++) {
            pdone_used_[g][ln - 1] = used[ln];
            if (used[ln]) V2H_HIP(hipEventRecord(ev_pdone_[g][ln - 1], pstream_[ln - 1]));
        }
        if (refine && seed) {
            for (int ln = 1; ln <= kSideStreams; ln++)
                if (used[ln] && ln != rl) V2H_HIP(hipStreamWaitEvent(rs, ev_pdone_[g][ln - 1], 0));
            tstart(kKindSeed, rs);
            proj_.refine_seed(B, rs);
            tstop(kKindSeed, rs);
            launches_++;
            seed_pending_ = true;
        }
        return;
    }
    for (int ln = 1; ln <= kSideStreams; ln++)
        if (used[ln]) {
            V2H_HIP(hipEventRecord(E.proj[ln - 1], pstream_[ln - 1]));
            V2H_HIP(hipStreamWaitEvent(stream_, E.proj[ln - 1], 0));
        }
    if (refine && seed) {
        // after every fallback launch, whichever lane (stream_ included) ran it
        V2H_HIP(hipEventRecord(ev_joined_, stream_));
        V2H_HIP(hipStreamWaitEvent(rs, ev_joined_, 0));
        tstart(kKindSeed, rs);
        proj_.refine_seed(B, rs);
        tstop(kKindSeed, rs);
        launches_++;
        seed_pending_ = true;
    }
}

void Batch::wait_projection(int g) {
    for (int k = 0; k < kSideStreams; k++)
        if (pdone_used_[g][k]) {
            V2H_HIP(hipStreamWaitEvent(stream_, ev_pdone_[g][k], 0));
            pdone_used_[g][k] = false;
        }
}

void Batch::set_groups(const std::vector<int>& grp) {
    if (grp == grp_) return;
    // the projection kernels and the eigenvector formation read the groups
    V2H_HIP(hipStreamSynchronize(stream_));
    for (hipStream_t p : pstream_) V2H_HIP(hipStreamSynchronize(p));
    grp_ = grp;
    V2H_HIP(hipMemcpy(d_grp_.p, grp_.data(), sizeof(int) * grp_.size(), hipMemcpyHostToDevice));
}

// The eigenvector formation after a replayed projection graph, which has
// joined every lane into stream_.
void Batch::launch_seed(const BatchDev& B) {
    if (!proj_.refine_on()) return;
    hipStream_t rs = pstream_[tuning_.refine_lane - 1];
    V2H_HIP(hipEventRecord(ev_eager_.post, stream_));
    V2H_HIP(hipStreamWaitEvent(rs, ev_eager_.post, 0));
    tstart(kKindSeed, rs);
    proj_.refine_seed(B, rs);
    tstop(kKindSeed, rs);
    launches_++;
    seed_pending_ = true;
}

Batch::GraphEntry& Batch::cg_graph(int steps, bool seed) {
    const std::string key = graph_key("cg" + std::to_string(steps) + (seed ? "s" : ""));
    for (auto& e : g_cg_)
        if (e.key == key) return e;
    if (g_cg_.size() >= kCgGraphs) {  // the oldest step count goes
        drop_graph(g_cg_.front());
        g_cg_.erase(g_cg_.begin());
    }
    g_cg_.emplace_back();
    return g_cg_.back();
}

// Graph key: what a capture bakes into its launches (the plan's instance
// selection, hence grid sizes and plan lengths) and the sequence's own tag.
std::string Batch::graph_key(const std::string& tag) const {
    std::string k(plan_sel_.begin(), plan_sel_.end());
    for (auto& ch : k) ch = ch ? '1' : '0';
    return k + "|" + tag;
}

// The eigenvector formation of the last iteration runs on the refinement
// lane, which stream_ does not wait for; join it before anything else touches
// the projection state.
void Batch::join_seed() {
    if (!seed_pending_) return;
    V2H_HIP(hipEventRecord(ev_seed_, pstream_[tuning_.refine_lane - 1]));
    V2H_HIP(hipStreamWaitEvent(stream_, ev_seed_, 0));
    seed_pending_ = false;
}

void Batch::set_refine(bool on) {
    V2H_HIP(hipSetDevice(device_));
    if (on && !proj_.refine_on())
        for (int i = 0; i < count(); i++) proj_.forget(i, stream_);  // eigenvectors from before are stale
    proj_.set_refine(on);
    drop_graphs();
    V2H_HIP(hipStreamSynchronize(stream_));
}

InstStatus Batch::status(int i) {
    const InstState& st = state_[i];
    InstStatus s{};
    s.mu = st.mu;
    s.energy_primal = st.e_primal;
    s.energy_dual = st.e_dual;
    s.eps_p = st.eps_p;
    s.eps_d = st.eps_d;
    s.oiter = st.oiter;
    s.iiter = st.iiter;
    s.converged = st.converged;
    s.active = st.active;
    s.status = st.status;
    s.cg_last = st.cg_last;
    return s;
}

static void d2h(double* dst, const double* src, int64_t n, hipStream_t s) {
    V2H_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    V2H_HIP(hipStreamSynchronize(s));
}
static void h2d(double* dst, const double* src, int64_t n, hipStream_t s) {
    V2H_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, s));
    V2H_HIP(hipStreamSynchronize(s));
}

void Batch::get_primal(int i, double* x) { d2h(x, x_.p + inst_[i].poff, n_primal(i), stream_); }
// y is the fourth field of the CG rows: gathered into a contiguous buffer
void Batch::get_dual(int i, double* y) {
    DBuf<double> t;
    t.alloc(n_dual(i));
    launch_copy_strided(t.p, 1, &cg_.p[inst_[i].doff].y, 4, n_dual(i), stream_);
    d2h(y, t.p, n_dual(i), stream_);
}
void Batch::get_slack(int i, double* z) { d2h(z, z_.p + inst_[i].poff, n_primal(i), stream_); }
void Batch::get_cost(int i, double* c) { d2h(c, c_.p + inst_[i].poff, n_primal(i), stream_); }
void Batch::get_b(int i, double* b) const {
    const auto& v = structs_[inst_sid_[i]]->b;
    std::copy(v.begin(), v.end(), b);
}
void Batch::set_dual(int i, const double* y) {
    const Structure& S = *structs_[inst_sid_[i]];
    const InstDev& I = inst_[i];
    DBuf<double> t;
    t.alloc(n_dual(i));
    h2d(t.p, y, n_dual(i), stream_);
    launch_copy_strided(&cg_.p[I.doff].y, 4, t.p, 1, n_dual(i), stream_);
    // the next outer iteration starts from A (A^T y): keep aty = A^T y with y
    launch_fill(aty_.p + I.poff, S.n_primal, 0.0, stream_);
    launch_spmv_plain(S.dev().AT, unitsAT_.p, hwgAT_begin_[i], hwgAT_begin_[i + 1] - hwgAT_begin_[i], t.p,
                      aty_.p + I.poff, stream_);
    V2H_HIP(hipStreamSynchronize(stream_));
}
void Batch::set_slack(int i, const double* z) {
    const InstDev& I = inst_[i];
    h2d(z_.p + I.poff, z, n_primal(i), stream_);
    // v1 = c - z
    std::vector<double> c(n_primal(i));
    get_cost(i, c.data());
    for (int64_t k = 0; k < n_primal(i); k++) c[k] -= z[k];
    h2d(v1_.p + I.poff, c.data(), n_primal(i), stream_);
}
void Batch::set_cg_tolerance_factor(double f) {
    if (!(f > 0.0 && f <= 1.0)) throw std::invalid_argument("CG tolerance factor must be in (0, 1]");
    cg_factor_ = f;
    drop_graphs();  // the options are kernel arguments
}

void Batch::set_options(const SolverOptions& o) {
    const bool tighter = o.e_convergence < opt_.e_convergence || o.r_convergence < opt_.r_convergence;
    opt_ = o;
    drop_graphs();  // the options are kernel arguments
    if (!tighter || !constructed_) return;
    sync_state_to_host();
    for (auto& st : state_) st.converged = 0;
    sync_state_to_device();
}
double Batch::get_mu(int i) {
    sync_state_to_host();
    return state_[i].mu;
}
void Batch::set_mu(int i, double mu) {
    sync_state_to_host();
    state_[i].mu = mu;
    sync_state_to_device();
}

void Batch::apply_A(int i, const double* u, double* out) {
    const Structure& S = *structs_[inst_sid_[i]];
    double* in = scratch_.p;
    double* o = scratch_.p + std::max(S.n_primal, S.n_dual);
    h2d(in, u, S.n_primal, stream_);
    launch_fill(o, S.n_dual, 0.0, stream_);
    launch_spmv_plain(S.dev().A, unitsA_.p, hwgA_begin_[i], hwgA_begin_[i + 1] - hwgA_begin_[i], in, o, stream_);
    d2h(out, o, S.n_dual, stream_);
}

void Batch::apply_A_device(int i, const double* in_dev, double* out_dev) {
    const Structure& S = *structs_[inst_sid_[i]];
    launch_fill(out_dev, S.n_dual, 0.0, stream_);
    launch_spmv_plain(S.dev().A, unitsA_.p, hwgA_begin_[i], hwgA_begin_[i + 1] - hwgA_begin_[i], in_dev, out_dev,
                      stream_);
}

void Batch::apply_AT(int i, const double* y, double* out) {
    const Structure& S = *structs_[inst_sid_[i]];
    double* in = scratch_.p;
    double* o = scratch_.p + std::max(S.n_primal, S.n_dual);
    h2d(in, y, S.n_dual, stream_);
    launch_fill(o, S.n_primal, 0.0, stream_);
    launch_spmv_plain(S.dev().AT, unitsAT_.p, hwgAT_begin_[i], hwgAT_begin_[i + 1] - hwgAT_begin_[i], in, o,
                      stream_);
    d2h(out, o, S.n_primal, stream_);
}

void Batch::cost_adjoint(int i, const double* xbar, double* dh, double* dg) {
    const Structure& S = *structs_[inst_sid_[i]];
    const InstDev& I = inst_[i];
    const double* xb = x_.p + I.poff;
    if (xbar) {
        h2d(scratch_.p, xbar, S.n_primal, stream_);
        xb = scratch_.p;
    }
    double* out = scratch_.p + S.n_primal;
    if (scratch_.n < (size_t)(S.n_primal + S.nsrc)) throw std::runtime_error("adjoint scratch too small");
    launch_cost_adjoint(S.adj_ptr.p, S.adj_ent.p, xb, out, S.nsrc, stream_);
    const int64_t n = S.L.n;
    d2h(dh, out, n * n, stream_);
    d2h(dg, out + n * n, n * n * n * n, stream_);
}

}  // namespace v2h
