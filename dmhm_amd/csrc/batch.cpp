// Batches of independent H-matrices of different geometries, worked on with
// one launch per phase for all members: the reference inverts such operators
// one by one (HMat3d::SchulzInvert per operator, src/HMat3d/Invert-incl.hpp:
// 127-188); here their plans are merged (merge_plans) and run together.
//
// A role of the batch (the operators, the iterates, Z) is one arena: one
// allocation holding every member's values at its own offset and one holding
// every member's ranks, so a merged plan binds each slot once and every
// member's tasks carry their member's offsets. Members of one class (same
// geometry and storage layout contents) share one host plan. Every member
// draws the seeds it would draw alone, so a member's result in the batch is
// its result alone, up to the tile engine each merged phase tunes to.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include "runtime.hpp"

namespace dmhm {

// Several H-matrices in one allocation; views are non-owning handles.
struct Arena {
    int prec = DMHM_F64;
    std::vector<std::unique_ptr<dmhm_hmat_s>> views;
    std::vector<int64_t> val, rank;     // member bases (elements, rank entries)
    int64_t totalVals = 0, totalRanks = 0;
    DeviceBuffer vals, ranks;
    size_t size() const { return views.size(); }
    dmhm_hmat_s* at(size_t i) const { return views[i].get(); }
    size_t esize() const { return prec == DMHM_F64 ? 8 : 4; }
    ~Arena() {
        views.clear();
        vals.release();
        ranks.release();
    }
};

// Allocates an arena of members with the geometries of `like` and layouts
// `lays`, zero-filled.
static void arena_alloc(Arena& a, const std::vector<dmhm_hmat_s*>& like,
                        const std::vector<std::shared_ptr<Layout>>& lays, int prec)
{
    a.views.clear();
    a.val.clear();
    a.rank.clear();
    a.prec = prec;
    int64_t v = 0, r = 0;
    for (size_t i = 0; i < like.size(); ++i) {
        a.val.push_back(v);
        a.rank.push_back(r);
        v += (lays[i]->totalElems() + 63) & ~int64_t(63);
        r += (std::max(lays[i]->nLowRank(), 1) + 63) & ~63;
    }
    const size_t es = a.esize();
    a.totalVals = v;
    a.totalRanks = r;
    a.vals.ensure(std::max<size_t>(v * es, 256));
    a.ranks.ensure(std::max<size_t>(r * sizeof(int), 256));
    HIPCHECK(hipMemsetAsync(a.vals.ptr, 0, v * es, g_lib.stream));
    HIPCHECK(hipMemsetAsync(a.ranks.ptr, 0, r * sizeof(int), g_lib.stream));
    for (size_t i = 0; i < like.size(); ++i) {
        std::unique_ptr<dmhm_hmat_s> h(new dmhm_hmat_s());
        h->owner = false;
        h->gc = like[i]->gc;
        h->lay = lays[i];
        h->stats = like[i]->stats;
        h->prec = prec;
        h->maxRank = like[i]->maxRank;
        h->vals = static_cast<char*>(a.vals.ptr) + a.val[i] * es;
        h->ranks = static_cast<int*>(a.ranks.ptr) + a.rank[i];
        a.views.push_back(std::move(h));
    }
}

static void copy_member(dmhm_hmat_s* dst, const dmhm_hmat_s* src)
{
    if (dst->lay->totalElems() != src->lay->totalElems() || dst->lay->cap != src->lay->cap)
        throw Error(DMHM_EINVAL, "member copy between different layouts");
    HIPCHECK(hipMemcpyAsync(dst->vals, src->vals, src->lay->totalElems() * src->esize(), hipMemcpyDeviceToDevice,
                            g_lib.stream));
    HIPCHECK(hipMemcpyAsync(dst->ranks, src->ranks, src->lay->nLowRank() * sizeof(int), hipMemcpyDeviceToDevice,
                            g_lib.stream));
    dst->full = src->full;
}

static std::vector<std::shared_ptr<Layout>> std_layouts(const Arena& a)
{
    std::vector<std::shared_ptr<Layout>> out;
    for (size_t i = 0; i < a.size(); ++i) out.push_back(std_layout(*a.at(i)->gc, a.at(i)->maxRank));
    return out;
}

static std::vector<dmhm_hmat_s*> members(const Arena& a)
{
    std::vector<dmhm_hmat_s*> out;
    for (size_t i = 0; i < a.size(); ++i) out.push_back(a.at(i));
    return out;
}

// Builds `count` host plans on worker threads (the planner is host-only and
// shares nothing between geometries' plans).
static void build_parallel(size_t count, const std::function<void(size_t)>& build)
{
    const size_t nt = std::min<size_t>(count, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::exception_ptr> err(nt);
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            try {
                for (size_t i = t; i < count; i += nt) build(i);
            } catch (...) {
                err[t] = std::current_exception();
            }
        });
    for (auto& x : th) x.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

// Merged plan of C_i := alpha A_i B_i (mode: plain product) for every member.
// Members whose product has a tile phase of at least this many tiles run
// unmerged (a chunk of their own).
constexpr size_t kAloneTiles = 65536;

// Member (class) product plans, shared by every batch whose members have the
// same geometry, layouts and operand forms: built, uploaded and tile-tuned
// once. (Allocated once and released by dmhm_finalize.)
static std::map<std::string, std::shared_ptr<Plan>>* g_classPlans = new std::map<std::string, std::shared_ptr<Plan>>();

static Plan* merged_product(const Arena& A, const Arena& B, const Arena& C, int oversample)
{
    const size_t M = A.size();
    struct Cls { const dmhm_hmat_s *a, *b, *c; std::shared_ptr<Plan> plan; };
    std::vector<Cls> cls;
    std::vector<size_t> of(M);
    for (size_t m = 0; m < M; ++m) {
        const dmhm_hmat_s *a = A.at(m), *b = B.at(m), *c = C.at(m);
        if (a->gc != b->gc || a->gc != c->gc) throw Error(DMHM_ENONCONF, "Attempted nonconformal matrix-matrix multiply");
        size_t k = 0;
        for (; k < cls.size(); ++k) {
            const Cls& x = cls[k];
            if (x.a->gc == a->gc && x.c->maxRank == c->maxRank && x.a->full == a->full && x.b->full == b->full &&
                x.a->lay->cap == a->lay->cap && x.b->lay->cap == b->lay->cap && x.c->lay->cap == c->lay->cap)
                break;
        }
        if (k == cls.size()) cls.push_back({a, b, c, nullptr});
        of[m] = k;
    }
    // Each class plan is the plan of a lone member's product (same workspace
    // budget, so the same level-1 batches and phases): members merge phase
    // for phase, and a batch launches what one member launches.
    const int64_t budget = hh_batch_budget();
    std::vector<std::string> ckey(cls.size());
    std::vector<size_t> todo;
    for (size_t k = 0; k < cls.size(); ++k) {
        const Cls& x = cls[k];
        ckey[k] = std::to_string(reinterpret_cast<uintptr_t>(x.a->gc.get())) + ":" + std::to_string(x.a->lay->id) + "," +
                  std::to_string(x.b->lay->id) + "," + std::to_string(x.c->lay->id) + ":" +
                  std::to_string((int)x.a->full) + std::to_string((int)x.b->full) + ":" +
                  std::to_string(x.c->maxRank) + ":" + std::to_string(oversample);
        auto it = g_classPlans->find(ckey[k]);
        if (it != g_classPlans->end()) cls[k].plan = it->second;
        else todo.push_back(k);
    }
    build_parallel(todo.size(), [&](size_t i) {
        Cls& x = cls[todo[i]];
        Planner pl(x.a->geo());
        pl.setWorkspaceBudget(budget);
        pl.setStacking(x.a->full, x.b->full);
        x.plan.reset(pl.hhProduct(*x.a->lay, *x.b->lay, *x.c->lay, x.c->maxRank, oversample));
        x.plan->keepHost = true;
    });
    std::unique_ptr<Plan> holder(new Plan());
    // Chunks of members of one class whose workspaces together fit the budget
    // (at least one member each), one merged plan per chunk; the chunks run
    // one after another in one shared workspace. Members of a class share
    // their plan, so a chunk's merged phases are the class plan's phases and
    // its tiles keep the class's shapes (a phase's engine is instantiated for
    // its largest tile, which a smaller member would otherwise run at).
    std::vector<size_t> order(M);
    for (size_t m = 0; m < M; ++m) order[m] = m;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return of[x] < of[y]; });
    // A member whose tile phases fill the chip many times over runs as a
    // chunk of its own: merging would interleave the members phase by phase,
    // and a phase's output (hundreds of MB) would leave the caches before
    // the next phase of the same member reads it. (C3's 3D operators carry
    // 0.8-1M tiles in their largest phase at every size from 8^3 to 32^3; 48
    // of them in chunks of several members took 1.34x the time of one by
    // one.)
    std::vector<char> alone(cls.size(), 0);
    for (size_t k = 0; k < cls.size(); ++k)
        for (const Phase& ph : cls[k].plan->phases)
            if (ph.kind == PH_TILE && ph.tiles.size() >= kAloneTiles) alone[k] = 1;
    const int64_t slots[NUM_SLOTS] = {A.totalVals, B.totalVals, C.totalVals, 0, 0, 0, 0, 0};
    const int64_t ranks[NUM_SLOTS] = {A.totalRanks, B.totalRanks, C.totalRanks, 0, 0, 0, 0, 0};
    for (size_t i0 = 0; i0 < M;) {
        int64_t ws = 0;
        size_t i1 = i0;
        while (i1 < M) {
            const size_t m = order[i1];
            const int64_t w = (cls[of[m]].plan->wsElems + 63) & ~int64_t(63);
            if (i1 > i0 && (ws + w > budget || of[m] != of[order[i0]] || alone[of[m]])) break;
            ws += w;
            ++i1;
        }
        std::vector<PlanPart> parts(i1 - i0);
        for (size_t i = i0; i < i1; ++i) {
            const size_t m = order[i];
            PlanPart& p = parts[i - i0];
            p.plan = cls[of[m]].plan.get();
            p.val[SLOT_A] = A.val[m]; p.rank[SLOT_A] = A.rank[m];
            p.val[SLOT_B] = B.val[m]; p.rank[SLOT_B] = B.rank[m];
            p.val[SLOT_C] = C.val[m]; p.rank[SLOT_C] = C.rank[m];
        }
        // Tile phases by segment (no copies of the class plans' tiles); the
        // smaller task lists (QR, SVD, sketches) are concatenated.
        std::unique_ptr<Plan> merged(merge_plans(parts, true));
        const std::string e = check_plan(*merged, slots, ranks);
        if (!e.empty()) throw Error(DMHM_EINVAL, "merged batch product: " + e);
        holder->wsElems = std::max(holder->wsElems, merged->wsElems);
        holder->flops += merged->flops;
        holder->chunks.push_back(std::move(merged));
        i0 = i1;
    }
    // Segmented tile phases read the class plans' arrays on the device, so
    // new class plans go up first (keeping their host arrays, which later
    // merged plans are checked against).
    for (size_t k : todo) {
        cls[k].plan->upload();
        (*g_classPlans)[ckey[k]] = cls[k].plan;
    }
    for (auto& x : cls) holder->members.push_back(x.plan);
    for (auto& c : holder->chunks) c->upload();
    return holder.release();
}

// Merged streaming H*x (one right-hand side) of every member; member i's x and
// y at vecBase[i] of their buffers.
static Plan* merged_matvec(const Arena& H, const std::vector<int64_t>& vecBase, int64_t vecElems)
{
    const size_t M = H.size();
    struct Cls { const dmhm_hmat_s* h; std::unique_ptr<Plan> plan; };
    std::vector<Cls> cls;
    std::vector<size_t> of(M);
    for (size_t m = 0; m < M; ++m) {
        const dmhm_hmat_s* h = H.at(m);
        size_t k = 0;
        for (; k < cls.size(); ++k)
            if (cls[k].h->gc == h->gc && cls[k].h->lay->cap == h->lay->cap) break;
        if (k == cls.size()) cls.push_back({h, nullptr});
        of[m] = k;
    }
    build_parallel(cls.size(), [&](size_t k) {
        const int n = cls[k].h->geo().n;
        cls[k].plan.reset(Planner(cls[k].h->geo()).matvecStream(*cls[k].h->lay, 1, n, n));
    });
    std::vector<PlanPart> parts(M);
    for (size_t m = 0; m < M; ++m) {
        PlanPart& p = parts[m];
        p.plan = cls[of[m]].plan.get();
        p.val[SLOT_A] = H.val[m]; p.rank[SLOT_A] = H.rank[m];
        p.val[SLOT_X] = vecBase[m]; p.val[SLOT_Y] = vecBase[m];
    }
    std::unique_ptr<Plan> merged(merge_plans(parts));
    for (const Phase& ph : merged->phases) {
        // Streaming phases: projection items, accumulation atoms and their
        // contributions inside the operator arena, the vectors and the workspace.
        for (const MvProjItem& it : ph.mvProj)
            if (it.voff < 0 || it.voff + (int64_t)it.ldv * (it.cap - 1) + it.rows > H.totalVals ||
                it.xrow < 0 || it.xrow + it.rows > vecElems || it.poff < 0 || it.poff + it.cap > merged->wsElems ||
                it.rankIdx >= H.totalRanks)
                throw Error(DMHM_EINVAL, "merged batch matvec: projection out of bounds");
        for (const MvAtom& at : ph.mvAtoms)
            if (at.yrow < 0 || at.yrow + at.m > vecElems || at.cb < 0 || at.ce > (int)ph.mvContribs.size())
                throw Error(DMHM_EINVAL, "merged batch matvec: atom out of bounds");
    }
    merged->upload();
    return merged.release();
}

} // namespace dmhm

using namespace dmhm;

struct dmhm_hbatch_s {
    Arena a;                                            // the members
    // Power iteration: per-member vector segments and buffers, log accumulators.
    DeviceBuffer segs, vecs, logAcc;
    std::vector<VecSegment> hsegs;
    std::vector<int64_t> vecBase;
    int64_t vecElems = 0, maxN = 0;
    double* hRead = nullptr;                            // pinned readback of per-member estimates
    // AddConstantToDiagonal over an arena of standard layouts: leaf offsets, sizes.
    DeviceBuffer diagOff, diagSize;
    int diagCount = 0;
    ~dmhm_hbatch_s() {
        segs.release(); vecs.release(); logAcc.release();
        diagOff.release(); diagSize.release();
        if (hRead) (void)hipHostFree(hRead);
    }
};

namespace dmhm {

// Per-member power-step counts of EstimateTwoNorm (hmat_tools.hpp:1504-1544).
static int power_steps(int n, double theta, double confidence)
{
    return (int)std::ceil(std::log(0.8 * std::sqrt((double)n) * std::pow(10.0, confidence)) / std::log(theta));
}

static void ensure_vectors(dmhm_hbatch_s& B, double theta, double confidence)
{
    const size_t M = B.a.size();
    B.hsegs.clear();
    B.vecBase.clear();
    int64_t off = 0;
    B.maxN = 0;
    for (size_t m = 0; m < M; ++m) {
        const int n = B.a.at(m)->geo().n;
        B.vecBase.push_back(off);
        B.hsegs.push_back({off, n, power_steps(n, theta, confidence)});
        off += (n + 63) & ~63;
        B.maxN = std::max<int64_t>(B.maxN, n);
    }
    B.vecElems = off;
    B.segs.ensure(M * sizeof(VecSegment));
    HIPCHECK(copy_sync(B.segs.ptr, B.hsegs.data(), M * sizeof(VecSegment), hipMemcpyHostToDevice));
    B.vecs.ensure(std::max<size_t>(2 * off * sizeof(double), 256));
    B.logAcc.ensure(std::max<size_t>(M * sizeof(double), 256));
    if (!B.hRead) HIPCHECK(hipHostMalloc((void**)&B.hRead, std::max<size_t>(M, 1) * sizeof(double), hipHostMallocDefault));
}

// EstimateTwoNorm of every member of H (launches only; the logs of the
// estimates accumulate in B.logAcc).
template <typename T>
static void launch_batch_two_norm(dmhm_hbatch_s& B, const Arena& H, Plan& mv, uint64_t seed)
{
    hipStream_t s = g_lib.stream;
    const int M = (int)H.size();
    const VecSegment* segs = static_cast<const VecSegment*>(B.segs.ptr);
    T* x = static_cast<T*>(B.vecs.ptr);
    T* y = x + B.vecElems;
    double* acc = static_cast<double*>(B.logAcc.ptr);
    int kmax = 0;
    for (const auto& sg : B.hsegs) kmax = std::max(kmax, sg.steps);
    launch_seg_gaussian<T>(x, segs, M, B.maxN, seed, 7, s, g_lib.seedOff);
    launch_zero_double(acc, M, s);
    launch_seg_normalize<T>(x, segs, M, nullptr, -1, s);
    Bases b{};
    b.val[SLOT_A] = H.vals.ptr;
    b.rank[SLOT_A] = static_cast<int*>(H.ranks.ptr);
    b.val[SLOT_WS] = g_lib.ws.ptr;
    for (int i = 0; i < kmax; ++i) {
        b.val[SLOT_X] = x;
        b.val[SLOT_Y] = y;
        run_plan<T>(mv, b, T(1), T(0), {0, 0}, 0);
        launch_seg_normalize<T>(y, segs, M, acc, i, s);
        std::swap(x, y);
    }
}

static void read_batch_two_norm(dmhm_hbatch_s& B, double theta, std::vector<double>& est)
{
    const size_t M = B.a.size();
    HIPCHECK(hipMemcpyAsync(B.hRead, B.logAcc.ptr, M * sizeof(double), hipMemcpyDeviceToHost, g_lib.stream));
    HIPCHECK(hipStreamSynchronize(g_lib.stream));
    est.resize(M);
    for (size_t m = 0; m < M; ++m) est[m] = theta * std::exp(B.hRead[m]);
}

// Merged plans are cached by the member layouts they were built for (layout
// ids are unique in the process and layouts live as long as a matrix uses
// them), so batches of the same operators share them.
// (Allocated once and released by dmhm_finalize, not by a static destructor
// that could run after the HIP runtime is gone.)
static std::map<std::string, std::unique_ptr<Plan>>* g_batchPlans = new std::map<std::string, std::unique_ptr<Plan>>();

void release_batch_plans()
{
    g_batchPlans->clear();
    g_classPlans->clear();
}

static Plan& batch_plan(dmhm_hbatch_s& B, const std::string& key, const std::function<Plan*()>& make)
{
    (void)B;
    auto it = g_batchPlans->find(key);
    if (it != g_batchPlans->end()) return *it->second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHECK(hipStreamIsCapturing(g_lib.stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
        throw Error(DMHM_EINVAL, "a batch plan would be built while the step is captured");
    std::unique_ptr<Plan> p(make());
    Plan& ref = *p;
    (*g_batchPlans)[key] = std::move(p);
    return ref;
}

// Identity of an arena's members for the plan cache: geometry, layout id
// (standard layouts are numbered by rank, so the geometry is part of it) and
// the full-rank flag, member by member.
// (The flag matters for operands only: it selects their plan form.)
static std::string arena_key(const Arena& a, bool withFull = true)
{
    std::string k;
    for (size_t i = 0; i < a.size(); ++i) {
        const Geometry& g = a.at(i)->geo();
        k += std::to_string(g.dim) + "/" + std::to_string(g.sizes[0]) + "x" + std::to_string(g.sizes[1]) + "x" +
             std::to_string(g.sizes[2]) + "/" + std::to_string(g.numLevels) + (g.strong ? "s/" : "w/") +
             std::to_string(a.at(i)->lay->id) + (withFull && a.at(i)->full ? "f," : ",");
    }
    return k;
}

// C_i := alpha A_i B_i into C (standard layouts), one merged plan.
static void batch_multiply_into(dmhm_hbatch_s& P, double alpha, const Arena& A, const Arena& B, Arena& C,
                                const dmhm_hh_opts& o)
{
    if (A.prec != B.prec || A.prec != C.prec) throw Error(DMHM_EINVAL, "operands must share a precision");
    const int over = o.oversample > 0 ? o.oversample : 4;
    if (o.rankModel) throw Error(DMHM_EINVAL, "batched products take no rank model");
    const std::string key = "hh:" + arena_key(A) + "|" + arena_key(B) + "|" + arena_key(C, false) + ":" +
                            std::to_string(over);
    Plan& p = batch_plan(P, key, [&] { return merged_product(A, B, C, over); });
    ensure_ws(p, A.esize());
    Bases b{};
    b.val[SLOT_A] = A.vals.ptr; b.rank[SLOT_A] = static_cast<int*>(A.ranks.ptr);
    b.val[SLOT_B] = B.vals.ptr; b.rank[SLOT_B] = static_cast<int*>(B.ranks.ptr);
    b.val[SLOT_C] = C.vals.ptr; b.rank[SLOT_C] = static_cast<int*>(C.ranks.ptr);
    b.val[SLOT_WS] = g_lib.ws.ptr;
    const TruncPolicy pol{o.tolMode, o.relTol};
    for (auto& chunk : p.chunks) {
        if (A.prec == DMHM_F64) run_plan<double>(*chunk, b, alpha, 0.0, pol, o.seed);
        else run_plan<float>(*chunk, b, (float)alpha, 0.0f, pol, o.seed);
    }
    for (size_t i = 0; i < C.size(); ++i) C.at(i)->full = true;
}

// Diagonal dense leaves of an arena of standard layouts (AddConstantToDiagonal).
static void ensure_diag(dmhm_hbatch_s& B, const Arena& Z)
{
    if (B.diagCount > 0) return;
    std::vector<int64_t> offs;
    std::vector<int> sizes;
    for (size_t m = 0; m < Z.size(); ++m) {
        const Geometry& g = Z.at(m)->geo();
        const Layout& L = *Z.at(m)->lay;
        for (size_t i = 0; i < g.denseBlocks.size(); ++i) {
            const Block& bl = g.blocks[g.denseBlocks[i]];
            if (bl.t != bl.s) continue;
            offs.push_back(Z.val[m] + L.denseOff[i]);
            sizes.push_back(g.tbox(g.denseBlocks[i]).size);
        }
    }
    B.diagOff.ensure(std::max<size_t>(offs.size() * sizeof(int64_t), 8));
    B.diagSize.ensure(std::max<size_t>(sizes.size() * sizeof(int), 4));
    if (!offs.empty()) {
        HIPCHECK(copy_sync(B.diagOff.ptr, offs.data(), offs.size() * sizeof(int64_t), hipMemcpyHostToDevice));
        HIPCHECK(copy_sync(B.diagSize.ptr, sizes.data(), sizes.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    B.diagCount = (int)offs.size();
}

static void batch_add_diag(dmhm_hbatch_s& B, const Arena& Z, double alpha)
{
    const int64_t* o = static_cast<const int64_t*>(B.diagOff.ptr);
    const int* s = static_cast<const int*>(B.diagSize.ptr);
    if (Z.prec == DMHM_F64) launch_add_diag<double>(static_cast<double*>(Z.vals.ptr), o, s, B.diagCount, alpha, g_lib.stream);
    else launch_add_diag<float>(static_cast<float*>(Z.vals.ptr), o, s, B.diagCount, (float)alpha, g_lib.stream);
}

// SchulzInvert of every member (Invert-incl.hpp:128-196 per member): the
// same iteration, seeds and stop rule as a lone inversion, each member
// stopping on its own estimate; every step of the batch is one merged launch
// per phase (the graph-replayed step once tuned, as for one matrix).
static void batch_schulz(dmhm_hbatch_s& B, const dmhm_schulz_opts& o, dmhm_schulz_report* reps)
{
    const auto t0 = std::chrono::steady_clock::now();
    Arena& A = B.a;
    const size_t M = A.size();
    if (M == 0) return;
    if (o.mixed || o.bf16 || o.start != 0 || o.hh.rankModel)
        throw Error(DMHM_EINVAL, "batched Schulz runs in the members' precision from alpha A^T, without a rank model");
    if (o.theta <= 1) throw Error(DMHM_EINVAL, "Theta must be > 1");
    if (o.confidence <= 0) throw Error(DMHM_EINVAL, "Confidence must be positive");
    bool stop = false;
    int numIt = o.numIterations;
    if (numIt <= 0) { numIt = o.maxIterations > 0 ? o.maxIterations : 30; stop = true; }
    if (numIt > 4096) throw Error(DMHM_EINVAL, "at most 4096 Schulz iterations");
    const int prec = A.prec;
    const uint64_t seed = o.hh.seed;
    hipStream_t s = g_lib.stream;

    ensure_vectors(B, o.theta, o.confidence);
    launch_zero_double(reinterpret_cast<double*>(g_lib.dCounter), 1, s);
    SeedOffsetScope seeds;
    // X_0 = alpha_i A_i^T with alpha_i = 2 / est_i^2.
    Plan& mvA = batch_plan(B, "mv:" + arena_key(A), [&] { return merged_matvec(A, B.vecBase, B.vecElems); });
    ensure_ws(mvA, A.esize());
    if (prec == DMHM_F64) launch_batch_two_norm<double>(B, A, mvA, seed);
    else launch_batch_two_norm<float>(B, A, mvA, seed);
    std::vector<double> normEst;
    read_batch_two_norm(B, o.theta, normEst);
    Arena X0;
    {
        std::vector<std::unique_ptr<dmhm_hmat_s>> tr;
        std::vector<std::shared_ptr<Layout>> lays;
        for (size_t m = 0; m < M; ++m) {
            tr.emplace_back(transpose_hmat(A.at(m)));
            lays.push_back(tr.back()->lay);
        }
        arena_alloc(X0, members(A), lays, prec);
        for (size_t m = 0; m < M; ++m) {
            copy_member(X0.at(m), tr[m].get());
            scale_hmat(X0.at(m), 2.0 / (normEst[m] * normEst[m]));
        }
        HIPCHECK(hipStreamSynchronize(s));
    }
    const auto LS = std_layouts(A);
    Arena X[2], Z, R;
    arena_alloc(X[0], members(A), LS, prec);
    arena_alloc(X[1], members(A), LS, prec);
    arena_alloc(Z, members(A), LS, prec);
    arena_alloc(R, members(A), LS, prec);
    ensure_diag(B, Z);
    Plan* mvZ = nullptr;

    std::vector<int> doneAt(M, 0);
    std::vector<std::vector<double>> hist(M);
    std::vector<double> est(M, -1.0);
    const Arena* Xc = &X0;
    int side = 0, it = 0, itGraph = 0, steps = 0;
    StepGraph graphs[2];
    bool captured[2] = {false, false}, tuned = false;
    auto step = [&](const Arena& Xin, Arena& Xout, uint64_t sd) {
        launch_bump(g_lib.dCounter, s);
        dmhm_hh_opts h = o.hh;
        h.seed = sd;
        batch_multiply_into(B, -1.0, Xin, A, Z, h);
        batch_add_diag(B, Z, 1.0);
        if (stop) {
            if (prec == DMHM_F64) launch_batch_two_norm<double>(B, Z, *mvZ, sd + 1);
            else launch_batch_two_norm<float>(B, Z, *mvZ, sd + 1);
        }
        batch_add_diag(B, Z, 1.0);
        h.seed = sd + 2;
        batch_multiply_into(B, 1.0, Z, Xin, Xout, h);
    };
    for (int kk = 0; kk < numIt; ++kk) {
        it = kk + 1;
        if (stop && !mvZ) {
            mvZ = &batch_plan(B, "mv:" + arena_key(Z), [&] { return merged_matvec(Z, B.vecBase, B.vecElems); });
            ensure_ws(*mvZ, Z.esize());
        }
        const uint64_t sd = seed + 1 + 3 * (uint64_t)(prec == DMHM_F64 ? 0 : 1);
        const bool steady = Xc == &X[0] || Xc == &X[1];
        if (o.graph && steady && steps >= 2 && tuned) {
            if (!captured[side]) {
                const Arena& xin = *Xc;
                Arena& xout = X[side];
                graphs[side].capture([&] { step(xin, xout, sd); });
                captured[side] = true;
            }
            graphs[side].replay();
            ++itGraph;
        } else {
            const long u0 = g_lib.untuned;
            step(*Xc, X[side], sd);
            if (steady && g_lib.untuned == u0) tuned = true;
        }
        ++steps;
        const Arena* Xprev = Xc;
        Xc = &X[side];
        side ^= 1;
        bool all = true;
        if (stop) {
            read_batch_two_norm(B, o.theta, est);
            for (size_t m = 0; m < M; ++m) {
                if (doneAt[m]) continue;
                hist[m].push_back(est[m]);
                const bool stalled = o.stallStop && stall_detected(hist[m]);
                if (std::fabs(est[m]) < o.stopTol || stalled) {
                    // At the floor the member keeps the step's input X_k, as
                    // the single-operator iteration does.
                    doneAt[m] = it;
                    copy_member(R.at(m), (stalled && !(std::fabs(est[m]) < o.stopTol) ? Xprev : Xc)->at(m));
                } else {
                    all = false;
                }
            }
        } else {
            all = false;
        }
        if (stop && all) break;
    }
    for (size_t m = 0; m < M; ++m)
        if (!doneAt[m]) copy_member(R.at(m), Xc->at(m));
    HIPCHECK(hipStreamSynchronize(s));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    bool conv = true;
    for (size_t m = 0; m < M; ++m) {
        if (reps) {
            dmhm_schulz_report& r = reps[m];
            std::memset(&r, 0, sizeof(r));
            r.iterations = doneAt[m] ? doneAt[m] : it;
            r.residualEstimate = hist[m].empty() ? -1.0 : hist[m].back();
            r.normEstimate = normEst[m];
            r.seconds = secs;
            r.graphIterations = std::min(itGraph, r.iterations);
            for (size_t k = 0; k < hist[m].size() && k < 128; ++k) r.history[k] = hist[m][k];
        }
        if (stop && !(hist[m].size() && std::fabs(hist[m].back()) < o.stopTol)) conv = false;
    }
    // The members become their inverses (standard layouts).
    std::swap(A.views, R.views);
    std::swap(A.val, R.val);
    std::swap(A.rank, R.rank);
    std::swap(A.vals, R.vals);
    std::swap(A.ranks, R.ranks);
    for (size_t m = 0; m < M; ++m) A.at(m)->full = true;
    B.diagCount = 0;
    if (!conv) throw Error(DMHM_ENOCONV, "Schulz iteration did not reach the stop tolerance in every member");
}

} // namespace dmhm

extern "C" {

int dmhm_hbatch_create(int count, const dmhm_hmat* mats, dmhm_hbatch* out)
{
    return guard([&] {
        require_ready();
        if (count < 1 || !mats || !out) throw Error(DMHM_EINVAL, "a batch needs members");
        std::vector<dmhm_hmat_s*> like;
        std::vector<std::shared_ptr<Layout>> lays;
        for (int i = 0; i < count; ++i) {
            if (!mats[i]) throw Error(DMHM_EINVAL, "null member");
            if (mats[i]->prec != mats[0]->prec) throw Error(DMHM_EINVAL, "members must share a precision");
            like.push_back(mats[i]);
            lays.push_back(mats[i]->lay);
        }
        std::unique_ptr<dmhm_hbatch_s> b(new dmhm_hbatch_s());
        arena_alloc(b->a, like, lays, mats[0]->prec);
        for (int i = 0; i < count; ++i) copy_member(b->a.at(i), mats[i]);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
        *out = b.release();
    });
}

int dmhm_hbatch_destroy(dmhm_hbatch b)
{
    return guard([&] { delete b; });
}

int dmhm_hbatch_count(dmhm_hbatch b, int* count)
{
    return guard([&] {
        if (!b || !count) throw Error(DMHM_EINVAL, "null pointer");
        *count = (int)b->a.size();
    });
}

int dmhm_hbatch_member(dmhm_hbatch b, int index, dmhm_hmat* out)
{
    return guard([&] {
        require_ready();
        if (!b || !out) throw Error(DMHM_EINVAL, "null pointer");
        if (index < 0 || index >= (int)b->a.size()) throw Error(DMHM_EINVAL, "member index out of range");
        const dmhm_hmat_s* v = b->a.at(index);
        std::unique_ptr<dmhm_hmat_s> h(alloc_hmat(v->gc, v->lay, v->prec, v->maxRank));
        h->stats = v->stats;
        copy_member(h.get(), v);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
        *out = h.release();
    });
}

int dmhm_hbatch_multiply_hmat(double alpha, dmhm_hbatch A, dmhm_hbatch B, const dmhm_hh_opts* opts, dmhm_hbatch* C)
{
    return guard([&] {
        require_ready();
        if (!A || !B || !C) throw Error(DMHM_EINVAL, "null pointer");
        if (A->a.size() != B->a.size()) throw Error(DMHM_ENONCONF, "batches of different sizes");
        dmhm_hh_opts o{4, 0, 1e-10, 1, nullptr, 0.0};
        if (opts) o = *opts;
        std::unique_ptr<dmhm_hbatch_s> c(new dmhm_hbatch_s());
        arena_alloc(c->a, members(A->a), std_layouts(A->a), A->a.prec);
        for (size_t i = 0; i < c->a.size(); ++i)
            c->a.at(i)->stats = A->a.at(i)->stats ? A->a.at(i)->stats : B->a.at(i)->stats;
        batch_multiply_into(*A, alpha, A->a, B->a, c->a, o);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
        *C = c.release();
    });
}

int dmhm_hbatch_estimate_two_norm(dmhm_hbatch b, double theta, double confidence, uint64_t seed, double* estimates)
{
    return guard([&] {
        require_ready();
        if (!b || !estimates) throw Error(DMHM_EINVAL, "null pointer");
        if (theta <= 1) throw Error(DMHM_EINVAL, "Theta must be > 1");
        if (confidence <= 0) throw Error(DMHM_EINVAL, "Confidence must be positive");
        ensure_vectors(*b, theta, confidence);
        Plan& mv = batch_plan(*b, "mv:" + arena_key(b->a), [&] { return merged_matvec(b->a, b->vecBase, b->vecElems); });
        ensure_ws(mv, b->a.esize());
        if (b->a.prec == DMHM_F64) launch_batch_two_norm<double>(*b, b->a, mv, seed);
        else launch_batch_two_norm<float>(*b, b->a, mv, seed);
        std::vector<double> est;
        read_batch_two_norm(*b, theta, est);
        std::copy(est.begin(), est.end(), estimates);
    });
}

int dmhm_hbatch_schulz_invert(dmhm_hbatch b, const dmhm_schulz_opts* opts, dmhm_schulz_report* reports)
{
    return guard([&] {
        require_ready();
        if (!b) throw Error(DMHM_EINVAL, "null handle");
        dmhm_schulz_opts o{};
        if (opts) o = *opts;
        else { o.numIterations = -1; o.maxIterations = 30; o.theta = 1.5; o.confidence = 6; o.stopTol = 1e-4;
               o.hh = {4, 0, 1e-10, 1, nullptr, 0.0}; }
        batch_schulz(*b, o, reports);
    });
}

} // extern "C"
