// Host runtime and C-ABI of libdmhm_amd.so: device storage of H-matrices,
// plan caches, and the operations of include/dmhm_amd.h.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <random>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "runtime.hpp"

namespace dmhm {

static thread_local std::string g_lastError;

Library g_lib;

// Copies on the library stream and waits for them: the stream is created
// non-blocking, so a plain hipMemcpy (null stream) would not be ordered after
// work already queued on it (for example the zero fill of a new matrix).
hipError_t copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, g_lib.stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(g_lib.stream);
}

// Uploads the block tree of a geometry as the flat int arrays of GeoDev (one
// allocation), once.
const GeoDev& geo_dev(GeoCtx& gc)
{
    if (gc.devReady) return gc.dev;
    const Geometry& g = gc.geo;
    const int nc = g.nc();
    std::vector<int> lvlBase, boxOff, boxSize, boxChild;
    int base = 0;
    for (int l = 0; l < g.numLevels; ++l) {
        lvlBase.push_back(base);
        base += (int)g.boxes[l].size();
    }
    for (int l = 0; l < g.numLevels; ++l)
        for (const Box& B : g.boxes[l]) {
            boxOff.push_back(B.off);
            boxSize.push_back(B.size);
            for (int c = 0; c < nc; ++c)
                boxChild.push_back(l + 1 < g.numLevels ? lvlBase[l + 1] + B.child[c] : -1);
        }
    const size_t nb = g.blocks.size();
    std::vector<int> type(nb), level(nb), bt(nb), bs(nb), cb(nb), lr(nb), dn(nb);
    for (size_t i = 0; i < nb; ++i) {
        const Block& B = g.blocks[i];
        type[i] = B.type; level[i] = B.level; bt[i] = B.t; bs[i] = B.s;
        cb[i] = B.childBegin; lr[i] = B.lrIndex; dn[i] = B.denseIndex;
    }
    std::vector<const std::vector<int>*> parts = {&lvlBase, &boxOff, &boxSize, &boxChild, &type, &level, &bt,
                                                  &bs, &cb, &lr, &dn, &g.children, &g.lrBlocks};
    size_t total = 0;
    for (auto* v : parts) total += std::max<size_t>(v->size(), 1);
    std::vector<int> flat;
    flat.reserve(total);
    std::vector<size_t> offs;
    for (auto* v : parts) {
        offs.push_back(flat.size());
        flat.insert(flat.end(), v->begin(), v->end());
        if (v->empty()) flat.push_back(0);
    }
    gc.devTree.ensure(flat.size() * sizeof(int));
    HIPCHECK(copy_sync(gc.devTree.ptr, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice));
    const int* d = static_cast<const int*>(gc.devTree.ptr);
    GeoDev& G = gc.dev;
    G.dim = g.dim; G.nc = nc; G.numLevels = g.numLevels; G.n = g.n;
    G.lvlBase = d + offs[0]; G.boxOff = d + offs[1]; G.boxSize = d + offs[2]; G.boxChild = d + offs[3];
    G.blkType = d + offs[4]; G.blkLevel = d + offs[5]; G.blkT = d + offs[6]; G.blkS = d + offs[7];
    G.blkChildBegin = d + offs[8]; G.blkLr = d + offs[9]; G.blkDense = d + offs[10]; G.children = d + offs[11];
    G.lrBlock = d + offs[12];
    gc.devReady = true;
    return G;
}

std::shared_ptr<GeoCtx> get_geo(int dim, const int* sizes, int L, bool strong)
{
    std::lock_guard<std::mutex> lk(g_lib.mu);
    for (auto& w : g_lib.geos) {
        auto g = w.lock();
        if (!g) continue;
        const Geometry& G = g->geo;
        bool same = G.dim == dim && G.numLevels == L && G.strong == strong;
        for (int a = 0; a < dim && same; ++a) same = G.sizes[a] == sizes[a];
        if (same) return g;
    }
    auto g = std::make_shared<GeoCtx>(dim, sizes, L, strong);
    g_lib.geos.push_back(g);
    return g;
}

std::shared_ptr<Layout> std_layout(GeoCtx& gc, int maxRank)
{
    auto it = gc.stdLayouts.find(maxRank);
    if (it != gc.stdLayouts.end()) return it->second;
    std::vector<int> caps(gc.geo.lrBlocks.size(), std::max(maxRank, 1));
    auto L = std::make_shared<Layout>(Layout::make(gc.geo, caps, maxRank + 1));
    gc.stdLayouts[maxRank] = L;
    return L;
}

} // namespace dmhm

using namespace dmhm;


// Rank model: the perceptron's parameters and Adam state on the device.
struct dmhm_rank_model_s {
    int nf = 0, np = 0, step = 0;
    dmhm::DeviceBuffer params, grads, m, v;
    ~dmhm_rank_model_s() { params.release(); grads.release(); m.release(); v.release(); }
};

namespace dmhm {

// Features per admissible block: kStaticFeatures from the geometry and the
// operator's coefficient statistics, then the previous rank.
constexpr int kStaticFeatures = 11;
constexpr int kRankFeatures = kStaticFeatures + 1;

// What the last width-indirect product did (dmhm_last_product_stats).
struct ProductStats { long long blocks = 0, widthSum = 0, fallbacks = 0, fullWidth = 0; };
static ProductStats g_lastProduct;

dmhm_hmat_s* alloc_hmat(std::shared_ptr<GeoCtx> gc, std::shared_ptr<Layout> lay, int prec,
                               int maxRank)
{
    std::unique_ptr<dmhm_hmat_s> h(new dmhm_hmat_s());
    h->gc = gc; h->lay = lay; h->prec = prec; h->maxRank = maxRank;
    const size_t bytes = std::max<size_t>(lay->totalElems() * h->esize(), 256);
    HIPCHECK(hipMalloc(&h->vals, bytes));
    HIPCHECK(hipMemsetAsync(h->vals, 0, bytes, g_lib.stream));
    const size_t rb = std::max<size_t>(lay->nLowRank() * sizeof(int), 4);
    HIPCHECK(hipMalloc((void**)&h->ranks, rb));
    HIPCHECK(hipMemsetAsync(h->ranks, 0, rb, g_lib.stream));
    return h.release();
}

static Bases bases_for(dmhm_hmat_s* A, dmhm_hmat_s* B, dmhm_hmat_s* C, void* X, void* Y)
{
    Bases b{};
    if (A) { b.val[SLOT_A] = A->vals; b.rank[SLOT_A] = A->ranks; }
    if (B) { b.val[SLOT_B] = B->vals; b.rank[SLOT_B] = B->ranks; }
    if (C) { b.val[SLOT_C] = C->vals; b.rank[SLOT_C] = C->ranks; }
    b.val[SLOT_WS] = g_lib.ws.ptr;
    b.val[SLOT_X] = X; b.val[SLOT_Y] = Y;
    return b;
}

// Tile-engine candidates per execution slot (FP64, FP32, FP32 with BF16
// operands allowed); FP32 engines stay candidates in the BF16 slot, since a
// more precise engine that is also faster is the better choice there.
// (The prefetching MFMA form is not a candidate in FP64: it produced wrong
// products there on gfx950 while the same source is exact in FP32.)
static const int kTuneCands[3][5] = {{TILE_LDS, TILE_MFMA, -1, -1, -1},
                                     {TILE_LDS, TILE_MFMA, TILE_MFMA_PF, -1, -1},
                                     {TILE_LDS, TILE_MFMA, TILE_MFMA_PF, TILE_MFMA_BF16, TILE_MFMA_BF16_PF}};
static const int kTuneCount[3] = {2, 3, 5};

// Launches a TILE phase with its engine. Autotuning (no engine forced): the
// first run of a phase in a slot is a warm-up with the first candidate (its
// operands' first touch and cold caches would otherwise be charged to that
// candidate alone), the next runs time one candidate each (a complete,
// correct execution every time, synchronised for the timing), after which the
// fastest candidate is kept for that slot.
template <typename T>
static void run_tile_phase(Phase& ph, const Bases& b, T alpha, T beta)
{
    hipStream_t s = g_lib.stream;
    const int slot = sizeof(T) == 8 ? 0 : (g_lib.bf16 ? 2 : 1);
    if (g_lib.tileEngine >= 0) {
        int e = g_lib.tileEngine;
        if ((e == TILE_MFMA_BF16 || e == TILE_MFMA_BF16_PF) && slot != 2) e = e == TILE_MFMA_BF16 ? TILE_MFMA : TILE_MFMA_PF;
        launch_tile_phase<T>(ph, b, alpha, beta, s, e);
        return;
    }
    // Segments of one source phase tune in that phase, timed per segment.
    const Phase* tp = &ph;
    if (!ph.segSrc.empty()) {
        tp = ph.segSrc[0].src;
        for (const auto& sg : ph.segSrc)
            if (sg.src != tp) { tp = &ph; break; }
    }
    const float per = ph.segSrc.empty() || tp == &ph ? 1.0f : (float)ph.segSrc.size();
    Phase::Tune& t = tp->tune[slot];
    if (t.choice >= 0) {
        launch_tile_phase<T>(ph, b, alpha, beta, s, t.choice);
        return;
    }
    if (t.next < 0) {
        launch_tile_phase<T>(ph, b, alpha, beta, s, kTuneCands[slot][0]);
        ++g_lib.untuned;
        t.next = 0;
        return;
    }
    const int cand = kTuneCands[slot][t.next];
    EventPair& ev = *g_lib.tuneEv;
    HIPCHECK(hipEventRecord(ev.a, s));
    launch_tile_phase<T>(ph, b, alpha, beta, s, cand);
    HIPCHECK(hipEventRecord(ev.b, s));
    HIPCHECK(hipEventSynchronize(ev.b));
    t.ms[t.next] = ev.ms() / per;
    ++g_lib.untuned;
    if (++t.next == kTuneCount[slot]) {
        int best = 0;
        for (int i = 1; i < kTuneCount[slot]; ++i)
            if (t.ms[i] < t.ms[best]) best = i;
        t.choice = kTuneCands[slot][best];
    }
}

template <typename T>
void run_phase(Phase& ph, const Bases& b, T alpha, T beta, TruncPolicy pol, uint64_t seed)
{
    hipStream_t s = g_lib.stream;
    switch (ph.kind) {
    case PH_TILE: run_tile_phase<T>(ph, b, alpha, beta); break;
    case PH_QR: launch_qr_phase<T>(ph, b, s); break;
    case PH_SVD: launch_svd_phase<T>(ph, b, pol, s); break;
    case PH_RNG: launch_rng_phase<T>(ph, b, seed, s, g_lib.seedOff); break;
    case PH_EYE:
        if (!ph.eyes.empty())
            launch_eye_multi<T>(reinterpret_cast<T*>(b.val[SLOT_WS]), static_cast<const int64_t*>(ph.dTasks),
                                (int)ph.eyes.size(), (int)ph.zeroCount, s);
        else
            launch_eye<T>(reinterpret_cast<T*>(b.val[SLOT_WS]) + ph.zeroOff, (int)ph.zeroCount, s);
        break;
    case PH_MV_ZERO: case PH_MV_PROJ: case PH_MV_ACC: launch_mv_phase<T>(ph, b, alpha, beta, s); break;
    default: break;
    }
}

template <typename T>
void run_plan(Plan& p, const Bases& b, T alpha, T beta, TruncPolicy pol, uint64_t seed)
{
    for (auto& ph : p.phases) run_phase<T>(ph, b, alpha, beta, pol, seed);
}
template void run_phase<double>(Phase&, const Bases&, double, double, TruncPolicy, uint64_t);
template void run_phase<float>(Phase&, const Bases&, float, float, TruncPolicy, uint64_t);
template void run_plan<double>(Plan&, const Bases&, double, double, TruncPolicy, uint64_t);
template void run_plan<float>(Plan&, const Bases&, float, float, TruncPolicy, uint64_t);

// Workspace budget (elements) of one H*H batch: 2^33 (64 GiB in FP64) by
// default, which keeps the 256^2 rank-16 product in 2-3 batches (103 ms FP64,
// against 116 ms with 2^31 and 185 ms with 2^29); DMHM_HH_BATCH_ELEMS
// overrides it (plans built before a change keep theirs).
int64_t hh_batch_budget()
{
    static const int64_t v = [] {
        const char* e = getenv("DMHM_HH_BATCH_ELEMS");
        const long long x = e ? atoll(e) : 0;
        return x > 0 ? (int64_t)x : (int64_t(1) << 33);
    }();
    return v;
}

Plan& get_plan(GeoCtx& gc, const std::string& key, const std::function<Plan*()>& make)
{
    auto it = gc.plans.find(key);
    if (it != gc.plans.end()) return *it->second;
    std::unique_ptr<Plan> p(make());
    p->upload();
    Plan& ref = *p;
    gc.plans[key] = std::move(p);
    return ref;
}

void ensure_ws(const Plan& p, size_t esize)
{
    g_lib.ws.ensure(std::max<size_t>(p.wsElems * esize, 256));
}

// ---------------------------------------------------------------------------
// Operations
// ---------------------------------------------------------------------------
static void multiply_dense(dmhm_hmat_s* h, bool trans, double alpha, const void* X, int ldx, int k,
                           double beta, void* Y, int ldy)
{
    GeoCtx& gc = *h->gc;
    const bool stream = !trans && k <= 8;
    const std::string key = std::string(stream ? "mvs:" : "mv:") + std::to_string(h->lay->id) + ":" +
                            std::to_string(trans) + ":" + std::to_string(k) + ":" + std::to_string(ldx) + ":" +
                            std::to_string(ldy);
    Plan& p = get_plan(gc, key, [&] {
        Planner pl(gc.geo);
        return stream ? pl.matvecStream(*h->lay, k, ldx, ldy) : pl.matvec(*h->lay, trans, k, ldx, ldy);
    });
    ensure_ws(p, h->esize());
    Bases b = bases_for(h, nullptr, nullptr, const_cast<void*>(X), Y);
    if (h->prec == DMHM_F64) run_plan<double>(p, b, alpha, beta, {0, 0}, 0);
    else run_plan<float>(p, b, (float)alpha, (float)beta, {0, 0}, 0);
}

// H*X (k <= 8, device operands) on the streaming plan with the projections
// in the device scratch array: no device allocation, no host staging. The
// plan's index tables must exist already (built by a multiply_dense call of
// the same shape), since building them allocates device memory.
static void multiply_dense_noalloc(dmhm_hmat_s* h, double alpha, const void* X, int ldx, int k, double beta,
                                   void* Y, int ldy)
{
    if (k < 1 || k > 8) throw Error(DMHM_EINVAL, "the allocation-free H*X takes 1 to 8 right-hand sides");
    GeoCtx& gc = *h->gc;
    const std::string key = "mvs:" + std::to_string(h->lay->id) + ":0:" + std::to_string(k) + ":" +
                            std::to_string(ldx) + ":" + std::to_string(ldy);
    auto it = gc.plans.find(key);
    if (it == gc.plans.end())
        throw Error(DMHM_EINVAL, "no prepared plan for this shape: call dmhm_hmat_multiply_dense once with op 'N' "
                                 "and the same k, ldx and ldy first (that call may allocate)");
    Plan& p = *it->second;
    size_t cap = 0;
    void* ws = scratch_address(g_lib.device, &cap);
    if (!ws) throw Error(DMHM_ENODEV, "device scratch is not set up (dmhm_init)");
    const size_t need = (size_t)p.wsElems * h->esize();
    if (need > cap)
        throw Error(DMHM_EINVAL, "workspace of " + std::to_string(need) + " bytes exceeds the scratch maximum of " +
                                 std::to_string(cap) + " bytes");
    Bases b = bases_for(h, nullptr, nullptr, const_cast<void*>(X), Y);
    b.val[SLOT_WS] = ws;
    if (h->prec == DMHM_F64) run_plan<double>(p, b, alpha, beta, {0, 0}, 0);
    else run_plan<float>(p, b, (float)alpha, (float)beta, {0, 0}, 0);
}

// Static features of every admissible block of a geometry (kStaticFeatures
// floats each, on the device, built once per operator statistics): level;
// log sizes of the target and source boxes; log of their distance in boxes;
// the signed box offsets per axis over the boxes per axis; mean (relative to
// the operator's) and spread of log|a_ii| over the target and source boxes.
static const float* static_features(GeoCtx& gc, std::shared_ptr<OperatorStats> st)
{
    const Geometry& g = gc.geo;
    if (!st) {
        if (!gc.zeroStats) {
            gc.zeroStats = std::make_shared<OperatorStats>();
            int base = 0;
            for (int l = 0; l < g.numLevels; ++l) {
                gc.zeroStats->levelBase.push_back(base);
                base += (int)g.boxes[l].size();
            }
            gc.zeroStats->mean.assign(base, 0.0f);
            gc.zeroStats->sdev.assign(base, 0.0f);
        }
        st = gc.zeroStats;
    }
    if (st->featReady) return static_cast<const float*>(st->feat.ptr);
    const int nl = (int)g.lrBlocks.size();
    std::vector<float> f((size_t)std::max(nl, 1) * kStaticFeatures, 0.0f);
    const double logn = std::log2((double)std::max(2, g.n));
    for (int i = 0; i < nl; ++i) {
        const Block& B = g.blocks[g.lrBlocks[i]];
        const Box& tb = g.boxes[B.level][B.t];
        const Box& sb = g.boxes[B.level][B.s];
        float* x = f.data() + (size_t)i * kStaticFeatures;
        int dist = 1;
        for (int a = 0; a < g.dim; ++a) dist = std::max(dist, std::abs(tb.coord[a] - sb.coord[a]));
        const double nbox = (double)(1 << B.level);
        x[0] = (float)B.level / (float)std::max(1, g.numLevels - 1);
        x[1] = (float)(std::log2((double)std::max(1, tb.size)) / logn);
        x[2] = (float)(std::log2((double)std::max(1, sb.size)) / logn);
        x[3] = (float)(std::log2((double)dist) / std::max(1, B.level));
        x[4] = (float)((tb.coord[0] - sb.coord[0]) / nbox);
        x[5] = (float)((tb.coord[1] - sb.coord[1]) / nbox);
        x[6] = g.dim == 3 ? (float)((tb.coord[2] - sb.coord[2]) / nbox) : 0.0f;
        const int it = st->levelBase[B.level] + B.t, is = st->levelBase[B.level] + B.s;
        x[7] = st->mean[it] - st->globalMean;
        x[8] = st->sdev[it];
        x[9] = st->mean[is] - st->globalMean;
        x[10] = st->sdev[is];
    }
    st->feat.ensure(f.size() * sizeof(float));
    HIPCHECK(copy_sync(st->feat.ptr, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    st->featReady = true;
    return static_cast<const float*>(st->feat.ptr);
}

// Features of the blocks of a product with left operand A into `out` (device,
// nLowRank x kRankFeatures): the static features of A's operator and the rank
// of each block in A (the "previous rank": in a Schulz step the left operand
// is the current iterate, or Z = I - X A).
static void rank_features(dmhm_hmat_s* A, float* out)
{
    const float* st = static_features(*A->gc, A->stats);
    launch_rank_features(st, kStaticFeatures, A->ranks, 1.0f / std::max(1, A->maxRank), out,
                         (int)A->geo().lrBlocks.size(), g_lib.stream);
}

// C := alpha A B into an existing C of the same geometry and precision (its
// layout is the plan's output layout). The plan writes every dense entry, every
// U/V column and every rank of C, so C's previous contents do not matter.
// With a rank model in tolerance mode the sketches are width-indirect: the
// model predicts every block's rank from its features, the sketch takes that
// plus the oversampling (at most maxRank + oversample columns), and blocks
// whose SVD flags the sketch as too narrow are recomputed at full width.
static void multiply_into(double alpha, dmhm_hmat_s* A, dmhm_hmat_s* B, dmhm_hmat_s* C, const dmhm_hh_opts& o)
{
    if (A->gc != B->gc || A->gc != C->gc) throw Error(DMHM_ENONCONF, "Attempted nonconformal matrix-matrix multiply");
    if (A->prec != B->prec || A->prec != C->prec) throw Error(DMHM_EINVAL, "operands must share a precision");
    GeoCtx& gc = *A->gc;
    const int maxRank = C->maxRank;
    const Layout& LC = *C->lay;
    const int over = o.oversample > 0 ? o.oversample : 4;
    dmhm_rank_model_s* rm = o.rankModel;
    const bool learned = rm != nullptr && o.tolMode == 1;
    const std::string key = "hh:" + std::to_string(A->lay->id) + ":" + std::to_string(B->lay->id) + ":" +
                            std::to_string(LC.id) + ":" + std::to_string(maxRank) + ":" +
                            std::to_string(over) + ":" + std::to_string(A->full) + std::to_string(B->full) +
                            (learned ? ":w" : "");
    Plan& p = get_plan(gc, key, [&] {
        Planner pl(gc.geo);
        pl.setWorkspaceBudget(hh_batch_budget());
        pl.setStacking(A->full, B->full);
        pl.setWidthIndirect(learned);
        return pl.hhProduct(*A->lay, *B->lay, LC, maxRank, over);
    });
    ensure_ws(p, A->esize());
    Bases b = bases_for(A, B, C, nullptr, nullptr);
    TruncPolicy pol{o.tolMode, o.relTol};
    auto run = [&] {
        if (A->prec == DMHM_F64) run_plan<double>(p, b, alpha, 0.0, pol, o.seed);
        else run_plan<float>(p, b, (float)alpha, 0.0f, pol, o.seed);
    };
    if (!learned) {
        run();
        C->full = true;
        return;
    }
    const int nl = LC.nLowRank();
    const int kS = maxRank + over;
    g_lib.featBuf.ensure((size_t)std::max(nl, 1) * kRankFeatures * sizeof(float));
    g_lib.widthBuf.ensure((size_t)std::max(nl, 1) * sizeof(int));
    g_lib.flagBuf.ensure((size_t)std::max(nl, 1) * sizeof(int));
    float* feat = static_cast<float*>(g_lib.featBuf.ptr);
    int* widths = static_cast<int*>(g_lib.widthBuf.ptr);
    int* flags = static_cast<int*>(g_lib.flagBuf.ptr);
    rank_features(A, feat);
    launch_mlp_infer(static_cast<const float*>(rm->params.ptr), rm->nf, feat, nl, nullptr, widths,
                     (float)o.rankMargin, over, over + 1, kS, g_lib.stream);
    launch_fill_int(flags, nl, 0, g_lib.stream);
    b.rank[SLOT_WIDTH] = widths;
    b.rank[SLOT_FLAG] = flags;
    run();
    // A-posteriori check: widen the flagged blocks to the full sketch and
    // recompute (the other blocks come out identical: same widths and seeds).
    int* cnt = reinterpret_cast<int*>(g_lib.dScalars + 4);
    long long* wsum = reinterpret_cast<long long*>(g_lib.dScalars + 5);
    launch_zero_double(g_lib.dScalars + 4, 2, g_lib.stream);
    launch_widen_flagged(widths, flags, nl, kS, cnt, wsum, g_lib.stream);
    struct { int cnt; int pad; long long wsum; } host{};
    HIPCHECK(copy_sync(&host, g_lib.dScalars + 4, 16, hipMemcpyDeviceToHost));
    if (host.cnt > 0) run();
    g_lastProduct.blocks = nl;
    g_lastProduct.widthSum = host.wsum;
    g_lastProduct.fallbacks = host.cnt;
    g_lastProduct.fullWidth = (long long)nl * kS;
    C->full = true;
}

static dmhm_hmat_s* multiply_hmat(double alpha, dmhm_hmat_s* A, dmhm_hmat_s* B, const dmhm_hh_opts& o)
{
    if (A->gc != B->gc) throw Error(DMHM_ENONCONF, "Attempted nonconformal matrix-matrix multiply");
    if (A->prec != B->prec) throw Error(DMHM_EINVAL, "operands must share a precision");
    std::unique_ptr<dmhm_hmat_s> C(alloc_hmat(A->gc, std_layout(*A->gc, A->maxRank), A->prec, A->maxRank));
    C->stats = A->stats ? A->stats : B->stats;
    multiply_into(alpha, A, B, C.get(), o);
    return C.release();
}


static void multiply_hmat_update(double alpha, dmhm_hmat_s* A, dmhm_hmat_s* B, double beta, dmhm_hmat_s* C,
                                 const dmhm_hh_opts& o)
{
    if (A->gc != B->gc || A->gc != C->gc) throw Error(DMHM_ENONCONF, "Attempted nonconformal matrix-matrix multiply");
    if (A->prec != B->prec || A->prec != C->prec) throw Error(DMHM_EINVAL, "operands must share a precision");
    GeoCtx& gc = *A->gc;
    const int maxRank = C->maxRank;
    for (int c : C->lay->cap)
        if (c > 64) throw Error(DMHM_EINVAL, "low-rank capacity above 64 in a product");
    std::unique_ptr<dmhm_hmat_s> Cold(copy_hmat(C, C->prec));
    const int over = o.oversample > 0 ? o.oversample : 4;
    const std::string key = "hhu:" + std::to_string(A->lay->id) + ":" + std::to_string(B->lay->id) + ":" +
                            std::to_string(C->lay->id) + ":" + std::to_string(maxRank) + ":" + std::to_string(over) +
                            ":" + std::to_string(A->full) + std::to_string(B->full);
    Plan& p = get_plan(gc, key, [&] {
        Planner pl(gc.geo);
        pl.setWorkspaceBudget(hh_batch_budget());
        pl.setStacking(A->full, B->full);
        return pl.hhProduct(*A->lay, *B->lay, *C->lay, maxRank, over, Planner::HH_UPDATE);
    });
    ensure_ws(p, A->esize());
    Bases b = bases_for(A, B, C, Cold->vals, nullptr);
    b.rank[SLOT_X] = Cold->ranks;
    TruncPolicy pol{o.tolMode, o.relTol};
    if (A->prec == DMHM_F64) run_plan<double>(p, b, alpha, beta, pol, o.seed);
    else run_plan<float>(p, b, (float)alpha, (float)beta, pol, o.seed);
    HIPCHECK(hipStreamSynchronize(g_lib.stream));
    C->full = true;
}

void update_with(dmhm_hmat_s* A, double alpha, dmhm_hmat_s* B, const dmhm_hh_opts& o)
{
    if (A->gc != B->gc) throw Error(DMHM_ENONCONF, "Attempted nonconformal update");
    if (A->prec != B->prec) throw Error(DMHM_EINVAL, "operands must share a precision");
    GeoCtx& gc = *A->gc;
    const int maxRank = A->maxRank;
    auto LC = std_layout(gc, maxRank);
    std::unique_ptr<dmhm_hmat_s> C(alloc_hmat(A->gc, LC, A->prec, maxRank));
    const int over = o.oversample > 0 ? o.oversample : 4;
    const std::string key = "hha:" + std::to_string(A->lay->id) + ":" + std::to_string(B->lay->id) + ":" +
                            std::to_string(LC->id) + ":" + std::to_string(maxRank) + ":" + std::to_string(over);
    Plan& p = get_plan(gc, key, [&] { Planner pl(gc.geo); return pl.hhAdd(*A->lay, *B->lay, *LC, maxRank, over); });
    ensure_ws(p, A->esize());
    Bases b = bases_for(A, B, C.get(), nullptr, nullptr);
    TruncPolicy pol{o.tolMode, o.relTol};
    const int64_t nd = A->lay->denseElems;
    if (A->prec == DMHM_F64) {
        run_plan<double>(p, b, alpha, 0.0, pol, o.seed);
        launch_axpby<double>((const double*)A->vals, (const double*)B->vals, alpha, (double*)C->vals, nd, g_lib.stream);
    } else {
        run_plan<float>(p, b, (float)alpha, 0.0f, pol, o.seed);
        launch_axpby<float>((const float*)A->vals, (const float*)B->vals, (float)alpha, (float*)C->vals, nd, g_lib.stream);
    }
    HIPCHECK(hipStreamSynchronize(g_lib.stream));
    A->full = true;
    std::swap(A->lay, C->lay);
    std::swap(A->vals, C->vals);
    std::swap(A->ranks, C->ranks);
}

// One-off device runs of the QR / SVD / tile kernels on host data (FP64).
struct OneOff {
    Plan plan;
    DeviceBuffer ws;
    int* rank = nullptr;
    ~OneOff() { ws.release(); if (rank) (void)hipFree(rank); }
    void run() {
        plan.upload();
        Bases b{};
        b.val[SLOT_WS] = ws.ptr;
        b.rank[SLOT_C] = rank;
        run_plan<double>(plan, b, 1.0, 0.0, {0, 0}, 0);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    }
};

// Tiles computing out (rows x k, ld rows) = L (rows x q, ld ldl) * R (q x k, ld ldr), all in the workspace.
static void add_gemm_tiles(Phase& ph, int64_t out, int rows, int k, int64_t L, int ldl, int q, int64_t R, int ldr,
                           uint8_t rflags)
{
    PhaseBuilder pb(ph);
    for (int r0 = 0; r0 < rows; r0 += 64)
        for (int c0 = 0; c0 < k; c0 += 64) {
            Tile t{};
            t.off = out + r0 + (int64_t)c0 * rows; t.ld = rows;
            t.m = std::min(64, rows - r0); t.k = std::min(64, k - c0);
            t.mrank = t.krank = -1; t.slot = SLOT_WS;
            const int id = pb.addTile(t);
            Contrib c{};
            c.loff = L + r0; c.ldl = ldl; c.q = q; c.qrank = -1;
            c.roff = (rflags & CF_TRANS_R) ? R + c0 : R + (int64_t)c0 * ldr; c.ldr = ldr;
            c.lslot = SLOT_WS; c.rslot = SLOT_WS; c.flags = rflags; c.alpha = 1.0f;
            pb.add(id, c);
        }
    pb.finalize();
}

static void rounded_update_dev(int maxRank, double alpha, int m, int n, int rA, const double* AU, const double* AV,
                               double beta, int rB, const double* BU, const double* BV, double* outU, double* outV,
                               int* outRank)
{
    const int r = rA + rB;
    const int cap = std::min(std::min(r, std::min(m, n)), maxRank);
    if (cap == r) {
        // Early exit: exact concatenation [beta BU, alpha AU], [BV, AV] (RoundedUpdate.cpp:34-53).
        for (int j = 0; j < rB; ++j) {
            for (int i = 0; i < m; ++i) outU[i + (size_t)j * m] = beta * BU[i + (size_t)j * m];
            for (int i = 0; i < n; ++i) outV[i + (size_t)j * n] = BV[i + (size_t)j * n];
        }
        for (int j = 0; j < rA; ++j) {
            for (int i = 0; i < m; ++i) outU[i + (size_t)(rB + j) * m] = alpha * AU[i + (size_t)j * m];
            for (int i = 0; i < n; ++i) outV[i + (size_t)(rB + j) * n] = AV[i + (size_t)j * n];
        }
        *outRank = r;
        return;
    }
    if (r > 64) throw Error(DMHM_EINVAL, "rounded update supports rA + rB <= 64");
    // Panels with fewer rows than columns are padded with zero rows so the
    // panel QR sees at least r rows; the padded rows of the result are dropped.
    const int mp = std::max(m, r), np = std::max(n, r);
    OneOff job;
    int64_t o = 0;
    auto take = [&](int64_t e) { const int64_t x = o; o += (e + 63) & ~int64_t(63); return x; };
    const int64_t Uc = take((int64_t)mp * r), Vc = take((int64_t)np * r), R1 = take(r * r), R2 = take(r * r),
                  Wm = take(r * r), G1 = take(r * r), G2 = take(r * r), oU = take((int64_t)mp * r),
                  oV = take((int64_t)np * r);
    std::vector<double> host(o, 0.0);
    for (int j = 0; j < rA; ++j) {
        for (int i = 0; i < m; ++i) host[Uc + i + (size_t)j * mp] = alpha * AU[i + (size_t)j * m];
        for (int i = 0; i < n; ++i) host[Vc + i + (size_t)j * np] = AV[i + (size_t)j * n];
    }
    for (int j = 0; j < rB; ++j) {
        for (int i = 0; i < m; ++i) host[Uc + i + (size_t)(rA + j) * mp] = beta * BU[i + (size_t)j * m];
        for (int i = 0; i < n; ++i) host[Vc + i + (size_t)(rA + j) * np] = BV[i + (size_t)j * n];
    }
    job.ws.ensure(o * sizeof(double));
    HIPCHECK(copy_sync(job.ws.ptr, host.data(), o * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMalloc((void**)&job.rank, sizeof(int)));
    job.plan.phases.resize(5);
    Phase& q = job.plan.phases[0];
    q.kind = PH_QR;
    q.qr.push_back(QrTask{Uc, R1, mp, r, mp, r, SLOT_WS});
    q.qr.push_back(QrTask{Vc, R2, np, r, np, r, SLOT_WS});
    job.plan.phases[1].kind = PH_TILE;                        // W = R1 R2^T
    add_gemm_tiles(job.plan.phases[1], Wm, r, r, R1, r, r, R2, r, CF_TRANS_R);
    Phase& sv = job.plan.phases[2];
    sv.kind = PH_SVD;
    sv.svd.push_back(SvdTask{Wm, G1, G2, r, r, 0, cap});      // W = U S V^T: G1 = V S, G2 = U
    job.plan.phases[3].kind = PH_TILE;                        // out U = Q1 U_W
    add_gemm_tiles(job.plan.phases[3], oU, mp, r, Uc, mp, r, G2, r, 0);
    job.plan.phases[4].kind = PH_TILE;                        // out V = Q2 V_W S
    add_gemm_tiles(job.plan.phases[4], oV, np, r, Vc, np, r, G1, r, 0);
    job.run();
    int keep = 0;
    HIPCHECK(copy_sync(&keep, job.rank, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHECK(copy_sync(host.data(), job.ws.ptr, o * sizeof(double), hipMemcpyDeviceToHost));
    for (int j = 0; j < keep; ++j) {
        for (int i = 0; i < m; ++i) outU[i + (size_t)j * m] = host[oU + i + (size_t)j * mp];
        for (int i = 0; i < n; ++i) outV[i + (size_t)j * n] = host[oV + i + (size_t)j * np];
    }
    *outRank = keep;
}

static void compress_dense_dev(int maxRank, int m, int n, const double* D, double* outU, double* outV, int* outRank)
{
    if (std::min(m, n) > 64) throw Error(DMHM_EINVAL, "dense compression supports min(m, n) <= 64");
    if (n > 64) {
        // Compress D^T and swap the factors.
        std::vector<double> Dt((size_t)m * n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) Dt[j + (size_t)i * n] = D[i + (size_t)j * m];
        compress_dense_dev(maxRank, n, m, Dt.data(), outV, outU, outRank);
        return;
    }
    const int k = n, cap = std::min(std::min(m, n), maxRank);
    const int mp = std::max(m, k);                            // zero rows pad a wide panel
    OneOff job;
    int64_t o = 0;
    auto take = [&](int64_t e) { const int64_t x = o; o += (e + 63) & ~int64_t(63); return x; };
    const int64_t Dd = take((int64_t)mp * n), R = take(k * k), G1 = take(k * k), G2 = take(k * k),
                  oU = take((int64_t)mp * k);
    std::vector<double> host(o, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) host[Dd + i + (size_t)j * mp] = D[i + (size_t)j * m];
    job.ws.ensure(o * sizeof(double));
    HIPCHECK(copy_sync(job.ws.ptr, host.data(), o * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMalloc((void**)&job.rank, sizeof(int)));
    job.plan.phases.resize(3);
    job.plan.phases[0].kind = PH_QR;
    job.plan.phases[0].qr.push_back(QrTask{Dd, R, mp, k, mp, k, SLOT_WS});
    job.plan.phases[1].kind = PH_SVD;                         // R = U S V^T: G1 = V S, G2 = U
    job.plan.phases[1].svd.push_back(SvdTask{R, G1, G2, k, k, 0, cap});
    job.plan.phases[2].kind = PH_TILE;                        // out U = Q U_R
    add_gemm_tiles(job.plan.phases[2], oU, mp, k, Dd, mp, k, G2, k, 0);
    job.run();
    int keep = 0;
    HIPCHECK(copy_sync(&keep, job.rank, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHECK(copy_sync(host.data(), job.ws.ptr, o * sizeof(double), hipMemcpyDeviceToHost));
    for (int j = 0; j < keep; ++j) {
        for (int i = 0; i < m; ++i) outU[i + (size_t)j * m] = host[oU + i + (size_t)j * mp];
        for (int i = 0; i < n; ++i) outV[i + (size_t)j * n] = host[G1 + i + (size_t)j * k];
    }
    *outRank = keep;
}

static dmhm_hmat_s* create_from_lowrank(int dim, const int* sizes, int L, int maxRank, bool strong, int prec, int r,
                                        const double* U, const double* V)
{
    auto gc = get_geo(dim, sizes, L, strong);
    const Geometry& g = gc->geo;
    std::vector<int> caps(g.lrBlocks.size(), std::max(std::max(maxRank, 1), r));
    std::shared_ptr<Layout> lay;
    if (r <= std::max(maxRank, 1)) lay = std_layout(*gc, maxRank);
    else {
        std::lock_guard<std::mutex> lk(g_lib.mu);
        lay = std::make_shared<Layout>(Layout::make(g, caps, g_lib.nextLayoutId++));
    }
    std::unique_ptr<dmhm_hmat_s> h(alloc_hmat(gc, lay, prec, maxRank));
    std::vector<double> host(lay->totalElems(), 0.0);
    const int n = g.n;
    for (size_t i = 0; i < g.denseBlocks.size(); ++i) {
        const int b = g.denseBlocks[i];
        const Box& tb = g.tbox(b); const Box& sb = g.sbox(b);
        double* Dp = host.data() + lay->denseOff[i];
        for (int c = 0; c < sb.size; ++c)
            for (int rr = 0; rr < tb.size; ++rr) {
                double sum = 0;
                for (int p = 0; p < r; ++p) sum += U[(tb.off + rr) + (size_t)p * n] * V[(sb.off + c) + (size_t)p * n];
                Dp[rr + (size_t)c * tb.size] = sum;
            }
    }
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        const int b = g.lrBlocks[i];
        const Box& tb = g.tbox(b); const Box& sb = g.sbox(b);
        for (int p = 0; p < r; ++p) {
            for (int rr = 0; rr < tb.size; ++rr) host[lay->uOff[i] + rr + (size_t)p * tb.size] = U[(tb.off + rr) + (size_t)p * n];
            for (int rr = 0; rr < sb.size; ++rr) host[lay->vOff[i] + rr + (size_t)p * sb.size] = V[(sb.off + rr) + (size_t)p * n];
        }
    }
    if (prec == DMHM_F64) HIPCHECK(copy_sync(h->vals, host.data(), host.size() * 8, hipMemcpyHostToDevice));
    else {
        std::vector<float> hf(host.begin(), host.end());
        HIPCHECK(copy_sync(h->vals, hf.data(), hf.size() * 4, hipMemcpyHostToDevice));
    }
    std::vector<int> ranks(g.lrBlocks.size(), r);
    if (!ranks.empty()) HIPCHECK(copy_sync(h->ranks, ranks.data(), ranks.size() * sizeof(int), hipMemcpyHostToDevice));
    h->full = r >= std::max(maxRank, 1);
    return h.release();
}

dmhm_hmat_s* copy_hmat(dmhm_hmat_s* s, int prec)
{
    std::unique_ptr<dmhm_hmat_s> d(alloc_hmat(s->gc, s->lay, prec, s->maxRank));
    const int64_t n = s->lay->totalElems();
    if (prec == s->prec) HIPCHECK(hipMemcpyAsync(d->vals, s->vals, n * s->esize(), hipMemcpyDeviceToDevice, g_lib.stream));
    else if (s->prec == DMHM_F64) launch_convert<double, float>((const double*)s->vals, (float*)d->vals, n, g_lib.stream);
    else launch_convert<float, double>((const float*)s->vals, (double*)d->vals, n, g_lib.stream);
    HIPCHECK(hipMemcpyAsync(d->ranks, s->ranks, s->lay->nLowRank() * sizeof(int), hipMemcpyDeviceToDevice,
                            g_lib.stream));
    d->full = s->full;
    d->stats = s->stats;
    return d.release();
}

template <typename T>
static void upload(std::vector<T>& v, T** d)
{
    HIPCHECK(hipMalloc((void**)d, std::max<size_t>(v.size(), 1) * sizeof(T)));
    if (!v.empty()) HIPCHECK(copy_sync(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

dmhm_hmat_s* transpose_hmat(dmhm_hmat_s* s)
{
    const Geometry& g = s->geo();
    const Layout& Ls = *s->lay;
    // Destination capacities: block b takes its partner's capacity.
    std::vector<int> caps(g.lrBlocks.size());
    bool same = true;
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        const int partner = g.blocks[g.blocks[g.lrBlocks[i]].transposeBlock].lrIndex;
        caps[i] = Ls.cap[partner];
        same = same && caps[i] == Ls.cap[i];
    }
    std::shared_ptr<Layout> Ld = s->lay;
    if (!same) {
        // One transposed layout per source layout, so that plans over
        // transposes (the Schulz start) are cached across calls.
        auto& cached = s->gc->transLayouts[Ls.id];
        if (!cached) {
            std::lock_guard<std::mutex> lk(g_lib.mu);
            cached = std::make_shared<Layout>(Layout::make(g, caps, g_lib.nextLayoutId++));
        }
        Ld = cached;
    }
    std::unique_ptr<dmhm_hmat_s> d(alloc_hmat(s->gc, Ld, s->prec, s->maxRank));
    d->full = s->full;
    d->stats = s->stats;
    std::vector<int64_t> so, dof; std::vector<int> mm, nn;
    for (size_t i = 0; i < g.denseBlocks.size(); ++i) {
        const int b = g.denseBlocks[i];
        const int pb = g.blocks[b].transposeBlock;
        so.push_back(Ls.denseOff[g.blocks[pb].denseIndex]);
        dof.push_back(Ld->denseOff[i]);
        mm.push_back(g.tbox(pb).size); nn.push_back(g.sbox(pb).size);
    }
    std::vector<int64_t> su, sv, du, dv; std::vector<int> mr, nr, cp, perm;
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        const int b = g.lrBlocks[i];
        const int pl = g.blocks[g.blocks[b].transposeBlock].lrIndex;
        su.push_back(Ls.uOff[pl]); sv.push_back(Ls.vOff[pl]);
        du.push_back(Ld->uOff[i]); dv.push_back(Ld->vOff[i]);
        mr.push_back(g.tbox(b).size); nr.push_back(g.sbox(b).size); cp.push_back(caps[i]);
        perm.push_back(pl);
    }
    int64_t *dso, *ddo, *dsu, *dsv, *ddu, *ddv; int *dmm, *dnn, *dmr, *dnr, *dcp, *dperm;
    upload(so, &dso); upload(dof, &ddo); upload(mm, &dmm); upload(nn, &dnn);
    upload(su, &dsu); upload(sv, &dsv); upload(du, &ddu); upload(dv, &ddv);
    upload(mr, &dmr); upload(nr, &dnr); upload(cp, &dcp); upload(perm, &dperm);
    hipStream_t st = g_lib.stream;
    if (s->prec == DMHM_F64) {
        launch_transpose_dense<double>((const double*)s->vals, (double*)d->vals, dso, ddo, dmm, dnn, (int)so.size(), st);
        launch_swap_lowrank<double>((const double*)s->vals, (double*)d->vals, dsu, dsv, ddu, ddv, dmr, dnr, dcp,
                                    (int)su.size(), st);
    } else {
        launch_transpose_dense<float>((const float*)s->vals, (float*)d->vals, dso, ddo, dmm, dnn, (int)so.size(), st);
        launch_swap_lowrank<float>((const float*)s->vals, (float*)d->vals, dsu, dsv, ddu, ddv, dmr, dnr, dcp,
                                   (int)su.size(), st);
    }
    launch_permute_int(s->ranks, d->ranks, dperm, (int)perm.size(), st);
    HIPCHECK(hipStreamSynchronize(st));
    for (void* p : {(void*)dso, (void*)ddo, (void*)dsu, (void*)dsv, (void*)ddu, (void*)ddv, (void*)dmm,
                    (void*)dnn, (void*)dmr, (void*)dnr, (void*)dcp, (void*)dperm})
        (void)hipFree(p);
    return d.release();
}

void scale_hmat(dmhm_hmat_s* h, double alpha)
{
    const Layout& L = *h->lay;
    hipStream_t st = g_lib.stream;
    if (alpha == 0.0) {
        h->full = false;
        // hmat_tools::Scale(0, LowRank) drops the rank (hmat_tools.hpp:1644-1648).
        HIPCHECK(hipMemsetAsync(h->vals, 0, L.totalElems() * h->esize(), st));
        launch_fill_int(h->ranks, L.nLowRank(), 0, st);
        return;
    }
    const int64_t n = L.denseElems + L.uElems;     // dense pool and U pool are contiguous
    if (h->prec == DMHM_F64) launch_scale<double>((double*)h->vals, n, alpha, st);
    else launch_scale<float>((float*)h->vals, n, (float)alpha, st);
}

void add_diag(dmhm_hmat_s* h, double alpha)
{
    GeoCtx& gc = *h->gc;
    const Geometry& g = gc.geo;
    const int id = h->lay->id;
    if (!gc.diagTables.count(id)) {
        std::vector<int64_t> offs; std::vector<int> sizes;
        for (size_t i = 0; i < g.denseBlocks.size(); ++i) {
            const Block& b = g.blocks[g.denseBlocks[i]];
            if (b.t == b.s) { offs.push_back(h->lay->denseOff[i]); sizes.push_back(g.tbox(g.denseBlocks[i]).size); }
        }
        int64_t* d1; int* d2;
        upload(offs, &d1); upload(sizes, &d2);
        gc.diagTables[id] = {d1, d2};
        gc.diagCount[id] = (int)offs.size();
    }
    auto t = gc.diagTables[id];
    const int cnt = gc.diagCount[id];
    if (h->prec == DMHM_F64) launch_add_diag<double>((double*)h->vals, t.first, t.second, cnt, alpha, g_lib.stream);
    else launch_add_diag<float>((float*)h->vals, t.first, t.second, cnt, (float)alpha, g_lib.stream);
}

static void set_to_random(dmhm_hmat_s* h, uint64_t seed)
{
    h->full = true;
    const Layout& L = *h->lay;
    hipStream_t st = g_lib.stream;
    // Dense leaves, then U and V at rank = maxRank (SetToRandom-incl.hpp:34-48);
    // columns beyond maxRank (import capacities) stay zero.
    if (h->prec == DMHM_F64) launch_gaussian<double>((double*)h->vals, L.denseElems, seed, 1, st);
    else launch_gaussian<float>((float*)h->vals, L.denseElems, seed, 1, st);
    const Geometry& g = h->geo();
    bool stdCaps = true;
    for (int c : L.cap) stdCaps = stdCaps && c == std::max(h->maxRank, 1);
    if (stdCaps) {
        if (h->prec == DMHM_F64)
            launch_gaussian<double>((double*)h->vals + L.denseElems, L.uElems + L.vElems, seed, 2, st);
        else
            launch_gaussian<float>((float*)h->vals + L.denseElems, L.uElems + L.vElems, seed, 2, st);
    } else {
        HIPCHECK(hipMemsetAsync((char*)h->vals + L.denseElems * h->esize(), 0, (L.uElems + L.vElems) * h->esize(), st));
        for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
            const int b = g.lrBlocks[i];
            const int64_t nu = (int64_t)g.tbox(b).size * h->maxRank, nv = (int64_t)g.sbox(b).size * h->maxRank;
            if (h->prec == DMHM_F64) {
                launch_gaussian<double>((double*)h->vals + L.uOff[i], nu, seed, 3 + 2 * i, st);
                launch_gaussian<double>((double*)h->vals + L.vOff[i], nv, seed, 4 + 2 * i, st);
            } else {
                launch_gaussian<float>((float*)h->vals + L.uOff[i], nu, seed, 3 + 2 * i, st);
                launch_gaussian<float>((float*)h->vals + L.vOff[i], nv, seed, 4 + 2 * i, st);
            }
        }
    }
    launch_fill_int(h->ranks, L.nLowRank(), h->maxRank, st);
}

// Power iteration of EstimateTwoNorm (hmat_tools.hpp:1504-1544) as launches
// only: the log of the estimate accumulates in dScalars[1] on the device, so
// the sequence can be captured in a HIP graph; read_two_norm fetches it.
template <typename T>
static void launch_two_norm_t(dmhm_hmat_s* h, double confidence, double theta, uint64_t seed)
{
    const Geometry& g = h->geo();
    const int n = g.n;
    const int k = (int)std::ceil(std::log(0.8 * std::sqrt((double)n) * std::pow(10.0, confidence)) /
                                 std::log(theta));
    g_lib.stageX.ensure(2 * (size_t)n * sizeof(double) + 64);
    T* x = (T*)g_lib.stageX.ptr;
    T* y = x + n;
    hipStream_t st = g_lib.stream;
    double* d = g_lib.dScalars;       // d[0] sumsq, d[1] log accumulator
    launch_gaussian<T>(x, n, seed, 7, st, g_lib.seedOff);
    launch_zero_double(d, 2, st);
    launch_sumsq<T>(x, n, d, st);
    launch_normalize<T>(x, n, d, nullptr, 0.0, st);
    for (int i = 0; i < k; ++i) {
        multiply_dense(h, false, 1.0, x, n, 1, 0.0, y, n);
        launch_zero_double(d, 1, st);
        launch_sumsq<T>(y, n, d, st);
        launch_normalize<T>(y, n, d, d + 1, 1.0 / k, st);
        std::swap(x, y);
    }
}

// The same estimate for I - X A applied through H*x products, (I - X A) v =
// v - X (A v), without forming the product X A.
template <typename T>
static void launch_residual_norm_t(dmhm_hmat_s* X, dmhm_hmat_s* A, double confidence, double theta, uint64_t seed)
{
    const int n = X->geo().n;
    const int k = (int)std::ceil(std::log(0.8 * std::sqrt((double)n) * std::pow(10.0, confidence)) /
                                 std::log(theta));
    g_lib.stageX.ensure(3 * (size_t)n * sizeof(double) + 64);
    T* x = (T*)g_lib.stageX.ptr;
    T* y = x + n;
    T* w = y + n;
    hipStream_t st = g_lib.stream;
    double* d = g_lib.dScalars;       // d[0] sumsq, d[1] log accumulator
    launch_gaussian<T>(x, n, seed, 7, st, g_lib.seedOff);
    launch_zero_double(d, 2, st);
    launch_sumsq<T>(x, n, d, st);
    launch_normalize<T>(x, n, d, nullptr, 0.0, st);
    for (int i = 0; i < k; ++i) {
        multiply_dense(A, false, 1.0, x, n, 1, 0.0, y, n);
        multiply_dense(X, false, 1.0, y, n, 1, 0.0, w, n);
        launch_axpby<T>(x, w, T(-1), y, n, st);
        launch_zero_double(d, 1, st);
        launch_sumsq<T>(y, n, d, st);
        launch_normalize<T>(y, n, d, d + 1, 1.0 / k, st);
        std::swap(x, y);
    }
}

static void launch_residual_norm(dmhm_hmat_s* X, dmhm_hmat_s* A, double theta, double confidence, uint64_t seed)
{
    if (X->prec == DMHM_F64) launch_residual_norm_t<double>(X, A, confidence, theta, seed);
    else launch_residual_norm_t<float>(X, A, confidence, theta, seed);
}

static void launch_two_norm(dmhm_hmat_s* h, double theta, double confidence, uint64_t seed)
{
    if (h->prec == DMHM_F64) launch_two_norm_t<double>(h, confidence, theta, seed);
    else launch_two_norm_t<float>(h, confidence, theta, seed);
}

static double read_two_norm(double theta)
{
    HIPCHECK(hipMemcpyAsync(g_lib.hReadback, g_lib.dScalars + 1, sizeof(double), hipMemcpyDeviceToHost,
                            g_lib.stream));
    HIPCHECK(hipStreamSynchronize(g_lib.stream));
    return theta * std::exp(g_lib.hReadback[0]);
}

double estimate_two_norm(dmhm_hmat_s* h, double theta, double confidence, uint64_t seed)
{
    if (theta <= 1) throw Error(DMHM_EINVAL, "Theta must be > 1");
    if (confidence <= 0) throw Error(DMHM_EINVAL, "Confidence must be positive");
    launch_two_norm(h, theta, confidence, seed);
    return read_two_norm(theta);
}


// One Schulz step X_out = (2I - X_in A) X_in (Invert-incl.hpp:160-192) as
// launches only, in two halves. head(): the counter bump, Z = -X_in A, Z += I
// and, in stop mode, the residual power iteration on Z; tail(): Z += I,
// X_out = Z X_in. The stop decision reads the estimate between the halves, so
// a step whose estimate shows the iteration at its floor skips the second
// product. Every seed is the step's fixed base plus the device counter, so a
// captured half replayed by a HIP graph launches exactly what the eager half
// launches.
struct SchulzStep {
    dmhm_hmat_s *A, *Xin, *Xout, *Z;
    bool stop;
    double theta, confidence;
    dmhm_hh_opts hh;
    uint64_t seed;
    void head() const {
        launch_bump(g_lib.dCounter, g_lib.stream);
        dmhm_hh_opts o = hh;
        o.seed = seed;
        multiply_into(-1.0, Xin, A, Z, o);
        add_diag(Z, 1.0);
        if (stop) launch_two_norm(Z, theta, confidence, seed + 1);
    }
    void tail() const {
        add_diag(Z, 1.0);
        dmhm_hh_opts o = hh;
        o.seed = seed + 2;
        multiply_into(1.0, Z, Xin, Xout, o);
    }
};

// The truncation floor of a Schulz iteration: the estimate is below 0.1 and
// either barely moved in the last step (above 0.9 of the previous one) or did
// not halve over the last two. Below 0.1 an iteration in its quadratic regime
// shrinks the estimate to at most theta times its square per step, and one
// whose products lose accuracy to truncation still contracts by more than
// half over two steps while it has room above the floor.
bool stall_detected(const std::vector<double>& h)
{
    const size_t n = h.size();
    if (n < 2 || !(h[n - 1] < 0.1)) return false;
    if (h[n - 1] > 0.9 * h[n - 2]) return true;
    return n >= 3 && h[n - 1] > 0.5 * h[n - 3];
}

// SchulzInvert (Invert-incl.hpp:128-196): X_0 = alpha A^T, alpha = 2/est^2,
// X <- (2I - X A) X. The iterates live in two preallocated buffers of the
// standard layout (plus Z and the transposed start), so the loop allocates
// nothing after its first two steps; with o.graph set, every later step of a
// precision is one HIP graph replay (two graphs, one per ping-pong
// direction) followed by the one readback of the residual estimate.
static void schulz(dmhm_hmat_s* A, const dmhm_schulz_opts& o, dmhm_schulz_report* rep)
{
    const auto t0 = std::chrono::steady_clock::now();
    if (o.theta <= 1) throw Error(DMHM_EINVAL, "Theta must be > 1");
    if (o.confidence <= 0) throw Error(DMHM_EINVAL, "Confidence must be positive");
    bool stop = false;
    int numIt = o.numIterations;
    if (numIt <= 0) { numIt = o.maxIterations > 0 ? o.maxIterations : 30; stop = true; }
    if (numIt > 4096) throw Error(DMHM_EINVAL, "at most 4096 Schulz iterations");
    uint64_t seed = o.hh.seed;
    GeoCtx& gc = *A->gc;
    const int maxRank = A->maxRank;
    auto LS = std_layout(gc, maxRank);

    // Operator and buffers per precision (index 0: FP64, 1: FP32).
    struct PrecState {
        dmhm_hmat_s* A = nullptr;
        std::unique_ptr<dmhm_hmat_s> Aown, X[2], Z;
        StepGraph g[2][2][2];               // [BF16 operands][ping-pong direction][half]
        bool captured[2][2][2] = {};
        bool tuned[2] = {false, false};     // an eager steady step ran with every phase tuned
    } ps[2];
    // The direct start already sits at the truncation floor of A's precision,
    // where lower-precision steps could only add rounding error: it runs every
    // step in A's precision.
    int cur = (o.mixed && o.start != 1 && A->prec == DMHM_F64) ? 1 : (A->prec == DMHM_F64 ? 0 : 1);
    const int target = A->prec == DMHM_F64 ? 0 : 1;    // precision of the result
    auto prepare = [&](int pi) {
        PrecState& P = ps[pi];
        if (P.A) return;
        const int prec = pi == 0 ? DMHM_F64 : DMHM_F32;
        if (A->prec == prec) P.A = A;
        else { P.Aown.reset(copy_hmat(A, prec)); P.A = P.Aown.get(); }
        for (auto& x : P.X) x.reset(alloc_hmat(A->gc, LS, prec, maxRank));
        P.Z.reset(alloc_hmat(A->gc, LS, prec, maxRank));
    };
    prepare(cur);
    launch_zero_double(reinterpret_cast<double*>(g_lib.dCounter), 1, g_lib.stream);   // counter = 0
    SeedOffsetScope seeds;
    // X_0: alpha A^T with alpha = 2 / est(||A||_2)^2 (Invert-incl.hpp:150-158),
    // or the recursive direct inverse (refined by the iteration).
    double est = 0;
    std::unique_ptr<dmhm_hmat_s> X0;
    if (o.start == 1) {
        X0.reset(copy_hmat(ps[cur].A, ps[cur].A->prec));
        direct_invert(X0.get(), o.hh);
    } else {
        est = estimate_two_norm(ps[cur].A, o.theta, o.confidence, seed);
        const double alpha = 2.0 / (est * est);
        X0.reset(transpose_hmat(ps[cur].A));
        scale_hmat(X0.get(), alpha);
    }

    int it = 0, itLow = 0, itGraph = 0, itBf16 = 0;
    double e = -1;
    std::vector<double> hist;
    auto stalled = [&]() { return o.stallStop && stall_detected(hist); };
    dmhm_hmat_s* Xc = X0.get();       // current iterate
    int side = 0;                     // X[side] receives the next iterate
    int stepsInPrec = 0;
    struct Bf16Scope {              // BF16 operands allowed in the tile phases while alive
        explicit Bf16Scope(bool on) { g_lib.bf16 = on; }
        ~Bf16Scope() { g_lib.bf16 = false; }
    };
    // From the direct start the iterate usually sits at the truncation floor
    // already. Each step then first estimates ||I - X_k A||_2 through H*x
    // products (the same estimator, applied to the operator instead of the
    // compressed Z) and decides before Z = I - X_k A is formed: at the floor
    // the step ends there, keeping X_k; below the tolerance it takes the
    // reference's last refinement step and stops.
    const bool direct = o.start == 1 && stop;
    if (direct) {
        PrecState& P = ps[cur];
        for (int kk = 0; kk < numIt; ++kk) {
            it = kk + 1;
            launch_residual_norm(Xc, P.A, o.theta, o.confidence, seed + 1 + 3 * (uint64_t)kk);
            e = read_two_norm(o.theta);
            if (rep && kk < 128) rep->history[kk] = e;
            hist.push_back(e);
            const bool conv = std::fabs(e) < o.stopTol;
            if (!conv && stalled()) break;
            SchulzStep st{P.A, Xc, P.X[side].get(), P.Z.get(), false, o.theta, o.confidence, o.hh,
                          seed + 1 + 3 * (uint64_t)cur};
            st.head();
            st.tail();
            Xc = P.X[side].get();
            side ^= 1;
            if (conv) break;
        }
    }
    for (int kk = 0; kk < (direct ? 0 : numIt); ++kk) {
        it = kk + 1;
        const bool low = cur != target;
        if (low) ++itLow;
        PrecState& P = ps[cur];
        // BF16 operands (FP32 accumulation) for the early low-precision steps:
        // until the residual estimate drops below bf16Tol.
        const int bf = (low && cur == 1 && o.bf16 && (hist.empty() || std::fabs(e) >= o.bf16Tol)) ? 1 : 0;
        if (bf) ++itBf16;
        Bf16Scope bscope(bf != 0);
        SchulzStep st{P.A, Xc, P.X[side].get(), P.Z.get(), stop, o.theta, o.confidence, o.hh,
                      seed + 1 + 3 * (uint64_t)cur};
        // Eager steps build the plans, size the workspace and autotune the
        // tile engines (the very first step reads X_0, whose layout and rank
        // profile differ); once a steady step of this precision and operand
        // mode ran with every phase tuned, later steps replay captured graphs
        // (one per half and ping-pong direction).
        const bool steady = Xc == P.X[0].get() || Xc == P.X[1].get();
        const bool replay = o.graph && steady && stepsInPrec >= 2 && P.tuned[bf];
        const long u0 = g_lib.untuned;
        auto half = [&](int h) {
            if (!replay) { h == 0 ? st.head() : st.tail(); return; }
            StepGraph& G = P.g[bf][side][h];
            if (!P.captured[bf][side][h]) {
                G.capture([&] { h == 0 ? st.head() : st.tail(); });
                P.captured[bf][side][h] = true;
            }
            G.replay();
        };
        half(0);
        if (replay) ++itGraph;
        ++stepsInPrec;
        bool done = false, skipTail = false;
        if (stop) {
            e = read_two_norm(o.theta);
            if (rep && kk < 128) rep->history[kk] = e;
            hist.push_back(e);
            const bool atFloor = stalled();
            // At the floor the second product cannot improve X: keep X_k.
            if (!low && std::fabs(e) < o.stopTol) done = true;          // the reference's stop (with X_k+1)
            else if (!low && atFloor) { done = true; skipTail = true; }
            else if (low && atFloor) skipTail = true;
        }
        if (!skipTail) {
            half(1);
            if (!replay && steady && g_lib.untuned == u0) P.tuned[bf] = true;
            Xc = P.X[side].get();
            side ^= 1;
        }
        if (done) break;
        if (low && stop && (std::fabs(e) < o.switchTol || skipTail)) {
            // Finish in the result's precision: convert the iterate into its buffers.
            prepare(target);
            PrecState& T = ps[target];
            launch_convert<float, double>((const float*)Xc->vals, (double*)T.X[0]->vals, LS->totalElems(),
                                          g_lib.stream);
            HIPCHECK(hipMemcpyAsync(T.X[0]->ranks, Xc->ranks, LS->nLowRank() * sizeof(int), hipMemcpyDeviceToDevice,
                                    g_lib.stream));
            T.X[0]->full = Xc->full;
            Xc = T.X[0].get();
            side = 1;
            cur = target;
            stepsInPrec = 0;
            hist.clear();
        }
    }
    // The result in A's precision.
    std::unique_ptr<dmhm_hmat_s> Xfinal;
    if (Xc->prec != A->prec) {
        prepare(target);
        Xfinal.reset(copy_hmat(Xc, A->prec));
        Xc = Xfinal.get();
    }
    HIPCHECK(hipStreamSynchronize(g_lib.stream));
    // Replace A's contents with X (CopyFrom, Invert-incl.hpp:195); A's old
    // storage is released with the buffer that held X.
    dmhm_hmat_s* holder = Xc;
    A->full = holder->full;
    std::swap(A->lay, holder->lay);
    std::swap(A->vals, holder->vals);
    std::swap(A->ranks, holder->ranks);
    if (rep) {
        rep->iterations = it;
        rep->iterationsLowPrecision = itLow;
        rep->residualEstimate = e;
        rep->normEstimate = est;
        rep->graphIterations = itGraph;
        rep->iterationsBf16 = itBf16;
        rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (stop && !(std::fabs(e) < o.stopTol))
        throw Error(DMHM_ENOCONV, "Schulz iteration did not reach the stop tolerance (estimate " +
                                  std::to_string(e) + ")");
}

template <typename T>
static void to_dense_t(dmhm_hmat_s* h, double* out, int ld)
{
    const Geometry& g = h->geo();
    const Layout& L = *h->lay;
    std::vector<T> v(L.totalElems());
    std::vector<int> r(L.nLowRank());
    HIPCHECK(hipStreamSynchronize(g_lib.stream));
    HIPCHECK(copy_sync(v.data(), h->vals, v.size() * sizeof(T), hipMemcpyDeviceToHost));
    if (!r.empty()) HIPCHECK(copy_sync(r.data(), h->ranks, r.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int j = 0; j < g.n; ++j)
        for (int i = 0; i < g.n; ++i) out[i + (size_t)j * ld] = 0.0;
    for (size_t i = 0; i < g.denseBlocks.size(); ++i) {
        const int b = g.denseBlocks[i];
        const Box& tb = g.tbox(b); const Box& sb = g.sbox(b);
        const T* D = v.data() + L.denseOff[i];
        for (int c = 0; c < sb.size; ++c)
            for (int rr = 0; rr < tb.size; ++rr)
                out[(tb.off + rr) + (size_t)(sb.off + c) * ld] = (double)D[rr + (size_t)c * tb.size];
    }
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        const int b = g.lrBlocks[i];
        const Box& tb = g.tbox(b); const Box& sb = g.sbox(b);
        const T* U = v.data() + L.uOff[i];
        const T* V = v.data() + L.vOff[i];
        for (int c = 0; c < sb.size; ++c)
            for (int rr = 0; rr < tb.size; ++rr) {
                double s = 0;
                for (int p = 0; p < r[i]; ++p) s += (double)U[rr + (size_t)p * tb.size] * (double)V[c + (size_t)p * sb.size];
                out[(tb.off + rr) + (size_t)(sb.off + c) * ld] = s;
            }
    }
}

// Per-box mean and standard deviation of log|a_ii| (the diagonal carries the
// local coefficient of a stencil operator: the sum of its face coefficients).
static std::shared_ptr<OperatorStats> operator_stats(const Geometry& g, const int* rowptr, const int* colind,
                                                     const double* vals)
{
    auto st = std::make_shared<OperatorStats>();
    std::vector<double> ld(g.n, 0.0);
    double tot = 0;
    for (int i = 0; i < g.n; ++i) {
        double d = 0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
            if (colind[k] == i) d = vals[k];
        ld[i] = std::log(std::max(std::fabs(d), 1e-300));
        tot += ld[i];
    }
    st->globalMean = (float)(tot / std::max(1, g.n));
    int base = 0;
    for (int l = 0; l < g.numLevels; ++l) {
        st->levelBase.push_back(base);
        for (const Box& B : g.boxes[l]) {
            double s1 = 0, s2 = 0;
            for (int i = B.off; i < B.off + B.size; ++i) { s1 += ld[i]; s2 += ld[i] * ld[i]; }
            const double m = s1 / std::max(1, B.size);
            st->mean.push_back((float)m);
            st->sdev.push_back((float)std::sqrt(std::max(0.0, s2 / std::max(1, B.size) - m * m)));
        }
        base += (int)g.boxes[l].size();
    }
    return st;
}

// HMat2d/HMat3d(Sparse, ...) + ImportSparse (HMat2d.cpp:144-162, 961-1033)
// on the device: every nonzero walks the block tree to its leaf block
// (csr_locate); the admissible blocks' nonzero counts (their exact import
// ranks, ConvertSubmatrix.cpp:57-119) size the layout; then dense values are
// written in place and each admissible block receives one unit column of U
// and one value column of V per nonzero, in CSR order (csr_fill).
static dmhm_hmat_s* create_from_csr(int dim, const int* sizes, int L, int maxRank, bool strong, int prec,
                                    const int* rowptr, const int* colind, const double* vals)
{
    auto gc = get_geo(dim, sizes, L, strong);
    const Geometry& g = gc->geo;
    const int n = g.n;
    if (rowptr[0] != 0) throw Error(DMHM_EINVAL, "rowOffsets[0] must be 0");
    for (int i = 0; i < n; ++i)
        if (rowptr[i + 1] < rowptr[i]) throw Error(DMHM_EINVAL, "rowOffsets must be nondecreasing");
    const int nnz = rowptr[n];
    const int nLr = (int)g.lrBlocks.size();
    const GeoDev& gd = geo_dev(*gc);
    hipStream_t st = g_lib.stream;
    // Device scratch of the import: CSR, block per nonzero, counts, staging.
    DeviceBuffer dRow, dCol, dVal, dBlk, dInt, dStage, dRowOf;
    dRow.ensure((size_t)(n + 1) * sizeof(int));
    dCol.ensure((size_t)std::max(nnz, 1) * sizeof(int));
    dVal.ensure((size_t)std::max(nnz, 1) * sizeof(double));
    dBlk.ensure((size_t)std::max(nnz, 1) * sizeof(int));
    dInt.ensure((size_t)(3 * nLr + 4) * sizeof(int));       // counts | bases | fill | bad flag
    HIPCHECK(copy_sync(dRow.ptr, rowptr, (size_t)(n + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (nnz > 0) {
        HIPCHECK(copy_sync(dCol.ptr, colind, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
        HIPCHECK(copy_sync(dVal.ptr, vals, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice));
    }
    int* cnt = static_cast<int*>(dInt.ptr);
    int* bases = cnt + nLr;
    int* fill = bases + nLr;
    int* bad = fill + nLr;
    launch_fill_int(cnt, 3 * nLr + 4, 0, st);
    launch_csr_locate(gd, (const int*)dRow.ptr, (const int*)dCol.ptr, n, (int*)dBlk.ptr, cnt, bad, st);
    std::vector<int> importRank(nLr + 1, 0);
    int badCount = 0;
    HIPCHECK(copy_sync(&badCount, bad, sizeof(int), hipMemcpyDeviceToHost));
    if (badCount > 0) throw Error(DMHM_EINVAL, "column index out of range");
    if (nLr > 0) HIPCHECK(copy_sync(importRank.data(), cnt, (size_t)nLr * sizeof(int), hipMemcpyDeviceToHost));
    importRank.resize(nLr);
    std::vector<int> caps(nLr), base(nLr + 1, 0);
    bool stdCaps = true;
    for (int i = 0; i < nLr; ++i) {
        caps[i] = std::max(std::max(maxRank, 1), importRank[i]);
        stdCaps = stdCaps && caps[i] == std::max(maxRank, 1);
        base[i + 1] = base[i] + importRank[i];
    }
    std::shared_ptr<Layout> lay;
    if (stdCaps) lay = std_layout(*gc, maxRank);
    else {
        std::lock_guard<std::mutex> lk(g_lib.mu);
        lay = std::make_shared<Layout>(Layout::make(g, caps, g_lib.nextLayoutId++));
    }
    std::unique_ptr<dmhm_hmat_s> h(alloc_hmat(gc, lay, prec, maxRank));
    const int lrTotal = base[nLr];
    dStage.ensure((size_t)std::max(lrTotal, 1) * sizeof(int));
    dRowOf.ensure((size_t)std::max(nnz, 1) * sizeof(int));
    if (nLr > 0) HIPCHECK(copy_sync(bases, base.data(), (size_t)nLr * sizeof(int), hipMemcpyHostToDevice));
    // Layout offsets for the scatter.
    DeviceBuffer dLay;
    const size_t nd = lay->denseOff.size();
    dLay.ensure((nd + 2 * (size_t)nLr + 1) * sizeof(int64_t));
    int64_t* lo = static_cast<int64_t*>(dLay.ptr);
    if (nd) HIPCHECK(copy_sync(lo, lay->denseOff.data(), nd * sizeof(int64_t), hipMemcpyHostToDevice));
    if (nLr) {
        HIPCHECK(copy_sync(lo + nd, lay->uOff.data(), (size_t)nLr * sizeof(int64_t), hipMemcpyHostToDevice));
        HIPCHECK(copy_sync(lo + nd + nLr, lay->vOff.data(), (size_t)nLr * sizeof(int64_t), hipMemcpyHostToDevice));
    }
    LayDev ld;
    ld.denseOff = lo; ld.uOff = lo + nd; ld.vOff = lo + nd + nLr;
    if (prec == DMHM_F64)
        launch_csr_fill<double>(gd, ld, (const int*)dRow.ptr, (const int*)dCol.ptr, (const double*)dVal.ptr, n,
                                (const int*)dBlk.ptr, bases, fill, cnt, (int*)dStage.ptr, (int*)dRowOf.ptr, nLr,
                                (double*)h->vals, st);
    else
        launch_csr_fill<float>(gd, ld, (const int*)dRow.ptr, (const int*)dCol.ptr, (const double*)dVal.ptr, n,
                               (const int*)dBlk.ptr, bases, fill, cnt, (int*)dStage.ptr, (int*)dRowOf.ptr, nLr,
                               (float*)h->vals, st);
    if (nLr > 0) HIPCHECK(hipMemcpyAsync(h->ranks, cnt, (size_t)nLr * sizeof(int), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipStreamSynchronize(st));
    for (DeviceBuffer* b : {&dRow, &dCol, &dVal, &dBlk, &dInt, &dStage, &dRowOf, &dLay}) b->release();
    h->stats = operator_stats(g, rowptr, colind, vals);
    return h.release();
}

static void matvec_bytes(dmhm_hmat_s* h, int k, double* bytes, double* flops)
{
    const Geometry& g = h->geo();
    const Layout& L = *h->lay;
    std::vector<int> r(L.nLowRank());
    HIPCHECK(hipStreamSynchronize(g_lib.stream));
    if (!r.empty()) HIPCHECK(copy_sync(r.data(), h->ranks, r.size() * sizeof(int), hipMemcpyDeviceToHost));
    double entries = 0;
    for (int b : g.denseBlocks) entries += (double)g.tbox(b).size * g.sbox(b).size;
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        const int b = g.lrBlocks[i];
        entries += (double)(g.tbox(b).size + g.sbox(b).size) * r[i];
    }
    const double es = (double)h->esize();
    if (bytes) *bytes = es * entries + 2.0 * es * g.n * k;
    if (flops) *flops = 2.0 * k * entries;
}



static void profile_matvec(dmhm_hmat_s* h, int k, int reps, double* ms, double* bytes)
{
    const Geometry& g = h->geo();
    const Layout& L = *h->lay;
    const int n = g.n;
    GeoCtx& gc = *h->gc;
    const bool stream = k <= 8;
    const std::string key = std::string(stream ? "mvs:" : "mv:") + std::to_string(L.id) + ":0:" +
                            std::to_string(k) + ":" + std::to_string(n) + ":" + std::to_string(n);
    Plan& p = get_plan(gc, key, [&] {
        Planner pl(gc.geo);
        return stream ? pl.matvecStream(L, k, n, n) : pl.matvec(L, false, k, n, n);
    });
    const size_t last = p.phases.size() - 1;
    ensure_ws(p, h->esize());
    g_lib.stageX.ensure((size_t)n * k * h->esize());
    g_lib.stageY.ensure((size_t)n * k * h->esize());
    if (h->prec == DMHM_F64) launch_gaussian<double>((double*)g_lib.stageX.ptr, (int64_t)n * k, 5, 5, g_lib.stream);
    else launch_gaussian<float>((float*)g_lib.stageX.ptr, (int64_t)n * k, 5, 5, g_lib.stream);
    Bases b = bases_for(h, nullptr, nullptr, g_lib.stageX.ptr, g_lib.stageY.ptr);
    EventPair e0, e1;
    double tot = 0, t0 = 0, t1 = 0;
    // Untimed runs first: enough for the tile-engine tuning of the non-
    // streaming form (a warm-up and up to five candidates).
    for (int r = -6; r < reps; ++r) {
        HIPCHECK(hipEventRecord(e0.a, g_lib.stream));
        for (size_t i = 0; i < last; ++i) {
            if (h->prec == DMHM_F64) run_phase<double>(p.phases[i], b, 1.0, 0.0, {0, 0}, 0);
            else run_phase<float>(p.phases[i], b, 1.0f, 0.0f, {0, 0}, 0);
        }
        HIPCHECK(hipEventRecord(e0.b, g_lib.stream));
        HIPCHECK(hipEventRecord(e1.a, g_lib.stream));
        if (h->prec == DMHM_F64) run_phase<double>(p.phases[last], b, 1.0, 0.0, {0, 0}, 0);
        else run_phase<float>(p.phases[last], b, 1.0f, 0.0f, {0, 0}, 0);
        HIPCHECK(hipEventRecord(e1.b, g_lib.stream));
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
        if (r >= 0) {
            const double a = e0.ms(), c = e1.ms();
            float whole = 0;
            HIPCHECK(hipEventElapsedTime(&whole, e0.a, e1.b));
            t0 += a; t1 += c; tot += whole;
        }
    }
    ms[0] = tot / reps; ms[1] = t0 / reps; ms[2] = t1 / reps;
    // Algorithmic bytes at the current ranks.
    std::vector<int> r(L.nLowRank());
    if (!r.empty()) HIPCHECK(copy_sync(r.data(), h->ranks, r.size() * sizeof(int), hipMemcpyDeviceToHost));
    double vEnt = 0, uEnt = 0, dEnt = 0;
    for (int bd : g.denseBlocks) dEnt += (double)g.tbox(bd).size * g.sbox(bd).size;
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        uEnt += (double)g.tbox(g.lrBlocks[i]).size * r[i];
        vEnt += (double)g.sbox(g.lrBlocks[i]).size * r[i];
    }
    const double es = (double)h->esize();
    bytes[0] = es * (vEnt + (double)n * k);
    bytes[1] = es * (dEnt + uEnt + (double)n * k);
}

struct PhaseTiming { int kind; double ms, flops; };
static std::vector<PhaseTiming> g_lastPhases;

static void profile_hh(double alpha, dmhm_hmat_s* A, dmhm_hmat_s* B, const dmhm_hh_opts& o, double* ms,
                       double* flops)
{
    if (A->gc != B->gc) throw Error(DMHM_ENONCONF, "Attempted nonconformal matrix-matrix multiply");
    if (A->prec != B->prec) throw Error(DMHM_EINVAL, "operands must share a precision");
    GeoCtx& gc = *A->gc;
    const int maxRank = A->maxRank;
    auto LC = std_layout(gc, maxRank);
    std::unique_ptr<dmhm_hmat_s> C(alloc_hmat(A->gc, LC, A->prec, maxRank));
    const int over = o.oversample > 0 ? o.oversample : 4;
    const std::string key = "hh:" + std::to_string(A->lay->id) + ":" + std::to_string(B->lay->id) + ":" +
                            std::to_string(LC->id) + ":" + std::to_string(maxRank) + ":" + std::to_string(over) +
                            ":" + std::to_string(A->full) + std::to_string(B->full);
    Plan& p = get_plan(gc, key, [&] {
        Planner pl(gc.geo);
        pl.setWorkspaceBudget(hh_batch_budget());
        pl.setStacking(A->full, B->full);
        return pl.hhProduct(*A->lay, *B->lay, *LC, maxRank, over);
    });
    ensure_ws(p, A->esize());
    Bases b = bases_for(A, B, C.get(), nullptr, nullptr);
    TruncPolicy pol{o.tolMode, o.relTol};
    std::vector<EventPair> ev(p.phases.size());
    EventPair all;
    HIPCHECK(hipStreamSynchronize(g_lib.stream));
    HIPCHECK(hipEventRecord(all.a, g_lib.stream));
    for (size_t i = 0; i < p.phases.size(); ++i) {
        HIPCHECK(hipEventRecord(ev[i].a, g_lib.stream));
        if (A->prec == DMHM_F64) run_phase<double>(p.phases[i], b, alpha, 0.0, pol, o.seed);
        else run_phase<float>(p.phases[i], b, (float)alpha, 0.0f, pol, o.seed);
        HIPCHECK(hipEventRecord(ev[i].b, g_lib.stream));
    }
    HIPCHECK(hipEventRecord(all.b, g_lib.stream));
    HIPCHECK(hipStreamSynchronize(g_lib.stream));
    for (int i = 0; i < 6; ++i) ms[i] = 0;
    flops[0] = flops[1] = 0;
    ms[0] = all.ms();
    g_lastPhases.clear();
    for (size_t i = 0; i < p.phases.size(); ++i) {
        const double t = ev[i].ms();
        g_lastPhases.push_back({(int)p.phases[i].kind, t, p.phases[i].flops});
        switch (p.phases[i].kind) {
        case PH_TILE:
            ms[1] += t;
            flops[0] += p.phases[i].flops;
            if (t > ms[5]) { ms[5] = t; flops[1] = p.phases[i].flops; }
            break;
        case PH_QR: ms[2] += t; break;
        case PH_SVD: ms[3] += t; break;
        case PH_RNG: ms[4] += t; break;
        default: break;
        }
    }
}

int guard(const std::function<void()>& fn)
{
    try {
        fn();
        return DMHM_OK;
    } catch (const Error& e) {
        g_lastError = e.what();
        return e.code;
    } catch (const std::invalid_argument& e) {
        g_lastError = e.what();
        return DMHM_EINVAL;
    } catch (const std::logic_error& e) {
        g_lastError = e.what();
        return DMHM_EINVAL;
    } catch (const std::exception& e) {
        g_lastError = e.what();
        return DMHM_EHIP;
    }
}

} // namespace dmhm

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int dmhm_init(int device)
{
    return guard([&] {
        if (g_lib.ready) return;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
            throw Error(DMHM_ENODEV, "no HIP device available");
        if (device < 0 || device >= count) throw Error(DMHM_EINVAL, "device index out of range");
        HIPCHECK(hipSetDevice(device));
        HIPCHECK(hipStreamCreateWithFlags(&g_lib.stream, hipStreamNonBlocking));
        HIPCHECK(hipMalloc((void**)&g_lib.dScalars, 16 * sizeof(double)));
        HIPCHECK(hipMalloc((void**)&g_lib.dCounter, sizeof(uint64_t)));
        HIPCHECK(hipHostMalloc((void**)&g_lib.hReadback, 4 * sizeof(double), hipHostMallocDefault));
        scratch_setup(device);
        if (const char* e = getenv("DMHM_TILE_ENGINE")) {
            const std::string v(e);
            if (v == "lds") g_lib.tileEngine = TILE_LDS;
            else if (v == "mfma") g_lib.tileEngine = TILE_MFMA;
            else if (v == "mfma_pf") g_lib.tileEngine = TILE_MFMA_PF;
            else if (v == "bf16") g_lib.tileEngine = TILE_MFMA_BF16;
            else if (v == "bf16_pf") g_lib.tileEngine = TILE_MFMA_BF16_PF;
            else if (v != "auto") throw Error(DMHM_EINVAL, "DMHM_TILE_ENGINE must be auto, lds, mfma, mfma_pf, bf16 or bf16_pf");
        }
        if (!g_lib.tuneEv) g_lib.tuneEv = new EventPair();
        g_lib.device = device;
        g_lib.ready = true;
    });
}

int dmhm_finalize(void)
{
    return guard([&] {
        if (!g_lib.ready) return;
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
        release_batch_plans();
        delete g_lib.tuneEv;
        g_lib.tuneEv = nullptr;
        g_lib.ws.release(); g_lib.stageX.release(); g_lib.stageY.release();
        g_lib.featBuf.release(); g_lib.widthBuf.release(); g_lib.flagBuf.release();
        (void)hipFree(g_lib.dScalars);
        (void)hipFree(g_lib.dCounter);
        (void)hipHostFree(g_lib.hReadback);
        (void)hipStreamDestroy(g_lib.stream);
        g_lib.ready = false;
    });
}

const char* dmhm_last_error(void) { return g_lastError.c_str(); }

int dmhm_synchronize(void)
{
    return guard([&] { require_ready(); HIPCHECK(hipStreamSynchronize(g_lib.stream)); });
}

int dmhm_build_natural_to_hierarchical_map(int dim, const int* sizes, int numLevels, int* map)
{
    return guard([&] {
        if (!sizes || !map) throw Error(DMHM_EINVAL, "null pointer");
        Geometry g(dim, sizes, numLevels, true);
        auto m = g.naturalToHierarchical();
        std::copy(m.begin(), m.end(), map);
    });
}

int dmhm_build_natural_to_hierarchical_map_device(int dim, const int* sizes, int numLevels, int* map)
{
    return guard([&] {
        require_ready();
        if (!sizes || !map) throw Error(DMHM_EINVAL, "null pointer");
        auto gc = get_geo(dim, sizes, numLevels, true);
        const GeoDev& gd = geo_dev(*gc);
        const Geometry& g = gc->geo;
        DeviceBuffer d;
        d.ensure((size_t)std::max(g.n, 1) * sizeof(int));
        launch_natural_map(gd, g.sizes[0], g.sizes[1], g.sizes[2], static_cast<int*>(d.ptr), g_lib.stream);
        HIPCHECK(copy_sync(map, d.ptr, (size_t)g.n * sizeof(int), hipMemcpyDeviceToHost));
        d.release();
    });
}

int dmhm_geometry_block_counts(int dim, const int* sizes, int numLevels, int strong, long long* dense,
                               long long* lowRank, int maxLevels)
{
    return guard([&] {
        if (!sizes || !dense || !lowRank) throw Error(DMHM_EINVAL, "null pointer");
        Geometry g(dim, sizes, numLevels, strong != 0);
        std::vector<long> d, f;
        g.counts(d, f);
        for (int l = 0; l < maxLevels; ++l) {
            dense[l] = l < (int)d.size() ? d[l] : 0;
            lowRank[l] = l < (int)f.size() ? f[l] : 0;
        }
    });
}

int dmhm_geometry_leaves(int dim, const int* sizes, int numLevels, int strong, int* out, int maxLeaves,
                         int* count)
{
    return guard([&] {
        if (!sizes || !count) throw Error(DMHM_EINVAL, "null pointer");
        Geometry g(dim, sizes, numLevels, strong != 0);
        *count = (int)g.leaves.size();
        for (int i = 0; i < (int)g.leaves.size() && i < maxLeaves && out; ++i) {
            const Block& b = g.blocks[g.leaves[i]];
            const Box& tb = g.tbox(g.leaves[i]);
            const Box& sb = g.sbox(g.leaves[i]);
            int* r = out + 7 * (size_t)i;
            r[0] = b.type == BT_DENSE ? 1 : 0; r[1] = b.level;
            r[2] = tb.off; r[3] = tb.size; r[4] = sb.off; r[5] = sb.size;
            r[6] = b.type == BT_DENSE ? b.denseIndex : b.lrIndex;
        }
    });
}

int dmhm_geometry_product_terms(int dim, const int* sizes, int numLevels, int strong, long long* out)
{
    return guard([&] {
        if (!sizes || !out) throw Error(DMHM_EINVAL, "null pointer");
        Geometry g(dim, sizes, numLevels, strong != 0);
        ProductTerms t = enumerateProductTerms(g);
        out[0] = t.sketchTerms; out[1] = t.ddToLowRank; out[2] = t.ddToDense;
        out[3] = t.lowRankTerms; out[4] = t.pushdownTerms;
    });
}

int dmhm_hmat_create_from_csr(int dim, const int* sizes, int numLevels, int maxRank, int strong,
                              int precision, const int* rowOffsets, const int* columnIndices,
                              const double* nonzeros, dmhm_hmat* out)
{
    return guard([&] {
        require_ready();
        if (!sizes || !rowOffsets || !columnIndices || !nonzeros || !out) throw Error(DMHM_EINVAL, "null pointer");
        if (precision != DMHM_F64 && precision != DMHM_F32) throw Error(DMHM_EINVAL, "precision must be 8 or 4");
        if (maxRank < 0 || maxRank > 64) throw Error(DMHM_EINVAL, "maxRank must be in [0, 64]");
        *out = create_from_csr(dim, sizes, numLevels, maxRank, strong != 0, precision, rowOffsets, columnIndices,
                               nonzeros);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_hmat_create_zero(int dim, const int* sizes, int numLevels, int maxRank, int strong, int precision,
                          dmhm_hmat* out)
{
    return guard([&] {
        require_ready();
        if (!sizes || !out) throw Error(DMHM_EINVAL, "null pointer");
        if (precision != DMHM_F64 && precision != DMHM_F32) throw Error(DMHM_EINVAL, "precision must be 8 or 4");
        if (maxRank < 0 || maxRank > 64) throw Error(DMHM_EINVAL, "maxRank must be in [0, 64]");
        auto gc = get_geo(dim, sizes, numLevels, strong != 0);
        *out = alloc_hmat(gc, std_layout(*gc, maxRank), precision, maxRank);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_hmat_destroy(dmhm_hmat h)
{
    return guard([&] {
        if (!h) return;
        if (g_lib.ready) HIPCHECK(hipStreamSynchronize(g_lib.stream));
        delete h;
    });
}

int dmhm_hmat_info(dmhm_hmat h, dmhm_info* info)
{
    return guard([&] {
        if (!h || !info) throw Error(DMHM_EINVAL, "null pointer");
        const Geometry& g = h->geo();
        info->dim = g.dim; info->n = g.n; info->numLevels = g.numLevels; info->maxRank = h->maxRank;
        info->strong = g.strong; info->precision = h->prec;
        info->numDense = (long long)g.denseBlocks.size();
        info->numLowRank = (long long)g.lrBlocks.size();
        info->storageBytes = (long long)(h->lay->totalElems() * h->esize());
        std::vector<int> r(h->lay->nLowRank());
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
        if (!r.empty()) HIPCHECK(copy_sync(r.data(), h->ranks, r.size() * sizeof(int), hipMemcpyDeviceToHost));
        long long s = 0; int mx = 0;
        for (int v : r) { s += v; mx = std::max(mx, v); }
        info->sumRank = s; info->maxBlockRank = mx;
    });
}

int dmhm_hmat_block_counts(dmhm_hmat h, long long* dense, long long* lowRank, int maxLevels)
{
    return guard([&] {
        if (!h || !dense || !lowRank) throw Error(DMHM_EINVAL, "null pointer");
        std::vector<long> d, f;
        h->geo().counts(d, f);
        for (int l = 0; l < maxLevels; ++l) {
            dense[l] = l < (int)d.size() ? d[l] : 0;
            lowRank[l] = l < (int)f.size() ? f[l] : 0;
        }
    });
}

int dmhm_hmat_set_to_random(dmhm_hmat h, uint64_t seed)
{
    return guard([&] {
        require_ready();
        if (!h) throw Error(DMHM_EINVAL, "null handle");
        set_to_random(h, seed);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_hmat_multiply_dense(dmhm_hmat h, char op, double alpha, const void* X, int ldx, int k, double beta,
                             void* Y, int ldy, int memory)
{
    return guard([&] {
        require_ready();
        if (!h || !X || !Y) throw Error(DMHM_EINVAL, "null pointer");
        const int n = h->geo().n;
        if (k < 0 || ldx < n || ldy < n) throw Error(DMHM_EINVAL, "bad leading dimension or column count");
        if (k == 0) return;
        const bool trans = (op == 'T' || op == 't' || op == 'C' || op == 'c');
        if (!trans && op != 'N' && op != 'n') throw Error(DMHM_EINVAL, "op must be N, T or C");
        if (memory == DMHM_DEVICE) {
            multiply_dense(h, trans, alpha, X, ldx, k, beta, Y, ldy);
        } else {
            const size_t es = h->esize();
            const size_t bx = (size_t)ldx * k * es, by = (size_t)ldy * k * es;
            g_lib.stageX.ensure(bx);
            g_lib.stageY.ensure(by);
            HIPCHECK(hipMemcpyAsync(g_lib.stageX.ptr, X, bx, hipMemcpyHostToDevice, g_lib.stream));
            if (beta != 0.0) HIPCHECK(hipMemcpyAsync(g_lib.stageY.ptr, Y, by, hipMemcpyHostToDevice, g_lib.stream));
            multiply_dense(h, trans, alpha, g_lib.stageX.ptr, ldx, k, beta, g_lib.stageY.ptr, ldy);
            HIPCHECK(hipMemcpyAsync(Y, g_lib.stageY.ptr, by, hipMemcpyDeviceToHost, g_lib.stream));
        }
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_rank_model_num_params(void) { return mlp_param_count(kRankFeatures); }

int dmhm_rank_model_create(uint64_t seed, dmhm_rank_model* out)
{
    return guard([&] {
        require_ready();
        if (!out) throw Error(DMHM_EINVAL, "null pointer");
        std::unique_ptr<dmhm_rank_model_s> m(new dmhm_rank_model_s());
        m->nf = kRankFeatures;
        m->np = mlp_param_count(m->nf);
        // He initialisation of the weights, zero biases (layout as in k_mlp.hip).
        std::vector<float> p(m->np, 0.0f);
        std::mt19937_64 gen(seed);
        std::normal_distribution<float> nd(0.0f, 1.0f);
        const int H = 64;
        float* W1 = p.data();
        float* W2 = W1 + m->nf * H + H;
        float* w3 = W2 + H * H + H;
        for (int e = 0; e < m->nf * H; ++e) W1[e] = nd(gen) * std::sqrt(2.0f / m->nf);
        for (int e = 0; e < H * H; ++e) W2[e] = nd(gen) * std::sqrt(2.0f / H);
        for (int e = 0; e < H; ++e) w3[e] = nd(gen) * std::sqrt(1.0f / H);
        const size_t bytes = (size_t)m->np * sizeof(float);
        m->params.ensure(bytes); m->grads.ensure(bytes); m->m.ensure(bytes); m->v.ensure(bytes);
        HIPCHECK(copy_sync(m->params.ptr, p.data(), bytes, hipMemcpyHostToDevice));
        launch_fill<float>((float*)m->grads.ptr, m->np, 0.0f, g_lib.stream);
        launch_fill<float>((float*)m->m.ptr, m->np, 0.0f, g_lib.stream);
        launch_fill<float>((float*)m->v.ptr, m->np, 0.0f, g_lib.stream);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
        *out = m.release();
    });
}

int dmhm_rank_model_destroy(dmhm_rank_model m)
{
    return guard([&] {
        if (g_lib.ready) HIPCHECK(hipStreamSynchronize(g_lib.stream));
        delete m;
    });
}

int dmhm_rank_model_get_params(dmhm_rank_model m, float* params, int count)
{
    return guard([&] {
        require_ready();
        if (!m || !params || count != m->np) throw Error(DMHM_EINVAL, "bad parameter buffer");
        HIPCHECK(copy_sync(params, m->params.ptr, (size_t)count * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int dmhm_rank_model_set_params(dmhm_rank_model m, const float* params, int count)
{
    return guard([&] {
        require_ready();
        if (!m || !params || count != m->np) throw Error(DMHM_EINVAL, "bad parameter buffer");
        HIPCHECK(copy_sync(m->params.ptr, params, (size_t)count * sizeof(float), hipMemcpyHostToDevice));
        m->step = 0;
        launch_fill<float>((float*)m->m.ptr, m->np, 0.0f, g_lib.stream);
        launch_fill<float>((float*)m->v.ptr, m->np, 0.0f, g_lib.stream);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_rank_model_gradient(dmhm_rank_model m, const float* X, const float* t, int n, float* grads, float* loss)
{
    return guard([&] {
        require_ready();
        if (!m || !X || !t || !grads || n <= 0) throw Error(DMHM_EINVAL, "bad arguments");
        DeviceBuffer dX, dt, dl;
        dX.ensure((size_t)n * m->nf * sizeof(float)); dt.ensure((size_t)n * sizeof(float)); dl.ensure(64);
        HIPCHECK(copy_sync(dX.ptr, X, (size_t)n * m->nf * sizeof(float), hipMemcpyHostToDevice));
        HIPCHECK(copy_sync(dt.ptr, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
        launch_fill<float>((float*)dl.ptr, 1, 0.0f, g_lib.stream);
        launch_mlp_grad((const float*)m->params.ptr, m->nf, (const float*)dX.ptr, (const float*)dt.ptr, n,
                        (float*)m->grads.ptr, (float*)dl.ptr, g_lib.stream);
        HIPCHECK(copy_sync(grads, m->grads.ptr, (size_t)m->np * sizeof(float), hipMemcpyDeviceToHost));
        if (loss) HIPCHECK(copy_sync(loss, dl.ptr, sizeof(float), hipMemcpyDeviceToHost));
        launch_fill<float>((float*)m->grads.ptr, m->np, 0.0f, g_lib.stream);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
        dX.release(); dt.release(); dl.release();
    });
}

int dmhm_rank_model_predict(dmhm_rank_model m, const float* X, int n, float* y)
{
    return guard([&] {
        require_ready();
        if (!m || !X || !y || n < 0) throw Error(DMHM_EINVAL, "bad arguments");
        if (n == 0) return;
        DeviceBuffer dX, dy;
        dX.ensure((size_t)n * m->nf * sizeof(float)); dy.ensure((size_t)n * sizeof(float));
        HIPCHECK(copy_sync(dX.ptr, X, (size_t)n * m->nf * sizeof(float), hipMemcpyHostToDevice));
        launch_mlp_infer((const float*)m->params.ptr, m->nf, (const float*)dX.ptr, n, (float*)dy.ptr, nullptr,
                         0.0f, 0, 0, 0, g_lib.stream);
        HIPCHECK(copy_sync(y, dy.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
        dX.release(); dy.release();
    });
}

int dmhm_rank_model_train(dmhm_rank_model m, const float* X, const float* t, int n, int steps, int batch, float lr,
                          uint64_t seed, float* losses)
{
    return guard([&] {
        require_ready();
        if (!m || !X || !t || n <= 0 || steps < 0 || batch <= 0 || !(lr > 0))
            throw Error(DMHM_EINVAL, "bad training arguments");
        // Seeded shuffle of the samples on the host, then everything on the device.
        std::vector<int> perm(n);
        for (int i = 0; i < n; ++i) perm[i] = i;
        std::mt19937_64 gen(seed);
        std::shuffle(perm.begin(), perm.end(), gen);
        std::vector<float> hx((size_t)n * m->nf), ht(n);
        for (int i = 0; i < n; ++i) {
            std::memcpy(&hx[(size_t)i * m->nf], X + (size_t)perm[i] * m->nf, m->nf * sizeof(float));
            ht[i] = t[perm[i]];
        }
        DeviceBuffer dX, dt, dl;
        dX.ensure(hx.size() * sizeof(float)); dt.ensure(ht.size() * sizeof(float));
        dl.ensure((size_t)std::max(steps, 1) * sizeof(float));
        HIPCHECK(copy_sync(dX.ptr, hx.data(), hx.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPCHECK(copy_sync(dt.ptr, ht.data(), ht.size() * sizeof(float), hipMemcpyHostToDevice));
        launch_fill<float>((float*)dl.ptr, std::max(steps, 1), 0.0f, g_lib.stream);
        const int b = std::min(batch, n);
        for (int st = 0; st < steps; ++st) {
            const int off = (int)(((int64_t)st * b) % n);
            const int nb = std::min(b, n - off);
            launch_mlp_grad((const float*)m->params.ptr, m->nf, (const float*)dX.ptr + (size_t)off * m->nf,
                            (const float*)dt.ptr + off, nb, (float*)m->grads.ptr, (float*)dl.ptr + st,
                            g_lib.stream);
            launch_mlp_adam((float*)m->params.ptr, (float*)m->grads.ptr, (float*)m->m.ptr, (float*)m->v.ptr, m->np,
                            lr, ++m->step, g_lib.stream);
        }
        if (losses && steps > 0)
            HIPCHECK(copy_sync(losses, dl.ptr, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost));
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
        dX.release(); dt.release(); dl.release();
    });
}

int dmhm_hmat_rank_features(dmhm_hmat A, float* out, int count)
{
    return guard([&] {
        require_ready();
        if (!A || !out) throw Error(DMHM_EINVAL, "null pointer");
        const int nl = (int)A->geo().lrBlocks.size();
        if (count != nl * kRankFeatures) throw Error(DMHM_EINVAL, "count must be numLowRank x DMHM_RANK_FEATURES");
        if (nl == 0) return;
        g_lib.featBuf.ensure((size_t)nl * kRankFeatures * sizeof(float));
        rank_features(A, static_cast<float*>(g_lib.featBuf.ptr));
        HIPCHECK(copy_sync(out, g_lib.featBuf.ptr, (size_t)count * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int dmhm_last_product_stats(long long* out)
{
    return guard([&] {
        if (!out) throw Error(DMHM_EINVAL, "null pointer");
        out[0] = g_lastProduct.blocks; out[1] = g_lastProduct.widthSum;
        out[2] = g_lastProduct.fallbacks; out[3] = g_lastProduct.fullWidth;
    });
}

int dmhm_set_tile_engine(int engine)
{
    return guard([&] {
        if (engine < -1 || engine >= TILE_ENGINES) throw Error(DMHM_EINVAL, "unknown tile engine");
        g_lib.tileEngine = engine;
    });
}

int dmhm_get_tile_engine(void) { return g_lib.tileEngine; }

int dmhm_hmat_multiply_dense_noalloc(dmhm_hmat h, double alpha, const void* X, int ldx, int k, double beta,
                                     void* Y, int ldy)
{
    return guard([&] {
        require_ready();
        if (!h || !X || !Y) throw Error(DMHM_EINVAL, "null pointer");
        const int n = h->geo().n;
        if (ldx < n || ldy < n) throw Error(DMHM_EINVAL, "bad leading dimension");
        multiply_dense_noalloc(h, alpha, X, ldx, k, beta, Y, ldy);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_hmat_multiply_hmat(double alpha, dmhm_hmat A, dmhm_hmat B, dmhm_hmat* C, const dmhm_hh_opts* opts)
{
    return guard([&] {
        require_ready();
        if (!A || !B || !C) throw Error(DMHM_EINVAL, "null pointer");
        dmhm_hh_opts o{4, 0, 1e-10, 1};
        if (opts) o = *opts;
        *C = multiply_hmat(alpha, A, B, o);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_hmat_multiply_hmat_update(double alpha, dmhm_hmat A, dmhm_hmat B, double beta, dmhm_hmat C,
                                   const dmhm_hh_opts* opts)
{
    return guard([&] {
        require_ready();
        if (!A || !B || !C) throw Error(DMHM_EINVAL, "null pointer");
        dmhm_hh_opts o{4, 0, 1e-10, 1};
        if (opts) o = *opts;
        multiply_hmat_update(alpha, A, B, beta, C, o);
    });
}

int dmhm_hmat_update_with(dmhm_hmat A, double alpha, dmhm_hmat B, const dmhm_hh_opts* opts)
{
    return guard([&] {
        require_ready();
        if (!A || !B) throw Error(DMHM_EINVAL, "null pointer");
        dmhm_hh_opts o{4, 0, 1e-10, 1};
        if (opts) o = *opts;
        update_with(A, alpha, B, o);
    });
}

int dmhm_hmat_create_from_lowrank(int dim, const int* sizes, int numLevels, int maxRank, int strong, int precision,
                                  int r, const double* U, const double* V, dmhm_hmat* out)
{
    return guard([&] {
        require_ready();
        if (!sizes || !out || (r > 0 && (!U || !V))) throw Error(DMHM_EINVAL, "null pointer");
        if (precision != DMHM_F64 && precision != DMHM_F32) throw Error(DMHM_EINVAL, "precision must be 8 or 4");
        if (maxRank < 0 || maxRank > 64 || r < 0) throw Error(DMHM_EINVAL, "bad rank");
        *out = create_from_lowrank(dim, sizes, numLevels, maxRank, strong != 0, precision, r, U, V);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_rounded_update(int maxRank, double alpha, int m, int n, int rA, const double* AU, const double* AV,
                        double beta, int rB, const double* BU, const double* BV, double* outU, double* outV,
                        int* outRank)
{
    return guard([&] {
        require_ready();
        if (m <= 0 || n <= 0 || rA < 0 || rB < 0 || maxRank < 0) throw Error(DMHM_EINVAL, "bad sizes");
        if (!outU || !outV || !outRank || (rA && (!AU || !AV)) || (rB && (!BU || !BV)))
            throw Error(DMHM_EINVAL, "null pointer");
        rounded_update_dev(maxRank, alpha, m, n, rA, AU, AV, beta, rB, BU, BV, outU, outV, outRank);
    });
}

int dmhm_compress_dense(int maxRank, int m, int n, const double* D, double* outU, double* outV, int* outRank)
{
    return guard([&] {
        require_ready();
        if (m <= 0 || n <= 0 || maxRank < 0) throw Error(DMHM_EINVAL, "bad sizes");
        if (!D || !outU || !outV || !outRank) throw Error(DMHM_EINVAL, "null pointer");
        compress_dense_dev(maxRank, m, n, D, outU, outV, outRank);
    });
}

int dmhm_hmat_copy(dmhm_hmat src, dmhm_hmat* dst)
{
    return guard([&] {
        require_ready();
        if (!src || !dst) throw Error(DMHM_EINVAL, "null pointer");
        *dst = copy_hmat(src, src->prec);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_hmat_convert(dmhm_hmat src, int precision, dmhm_hmat* dst)
{
    return guard([&] {
        require_ready();
        if (!src || !dst) throw Error(DMHM_EINVAL, "null pointer");
        if (precision != DMHM_F64 && precision != DMHM_F32) throw Error(DMHM_EINVAL, "precision must be 8 or 4");
        *dst = copy_hmat(src, precision);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_hmat_transpose(dmhm_hmat src, dmhm_hmat* dst)
{
    return guard([&] {
        require_ready();
        if (!src || !dst) throw Error(DMHM_EINVAL, "null pointer");
        *dst = transpose_hmat(src);
    });
}

int dmhm_hmat_scale(dmhm_hmat h, double alpha)
{
    return guard([&] {
        require_ready();
        if (!h) throw Error(DMHM_EINVAL, "null handle");
        scale_hmat(h, alpha);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_hmat_set_to_identity(dmhm_hmat h)
{
    return guard([&] {
        require_ready();
        if (!h) throw Error(DMHM_EINVAL, "null handle");
        scale_hmat(h, 0.0);
        add_diag(h, 1.0);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_hmat_add_constant_to_diagonal(dmhm_hmat h, double alpha)
{
    return guard([&] {
        require_ready();
        if (!h) throw Error(DMHM_EINVAL, "null handle");
        add_diag(h, alpha);
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
    });
}

int dmhm_hmat_estimate_two_norm(dmhm_hmat h, double theta, double confidence, uint64_t seed, double* estimate)
{
    return guard([&] {
        require_ready();
        if (!h || !estimate) throw Error(DMHM_EINVAL, "null pointer");
        *estimate = estimate_two_norm(h, theta, confidence, seed);
    });
}

int dmhm_hmat_schulz_invert(dmhm_hmat h, const dmhm_schulz_opts* opts, dmhm_schulz_report* report)
{
    return guard([&] {
        require_ready();
        if (!h) throw Error(DMHM_EINVAL, "null handle");
        if (h->geo().n == 0) throw Error(DMHM_EINVAL, "empty matrix");
        dmhm_schulz_opts o{};
        if (opts) o = *opts;
        else { o.numIterations = -1; o.maxIterations = 30; o.theta = 1.5; o.confidence = 6; o.stopTol = 1e-4;
               o.hh = {4, 0, 1e-10, 1}; }
        if (report) std::memset(report, 0, sizeof(*report));
        schulz(h, o, report);
    });
}

int dmhm_hmat_direct_invert(dmhm_hmat h, const dmhm_hh_opts* opts)
{
    return guard([&] {
        require_ready();
        if (!h) throw Error(DMHM_EINVAL, "null handle");
        dmhm_hh_opts o{4, 0, 1e-10, 1, nullptr, 0.0};
        if (opts) o = *opts;
        direct_invert(h, o);
    });
}

int dmhm_check_direct_invert_program(int dim, const int* sizes, int numLevels, int maxRank, int strong,
                                     int oversample, long long* stats)
{
    return guard([&] {
        if (!sizes || !stats) throw Error(DMHM_EINVAL, "null pointer");
        const int over = oversample > 0 ? oversample : 4;
        if (maxRank < 1 || maxRank + over > 64) throw Error(DMHM_EINVAL, "maxRank + oversample must be <= 64");
        Geometry g(dim, sizes, numLevels, strong != 0);
        check_direct_invert_program(g, maxRank, over, stats);
    });
}

int dmhm_direct_invert_program_stats(dmhm_hmat h, int oversample, long long* products, long long* plans,
                                     double* seconds)
{
    return guard([&] {
        require_ready();
        if (!h || !products || !plans || !seconds) throw Error(DMHM_EINVAL, "null pointer");
        long p = 0, q = 0;
        direct_invert_stats(h, oversample, &p, &q, seconds);
        *products = p;
        *plans = q;
    });
}

int dmhm_hmat_to_dense(dmhm_hmat h, double* out, int ld)
{
    return guard([&] {
        require_ready();
        if (!h || !out) throw Error(DMHM_EINVAL, "null pointer");
        if (ld < h->geo().n) throw Error(DMHM_EINVAL, "ld too small");
        if (h->prec == DMHM_F64) to_dense_t<double>(h, out, ld);
        else to_dense_t<float>(h, out, ld);
    });
}

// Packed form: a self-describing little-endian byte image.
//   char[8] "DMHMAMD1"; int32 dim, sizes[3], numLevels, maxRank, strong,
//   precision, full; int64 nLowRank, denseElems; int32 caps[nLowRank];
//   int32 ranks[nLowRank]; dense leaves (denseElems values, tree order);
//   then per admissible block U (m x rank) and V (n x rank), column-major.
// Only the first `rank` columns of each factor are stored.
static const char kPackMagic[8] = {'D', 'M', 'H', 'M', 'A', 'M', 'D', '1'};

struct PackHeader {
    char magic[8];
    int32_t dim, sizes[3], numLevels, maxRank, strong, precision, full;
    int64_t nLowRank, denseElems;
};

static void read_back(dmhm_hmat_s* h, std::vector<char>& vals, std::vector<int>& ranks)
{
    const Layout& L = *h->lay;
    HIPCHECK(hipStreamSynchronize(g_lib.stream));
    vals.resize(L.totalElems() * h->esize());
    ranks.resize(L.nLowRank());
    if (!vals.empty()) HIPCHECK(copy_sync(vals.data(), h->vals, vals.size(), hipMemcpyDeviceToHost));
    if (!ranks.empty()) HIPCHECK(copy_sync(ranks.data(), h->ranks, ranks.size() * sizeof(int), hipMemcpyDeviceToHost));
}

static uint64_t packed_size(dmhm_hmat_s* h, const std::vector<int>& ranks)
{
    const Geometry& g = h->geo();
    const Layout& L = *h->lay;
    uint64_t elems = L.denseElems;
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        const int b = g.lrBlocks[i];
        elems += (uint64_t)(g.tbox(b).size + g.sbox(b).size) * std::min(ranks[i], L.cap[i]);
    }
    return sizeof(PackHeader) + 2 * sizeof(int32_t) * L.nLowRank() + elems * h->esize();
}

static uint64_t pack_hmat(dmhm_hmat_s* h, char* out, uint64_t cap)
{
    const Geometry& g = h->geo();
    const Layout& L = *h->lay;
    std::vector<char> vals;
    std::vector<int> ranks;
    read_back(h, vals, ranks);
    const uint64_t need = packed_size(h, ranks);
    if (!out) return need;
    if (cap < need) throw Error(DMHM_EINVAL, "pack buffer too small");
    PackHeader hd{};
    std::memcpy(hd.magic, kPackMagic, 8);
    hd.dim = g.dim;
    for (int a = 0; a < 3; ++a) hd.sizes[a] = g.sizes[a];
    hd.numLevels = g.numLevels; hd.maxRank = h->maxRank; hd.strong = g.strong ? 1 : 0;
    hd.precision = h->prec; hd.full = h->full ? 1 : 0;
    hd.nLowRank = L.nLowRank(); hd.denseElems = L.denseElems;
    char* p = out;
    std::memcpy(p, &hd, sizeof(hd)); p += sizeof(hd);
    std::memcpy(p, L.cap.data(), sizeof(int32_t) * L.nLowRank()); p += sizeof(int32_t) * L.nLowRank();
    std::memcpy(p, ranks.data(), sizeof(int32_t) * L.nLowRank()); p += sizeof(int32_t) * L.nLowRank();
    const size_t es = h->esize();
    std::memcpy(p, vals.data(), L.denseElems * es); p += L.denseElems * es;
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        const int b = g.lrBlocks[i];
        const int r = std::min(ranks[i], L.cap[i]);
        const size_t u = (size_t)g.tbox(b).size * r * es, v = (size_t)g.sbox(b).size * r * es;
        std::memcpy(p, vals.data() + L.uOff[i] * es, u); p += u;
        std::memcpy(p, vals.data() + L.vOff[i] * es, v); p += v;
    }
    return (uint64_t)(p - out);
}

static dmhm_hmat_s* unpack_hmat(const char* in, uint64_t size)
{
    PackHeader hd;
    if (size < sizeof(hd)) throw Error(DMHM_EINVAL, "packed image too small");
    std::memcpy(&hd, in, sizeof(hd));
    if (std::memcmp(hd.magic, kPackMagic, 8) != 0) throw Error(DMHM_EINVAL, "not a packed dmhm_amd H-matrix");
    if (hd.precision != DMHM_F64 && hd.precision != DMHM_F32) throw Error(DMHM_EINVAL, "bad precision in image");
    auto gc = get_geo(hd.dim, hd.sizes, hd.numLevels, hd.strong != 0);
    const Geometry& g = gc->geo;
    if (hd.nLowRank != (int64_t)g.lrBlocks.size()) throw Error(DMHM_EINVAL, "image does not match its geometry");
    const char* p = in + sizeof(hd);
    const char* end = in + size;
    auto take = [&](size_t n) {
        if ((size_t)(end - p) < n) throw Error(DMHM_EINVAL, "packed image truncated");
        const char* q = p; p += n; return q;
    };
    std::vector<int> caps(hd.nLowRank), ranks(hd.nLowRank);
    std::memcpy(caps.data(), take(sizeof(int32_t) * hd.nLowRank), sizeof(int32_t) * hd.nLowRank);
    std::memcpy(ranks.data(), take(sizeof(int32_t) * hd.nLowRank), sizeof(int32_t) * hd.nLowRank);
    bool stdCaps = true;
    for (int c : caps) stdCaps = stdCaps && c == std::max(hd.maxRank, 1);
    std::shared_ptr<Layout> lay;
    if (stdCaps) lay = std_layout(*gc, hd.maxRank);
    else {
        std::lock_guard<std::mutex> lk(g_lib.mu);
        lay = std::make_shared<Layout>(Layout::make(g, caps, g_lib.nextLayoutId++));
    }
    if (lay->denseElems != hd.denseElems) throw Error(DMHM_EINVAL, "image does not match its geometry");
    std::unique_ptr<dmhm_hmat_s> h(alloc_hmat(gc, lay, hd.precision, hd.maxRank));
    const size_t es = h->esize();
    std::vector<char> vals(lay->totalElems() * es, 0);
    std::memcpy(vals.data(), take(lay->denseElems * es), lay->denseElems * es);
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        const int b = g.lrBlocks[i];
        if (ranks[i] < 0 || ranks[i] > caps[i]) throw Error(DMHM_EINVAL, "bad rank in image");
        const size_t u = (size_t)g.tbox(b).size * ranks[i] * es, v = (size_t)g.sbox(b).size * ranks[i] * es;
        std::memcpy(vals.data() + lay->uOff[i] * es, take(u), u);
        std::memcpy(vals.data() + lay->vOff[i] * es, take(v), v);
    }
    HIPCHECK(copy_sync(h->vals, vals.data(), vals.size(), hipMemcpyHostToDevice));
    if (!ranks.empty()) HIPCHECK(copy_sync(h->ranks, ranks.data(), ranks.size() * sizeof(int), hipMemcpyHostToDevice));
    h->full = hd.full != 0;
    return h.release();
}

// Length of a packed image, from its header and rank table (the reference's
// Unpack(const byte*) reads an image of implicit length).
static uint64_t image_size(const char* in)
{
    PackHeader hd;
    std::memcpy(&hd, in, sizeof(hd));
    if (std::memcmp(hd.magic, kPackMagic, 8) != 0) throw Error(DMHM_EINVAL, "not a packed dmhm_amd H-matrix");
    if (hd.precision != DMHM_F64 && hd.precision != DMHM_F32) throw Error(DMHM_EINVAL, "bad precision in image");
    auto gc = get_geo(hd.dim, hd.sizes, hd.numLevels, hd.strong != 0);
    const Geometry& g = gc->geo;
    if (hd.nLowRank != (int64_t)g.lrBlocks.size() || hd.denseElems < 0)
        throw Error(DMHM_EINVAL, "image does not match its geometry");
    const char* rk = in + sizeof(hd) + sizeof(int32_t) * hd.nLowRank;
    uint64_t elems = (uint64_t)hd.denseElems;
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        int32_t r = 0;
        std::memcpy(&r, rk + sizeof(int32_t) * i, sizeof(r));
        if (r < 0) throw Error(DMHM_EINVAL, "bad rank in image");
        const int b = g.lrBlocks[i];
        elems += (uint64_t)(g.tbox(b).size + g.sbox(b).size) * r;
    }
    const uint64_t es = hd.precision == DMHM_F64 ? 8 : 4;
    return sizeof(hd) + 2 * sizeof(int32_t) * hd.nLowRank + elems * es;
}

int dmhm_hmat_unpack_image(const void* buf, uint64_t* consumed, dmhm_hmat* out)
{
    return guard([&] {
        require_ready();
        if (!buf || !out) throw Error(DMHM_EINVAL, "null pointer");
        const uint64_t n = image_size(static_cast<const char*>(buf));
        *out = unpack_hmat(static_cast<const char*>(buf), n);
        if (consumed) *consumed = n;
    });
}

int dmhm_hmat_packed_size(dmhm_hmat h, uint64_t* bytes)
{
    return guard([&] {
        require_ready();
        if (!h || !bytes) throw Error(DMHM_EINVAL, "null pointer");
        *bytes = pack_hmat(h, nullptr, 0);
    });
}

int dmhm_hmat_pack(dmhm_hmat h, void* buf, uint64_t capacity, uint64_t* written)
{
    return guard([&] {
        require_ready();
        if (!h || !buf || !written) throw Error(DMHM_EINVAL, "null pointer");
        *written = pack_hmat(h, static_cast<char*>(buf), capacity);
    });
}

int dmhm_hmat_unpack(const void* buf, uint64_t size, dmhm_hmat* out)
{
    return guard([&] {
        require_ready();
        if (!buf || !out) throw Error(DMHM_EINVAL, "null pointer");
        *out = unpack_hmat(static_cast<const char*>(buf), size);
    });
}

int dmhm_hmat_ranks(dmhm_hmat h, int* ranks, int count)
{
    return guard([&] {
        require_ready();
        if (!h || !ranks) throw Error(DMHM_EINVAL, "null pointer");
        const int n = std::min(count, h->lay->nLowRank());
        HIPCHECK(hipStreamSynchronize(g_lib.stream));
        if (n > 0) HIPCHECK(copy_sync(ranks, h->ranks, n * sizeof(int), hipMemcpyDeviceToHost));
    });
}

int dmhm_profile_matvec(dmhm_hmat h, int k, int reps, double* ms, double* bytes)
{
    return guard([&] {
        require_ready();
        if (!h || !ms || !bytes) throw Error(DMHM_EINVAL, "null pointer");
        if (k < 1 || reps < 1) throw Error(DMHM_EINVAL, "k and reps must be positive");
        profile_matvec(h, k, reps, ms, bytes);
    });
}

int dmhm_profile_hh(double alpha, dmhm_hmat A, dmhm_hmat B, const dmhm_hh_opts* opts, double* ms, double* flops)
{
    return guard([&] {
        require_ready();
        if (!A || !B || !ms || !flops) throw Error(DMHM_EINVAL, "null pointer");
        dmhm_hh_opts o{4, 0, 1e-10, 1};
        if (opts) o = *opts;
        profile_hh(alpha, A, B, o, ms, flops);
    });
}

int dmhm_profile_hh_phases(int cap, double* ms, double* flops, int* kinds, int* count)
{
    return guard([&] {
        if (!count) throw Error(DMHM_EINVAL, "null pointer");
        *count = (int)g_lastPhases.size();
        for (int i = 0; i < cap && i < (int)g_lastPhases.size(); ++i) {
            if (ms) ms[i] = g_lastPhases[i].ms;
            if (flops) flops[i] = g_lastPhases[i].flops;
            if (kinds) kinds[i] = g_lastPhases[i].kind;
        }
    });
}

int dmhm_hmat_matvec_bytes(dmhm_hmat h, int k, double* bytes, double* flops)
{
    return guard([&] {
        require_ready();
        if (!h) throw Error(DMHM_EINVAL, "null handle");
        matvec_bytes(h, k, bytes, flops);
    });
}

} // extern "C"

// Plan upload/release (kept here so plan.cpp stays host-only C++).
namespace dmhm {
void Plan::upload()
{
    for (auto& ph : phases) {
        if (!ph.segSrc.empty()) {
            // Segment table of a merged tile phase: the members' source
            // arrays must already be on the device.
            std::vector<TileSeg> segs;
            int b1 = 0, b4 = 0;
            for (const auto& s : ph.segSrc) {
                const int n = (int)s.src->tiles.size();
                if (n > 0 && (!s.src->dTiles || !s.src->dContribs))
                    throw Error(DMHM_EINVAL, "segment source plan not uploaded");
                TileSeg t{};
                t.tiles = static_cast<const Tile*>(s.src->dTiles);
                t.cs = static_cast<const Contrib*>(s.src->dContribs);
                for (int k = 0; k < NUM_SLOTS; ++k) { t.val[k] = s.val[k]; t.rank[k] = s.rank[k]; }
                t.count = n;
                t.first1 = b1;
                t.first4 = b4;
                b1 += n;
                b4 += (n + 3) / 4;
                segs.push_back(t);
            }
            ph.segBlocks1 = b1;
            ph.segBlocks4 = b4;
            HIPCHECK(hipMalloc(&ph.dSegs, segs.size() * sizeof(TileSeg)));
            HIPCHECK(copy_sync(ph.dSegs, segs.data(), segs.size() * sizeof(TileSeg), hipMemcpyHostToDevice));
        }
        if (!ph.tiles.empty()) {
            HIPCHECK(hipMalloc(&ph.dTiles, ph.tiles.size() * sizeof(Tile)));
            HIPCHECK(copy_sync(ph.dTiles, ph.tiles.data(), ph.tiles.size() * sizeof(Tile), hipMemcpyHostToDevice));
        }
        if (!ph.contribs.empty()) {
            HIPCHECK(hipMalloc(&ph.dContribs, ph.contribs.size() * sizeof(Contrib)));
            HIPCHECK(copy_sync(ph.dContribs, ph.contribs.data(), ph.contribs.size() * sizeof(Contrib),
                               hipMemcpyHostToDevice));
        }
        const void* src = nullptr; size_t bytes = 0;
        if (ph.kind == PH_QR && !ph.qr.empty()) {
            // Sort the (independent) tasks by panel class: register-resident
            // kernels for m <= 64 / 256 / 512 / 1024 with k <= 40, then the rest.
            auto cls = [](const QrTask& t) {
                if (t.k > 40) return 4;
                if (t.m <= 64) return 0;
                if (t.m <= 256) return 1;
                if (t.m <= 512) return 2;
                if (t.m <= 1024 && t.k <= 24) return 3;
                if (t.ld != t.m || t.rld != t.k) throw Error(DMHM_EINVAL, "strided QR panel outside the register classes");
                return 4;
            };
            std::stable_sort(ph.qr.begin(), ph.qr.end(),
                             [&](const QrTask& x, const QrTask& y) { return cls(x) < cls(y); });
            for (int c = 0; c < 4; ++c) ph.qrClassEnd[c] = 0;
            ph.qrRegK = 0;
            int pos = 0;
            for (int c = 0; c < 4; ++c) {
                while (pos < (int)ph.qr.size() && cls(ph.qr[pos]) == c) {
                    if (c < 3) ph.qrRegK = std::max(ph.qrRegK, ph.qr[pos].k);
                    ++pos;
                }
                ph.qrClassEnd[c] = pos;
            }
            src = ph.qr.data(); bytes = ph.qr.size() * sizeof(QrTask);
        }
        if (ph.kind == PH_SVD && !ph.svd.empty()) {
            ph.svdK = 0;
            for (const auto& t : ph.svd) ph.svdK = std::max(ph.svdK, t.k);
            src = ph.svd.data(); bytes = ph.svd.size() * sizeof(SvdTask);
        }
        if (ph.kind == PH_RNG && !ph.rng.empty()) { src = ph.rng.data(); bytes = ph.rng.size() * sizeof(RngTask); }
        if (ph.kind == PH_EYE && !ph.eyes.empty()) { src = ph.eyes.data(); bytes = ph.eyes.size() * sizeof(int64_t); }
        if (ph.kind == PH_MV_PROJ && !ph.mvProj.empty()) {
            src = ph.mvProj.data(); bytes = ph.mvProj.size() * sizeof(MvProjItem);
        }
        if (ph.kind == PH_MV_ACC && !ph.mvAtoms.empty()) {
            src = ph.mvAtoms.data(); bytes = ph.mvAtoms.size() * sizeof(MvAtom);
            HIPCHECK(hipMalloc(&ph.dAux, std::max<size_t>(ph.mvContribs.size(), 1) * sizeof(MvContrib)));
            if (!ph.mvContribs.empty())
                HIPCHECK(copy_sync(ph.dAux, ph.mvContribs.data(), ph.mvContribs.size() * sizeof(MvContrib),
                                   hipMemcpyHostToDevice));
            std::vector<MvContrib>().swap(ph.mvContribs);
        }
        if (src) {
            HIPCHECK(hipMalloc(&ph.dTasks, bytes));
            HIPCHECK(copy_sync(ph.dTasks, src, bytes, hipMemcpyHostToDevice));
        }
        // Host copies of the big arrays are no longer needed (unless merged
        // plans will be checked against them).
        if (!keepHost) std::vector<Contrib>().swap(ph.contribs);
    }
}

void Plan::release()
{
    for (auto& ph : phases) {
        if (ph.dTiles) (void)hipFree(ph.dTiles);
        if (ph.dContribs) (void)hipFree(ph.dContribs);
        if (ph.dTasks) (void)hipFree(ph.dTasks);
        if (ph.dAux) (void)hipFree(ph.dAux);
        if (ph.dSegs) (void)hipFree(ph.dSegs);
        ph.dTiles = ph.dContribs = ph.dTasks = ph.dAux = ph.dSegs = nullptr;
    }
}
} // namespace dmhm
