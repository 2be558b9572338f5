// Recursive direct inversion of an H-matrix (HMat2d::DirectInvert,
// src/HMat2d/Invert-incl.hpp:17-125; HMat3d::DirectInvert, src/HMat3d/
// Invert-incl.hpp) lowered to a replayable program.
//
// The reference inverts a NODE block B (children B_ij, nc x nc of them) by
// Gauss-Jordan elimination into A, which starts as the identity:
//   for l = 0..nc-1:
//     A_ll := inv(B_ll)                                  (recursively)
//     A_lj := A_ll A_lj (j < l),   B_lj := A_ll B_lj (j > l)
//     for i > l:  A_ij -= B_il A_lj (j <= l),  B_ij -= B_il B_lj (j > l)
//   for l = nc-1..1, i < l, all j:  A_ij -= B_il A_lj
// with dense leaves inverted by LAPACK and every product an H*H product
// truncated to maxRank.
//
// Here three H-matrices of the standard layout share one allocation (the
// arena): X (the inverse being formed, A above), W (the working copy, B above)
// and K (copies of the blocks a step overwrites). The recursion needs no
// copies: inverting the diagonal child c of a node works on W's own block c,
// which the parent no longer reads once it is inverted. Every step's
// products are independent (distinct outputs, inputs not written in the
// step), so each step is one merged plan: sub-matrix product plans (roots
// set to the three child blocks) concatenated phase by phase, one launch per
// phase for all of them. A subtree of the block tree is one contiguous range
// of each storage pool, so zeroing and copying sub-matrices are range
// launches. The program is built once per geometry, rank and oversampling
// and contains only launches (no host round trip until the final pivot
// check).
#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <cstring>

#include "runtime.hpp"

namespace dmhm {

struct DinvStep {
    enum Kind { RANGES, PLAN, INVERT } kind = RANGES;
    // RANGES: value ranges then rank ranges (offsets into the range table).
    bool fill = false;
    int64_t valAt = 0, rankAt = 0;
    int nVal = 0, nRank = 0;
    int64_t maxVal = 0, maxRankLen = 0;
    // PLAN: a merged plan and its run-time scalars.
    std::unique_ptr<Plan> plan;
    double alpha = 1, beta = 0;
    int seq = 0;
    // INVERT: dense leaf of W at src -> X at dst (arena element offsets).
    int64_t src = 0, dst = 0;
    int m = 0;
};

struct DinvProgram {
    int64_t memberElems = 0;        // values per arena member
    int64_t memberRanks = 0;        // ranks per arena member
    int64_t wsElems = 0;
    std::vector<DinvStep> steps;
    std::vector<CopyRange> ranges;
    DeviceBuffer dRanges;
    DeviceBuffer arenaV, arenaR, flag;   // the arena (X, W, K) and the singular flag, kept between calls
    long products = 0, plans = 0;   // sub-matrix products, merged plans
    double buildSeconds = 0;
    ~DinvProgram()
    {
        dRanges.release();
        arenaV.release();
        arenaR.release();
        flag.release();
    }
};

namespace {

enum Member { MX = 0, MW = 1, MK = 2 };

// Storage ranges of the subtree of block b in layout L.
struct Span {
    int64_t d0 = 0, d1 = 0, u0 = 0, u1 = 0, v0 = 0, v1 = 0;
    int r0 = 0, r1 = 0;
};

Span span_of(const Geometry& g, const Layout& L, int b)
{
    const Block& B = g.blocks[b];
    int dLo = INT_MAX, dHi = -1, lLo = INT_MAX, lHi = -1;
    for (int li = B.leafBegin; li < B.leafEnd; ++li) {
        const Block& x = g.blocks[g.leaves[li]];
        if (x.type == BT_DENSE) { dLo = std::min(dLo, x.denseIndex); dHi = std::max(dHi, x.denseIndex); }
        else { lLo = std::min(lLo, x.lrIndex); lHi = std::max(lHi, x.lrIndex); }
    }
    Span s;
    if (dHi >= 0) {
        const int lastBlock = g.denseBlocks[dHi];
        s.d0 = L.denseOff[dLo];
        s.d1 = L.denseOff[dHi] + (int64_t)g.tbox(lastBlock).size * g.sbox(lastBlock).size;
    }
    if (lHi >= 0) {
        const int lastBlock = g.lrBlocks[lHi];
        s.u0 = L.uOff[lLo];
        s.u1 = L.uOff[lHi] + (int64_t)g.tbox(lastBlock).size * L.cap[lHi];
        s.v0 = L.vOff[lLo];
        s.v1 = L.vOff[lHi] + (int64_t)g.sbox(lastBlock).size * L.cap[lHi];
        s.r0 = lLo;
        s.r1 = lHi + 1;
    }
    return s;
}

struct Prod {
    int bA, bB, bC;     // blocks of the left operand, right operand and output
    int mA, mB, mC;     // their arena members
};

class Builder {
public:
    Builder(const Geometry& g, const Layout& L, int maxRank, int over, DinvProgram& P)
        : g_(g), L_(L), maxRank_(maxRank), over_(over), P_(P) {}

    void invert(int b)
    {
        const Block& B = g_.blocks[b];
        if (B.type == BT_DENSE) {
            DinvStep st;
            st.kind = DinvStep::INVERT;
            st.src = val(MW) + L_.denseOff[B.denseIndex];
            st.dst = val(MX) + L_.denseOff[B.denseIndex];
            st.m = g_.tbox(b).size;
            P_.steps.push_back(std::move(st));
            return;
        }
        if (B.type != BT_NODE) throw Error(DMHM_EINVAL, "Cannot invert low-rank matrices");
        const int nc = g_.nc();
        auto ch = [&](int i, int j) { return g_.child(b, i, j); };
        // X := 0 on the node (its diagonal children are then written by the
        // recursion; the identity's off-diagonal blocks are zero).
        ranges({{MX, MX, b}}, true, true);
        for (int l = 0; l < nc; ++l) {
            invert(ch(l, l));
            // Row l scaled by the new pivot inverse X_ll; the right operands
            // are copies in K of the blocks being overwritten.
            std::vector<std::pair<int, int>> cp;        // (member, block) copied to K
            std::vector<Prod> pr;
            for (int j = 0; j < nc; ++j) {
                if (j == l) continue;
                const int m = j < l ? MX : MW;
                cp.push_back({m, ch(l, j)});
                pr.push_back({ch(l, l), ch(l, j), ch(l, j), MX, MK, m});
            }
            copies_to_k(cp, true);
            plan(pr, Planner::HH_PRODUCT, 1.0, 0.0);
            if (l == nc - 1) break;
            // Elimination below the pivot (update form, C_old from K).
            cp.clear();
            pr.clear();
            for (int i = l + 1; i < nc; ++i)
                for (int j = 0; j < nc; ++j) {
                    const int m = j <= l ? MX : MW;
                    cp.push_back({m, ch(i, j)});
                    pr.push_back({ch(i, l), ch(l, j), ch(i, j), MW, m, m});
                }
            copies_to_k(cp, false);
            plan(pr, Planner::HH_UPDATE, -1.0, 1.0);
        }
        // Back substitution.
        for (int l = nc - 1; l >= 1; --l) {
            std::vector<std::pair<int, int>> cp;
            std::vector<Prod> pr;
            for (int i = l - 1; i >= 0; --i)
                for (int j = 0; j < nc; ++j) {
                    cp.push_back({MX, ch(i, j)});
                    pr.push_back({ch(i, l), ch(l, j), ch(i, j), MW, MX, MX});
                }
            copies_to_k(cp, false);
            plan(pr, Planner::HH_UPDATE, -1.0, 1.0);
        }
    }

private:
    int64_t val(int m) const { return m * P_.memberElems; }
    int64_t rnk(int m) const { return m * P_.memberRanks; }

    // Copies (src member -> dst member) or zero fills (dst member) of whole
    // subtrees: values (the dense pool only if withDense) and ranks.
    void ranges(const std::vector<std::array<int, 3>>& items, bool fill, bool withDense)
    {
        DinvStep st;
        st.kind = DinvStep::RANGES;
        st.fill = fill;
        st.valAt = (int64_t)P_.ranges.size();
        for (const auto& it : items) {
            const Span s = span_of(g_, L_, it[2]);
            const int64_t so = val(it[0]), dof = val(it[1]);
            auto add = [&](int64_t a, int64_t e) {
                if (e <= a) return;
                P_.ranges.push_back({so + a, dof + a, e - a});
                st.maxVal = std::max(st.maxVal, e - a);
            };
            if (withDense) add(s.d0, s.d1);
            add(s.u0, s.u1);
            add(s.v0, s.v1);
        }
        st.nVal = (int)((int64_t)P_.ranges.size() - st.valAt);
        st.rankAt = (int64_t)P_.ranges.size();
        for (const auto& it : items) {
            const Span s = span_of(g_, L_, it[2]);
            if (s.r1 <= s.r0) continue;
            P_.ranges.push_back({rnk(it[0]) + s.r0, rnk(it[1]) + s.r0, (int64_t)(s.r1 - s.r0)});
            st.maxRankLen = std::max<int64_t>(st.maxRankLen, s.r1 - s.r0);
        }
        st.nRank = (int)((int64_t)P_.ranges.size() - st.rankAt);
        if (st.nVal + st.nRank > 0) P_.steps.push_back(std::move(st));
    }

    void copies_to_k(const std::vector<std::pair<int, int>>& cp, bool withDense)
    {
        std::vector<std::array<int, 3>> items;
        for (const auto& c : cp) items.push_back({c.first, MK, c.second});
        ranges(items, false, withDense);
    }

    void plan(const std::vector<Prod>& prods, int mode, double alpha, double beta)
    {
        if (prods.empty()) return;
        std::vector<std::unique_ptr<Plan>> subs;
        std::vector<PlanPart> parts;
        const int64_t budget = std::max<int64_t>(hh_batch_budget() / (int64_t)prods.size(), int64_t(1) << 24);
        for (const Prod& p : prods) {
            Planner pl(g_);
            pl.setRoots(p.bA, p.bB, p.bC);
            pl.setStacking(false, false);
            pl.setWorkspaceBudget(budget);
            subs.emplace_back(pl.hhProduct(L_, L_, L_, maxRank_, over_, mode));
            PlanPart part;
            part.plan = subs.back().get();
            part.val[SLOT_A] = val(p.mA); part.rank[SLOT_A] = rnk(p.mA);
            part.val[SLOT_B] = val(p.mB); part.rank[SLOT_B] = rnk(p.mB);
            part.val[SLOT_C] = val(p.mC); part.rank[SLOT_C] = rnk(p.mC);
            part.val[SLOT_X] = val(MK); part.rank[SLOT_X] = rnk(MK);
            parts.push_back(part);
        }
        DinvStep st;
        st.kind = DinvStep::PLAN;
        st.plan.reset(merge_plans(parts));
        subs.clear();
        st.alpha = alpha;
        st.beta = beta;
        st.seq = (int)P_.plans;
        P_.wsElems = std::max(P_.wsElems, st.plan->wsElems);
        P_.products += (long)prods.size();
        P_.plans += 1;
        P_.steps.push_back(std::move(st));
    }

    const Geometry& g_;
    const Layout& L_;
    int maxRank_, over_;
    DinvProgram& P_;
};

// Lowers the inversion of the whole matrix (host only; plans not uploaded).
std::shared_ptr<DinvProgram> build_program(const Geometry& g, const Layout& L, int maxRank, int over)
{
    const auto t0 = std::chrono::steady_clock::now();
    auto P = std::make_shared<DinvProgram>();
    P->memberElems = (L.totalElems() + 63) & ~int64_t(63);
    P->memberRanks = (L.nLowRank() + 63) & ~int64_t(63);
    Builder(g, L, maxRank, over, *P).invert(g.root());
    P->buildSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return P;
}

// Every task of every step inside the arena and the workspace (host check).
void check_program(const DinvProgram& P)
{
    const int64_t arena = 3 * P.memberElems, arenaRanks = 3 * P.memberRanks;
    int64_t slots[NUM_SLOTS] = {arena, arena, arena, 0, arena, 0, 0, 0};
    int64_t ranks[NUM_SLOTS] = {arenaRanks, arenaRanks, arenaRanks, 0, arenaRanks, 0, 0, 0};
    for (size_t i = 0; i < P.steps.size(); ++i) {
        const DinvStep& st = P.steps[i];
        const std::string at = "direct-inversion step " + std::to_string(i) + ": ";
        if (st.kind == DinvStep::PLAN) {
            const std::string e = check_plan(*st.plan, slots, ranks);
            if (!e.empty()) throw Error(DMHM_EINVAL, at + e);
        } else if (st.kind == DinvStep::RANGES) {
            for (int r = 0; r < st.nVal + st.nRank; ++r) {
                const CopyRange& c = P.ranges[st.valAt + r];
                const int64_t lim = r < st.nVal ? arena : arenaRanks;
                if (c.len <= 0 || c.src < 0 || c.dst < 0 || c.src + c.len > lim || c.dst + c.len > lim)
                    throw Error(DMHM_EINVAL, at + "range out of bounds");
            }
        } else if (st.src < 0 || st.dst < 0 || st.m <= 0 || st.src + (int64_t)st.m * st.m > arena ||
                   st.dst + (int64_t)st.m * st.m > arena) {
            throw Error(DMHM_EINVAL, at + "dense leaf out of bounds");
        }
    }
}

std::shared_ptr<DinvProgram> get_program(GeoCtx& gc, const Layout& L, int maxRank, int over)
{
    const std::string key = std::to_string(L.id) + ":" + std::to_string(maxRank) + ":" + std::to_string(over);
    auto it = gc.dinvPrograms.find(key);
    if (it != gc.dinvPrograms.end()) return it->second;
    auto P = build_program(gc.geo, L, maxRank, over);
    check_program(*P);
    for (DinvStep& st : P->steps)
        if (st.kind == DinvStep::PLAN) st.plan->upload();
    if (!P->ranges.empty()) {
        P->dRanges.ensure(P->ranges.size() * sizeof(CopyRange));
        HIPCHECK(copy_sync(P->dRanges.ptr, P->ranges.data(), P->ranges.size() * sizeof(CopyRange),
                           hipMemcpyHostToDevice));
    }
    gc.dinvPrograms[key] = P;
    return P;
}

template <typename T>
void run_program(DinvProgram& P, T* vals, int* ranks, int* singular, const dmhm_hh_opts& o)
{
    hipStream_t s = g_lib.stream;
    const CopyRange* tab = static_cast<const CopyRange*>(P.dRanges.ptr);
    Bases b{};
    for (int sl : {SLOT_A, SLOT_B, SLOT_C, SLOT_X}) { b.val[sl] = vals; b.rank[sl] = ranks; }
    b.val[SLOT_WS] = g_lib.ws.ptr;
    const TruncPolicy pol{o.tolMode, o.relTol};
    for (DinvStep& st : P.steps) {
        switch (st.kind) {
        case DinvStep::RANGES:
            // The rank ranges follow the value ranges in the table.
            launch_arena_ranges<T>(vals, ranks, tab + st.valAt, st.nVal, st.nRank,
                                   std::max(st.maxVal, st.maxRankLen), st.fill, s);
            break;
        case DinvStep::PLAN:
            // Each merged plan draws its own sketches.
            run_plan<T>(*st.plan, b, (T)st.alpha, (T)st.beta, pol, o.seed + 0x9E3779B97F4A7C15ull * (uint64_t)(st.seq + 1));
            break;
        case DinvStep::INVERT:
            launch_dense_inverse<T>(vals + st.src, vals + st.dst, st.m, singular, s);
            break;
        }
    }
}

} // namespace

// A := inv(A) (in A's precision). The result takes the standard layout.
void direct_invert(dmhm_hmat_s* A, const dmhm_hh_opts& o)
{
    if (A->geo().n == 0) throw Error(DMHM_EINVAL, "empty matrix");
    GeoCtx& gc = *A->gc;
    const int maxRank = A->maxRank;
    const int over = o.oversample > 0 ? o.oversample : 4;
    if (maxRank + over > 64) throw Error(DMHM_EINVAL, "maxRank + oversample must be <= 64");
    auto LS = std_layout(gc, maxRank);
    // The working copy W in the standard layout (imports are re-laid out by
    // the blockwise sum W = A + 0 A, which recompresses every admissible block).
    std::unique_ptr<dmhm_hmat_s> Wsrc;
    dmhm_hmat_s* W = A;
    if (A->lay.get() != LS.get()) {
        Wsrc.reset(copy_hmat(A, A->prec));
        update_with(Wsrc.get(), 0.0, A, o);
        W = Wsrc.get();
    }
    auto P = get_program(gc, *LS, maxRank, over);
    const size_t es = A->esize();
    g_lib.ws.ensure(std::max<size_t>(P->wsElems * es, 256));
    hipStream_t s = g_lib.stream;
    // The arena stays with the program (every call on this geometry reuses
    // it; calls run one after another on the library stream).
    P->arenaV.ensure(3 * P->memberElems * es);
    P->arenaR.ensure(3 * P->memberRanks * sizeof(int));
    P->flag.ensure(sizeof(int));
    char* av = static_cast<char*>(P->arenaV.ptr);
    int* ar = static_cast<int*>(P->arenaR.ptr);
    HIPCHECK(hipMemcpyAsync(av + MW * P->memberElems * es, W->vals, LS->totalElems() * es, hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipMemcpyAsync(ar + MW * P->memberRanks, W->ranks, LS->nLowRank() * sizeof(int), hipMemcpyDeviceToDevice, s));
    int* sing = static_cast<int*>(P->flag.ptr);
    launch_fill_int(sing, 1, 0, s);
    if (A->prec == DMHM_F64) run_program<double>(*P, reinterpret_cast<double*>(av), ar, sing, o);
    else run_program<float>(*P, reinterpret_cast<float*>(av), ar, sing, o);
    std::unique_ptr<dmhm_hmat_s> R(alloc_hmat(A->gc, LS, A->prec, maxRank));
    HIPCHECK(hipMemcpyAsync(R->vals, av + MX * P->memberElems * es, LS->totalElems() * es, hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipMemcpyAsync(R->ranks, ar + MX * P->memberRanks, LS->nLowRank() * sizeof(int), hipMemcpyDeviceToDevice, s));
    int hostSing = 0;
    HIPCHECK(copy_sync(&hostSing, sing, sizeof(int), hipMemcpyDeviceToHost));
    if (hostSing) throw Error(DMHM_ESINGULAR, "singular pivot block in the direct inversion");
    R->stats = A->stats;
    A->full = true;
    std::swap(A->lay, R->lay);
    std::swap(A->vals, R->vals);
    std::swap(A->ranks, R->ranks);
}

// Host-only lowering and bounds check of the program of a geometry at a
// rank (standard layout); stats: sub-matrix products, merged plans, steps,
// dense leaf inversions, workspace elements, plan phases (launches), identity
// phases among them.
void check_direct_invert_program(const Geometry& g, int maxRank, int over, long long* stats)
{
    std::vector<int> caps(g.lrBlocks.size(), std::max(maxRank, 1));
    const Layout L = Layout::make(g, caps, 0);
    auto P = build_program(g, L, maxRank, over);
    check_program(*P);
    long long inv = 0, phases = 0, eyes = 0;
    for (const DinvStep& st : P->steps) {
        inv += st.kind == DinvStep::INVERT;
        if (st.kind == DinvStep::PLAN) {
            phases += (long long)st.plan->phases.size();
            for (const Phase& ph : st.plan->phases) eyes += ph.kind == PH_EYE;
        }
    }
    stats[0] = P->products;
    stats[1] = P->plans;
    stats[2] = (long long)P->steps.size();
    stats[3] = inv;
    stats[4] = P->wsElems;
    stats[5] = phases;
    stats[6] = eyes;
}

// Statistics of the cached program (products, merged plans, build seconds).
void direct_invert_stats(dmhm_hmat_s* A, int oversample, long* products, long* plans, double* seconds)
{
    auto LS = std_layout(*A->gc, A->maxRank);
    auto P = get_program(*A->gc, *LS, A->maxRank, oversample > 0 ? oversample : 4);
    *products = P->products;
    *plans = P->plans;
    *seconds = P->buildSeconds;
}

} // namespace dmhm
