#include "plan.hpp"
#include <algorithm>
#include <climits>
#include <map>
#include <memory>
#include <stdexcept>

namespace dmhm {

void PhaseBuilder::finalize()
{
    // Stable counting sort of the pending contributions by tile id.
    const int nt = (int)ph_.tiles.size();
    std::vector<int> count(nt + 1, 0);
    for (auto& p : pending_) count[p.first + 1]++;
    for (int i = 0; i < nt; ++i) count[i + 1] += count[i];
    const int base = (int)ph_.contribs.size();
    ph_.contribs.resize(base + pending_.size());
    std::vector<int> pos(count.begin(), count.end() - 1);
    for (auto& p : pending_) ph_.contribs[base + pos[p.first]++] = p.second;
    for (int i = 0; i < nt; ++i) {
        Tile& t = ph_.tiles[i];
        t.cb = base + count[i];
        t.ce = base + count[i + 1];
        ph_.maxKTile = std::max(ph_.maxKTile, t.k);
        ph_.maxMTile = std::max(ph_.maxMTile, t.m);
        for (int c = t.cb; c < t.ce; ++c) ph_.flops += 2.0 * t.m * ph_.contribs[c].q * t.k;
    }
    pending_.clear();
    pending_.shrink_to_fit();
}

static Contrib mkc(uint8_t lslot, int64_t loff, int ldl, uint8_t rslot, int64_t roff, int ldr,
                   int q, int qrank, uint8_t qslot, uint8_t flags, float alpha)
{
    Contrib c;
    c.loff = loff; c.roff = roff; c.ldl = ldl; c.ldr = ldr; c.q = q; c.qrank = qrank;
    c.lslot = lslot; c.rslot = rslot; c.qslot = qslot; c.flags = flags; c.alpha = alpha;
    return c;
}

Planner::PanelTiles Planner::addPanelTiles(PhaseBuilder& pb, const Operand& out, int level,
                                           int box, int k, int krank, uint8_t kslot,
                                           uint8_t flags)
{
    const Box& B = g_.boxes[level][box];
    PanelTiles pt;
    pt.level = level; pt.box = box;
    pt.kChunks = (k + 63) / 64;
    pt.tileOfAtom.resize((size_t)(B.atomEnd - B.atomBegin) * pt.kChunks);
    for (int a = B.atomBegin; a < B.atomEnd; ++a)
        for (int c = 0; c < pt.kChunks; ++c) {
            Tile t{};
            t.off = out.off + (g_.atomStart[a] - B.off) + (int64_t)c * 64 * out.ld;
            t.ld = out.ld;
            t.m = g_.atomSize[a];
            t.k = std::min(64, k - 64 * c);
            t.mrank = -1; t.moff = 0;
            t.krank = krank; t.koff = 64 * c; t.kslot = kslot;
            t.slot = out.slot; t.mslot = 0;
            t.flags = flags;
            pt.tileOfAtom[(size_t)(a - B.atomBegin) * pt.kChunks + c] = pb.addTile(t);
        }
    return pt;
}

void Planner::subh(PhaseBuilder& proj, PhaseBuilder& acc, uint8_t mslot, const Layout& LM, int b,
                   bool trans, const Operand& in, int inLevel, int inBox, const PanelTiles& out,
                   int k, int krank, uint8_t kslot, float alpha, uint8_t cflags, double* flops)
{
    const Box& InB = g_.boxes[inLevel][inBox];
    const Box& OutB = g_.boxes[out.level][out.box];
    const Block& blk = g_.blocks[b];
    for (int li = blk.leafBegin; li < blk.leafEnd; ++li) {
        const int lb = g_.leaves[li];
        const Block& L = g_.blocks[lb];
        const Box& tb = g_.tbox(lb);
        const Box& sb = g_.sbox(lb);
        const Box& rowB = trans ? sb : tb;
        const Box& colB = trans ? tb : sb;
        if (colB.off < InB.off || colB.off + colB.size > InB.off + InB.size ||
            rowB.off < OutB.off || rowB.off + rowB.size > OutB.off + OutB.size)
            throw std::logic_error("sub-H product operand outside its panel");
        const int64_t inRow = colB.off - InB.off;
        if (L.type == BT_DENSE) {
            const int64_t D = LM.denseOff[L.denseIndex];
            const int m = tb.size, n = sb.size;
            for (int a = rowB.atomBegin; a < rowB.atomEnd; ++a) {
                const int ra = g_.atomStart[a] - rowB.off;
                for (int c = 0; c < out.kChunks; ++c) {
                    const int tile = out.tileOfAtom[(size_t)(a - OutB.atomBegin) * out.kChunks + c];
                    const int64_t roff = in.off + inRow + (int64_t)c * 64 * in.ld;
                    if (!trans)
                        acc.add(tile, mkc(mslot, D + ra, m, in.slot, roff, in.ld, n, -1, 0, cflags, alpha));
                    else
                        acc.add(tile, mkc(mslot, D + (int64_t)ra * m, m, in.slot, roff, in.ld, m, -1, 0,
                                          cflags | CF_TRANS_L, alpha));
                }
                if (flops) *flops += 2.0 * g_.atomSize[a] * (trans ? m : n) * k;
            }
        } else {
            const int li2 = L.lrIndex;
            const int cap = LM.cap[li2];
            const int64_t U = LM.uOff[li2], V = LM.vOff[li2];
            const int64_t inSide = trans ? U : V;          // rows = colB
            const int inLd = trans ? tb.size : sb.size;
            const int64_t outSide = trans ? V : U;         // rows = rowB
            const int outLd = trans ? sb.size : tb.size;
            // Projection P = inSide^T * in(colB rows) : cap x k (ld cap)
            const int64_t P = wsAlloc((int64_t)cap * k);
            const int kCh = (k + 63) / 64, rCh = (cap + 63) / 64;
            for (int rc = 0; rc < rCh; ++rc)
                for (int c = 0; c < kCh; ++c) {
                    Tile t{};
                    t.off = P + rc * 64 + (int64_t)c * 64 * cap;
                    t.ld = cap;
                    t.m = std::min(64, cap - 64 * rc); t.mrank = li2; t.moff = 64 * rc; t.mslot = mslot;
                    t.k = std::min(64, k - 64 * c); t.krank = krank; t.koff = 64 * c; t.kslot = kslot;
                    t.slot = SLOT_WS; t.flags = 0;
                    const int tile = proj.addTile(t);
                    for (int a = colB.atomBegin; a < colB.atomEnd; ++a) {
                        const int ra = g_.atomStart[a] - colB.off;
                        proj.add(tile, mkc(mslot, inSide + ra + (int64_t)rc * 64 * inLd, inLd, in.slot,
                                           in.off + inRow + ra + (int64_t)c * 64 * in.ld, in.ld,
                                           g_.atomSize[a], -1, 0, CF_TRANS_L, 1.0f));
                    }
                }
            if (flops) *flops += 2.0 * colB.size * cap * k;
            for (int a = rowB.atomBegin; a < rowB.atomEnd; ++a) {
                const int ra = g_.atomStart[a] - rowB.off;
                for (int c = 0; c < out.kChunks; ++c) {
                    const int tile = out.tileOfAtom[(size_t)(a - OutB.atomBegin) * out.kChunks + c];
                    acc.add(tile, mkc(mslot, outSide + ra, outLd, SLOT_WS, P + (int64_t)c * 64 * cap, cap,
                                      cap, li2, mslot, cflags, alpha));
                }
            }
            if (flops) *flops += 2.0 * rowB.size * cap * k;
        }
    }
}

Plan* Planner::matvec(const Layout& L, bool trans, int k, int ldx, int ldy)
{
    Plan* plan = new Plan();
    plan->phases.resize(2);
    plan->phases[0].kind = PH_TILE;
    plan->phases[1].kind = PH_TILE;
    plan->phases[0].role = 1;
    PhaseBuilder proj(plan->phases[0]), acc(plan->phases[1]);
    Operand out{SLOT_Y, 0, ldy}, in{SLOT_X, 0, ldx};
    PanelTiles pt = addPanelTiles(acc, out, 0, 0, k, -1, 0, TF_RUNTIME_BETA);
    subh(proj, acc, SLOT_A, L, g_.root(), trans, in, 0, 0, pt, k, -1, 0, 1.0f, CF_RUNTIME_ALPHA,
         &plan->flops);
    proj.finalize();
    acc.finalize();
    plan->wsElems = wsUsed();
    for (auto& ph : plan->phases) { plan->tileCount += ph.tiles.size(); plan->contribCount += ph.contribs.size(); }
    return plan;
}

Plan* Planner::matvecStream(const Layout& L, int k, int ldx, int ldy)
{
    const Geometry& g = g_;
    Plan* plan = new Plan();
    plan->phases.resize(3);
    Phase& z = plan->phases[0];
    Phase& pr = plan->phases[1];
    Phase& ac = plan->phases[2];
    z.kind = PH_MV_ZERO; pr.kind = PH_MV_PROJ; ac.kind = PH_MV_ACC;
    for (Phase* ph : {&z, &pr, &ac}) { ph->mvK = k; ph->ldx = ldx; ph->ldy = ldy; }
    // P per low-rank block: cap x k.
    std::vector<int64_t> P(g.lrBlocks.size());
    int64_t pTotal = 0;
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        P[i] = wsAlloc((int64_t)L.cap[i] * k);
        pTotal = wsUsed();
    }
    z.zeroOff = 0; z.zeroCount = pTotal;
    const int chunkAtoms = 8;
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        const int b = g.lrBlocks[i];
        const Box& sb = g.sbox(b);
        const int natoms = sb.atomEnd - sb.atomBegin;
        const bool atomic = natoms > chunkAtoms;
        for (int a0 = sb.atomBegin; a0 < sb.atomEnd; a0 += chunkAtoms) {
            const int a1 = std::min(sb.atomEnd, a0 + chunkAtoms);
            const int r0 = g.atomStart[a0] - sb.off;
            const int rows = g.atomStart[a1 - 1] + g.atomSize[a1 - 1] - g.atomStart[a0];
            MvProjItem it{};
            it.voff = L.vOff[i] + r0; it.ldv = sb.size; it.rows = rows;
            it.xrow = sb.off + r0; it.poff = P[i]; it.rankIdx = (int)i; it.cap = L.cap[i];
            it.atomic = atomic ? 1 : 0;
            pr.mvProj.push_back(it);
        }
        plan->flops += 2.0 * sb.size * L.cap[i] * k;
    }
    // Longest chunks first: the grid's tail is then made of short items.
    std::stable_sort(pr.mvProj.begin(), pr.mvProj.end(),
                     [](const MvProjItem& a, const MvProjItem& b) { return a.rows > b.rows; });
    // Contributions per output atom.
    std::vector<std::vector<MvContrib>> per(g.atomStart.size());
    for (int b : g.leaves) {
        const Block& B = g.blocks[b];
        const Box& tb = g.tbox(b);
        const Box& sb = g.sbox(b);
        for (int a = tb.atomBegin; a < tb.atomEnd; ++a) {
            const int ra = g.atomStart[a] - tb.off;
            MvContrib c{};
            if (B.type == BT_DENSE) {
                c.kind = 0; c.loff = L.denseOff[B.denseIndex] + ra; c.ld = tb.size;
                c.q = sb.size; c.rankIdx = -1; c.roff = sb.off; c.ldr = ldx;
                plan->flops += 2.0 * g.atomSize[a] * sb.size * k;
            } else {
                c.kind = 1; c.loff = L.uOff[B.lrIndex] + ra; c.ld = tb.size;
                c.q = L.cap[B.lrIndex]; c.rankIdx = B.lrIndex; c.roff = P[B.lrIndex]; c.ldr = L.cap[B.lrIndex];
                plan->flops += 2.0 * g.atomSize[a] * L.cap[B.lrIndex] * k;
            }
            per[a].push_back(c);
        }
    }
    // Atoms with the most contributions first (same tail argument).
    std::vector<int> order(per.size());
    for (size_t a = 0; a < per.size(); ++a) order[a] = (int)a;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return per[x].size() > per[y].size(); });
    for (int a : order) {
        MvAtom at{};
        at.yrow = g.atomStart[a]; at.m = g.atomSize[a];
        at.cb = (int)ac.mvContribs.size();
        ac.mvContribs.insert(ac.mvContribs.end(), per[a].begin(), per[a].end());
        at.ce = (int)ac.mvContribs.size();
        ac.mvAtoms.push_back(at);
    }
    plan->wsElems = wsUsed();
    plan->tileCount = ac.mvAtoms.size() + pr.mvProj.size();
    plan->contribCount = ac.mvContribs.size();
    return plan;
}

// ---------------------------------------------------------------------------
// H*H product with randomized sum-sketch compression.
// ---------------------------------------------------------------------------
ProductTerms enumerateProductTerms(const Geometry& g, int rootA, int rootB, int rootC)
{
    ProductTerms T;
    T.cBase = rootC;
    const size_t nb = g.blocks[rootC].blockEnd - rootC;
    T.pushdowns.resize(nb); T.lrTerms.resize(nb); T.nnTerms.resize(nb); T.ddTerms.resize(nb);
    auto addW = [&](int a, int b, int c) {
        T.w.push_back({a, b, c, g.blocks[a].type == BT_LOWRANK});
        return (int)T.w.size() - 1;
    };
    struct Tr { int a, b, c; };
    std::vector<Tr> stack{{rootA, rootB, rootC}};
    while (!stack.empty()) {
        const Tr tr = stack.back();
        stack.pop_back();
        const Block& A = g.blocks[tr.a];
        const Block& B = g.blocks[tr.b];
        const Block& C = g.blocks[tr.c];
        if (C.type == BT_NODE) {
            if (A.type == BT_NODE && B.type == BT_NODE) {
                const int nc = g.nc();
                for (int tc = 0; tc < nc; ++tc)
                    for (int uc = 0; uc < nc; ++uc)
                        for (int sc = 0; sc < nc; ++sc)
                            stack.push_back({g.child(tr.a, tc, uc), g.child(tr.b, uc, sc), g.child(tr.c, tc, sc)});
            } else {
                T.pushdowns[tr.c - rootC].push_back(addW(tr.a, tr.b, tr.c));
                T.pushdownTerms++;
            }
        } else if (C.type == BT_LOWRANK) {
            if (A.type == BT_LOWRANK || B.type == BT_LOWRANK) {
                T.lrTerms[tr.c - rootC].push_back(addW(tr.a, tr.b, tr.c));
                T.lowRankTerms++;
            } else {
                T.nnTerms[tr.c - rootC].push_back({tr.a, tr.b});
                if (A.type == BT_DENSE) T.ddToLowRank++;
                else T.sketchTerms++;
            }
        } else {
            if (A.type == BT_DENSE && B.type == BT_DENSE) {
                T.ddTerms[tr.c - rootC].push_back({tr.a, tr.b});
                T.ddToDense++;
            } else {
                T.lrTerms[tr.c - rootC].push_back(addW(tr.a, tr.b, tr.c));
                T.lowRankTerms++;
            }
        }
    }
    return T;
}

namespace {
struct WFactor {            // exact low-rank term W = PU PV^T (times the run-time scalar in cflag)
    Operand PU, PV;
    int level, t, s;        // boxes of PU rows (t) and PV rows (s) at `level`
    int rankIndex;          // rank array entry of the low-rank operand
    uint8_t rankSlot;
    int cap;
    uint8_t cflag = 0;      // CF_RUNTIME_ALPHA / CF_RUNTIME_BETA scaling of the term
};
} // namespace

Plan* Planner::hhAdd(const Layout& LA, const Layout& LB, const Layout& LC, int maxRank, int oversample)
{
    return hhProduct(LA, LB, LC, maxRank, oversample, HH_ADD);
}

Plan* Planner::hhProduct(const Layout& LA, const Layout& LB, const Layout& LC, int maxRank,
                         int oversample, int mode)
{
    const bool update = (mode == HH_UPDATE);
    const bool add = (mode == HH_ADD);
    const Geometry& g = g_;
    const int kS = maxRank + oversample;
    if (kS > 64) throw std::invalid_argument("maxRank + oversample must be <= 64");
    Plan* plan = new Plan();
    // newPhase hands out references into plan->phases: capacity is reserved
    // for every phase once the batches are known, so they stay valid.
    auto newPhase = [&](PhaseKind k) -> Phase& {
        if (plan->phases.size() == plan->phases.capacity())
            throw std::logic_error("phase capacity exhausted");
        plan->phases.emplace_back();
        plan->phases.back().kind = k;
        return plan->phases.back();
    };
    auto key2 = [](int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; };

    // The blocks of C written by this plan: the subtree of rootC (all of C
    // unless sub-matrix roots were set); leaves in DFS order.
    const Block& rootBlk = g.blocks[rootC_];
    std::vector<int> cLr, cDense;
    for (int li = rootBlk.leafBegin; li < rootBlk.leafEnd; ++li) {
        const int b = g.leaves[li];
        (g.blocks[b].type == BT_LOWRANK ? cLr : cDense).push_back(b);
    }
    if (add && rootC_ != 0) throw std::invalid_argument("the blockwise sum takes whole matrices");

    // 1. Enumerate the product recursion; describe every low-rank term.
    ProductTerms terms;
    // Arrays over C's blocks are indexed by b - cBase.
    const int cBase = rootC_;
    const size_t nC = rootBlk.blockEnd - cBase;
    if (!add) terms = enumerateProductTerms(g, rootA_, rootB_, rootC_);
    else {
        terms.pushdowns.resize(nC); terms.lrTerms.resize(nC); terms.nnTerms.resize(nC); terms.ddTerms.resize(nC);
    }
    const auto& pushdowns = terms.pushdowns;
    const auto& lrTerms = terms.lrTerms;
    const auto& nnTerms = terms.nnTerms;
    const auto& ddTerms = terms.ddTerms;
    const uint8_t abFlag = update ? CF_RUNTIME_ALPHA : 0;     // scaling of the A B terms in sketches
    const uint8_t formFlag = (update || add) ? 0 : CF_RUNTIME_ALPHA;    // scaling applied when forming C.U
    std::vector<WFactor> W(terms.w.size());
    struct Pending { int bA, bB, bC; int wIndex; bool left; };
    std::vector<Pending> wPending;
    for (size_t wi = 0; wi < terms.w.size(); ++wi) {
        const auto& tw = terms.w[wi];
        const Block& C = g.blocks[tw.bC];
        WFactor& w = W[wi];
        w.level = C.level; w.t = C.t; w.s = C.s;
        w.cflag = abFlag;
        if (tw.leftLowRank) {
            const int la = g.blocks[tw.bA].lrIndex;
            w.cap = LA.cap[la];
            w.PU = {SLOT_A, LA.uOff[la], g.boxes[C.level][C.t].size};
            w.PV = {SLOT_WS, -1, g.boxes[C.level][C.s].size};      // allocated with its batch
            w.rankIndex = la; w.rankSlot = SLOT_A;
        } else {
            const int lb = g.blocks[tw.bB].lrIndex;
            w.cap = LB.cap[lb];
            w.PU = {SLOT_WS, -1, g.boxes[C.level][C.t].size};
            w.PV = {SLOT_B, LB.vOff[lb], g.boxes[C.level][C.s].size};
            w.rankIndex = lb; w.rankSlot = SLOT_B;
        }
        wPending.push_back({tw.bA, tw.bB, tw.bC, (int)wi, tw.leftLowRank});
    }
    // beta C_old for every admissible C block (update form); A and alpha B for
    // the blockwise sum (HMat2d::UpdateWith).
    std::vector<int> oldTerm(nC, -1), addTerm(nC, -1);
    if (add)
        for (int lb : cLr) {
            const Block& C = g.blocks[lb];
            for (int which = 0; which < 2; ++which) {
                const Layout& LX = which ? LB : LA;
                WFactor w;
                w.level = C.level; w.t = C.t; w.s = C.s;
                w.cap = LX.cap[C.lrIndex];
                const uint8_t slot = which ? SLOT_B : SLOT_A;
                w.PU = {slot, LX.uOff[C.lrIndex], g.boxes[C.level][C.t].size};
                w.PV = {slot, LX.vOff[C.lrIndex], g.boxes[C.level][C.s].size};
                w.rankIndex = C.lrIndex; w.rankSlot = slot; w.cflag = which ? CF_RUNTIME_ALPHA : 0;
                if (which == 0) addTerm[lb - cBase] = (int)W.size();
                W.push_back(w);
            }
        }
    if (update)
        for (int lb : cLr) {
            const Block& C = g.blocks[lb];
            WFactor w;
            w.level = C.level; w.t = C.t; w.s = C.s;
            w.cap = LC.cap[C.lrIndex];
            w.PU = {SLOT_X, LC.uOff[C.lrIndex], g.boxes[C.level][C.t].size};
            w.PV = {SLOT_X, LC.vOff[C.lrIndex], g.boxes[C.level][C.s].size};
            w.rankIndex = C.lrIndex; w.rankSlot = SLOT_X; w.cflag = CF_RUNTIME_BETA;
            oldTerm[lb - cBase] = (int)W.size();
            W.push_back(w);
        }
    // Ancestor pushdowns of every C leaf (ImportLowRank / UpdateWithLowRank
    // distribute a low-rank W over the subtree of a NODE, HMat2d.cpp:811-959).
    std::vector<std::vector<int>> leafWrel(nC);
    for (size_t r = 0; r < nC; ++r) {
        if (pushdowns[r].empty()) continue;
        const Block& blk = g.blocks[cBase + r];
        for (int li = blk.leafBegin; li < blk.leafEnd; ++li) {
            auto& dst = leafWrel[g.leaves[li] - cBase];
            dst.insert(dst.end(), pushdowns[r].begin(), pushdowns[r].end());
        }
    }
    for (size_t r = 0; r < nC; ++r) {
        auto& dst = leafWrel[r];
        dst.insert(dst.end(), lrTerms[r].begin(), lrTerms[r].end());
        if (oldTerm[r] >= 0) dst.push_back(oldTerm[r]);
        if (addTerm[r] >= 0) { dst.push_back(addTerm[r]); dst.push_back(addTerm[r] + 1); }
    }
    auto leafW = [&](int b) -> const std::vector<int>& { return leafWrel[b - cBase]; };
    auto nnOf = [&](int b) -> const std::vector<std::pair<int, int>>& { return nnTerms[b - cBase]; };

    // 2. Batches. Every term lands in one C block. The C blocks are grouped
    //    into units: the subtrees rooted at level `unitLevel`, and the leaves
    //    and admissible blocks above that level. Batches of consecutive units
    //    run one after the other and reuse one workspace region; only the
    //    sketch matrices Omega, T1 = B Omega and the low-rank terms pushed
    //    down from NODE blocks above unitLevel are shared by all batches.
    //    unitLevel starts at 1 and deepens until every unit's estimated
    //    workspace fits the budget.
    const double budget = (double)wsBudget_;
    std::vector<int> unitRel(nC, -1);
    auto unit = [&](int b) -> int& { return unitRel[b - cBase]; };
    int nUnits = 0;
    std::vector<double> est;
    for (int unitLevel = rootBlk.level + 1;; ++unitLevel) {
        std::fill(unitRel.begin(), unitRel.end(), -1);
        nUnits = 0;
        // DFS from C's root: a block at unitLevel (or a leaf above it) opens a unit.
        std::vector<std::pair<int, int>> stack{{rootC_, -1}};
        while (!stack.empty()) {
            const auto [b, u] = stack.back();
            stack.pop_back();
            const Block& B = g.blocks[b];
            int ub = u;
            if (ub < 0 && (B.level >= unitLevel || B.type != BT_NODE)) ub = nUnits++;
            unit(b) = ub;
            if (B.type == BT_NODE)
                for (int c = g.nc() * g.nc() - 1; c >= 0; --c) stack.push_back({g.children[B.childBegin + c], ub});
        }
        // Workspace estimate per unit (elements), used only to group units.
        est.assign(nUnits, 0.0);
        for (int lbk : cLr) {
            const Block& C = g.blocks[lbk];
            const int m = g.boxes[C.level][C.t].size, n = g.boxes[C.level][C.s].size;
            double e = (double)(m + n) * kS + 3.0 * kS * kS;
            if (C.level == g.numLevels - 1) e += (double)m * n;
            e += 2.0 * kS * 64 * leafW(lbk).size();
            for (const auto& tm : nnOf(lbk)) {
                const Block& A = g.blocks[tm.first];
                e += (double)g.boxes[A.level][A.s].size * kS + 2.0 * 64 * kS * (A.leafEnd - A.leafBegin);
            }
            est[unit(lbk)] += e;
        }
        for (const auto& wp : wPending) {
            if (unit(wp.bC) < 0) continue;
            const WFactor& w = W[wp.wIndex];
            const Block& Bl = g.blocks[wp.left ? wp.bB : wp.bA];
            est[unit(wp.bC)] += (double)(wp.left ? w.PV.ld : w.PU.ld) * w.cap +
                                64.0 * kS * (Bl.leafEnd - Bl.leafBegin);
        }
        const double worst = est.empty() ? 0.0 : *std::max_element(est.begin(), est.end());
        if (worst <= budget || unitLevel >= g.numLevels - 1) break;
    }
    std::vector<std::vector<char>> groups;
    {
        double acc = 0;
        for (int i = 0; i < nUnits; ++i) {
            if (groups.empty() || (acc > 0 && acc + est[i] > budget)) {
                groups.emplace_back(nUnits, 0);
                acc = 0;
            }
            groups.back()[i] = 1;
            acc += est[i];
        }
    }
    auto inBatch = [&](const std::vector<char>& grp, int b) { return unit(b) >= 0 && grp[unit(b)]; };

    plan->phases.reserve(4 + 18 * groups.size());
    // Exact factor of a low-rank term: the side that is not a stored factor
    // of A or B is a sub-H product, written to a fresh workspace panel.
    auto buildW = [&](PhaseBuilder& proj, PhaseBuilder& acc, const Pending& wp) {
        WFactor& w = W[wp.wIndex];
        const Block& A = g.blocks[wp.bA];
        const Block& B = g.blocks[wp.bB];
        if (wp.left) {
            // PV (|s| x cap) = B(u,s)^T V_A, V_A rows = u box
            w.PV.off = wsAlloc((int64_t)w.PV.ld * w.cap);
            PanelTiles pt = addPanelTiles(acc, w.PV, w.level, w.s, w.cap, w.rankIndex, w.rankSlot, 0);
            Operand in{SLOT_A, LA.vOff[A.lrIndex], g.boxes[A.level][A.s].size};
            subh(proj, acc, SLOT_B, LB, wp.bB, true, in, A.level, A.s, pt, w.cap, w.rankIndex,
                 w.rankSlot, 1.0f, 0, &plan->flops);
        } else {
            // PU (|t| x cap) = A(t,u) U_B, U_B rows = u box
            w.PU.off = wsAlloc((int64_t)w.PU.ld * w.cap);
            PanelTiles pt = addPanelTiles(acc, w.PU, w.level, w.t, w.cap, w.rankIndex, w.rankSlot, 0);
            Operand in{SLOT_B, LB.uOff[B.lrIndex], g.boxes[B.level][B.t].size};
            subh(proj, acc, SLOT_A, LA, wp.bA, false, in, B.level, B.t, pt, w.cap, w.rankIndex,
                 w.rankSlot, 1.0f, 0, &plan->flops);
        }
    };

    // 3. Shared sketch matrices Omega_s (per level and source box) and
    //    T1 = B(u,s) Omega_s for the NODE*NODE terms of non-leaf blocks.
    std::unordered_map<uint64_t, int64_t> omegaOf;     // (level, s) -> panel
    std::unordered_map<uint64_t, int64_t> t1Of;        // B block -> T1 panel
    // A 64 x 64 identity, the right operand of the gathers below.
    const int64_t eyeOff = wsAlloc(64 * 64);
    bool eyeUsed = false;       // set by the first gather; the phase is dropped otherwise
    {
        Phase& eye = newPhase(PH_EYE);
        eye.zeroOff = eyeOff;
        eye.zeroCount = 64;
    }
    {
        Phase& rngPh = newPhase(PH_RNG);
        for (int lbk : cLr) {
            const Block& C = g.blocks[lbk];
            if (omegaOf.count(key2(C.level, C.s))) continue;
            const int n = g.boxes[C.level][C.s].size;
            const int64_t o = wsAlloc((int64_t)n * kS);
            omegaOf[key2(C.level, C.s)] = o;
            RngTask r{};
            r.off = o; r.m = n; r.k = kS; r.stream = key2(C.level, C.s);
            rngPh.rng.push_back(r);
        }
        Phase& p0 = newPhase(PH_TILE);
        p0.role = 1;
        Phase& a0 = newPhase(PH_TILE);
        PhaseBuilder proj(p0), acc(a0);
        for (const auto& wp : wPending)
            if (unit(wp.bC) < 0) buildW(proj, acc, wp);
        for (int lbk : cLr) {
            const Block& C = g.blocks[lbk];
            if (C.level == g.numLevels - 1) continue;
            for (const auto& tm : nnOf(lbk)) {
                if (t1Of.count((uint64_t)tm.second)) continue;
                const Block& B = g.blocks[tm.second];
                const Box& ub = g.boxes[B.level][B.t];
                const int64_t T1 = wsAlloc((int64_t)ub.size * kS);
                t1Of[(uint64_t)tm.second] = T1;
                PanelTiles pt = addPanelTiles(acc, {SLOT_WS, T1, ub.size}, B.level, B.t, kS, -1, 0, 0);
                Operand in{SLOT_WS, omegaOf[key2(B.level, B.s)], g.boxes[B.level][B.s].size};
                subh(proj, acc, SLOT_B, LB, tm.second, false, in, B.level, B.s, pt, kS, -1, 0, 1.0f, 0,
                     &plan->flops);
            }
        }
        proj.finalize();
        acc.finalize();
        if (p0.tiles.empty() && a0.tiles.empty()) plan->phases.resize(plan->phases.size() - 2);
    }
    const int64_t globalEnd = wsUsed();
    int64_t wsPeak = globalEnd;

    // Low-rank term restricted to C block (t,s): PU rows of t, PV rows of s.
    auto slicePU = [&](const WFactor& w, const Box& tb) {
        return Operand{w.PU.slot, w.PU.off + (tb.off - g.boxes[w.level][w.t].off), w.PU.ld};
    };
    auto slicePV = [&](const WFactor& w, const Box& sbx) {
        return Operand{w.PV.slot, w.PV.off + (sbx.off - g.boxes[w.level][w.s].off), w.PV.ld};
    };
    // Projection tile P = F^T X over the atoms of box `bx` (F rows = bx rows).
    auto projTile = [&](PhaseBuilder& pb, const Operand& F, const Box& bx, const Operand& X,
                        const WFactor& w, int64_t P, int wr) {
        Tile t{};
        t.off = P; t.ld = w.cap;
        t.m = std::min(64, w.cap); t.mrank = w.rankIndex; t.moff = 0; t.mslot = w.rankSlot;
        t.k = kS; t.krank = wr; t.koff = 0; t.kslot = SLOT_WIDTH; t.slot = SLOT_WS; t.flags = 0;
        if (w.cap > 64) throw std::invalid_argument("low-rank capacity above 64 in a product");
        const int tile = pb.addTile(t);
        for (int a = bx.atomBegin; a < bx.atomEnd; ++a) {
            const int ra = g.atomStart[a] - bx.off;
            pb.add(tile, mkc(F.slot, F.off + ra, F.ld, X.slot, X.off + ra, X.ld, g.atomSize[a], -1, 0,
                             CF_TRANS_L, 1.0f));
        }
        plan->flops += 2.0 * bx.size * w.cap * kS;
    };
    // Dense accumulation of a leaf-level C block: out (|t| x |s|) =
    // sum D_A D_B + sum PU PV^T over its dense*dense and low-rank terms.
    auto denseAccumulate = [&](PhaseBuilder& acc, int b, const Operand& out,
                               const std::vector<std::pair<int, int>>& dd, uint8_t cflags, uint8_t tflags) {
        const Block& C = g.blocks[b];
        const Box& tb = g.boxes[C.level][C.t];
        const Box& sb = g.boxes[C.level][C.s];
        PanelTiles pt = addPanelTiles(acc, out, C.level, C.t, sb.size, -1, 0, tflags);
        for (const auto& tm : dd) {
            const Block& A = g.blocks[tm.first];
            const Block& B = g.blocks[tm.second];
            const int usz = g.boxes[A.level][A.s].size;
            for (int a = tb.atomBegin; a < tb.atomEnd; ++a) {
                const int ra = g.atomStart[a] - tb.off;
                for (int c = 0; c < pt.kChunks; ++c)
                    acc.add(pt.tileOfAtom[(size_t)(a - tb.atomBegin) * pt.kChunks + c],
                            mkc(SLOT_A, LA.denseOff[A.denseIndex] + ra, tb.size, SLOT_B,
                                LB.denseOff[B.denseIndex] + (int64_t)c * 64 * usz, usz, usz, -1, 0, cflags, 1.0f));
            }
            plan->flops += 2.0 * tb.size * usz * sb.size;
        }
        for (int wi : leafW(b)) {
            const WFactor& w = W[wi];
            const Operand pu = slicePU(w, tb), pv = slicePV(w, sb);
            for (int a = tb.atomBegin; a < tb.atomEnd; ++a) {
                const int ra = g.atomStart[a] - tb.off;
                for (int c = 0; c < pt.kChunks; ++c)
                    acc.add(pt.tileOfAtom[(size_t)(a - tb.atomBegin) * pt.kChunks + c],
                            mkc(pu.slot, pu.off + ra, pu.ld, pv.slot, pv.off + 64 * c, pv.ld, w.cap, w.rankIndex,
                                w.rankSlot, CF_TRANS_R | (w.cflag ? w.cflag : cflags), 1.0f));
            }
            plan->flops += 2.0 * tb.size * w.cap * sb.size;
        }
    };

    // Householder QR of an m x kS panel of box (level, box) at `off` (Q in
    // place, R to roff unless -1). Panels taller than one register-resident
    // workgroup are factored as a TSQR: QR of row blocks of whole atoms
    // (their R factors stacked), QR of the stack, then Q_i := Q_i Q_stack,i.
    const int tsqrRows = tsqrRows_ > 0 ? tsqrRows_ : (kS <= 24 ? 1024 : 512);
    auto addQr = [&](Phase& q1, Phase& q2, PhaseBuilder& comb, int64_t off, int level, int box, int64_t roff,
                     int wr) {
        const Box& B = g.boxes[level][box];
        const int m = B.size;
        plan->flops += 4.0 * m * kS * kS;
        if (m <= tsqrRows || kS > 40) {
            QrTask t{};
            t.off = off; t.roff = roff; t.m = m; t.k = kS; t.ld = m; t.rld = kS; t.slot = SLOT_WS;
            t.krank = wr; t.kslot = SLOT_WIDTH;
            q1.qr.push_back(t);
            return;
        }
        std::vector<int> blockAtom{B.atomBegin};
        for (int a = B.atomBegin, rows = 0; a < B.atomEnd; ++a) {
            if (rows + g.atomSize[a] > tsqrRows) { blockAtom.push_back(a); rows = 0; }
            rows += g.atomSize[a];
        }
        blockAtom.push_back(B.atomEnd);
        const int nb = (int)blockAtom.size() - 1;
        const int64_t Rs = wsAlloc((int64_t)nb * kS * kS);
        for (int i = 0; i < nb; ++i) {
            const int r0 = g.atomStart[blockAtom[i]] - B.off;
            const int r1 = g.atomStart[blockAtom[i + 1] - 1] + g.atomSize[blockAtom[i + 1] - 1] - B.off;
            QrTask t{};
            t.off = off + r0; t.roff = Rs + (int64_t)i * kS; t.m = r1 - r0; t.k = kS;
            t.ld = m; t.rld = nb * kS; t.slot = SLOT_WS;
            t.krank = wr; t.kslot = SLOT_WIDTH;
            q1.qr.push_back(t);
        }
        QrTask t{};
        t.off = Rs; t.roff = roff; t.m = nb * kS; t.k = kS; t.ld = nb * kS; t.rld = kS; t.slot = SLOT_WS;
        t.krank = wr; t.kslot = SLOT_WIDTH;
        q2.qr.push_back(t);
        for (int i = 0; i < nb; ++i)
            for (int a = blockAtom[i]; a < blockAtom[i + 1]; ++a) {
                const int ra = g.atomStart[a] - B.off;
                Tile tl{};
                tl.off = off + ra; tl.ld = m; tl.m = g.atomSize[a]; tl.k = kS;
                tl.mrank = -1; tl.krank = wr; tl.koff = 0; tl.kslot = SLOT_WIDTH; tl.slot = SLOT_WS;
                const int id = comb.addTile(tl);
                comb.add(id, mkc(SLOT_WS, off + ra, m, SLOT_WS, Rs + (int64_t)i * kS, nb * kS, kS, wr, SLOT_WIDTH,
                                 0, 1.0f));
                plan->flops += 2.0 * g.atomSize[a] * kS * kS;
            }
    };

    // 4. Per batch: W factors, dense accumulation, Y = T Omega, QR, Z = T^T Q,
    //    QR + SVD, formation of C.U and C.V.
    struct Sk { int b; int64_t omega, Y, Z, R, G1, G2, D; int m, n; bool leaf; int wr; };
    for (const auto& inGroup : groups) {
        ws_ = globalEnd;
        std::vector<Sk> sk;
        for (int lbk : cLr) {
            if (!inBatch(inGroup, lbk)) continue;
            const Block& C = g.blocks[lbk];
            const Box& tb = g.boxes[C.level][C.t];
            const Box& sb = g.boxes[C.level][C.s];
            Sk s;
            s.b = lbk; s.m = tb.size; s.n = sb.size;
            s.wr = widthIndirect_ ? C.lrIndex : -1;      // per-block sketch width (rank array index)
            // Leaf-level admissible blocks are accumulated densely, then compressed.
            s.leaf = (C.level == g.numLevels - 1);
            s.D = s.leaf ? wsAlloc((int64_t)tb.size * sb.size) : -1;
            s.omega = omegaOf[key2(C.level, C.s)];
            s.Y = wsAlloc((int64_t)tb.size * kS);
            s.Z = wsAlloc((int64_t)sb.size * kS);
            s.R = wsAlloc((int64_t)kS * kS);
            s.G1 = wsAlloc((int64_t)kS * LC.cap[C.lrIndex]);
            s.G2 = wsAlloc((int64_t)kS * LC.cap[C.lrIndex]);
            sk.push_back(s);
        }
        // W factors of this batch. Terms that share their sub-H block (the
        // same B(u,s) for a low-rank A(t,u), or A(t,u) for a low-rank B(u,s))
        // are stacked: their stored factors are gathered side by side into
        // one panel, and one sub-H product of width cap * n computes all
        // their exact factors at once (wider tiles, fewer contributions).
        {
            Phase& gp = newPhase(PH_TILE);
            Phase& p1 = newPhase(PH_TILE);
            p1.role = 1;
            Phase& a1 = newPhase(PH_TILE);
            PhaseBuilder gather(gp), proj(p1), acc(a1);
            std::map<std::pair<int, int>, std::vector<int>> stacks;
            std::vector<std::pair<int, int>> order;
            for (size_t i = 0; i < wPending.size(); ++i) {
                const auto& wp = wPending[i];
                if (!inBatch(inGroup, wp.bC)) continue;
                const std::pair<int, int> key{wp.left ? 1 : 0, wp.left ? wp.bB : wp.bA};
                auto& v = stacks[key];
                if (v.empty()) order.push_back(key);
                v.push_back((int)i);
            }
            for (const auto& key : order) {
                const auto& idx = stacks[key];
                const Pending& w0 = wPending[idx[0]];
                const int cap = W[w0.wIndex].cap;
                // Stacking computes every column of every factor; it pays for
                // factors of products and iterates (full rank), while imported
                // operands keep the rank-indirect per-term form (rank-0 blocks
                // then cost nothing).
                bool same = cap <= 64 && (w0.left ? stackA_ : stackB_);
                for (int i : idx) same = same && W[wPending[i].wIndex].cap == cap;
                if (idx.size() < 2 || !same) {
                    for (int i : idx) buildW(proj, acc, wPending[i]);
                    continue;
                }
                const int n = (int)idx.size();
                const bool left = w0.left;
                const Block& sub = g.blocks[left ? w0.bB : w0.bA];
                // Input rows: the u box (B's target box / A's source box).
                const int inLevel = sub.level, inBoxId = left ? sub.t : sub.s;
                const Box& inB = g.boxes[inLevel][inBoxId];
                const WFactor& wf0 = W[w0.wIndex];
                const int outBoxId = left ? wf0.s : wf0.t;
                const Box& outB = g.boxes[wf0.level][outBoxId];
                const int64_t IN = wsAlloc((int64_t)inB.size * cap * n);
                const int64_t OUT = wsAlloc((int64_t)outB.size * cap * n);
                eyeUsed = true;
                for (int i = 0; i < n; ++i) {
                    const Pending& wp = wPending[idx[i]];
                    const Block& lr = g.blocks[left ? wp.bA : wp.bB];
                    const int li = lr.lrIndex;
                    const int64_t src = left ? LA.vOff[li] : LB.uOff[li];
                    const uint8_t slot = left ? SLOT_A : SLOT_B;
                    for (int a = inB.atomBegin; a < inB.atomEnd; ++a) {
                        const int ra = g.atomStart[a] - inB.off;
                        Tile t{};
                        t.off = IN + ra + (int64_t)i * cap * inB.size; t.ld = inB.size;
                        t.m = g.atomSize[a]; t.k = cap; t.mrank = t.krank = -1; t.slot = SLOT_WS;
                        const int tile = gather.addTile(t);
                        gather.add(tile, mkc(slot, src + ra, inB.size, SLOT_WS, eyeOff, 64, cap, -1, 0, 0, 1.0f));
                    }
                    WFactor& w = W[wp.wIndex];
                    if (left) w.PV.off = OUT + (int64_t)i * cap * outB.size;
                    else w.PU.off = OUT + (int64_t)i * cap * outB.size;
                }
                PanelTiles pt = addPanelTiles(acc, {SLOT_WS, OUT, outB.size}, wf0.level, outBoxId, cap * n, -1, 0, 0);
                subh(proj, acc, left ? SLOT_B : SLOT_A, left ? LB : LA, left ? w0.bB : w0.bA, left,
                     {SLOT_WS, IN, inB.size}, inLevel, inBoxId, pt, cap * n, -1, 0, 1.0f, 0, &plan->flops);
            }
            gather.finalize();
            proj.finalize();
            acc.finalize();
        }
        {
            Phase& ad = newPhase(PH_TILE);
            PhaseBuilder acc(ad);
            if (!add)
                for (int bd : cDense) {
                    if (!inBatch(inGroup, bd)) continue;
                    const Block& C = g.blocks[bd];
                    // Dense C leaves: C = alpha (sum of terms) (+ beta C_old, read in place).
                    denseAccumulate(acc, bd, {SLOT_C, LC.denseOff[C.denseIndex], g.boxes[C.level][C.t].size},
                                    ddTerms[bd - cBase], CF_RUNTIME_ALPHA, update ? TF_RUNTIME_BETA : 0);
                }
            for (const auto& s : sk)
                if (s.leaf) denseAccumulate(acc, s.b, {SLOT_WS, s.D, s.m}, nnOf(s.b), abFlag, 0);
            acc.finalize();
        }
        // Y = T Omega for admissible C blocks.
        {
            Phase& p2 = newPhase(PH_TILE);
            p2.role = 1;
            Phase& a2 = newPhase(PH_TILE);
            PhaseBuilder proj(p2), acc(a2);
            for (const Sk& s : sk) {
                const Block& C = g.blocks[s.b];
                const Box& tb = g.boxes[C.level][C.t];
                const Box& sb = g.boxes[C.level][C.s];
                PanelTiles pt = addPanelTiles(acc, {SLOT_WS, s.Y, s.m}, C.level, C.t, kS, s.wr, SLOT_WIDTH, 0);
                const Operand om{SLOT_WS, s.omega, s.n};
                if (s.leaf) {
                    for (int a = tb.atomBegin; a < tb.atomEnd; ++a) {
                        const int ra = g.atomStart[a] - tb.off;
                        acc.add(pt.tileOfAtom[a - tb.atomBegin],
                                mkc(SLOT_WS, s.D + ra, s.m, SLOT_WS, s.omega, s.n, s.n, -1, 0, 0, 1.0f));
                    }
                    plan->flops += 2.0 * s.m * s.n * kS;
                    continue;
                }
                for (int wi : leafW(s.b)) {
                    const WFactor& w = W[wi];
                    const int64_t P = wsAlloc((int64_t)w.cap * kS);
                    projTile(proj, slicePV(w, sb), sb, om, w, P, s.wr);
                    const Operand pu = slicePU(w, tb);
                    for (int a = tb.atomBegin; a < tb.atomEnd; ++a) {
                        const int ra = g.atomStart[a] - tb.off;
                        acc.add(pt.tileOfAtom[a - tb.atomBegin],
                                mkc(pu.slot, pu.off + ra, pu.ld, SLOT_WS, P, w.cap, w.cap, w.rankIndex,
                                    w.rankSlot, w.cflag, 1.0f));
                    }
                    plan->flops += 2.0 * tb.size * w.cap * kS;
                }
                for (const auto& tm : nnOf(s.b)) {
                    const Block& B = g.blocks[tm.second];
                    const Operand in{SLOT_WS, t1Of[(uint64_t)tm.second], g.boxes[B.level][B.t].size};
                    subh(proj, acc, SLOT_A, LA, tm.first, false, in, B.level, B.t, pt, kS, s.wr, SLOT_WIDTH, 1.0f,
                         abFlag, &plan->flops);
                }
            }
            proj.finalize();
            acc.finalize();
        }
        // Q = qr(Y)
        {
            Phase& q = newPhase(PH_QR);
            Phase& q2 = newPhase(PH_QR);
            Phase& cb = newPhase(PH_TILE);
            q.qrK = q2.qrK = kS;
            PhaseBuilder comb(cb);
            for (const Sk& s : sk) {
                const Block& C = g.blocks[s.b];
                addQr(q, q2, comb, s.Y, C.level, C.t, -1, s.wr);
            }
            comb.finalize();
        }
        // Z = T^T Q: low-rank terms directly, NN terms via T2 = A(t,u)^T Q.
        {
            Phase& p3 = newPhase(PH_TILE);
            p3.role = 1;
            Phase& a3 = newPhase(PH_TILE);
            Phase& p4 = newPhase(PH_TILE);
            p4.role = 1;
            Phase& a4 = newPhase(PH_TILE);
            PhaseBuilder proj3(p3), acc3(a3), proj4(p4), acc4(a4);
            for (const Sk& s : sk) {
                const Block& C = g.blocks[s.b];
                const Box& tb = g.boxes[C.level][C.t];
                const Box& sb = g.boxes[C.level][C.s];
                PanelTiles zt = addPanelTiles(acc4, {SLOT_WS, s.Z, s.n}, C.level, C.s, kS, s.wr, SLOT_WIDTH, 0);
                const Operand Q{SLOT_WS, s.Y, s.m};
                if (s.leaf) {
                    // Z = D^T Q: L(i, p) = D(p, i)
                    for (int a = sb.atomBegin; a < sb.atomEnd; ++a) {
                        const int ra = g.atomStart[a] - sb.off;
                        acc4.add(zt.tileOfAtom[a - sb.atomBegin],
                                 mkc(SLOT_WS, s.D + (int64_t)ra * s.m, s.m, SLOT_WS, s.Y, s.m, s.m, -1, 0,
                                     CF_TRANS_L, 1.0f));
                    }
                    plan->flops += 2.0 * s.m * s.n * kS;
                    continue;
                }
                for (int wi : leafW(s.b)) {
                    const WFactor& w = W[wi];
                    const int64_t P = wsAlloc((int64_t)w.cap * kS);
                    projTile(proj3, slicePU(w, tb), tb, Q, w, P, s.wr);
                    const Operand pv = slicePV(w, sb);
                    for (int a = sb.atomBegin; a < sb.atomEnd; ++a) {
                        const int ra = g.atomStart[a] - sb.off;
                        acc4.add(zt.tileOfAtom[a - sb.atomBegin],
                                 mkc(pv.slot, pv.off + ra, pv.ld, SLOT_WS, P, w.cap, w.cap, w.rankIndex,
                                     w.rankSlot, w.cflag, 1.0f));
                    }
                    plan->flops += 2.0 * sb.size * w.cap * kS;
                }
                for (const auto& tm : nnOf(s.b)) {
                    const Block& A = g.blocks[tm.first];
                    const Box& ub = g.boxes[A.level][A.s];
                    const int64_t T2 = wsAlloc((int64_t)ub.size * kS);
                    PanelTiles pt = addPanelTiles(acc3, {SLOT_WS, T2, ub.size}, A.level, A.s, kS, s.wr,
                                                  SLOT_WIDTH, 0);
                    subh(proj3, acc3, SLOT_A, LA, tm.first, true, Q, C.level, C.t, pt, kS, s.wr, SLOT_WIDTH, 1.0f, 0,
                         &plan->flops);
                    subh(proj4, acc4, SLOT_B, LB, tm.second, true, {SLOT_WS, T2, ub.size}, A.level, A.s, zt, kS,
                         s.wr, SLOT_WIDTH, 1.0f, abFlag, &plan->flops);
                }
            }
            proj3.finalize(); acc3.finalize(); proj4.finalize(); acc4.finalize();
        }
        // QR(Z) keeping R, SVD(R) with rank decision, form C.U, C.V.
        {
            Phase& q = newPhase(PH_QR);
            Phase& q2 = newPhase(PH_QR);
            Phase& cb = newPhase(PH_TILE);
            q.qrK = q2.qrK = kS;
            {
                PhaseBuilder comb(cb);
                for (const Sk& s : sk) {
                    const Block& C = g.blocks[s.b];
                    addQr(q, q2, comb, s.Z, C.level, C.s, s.R, s.wr);
                }
                comb.finalize();
            }
            Phase& sv = newPhase(PH_SVD);
            for (const Sk& s : sk) {
                const Block& C = g.blocks[s.b];
                SvdTask v{};
                v.roff = s.R; v.g1off = s.G1; v.g2off = s.G2; v.k = kS; v.ld = kS;
                v.cap = LC.cap[C.lrIndex];
                v.rankIndex = C.lrIndex;
                v.krank = s.wr; v.spare = oversample;
                v.maxKeep = std::min(std::min(maxRank, v.cap), std::min(std::min(s.m, s.n), kS));
                sv.svd.push_back(v);
                plan->flops += 30.0 * kS * kS * kS;
            }
            Phase& f = newPhase(PH_TILE);
            PhaseBuilder acc(f);
            for (const Sk& s : sk) {
                const Block& C = g.blocks[s.b];
                const int cap = LC.cap[C.lrIndex];
                const Box& tb = g.boxes[C.level][C.t];
                const Box& sb = g.boxes[C.level][C.s];
                PanelTiles ut = addPanelTiles(acc, {SLOT_C, LC.uOff[C.lrIndex], s.m}, C.level, C.t, cap, -1, 0, 0);
                PanelTiles vt = addPanelTiles(acc, {SLOT_C, LC.vOff[C.lrIndex], s.n}, C.level, C.s, cap, -1, 0, 0);
                for (int a = tb.atomBegin; a < tb.atomEnd; ++a)
                    for (int c = 0; c < ut.kChunks; ++c)
                        acc.add(ut.tileOfAtom[(size_t)(a - tb.atomBegin) * ut.kChunks + c],
                                mkc(SLOT_WS, s.Y + (g.atomStart[a] - tb.off), s.m, SLOT_WS,
                                    s.G1 + (int64_t)c * 64 * kS, kS, kS, s.wr, SLOT_WIDTH, formFlag, 1.0f));
                for (int a = sb.atomBegin; a < sb.atomEnd; ++a)
                    for (int c = 0; c < vt.kChunks; ++c)
                        acc.add(vt.tileOfAtom[(size_t)(a - sb.atomBegin) * vt.kChunks + c],
                                mkc(SLOT_WS, s.Z + (g.atomStart[a] - sb.off), s.n, SLOT_WS,
                                    s.G2 + (int64_t)c * 64 * kS, kS, kS, s.wr, SLOT_WIDTH, 0, 1.0f));
                plan->flops += 2.0 * (s.m + s.n) * kS * cap;
            }
            acc.finalize();
        }
        wsPeak = std::max(wsPeak, wsUsed());
    }
    // Drop phases with nothing to do (small geometries, empty batches).
    std::vector<Phase> kept;
    kept.reserve(plan->phases.size());
    for (auto& ph : plan->phases) {
        const bool empty = ph.kind == PH_EYE ? !eyeUsed
                                             : ph.tiles.empty() && ph.qr.empty() && ph.svd.empty() && ph.rng.empty();
        if (!empty) kept.push_back(std::move(ph));
    }
    plan->phases.swap(kept);
    plan->wsElems = wsPeak;
    plan->wsShared = globalEnd;
    plan->batches = (int)groups.size();
    for (auto& ph : plan->phases) {
        plan->tileCount += ph.tiles.size();
        plan->contribCount += ph.contribs.size();
    }
    return plan;
}

// ---------------------------------------------------------------------------
// Merged plans
// ---------------------------------------------------------------------------
Plan* merge_plans(const std::vector<PlanPart>& parts, bool segmentTiles)
{
    std::unique_ptr<Plan> out(new Plan());
    const size_t np = parts.size();
    std::vector<int64_t> wsBase(np);
    int64_t ws = 0;
    for (size_t i = 0; i < np; ++i) {
        wsBase[i] = ws;
        ws += (parts[i].plan->wsElems + 63) & ~int64_t(63);
        out->flops += parts[i].plan->flops;
        out->batches = std::max(out->batches, parts[i].plan->batches);
    }
    out->wsElems = ws;
    // Shifts of part i: the workspace slot goes to the part's region.
    auto vsh = [&](size_t i, int slot) { return slot == SLOT_WS ? wsBase[i] : parts[i].val[slot]; };
    auto rsh = [&](size_t i, int slot, int idx) { return idx < 0 ? idx : idx + (int)parts[i].rank[slot]; };

    std::vector<size_t> pos(np, 0);
    for (;;) {
        // The next merged phase takes the kind of the part with the most
        // phases left, and every part whose next phase has that kind.
        int lead = -1;
        size_t most = 0;
        for (size_t i = 0; i < np; ++i) {
            const size_t left = parts[i].plan->phases.size() - pos[i];
            if (left > most) { most = left; lead = (int)i; }
        }
        if (lead < 0) break;
        const Phase& lp = parts[lead].plan->phases[pos[lead]];
        out->phases.emplace_back();
        Phase& m = out->phases.back();
        m.kind = lp.kind;
        m.mvK = lp.mvK; m.ldx = lp.ldx; m.ldy = lp.ldy;
        m.zeroCount = lp.zeroCount;
        m.role = lp.role;
        bool first = true;
        for (size_t i = 0; i < np; ++i) {
            if (pos[i] >= parts[i].plan->phases.size()) continue;
            const Phase& ph = parts[i].plan->phases[pos[i]];
            if (ph.kind != m.kind) continue;
            ++pos[i];
            m.flops += ph.flops;
            m.qrK = std::max(m.qrK, ph.qrK);
            switch (ph.kind) {
            case PH_TILE: {
                if (segmentTiles) {
                    Phase::SegSrc sg;
                    sg.src = &ph;
                    for (int s = 0; s < NUM_SLOTS; ++s) { sg.val[s] = vsh(i, s); sg.rank[s] = parts[i].rank[s]; }
                    m.segSrc.push_back(sg);
                    m.maxKTile = std::max(m.maxKTile, ph.maxKTile);
                    m.maxMTile = std::max(m.maxMTile, ph.maxMTile);
                    break;
                }
                const int cbase = (int)m.contribs.size();
                for (Tile t : ph.tiles) {
                    t.off += vsh(i, t.slot);
                    t.mrank = rsh(i, t.mslot, t.mrank);
                    t.krank = rsh(i, t.kslot, t.krank);
                    t.cb += cbase; t.ce += cbase;
                    m.tiles.push_back(t);
                    m.maxKTile = std::max(m.maxKTile, t.k);
                    m.maxMTile = std::max(m.maxMTile, t.m);
                }
                for (Contrib c : ph.contribs) {
                    c.loff += vsh(i, c.lslot);
                    c.roff += vsh(i, c.rslot);
                    c.qrank = rsh(i, c.qslot, c.qrank);
                    m.contribs.push_back(c);
                }
                break;
            }
            case PH_QR:
                for (QrTask t : ph.qr) {
                    t.off += vsh(i, t.slot);
                    if (t.roff >= 0) t.roff += wsBase[i];
                    t.krank = rsh(i, t.kslot, t.krank);
                    m.qr.push_back(t);
                }
                break;
            case PH_SVD:
                for (SvdTask t : ph.svd) {
                    t.roff += wsBase[i]; t.g1off += wsBase[i]; t.g2off += wsBase[i];
                    t.rankIndex = rsh(i, SLOT_C, t.rankIndex);
                    t.krank = rsh(i, SLOT_WIDTH, t.krank);
                    m.svd.push_back(t);
                }
                break;
            case PH_RNG:
                for (RngTask t : ph.rng) {
                    t.off += wsBase[i];
                    m.rng.push_back(t);
                }
                break;
            case PH_EYE:
                if (ph.zeroCount != m.zeroCount) throw std::logic_error("merged identities of different orders");
                if (ph.eyes.empty()) m.eyes.push_back(ph.zeroOff + wsBase[i]);
                for (int64_t e : ph.eyes) m.eyes.push_back(e + wsBase[i]);
                break;
            case PH_MV_ZERO:
                // A streaming product zeroes the front of its workspace (its
                // projections); the merged phase zeroes the whole workspace.
                if (ph.zeroOff != 0 || ph.zeroCount > parts[i].plan->wsElems)
                    throw std::logic_error("merged zero range is not inside the part's workspace");
                m.zeroOff = 0;
                m.zeroCount = -1;
                break;
            case PH_MV_PROJ:
                if (!first && (ph.mvK != m.mvK || (ph.mvK > 1 && (ph.ldx != m.ldx || ph.ldy != m.ldy))))
                    throw std::logic_error("merged streaming products of different shapes");
                for (MvProjItem it : ph.mvProj) {
                    it.voff += vsh(i, SLOT_A);
                    it.xrow += vsh(i, SLOT_X);
                    it.poff += wsBase[i];
                    it.rankIdx = rsh(i, SLOT_A, it.rankIdx);
                    m.mvProj.push_back(it);
                }
                break;
            case PH_MV_ACC: {
                if (!first && (ph.mvK != m.mvK || (ph.mvK > 1 && (ph.ldx != m.ldx || ph.ldy != m.ldy))))
                    throw std::logic_error("merged streaming products of different shapes");
                const int cbase = (int)m.mvContribs.size();
                for (MvAtom a : ph.mvAtoms) {
                    a.yrow += vsh(i, SLOT_Y);
                    a.cb += cbase; a.ce += cbase;
                    m.mvAtoms.push_back(a);
                }
                for (MvContrib c : ph.mvContribs) {
                    c.loff += vsh(i, SLOT_A);
                    c.roff += c.kind == 0 ? vsh(i, SLOT_X) : wsBase[i];
                    c.rankIdx = rsh(i, SLOT_A, c.rankIdx);
                    m.mvContribs.push_back(c);
                }
                break;
            }
            default:
                throw std::logic_error("phase kind cannot be merged");
            }
            first = false;
        }
    }
    for (auto& ph : out->phases) {
        if (ph.kind == PH_MV_ZERO) ph.zeroCount = out->wsElems;
        out->tileCount += ph.tiles.size();
        out->contribCount += ph.contribs.size();
        for (const auto& sg : ph.segSrc) {
            out->tileCount += sg.src->tiles.size();
            out->contribCount += sg.src->contribs.size();
        }
    }
    return out.release();
}

std::string check_plan(const Plan& p, const int64_t* slotElems, const int64_t* rankElems)
{
    auto size = [&](int slot) { return slot == SLOT_WS ? p.wsElems : slotElems[slot]; };
    auto where = [&](size_t ph, const char* what) {
        return "phase " + std::to_string(ph) + " " + what;
    };
    // Extent of an r x c column-major region at off with leading dimension ld.
    auto inside = [&](int slot, int64_t off, int64_t ld, int64_t r, int64_t c) {
        if (r <= 0 || c <= 0) return true;
        if (slot < 0 || slot >= NUM_SLOTS || off < 0 || ld < r) return false;
        return off + (c - 1) * ld + r <= size(slot);
    };
    auto rankOk = [&](int slot, int idx) { return idx < 0 || (slot < NUM_SLOTS && idx < rankElems[slot]); };
    const int64_t zero[NUM_SLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
    // Tiles of a phase (or of a member's source phase, with its shifts).
    auto checkTiles = [&](size_t i, const Phase& ph, const int64_t* vs, const int64_t* rs) -> std::string {
        auto rk = [&](int slot, int idx) { return rankOk(slot, idx < 0 ? idx : (int)(idx + rs[slot])); };
        for (const Tile& t : ph.tiles) {
            if (!inside(t.slot, t.off + vs[t.slot], t.ld, t.m, t.k)) return where(i, "tile output out of bounds");
            if (!rk(t.mslot, t.mrank) || !rk(t.kslot, t.krank)) return where(i, "tile rank index");
            if (t.cb < 0 || t.ce > (int)ph.contribs.size() || t.cb > t.ce) return where(i, "tile contribution range");
            for (int q = t.cb; q < t.ce; ++q) {
                const Contrib& c = ph.contribs[q];
                const bool tl = c.flags & CF_TRANS_L, tr = c.flags & CF_TRANS_R;
                const int64_t lo = c.loff + vs[c.lslot], ro = c.roff + vs[c.rslot];
                const bool okL = tl ? inside(c.lslot, lo, c.ldl, c.q, t.m) : inside(c.lslot, lo, c.ldl, t.m, c.q);
                const bool okR = tr ? inside(c.rslot, ro, c.ldr, t.k, c.q) : inside(c.rslot, ro, c.ldr, c.q, t.k);
                if (!okL) return where(i, "left operand out of bounds");
                if (!okR) return where(i, "right operand out of bounds");
                if (!rk(c.qslot, c.qrank)) return where(i, "contribution rank index");
            }
        }
        return std::string();
    };
    // Footprint of a member's source phase (per slot: lowest offset, highest
    // end, highest rank index), computed once per source and checked against
    // every member's shifts.
    struct Foot { int64_t lo[NUM_SLOTS], hi[NUM_SLOTS], rk[NUM_SLOTS]; bool bad; };
    std::map<const Phase*, Foot> feet;
    auto footprint = [&](const Phase& ph) -> const Foot& {
        auto it = feet.find(&ph);
        if (it != feet.end()) return it->second;
        Foot f;
        for (int s = 0; s < NUM_SLOTS; ++s) { f.lo[s] = INT64_MAX; f.hi[s] = -1; f.rk[s] = -1; }
        f.bad = false;
        auto region = [&](int slot, int64_t off, int64_t ld, int64_t r, int64_t cc) {
            if (r <= 0 || cc <= 0) return;
            if (slot < 0 || slot >= NUM_SLOTS || ld < r) { f.bad = true; return; }
            f.lo[slot] = std::min(f.lo[slot], off);
            f.hi[slot] = std::max(f.hi[slot], off + (cc - 1) * ld + r);
        };
        auto rank = [&](int slot, int idx) {
            if (idx < 0) return;
            if (slot < 0 || slot >= NUM_SLOTS) { f.bad = true; return; }
            f.rk[slot] = std::max<int64_t>(f.rk[slot], idx);
        };
        for (const Tile& t : ph.tiles) {
            region(t.slot, t.off, t.ld, t.m, t.k);
            rank(t.mslot, t.mrank);
            rank(t.kslot, t.krank);
            if (t.cb < 0 || t.ce > (int)ph.contribs.size() || t.cb > t.ce) { f.bad = true; continue; }
            for (int q = t.cb; q < t.ce; ++q) {
                const Contrib& x = ph.contribs[q];
                const bool tl = x.flags & CF_TRANS_L, tr = x.flags & CF_TRANS_R;
                if (tl) region(x.lslot, x.loff, x.ldl, x.q, t.m); else region(x.lslot, x.loff, x.ldl, t.m, x.q);
                if (tr) region(x.rslot, x.roff, x.ldr, t.k, x.q); else region(x.rslot, x.roff, x.ldr, x.q, t.k);
                rank(x.qslot, x.qrank);
            }
        }
        return feet.emplace(&ph, f).first->second;
    };
    for (size_t i = 0; i < p.phases.size(); ++i) {
        const Phase& ph = p.phases[i];
        std::string e = checkTiles(i, ph, zero, zero);
        if (!e.empty()) return e;
        for (const auto& sg : ph.segSrc) {
            const Foot& f = footprint(*sg.src);
            if (f.bad) return where(i, "member tile malformed");
            for (int s = 0; s < NUM_SLOTS; ++s) {
                if (f.hi[s] >= 0 && (f.lo[s] + sg.val[s] < 0 || f.hi[s] + sg.val[s] > size(s)))
                    return where(i, "member tile operand out of bounds");
                if (f.rk[s] >= 0 && f.rk[s] + sg.rank[s] >= rankElems[s])
                    return where(i, "member tile rank index");
            }
        }
        for (const QrTask& t : ph.qr) {
            if (!inside(t.slot, t.off, t.ld, t.m, t.k)) return where(i, "QR panel out of bounds");
            if (t.roff >= 0 && !inside(SLOT_WS, t.roff, t.rld, t.k, t.k)) return where(i, "QR R factor out of bounds");
            if (!rankOk(t.kslot, t.krank)) return where(i, "QR rank index");
        }
        for (const SvdTask& t : ph.svd) {
            const int ld = t.ld > 0 ? t.ld : t.k;
            if (!inside(SLOT_WS, t.roff, ld, t.k, t.k) || !inside(SLOT_WS, t.g1off, ld, t.k, t.cap) ||
                !inside(SLOT_WS, t.g2off, ld, t.k, t.cap))
                return where(i, "SVD factor out of bounds");
            if (!rankOk(SLOT_C, t.rankIndex) || t.rankIndex < 0) return where(i, "SVD rank index");
        }
        for (const RngTask& t : ph.rng)
            if (!inside(SLOT_WS, t.off, t.m, t.m, t.k)) return where(i, "sketch panel out of bounds");
        if (ph.kind == PH_EYE) {
            std::vector<int64_t> eyes = ph.eyes;
            if (eyes.empty()) eyes.push_back(ph.zeroOff);
            for (int64_t e : eyes)
                if (!inside(SLOT_WS, e, ph.zeroCount, ph.zeroCount, ph.zeroCount)) return where(i, "identity out of bounds");
        }
    }
    return std::string();
}

} // namespace dmhm
