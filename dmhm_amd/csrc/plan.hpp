// Execution plans: every operation on H-matrices (H*X, H^T*X, H*H with
// randomized compression) is lowered on the host, once per geometry and
// storage layout, into a short sequence of device phases:
//
//  * TILE phases: grouped, output-owned small GEMMs. An output tile is a
//    row atom (<= 64 rows) x <= 64 columns of some panel; it owns a list of
//    contributions op(L) * op(R) that the kernel sums in registers and writes
//    once (no atomics, deterministic order).
//  * QR phases: batched Householder QR of tall-skinny panels (Q formed).
//  * SVD phases: batched one-sided Jacobi SVD of small k x k factors with the
//    rank decision written to the output matrix's rank array.
//  * RNG phases: Gaussian sketch panels.
//
// Operands are (slot, element offset, leading dimension): the slot selects a
// base pointer bound at launch (A, B, C matrices, workspace, user X / Y), so
// one plan serves every matrix with the same layout and both precisions.
// Inner dimensions and tile shapes may be "rank-indirect": read at run time
// from a matrix's device rank array, so plans never need host round trips
// and whole iterations can be captured in a HIP graph.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>
#include "geometry.hpp"

namespace dmhm {

// Slots 6 and 7 carry only integer arrays: per-block sketch widths (read
// where a plan is width-indirect) and per-block flags of the a-posteriori
// sketch check.
enum Slot : uint8_t { SLOT_A = 0, SLOT_B = 1, SLOT_C = 2, SLOT_WS = 3, SLOT_X = 4, SLOT_Y = 5,
                      SLOT_WIDTH = 6, SLOT_FLAG = 7, NUM_SLOTS = 8 };

enum ContribFlags : uint8_t { CF_TRANS_L = 1, CF_TRANS_R = 2, CF_RUNTIME_ALPHA = 4, CF_RUNTIME_BETA = 8 };
enum TileFlags : uint8_t { TF_ACCUMULATE = 1, TF_RUNTIME_BETA = 2 };

// Plain-old-data records shared with the kernels.
struct Contrib {
    int64_t loff, roff;     // element offsets of L(0,0) and R(0,0) in their slots
    int32_t ldl, ldr;
    int32_t q;              // inner dimension (if qrank < 0)
    int32_t qrank;          // rank-array index giving the inner dimension, or -1
    uint8_t lslot, rslot, qslot, flags;
    float alpha;
};

struct Tile {
    int64_t off;            // element offset of the tile's (0,0) in its slot
    int32_t ld;
    int32_t m, k;           // static sizes (<= 64); clamps when rank-indirect
    int32_t mrank, krank;   // rank-array index (or -1): m = min(m, rank - moff)
    int32_t moff, koff;
    int32_t cb, ce;         // contribution range
    uint8_t slot, mslot, kslot, flags;
    int32_t pad;
};

struct QrTask {              // Householder QR of an m x k panel, Q in place
    int64_t off;            // panel offset (slot)
    int64_t roff;           // R output (k x k upper triangle, ld rld) offset in the workspace, or -1
    int32_t m, k;           // k: static width (R rows written: k; columns: the effective width)
    int32_t ld, rld;        // panel and R leading dimensions
    uint8_t slot = 0, kslot = 0, pad1 = 0, pad2 = 0;
    int32_t krank = -1;     // rank-array index of the effective width min(k, rank[kslot][krank]), or -1
};

struct SvdTask {             // SVD of R (k x k) -> G1 = V S (k x cap), G2 = U (k x cap)
    int64_t roff, g1off, g2off;   // workspace offsets (R, G1, G2 with leading dimension ld)
    int32_t k, cap;
    int32_t rankIndex;      // rank slot entry written with the kept rank
    int32_t maxKeep;        // min(maxRank, m, n, k)
    int32_t ld = 0;         // leading dimension of R, G1, G2 (the static width; 0: k)
    int32_t krank = -1;     // effective width min(k, rank[SLOT_WIDTH][krank]), or -1
    int32_t spare = 0;      // oversampling columns required below the tolerance (a-posteriori check)
    int32_t pad = 0;
};

struct RngTask {
    int64_t off;            // workspace offset of an m x k panel (ld m)
    int32_t m, k;
    uint64_t stream;        // stream id mixed into the counter
};

// Streaming H*X for few right-hand sides (HBM-bound): projections
// P = V^T X per (low-rank block, row chunk) by one wave each, then one
// workgroup per output row atom summing dense-leaf and U*P contributions.
struct MvProjItem {
    int64_t voff;           // slot A offset of V(row0, 0) (ld = ldv)
    int64_t xrow;           // first X row of the chunk
    int64_t poff;           // workspace offset of P (cap x k, ld cap)
    int32_t ldv, rows;      // V leading dimension, rows in this chunk
    int32_t rankIdx, cap;
    int32_t atomic;         // several chunks share P: accumulate atomically
    int32_t pad;
};
struct MvAtom {
    int64_t yrow;
    int32_t m, cb, ce, pad;
};
struct MvContrib {
    int64_t loff;           // slot A offset of the first row of L (dense leaf or U)
    int64_t roff;           // X row offset (kind 0) or workspace offset of P (kind 1)
    int32_t ld;             // leading dimension of L
    int32_t q;              // inner dimension (static) ...
    int32_t rankIdx;        // ... or the block's rank (kind 1)
    int32_t kind;           // 0: dense leaf times X rows, 1: U rows times P
    int32_t ldr;            // P leading dimension (cap) for kind 1
    int32_t pad;
};

enum PhaseKind { PH_TILE, PH_QR, PH_SVD, PH_RNG, PH_SCALE_C, PH_MV_ZERO, PH_MV_PROJ, PH_MV_ACC, PH_EYE };

// One member's share of a merged tile phase that is not copied: the member
// plan's tile and contribution arrays (device) and the shifts of its slots
// (elements, rank indices; the workspace slot is shifted to the member's
// workspace region). first1 / first4: the member's first workgroup when a
// workgroup takes one / four tiles.
struct TileSeg {
    const Tile* tiles;
    const Contrib* cs;
    int64_t val[NUM_SLOTS];
    int64_t rank[NUM_SLOTS];
    int count, first1, first4, pad;
};

struct Phase {
    PhaseKind kind;
    std::vector<Tile> tiles;
    std::vector<Contrib> contribs;
    std::vector<QrTask> qr;
    std::vector<SvdTask> svd;
    std::vector<RngTask> rng;
    std::vector<MvProjItem> mvProj;
    std::vector<MvAtom> mvAtoms;
    std::vector<MvContrib> mvContribs;
    int64_t zeroOff = 0, zeroCount = 0;   // PH_MV_ZERO range in the workspace; PH_EYE: offset, order
    std::vector<int64_t> eyes;            // PH_EYE of a merged plan: every identity's workspace offset
    // PH_TILE of a merged plan built without copying: per member, the source
    // phase (host, uploaded with its plan) and the member's shifts; the
    // device table of TileSeg is built at upload.
    struct SegSrc { const Phase* src; int64_t val[NUM_SLOTS]; int64_t rank[NUM_SLOTS]; };
    std::vector<SegSrc> segSrc;
    void* dSegs = nullptr;
    int segBlocks1 = 0, segBlocks4 = 0;   // workgroups of a segmented launch (1 / 4 tiles each)
    int mvK = 0, ldx = 0, ldy = 0;
    void* dAux = nullptr;
    int qrK = 0;            // panel width of the QR phase
    int qrClassEnd[4] = {0, 0, 0, 0};   // QR tasks sorted by panel class at upload (see launch_qr_phase)
    int qrRegK = 0;         // widest task of classes 0..3
    int svdK = 0;           // largest SVD factor of the phase (set by upload)
    // device copies (filled by upload)
    void* dTiles = nullptr; void* dContribs = nullptr; void* dTasks = nullptr;
    int maxKTile = 0;       // widest tile (selects the kernel variant)
    int maxMTile = 0;       // tallest tile
    int role = 0;           // 0 accumulation, 1 projection (profiling identity)
    double flops = 0;       // algorithmic flops (2 m q k per contribution at static sizes)
    // Tile-engine autotuning state per execution slot (FP64, FP32, FP32 with
    // BF16 operands allowed): candidates timed so far (per member segment)
    // and the choice (-1 while tuning). A merged phase whose segments all
    // come from one member plan's phase tunes in that phase (mutable: merged
    // plans hold their sources const), so every chunk of a class shares one
    // tuning, as the members' lone runs share their class plan's.
    struct Tune { int choice = -1; int next = -1; float ms[8] = {0, 0, 0, 0, 0, 0, 0, 0}; };   // next -1: warm-up run
    mutable Tune tune[3];
};

struct Plan {
    std::vector<Phase> phases;
    int64_t wsElems = 0;    // workspace elements (of the run-time precision)
    double flops = 0;       // algorithmic flops (GEMM shapes)
    int64_t contribCount = 0, tileCount = 0;
    int batches = 1;        // H*H: number of level-1 subtree batches sharing the workspace
    int64_t wsShared = 0;   // H*H: workspace elements shared by all batches (Omega, T1)
    bool keepHost = false;  // keep the host tile/contribution arrays after upload (plans that
                            // merged plans are built from and checked against)
    std::vector<std::shared_ptr<Plan>> members;   // member plans a segmented merged plan reads
    std::vector<std::unique_ptr<Plan>> chunks;    // merged plans run one after another (a batch
                                                  // split where the members' workspaces exceed the budget)
    void upload();          // copy tile/contrib/task arrays to the device
    void release();
    ~Plan() { release(); }
};

// Builds the phases. Contributions are collected per tile and sorted at
// finalize so that each tile owns a contiguous contribution range.
class PhaseBuilder {
public:
    explicit PhaseBuilder(Phase& ph) : ph_(ph) {}
    int addTile(const Tile& t) { ph_.tiles.push_back(t); return (int)ph_.tiles.size() - 1; }
    void add(int tile, const Contrib& c) { pending_.push_back({tile, c}); }
    void finalize();
private:
    Phase& ph_;
    std::vector<std::pair<int, Contrib>> pending_;
};

struct Operand {            // a panel or matrix block region
    uint8_t slot;
    int64_t off;
    int32_t ld;
};

// Terms of C = A B for three H-matrices of one geometry, from the reference's
// recursion (MultiplyHMat-incl.hpp:12-380): a triple (A(t,u), B(u,s), C(t,s))
// recurses while all three are NODEs; otherwise it is
//  * a low-rank term (A or B low rank): exact factors W = PU PV^T, added to an
//    admissible or dense C block, or pushed down into a C NODE
//    (ImportLowRank / UpdateWithLowRank, HMat2d.cpp:811-959);
//  * a NODE*NODE term into an admissible C block (the randomized sketch case,
//    hmat_tools.hpp:1657-1746);
//  * a dense*dense term into an admissible C block (truncated D*D,
//    MultiplyMatrix.cpp:405-449) or into a dense C leaf.
struct ProductTerms {
    struct W { int bA, bB, bC; bool leftLowRank; };
    std::vector<W> w;                              // all low-rank terms
    // Per C block b of the subtree of rootC, at index b - cBase:
    int cBase = 0;
    std::vector<std::vector<int>> pushdowns;       // per C NODE block: w ids
    std::vector<std::vector<int>> lrTerms;         // per C leaf block: w ids
    std::vector<std::vector<std::pair<int, int>>> nnTerms, ddTerms;  // per C block (bA, bB)
    long sketchTerms = 0, ddToLowRank = 0, ddToDense = 0, pushdownTerms = 0, lowRankTerms = 0;
};
// The recursion starts at the block triple (rootA, rootB, rootC) (the whole
// matrices: the root block three times).
ProductTerms enumerateProductTerms(const Geometry& g, int rootA = 0, int rootB = 0, int rootC = 0);

// One part of a merged plan: a plan and, per slot, the element offset added to
// every operand offset in that slot and the index offset added to every rank
// index read from that slot's rank array (SLOT_WS is placed by the merge).
struct PlanPart {
    const Plan* plan = nullptr;
    int64_t val[NUM_SLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t rank[NUM_SLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
};
// Merges plans over disjoint data into one plan with one launch per merged
// phase: each part keeps its phase order, phases of the same kind from
// different parts share a launch (tasks concatenated and shifted), and every
// part gets its own workspace region. The parts must not be uploaded (their
// host task arrays are read) and must not write what another part reads.
// With segmentTiles, tile phases are not copied: each merged tile phase
// lists its members' source phases and shifts (launched through a segment
// table); the source plans must then stay alive and be uploaded before the
// merged plan.
Plan* merge_plans(const std::vector<PlanPart>& parts, bool segmentTiles = false);

// Host-side bounds check of a plan that has not been uploaded: every operand,
// output, QR panel, SVD factor, sketch panel and identity of every task lies
// inside its slot (slotElems[slot] elements; SLOT_WS: the plan's workspace)
// and every rank index inside rankElems[slot]. Returns an empty string or the
// first violation.
std::string check_plan(const Plan& p, const int64_t* slotElems, const int64_t* rankElems);

// Plan builders.
class Planner {
public:
    Planner(const Geometry& g) : g_(g) {}

    // Y (n x k, ld ldy) := alpha op(H) X + beta Y, H in SLOT_A with layout L.
    Plan* matvec(const Layout& L, bool trans, int k, int ldx, int ldy);
    // Streaming form of the non-transposed product for k <= 8.
    Plan* matvecStream(const Layout& L, int k, int ldx, int ldy);

    // C := alpha A B with randomized sum-sketch compression of every admissible
    // block (sketch width kS = maxRank + oversample), A/B/C layouts given.
    // tolMode 0: keep min(maxRank, kS) singular values; 1: keep those above
    // relTol * sigma_0 (at most maxRank).
    // update = true: C := alpha A B + beta C_old with C_old (layout LC) bound in
    // SLOT_X; every term of A B is scaled by the run-time alpha and C_old by the
    // run-time beta inside the sketch.
    enum HHMode { HH_PRODUCT = 0, HH_UPDATE = 1, HH_ADD = 2 };
    Plan* hhProduct(const Layout& LA, const Layout& LB, const Layout& LC, int maxRank,
                    int oversample, int mode = HH_PRODUCT);
    // C := A + alpha B blockwise (HMat2d::UpdateWith): dense leaves summed,
    // admissible blocks recompressed from a sketch of the exact sum.
    Plan* hhAdd(const Layout& LA, const Layout& LB, const Layout& LC, int maxRank, int oversample);

    int64_t wsAlloc(int64_t elems) {
        const int64_t o = ws_;
        ws_ += (elems + 63) & ~int64_t(63);
        return o;
    }
    int64_t wsUsed() const { return ws_; }
    // Workspace budget (elements) per H*H batch; the product is split into
    // batches of level-1 subtrees whose estimated workspace fits it.
    void setWorkspaceBudget(int64_t elems) { wsBudget_ = elems; }
    // Row-block height of the TSQR of tall sketch panels (0: 1024 for sketch
    // widths <= 24, else 512, the register-resident QR limits).
    void setTsqrRows(int rows) { tsqrRows_ = rows; }
    // Whether the low-rank factors of A / B are expected at full rank, so that
    // the exact-factor terms they enter are stacked (see hhProduct).
    void setStacking(bool a, bool b) { stackA_ = a; stackB_ = b; }
    // Sub-matrix products: C(rootC) = A(rootA) B(rootB) for blocks (t,u), (u,s),
    // (t,s) of one level (HMat2d::Multiply on child matrices,
    // MultiplyHMat-incl.hpp:12-380); only the blocks of C's subtree are written.
    void setRoots(int a, int b, int c) { rootA_ = a; rootB_ = b; rootC_ = c; }
    // Width-indirect sketches: every admissible C block sketches with
    // min(maxRank + oversample, widths[block]) columns, widths read at run time
    // from the SLOT_WIDTH array (indexed like C's ranks); the SVD then flags
    // (SLOT_FLAG array) blocks whose kept rank leaves fewer than `oversample`
    // sketch columns below the truncation tolerance.
    void setWidthIndirect(bool on) { widthIndirect_ = on; }

    // Sub-H product: out(|outBox| x k) (+)= op(M(b)) in(|inBox| x k). Leaves of
    // M inside block b contribute to the tiles of `out` (registered per atom in
    // outTiles); low-rank leaves first project into workspace in `proj`.
    // k may be rank-indirect (krank >= 0, kslot) with static maximum k.
    struct PanelTiles {
        int level, box;         // the box whose atoms the panel's rows cover
        std::vector<int> tileOfAtom; // tile id per atom of the box (atom - box.atomBegin)
        int kChunks;            // number of 64-column chunks
    };
    PanelTiles addPanelTiles(PhaseBuilder& pb, const Operand& out, int level, int box, int k,
                             int krank, uint8_t kslot, uint8_t flags);
    void subh(PhaseBuilder& proj, PhaseBuilder& acc, uint8_t mslot, const Layout& LM, int b,
              bool trans, const Operand& in, int inLevel, int inBox, const PanelTiles& out,
              int k, int krank, uint8_t kslot, float alpha, uint8_t cflags, double* flops);
    // out tiles += U(rows of t-atoms) * P for a low-rank factor pair given as
    // operands (generic helper used for pushdowns and exact low-rank terms).
private:
    const Geometry& g_;
    int64_t ws_ = 0;
    int64_t wsBudget_ = int64_t(1) << 31;
    int tsqrRows_ = 0;
    bool stackA_ = true, stackB_ = true;
    bool widthIndirect_ = false;
    int rootA_ = 0, rootB_ = 0, rootC_ = 0;
};

} // namespace dmhm
