// Internal runtime of libdmhm_amd.so shared by its host translation units:
// the library state (device, stream, workspace), per-geometry contexts
// with their plan caches, device H-matrix storage, and the plan runner.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dmhm_amd.h"
#include "geometry.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace dmhm {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(x)                                                                       \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess)                                                             \
            throw Error(e_ == hipErrorOutOfMemory ? DMHM_ENOMEM : DMHM_EHIP,              \
                        std::string(#x) + ": " + hipGetErrorString(e_));                  \
    } while (0)

// ---------------------------------------------------------------------------
// Library state: one device, one stream (default priority), a growing
// workspace and per-geometry plan caches.
// ---------------------------------------------------------------------------
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    void ensure(size_t b) {
        if (b <= bytes) return;
        if (ptr) HIPCHECK(hipFree(ptr));
        ptr = nullptr; bytes = 0;
        HIPCHECK(hipMalloc(&ptr, b));
        bytes = b;
    }
    void release() { if (ptr) (void)hipFree(ptr); ptr = nullptr; bytes = 0; }
};

struct GeoCtx;

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() { HIPCHECK(hipEventCreate(&a)); HIPCHECK(hipEventCreate(&b)); }
    ~EventPair() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    float ms() { float t = 0; HIPCHECK(hipEventElapsedTime(&t, a, b)); return t; }
};

struct Library {
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    DeviceBuffer ws;            // plan workspace
    DeviceBuffer stageX, stageY;
    DeviceBuffer featBuf, widthBuf, flagBuf;   // rank-model features, sketch widths, check flags
    double* dScalars = nullptr; // power-iteration scalars
    std::vector<std::weak_ptr<GeoCtx>> geos;
    std::mutex mu;
    int nextLayoutId = 100;
    // Tile engine of every TILE phase: -1 autotunes per phase (default);
    // DMHM_TILE_ENGINE=lds|mfma|mfma_pf|bf16|bf16_pf forces one.
    int tileEngine = -1;
    bool bf16 = false;          // BF16 operands allowed (set inside the mixed Schulz loop)
    long untuned = 0;           // tile launches made while tuning (graph capture waits for none)
    EventPair* tuneEv = nullptr;
    // Device iteration counter added to every sketch and start-vector seed
    // while a replayable loop runs (null otherwise), and a pinned word for the
    // loop's one readback per iteration.
    const uint64_t* seedOff = nullptr;
    uint64_t* dCounter = nullptr;
    double* hReadback = nullptr;
};
extern Library g_lib;

// Copies on the library stream and waits for them: the stream is created
// non-blocking, so a plain hipMemcpy (null stream) would not be ordered after
// work already queued on it (for example the zero fill of a new matrix).
hipError_t copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);

inline void require_ready()
{
    if (!g_lib.ready) throw Error(DMHM_ENODEV, "dmhm_init has not been called");
}

struct DeviceTables {        // small index tables for per-leaf kernels
    void* ptr = nullptr;
    ~DeviceTables() { if (ptr) (void)hipFree(ptr); }
};

struct OperatorStats;
struct DinvProgram;
struct GeoCtx {
    Geometry geo;
    std::shared_ptr<OperatorStats> zeroStats;     // for matrices without an operator's statistics
    std::map<std::string, std::unique_ptr<Plan>> plans;
    std::map<int, std::shared_ptr<Layout>> stdLayouts;      // by maxRank
    // diagonal dense leaves (t == s) for AddConstantToDiagonal, per layout id
    std::map<int, std::pair<int64_t*, int*>> diagTables;
    std::map<int, int> diagCount;
    // Layout of the transpose of a matrix of layout id (when the capacities move).
    std::map<int, std::shared_ptr<Layout>> transLayouts;
    // Direct-inversion programs by (layout, rank, oversample).
    std::map<std::string, std::shared_ptr<DinvProgram>> dinvPrograms;
    // Device copy of the block tree (CSR import, index map), built on first use.
    DeviceBuffer devTree;
    GeoDev dev;
    bool devReady = false;
    GeoCtx(int dim, const int* sizes, int L, bool strong) : geo(dim, sizes, L, strong) {}
    ~GeoCtx() {
        for (auto& kv : diagTables) { (void)hipFree(kv.second.first); (void)hipFree(kv.second.second); }
        devTree.release();
    }
};
// Coefficient statistics of an operator for the rank model: per box of the
// geometry (every level; box (level, id) at levelBase[level] + id), the mean
// and standard deviation of log|a_ii| over the box's rows, taken from the
// diagonal of the imported CSR matrix; and, built on first use, the static
// features of every admissible block on the device.
struct OperatorStats {
    std::vector<int> levelBase;
    std::vector<float> mean, sdev;
    float globalMean = 0;
    DeviceBuffer feat;          // nLowRank x kStaticFeatures floats
    bool featReady = false;
    ~OperatorStats() { feat.release(); }
};
} // namespace dmhm

// A device H-matrix: one allocation of [dense | U | V] values in its layout
// and one rank per admissible block.
struct dmhm_hmat_s {
    std::shared_ptr<dmhm::GeoCtx> gc;
    std::shared_ptr<dmhm::Layout> lay;
    std::shared_ptr<dmhm::OperatorStats> stats;   // of the operator this matrix derives from (or null)
    int prec = DMHM_F64;
    int maxRank = 0;
    void* vals = nullptr;
    int* ranks = nullptr;
    // Admissible blocks expected at full capacity rank (products, iterates,
    // SetToRandom); false for imports and zero matrices, whose rank-0 blocks
    // the product plans then skip at run time instead of stacking them.
    bool full = false;
    size_t esize() const { return prec == DMHM_F64 ? 8 : 4; }
    // Views into a larger allocation (a batch arena) do not own their storage.
    bool owner = true;
    ~dmhm_hmat_s() {
        if (!owner) return;
        if (vals) (void)hipFree(vals);
        if (ranks) (void)hipFree(ranks);
    }
    const dmhm::Geometry& geo() const { return gc->geo; }
};

namespace dmhm {

const GeoDev& geo_dev(GeoCtx& gc);
std::shared_ptr<GeoCtx> get_geo(int dim, const int* sizes, int L, bool strong);
// The layout of products and iterates (every low-rank capacity = maxRank).
std::shared_ptr<Layout> std_layout(GeoCtx& gc, int maxRank);
dmhm_hmat_s* alloc_hmat(std::shared_ptr<GeoCtx> gc, std::shared_ptr<Layout> lay, int prec, int maxRank);
dmhm_hmat_s* copy_hmat(dmhm_hmat_s* s, int prec);
dmhm_hmat_s* transpose_hmat(dmhm_hmat_s* s);
void scale_hmat(dmhm_hmat_s* h, double alpha);
void add_diag(dmhm_hmat_s* h, double alpha);
void update_with(dmhm_hmat_s* A, double alpha, dmhm_hmat_s* B, const dmhm_hh_opts& o);
bool stall_detected(const std::vector<double>& history);
double estimate_two_norm(dmhm_hmat_s* h, double theta, double confidence, uint64_t seed);
// HMat2d::DirectInvert (dinv.cpp): A := inv(A) in A's precision.
void direct_invert(dmhm_hmat_s* A, const dmhm_hh_opts& o);
void direct_invert_stats(dmhm_hmat_s* A, int oversample, long* products, long* plans, double* seconds);
void check_direct_invert_program(const Geometry& g, int maxRank, int over, long long* stats);
int guard(const std::function<void()>& fn);
// Frees the merged plans of batches (batch.cpp), at dmhm_finalize.
void release_batch_plans();

// Plan cache of a geometry (built and uploaded on first use) and the shared
// workspace (grown to a plan's need).
Plan& get_plan(GeoCtx& gc, const std::string& key, const std::function<Plan*()>& make);
void ensure_ws(const Plan& p, size_t esize);
int64_t hh_batch_budget();
// Launches the phases of a plan on the library stream (tile phases with the
// forced or autotuned engine).
template <typename T>
void run_phase(Phase& ph, const Bases& b, T alpha, T beta, TruncPolicy pol, uint64_t seed);
template <typename T>
void run_plan(Plan& p, const Bases& b, T alpha, T beta, TruncPolicy pol, uint64_t seed);

// Makes the device iteration counter offset every seed for its lifetime.
struct SeedOffsetScope {
    SeedOffsetScope() { g_lib.seedOff = g_lib.dCounter; }
    ~SeedOffsetScope() { g_lib.seedOff = nullptr; }
};

// A captured sequence of launches on the library stream and its executable
// graph (a Schulz step, per direction of the iterate ping-pong).
struct StepGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    StepGraph() = default;
    StepGraph(const StepGraph&) = delete;
    StepGraph& operator=(const StepGraph&) = delete;
    ~StepGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
    void capture(const std::function<void()>& launch) {
        HIPCHECK(hipStreamBeginCapture(g_lib.stream, hipStreamCaptureModeThreadLocal));
        try {
            launch();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(g_lib.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHECK(hipStreamEndCapture(g_lib.stream, &graph));
        HIPCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    void replay() { HIPCHECK(hipGraphLaunch(exec, g_lib.stream)); }
};

} // namespace dmhm
