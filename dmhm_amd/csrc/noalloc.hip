// Device scratch for the entry point that allocates nothing
// (dmhm_hmat_multiply_dense_noalloc): the streaming H*X plan's workspace
// (the projections P = V^T X of the low-rank blocks) lives in a file-scope
// device array instead of library-allocated memory. Every GPU that loads the
// library holds its own copy of the array until the process ends; its
// address on a device comes from hipGetSymbolAddress, taken once per device
// at dmhm_init with that device current. In host code the array's name is
// only a host-side stand-in, so it is never passed to a kernel or copy.
#include "kcommon.hpp"
#include <map>
#include <mutex>

namespace dmhm {

static constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB per GPU
static __device__ __align__(256) char scratch[kScratchBytes];

static std::mutex g_scratchMu;
static std::map<int, void*> g_scratchPtr;      // device ordinal -> address of `scratch` on it

void scratch_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (g_scratchPtr.count(device)) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
        throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratchPtr[device] = p;
}

void* scratch_address(int device, size_t* bytes)
{
    std::lock_guard<std::mutex> lk(g_scratchMu);
    auto it = g_scratchPtr.find(device);
    if (bytes) *bytes = kScratchBytes;
    return it == g_scratchPtr.end() ? nullptr : it->second;
}

} // namespace dmhm
