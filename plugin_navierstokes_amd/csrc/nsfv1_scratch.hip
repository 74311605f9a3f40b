// nsfv1_scratch.hip -- the fixed device scratch of the no-allocation
// assembly entry point (nsfv1_assemble_noalloc).
//
// The split and coupled assembly paths keep per-SCVF records between their
// record kernel and their row gather.  nsfv1_assemble allocates that buffer
// with hipMalloc on first use; the no-allocation entry takes it from the
// file-scope array `scratch` below instead.  Each device that loads the
// library holds its own copy of the array until the process ends; its
// device address is looked up once per device with hipGetSymbolAddress
// (the host-side name `scratch` is a stand-in that must never reach a
// kernel) and cached here.
#include <hip/hip_runtime.h>

#include <mutex>

#include "nsfv1_internal.hpp"

namespace nsfv1 {

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
void* g_ptr[MAX_DEVICES] = {};
}  // namespace

hipError_t noalloc_scratch_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the array is looked up with its device current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_ptr[device] = p;
    return hipSuccess;
}

void* noalloc_scratch(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_mutex);
    return g_ptr[device];
}

}  // namespace nsfv1
