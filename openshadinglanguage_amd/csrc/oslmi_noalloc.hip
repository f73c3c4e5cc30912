// The fixed device workspace of oslmi_batch_launch_noalloc: a file-scope
// device array that every call on a GPU shares, so that the entry point
// itself never allocates device memory.
//
// The array lives in the library's code object; its device address differs
// per GPU and is looked up once per device (hipGetSymbolAddress, from
// oslmi_init with that device current). Host code only ever uses that
// address: the array's name in host code is a stand-in, not a device
// pointer.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "oslmi.h"
#include "oslmi_group.h"
#include "oslmi_internal.h"

namespace oslmi {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// zero-fills n ints (the tile counters a launch starts from); a kernel
// rather than a memset, so the work stays an ordinary kernel on the stream
__global__ __launch_bounds__(256) void k_noalloc_clear(int32_t* __restrict__ p, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        p[i] = 0;
}

namespace {
std::mutex g_mu;
std::vector<void*> g_scratch;  // per device: the array's device address
}  // namespace

int noalloc_scratch_init()
{
    int dev = 0;
    if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice"))
        return rc;
    std::lock_guard<std::mutex> lock(g_mu);
    if ((int)g_scratch.size() <= dev)
        g_scratch.resize(dev + 1, nullptr);
    if (g_scratch[dev])
        return OSLMI_OK;
    void* p = nullptr;
    if (int rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                           "hipGetSymbolAddress(no-alloc scratch)"))
        return rc;
    if (!p)
        return set_error(OSLMI_ERR_HIP, "hipGetSymbolAddress(no-alloc scratch): null address");
    g_scratch[dev] = p;
    return OSLMI_OK;
}

int noalloc_scratch(unsigned char** out)
{
    int dev = 0;
    if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice"))
        return rc;
    std::lock_guard<std::mutex> lock(g_mu);
    if ((int)g_scratch.size() <= dev || !g_scratch[dev])
        return set_error(OSLMI_ERR_INVALID,
                         "no-alloc workspace of this device not set up: call oslmi_init with the "
                         "device current first");
    *out = (unsigned char*)g_scratch[dev];
    return OSLMI_OK;
}

int noalloc_clear(int32_t* p, int n, hipStream_t stream)
{
    if (n <= 0)
        return OSLMI_OK;
    hipLaunchKernelGGL(k_noalloc_clear, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p,
                       n);
    return check_launch("oslmi_batch_launch_noalloc (clear counters)");
}

}  // namespace oslmi
