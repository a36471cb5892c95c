// scratch.hip -- the static device workspace of the allocation-free KPM entry
// point (leo2d_kpm_moments_noalloc).
//
// `scratch` is a file-scope device array of kScratchBytes: every GPU the
// library runs on holds one copy of it, at its own address, from the first use
// of this code object until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in symbol); it asks
// for the device address once per device with hipGetSymbolAddress, with that
// device current, and uses that address everywhere afterwards.
//
// Every call on one GPU shares that GPU's copy and finds the previous call's
// contents in it: the entry point clears what must start at zero.  Calls that
// overlap on one GPU (two host threads, two handles on one device) race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "kernels.h"

namespace leo2d {
namespace {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void* g_ptr[kMaxDevices] = {};

__global__ void zero_f64_kernel(double* p, int64_t count)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

}  // namespace

hipError_t scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_lock);
    if (g_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_ptr[device] = p;
    return hipSuccess;
}

void* scratch_ptr(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_lock);
    return g_ptr[device];
}

hipError_t launch_zero_f64(double* p, int64_t count, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_f64_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, count);
    return hipGetLastError();
}

}  // namespace leo2d
