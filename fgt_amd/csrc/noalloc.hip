// noalloc.hip -- the device workspace of the no-allocation entry point.
//
// fgt_boxfgt_noalloc (api.cpp) takes every device buffer it needs beyond
// the caller's inputs and outputs -- plan tables, work lists and plane-wave
// expansion slots -- from the file-scope array below instead of hipMalloc.
// Each GPU the library runs on holds its own copy of the array, at its own
// address, until the process ends.  Host code never uses the array's name
// as an address: scratch_region() looks the address up once per device
// with hipGetSymbolAddress and hands out that pointer.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "kernels.h"

namespace fgt {

namespace {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

__global__ void __launch_bounds__(256) k_zero(unsigned long long* __restrict__ p, size_t nwords) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nwords) p[i] = 0ULL;
}

}  // namespace

void* scratch_region(size_t* bytes) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[dev] = p;
    }
    if (bytes) *bytes = kScratchBytes;
    return g_ptr[dev];
}

void launch_zero(void* p, size_t bytes, hipStream_t s) {
    if (bytes == 0) return;
    if ((reinterpret_cast<size_t>(p) & 7) || (bytes & 7)) throw std::runtime_error("launch_zero: unaligned range");
    const size_t nw = bytes / 8;
    const size_t nb = (nw + 255) / 256;
    if (nb > 0x7fffffffULL) throw std::runtime_error("launch_zero: range too large");
    hipLaunchKernelGGL(k_zero, dim3((unsigned)nb), dim3(256), 0, s, static_cast<unsigned long long*>(p), nw);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("k_zero launch: ") + hipGetErrorString(e));
}

}  // namespace fgt
