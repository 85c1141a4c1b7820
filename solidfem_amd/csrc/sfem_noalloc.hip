// sfem_noalloc.hip -- the allocation-free Hessian-vector entry point,
// sfem_matvec_noalloc, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// sfem_matvec stages the caller's direction and result in context buffers and
// host vectors and sums shared-node partial rows in the context's partial
// buffer.  This variant keeps all of that -- the FP64 host staging, the FP32
// direction and result, and the partial rows -- in one file-scope device
// array, `scratch`, of NOALLOC_SCRATCH_BYTES per device, and calls no
// allocating routine (device or host).  The array's device address is looked
// up once per device (hipGetSymbolAddress, at context creation with the
// context's device current) and only that address is handed to kernels and
// copies.  Every call on one device shares that device's array: calls that
// overlap on one device (several threads, several contexts) race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "sfem_capi.h"
#include "sfem_internal.hpp"
#include "sfem_kernels.hpp"

namespace sfem {

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;

// per device: the FP64 staging (3 nn), the FP32 direction and result (3 nn
// each) and the shared-node partial rows (3 per slot)
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};   // device address of `scratch`, per device

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// d32 = (float) d64 on free rows, 0 on fixed rows; q32 = 0 (rows of nodes that
// no element references are written by neither element kernel)
__global__ __launch_bounds__(256) void k_na_in(int64_t nn, const double* __restrict__ d64, const uint8_t* __restrict__ fixed,
                                               float* __restrict__ d32, float* __restrict__ q32)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x) {
        const bool f = fixed[i] != 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            d32[3 * i + c] = f ? 0.f : (float)d64[3 * i + c];
            q32[3 * i + c] = 0.f;
        }
    }
}

__global__ __launch_bounds__(256) void k_na_out(int64_t n3, const float* __restrict__ q32, double* __restrict__ out64)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x)
        out64[i] = (double)q32[i];
}

int grid256(int64_t n)
{
    int64_t g = (n + 255) / 256;
    return (int)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

}  // namespace

size_t noalloc_matvec_bytes(const sfem_ctx* c)
{
    const size_t n3 = 3 * (size_t)c->nn;
    return align256(n3 * sizeof(double)) + 2 * align256(n3 * sizeof(float)) + align256(3 * (size_t)c->nslots * sizeof(float));
}

// the scratch array's address on the current device, looked up once per device
int noalloc_init(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(SFEM_EDEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return SFEM_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SFEM_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
    return SFEM_OK;
}

}  // namespace sfem

extern "C" int sfem_matvec_noalloc(sfem_ctx* c, const double* d, double* Kd)
{
    using namespace sfem;
    if (!c || !d || !Kd) return fail(SFEM_EINVAL, "sfem_matvec_noalloc: null argument");
    const size_t need = noalloc_matvec_bytes(c);
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(SFEM_EINVAL, "sfem_matvec_noalloc: mesh needs " + std::to_string(need) + " bytes of scratch, above the " +
                                     std::to_string((size_t)NOALLOC_SCRATCH_BYTES) + "-byte limit (use sfem_matvec)");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES) base = g_scratch_ptr[c->device];
    }
    if (!base) return fail(SFEM_ESTATE, "sfem_matvec_noalloc: scratch not set up on the context's device");
    return guard_call("sfem_matvec_noalloc", [&]() -> int {
        const int64_t nn = c->nn;
        const size_t n3 = 3 * (size_t)nn;
        unsigned char* b = static_cast<unsigned char*>(base);
        double* stage64 = reinterpret_cast<double*>(b);
        float* d32 = reinterpret_cast<float*>(b + align256(n3 * sizeof(double)));
        float* q32 = reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(d32) + align256(n3 * sizeof(float)));
        float* part = reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(q32) + align256(n3 * sizeof(float)));
        hipStream_t s = c->stream;
        HIPCK(hipMemcpyAsync(stage64, d, n3 * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_na_in, dim3(grid256(nn)), dim3(256), 0, s, nn, stage64, c->d_fixed.as<uint8_t>(), d32, q32);
        HIPCK(hipGetLastError());
        HIPCK(launch_hessvec(c->material, c->blk, c->shr, c->d_u32.as<float>(), d32, q32, part, nullptr, s));
        hipLaunchKernelGGL(k_na_out, dim3(grid256((int64_t)n3)), dim3(256), 0, s, (int64_t)n3, q32, stage64);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(Kd, stage64, n3 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        return SFEM_OK;
    });
}
