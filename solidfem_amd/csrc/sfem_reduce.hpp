// sfem_reduce.hpp -- device helpers shared by the kernel files: fixed-order
// FP64 workgroup reductions and their cross-workgroup partials.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfem {

// an int read through the vector memory path even where its address is
// uniform: a scalar load would be waited for with every scalar kernel-argument
// load after it (scalar loads complete out of order, so the compiler waits for
// all of them), which puts its round trip in front of unrelated loads
__device__ __forceinline__ int vload(const int* f)
{
    return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// one 64-bit DPP move of a double (two 32-bit moves); lanes whose source is
// outside the pattern, and rows that row_mask leaves out, read 0
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, true);
    return __hiloint2double(hi, lo);
}

// the sum of v over the wave's 64 lanes, returned to every lane, in a fixed
// order: row_shr 1, 2, 4, 8 leave each 16-lane row's sum in its lane 15,
// row_bcast 15 and 31 add the rows into lane 63 (VALU moves: no LDS round trip)
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64<0x111, 0xf>(v);    // row_shr:1
    v += dpp_f64<0x112, 0xf>(v);    // row_shr:2
    v += dpp_f64<0x114, 0xf>(v);    // row_shr:4
    v += dpp_f64<0x118, 0xf>(v);    // row_shr:8
    v += dpp_f64<0x142, 0xa>(v);    // row_bcast:15 into rows 1, 3
    v += dpp_f64<0x143, 0xc>(v);    // row_bcast:31 into rows 2, 3
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// BS: the workgroup size the kernel is launched with (a constant, so that no
// dispatch-packet load and no wait on it precede the reduction's loads).
// Thread 0 ends with the workgroup's sums: each wave's sum, then the waves'
// sums in wave order.
template <int NV, int BS>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /* >= NV*16 */)
{
    static_assert(BS % 64 == 0 && BS <= 1024, "whole waves");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = BS >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = wave_sum(v[q]);
    if (nw == 1) return;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) lds[q * 16 + w] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double t = lds[q * 16];
            for (int k = 1; k < nw; k++) t += lds[q * 16 + k];
            v[q] = t;
        }
    }
    __syncthreads();
}

// this workgroup's NV sums -> part[q * gridDim.x + blockIdx.x]
template <int NV, int BS>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* part)
{
    __shared__ double lds[NV * 16];
    block_sum<NV, BS>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) part[(size_t)q * gridDim.x + blockIdx.x] = v[q];
}

// every thread of the workgroup gets sum_i part[q * stride + i], i < n, in a
// fixed order (thread-strided partial sums, then the fixed block tree).
// The partials were written by workgroups on every XCD, so each load misses
// this XCD's L2: U loads per thread are issued before the first add, so a
// thread waits for ceil(n / (U * BS)) round trips, not n / BS.
// In two steps, so that a kernel can issue other loads between the first
// round of partial loads and the wait for them: partials_issue loads the first
// U * BS partials, partials_finish adds them, sums the rest and the block tree.
template <int NV, int BS, int U>
__device__ __forceinline__ void partials_issue(const double* part, int n, int stride, double (&v)[NV][U])
{
#pragma unroll
    for (int q = 0; q < NV; q++)
#pragma unroll
        for (int k = 0; k < U; k++) {
            const int i = threadIdx.x + k * BS;
            v[q][k] = i < n ? part[(size_t)q * stride + i] : 0.0;
        }
}

template <int NV, int BS, int U>
__device__ __forceinline__ void partials_finish(const double* part, int n, int stride, const double (&v0)[NV][U],
                                                double (&tot)[NV])
{
    __shared__ double lds[NV * 16];
    __shared__ double res[NV];
    // starting from v0[q][0] rather than 0.0 + v0[q][0]: an add of a constant
    // to a conditional load is folded into the load's branch, which then waits
    // for the load where it is issued
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        acc[q] = v0[q][0];
#pragma unroll
        for (int k = 1; k < U; k++) acc[q] += v0[q][k];
    }
    for (int base = threadIdx.x + BS * U; base < n; base += BS * U) {
        double v[NV][U];
#pragma unroll
        for (int q = 0; q < NV; q++)
#pragma unroll
            for (int k = 0; k < U; k++) {
                const int i = base + k * BS;
                v[q][k] = i < n ? part[(size_t)q * stride + i] : 0.0;
            }
#pragma unroll
        for (int q = 0; q < NV; q++)
#pragma unroll
            for (int k = 0; k < U; k++) acc[q] += v[q][k];
    }
    block_sum<NV, BS>(acc, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) res[q] = acc[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; q++) tot[q] = res[q];
}

template <int NV, int BS, int U = 16>
__device__ __forceinline__ void sum_partials(const double* part, int n, int stride, double (&tot)[NV])
{
    double v[NV][U];
    partials_issue<NV, BS, U>(part, n, stride, v);
    partials_finish<NV, BS, U>(part, n, stride, v, tot);
}

}  // namespace sfem
