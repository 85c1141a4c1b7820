// sfem_mg.hip -- gfx950 kernels of the multigrid preconditioner: a V-cycle
// over a hierarchy of nested tetrahedral meshes, each level with its own
// matrix-free operator (the Hessian-vector element kernel at the level's
// state), Chebyshev smoothing on the level's 3x3 block preconditioner, and
// prolongation / restriction as fixed-order CSR gathers (no atomics, so the
// preconditioner is bitwise reproducible).
//
// One thread per node in every kernel.  A level's operator image of node n is
// assembled as in the CG update: the element kernel's row for nodes of one
// element block, the sum of the node's partial rows (slot order) for nodes
// shared by several blocks, 0 for fixed nodes.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sfem_kernels.hpp"
#include "sfem_mg.hpp"
#include "sfem_reduce.hpp"

namespace sfem {

constexpr int MG_BLOCK = 256;

static int mg_grid(int64_t n) { return (int)((n + MG_BLOCK - 1) / MG_BLOCK); }

// node_q: {first slot, end slot} of a shared node; {-1, 0} node of one block;
// {-2, 0} fixed node or node of no element
__device__ __forceinline__ void operator_row(int64_t n, const float* __restrict__ q, const int2* __restrict__ node_q,
                                             const float* __restrict__ partial, float (&qv)[3])
{
    const int2 nq = node_q[n];
#pragma unroll
    for (int c = 0; c < 3; c++) qv[c] = q[3 * n + c];
#pragma unroll
    for (int c = 0; c < 3; c++) qv[c] = nq.x == -1 ? qv[c] : 0.f;
    if (nq.x >= 0) {
        const int k1 = nq.y;
        for (int k = nq.x; k < k1; k += 4) {
            float t[4][3];
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int c = 0; c < 3; c++) t[j][c] = k + j < k1 ? partial[3 * (size_t)(k + j) + c] : 0.f;
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int c = 0; c < 3; c++) qv[c] += t[j][c];
        }
    }
}

__device__ __forceinline__ void block_apply(const float* __restrict__ minv, int64_t n, const float (&r)[3], float (&z)[3])
{
    float M[9];
#pragma unroll
    for (int k = 0; k < 9; k++) M[k] = minv[9 * n + k];
#pragma unroll
    for (int c = 0; c < 3; c++) z[c] = M[3 * c] * r[0] + M[3 * c + 1] * r[1] + M[3 * c + 2] * r[2];
}

// One Chebyshev step (q = the level operator applied to the previous d, or to
// the coarse correction e for MODE 2):
//   MODE 0: r = b;      d = c2 M r;         x = d        first step from x = 0
//   MODE 1: r -= q;     d = c1 d + c2 M r;  x += d       next step
//   MODE 2: r -= q;     d = c2 M r;         x += d       first step from x (after x += e)
//   MODE 3: r -= q                                       residual after the last step
template <int MODE>
__global__ __launch_bounds__(MG_BLOCK) void k_mg_cheb(int64_t nn, const float* __restrict__ b, float* __restrict__ r,
                                                      float* __restrict__ d, float* __restrict__ x,
                                                      const float* __restrict__ q, const int2* __restrict__ node_q,
                                                      const float* __restrict__ partial, const float* __restrict__ minv,
                                                      float c1, float c2)
{
    const int64_t i = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (i >= nn) return;
    float rv[3];
    if (MODE == 0) {
#pragma unroll
        for (int c = 0; c < 3; c++) rv[c] = b[3 * i + c];
    } else {
        float qv[3];
        operator_row(i, q, node_q, partial, qv);
#pragma unroll
        for (int c = 0; c < 3; c++) rv[c] = r[3 * i + c] - qv[c];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) r[3 * i + c] = rv[c];
    if (MODE == 3) return;
    float mr[3];
    block_apply(minv, i, rv, mr);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float dv = MODE == 1 ? c1 * d[3 * i + c] + c2 * mr[c] : c2 * mr[c];
        d[3 * i + c] = dv;
        x[3 * i + c] = MODE == 0 ? dv : x[3 * i + c] + dv;
    }
}

// e = P xc on the level above (0 at its fixed nodes); x += e
__global__ __launch_bounds__(MG_BLOCK) void k_mg_prolong(int64_t nn, const int* __restrict__ ptr, const int* __restrict__ col,
                                                         const float* __restrict__ val, const float* __restrict__ xc,
                                                         const uint8_t* __restrict__ fixed, float* __restrict__ e,
                                                         float* __restrict__ x)
{
    const int64_t i = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (i >= nn) return;
    float ev[3] = {0.f, 0.f, 0.f};
    const int k1 = ptr[i + 1];
    for (int k = ptr[i]; k < k1; k++) {
        const float w = val[k];
        const int j = col[k];
#pragma unroll
        for (int c = 0; c < 3; c++) ev[c] += w * xc[3 * (size_t)j + c];
    }
    const bool f = fixed[i] != 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float v = f ? 0.f : ev[c];
        e[3 * i + c] = v;
        x[3 * i + c] += v;
    }
}

// bc = P^T r: rows of the transpose (coarse nodes) in increasing fine-node
// order; 0 at fixed coarse nodes
__global__ __launch_bounds__(MG_BLOCK) void k_mg_restrict(int64_t nnc, const int* __restrict__ ptr, const int* __restrict__ row,
                                                          const float* __restrict__ val, const float* __restrict__ r,
                                                          const uint8_t* __restrict__ fixed, float* __restrict__ bc)
{
    const int64_t k = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (k >= nnc) return;
    float s[3] = {0.f, 0.f, 0.f};
    const int j1 = ptr[k + 1];
    for (int j = ptr[k]; j < j1; j++) {
        const float w = val[j];
        const int i = row[j];
#pragma unroll
        for (int c = 0; c < 3; c++) s[c] += w * r[3 * (size_t)i + c];
    }
    const bool f = fixed[k] != 0;
#pragma unroll
    for (int c = 0; c < 3; c++) bc[3 * k + c] = f ? 0.f : s[c];
}

// the coarse level's state: the displacement of the level above at the node
// that coincides with each coarse node
__global__ __launch_bounds__(MG_BLOCK) void k_mg_inject(int64_t nnc, const int* __restrict__ inj, const double* __restrict__ ua,
                                                        double* __restrict__ u64, float* __restrict__ u32)
{
    const int64_t k = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (k >= nnc) return;
    const int64_t i = inj[k];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double v = ua[3 * i + c];
        u64[3 * k + c] = v;
        u32[3 * k + c] = (float)v;
    }
}

// power iteration step for the largest eigenvalue of M A: y = M (A v)
__global__ __launch_bounds__(MG_BLOCK) void k_mg_power(int64_t nn, const float* __restrict__ q, const int2* __restrict__ node_q,
                                                       const float* __restrict__ partial, const float* __restrict__ minv,
                                                       float* __restrict__ y)
{
    const int64_t i = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (i >= nn) return;
    float qv[3], yv[3];
    operator_row(i, q, node_q, partial, qv);
    block_apply(minv, i, qv, yv);
#pragma unroll
    for (int c = 0; c < 3; c++) y[3 * i + c] = yv[c];
}

// a start vector for the power iteration: a hash of the dof index in [-1, 1),
// 0 at fixed nodes
__global__ __launch_bounds__(MG_BLOCK) void k_mg_start(int64_t nn, const uint8_t* __restrict__ fixed, float* __restrict__ v)
{
    const int64_t i = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (i >= nn) return;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        uint64_t z = (uint64_t)(3 * i + c) + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const float u = (float)(z >> 40) * (1.0f / 16777216.0f);
        v[3 * i + c] = fixed[i] ? 0.f : 2.f * u - 1.f;
    }
}

// partials of a.a (FP64) per workgroup, then their fixed-order sum to out[0]
__global__ __launch_bounds__(MG_BLOCK) void k_mg_sqnorm(int64_t n3, const float* __restrict__ a, double* __restrict__ part)
{
    double v[1] = {0.0};
    for (int64_t i = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x; i < n3; i += (int64_t)gridDim.x * MG_BLOCK)
        v[0] += (double)a[i] * a[i];
    block_partials<1, MG_BLOCK>(v, part);
}

__global__ __launch_bounds__(256) void k_mg_finish(const double* __restrict__ part, int n, double* __restrict__ out)
{
    double t[1];
    sum_partials<1, 256>(part, n, n, t);
    if (threadIdx.x == 0) out[0] = t[0];
}

// r.z and r.r partials in the CG update's layout (2 x gridDim.x)
__global__ __launch_bounds__(CG_BLOCK) void k_mg_rz(int64_t nn, const float* __restrict__ r, const float* __restrict__ z,
                                                    double* __restrict__ part)
{
    double v[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)CG_BLOCK + threadIdx.x; i < nn; i += (int64_t)gridDim.x * CG_BLOCK) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double rv = r[3 * i + c];
            v[0] += rv * z[3 * i + c];
            v[1] += rv * rv;
        }
    }
    block_partials<2, CG_BLOCK>(v, part);
}

__global__ __launch_bounds__(MG_BLOCK) void k_mg_copy(int64_t n, const float* __restrict__ a, float* __restrict__ b)
{
    const int64_t i = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// ------------------------------------------------------------------ launchers

hipError_t launch_mg_cheb(int mode, int64_t nn, const float* b, float* r, float* d, float* x, const float* q,
                          const int2* node_q, const float* partial, const float* minv, float c1, float c2, hipStream_t st)
{
    const dim3 g(mg_grid(nn)), bl(MG_BLOCK);
    switch (mode) {
    case 0: hipLaunchKernelGGL(k_mg_cheb<0>, g, bl, 0, st, nn, b, r, d, x, q, node_q, partial, minv, c1, c2); break;
    case 1: hipLaunchKernelGGL(k_mg_cheb<1>, g, bl, 0, st, nn, b, r, d, x, q, node_q, partial, minv, c1, c2); break;
    case 2: hipLaunchKernelGGL(k_mg_cheb<2>, g, bl, 0, st, nn, b, r, d, x, q, node_q, partial, minv, c1, c2); break;
    default: hipLaunchKernelGGL(k_mg_cheb<3>, g, bl, 0, st, nn, b, r, d, x, q, node_q, partial, minv, c1, c2); break;
    }
    return hipGetLastError();
}

hipError_t launch_mg_prolong(int64_t nn, const int* ptr, const int* col, const float* val, const float* xc,
                             const uint8_t* fixed, float* e, float* x, hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_prolong, dim3(mg_grid(nn)), dim3(MG_BLOCK), 0, st, nn, ptr, col, val, xc, fixed, e, x);
    return hipGetLastError();
}

hipError_t launch_mg_restrict(int64_t nnc, const int* ptr, const int* row, const float* val, const float* r,
                              const uint8_t* fixed, float* bc, hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_restrict, dim3(mg_grid(nnc)), dim3(MG_BLOCK), 0, st, nnc, ptr, row, val, r, fixed, bc);
    return hipGetLastError();
}

hipError_t launch_mg_inject(int64_t nnc, const int* inj, const double* ua, double* u64, float* u32, hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_inject, dim3(mg_grid(nnc)), dim3(MG_BLOCK), 0, st, nnc, inj, ua, u64, u32);
    return hipGetLastError();
}

hipError_t launch_mg_power(int64_t nn, const float* q, const int2* node_q, const float* partial, const float* minv, float* y,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_power, dim3(mg_grid(nn)), dim3(MG_BLOCK), 0, st, nn, q, node_q, partial, minv, y);
    return hipGetLastError();
}

hipError_t launch_mg_start(int64_t nn, const uint8_t* fixed, float* v, hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_start, dim3(mg_grid(nn)), dim3(MG_BLOCK), 0, st, nn, fixed, v);
    return hipGetLastError();
}

int mg_sqnorm_grid(int64_t n3)
{
    const int64_t g = (n3 + MG_BLOCK - 1) / MG_BLOCK;
    return (int)(g < 1 ? 1 : (g > 1024 ? 1024 : g));
}

hipError_t launch_mg_sqnorm(int64_t n3, const float* a, double* part, double* out, hipStream_t st)
{
    const int g = mg_sqnorm_grid(n3);
    hipLaunchKernelGGL(k_mg_sqnorm, dim3(g), dim3(MG_BLOCK), 0, st, n3, a, part);
    hipLaunchKernelGGL(k_mg_finish, dim3(1), dim3(256), 0, st, part, g, out);
    return hipGetLastError();
}

hipError_t launch_mg_rz(int64_t nn, const float* r, const float* z, double* part, hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_rz, dim3(cg_grid(nn)), dim3(CG_BLOCK), 0, st, nn, r, z, part);
    return hipGetLastError();
}

hipError_t launch_mg_copy(int64_t n, const float* a, float* b, hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_copy, dim3(mg_grid(n)), dim3(MG_BLOCK), 0, st, n, a, b);
    return hipGetLastError();
}

}  // namespace sfem
