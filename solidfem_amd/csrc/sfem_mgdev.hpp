// sfem_mgdev.hpp -- device helpers of the multigrid V-cycle shared by its
// node kernels (sfem_mg.hip) and by the element kernel's multigrid staging
// (sfem_kernels.hip).  Both use these functions, so a Chebyshev step, a
// residual or a prolongation computed in either place is the same FP32
// expression and gives bitwise the same result.  Floating-point contraction
// is off inside them and every fused multiply-add is written out: left to
// the compiler, a product here could be fused with an add in the caller in
// one kernel and not in another.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfem {

// A level's operator image of node n: the element kernel's row for a node of
// one element block, the sum of the node's partial rows in slot order for a
// node shared by several blocks, 0 for a fixed node or a node of no element.
// node_q: {first slot, end slot} of a shared node; {-1, 0} node of one block;
// {-2, 0} fixed node or node of no element.  qrow: the node's row of the image.
__device__ __forceinline__ void complete_row(int2 nq, const float (&qrow)[3], const float* __restrict__ partial,
                                             float (&qv)[3])
{
#pragma unroll
    for (int c = 0; c < 3; c++) qv[c] = nq.x == -1 ? qrow[c] : 0.f;
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

__device__ __forceinline__ void operator_row(int64_t n, const float* __restrict__ q, const int2* __restrict__ node_q,
                                             const float* __restrict__ partial, float (&qv)[3])
{
    const int2 nq = node_q[n];
    float qr[3];
#pragma unroll
    for (int c = 0; c < 3; c++) qr[c] = q[3 * n + c];
    complete_row(nq, qr, partial, qv);
}

// z = M r for one node's 3 x 3 block (row-major)
__device__ __forceinline__ void block_mul(const float (&M)[9], const float (&r)[3], float (&z)[3])
{
#pragma clang fp contract(off)
#pragma unroll
    for (int c = 0; c < 3; c++) z[c] = fmaf(M[3 * c + 2], r[2], fmaf(M[3 * c + 1], r[1], M[3 * c] * r[0]));
}

__device__ __forceinline__ void block_apply(const float* __restrict__ minv, int64_t n, const float (&r)[3], float (&z)[3])
{
    float M[9];
#pragma unroll
    for (int k = 0; k < 9; k++) M[k] = minv[9 * n + k];
    block_mul(M, r, z);
}

// the Chebyshev direction of one dof: c1 d_old + c2 (M r) for a next step,
// c2 (M r) for a first step
__device__ __forceinline__ float cheb_dir(bool next, float c1, float d_old, float c2, float mr)
{
#pragma clang fp contract(off)
    const float t = c2 * mr;
    return next ? fmaf(c1, d_old, t) : t;
}

// acc += sum over entries j0 .. j1-1 (in order) of val[j] * v[idx[j]] (3-vectors);
// eight entries' index loads, then their row loads, are in flight at once
__device__ __forceinline__ void gather_rows(int j0, int j1, const int* __restrict__ idx, const float* __restrict__ val,
                                            const float* __restrict__ v, float (&acc)[3])
{
    for (int j = j0; j < j1; j += 8) {
        int id[8];
        float w[8], t[8][3];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const bool ok = j + q < j1;
            id[q] = ok ? idx[j + q] : -1;
            w[q] = ok ? val[j + q] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < 8; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) t[q][c] = id[q] >= 0 ? v[3 * (size_t)id[q] + c] : 0.f;
#pragma unroll
        for (int q = 0; q < 8; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) acc[c] = fmaf(w[q], t[q][c], acc[c]);
    }
}

}  // namespace sfem
