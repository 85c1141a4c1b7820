// sfem_gnn.hip -- kernels of the learned block preconditioner: a message-passing
// network over the mesh node graph (3 rounds, width 32) whose MLP layers are
// BF16 x BF16 -> FP32 GEMMs on the matrix cores (gfx950's double-K
// v_mfma_f32_16x16x32_bf16 where the reduction length is a multiple of 32,
// v_mfma_f32_16x16x16_bf16 for 16), its backward pass, the SPD block
// construction, the training loss and Adam.
//
// MFMA operand layouts (one wave, lane l):
//   16x16x16, item q = 0..3:  A[i][k]: i = l % 16, k = 4 (l / 16) + q;  B[k][j]: j = l % 16, k = 4 (l / 16) + q
//   16x16x32, item q = 0..7:  A[i][k]: i = l % 16, k = 8 (l / 16) + q;  B[k][j]: j = l % 16, k = 8 (l / 16) + q
//   both:                     D[i][j]: j = l % 16, i = 4 (l / 16) + q (q = 0..3)
// Every reduction over nodes writes per-workgroup partials that one kernel
// sums in index order, so training is bitwise reproducible run to run.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sfem_device.hpp"
#include "sfem_gnn.hpp"

namespace sfem {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ short f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ __bf16 bits_bf16(short h) { return __builtin_bit_cast(__bf16, h); }

// one step of the reduction: D += A B over 32 (KS = 32) or 16 (KS = 16) values
// of k; a8 / b8 hold the lane's 8 (or first 4) items
template <int KS>
__device__ __forceinline__ floatx4 mfma_step(const short (&a8)[8], const short (&b8)[8], floatx4 acc)
{
    if constexpr (KS == 32) {
        bf16x8 a, b;
#pragma unroll
        for (int q = 0; q < 8; q++) { a[q] = bits_bf16(a8[q]); b[q] = bits_bf16(b8[q]); }
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    } else {
        bf16x4 a, b;
#pragma unroll
        for (int q = 0; q < 4; q++) { a[q] = a8[q]; b[q] = b8[q]; }
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
}

// ------------------------------------------------------------------- GEMMs
// Y[N,NO] = [X1 | X2] W + b (+ R when non-null), X1: N x K1, X2: N x (K - K1),
// W: K x NO bf16 row-major; Y and Yact = relu(Y) each written when non-null.
// K1 is a multiple of 4 (each lane loads 4 consecutive inputs of one row as
// one float4, from X1 or from X2).  One wave = 16 rows x NO columns.
template <int K, int NO>
__global__ __launch_bounds__(256) void k_gemm(int64_t N, const float* __restrict__ X1, int K1, const float* __restrict__ X2,
                                              const uint16_t* __restrict__ W, const float* __restrict__ bias,
                                              const float* __restrict__ R, float* __restrict__ Y, float* __restrict__ Yact)
{
    constexpr int KS = K % 32 == 0 ? 32 : 16, PQ = KS / 4;   // k values per step, per lane
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
    if (row0 >= N) return;
    const int i = lane & 15, kq = lane >> 4;
    floatx4 acc[NO / 16];
#pragma unroll
    for (int t = 0; t < NO / 16; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int64_t ar = row0 + i;
    const int K2 = K - K1;
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += KS) {
        short a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int h = 0; h < PQ / 4; h++) {
            const int k = k0 + PQ * kq + 4 * h;
            float4 xv = make_float4(0.f, 0.f, 0.f, 0.f);
            if (ar < N)
                xv = k < K1 ? *reinterpret_cast<const float4*>(X1 + ar * K1 + k)
                            : *reinterpret_cast<const float4*>(X2 + ar * K2 + (k - K1));
            a[4 * h] = f2bf(xv.x);
            a[4 * h + 1] = f2bf(xv.y);
            a[4 * h + 2] = f2bf(xv.z);
            a[4 * h + 3] = f2bf(xv.w);
        }
#pragma unroll
        for (int t = 0; t < NO / 16; t++) {
            short b[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < PQ; q++) b[q] = (short)W[(k0 + PQ * kq + q) * NO + t * 16 + i];
            acc[t] = mfma_step<KS>(a, b, acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < NO / 16; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int64_t r = row0 + 4 * kq + q;
            const int j = t * 16 + i;
            if (r < N) {
                float v = acc[t][q] + (bias ? bias[j] : 0.f);
                if (R) v = R[r * NO + j] + v;
                if (Y) Y[r * NO + j] = v;
                if (Yact) Yact[r * NO + j] = v > 0.f ? v : 0.f;
            }
        }
}

// partial[b][K][NO] = sum over block b's rows of [X1|X2]^T dY  (weight gradient)
template <int K, int NO>
__global__ __launch_bounds__(256) void k_gemm_wgrad(int64_t N, const float* __restrict__ X1, int K1, const float* __restrict__ X2,
                                                    const float* __restrict__ dY, int64_t rows_per_block,
                                                    float* __restrict__ partial)
{
    // the reduction runs over rows: 32 per MFMA step (rows_per_block is a multiple of 32)
    constexpr int TN = NO / 16, NT = (K / 16) * TN, PER = (NT + 3) / 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kq = lane >> 4;
    const int K2 = K - K1;
    const int64_t rbeg = (int64_t)blockIdx.x * rows_per_block;
    const int64_t rend = rbeg + rows_per_block < N ? rbeg + rows_per_block : N;
    floatx4 acc[PER];
#pragma unroll
    for (int tt = 0; tt < PER; tt++) acc[tt] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int64_t r0 = rbeg; r0 < rend; r0 += 32) {
#pragma unroll
        for (int tt = 0; tt < PER; tt++) {
            const int t = wave + 4 * tt;
            if (t < NT) {
                const int kt = t / TN, nt = t % TN;
                short a[8], b[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int64_t r = r0 + 8 * kq + q;
                    const int k = kt * 16 + i;
                    float xv = 0.f, yv = 0.f;
                    if (r < rend) {
                        xv = k < K1 ? X1[r * K1 + k] : X2[r * K2 + (k - K1)];
                        yv = dY[r * NO + nt * 16 + i];
                    }
                    a[q] = f2bf(xv);
                    b[q] = f2bf(yv);
                }
                acc[tt] = mfma_step<32>(a, b, acc[tt]);
            }
        }
    }
    float* out = partial + (size_t)blockIdx.x * K * NO;
#pragma unroll
    for (int tt = 0; tt < PER; tt++) {
        const int t = wave + 4 * tt;
        if (t < NT) {
            const int kt = t / TN, nt = t % TN;
#pragma unroll
            for (int q = 0; q < 4; q++) out[(kt * 16 + 4 * kq + q) * NO + nt * 16 + i] = acc[tt][q];
        }
    }
}

// partial[b][NO] = column sums of dY over block b's rows (bias gradient)
__global__ __launch_bounds__(256) void k_colsum(int64_t N, int NO, const float* __restrict__ dY, int64_t rows_per_block,
                                                float* __restrict__ partial)
{
    __shared__ float s[256];
    const int tid = threadIdx.x;
    const int j = tid % NO, slice = tid / NO, nslice = 256 / NO;
    const int64_t rbeg = (int64_t)blockIdx.x * rows_per_block;
    const int64_t rend = rbeg + rows_per_block < N ? rbeg + rows_per_block : N;
    float acc = 0.f;
    if (slice < nslice)
        for (int64_t r = rbeg + slice; r < rend; r += nslice) acc += dY[r * NO + j];
    s[tid] = acc;
    __syncthreads();
    if (tid < NO) {
        float t = 0.f;
        for (int k = 0; k < nslice; k++) t += s[k * NO + tid];
        partial[(size_t)blockIdx.x * NO + tid] = t;
    }
}

// out[j] (+)= sum_b partial[b][j], fixed order
__global__ __launch_bounds__(256) void k_reduce_partials(int nb, int n, const float* __restrict__ partial,
                                                         float* __restrict__ out, int accumulate)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float acc = 0.f;
    for (int b = 0; b < nb; b++) acc += partial[(size_t)b * n + j];
    out[j] = accumulate ? out[j] + acc : acc;
}

// ------------------------------------------------------------- elementwise
__global__ __launch_bounds__(256) void k_relu_bwd(int64_t n, const float* __restrict__ pre, const float* __restrict__ dpost,
                                                  float* __restrict__ dpre)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dpre[i] = pre[i] > 0.f ? dpost[i] : 0.f;
}

// dU (N x 64) -> dh (+)= dU[:, :32], dm = dU[:, 32:]
__global__ __launch_bounds__(256) void k_split_concat(int64_t N, const float* __restrict__ dU, float* __restrict__ dh,
                                                      float* __restrict__ dm)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= N * 32) return;
    const int64_t r = i / 32;
    const int c = (int)(i % 32);
    dh[i] += dU[r * 64 + c];
    dm[i] = dU[r * 64 + 32 + c];
}

// m_i = mean over neighbours of h_j (32 channels); one thread per (node, channel)
__global__ __launch_bounds__(256) void k_mean_fwd(int64_t N, const int* __restrict__ ptr, const int* __restrict__ idx,
                                                  const float* __restrict__ h, float* __restrict__ m)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= N * 32) return;
    const int64_t i = t / 32;
    const int c = (int)(t % 32);
    const int b = ptr[i], e = ptr[i + 1];
    float acc = 0.f;
    // eight neighbours' ids, then their eight feature loads, in flight
    // together; the sum keeps the neighbour order
    for (int k0 = b; k0 < e; k0 += 8) {
        int id[8];
#pragma unroll
        for (int q = 0; q < 8; q++) id[q] = k0 + q < e ? idx[k0 + q] : -1;
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = id[q] >= 0 ? h[(int64_t)id[q] * 32 + c] : 0.f;
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (k0 + q < e) acc += v[q];
    }
    m[t] = e > b ? acc / (float)(e - b) : 0.f;
}

// dh_j += sum_{i in N(j)} dm_i / deg_i  (the graph is symmetric)
__global__ __launch_bounds__(256) void k_mean_bwd(int64_t N, const int* __restrict__ ptr, const int* __restrict__ idx,
                                                  const float* __restrict__ dm, float* __restrict__ dh)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= N * 32) return;
    const int64_t j = t / 32;
    const int c = (int)(t % 32);
    float acc = 0.f;
    for (int k = ptr[j]; k < ptr[j + 1]; k++) {
        const int i = idx[k];
        acc += dm[(int64_t)i * 32 + c] / (float)(ptr[i + 1] - ptr[i]);
    }
    dh[t] += acc;
}

// --------------------------------------------------------------- features
// Inputs per node (16): the diagonal block D = sym(K_ii) scaled by t = tr(D)/3
// (6 values), log t minus its neighbour mean, degree / 16, fixed flag,
// fraction of fixed neighbours, 6 zeros.  cbase = lower Cholesky factor of
// D^-1 (6 values), 0 on fixed nodes.
__global__ __launch_bounds__(256) void k_features(int64_t N, const float* __restrict__ B, const uint8_t* __restrict__ fixed,
                                                  const int* __restrict__ ptr, const int* __restrict__ idx,
                                                  float* __restrict__ feat, float* __restrict__ cbase)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float* b = B + 9 * i;
    double D[9];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) D[3 * r + c] = 0.5 * ((double)b[3 * r + c] + (double)b[3 * c + r]);
    double t = (D[0] + D[4] + D[8]) / 3.0;
    if (!(t > 0)) t = 1.0;
    float* f = feat + 16 * i;
    f[0] = (float)(D[0] / t); f[1] = (float)(D[4] / t); f[2] = (float)(D[8] / t);
    f[3] = (float)(D[1] / t); f[4] = (float)(D[2] / t); f[5] = (float)(D[5] / t);
    double lt = log(t), ln = 0.0;
    int nf = 0;
    const int e0 = ptr[i], e1 = ptr[i + 1];
    for (int k = e0; k < e1; k++) {
        const float* bj = B + 9 * (int64_t)idx[k];
        double tj = ((double)bj[0] + bj[4] + bj[8]) / 3.0;
        ln += log(tj > 0 ? tj : 1.0);
        nf += fixed[idx[k]] ? 1 : 0;
    }
    const int deg = e1 - e0;
    f[6] = deg ? (float)(lt - ln / deg) : 0.f;
    f[7] = deg / 16.f;
    f[8] = fixed[i] ? 1.f : 0.f;
    f[9] = deg ? (float)nf / deg : 0.f;
    for (int k = 10; k < 16; k++) f[k] = 0.f;
    // Cholesky of D^-1
    double Di[9];
    inv3(D, Di);
    double c[6] = {0, 0, 0, 0, 0, 0};
    bool ok = !fixed[i];
    if (ok) {
        const double l00 = Di[0] > 0 ? sqrt(Di[0]) : 0.0;
        ok = l00 > 0;
        if (ok) {
            const double l10 = Di[3] / l00, l20 = Di[6] / l00;
            const double s11 = Di[4] - l10 * l10;
            ok = s11 > 0;
            if (ok) {
                const double l11 = sqrt(s11);
                const double l21 = (Di[7] - l20 * l10) / l11;
                const double s22 = Di[8] - l20 * l20 - l21 * l21;
                ok = s22 > 0;
                if (ok) { c[0] = l00; c[1] = l11; c[2] = sqrt(s22); c[3] = l10; c[4] = l20; c[5] = l21; }
            }
        }
        if (!ok) {   // not SPD (possible for the corotational tangent): diagonal fallback
            c[0] = 1.0 / sqrt(fabs(D[0]) + 1e-30); c[1] = 1.0 / sqrt(fabs(D[4]) + 1e-30); c[2] = 1.0 / sqrt(fabs(D[8]) + 1e-30);
        }
    }
    float* cb = cbase + 6 * i;
    for (int k = 0; k < 6; k++) cb[k] = (float)c[k];
}

// lower-triangular 3x3 from 6 values (diag 0..2, then (1,0), (2,0), (2,1))
__device__ __forceinline__ void lower3(const float* v, float* L)
{
    L[0] = v[0]; L[1] = 0.f; L[2] = 0.f;
    L[3] = v[3]; L[4] = v[1]; L[5] = 0.f;
    L[6] = v[4]; L[7] = v[5]; L[8] = v[2];
}

// Minv = C (L L^T + eps I) C^T, L = lower(exp(o0..2), o3..5)
__global__ __launch_bounds__(256) void k_blocks_fwd(int64_t N, const float* __restrict__ O, const float* __restrict__ cbase,
                                                    float eps, float* __restrict__ minv)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float* o = O + 16 * i;
    float lv[6] = {__expf(fminf(fmaxf(o[0], -8.f), 8.f)), __expf(fminf(fmaxf(o[1], -8.f), 8.f)),
                   __expf(fminf(fmaxf(o[2], -8.f), 8.f)), o[3], o[4], o[5]};
    float L[9], C[9], S[9], T[9], M[9];
    lower3(lv, L);
    lower3(cbase + 6 * i, C);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) S[3 * r + c] = L[3 * r] * L[3 * c] + L[3 * r + 1] * L[3 * c + 1] + L[3 * r + 2] * L[3 * c + 2] + (r == c ? eps : 0.f);
    mul3(C, S, T);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) M[3 * r + c] = T[3 * r] * C[3 * c] + T[3 * r + 1] * C[3 * c + 1] + T[3 * r + 2] * C[3 * c + 2];
    for (int k = 0; k < 9; k++) minv[9 * i + k] = M[k];
}

// dO from dMinv: dS = C^T dM C, dL = (dS + dS^T) L, chain through exp
__global__ __launch_bounds__(256) void k_blocks_bwd(int64_t N, const float* __restrict__ O, const float* __restrict__ cbase,
                                                    const float* __restrict__ dM, float* __restrict__ dO)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float* o = O + 16 * i;
    float ex[3];
    bool clamp[3];
    for (int k = 0; k < 3; k++) {
        clamp[k] = o[k] < -8.f || o[k] > 8.f;
        ex[k] = __expf(fminf(fmaxf(o[k], -8.f), 8.f));
    }
    float lv[6] = {ex[0], ex[1], ex[2], o[3], o[4], o[5]};
    float L[9], C[9], G[9], T[9], dS[9], dL[9];
    lower3(lv, L);
    lower3(cbase + 6 * i, C);
    const float* g = dM + 9 * i;
    for (int r = 0; r < 3; r++)   // T = C^T dM
        for (int c = 0; c < 3; c++) T[3 * r + c] = C[r] * g[c] + C[3 + r] * g[3 + c] + C[6 + r] * g[6 + c];
    mul3(T, C, dS);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) G[3 * r + c] = dS[3 * r + c] + dS[3 * c + r];
    mul3(G, L, dL);
    float* d = dO + 16 * i;
    d[0] = clamp[0] ? 0.f : dL[0] * ex[0];
    d[1] = clamp[1] ? 0.f : dL[4] * ex[1];
    d[2] = clamp[2] ? 0.f : dL[8] * ex[2];
    d[3] = dL[3];
    d[4] = dL[6];
    d[5] = dL[7];
    for (int k = 6; k < 16; k++) d[k] = 0.f;
}

// z_s = Minv r_s for a batch of S vectors (S x nn x 3)
__global__ __launch_bounds__(256) void k_apply_blocks(int64_t N, int S, const float* __restrict__ minv,
                                                      const float* __restrict__ r, float* __restrict__ z)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float* M = minv + 9 * i;
    for (int s = 0; s < S; s++) {
        const float* rv = r + ((size_t)s * N + i) * 3;
        float* zv = z + ((size_t)s * N + i) * 3;
        for (int c = 0; c < 3; c++) zv[c] = M[3 * c] * rv[0] + M[3 * c + 1] * rv[1] + M[3 * c + 2] * rv[2];
    }
}

// dMinv_i = sum_s (c1_s q_s,i - c2_s r_s,i) r_s,i^T
__global__ __launch_bounds__(256) void k_loss_grad_blocks(int64_t N, int S, const float* __restrict__ r, const float* __restrict__ q,
                                                          const float* __restrict__ c1, const float* __restrict__ c2,
                                                          float* __restrict__ dM)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= N) return;
    float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = 0; s < S; s++) {
        const float* rv = r + ((size_t)s * N + i) * 3;
        const float* qv = q + ((size_t)s * N + i) * 3;
        float gv[3];
        for (int c = 0; c < 3; c++) gv[c] = c1[s] * qv[c] - c2[s] * rv[c];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) acc[3 * a + b] += gv[a] * rv[b];
    }
    for (int k = 0; k < 9; k++) dM[9 * i + k] = acc[k];
}

// per-workgroup partials of (z_s . r_s, z_s . q_s) for s < S, FP64
__global__ __launch_bounds__(256) void k_pair_dots(int64_t n3, int S, const float* __restrict__ z, const float* __restrict__ r,
                                                   const float* __restrict__ q, double* __restrict__ partial)
{
    __shared__ double lds[256];
    for (int s = 0; s < S; s++) {
        double a = 0.0, b = 0.0;
        for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n3; k += (int64_t)gridDim.x * blockDim.x) {
            const double zv = z[(size_t)s * n3 + k];
            a += zv * r[(size_t)s * n3 + k];
            b += zv * q[(size_t)s * n3 + k];
        }
        for (int pass = 0; pass < 2; pass++) {
            lds[threadIdx.x] = pass ? b : a;
            __syncthreads();
            for (int w = 128; w > 0; w >>= 1) {
                if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
                __syncthreads();
            }
            if (threadIdx.x == 0) partial[((size_t)blockIdx.x * S + s) * 2 + pass] = lds[0];
            __syncthreads();
        }
    }
}

// sum_k (e_k - x_k) r_k partials, FP64
__global__ __launch_bounds__(256) void k_err_dot(int64_t n3, const float* __restrict__ e, const float* __restrict__ x,
                                                 const float* __restrict__ r, double* __restrict__ partial)
{
    __shared__ double lds[256];
    double a = 0.0;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n3; k += (int64_t)gridDim.x * blockDim.x)
        a += ((double)e[k] - (double)x[k]) * (double)r[k];
    lds[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

// Adam (Kingma & Ba), bias correction folded into lr_t by the host
__global__ __launch_bounds__(256) void k_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, float lr_t, float b1, float b2, float eps, float clip)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = fminf(fmaxf(g[i], -clip), clip);
    m[i] = b1 * m[i] + (1.f - b1) * gi;
    v[i] = b2 * v[i] + (1.f - b2) * gi * gi;
    w[i] -= lr_t * m[i] / (sqrtf(v[i]) + eps);
}

// bf16 copies of a K x NO weight: W (row-major) and W^T
__global__ __launch_bounds__(256) void k_to_bf16_pair(int K, int NO, const float* __restrict__ w, uint16_t* __restrict__ wb,
                                                      uint16_t* __restrict__ wtb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K * NO) return;
    const int k = i / NO, j = i % NO;
    const uint16_t b = (uint16_t)f2bf(w[i]);
    wb[i] = b;
    wtb[j * K + k] = b;
}

// ---------------------------------------------------------------- launchers
static inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

hipError_t gnn_gemm(int K, int NO, int64_t N, const float* X1, int K1, const float* X2, const uint16_t* W, const float* bias,
                    float* Y, float* Yact, hipStream_t s, const float* R)
{
    if (K1 % 4 != 0 || (K1 < K && !X2)) return hipErrorInvalidValue;
    const dim3 g((unsigned)((N + 63) / 64)), b(256);
#define SFEM_GEMM_CASE(KK, NN)                                                                            \
    if (K == KK && NO == NN) {                                                                             \
        hipLaunchKernelGGL((k_gemm<KK, NN>), g, b, 0, s, N, X1, K1, X2, W, bias, R, Y, Yact);              \
        return hipGetLastError();                                                                          \
    }
    SFEM_GEMM_CASE(16, 32)
    SFEM_GEMM_CASE(32, 32)
    SFEM_GEMM_CASE(64, 32)
    SFEM_GEMM_CASE(32, 16)
    SFEM_GEMM_CASE(32, 64)
#undef SFEM_GEMM_CASE
    return hipErrorInvalidValue;
}

hipError_t gnn_wgrad(int K, int NO, int64_t N, const float* X1, int K1, const float* X2, const float* dY, float* partial,
                     int nparts, float* out, hipStream_t s)
{
    const int64_t rpb = ((N + nparts - 1) / nparts + 31) / 32 * 32;
    const dim3 g((unsigned)nparts), b(256);
#define SFEM_WG_CASE(KK, NN)                                                                              \
    if (K == KK && NO == NN) {                                                                             \
        hipLaunchKernelGGL((k_gemm_wgrad<KK, NN>), g, b, 0, s, N, X1, K1, X2, dY, rpb, partial);           \
        goto reduce;                                                                                       \
    }
    SFEM_WG_CASE(16, 32)
    SFEM_WG_CASE(32, 32)
    SFEM_WG_CASE(64, 32)
    SFEM_WG_CASE(32, 16)
#undef SFEM_WG_CASE
    return hipErrorInvalidValue;
reduce:
    hipLaunchKernelGGL(k_reduce_partials, dim3(nblk(K * NO)), dim3(256), 0, s, nparts, K * NO, partial, out, 0);
    return hipGetLastError();
}

hipError_t gnn_bgrad(int NO, int64_t N, const float* dY, float* partial, int nparts, float* out, hipStream_t s)
{
    const int64_t rpb = (N + nparts - 1) / nparts;
    hipLaunchKernelGGL(k_colsum, dim3((unsigned)nparts), dim3(256), 0, s, N, NO, dY, rpb, partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(nblk(NO)), dim3(256), 0, s, nparts, NO, partial, out, 0);
    return hipGetLastError();
}

hipError_t gnn_relu_bwd(int64_t n, const float* pre, const float* dpost, float* dpre, hipStream_t s)
{
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(n)), dim3(256), 0, s, n, pre, dpost, dpre);
    return hipGetLastError();
}

hipError_t gnn_split_concat(int64_t N, const float* dU, float* dh, float* dm, hipStream_t s)
{
    hipLaunchKernelGGL(k_split_concat, dim3(nblk(N * 32)), dim3(256), 0, s, N, dU, dh, dm);
    return hipGetLastError();
}

hipError_t gnn_mean_fwd(int64_t N, const int* ptr, const int* idx, const float* h, float* m, hipStream_t s)
{
    hipLaunchKernelGGL(k_mean_fwd, dim3(nblk(N * 32)), dim3(256), 0, s, N, ptr, idx, h, m);
    return hipGetLastError();
}

hipError_t gnn_mean_bwd(int64_t N, const int* ptr, const int* idx, const float* dm, float* dh, hipStream_t s)
{
    hipLaunchKernelGGL(k_mean_bwd, dim3(nblk(N * 32)), dim3(256), 0, s, N, ptr, idx, dm, dh);
    return hipGetLastError();
}

hipError_t gnn_features(int64_t N, const float* B, const uint8_t* fixed, const int* ptr, const int* idx, float* feat,
                        float* cbase, hipStream_t s)
{
    hipLaunchKernelGGL(k_features, dim3(nblk(N)), dim3(256), 0, s, N, B, fixed, ptr, idx, feat, cbase);
    return hipGetLastError();
}

hipError_t gnn_blocks_fwd(int64_t N, const float* O, const float* cbase, float eps, float* minv, hipStream_t s)
{
    hipLaunchKernelGGL(k_blocks_fwd, dim3(nblk(N)), dim3(256), 0, s, N, O, cbase, eps, minv);
    return hipGetLastError();
}

hipError_t gnn_blocks_bwd(int64_t N, const float* O, const float* cbase, const float* dM, float* dO, hipStream_t s)
{
    hipLaunchKernelGGL(k_blocks_bwd, dim3(nblk(N)), dim3(256), 0, s, N, O, cbase, dM, dO);
    return hipGetLastError();
}

hipError_t gnn_apply_blocks(int64_t N, int S, const float* minv, const float* r, float* z, hipStream_t s)
{
    hipLaunchKernelGGL(k_apply_blocks, dim3(nblk(N)), dim3(256), 0, s, N, S, minv, r, z);
    return hipGetLastError();
}

hipError_t gnn_loss_grad_blocks(int64_t N, int S, const float* r, const float* q, const float* c1, const float* c2, float* dM,
                                hipStream_t s)
{
    hipLaunchKernelGGL(k_loss_grad_blocks, dim3(nblk(N)), dim3(256), 0, s, N, S, r, q, c1, c2, dM);
    return hipGetLastError();
}

hipError_t gnn_pair_dots(int64_t n3, int S, const float* z, const float* r, const float* q, double* partial, int nparts,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_pair_dots, dim3((unsigned)nparts), dim3(256), 0, s, n3, S, z, r, q, partial);
    return hipGetLastError();
}

hipError_t gnn_err_dot(int64_t n3, const float* e, const float* x, const float* r, double* partial, int nparts, hipStream_t s)
{
    hipLaunchKernelGGL(k_err_dot, dim3((unsigned)nparts), dim3(256), 0, s, n3, e, x, r, partial);
    return hipGetLastError();
}

hipError_t gnn_adam(int n, float* w, const float* g, float* m, float* v, float lr_t, float b1, float b2, float eps, float clip,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, s, n, w, g, m, v, lr_t, b1, b2, eps, clip);
    return hipGetLastError();
}

hipError_t gnn_to_bf16_pair(int K, int NO, const float* w, uint16_t* wb, uint16_t* wtb, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf16_pair, dim3(nblk(K * NO)), dim3(256), 0, s, K, NO, w, wb, wtb);
    return hipGetLastError();
}

}  // namespace sfem
