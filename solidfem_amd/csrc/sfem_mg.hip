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

#include "sfem_device.hpp"
#include "sfem_kernels.hpp"
#include "sfem_mg.hpp"
#include "sfem_mgdev.hpp"
#include "sfem_reduce.hpp"

namespace sfem {

constexpr int MG_BLOCK = 256;

static int mg_grid(int64_t n) { return (int)((n + MG_BLOCK - 1) / MG_BLOCK); }

// One Chebyshev step (q = the level operator applied to the previous d, or to
// the coarse correction e for MODE 2):
//   MODE 0: r = b;      d = c2 M r;         x = d        first step from x = 0
//   MODE 1: r -= q;     d = c1 d + c2 M r;  x += d       next step
//   MODE 2: r -= q;     d = c2 M r;         x += d       first step from x (after x += e)
//   MODE 3: r -= q                                       residual after the last step
//   MODE 4: r = b - q;  d = c1 d + c2 M r;  x += d       second step, r not yet copied from b
// done: the PCG's flag (a flag that stays 0 outside PCG), read with the first
// loads; once the solve has stopped, the V-cycles enqueued after it return at
// once (every multigrid kernel tests it before its first store)
template <int MODE>
__global__ __launch_bounds__(MG_BLOCK) void k_mg_cheb(int64_t nn, const float* __restrict__ b, float* __restrict__ r,
                                                      float* __restrict__ d, float* __restrict__ x,
                                                      const float* __restrict__ q, const int2* __restrict__ node_q,
                                                      const float* __restrict__ partial, const float* __restrict__ minv,
                                                      float c1, float c2, const int* __restrict__ done)
{
    const int stop = vload(done);
    const int64_t i = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (i >= nn) return;
    float rv[3];
    if (MODE == 0) {
#pragma unroll
        for (int c = 0; c < 3; c++) rv[c] = b[3 * i + c];
    } else {
        float qv[3];
        operator_row(i, q, node_q, partial, qv);
        const float* rin = MODE == 4 ? b : r;
#pragma unroll
        for (int c = 0; c < 3; c++) rv[c] = rin[3 * i + c] - qv[c];
    }
    if (stop) return;
#pragma unroll
    for (int c = 0; c < 3; c++) r[3 * i + c] = rv[c];
    if (MODE == 3) return;
    float mr[3];
    block_apply(minv, i, rv, mr);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        constexpr bool next = MODE == 1 || MODE == 4;
        const float dv = cheb_dir(next, c1, next ? d[3 * i + c] : 0.f, c2, mr[c]);
        d[3 * i + c] = dv;
        x[3 * i + c] = MODE == 0 ? dv : x[3 * i + c] + dv;
    }
}

// The finest level's last Chebyshev step (k_mg_cheb<1>) and the PCG's r.z and
// r.r partials of its result z = x against b (the PCG residual), in the CG
// update's layout: CG_BLOCK threads, cg_grid(nn) workgroups, grid-stride
// (k_mg_rz's order)
__global__ __launch_bounds__(CG_BLOCK) void k_mg_cheb_rz(int64_t nn, const float* __restrict__ b, float* __restrict__ r,
                                                         float* __restrict__ d, float* __restrict__ x,
                                                         const float* __restrict__ q, const int2* __restrict__ node_q,
                                                         const float* __restrict__ partial, const float* __restrict__ minv,
                                                         float c1, float c2, double* __restrict__ part,
                                                         const int* __restrict__ done)
{
    const int stop = vload(done);
    double v[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)CG_BLOCK + threadIdx.x; i < nn; i += (int64_t)gridDim.x * CG_BLOCK) {
        float qv[3], rv[3], mr[3];
        operator_row(i, q, node_q, partial, qv);
        if (stop) break;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            rv[c] = r[3 * i + c] - qv[c];
            r[3 * i + c] = rv[c];
        }
        block_apply(minv, i, rv, mr);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float dv = cheb_dir(true, c1, d[3 * i + c], c2, mr[c]);
            d[3 * i + c] = dv;
            const float xv = x[3 * i + c] + dv;
            x[3 * i + c] = xv;
            const double rb = b[3 * i + c];
            v[0] = fma(rb, (double)xv, v[0]);
            v[1] = fma(rb, rb, v[1]);
        }
    }
    if (stop) return;
    block_partials<2, CG_BLOCK>(v, part);
}

// e = P xc on the level above (0 at its fixed nodes); x += e
__global__ __launch_bounds__(MG_BLOCK) void k_mg_prolong(int64_t nn, const int* __restrict__ ptr, const int* __restrict__ col,
                                                         const float* __restrict__ val, const float* __restrict__ xc,
                                                         const uint8_t* __restrict__ fixed, float* __restrict__ e,
                                                         float* __restrict__ x, const int* __restrict__ done)
{
    const int stop = vload(done);
    const int64_t i = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (i >= nn) return;
    float ev[3] = {0.f, 0.f, 0.f};
    gather_rows(ptr[i], ptr[i + 1], col, val, xc, ev);
    if (stop) return;
    const bool f = fixed[i] != 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float v = f ? 0.f : ev[c];
        e[3 * i + c] = v;
        x[3 * i + c] += v;
    }
}

// bc = P^T r: rows of the transpose (coarse nodes) in increasing fine-node
// order; 0 at fixed coarse nodes.  Four lanes per coarse node: each sums a
// contiguous quarter of the row, then (s0 + s1) + (s2 + s3).
// With minv non-null, the coarse level's first Chebyshev step from x = 0
// follows on the new right-hand side (k_mg_cheb<0>: rc = bc, dc = c2 M bc,
// xc = dc), which saves its launch.
__global__ __launch_bounds__(MG_BLOCK) void k_mg_restrict(int64_t nnc, const int* __restrict__ ptr, const int* __restrict__ row,
                                                          const float* __restrict__ val, const float* __restrict__ r,
                                                          const uint8_t* __restrict__ fixed, float* __restrict__ bc,
                                                          const float* __restrict__ minv, float* __restrict__ rc,
                                                          float* __restrict__ dc, float* __restrict__ xc, float c2,
                                                          const int* __restrict__ done)
{
    const int stop = vload(done);
    const int64_t t = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    const int64_t k = t >> 2;
    const int lane = (int)(t & 3);
    float s[3] = {0.f, 0.f, 0.f};
    if (k < nnc) {
        const int p0 = ptr[k], p1 = ptr[k + 1];
        const int cs = (p1 - p0 + 3) >> 2;
        const int a = p0 + lane * cs, e = a + cs < p1 ? a + cs : p1;
        gather_rows(a, e, row, val, r, s);
    }
    if (stop) return;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        s[c] += __shfl_xor(s[c], 1, 64);
        s[c] += __shfl_xor(s[c], 2, 64);
    }
    if (k < nnc && lane == 0) {
        const bool f = fixed[k] != 0;
        float rv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            rv[c] = f ? 0.f : s[c];
            bc[3 * k + c] = rv[c];
        }
        if (minv) {
            float mr[3];
            block_apply(minv, k, rv, mr);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float dv = cheb_dir(false, 0.f, 0.f, c2, mr[c]);
                rc[3 * k + c] = rv[c];
                dc[3 * k + c] = dv;
                xc[3 * k + c] = dv;
            }
        }
    }
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
                                                    double* __restrict__ part, const int* __restrict__ done)
{
    if (vload(done)) return;
    double v[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)CG_BLOCK + threadIdx.x; i < nn; i += (int64_t)gridDim.x * CG_BLOCK) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double rv = r[3 * i + c];
            v[0] = fma(rv, (double)z[3 * i + c], v[0]);
            v[1] = fma(rv, rv, v[1]);
        }
    }
    block_partials<2, CG_BLOCK>(v, part);
}

__global__ __launch_bounds__(MG_BLOCK) void k_mg_copy(int64_t n, const float* __restrict__ a, float* __restrict__ b)
{
    const int64_t i = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ __launch_bounds__(MG_BLOCK) void k_mg_scale(int64_t n, const float* __restrict__ a, float sc,
                                                       float* __restrict__ b)
{
    const int64_t i = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (i < n) b[i] = sc * a[i];
}

// ------------------------------------------- direct solve on the coarsest level
// Its dense operator (n = 3 nn dofs, n <= MG_DENSE_MAX) at the current state:
// element matrices, a fixed-order gather into the dense matrix, and an
// in-LDS inverse; each V-cycle then applies the inverse (one dense product).

// Ke[e] (12 x 12, row (3a + c), column (3b + k)) = V dP(e_k (x) g_b) g_a, with
// dP the material tangent at F = I + grad u (the Hessian-vector kernel's
// element operator applied to unit corner displacements)
template <int MAT>
__global__ __launch_bounds__(MG_BLOCK) void k_mg_elem_matrix(int64_t ne, const int4* __restrict__ tets,
                                                             const float* __restrict__ u, const float* __restrict__ dminv,
                                                             const float* __restrict__ cvol, float* __restrict__ Ke)
{
    const int64_t e = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (e >= ne) return;
    const int4 t = tets[e];
    const int id[4] = {t.x, t.y, t.z, t.w};
    float g[9], uu[4][3], F[9];
#pragma unroll
    for (int k = 0; k < 9; k++) g[k] = dminv[9 * e + k];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) uu[a][c] = u[3 * (int64_t)id[a] + c];
    defgrad(g, uu, F);
    Tangent<float, MAT> tg;
    tg.init(F, cvol[2 * e], cvol[2 * e + 1]);
    float ga[4][3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        ga[1][c] = g[c]; ga[2][c] = g[3 + c]; ga[3][c] = g[6 + c];
        ga[0][c] = -(g[c] + g[3 + c] + g[6 + c]);
    }
    float* out = Ke + 144 * e;
    for (int b = 0; b < 4; b++)
        for (int k = 0; k < 3; k++) {
            float dF[9], dP[9];
#pragma unroll
            for (int q = 0; q < 9; q++) dF[q] = 0.f;
            dF[3 * k] = ga[b][0]; dF[3 * k + 1] = ga[b][1]; dF[3 * k + 2] = ga[b][2];
            tg.apply(dF, dP);
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int c = 0; c < 3; c++)
                    out[(3 * a + c) * 12 + 3 * b + k] =
                        dP[3 * c] * ga[a][0] + dP[3 * c + 1] * ga[a][1] + dP[3 * c + 2] * ga[a][2];
        }
}

// A[i][j] (n x n): the sum over the elements of node I = i / 3 (in element
// order) that also hold node J = j / 3; identity rows and columns at fixed nodes
__global__ __launch_bounds__(MG_BLOCK) void k_mg_dense_assemble(int nn, const int* __restrict__ inc_ptr,
                                                                const int2* __restrict__ inc, const int4* __restrict__ tets,
                                                                const float* __restrict__ Ke,
                                                                const uint8_t* __restrict__ fixed, float* __restrict__ A)
{
    const int n = 3 * nn;
    const int64_t t = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (t >= (int64_t)n * n) return;
    const int i = (int)(t / n), j = (int)(t % n);
    const int I = i / 3, c = i % 3, J = j / 3, k = j % 3;
    float s = 0.f;
    if (fixed[I] || fixed[J]) {
        s = i == j ? 1.f : 0.f;
    } else {
        for (int q = inc_ptr[I]; q < inc_ptr[I + 1]; q++) {
            const int2 ea = inc[q];   // (element, corner of I)
            const int4 tt = tets[ea.x];
            const int id[4] = {tt.x, tt.y, tt.z, tt.w};
#pragma unroll
            for (int b = 0; b < 4; b++)
                if (id[b] == J) s += Ke[144 * (int64_t)ea.x + (3 * ea.y + c) * 12 + 3 * b + k];
        }
    }
    A[t] = s;
}

// The coarsest operator's inverse by the sweep operator (Goodnight, The
// American Statistician 33 (1979) 149-158): sweeping pivot k (d = a_kk;
// a_kk <- -1/d, a_ik <- a_ik / d, a_ij <- a_ij - a_ik a_kj / d) for every k
// turns a symmetric positive definite A into -A^-1.  Blocked in panels of
// MG_TB pivots: with B the panel, S = the swept diagonal block (= -A_BB^-1),
// C = the panel's columns A_iB before the sweep and G_i = -C_i S, sweeping B
// sets A_BB <- S, A_iB <- G_i and A_ij <- A_ij - G_i . C_j (i, j outside B).
// k_mg_sym symmetrises W = (A + A^T) / 2 and copies panel 0's columns to C;
// per panel, k_mg_panel (one workgroup) sweeps the diagonal block (read from
// C), k_mg_gmat forms G, k_mg_tiles updates every upper 32 x 32 tile, writes
// it and its mirror (W stays exactly symmetric) and copies the next panel's
// columns as it writes them (C is double-buffered, C[p & 1] for panel p,
// zero past n).
// The sweep runs in FP64 (the coarse operators of nearly incompressible
// material have condition numbers ~1e6, where an FP32 sweep loses percents);
// k_mg_negate_f32 stores -W = A^-1 in FP32 for the V-cycle's product.
constexpr int MG_TB = 32;
typedef double MgT;   // arithmetic of the dense inverse

__global__ __launch_bounds__(MG_BLOCK) void k_mg_sym(int n, const float* __restrict__ A, MgT* __restrict__ W,
                                                     MgT* __restrict__ C0)
{
    const int64_t t = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (t < (int64_t)n * n) {
        const int i = (int)(t / n), j = (int)(t % n);
        const MgT v = MgT(0.5) * ((MgT)A[t] + (MgT)A[(int64_t)j * n + i]);
        W[t] = v;
        if (j < MG_TB) C0[i * MG_TB + j] = v;
    }
    if (t < (int64_t)n * MG_TB && (int)(t % MG_TB) >= n) C0[t] = MgT(0);   // n < MG_TB: zero past n
}

__global__ __launch_bounds__(MG_BLOCK) void k_mg_negate_f32(int64_t nn2, const MgT* __restrict__ W, float* __restrict__ Ainv)
{
    const int64_t t = blockIdx.x * (int64_t)MG_BLOCK + threadIdx.x;
    if (t < nn2) Ainv[t] = (float)(-W[t]);
}

constexpr int MG_PANEL_BLOCK = 1024;

__global__ __launch_bounds__(MG_PANEL_BLOCK) void k_mg_panel(int k0, int bp, const MgT* __restrict__ Cp,
                                                             MgT* __restrict__ S)
{
    __shared__ MgT col[2][MG_TB];
    const int tid = threadIdx.x;
    // thread -> entry (a, b), a <= b, of the block's upper triangle
    constexpr int NU = MG_TB * (MG_TB + 1) / 2;
    static_assert(NU <= MG_PANEL_BLOCK, "one entry per thread");
    int a = 0, b = -1;
    if (tid < NU) {
        b = (int)((sqrtf(8.f * (float)tid + 1.f) - 1.f) * 0.5f);
        while (b * (b + 1) / 2 > tid) b--;
        while ((b + 1) * (b + 2) / 2 <= tid) b++;
        a = tid - b * (b + 1) / 2;
    }
    const bool own = b >= 0 && b < bp;
    MgT v = 0;
    if (own) {
        v = Cp[(k0 + a) * MG_TB + b];   // C row k0 + a holds W row k0 + a at the panel's columns
        if (a == 0) col[0][b] = v;       // column 0 = row 0 of the upper triangle
    }
    __syncthreads();
    for (int k = 0; k < bp; k++) {
        const MgT* c = col[k & 1];
        MgT* cn = col[(k + 1) & 1];
        const MgT rd = MgT(1) / c[k];
        if (own) {
            if (b == k) v = a == k ? -rd : c[a] * rd;
            else if (a == k) v = c[b] * rd;
            else v = fma(-(c[a] * rd), c[b], v);
            // publish column k + 1: entries (a, k+1) and (k+1, b)
            if (b == k + 1) cn[a] = v;
            else if (a == k + 1) cn[b] = v;
        }
        __syncthreads();
    }
    if (b >= 0) {
        S[a * MG_TB + b] = v;   // zero outside the panel
        S[b * MG_TB + a] = v;
    }
}

// G = -C S for 32 rows of C per workgroup (rows past n are not written)
__global__ __launch_bounds__(MG_BLOCK) void k_mg_gmat(int n, const MgT* __restrict__ Cp, const MgT* __restrict__ S,
                                                      MgT* __restrict__ G)
{
    __shared__ MgT sS[MG_TB][MG_TB + 1], sC[MG_TB][MG_TB + 1];
    const int r0 = blockIdx.x * MG_TB, tid = threadIdx.x;
    for (int e = tid; e < MG_TB * MG_TB; e += MG_BLOCK) {
        const int r = e / MG_TB, k = e % MG_TB;
        sS[r][k] = S[e];
        sC[r][k] = r0 + r < n ? Cp[(r0 + r) * MG_TB + k] : MgT(0);
    }
    __syncthreads();
    for (int e = tid; e < MG_TB * MG_TB; e += MG_BLOCK) {
        const int r = e / MG_TB, k = e % MG_TB;
        MgT g = 0;
        for (int m = 0; m < MG_TB; m++) g = fma(sC[r][m], sS[m][k], g);
        if (r0 + r < n) G[(r0 + r) * MG_TB + k] = -g;
    }
}

__global__ __launch_bounds__(MG_BLOCK) void k_mg_tiles(int n, int p, int bp, MgT* __restrict__ W,
                                                       const MgT* __restrict__ Cp, MgT* __restrict__ Cn,
                                                       const MgT* __restrict__ S, const MgT* __restrict__ G)
{
    __shared__ MgT sA[MG_TB][MG_TB + 1], sB[MG_TB][MG_TB + 1], sT[MG_TB][MG_TB + 1];
    // upper tile (I, J), I <= J, enumerated column by column
    const int t = blockIdx.x;
    int J = (int)((sqrtf(8.f * (float)t + 1.f) - 1.f) * 0.5f);
    while (J * (J + 1) / 2 > t) J--;
    while ((J + 1) * (J + 2) / 2 <= t) J++;
    const int I = t - J * (J + 1) / 2;
    const int r0 = I * MG_TB, c0 = J * MG_TB, kn = (p + 1) * MG_TB;   // kn: the next panel's first column
    const int tid = threadIdx.x;
    // sA: the diagonal block S (panel tile), G of the columns (panel row of
    // tiles) or G of the rows; sB: C of the columns (general tiles)
    const bool diag = I == p && J == p, gcol = I == p;
    for (int e = tid; e < MG_TB * MG_TB; e += MG_BLOCK) {
        const int r = e / MG_TB, k = e % MG_TB;
        const int ra = gcol ? c0 + r : r0 + r;
        sA[r][k] = diag ? S[e] : (ra < n ? G[ra * MG_TB + k] : MgT(0));
        if (!gcol && J != p) sB[r][k] = c0 + r < n ? Cp[(c0 + r) * MG_TB + k] : MgT(0);
    }
    __syncthreads();
    // thread -> row r of the tile, columns cq .. cq + 3
    const int r = tid / 8, cq = (tid % 8) * 4;
    const int i = r0 + r;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int cc = cq + q, j = c0 + cc;
        MgT v = 0;
        if (i < n && j < n) {
            if (diag) v = sA[r][cc];
            else if (J == p) v = sA[r][cc];
            else if (I == p) v = sA[cc][r];
            else {
                MgT s = 0;
                for (int k = 0; k < bp; k++) s = fma(sA[r][k], sB[cc][k], s);
                v = W[(int64_t)i * n + j] - s;
            }
            if (I != J || j >= i) {
                W[(int64_t)i * n + j] = v;
                if (J == p + 1) Cn[i * MG_TB + (j - kn)] = v;
            }
        } else if (i < n && J == p + 1) {
            Cn[i * MG_TB + (j - kn)] = MgT(0);   // the next panel is the last and ends at n
        }
        sT[r][cc] = v;
    }
    __syncthreads();
    // the mirror: W(j, i) = W(i, j); on a diagonal tile its lower half
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int cc = cq + q;   // row r of the mirrored tile is column r of (I, J)
        if (I != J) {
            const int jj = c0 + r, ii = r0 + cc;
            if (jj < n && ii < n) {
                W[(int64_t)jj * n + ii] = sT[cc][r];
                if (I == p + 1) Cn[jj * MG_TB + (ii - kn)] = sT[cc][r];
            }
        } else if (cc < r) {
            const int ii = r0 + r, jj = r0 + cc;
            if (ii < n) {
                W[(int64_t)ii * n + jj] = sT[cc][r];
                if (I == p + 1) Cn[ii * MG_TB + (jj - kn)] = sT[cc][r];
            }
        }
    }
}

// x = Ainv b: one workgroup per row, each thread a strided partial sum, then
// a fixed-order tree
__global__ __launch_bounds__(MG_BLOCK) void k_mg_dense_solve(int n, const float* __restrict__ Ainv,
                                                             const float* __restrict__ b, float* __restrict__ x,
                                                             const int* __restrict__ done)
{
    __shared__ float red[MG_BLOCK / 64];
    const int stop = vload(done);
    const int i = blockIdx.x, tid = threadIdx.x;
    float s = 0.f;
    for (int j = tid; j < n; j += MG_BLOCK) s = fmaf(Ainv[(int64_t)i * n + j], b[j], s);
    if (stop) return;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        float t = 0.f;
#pragma unroll
        for (int w = 0; w < MG_BLOCK / 64; w++) t += red[w];
        x[i] = t;
    }
}

// ------------------------------------------------------------------ launchers

hipError_t launch_mg_cheb(int mode, int64_t nn, const float* b, float* r, float* d, float* x, const float* q,
                          const int2* node_q, const float* partial, const float* minv, float c1, float c2, const int* done,
                          hipStream_t st)
{
    const dim3 g(mg_grid(nn)), bl(MG_BLOCK);
    switch (mode) {
    case 0: hipLaunchKernelGGL(k_mg_cheb<0>, g, bl, 0, st, nn, b, r, d, x, q, node_q, partial, minv, c1, c2, done); break;
    case 1: hipLaunchKernelGGL(k_mg_cheb<1>, g, bl, 0, st, nn, b, r, d, x, q, node_q, partial, minv, c1, c2, done); break;
    case 2: hipLaunchKernelGGL(k_mg_cheb<2>, g, bl, 0, st, nn, b, r, d, x, q, node_q, partial, minv, c1, c2, done); break;
    case 4: hipLaunchKernelGGL(k_mg_cheb<4>, g, bl, 0, st, nn, b, r, d, x, q, node_q, partial, minv, c1, c2, done); break;
    default: hipLaunchKernelGGL(k_mg_cheb<3>, g, bl, 0, st, nn, b, r, d, x, q, node_q, partial, minv, c1, c2, done); break;
    }
    return hipGetLastError();
}

hipError_t launch_mg_prolong(int64_t nn, const int* ptr, const int* col, const float* val, const float* xc,
                             const uint8_t* fixed, float* e, float* x, const int* done, hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_prolong, dim3(mg_grid(nn)), dim3(MG_BLOCK), 0, st, nn, ptr, col, val, xc, fixed, e, x, done);
    return hipGetLastError();
}

hipError_t launch_mg_restrict(int64_t nnc, const int* ptr, const int* row, const float* val, const float* r,
                              const uint8_t* fixed, float* bc, const int* done, hipStream_t st, const float* minv, float* rc,
                              float* dc, float* xc, float c2)
{
    hipLaunchKernelGGL(k_mg_restrict, dim3(mg_grid(4 * nnc)), dim3(MG_BLOCK), 0, st, nnc, ptr, row, val, r, fixed, bc, minv,
                       rc, dc, xc, c2, done);
    return hipGetLastError();
}

hipError_t launch_mg_cheb_rz(int64_t nn, const float* b, float* r, float* d, float* x, const float* q, const int2* node_q,
                             const float* partial, const float* minv, float c1, float c2, double* part, const int* done,
                             hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_cheb_rz, dim3(cg_grid(nn)), dim3(CG_BLOCK), 0, st, nn, b, r, d, x, q, node_q, partial, minv, c1,
                       c2, part, done);
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

hipError_t launch_mg_rz(int64_t nn, const float* r, const float* z, double* part, const int* done, hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_rz, dim3(cg_grid(nn)), dim3(CG_BLOCK), 0, st, nn, r, z, part, done);
    return hipGetLastError();
}

hipError_t launch_mg_elem_matrix(int mat, int64_t ne, const int4* tets, const float* u, const float* dminv,
                                 const float* cvol, float* Ke, hipStream_t st)
{
    const dim3 g(mg_grid(ne)), bl(MG_BLOCK);
    switch (mat) {
    case LINEAR: hipLaunchKernelGGL(k_mg_elem_matrix<LINEAR>, g, bl, 0, st, ne, tets, u, dminv, cvol, Ke); break;
    case STVK: hipLaunchKernelGGL(k_mg_elem_matrix<STVK>, g, bl, 0, st, ne, tets, u, dminv, cvol, Ke); break;
    case COROTATIONAL: hipLaunchKernelGGL(k_mg_elem_matrix<COROTATIONAL>, g, bl, 0, st, ne, tets, u, dminv, cvol, Ke); break;
    case COROTATIONAL_POLAR:
        hipLaunchKernelGGL(k_mg_elem_matrix<COROTATIONAL_POLAR>, g, bl, 0, st, ne, tets, u, dminv, cvol, Ke);
        break;
    case NEO_HOOKEAN_COARSE:
        hipLaunchKernelGGL(k_mg_elem_matrix<NEO_HOOKEAN_COARSE>, g, bl, 0, st, ne, tets, u, dminv, cvol, Ke);
        break;
    default: hipLaunchKernelGGL(k_mg_elem_matrix<NEO_HOOKEAN>, g, bl, 0, st, ne, tets, u, dminv, cvol, Ke); break;
    }
    return hipGetLastError();
}

hipError_t launch_mg_dense_assemble(int nn, const int* inc_ptr, const int2* inc, const int4* tets, const float* Ke,
                                    const uint8_t* fixed, float* A, hipStream_t st)
{
    const int64_t n = 3 * (int64_t)nn;
    hipLaunchKernelGGL(k_mg_dense_assemble, dim3(mg_grid(n * n)), dim3(MG_BLOCK), 0, st, nn, inc_ptr, inc, tets, Ke, fixed, A);
    return hipGetLastError();
}

hipError_t launch_mg_dense_inverse(int n, const float* A, float* Ainv, double* W, double* C, double* S, hipStream_t st)
{
    if (n <= 0 || n > MG_DENSE_MAX) return hipErrorInvalidValue;
    const int64_t nsym = (int64_t)n * n > (int64_t)n * MG_TB ? (int64_t)n * n : (int64_t)n * MG_TB;
    hipLaunchKernelGGL(k_mg_sym, dim3(mg_grid(nsym)), dim3(MG_BLOCK), 0, st, n, A, W, C);
    const int nt = (n + MG_TB - 1) / MG_TB;
    double* G = C + (size_t)2 * n * MG_TB;
    for (int p = 0; p < nt; p++) {
        const int k0 = p * MG_TB, bp = n - k0 < MG_TB ? n - k0 : MG_TB;
        double* Cp = C + (size_t)(p & 1) * n * MG_TB;
        double* Cn = C + (size_t)((p + 1) & 1) * n * MG_TB;
        hipLaunchKernelGGL(k_mg_panel, dim3(1), dim3(MG_PANEL_BLOCK), 0, st, k0, bp, Cp, S);
        hipLaunchKernelGGL(k_mg_gmat, dim3(nt), dim3(MG_BLOCK), 0, st, n, Cp, S, G);
        hipLaunchKernelGGL(k_mg_tiles, dim3(nt * (nt + 1) / 2), dim3(MG_BLOCK), 0, st, n, p, bp, W, Cp, Cn, S, G);
    }
    hipLaunchKernelGGL(k_mg_negate_f32, dim3(mg_grid((int64_t)n * n)), dim3(MG_BLOCK), 0, st, (int64_t)n * n, W, Ainv);
    return hipGetLastError();
}

hipError_t launch_mg_dense_solve(int n, const float* Ainv, const float* b, float* x, const int* done, hipStream_t st)
{
    if (n <= 0 || n > MG_DENSE_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mg_dense_solve, dim3(n), dim3(MG_BLOCK), 0, st, n, Ainv, b, x, done);
    return hipGetLastError();
}

hipError_t launch_mg_copy(int64_t n, const float* a, float* b, hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_copy, dim3(mg_grid(n)), dim3(MG_BLOCK), 0, st, n, a, b);
    return hipGetLastError();
}

hipError_t launch_mg_scale(int64_t n, const float* a, float sc, float* b, hipStream_t st)
{
    hipLaunchKernelGGL(k_mg_scale, dim3(mg_grid(n)), dim3(MG_BLOCK), 0, st, n, a, sc, b);
    return hipGetLastError();
}

}  // namespace sfem
