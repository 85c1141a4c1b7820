// sfem_kernels.hip -- gfx950 kernels of the matrix-free Newton-CG path.
//
// Element work is organised in BLOCKS of BE consecutive elements of a
// space-filling-curve order.  One workgroup owns one block:
//   1. it stages the block's nodes (u, d) from HBM into LDS, one gather per
//      node, not per element corner;
//   2. one thread per element computes the four corner vectors and writes them
//      to LDS (component-major, conflict-free);
//   3. one thread per block-local node sums its incident (element, corner)
//      slots in a fixed order read from a per-block incidence list.
// Nodes touched by this block only are written to the output directly; nodes
// shared with other blocks write a partial row, and a second kernel sums each
// shared node's partial rows in block order.  No atomics touch the results,
// so every output is bitwise reproducible run to run.
//
// Reductions (dot products, norms) are FP64: a wave-shuffle tree per
// workgroup, per-workgroup partials, and the last workgroup to finish (ticket
// counter) sums the partials in index order and derives the CG scalars on
// the device.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sfem_device.hpp"
#include "sfem_kernels.hpp"

namespace sfem {

constexpr uint32_t NODE_MASK = 0x3fffffffu;
constexpr uint32_t FLAG_SHARED = 0x80000000u;
constexpr uint32_t FLAG_FIXED = 0x40000000u;

// --------------------------------------------------------------------- LDS
// slot of (component q, element el) in the corner-output array
__device__ __forceinline__ int sfo(int q, int el, int BE) { return q * BE + el; }

// ------------------------------------------------------- element block kernel
// KIND 0: Hessian-vector (FP32): out = K d            (u, d staged)
// KIND 1: internal force (FP64):  out = f_int (+ ext) (u staged)
// KIND 2: diagonal 3x3 blocks (FP32): out = K_ii      (u staged), W = 9
template <int KIND> struct KindTraits;
template <> struct KindTraits<0> { typedef float T; static constexpr int W = 3; static constexpr bool HAS_D = true; };
template <> struct KindTraits<1> { typedef double T; static constexpr int W = 3; static constexpr bool HAS_D = false; };
template <> struct KindTraits<2> { typedef float T; static constexpr int W = 9; static constexpr bool HAS_D = false; };

template <int KIND, int MAT, int BE>
__global__ __launch_bounds__(BE) void k_elem_block(BlockArgs a, const typename KindTraits<KIND>::T* __restrict__ u,
                                                    const typename KindTraits<KIND>::T* __restrict__ d,
                                                    typename KindTraits<KIND>::T* __restrict__ out,
                                                    typename KindTraits<KIND>::T* __restrict__ partial,
                                                    const double* __restrict__ ext, int mask_fixed,
                                                    const int* __restrict__ done)
{
    typedef typename KindTraits<KIND>::T T;
    constexpr int W = KindTraits<KIND>::W;
    constexpr bool HAS_D = KindTraits<KIND>::HAS_D;
    if (done && *done) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* sf = reinterpret_cast<T*>(smem_raw);          // 4*W*BE corner outputs
    T* su = sf + 4 * W * BE;                          // 3*nmax
    T* sd = su + 3 * a.nmax;                          // 3*nmax (HAS_D)

    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const int n0 = a.bnode_ptr[b];
    const int nb = a.bnode_ptr[b + 1] - n0;

    for (int j = tid; j < nb; j += BE) {
        const uint32_t node = a.bnodes[n0 + j] & NODE_MASK;
        const T* up = u + 3 * (size_t)node;
        su[3 * j] = up[0]; su[3 * j + 1] = up[1]; su[3 * j + 2] = up[2];
        if (HAS_D) {
            const T* dp = d + 3 * (size_t)node;
            sd[3 * j] = dp[0]; sd[3 * j + 1] = dp[1]; sd[3 * j + 2] = dp[2];
        }
    }
    __syncthreads();

    const int64_t e = (int64_t)b * BE + tid;
    if (e < a.ne) {
        const size_t st = (size_t)a.ne_pad;
        T g[9];
        T cmu, clam;
        if (sizeof(T) == 4) {
#pragma unroll
            for (int k = 0; k < 9; k++) g[k] = (T)a.dminv32[k * st + e];
            cmu = (T)a.coef32[e];
            clam = (T)a.coef32[st + e];
        } else {
#pragma unroll
            for (int k = 0; k < 9; k++) g[k] = (T)a.dminv64[k * st + e];
            cmu = (T)a.coef64[e];
            clam = (T)a.coef64[st + e];
        }
        const uint2 li = a.lidx[e];
        const int l[4] = {(int)(li.x & 0xffffu), (int)(li.x >> 16), (int)(li.y & 0xffffu), (int)(li.y >> 16)};
        T uu[4][3];
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int c = 0; c < 3; c++) uu[k][c] = su[3 * l[k] + c];
        T F[9];
        defgrad(g, uu, F);
        if (KIND == 0) {
            T dd[4][3];
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int c = 0; c < 3; c++) dd[k][c] = sd[3 * l[k] + c];
            T dF[9], dP[9], o[4][3];
            grad4(g, dd, dF);
            Tangent<T, MAT> tg;
            tg.init(F, cmu, clam);
            tg.apply(dF, dP);
            corners(dP, g, T(1), o);
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int c = 0; c < 3; c++) sf[sfo(3 * k + c, tid, BE)] = o[k][c];
        } else if (KIND == 1) {
            T P[9], o[4][3];
            stress<T, MAT>(F, cmu, clam, P);
            corners(P, g, T(-1), o);
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int c = 0; c < 3; c++) sf[sfo(3 * k + c, tid, BE)] = o[k][c];
        } else {
            // K_aa(:,k) = V dP(e_k (x) g_a) g_a,  g_a = gradient of shape function a
            Tangent<T, MAT> tg;
            tg.init(F, cmu, clam);
            T ga[4][3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                ga[1][c] = g[c]; ga[2][c] = g[3 + c]; ga[3][c] = g[6 + c];
                ga[0][c] = -(g[c] + g[3 + c] + g[6 + c]);
            }
#pragma unroll
            for (int aa = 0; aa < 4; aa++) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    T dF[9], dP[9];
#pragma unroll
                    for (int q = 0; q < 9; q++) dF[q] = T(0);
                    dF[3 * k] = ga[aa][0]; dF[3 * k + 1] = ga[aa][1]; dF[3 * k + 2] = ga[aa][2];
                    tg.apply(dF, dP);
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        sf[sfo(aa * 9 + 3 * c + k, tid, BE)] =
                            dP[3 * c] * ga[aa][0] + dP[3 * c + 1] * ga[aa][1] + dP[3 * c + 2] * ga[aa][2];
                }
            }
        }
    }
    __syncthreads();

    const uint16_t* lp = a.lptr + (n0 + b);
    const uint16_t* le = a.lent + (size_t)b * 4 * BE;
    for (int j = tid; j < nb; j += BE) {
        const int s0 = lp[j], s1 = lp[j + 1];
        T acc[W];
#pragma unroll
        for (int q = 0; q < W; q++) acc[q] = T(0);
        for (int k = s0; k < s1; k++) {
            const int ent = le[k];
            const int el = ent >> 2, cn = ent & 3;
#pragma unroll
            for (int q = 0; q < W; q++) acc[q] += sf[sfo(cn * W + q, el, BE)];
        }
        const uint32_t fl = a.bnodes[n0 + j];
        const uint32_t node = fl & NODE_MASK;
        if (fl & FLAG_SHARED) {
            T* pp = partial + (size_t)W * (n0 + j);
#pragma unroll
            for (int q = 0; q < W; q++) pp[q] = acc[q];
        } else {
            T* op = out + (size_t)W * node;
            if ((fl & FLAG_FIXED) && mask_fixed) {
#pragma unroll
                for (int q = 0; q < W; q++) op[q] = (KIND == 2 && (q == 0 || q == 4 || q == 8)) ? T(1) : T(0);
            } else {
#pragma unroll
                for (int q = 0; q < W; q++) op[q] = acc[q] + (ext ? (T)ext[3 * (size_t)node + q % 3] : T(0));
            }
        }
    }
}

// shared nodes: sum the partial rows in block order
template <typename T, int W>
__global__ __launch_bounds__(256) void k_shared_gather(SharedArgs s, const T* __restrict__ partial, T* __restrict__ out,
                                                       const double* __restrict__ ext, int mask_fixed,
                                                       const int* __restrict__ done)
{
    if (done && *done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nshared) return;
    const uint32_t fl = s.node[i];
    const uint32_t node = fl & NODE_MASK;
    T acc[W];
#pragma unroll
    for (int q = 0; q < W; q++) acc[q] = T(0);
    for (int k = s.ptr[i]; k < s.ptr[i + 1]; k++) {
        const T* pp = partial + (size_t)W * s.idx[k];
#pragma unroll
        for (int q = 0; q < W; q++) acc[q] += pp[q];
    }
    T* op = out + (size_t)W * node;
    if ((fl & FLAG_FIXED) && mask_fixed) {
#pragma unroll
        for (int q = 0; q < W; q++) op[q] = (W == 9 && (q == 0 || q == 4 || q == 8)) ? T(1) : T(0);
    } else {
#pragma unroll
        for (int q = 0; q < W; q++) op[q] = acc[q] + (ext ? (T)ext[3 * (size_t)node + q % 3] : T(0));
    }
}

// ------------------------------------------------------------ reductions
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /* >= NV*16 */)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++)
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) lds[q * 16 + w] = v[q];
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double t = lane < nw ? lds[q * 16 + lane] : 0.0;
            for (int off = 8; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
            v[q] = t;
        }
    }
    __syncthreads();
}

// Writes this workgroup's NV sums; returns true in the one workgroup that
// finishes last, after which `tot` holds the totals (all threads).
template <int NV>
__device__ __forceinline__ bool grid_sum(double (&v)[NV], double* partials, unsigned* counter, double (&tot)[NV])
{
    __shared__ double lds[NV * 16];
    __shared__ int is_last;
    block_sum<NV>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) partials[(size_t)q * gridDim.x + blockIdx.x] = v[q];
        __threadfence();
        unsigned t = atomicAdd(counter, 1u);
        is_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        acc[q] = 0.0;
        double* pv = partials + (size_t)q * gridDim.x;
        for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x)
            acc[q] += __hip_atomic_load(pv + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum<NV>(acc, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) lds[q] = acc[q];
        *counter = 0u;     // ready for the next launch
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; q++) tot[q] = lds[q];
    return true;
}

// ---------------------------------------------------------------- CG kernels
// scalars (double) layout: see Scal in sfem_kernels.hpp

// x = 0, r = b, z = M r, p = z; rz = r.z, rr = r.r
template <int PC>
__global__ __launch_bounds__(256) void k_cg_init(int64_t n3, const float* __restrict__ b, float* __restrict__ x,
                                                 float* __restrict__ r, float* __restrict__ z, float* __restrict__ p,
                                                 const float* __restrict__ minv, const uint8_t* __restrict__ fixed,
                                                 double* scal, double* partials, unsigned* counter, int* done)
{
    const int64_t nn = n3 / 3;
    double v[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x) {
        float rv[3], zv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) { rv[c] = fixed[i] ? 0.f : b[3 * i + c]; x[3 * i + c] = 0.f; r[3 * i + c] = rv[c]; }
        if (PC) {
            const float* M = minv + 9 * i;
#pragma unroll
            for (int c = 0; c < 3; c++) zv[c] = M[3 * c] * rv[0] + M[3 * c + 1] * rv[1] + M[3 * c + 2] * rv[2];
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++) zv[c] = rv[c];
        }
#pragma unroll
        for (int c = 0; c < 3; c++) { z[3 * i + c] = zv[c]; p[3 * i + c] = zv[c]; }
        v[0] += (double)rv[0] * zv[0] + (double)rv[1] * zv[1] + (double)rv[2] * zv[2];
        v[1] += (double)rv[0] * rv[0] + (double)rv[1] * rv[1] + (double)rv[2] * rv[2];
    }
    double tot[2];
    if (grid_sum<2>(v, partials, counter, tot) && threadIdx.x == 0) {
        scal[S_RZ] = tot[0];
        scal[S_RZ0] = tot[0];
        scal[S_RR] = tot[1];
        scal[S_RR0] = tot[1];
        scal[S_ITER] = 0;
        *done = (tot[0] == 0.0) ? 1 : 0;    // IterativeSolvers.h:156-157
    }
}

// pq = p.q ; alpha = rz / pq
__global__ __launch_bounds__(256) void k_cg_alpha(int64_t n3, const float* __restrict__ p, const float* __restrict__ q,
                                                  double* scal, double* partials, unsigned* counter, int* done,
                                                  int guard)
{
    if (*done) return;
    const int64_t nn = n3 / 3;
    double v[1] = {0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x)
        v[0] += (double)p[3 * i] * q[3 * i] + (double)p[3 * i + 1] * q[3 * i + 1] + (double)p[3 * i + 2] * q[3 * i + 2];
    double tot[1];
    if (grid_sum<1>(v, partials, counter, tot) && threadIdx.x == 0) {
        scal[S_PQ] = tot[0];
        scal[S_ALPHA] = scal[S_RZ] / tot[0];
        if (guard && !(tot[0] > 0.0)) *done = 3;   // breakdown: not positive definite
    }
}

// x += a p; r -= a q; z = M r; rz' = r.z; rr = r.r; beta = rz'/rz; test
template <int PC>
__global__ __launch_bounds__(256) void k_cg_update(int64_t n3, float* __restrict__ x, float* __restrict__ r,
                                                   float* __restrict__ z, const float* __restrict__ p,
                                                   const float* __restrict__ q, const float* __restrict__ minv,
                                                   double* scal, double* partials, unsigned* counter, int* done,
                                                   int test_kind, double tol2)
{
    if (*done) return;
    const int64_t nn = n3 / 3;
    const float alpha = (float)scal[S_ALPHA];
    double v[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x) {
        float rv[3], zv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            x[3 * i + c] += alpha * p[3 * i + c];
            rv[c] = r[3 * i + c] - alpha * q[3 * i + c];
            r[3 * i + c] = rv[c];
        }
        if (PC) {
            const float* M = minv + 9 * i;
#pragma unroll
            for (int c = 0; c < 3; c++) zv[c] = M[3 * c] * rv[0] + M[3 * c + 1] * rv[1] + M[3 * c + 2] * rv[2];
#pragma unroll
            for (int c = 0; c < 3; c++) z[3 * i + c] = zv[c];
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++) zv[c] = rv[c];
        }
        v[0] += (double)rv[0] * zv[0] + (double)rv[1] * zv[1] + (double)rv[2] * zv[2];
        v[1] += (double)rv[0] * rv[0] + (double)rv[1] * rv[1] + (double)rv[2] * rv[2];
    }
    double tot[2];
    if (grid_sum<2>(v, partials, counter, tot) && threadIdx.x == 0) {
        const double rz1 = tot[0];
        scal[S_BETA] = rz1 / scal[S_RZ];
        scal[S_RZ] = rz1;
        scal[S_RR] = tot[1];
        scal[S_ITER] += 1.0;
        // test 0: the reference criterion |r.z| <= eps^2 |r0.z0| (IterativeSolvers.h:180-185)
        // test 1: true residual ||r||^2 <= tol^2 ||r0||^2
        const bool conv = test_kind == 0 ? (fabs(rz1) <= tol2 * fabs(scal[S_RZ0])) : (tot[1] <= tol2 * scal[S_RR0]);
        if (conv) *done = 1;
        else if (!(tot[1] == tot[1]) || !(rz1 == rz1)) *done = 4;   // NaN
    }
}

// p = z + beta p
template <int PC>
__global__ __launch_bounds__(256) void k_cg_p(int64_t n3, const float* __restrict__ z, const float* __restrict__ r,
                                              float* __restrict__ p, const double* scal, const int* done)
{
    if (*done) return;
    const float beta = (float)scal[S_BETA];
    const float* zz = PC ? z : r;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = zz[i] + beta * p[i];
}

// ---------------------------------------------------------------- FP64 vectors
// a.a, b.b and d.a for FP64 node vectors a, b and an FP32 direction d
// -> scal[slot], scal[slot+1], scal[slot+2]  (b, d may be null)
__global__ __launch_bounds__(256) void k_dots3(int64_t n3, const double* __restrict__ a, const double* __restrict__ b,
                                               const float* __restrict__ d, double* scal, int slot, double* partials,
                                               unsigned* counter)
{
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3 / 3; i += (int64_t)gridDim.x * blockDim.x) {
        v[0] += a[3 * i] * a[3 * i] + a[3 * i + 1] * a[3 * i + 1] + a[3 * i + 2] * a[3 * i + 2];
        if (b) v[1] += b[3 * i] * b[3 * i] + b[3 * i + 1] * b[3 * i + 1] + b[3 * i + 2] * b[3 * i + 2];
        if (d) v[2] += (double)d[3 * i] * a[3 * i] + (double)d[3 * i + 1] * a[3 * i + 1] + (double)d[3 * i + 2] * a[3 * i + 2];
    }
    double tot[3];
    if (grid_sum<3>(v, partials, counter, tot) && threadIdx.x == 0) {
        scal[slot] = tot[0];
        scal[slot + 1] = tot[1];
        scal[slot + 2] = tot[2];
    }
}

// ut = u + alpha * dx (free rows); u32 = (float) ut
__global__ __launch_bounds__(256) void k_axpy_state(int64_t n3, const double* __restrict__ u, const float* __restrict__ dx,
                                                    double alpha, double* __restrict__ ut, float* __restrict__ u32)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = u[i] + alpha * (double)dx[i];
        ut[i] = v;
        u32[i] = (float)v;
    }
}

__global__ __launch_bounds__(256) void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

// external load per node: ext = frac (m g + f_dead) + pressure gather; 0 on fixed
__global__ __launch_bounds__(256) void k_external(int64_t nn, const double* __restrict__ mass, const double* __restrict__ fdead,
                                                  double gx, double gy, double gz, double frac,
                                                  const int* __restrict__ pptr, const int* __restrict__ pidx,
                                                  const double* __restrict__ ftri, const uint8_t* __restrict__ fixed,
                                                  double* __restrict__ ext)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nn) return;
    double e[3] = {0.0, 0.0, 0.0};
    if (!fixed[i]) {
        // ComputeGradients (FemPhysicsMatrixFree.cpp:547-552): m frac g
        const double m = mass[i] * frac;
        e[0] = m * gx; e[1] = m * gy; e[2] = m * gz;
        if (fdead) {
#pragma unroll
            for (int c = 0; c < 3; c++) e[c] += frac * fdead[3 * i + c];
        }
        if (pptr) {
            for (int k = pptr[i]; k < pptr[i + 1]; k++) {
#pragma unroll
                for (int c = 0; c < 3; c++) e[c] += ftri[3 * (size_t)pidx[k] + c];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) ext[3 * i + c] = e[c];
}

// follower pressure per triangle (FemPhysicsMatrixFree.cpp:979-1013) at x = X0 + u
__global__ __launch_bounds__(256) void k_pressure(int64_t ntri, const int* __restrict__ tris, const double* __restrict__ X0,
                                                  const double* __restrict__ u, double pressure, double frac,
                                                  double* __restrict__ ftri)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    double p[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int n = tris[3 * t + k];
#pragma unroll
        for (int c = 0; c < 3; c++) p[k][c] = X0[3 * (size_t)n + c] + u[3 * (size_t)n + c];
    }
    double a[3], b[3];
#pragma unroll
    for (int c = 0; c < 3; c++) { a[c] = p[1][c] - p[0][c]; b[c] = p[2][c] - p[0][c]; }
    double n[3] = {a[1] * b[2] - b[1] * a[2], -a[0] * b[2] + a[2] * b[0], a[0] * b[1] - b[0] * a[1]};
    const double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double area = 0.5 * len;
    const double s = 1.0 / len;
#pragma unroll
    for (int c = 0; c < 3; c++) ftri[3 * t + c] = frac * ((area / 3.0) * (pressure * (n[c] * s)));
}

// per-node 3x3 block inverse (FP64 arithmetic), fixed nodes -> 0
__global__ __launch_bounds__(256) void k_invert_blocks(int64_t nn, const float* __restrict__ B, const uint8_t* __restrict__ fixed,
                                                       float* __restrict__ Minv)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nn) return;
    double m[9], v[9];
#pragma unroll
    for (int k = 0; k < 9; k++) m[k] = B[9 * i + k];
    inv3(m, v);
#pragma unroll
    for (int k = 0; k < 9; k++) Minv[9 * i + k] = fixed[i] ? 0.f : (float)v[k];
}

// identity blocks on free nodes (unpreconditioned CG written as PCG)
__global__ __launch_bounds__(256) void k_identity_blocks(int64_t nn, const uint8_t* __restrict__ fixed, float* __restrict__ Minv)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nn) return;
#pragma unroll
    for (int k = 0; k < 9; k++) Minv[9 * i + k] = (!fixed[i] && (k == 0 || k == 4 || k == 8)) ? 1.f : 0.f;
}

// ------------------------------------------------------------------ launchers
static int grid_for(int64_t n, int bs, int cap)
{
    int64_t g = (n + bs - 1) / bs;
    if (g < 1) g = 1;
    return (int)(g > cap ? cap : g);
}

template <int KIND, int BE>
static hipError_t launch_elem_mat(int mat, const BlockArgs& a, const void* u, const void* d, void* out, void* partial,
                                  const double* ext, int mask_fixed, const int* done, size_t lds, hipStream_t s)
{
    typedef typename KindTraits<KIND>::T T;
    dim3 grid(a.nblk), block(BE);
    switch (mat) {
    case LINEAR:
        hipLaunchKernelGGL((k_elem_block<KIND, LINEAR, BE>), grid, block, lds, s, a, (const T*)u, (const T*)d, (T*)out, (T*)partial, ext, mask_fixed, done);
        break;
    case STVK:
        hipLaunchKernelGGL((k_elem_block<KIND, STVK, BE>), grid, block, lds, s, a, (const T*)u, (const T*)d, (T*)out, (T*)partial, ext, mask_fixed, done);
        break;
    case COROTATIONAL:
        hipLaunchKernelGGL((k_elem_block<KIND, COROTATIONAL, BE>), grid, block, lds, s, a, (const T*)u, (const T*)d, (T*)out, (T*)partial, ext, mask_fixed, done);
        break;
    default:
        hipLaunchKernelGGL((k_elem_block<KIND, NEO_HOOKEAN, BE>), grid, block, lds, s, a, (const T*)u, (const T*)d, (T*)out, (T*)partial, ext, mask_fixed, done);
        break;
    }
    return hipGetLastError();
}

template <int KIND>
static size_t lds_bytes(const BlockArgs& a)
{
    typedef typename KindTraits<KIND>::T T;
    constexpr int W = KindTraits<KIND>::W;
    return sizeof(T) * ((size_t)4 * W * BLOCK_E + (size_t)3 * a.nmax * (KindTraits<KIND>::HAS_D ? 2 : 1));
}

template <typename T, int W>
static hipError_t launch_shared(const SharedArgs& s, const void* partial, void* out, const double* ext, int mask_fixed,
                                const int* done, hipStream_t st)
{
    if (s.nshared == 0) return hipSuccess;
    hipLaunchKernelGGL((k_shared_gather<T, W>), dim3((s.nshared + 255) / 256), dim3(256), 0, st, s, (const T*)partial, (T*)out, ext, mask_fixed, done);
    return hipGetLastError();
}

hipError_t launch_hessvec(int mat, const BlockArgs& a, const SharedArgs& s, const float* u32, const float* d32, float* out,
                          float* partial, const int* done, hipStream_t st)
{
    hipError_t e = launch_elem_mat<0, BLOCK_E>(mat, a, u32, d32, out, partial, nullptr, 1, done, lds_bytes<0>(a), st);
    if (e != hipSuccess) return e;
    return launch_shared<float, 3>(s, partial, out, nullptr, 1, done, st);
}

hipError_t launch_shared_only(const SharedArgs& s, const float* partial, float* out, hipStream_t st)
{
    return launch_shared<float, 3>(s, partial, out, nullptr, 1, nullptr, st);
}

hipError_t launch_force(int mat, const BlockArgs& a, const SharedArgs& s, const double* u64, double* out, double* partial,
                        const double* ext, int mask_fixed, hipStream_t st)
{
    hipError_t e = launch_elem_mat<1, BLOCK_E>(mat, a, u64, nullptr, out, partial, ext, mask_fixed, nullptr, lds_bytes<1>(a), st);
    if (e != hipSuccess) return e;
    return launch_shared<double, 3>(s, partial, out, ext, mask_fixed, nullptr, st);
}

hipError_t launch_blockdiag(int mat, const BlockArgs& a, const SharedArgs& s, const float* u32, float* out, float* partial,
                            hipStream_t st)
{
    hipError_t e = launch_elem_mat<2, BLOCK_E>(mat, a, u32, nullptr, out, partial, nullptr, 1, nullptr, lds_bytes<2>(a), st);
    if (e != hipSuccess) return e;
    return launch_shared<float, 9>(s, partial, out, nullptr, 1, nullptr, st);
}

size_t elem_lds_max(const BlockArgs& a)
{
    size_t m = lds_bytes<0>(a);
    if (lds_bytes<1>(a) > m) m = lds_bytes<1>(a);
    if (lds_bytes<2>(a) > m) m = lds_bytes<2>(a);
    return m;
}

hipError_t set_elem_lds_limits(const BlockArgs& a)
{
    // allow dynamic LDS above the 64 KiB default for unstructured meshes
    const int lim = 160 * 1024;
    const void* fns[] = {
        (const void*)k_elem_block<0, LINEAR, BLOCK_E>, (const void*)k_elem_block<0, STVK, BLOCK_E>,
        (const void*)k_elem_block<0, COROTATIONAL, BLOCK_E>, (const void*)k_elem_block<0, NEO_HOOKEAN, BLOCK_E>,
        (const void*)k_elem_block<1, LINEAR, BLOCK_E>, (const void*)k_elem_block<1, STVK, BLOCK_E>,
        (const void*)k_elem_block<1, COROTATIONAL, BLOCK_E>, (const void*)k_elem_block<1, NEO_HOOKEAN, BLOCK_E>,
        (const void*)k_elem_block<2, LINEAR, BLOCK_E>, (const void*)k_elem_block<2, STVK, BLOCK_E>,
        (const void*)k_elem_block<2, COROTATIONAL, BLOCK_E>, (const void*)k_elem_block<2, NEO_HOOKEAN, BLOCK_E>,
    };
    if (elem_lds_max(a) <= 64 * 1024) return hipSuccess;
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_cg_init(int pc, int64_t n3, const float* b, float* x, float* r, float* z, float* p, const float* minv,
                          const uint8_t* fixed, double* scal, double* partials, unsigned* counter, int* done, hipStream_t st)
{
    const int g = grid_for(n3 / 3, 256, RED_GRID);
    if (pc) hipLaunchKernelGGL(k_cg_init<1>, dim3(g), dim3(256), 0, st, n3, b, x, r, z, p, minv, fixed, scal, partials, counter, done);
    else hipLaunchKernelGGL(k_cg_init<0>, dim3(g), dim3(256), 0, st, n3, b, x, r, z, p, minv, fixed, scal, partials, counter, done);
    return hipGetLastError();
}

hipError_t launch_cg_iter_tail(int pc, int64_t n3, float* x, float* r, float* z, float* p, const float* q, const float* minv,
                               double* scal, double* partials, unsigned* counter, int* done, int test_kind, double tol2,
                               int guard, hipStream_t st)
{
    const int g = grid_for(n3 / 3, 256, RED_GRID);
    hipLaunchKernelGGL(k_cg_alpha, dim3(g), dim3(256), 0, st, n3, p, q, scal, partials, counter, done, guard);
    if (pc) hipLaunchKernelGGL(k_cg_update<1>, dim3(g), dim3(256), 0, st, n3, x, r, z, p, q, minv, scal, partials, counter, done, test_kind, tol2);
    else hipLaunchKernelGGL(k_cg_update<0>, dim3(g), dim3(256), 0, st, n3, x, r, z, p, q, minv, scal, partials, counter, done, test_kind, tol2);
    const int gp = grid_for(n3, 256, 4096);
    if (pc) hipLaunchKernelGGL(k_cg_p<1>, dim3(gp), dim3(256), 0, st, n3, z, r, p, scal, done);
    else hipLaunchKernelGGL(k_cg_p<0>, dim3(gp), dim3(256), 0, st, n3, z, r, p, scal, done);
    return hipGetLastError();
}

hipError_t launch_dots3(int64_t n3, const double* a, const double* b, const float* d, double* scal, int slot,
                        double* partials, unsigned* counter, hipStream_t st)
{
    const int g = grid_for(n3 / 3, 256, RED_GRID);
    hipLaunchKernelGGL(k_dots3, dim3(g), dim3(256), 0, st, n3, a, b, d, scal, slot, partials, counter);
    return hipGetLastError();
}

hipError_t launch_axpy_state(int64_t n3, const double* u, const float* dx, double alpha, double* ut, float* u32, hipStream_t st)
{
    hipLaunchKernelGGL(k_axpy_state, dim3(grid_for(n3, 256, 8192)), dim3(256), 0, st, n3, u, dx, alpha, ut, u32);
    return hipGetLastError();
}

hipError_t launch_to_f32(int64_t n, const double* a, float* b, hipStream_t st)
{
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, n, a, b);
    return hipGetLastError();
}

hipError_t launch_external(int64_t nn, const double* mass, const double* fdead, const double* g, double frac, const int* pptr,
                           const int* pidx, const double* ftri, const uint8_t* fixed, double* ext, hipStream_t st)
{
    hipLaunchKernelGGL(k_external, dim3((nn + 255) / 256), dim3(256), 0, st, nn, mass, fdead, g[0], g[1], g[2], frac, pptr, pidx, ftri, fixed, ext);
    return hipGetLastError();
}

hipError_t launch_pressure(int64_t ntri, const int* tris, const double* X0, const double* u, double pressure, double frac,
                           double* ftri, hipStream_t st)
{
    if (ntri == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pressure, dim3((ntri + 255) / 256), dim3(256), 0, st, ntri, tris, X0, u, pressure, frac, ftri);
    return hipGetLastError();
}

hipError_t launch_invert_blocks(int64_t nn, const float* B, const uint8_t* fixed, float* Minv, hipStream_t st)
{
    hipLaunchKernelGGL(k_invert_blocks, dim3((nn + 255) / 256), dim3(256), 0, st, nn, B, fixed, Minv);
    return hipGetLastError();
}

hipError_t launch_identity_blocks(int64_t nn, const uint8_t* fixed, float* Minv, hipStream_t st)
{
    hipLaunchKernelGGL(k_identity_blocks, dim3((nn + 255) / 256), dim3(256), 0, st, nn, fixed, Minv);
    return hipGetLastError();
}

}  // namespace sfem
