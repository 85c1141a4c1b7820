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
// workgroup, per-workgroup partials, and every workgroup of the next kernel
// that needs the total sums the partials itself, in the same fixed order, and
// derives the CG scalars on the device.  (A last-workgroup ticket scheme needs
// a device-scope fence per workgroup, which writes back the XCD's L2: far
// slower here than the redundant sums.)
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sfem_device.hpp"
#include "sfem_kernels.hpp"
#include "sfem_mgdev.hpp"
#include "sfem_reduce.hpp"

namespace sfem {

constexpr uint32_t NODE_MASK = 0x1fffffffu;
constexpr uint32_t FLAG_SHARED = 0x80000000u;
constexpr uint32_t FLAG_FIXED = 0x40000000u;
constexpr uint32_t FLAG_OWNER = 0x20000000u;   // the node's only block, or the first of its blocks

// one workgroup: scal[slot + q] = sum of the NV partial arrays
template <int NV>
__global__ __launch_bounds__(256) void k_finish(const double* __restrict__ part, int n, double* scal, int slot)
{
    double tot[NV];
    sum_partials<NV, 256>(part, n, n, tot);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) scal[slot + q] = tot[q];
}

// ------------------------------------------------------- element block kernel
// KIND 0: Hessian-vector (FP32): out = K d            (u, d staged)
// KIND 1: internal force (FP64):  out = f_int (+ ext) (u staged)
// KIND 2: diagonal 3x3 blocks (FP32): out = K_ii      (u staged), W = 9
template <int KIND> struct KindTraits;
// HAS_X: Dm^-1 and V from the staged rest coordinates instead of per-element
// arrays (the per-element bytes dominate this kernel's HBM traffic)
template <> struct KindTraits<0> {
    typedef float T; static constexpr int W = 3; static constexpr bool HAS_D = true; static constexpr bool HAS_X = true;
};
template <> struct KindTraits<1> {
    typedef double T; static constexpr int W = 3; static constexpr bool HAS_D = false; static constexpr bool HAS_X = false;
};
template <> struct KindTraits<2> {
    typedef float T; static constexpr int W = 9; static constexpr bool HAS_D = false; static constexpr bool HAS_X = false;
};
// KIND 3: the Neo-Hookean tangent's state per element (nh_lin_state) into
// out in the BlockArgs::lin32 layout; no corner or node phase
template <> struct KindTraits<3> {
    typedef float T; static constexpr int W = 3; static constexpr bool HAS_D = false; static constexpr bool HAS_X = true;
};

// LDS layout (vec4 = 4 x T, one ds_read/ds_write_b128 per row):
//   corner outputs: CW vec4 per incidence slot k, in the block's node-sorted
//       incidence order, sf4[k * CW + j]  (CW = 1 for W = 3, 3 for W = 9); an
//       element writes its corner c to the slot lperm[element] gives it, so a
//       node's incidences are the contiguous slots [s_ptr[j], s_ptr[j+1])
//   staged node rows: su4[nmax] (not on the cached state; and sd4[nmax],
//   sx4[nmax]) = (x, y, z, 0);
//   sx4: rest coordinates relative to the block's centre
//   per row m / h^2 for dynamics s_md[nmax] (T, dynamics instances only),
//   the row's output s_out[nmax]: its partial slot | FLAG_SHARED for a shared
//   node, else its bnodes entry (node id and flags)
//   incidence offsets s_ptr[nmax+1]
template <typename T> struct Vec4;
template <> struct Vec4<float> { typedef float4 V; };
template <> struct Vec4<double> { typedef double4 V; };

template <int KIND>
__host__ __device__ constexpr int corner_vec4(int W_)
{
    return (W_ + 3) / 4;
}

// lc: the kernel on the cached Neo-Hookean state (no staged displacement
// rows); dyn: the implicit-dynamics instances (m / h^2 per row)
template <int KIND>
size_t lds_layout_bytes(int BE, int nmax, bool lc = false, bool dyn = false)
{
    typedef typename KindTraits<KIND>::T T;
    constexpr int W = KindTraits<KIND>::W;
    const size_t v4 = 4 * sizeof(T);
    const size_t rows =
        (size_t)nmax * ((lc ? 0 : 1) + (KindTraits<KIND>::HAS_D ? 1 : 0) + (KindTraits<KIND>::HAS_X ? 1 : 0));
    return v4 * ((size_t)4 * corner_vec4<KIND>(W) * BE + rows) + (sizeof(uint32_t) + (dyn ? sizeof(T) : 0)) * (size_t)nmax +
           sizeof(uint16_t) * ((size_t)nmax + 1) + 16;
}

// the CG scalars and status of one convergence test (one thread of the grid):
// the stagnation reference is read and written by that thread only
// (returns the flag's new value; the outcome also into slot when given)
__device__ __forceinline__ int cg_test_record(double* scal, int* done, double rz1, double rr, double beta, bool conv,
                                              bool bad, bool indef, int stall_window, CgSlot* slot = nullptr,
                                              double rr0 = 1.0)
{
    const double it = scal[S_ITER] + 1.0;
    scal[S_RZ] = rz1;
    scal[S_RR] = rr;
    scal[S_BETA] = beta;
    scal[S_ITER] = it;
    int st = CG_RUNNING;
    if (conv) st = CG_CONVERGED;
    else if (bad) st = CG_NOT_FINITE;
    else if (indef) st = CG_BREAKDOWN;
    else if (stall_window > 0) {
        if (rr <= 0.5 * scal[S_RR_BEST]) {
            scal[S_RR_BEST] = rr;
            scal[S_IT_BEST] = it;
        } else if (it - scal[S_IT_BEST] >= (double)stall_window) {
            st = CG_STAGNATED;
        }
    }
    if (st != CG_RUNNING) *done = st;
    if (slot) {
        slot->rr_rel = rr / rr0;
        slot->iter = (int)it;
        slot->done = st;
    }
    return st;
}

// workgroups are dispatched to the 8 XCDs round robin (workgroup w to XCD
// w % 8); each XCD gets a contiguous range of the space-filling-curve blocks,
// so that the node rows that neighbouring blocks share are fetched into its
// own L2 once
__device__ __forceinline__ int xcd_block(int w, int nblk)
{
    const int per = nblk >> 3, rem = nblk & 7, xcd = w & 7, loc = w >> 3;
    return xcd < rem ? xcd * (per + 1) + loc : rem * (per + 1) + (xcd - rem) * per + loc;
}

// DYN: implicit dynamics (mdiag non-null); a separate instance so that static
// solves carry none of its code.  SG (KIND 0): how the block forms the rows
// it applies K to -- 0 the given d; 1 (PF) the CG direction update of the
// previous iteration (PFuse), d is not read; MG_STAGE_CHEB and
// MG_STAGE_PROLONG a multigrid smoothing step or prolongation (MgStage).
template <int KIND, int MAT, int BE, bool DYN, int SG>
__global__ __launch_bounds__(BE) void k_elem_block(BlockArgs a, const typename KindTraits<KIND>::T* __restrict__ u,
                                                    const typename KindTraits<KIND>::T* __restrict__ d,
                                                    typename KindTraits<KIND>::T* __restrict__ out,
                                                    typename KindTraits<KIND>::T* __restrict__ partial,
                                                    const double* __restrict__ ext, int mask_fixed,
                                                    int* __restrict__ done, double* __restrict__ dot_part,
                                                    const typename KindTraits<KIND>::T* __restrict__ mdiag, PFuse pf,
                                                    MgStage mg)
{
    constexpr bool PF = SG == 1;
    constexpr bool MGC = SG == MG_STAGE_CHEB;
    constexpr bool MGP = SG == MG_STAGE_PROLONG;
    // LC: the Hessian-vector product on the cached Neo-Hookean state (u not read)
    constexpr bool LC = MAT == NEO_HOOKEAN_LIN;
    static_assert(SG == 0 || KIND == 0, "staged directions only in the Hessian-vector kernel");
    static_assert(!LC || (KIND == 0 && !DYN), "cached state: static Hessian-vector kernel only");
    static_assert(!(DYN && (MGC || MGP)), "multigrid staging is static only");
    typedef typename KindTraits<KIND>::T T;
    typedef typename Vec4<T>::V V4;
    constexpr int W = KindTraits<KIND>::W;
    constexpr int CW = corner_vec4<KIND>(W);
    constexpr bool HAS_D = KindTraits<KIND>::HAS_D;
    constexpr bool HAS_X = KindTraits<KIND>::HAS_X;
    // the done flag (never null: launches outside CG pass a flag that stays 0)
    // is read with the first-round loads and tested before the staging loop,
    // so its round trip does not precede theirs
    const int stop = vload(done);
    extern __shared__ __align__(32) unsigned char smem_raw[];
    V4* sf4 = reinterpret_cast<V4*>(smem_raw);        // 4*CW*BE corner rows
    V4* su4 = sf4 + 4 * CW * BE;                      // nmax (not LC)
    V4* sd4 = su4 + (LC ? 0 : a.nmax);                // nmax (HAS_D)
    V4* sx4 = sd4 + (HAS_D ? a.nmax : 0);             // nmax (HAS_X)
    T* s_md = reinterpret_cast<T*>(sx4 + (HAS_X ? a.nmax : 0));                  // nmax (DYN)
    uint32_t* s_out = reinterpret_cast<uint32_t*>(s_md + (DYN ? a.nmax : 0));     // nmax
    uint16_t* s_ptr = reinterpret_cast<uint16_t*>(s_out + a.nmax);                // nmax+1

    const int b = xcd_block(blockIdx.x, a.nblk);
    const int tid = threadIdx.x;
    // PF: the update's partials, issued before every other load
    constexpr int PU = 4;
    double rzv[2][PU];
    double rz0 = 0.0, rr0 = 0.0, rz_prev = 0.0;
    if (PF) {
        partials_issue<2, BE, PU>(pf.rz_part, pf.n_rz, pf.n_rz, rzv);
        // written by the previous kernels, read before the reduction needs them
        rz0 = pf.scal[S_RZ0];
        rr0 = pf.scal[S_RR0];
        rz_prev = pf.scal[S_RZ_PREV];
    }
    const int n0 = vload(a.bnode_ptr + b);
    const int nb = vload(a.bnode_ptr + b + 1) - n0;

    // element data first: independent of the node staging, so its latency
    // overlaps the gathers below
    const int64_t e = (int64_t)b * BE + tid;
    const bool live = e < a.ne;
    T g[9];
    T cmu = T(0), clam = T(0);
    T lE[9], lc = T(0);
    uint2 li = make_uint2(0, 0), lp4 = make_uint2(0, 0);
    if (live && HAS_X) {
        if (a.mat32) {
            cmu = (T)a.mat32[e];
            clam = (T)a.mat32[(size_t)a.ne_pad + e];
        } else {
            cmu = (T)a.mu0;
            clam = (T)a.lam0;
        }
        li = a.lidx[e];
        lp4 = a.lperm[e];
    } else if (live) {
        const size_t st = (size_t)a.ne_pad;
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
            if (KIND == 1 && a.coef_mode) {
                // d f / d mu or d f / d lambda: coefficients (V, 0) or (0, V),
                // V = 1 / (6 det Dm^-1) (ElasticEnergy.cpp:284-315)
                const T dg = g[0] * (g[4] * g[8] - g[5] * g[7]) - g[1] * (g[3] * g[8] - g[5] * g[6]) +
                             g[2] * (g[3] * g[7] - g[4] * g[6]);
                const T V = T(1) / (T(6) * dg);
                cmu = a.coef_mode == 1 ? V : T(0);
                clam = a.coef_mode == 2 ? V : T(0);
            }
        }
        li = a.lidx[e];
        lp4 = a.lperm[e];
    }
    // the block's incidence offsets: this thread's first one is loaded now and
    // stored to LDS after the node staging
    const uint16_t* lp = a.lptr + (n0 + b);
    const uint16_t lp0 = tid <= nb ? lp[tid] : uint16_t(0);
    if (stop) {
        if (PF && pf.slot && blockIdx.x == 0 && tid == 0) {
            pf.slot->rr_rel = 0.0;
            pf.slot->iter = (int)pf.scal[S_ITER];
            pf.slot->done = stop;
        }
        return;
    }
    // stage block nodes (rows, flags, partial slots): a row's loads, then its
    // LDS stores; the first round of loads is issued before the PF test
    struct Row {
        uint32_t fl, node;
        int slot;
        V4 uv, dv, xv;
        T zr[3], po[3], md;
        // multigrid staging: the node's slot range, its r, d, M, x and the
        // rows of the previous images (MGC); its prolongation range (MGP)
        int2 nq;
        float ro[3], dold[3], mb[9], qa[3], xo[3];
        int pp0, pp1;
    };
    // every load of a row is issued before any of its values is tested
    auto load_row = [&](int j, Row& w) {
        w.fl = a.bnodes[n0 + j];
        w.node = w.fl & NODE_MASK;
        if (!LC) {
            const T* up = u + 3 * (size_t)w.node;
            w.uv.x = up[0]; w.uv.y = up[1]; w.uv.z = up[2]; w.uv.w = T(0);
        }
        if (PF) {
            const float* zp = pf.z + 3 * (size_t)w.node;
            const float* pp = pf.p_old + 3 * (size_t)w.node;
#pragma unroll
            for (int c = 0; c < 3; c++) { w.zr[c] = zp[c]; w.po[c] = pp[c]; }
        } else if (MGC) {
            const size_t n3 = 3 * (size_t)w.node;
            w.nq = a.node_q[w.node];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                w.ro[c] = mg.r_old[n3 + c]; w.dold[c] = mg.d_old[n3 + c]; w.qa[c] = mg.qa[n3 + c];
                w.xo[c] = mg.x[n3 + c];
            }
#pragma unroll
            for (int k = 0; k < 9; k++) w.mb[k] = mg.minv[9 * (size_t)w.node + k];
        } else if (MGP) {
            w.pp0 = mg.p_ptr[w.node];
            w.pp1 = mg.p_ptr[w.node + 1];
#pragma unroll
            for (int c = 0; c < 3; c++) w.xo[c] = mg.x[3 * (size_t)w.node + c];
        } else if (HAS_D) {
            const T* dp = d + 3 * (size_t)w.node;
            w.dv.x = dp[0]; w.dv.y = dp[1]; w.dv.z = dp[2]; w.dv.w = T(0);
        }
        if (HAS_X) {
            const float* xp = a.rel32 + 3 * (size_t)(n0 + j);
            w.xv.x = (T)xp[0]; w.xv.y = (T)xp[1]; w.xv.z = (T)xp[2]; w.xv.w = T(0);
        }
        // unconditional loads (bslot is -1 on unshared rows): a load under a
        // branch would be waited for where the branch joins
        w.slot = a.bslot[n0 + j];
        w.md = DYN ? mdiag[w.node] : T(0);
    };
    auto store_row = [&](int j, Row& w, float beta) {
        // the row's owner: its only block, or the first block of a shared node
        // (dynamics: that row carries m / h^2, so the node's partial rows add
        // up to one mass term; PF: that block stores p)
        const bool owner = (w.fl & FLAG_OWNER) != 0;
        if (PF) {
            w.dv.x = w.zr[0] + beta * w.po[0]; w.dv.y = w.zr[1] + beta * w.po[1]; w.dv.z = w.zr[2] + beta * w.po[2];
            w.dv.w = T(0);
            if (owner) {
                float* pn = pf.p_new + 3 * (size_t)w.node;
                pn[0] = w.dv.x; pn[1] = w.dv.y; pn[2] = w.dv.z;
            }
        }
        if (MGC) {
            // the k_mg_cheb step of this node (same expressions): complete the
            // previous image, r, then d = c1 d_old + c2 M r and x += d
            float qv[3], r1[3], mr[3], dn[3];
            complete_row(w.nq, w.qa, mg.pa, qv);
#pragma unroll
            for (int c = 0; c < 3; c++) r1[c] = w.ro[c] - qv[c];
            block_mul(w.mb, r1, mr);
#pragma unroll
            for (int c = 0; c < 3; c++) dn[c] = cheb_dir(mg.has_dold != 0, mg.c1, w.dold[c], mg.c2, mr[c]);
            if (owner) {
                const size_t n3 = 3 * (size_t)w.node;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    mg.r_new[n3 + c] = r1[c];
                    mg.d_new[n3 + c] = dn[c];
                    mg.x[n3 + c] = w.xo[c] + dn[c];
                }
            }
            w.dv.x = dn[0]; w.dv.y = dn[1]; w.dv.z = dn[2]; w.dv.w = T(0);
        } else if (MGP) {
            // the k_mg_prolong row of this node: e = P xc, 0 at a fixed node
            float ev[3] = {0.f, 0.f, 0.f};
            gather_rows(w.pp0, w.pp1, mg.p_col, mg.p_val, mg.xc, ev);
            const bool f = (w.fl & FLAG_FIXED) != 0;
#pragma unroll
            for (int c = 0; c < 3; c++) ev[c] = f ? 0.f : ev[c];
            if (owner) {
                const size_t n3 = 3 * (size_t)w.node;
#pragma unroll
                for (int c = 0; c < 3; c++) mg.x[n3 + c] = w.xo[c] + ev[c];
            }
            w.dv.x = ev[0]; w.dv.y = ev[1]; w.dv.z = ev[2]; w.dv.w = T(0);
        }
        if (!LC) su4[j] = w.uv;
        if (HAS_D) sd4[j] = w.dv;
        if (HAS_X) sx4[j] = w.xv;
        if (DYN) s_md[j] = owner ? w.md : T(0);
        s_out[j] = (w.fl & FLAG_SHARED) ? ((uint32_t)w.slot | FLAG_SHARED) : w.fl;
    };
    Row r0;
    if (tid < nb) load_row(tid, r0);
    float beta = 0.f;
    if (PF) {
        // beta and the test of the previous iteration's residual, as k_cg_p
        // makes them (every workgroup reaches the same decision)
        double t[2];
        partials_finish<2, BE, PU>(pf.rz_part, pf.n_rz, pf.n_rz, rzv, t);
        const double rz1 = t[0], rr = t[1];
        const bool conv = pf.test_kind == 0 ? (fabs(rz1) <= pf.tol2 * fabs(rz0)) : (rr <= pf.tol2 * rr0);
        const bool bad = !(rr == rr) || !(rz1 == rz1);
        const bool indef = pf.rz_guard && !conv && !bad && !(rz1 > 0.0);
        const double beta_d = rz1 / rz_prev;
        if (blockIdx.x == 0 && tid == 0)
            cg_test_record(pf.scal, done, rz1, rr, beta_d, conv, bad, indef, pf.stall_window, pf.slot, rr0);
        if (conv || bad || indef) return;
        beta = (float)beta_d;
    }
    // LC: the element's cached state, issued after the first row loads and the
    // PF reduction (loads complete in order: the row stores and the reduction
    // do not wait for these) and unconditionally (in bounds for every thread
    // of the block; a load under a branch is waited for where the branch
    // joins); the block's states are contiguous, one 256 B run per wave and
    // value
    if (LC) {
        const float* lb = a.lin32 + (size_t)b * 10 * BE + tid;
#pragma unroll
        for (int k = 0; k < 9; k++) lE[k] = lb[k * BE];
        lc = lb[9 * BE];
    }
    if (tid < nb) store_row(tid, r0, beta);
    for (int j = tid + BE; j < nb; j += BE) {
        Row w;
        load_row(j, w);
        store_row(j, w, beta);
    }
    if (tid <= nb) s_ptr[tid] = lp0;
    for (int j = tid + BE; j <= nb; j += BE) s_ptr[j] = lp[j];
    __syncthreads();

    if (live) {
        const int l[4] = {(int)(li.x & 0xffffu), (int)(li.x >> 16), (int)(li.y & 0xffffu), (int)(li.y >> 16)};
        if (HAS_X) {
            // Dm = [X1 - X0, X2 - X0, X3 - X0] (columns), Dm^-1 and V = det / 6 as
            // at set-up (FemPhysicsMatrixFree.cpp:126-146), in FP32
            T X[4][3];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const V4 v = sx4[l[k]];
                X[k][0] = v.x; X[k][1] = v.y; X[k][2] = v.z;
            }
            T D[9];
#pragma unroll
            for (int k = 1; k < 4; k++)
#pragma unroll
                for (int r = 0; r < 3; r++) D[3 * r + (k - 1)] = X[k][r] - X[0][r];
            const T det = inv3(D, g);
            const T V = det * T(1.0 / 6.0);
            cmu *= V;
            clam *= V;
        }
        if constexpr (LC) {
            T dd[4][3], o[4][3];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const V4 v = sd4[l[k]];
                dd[k][0] = v.x; dd[k][1] = v.y; dd[k][2] = v.z;
            }
            nh_lin_corners(g, lE, cmu, clam, lc, dd, o);
            const int sl[4] = {(int)(lp4.x & 0xffffu), (int)(lp4.x >> 16), (int)(lp4.y & 0xffffu), (int)(lp4.y >> 16)};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                V4 v;
                v.x = o[k][0]; v.y = o[k][1]; v.z = o[k][2]; v.w = T(0);
                sf4[sl[k]] = v;
            }
        } else {
            T uu[4][3];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const V4 v = su4[l[k]];
                uu[k][0] = v.x; uu[k][1] = v.y; uu[k][2] = v.z;
            }
            T F[9];
            defgrad(g, uu, F);
            if constexpr (KIND == 3) {
                T E[9], cc;
                nh_lin_state(g, F, cmu, clam, E, cc);
                T* ob = out + (size_t)b * 10 * BE + tid;
#pragma unroll
                for (int k = 0; k < 9; k++) ob[k * BE] = E[k];
                ob[9 * BE] = cc;
            } else if (KIND == 0 || KIND == 1) {
                T o[4][3];
                if (KIND == 0) {
                    T dd[4][3];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const V4 v = sd4[l[k]];
                        dd[k][0] = v.x; dd[k][1] = v.y; dd[k][2] = v.z;
                    }
                    T dF[9], dP[9];
                    grad4(g, dd, dF);
                    Tangent<T, MAT> tg;
                    tg.init(F, cmu, clam);
                    tg.apply(dF, dP);
                    corners(dP, g, T(1), o);
                } else {
                    T P[9];
                    stress<T, MAT>(F, cmu, clam, P);
                    corners(P, g, T(-1), o);
                }
                const int sl[4] = {(int)(lp4.x & 0xffffu), (int)(lp4.x >> 16), (int)(lp4.y & 0xffffu), (int)(lp4.y >> 16)};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    V4 v;
                    v.x = o[k][0]; v.y = o[k][1]; v.z = o[k][2]; v.w = T(0);
                    sf4[sl[k]] = v;
                }
            } else {
                // K_aa(:,k) = V dP(e_k (x) g_a) g_a,  g_a = gradient of shape function a
                Tangent<T, MAT> tg;
                tg.init(F, cmu, clam);
                const int sl[4] = {(int)(lp4.x & 0xffffu), (int)(lp4.x >> 16), (int)(lp4.y & 0xffffu), (int)(lp4.y >> 16)};
                T ga[4][3];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    ga[1][c] = g[c]; ga[2][c] = g[3 + c]; ga[3][c] = g[6 + c];
                    ga[0][c] = -(g[c] + g[3 + c] + g[6 + c]);
                }
#pragma unroll
                for (int aa = 0; aa < 4; aa++) {
                    T blk[12];
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        T dF[9], dP[9];
#pragma unroll
                        for (int q = 0; q < 9; q++) dF[q] = T(0);
                        dF[3 * k] = ga[aa][0]; dF[3 * k + 1] = ga[aa][1]; dF[3 * k + 2] = ga[aa][2];
                        tg.apply(dF, dP);
#pragma unroll
                        for (int c = 0; c < 3; c++)
                            blk[3 * c + k] = dP[3 * c] * ga[aa][0] + dP[3 * c + 1] * ga[aa][1] + dP[3 * c + 2] * ga[aa][2];
                    }
                    blk[9] = blk[10] = blk[11] = T(0);
#pragma unroll
                    for (int jv = 0; jv < CW; jv++) {
                        V4 v;
                        v.x = blk[4 * jv]; v.y = blk[4 * jv + 1]; v.z = blk[4 * jv + 2]; v.w = blk[4 * jv + 3];
                        sf4[sl[aa] * CW + jv] = v;
                    }
                }
            }
        }
    }
    if constexpr (KIND == 3) return;
    __syncthreads();

    // node phase: two adjacent lanes per node split its contiguous incidence
    // slots (even / odd, two loads in flight each) and combine by a lane swap
    double dot = 0.0;
    const int half = tid & 1;
    for (int jb = 0; jb < nb; jb += BE / 2) {
        const int j = jb + (tid >> 1);
        const bool has = j < nb;
        T acc[4 * CW];
#pragma unroll
        for (int q = 0; q < 4 * CW; q++) acc[q] = T(0);
        if (has) {
            const int s1 = s_ptr[j + 1];
            int k = s_ptr[j] + half;
            for (; k + 2 < s1; k += 4) {
#pragma unroll
                for (int jv = 0; jv < CW; jv++) {
                    const V4 v0 = sf4[k * CW + jv];
                    const V4 v1 = sf4[(k + 2) * CW + jv];
                    acc[4 * jv] += v0.x + v1.x; acc[4 * jv + 1] += v0.y + v1.y;
                    acc[4 * jv + 2] += v0.z + v1.z; acc[4 * jv + 3] += v0.w + v1.w;
                }
            }
            if (k < s1) {
#pragma unroll
                for (int jv = 0; jv < CW; jv++) {
                    const V4 v0 = sf4[k * CW + jv];
                    acc[4 * jv] += v0.x; acc[4 * jv + 1] += v0.y; acc[4 * jv + 2] += v0.z; acc[4 * jv + 3] += v0.w;
                }
            }
        }
        // both lanes of the pair are in the same wave: add the partner's sums
#pragma unroll
        for (int q = 0; q < 4 * CW; q++) acc[q] += __shfl_xor(acc[q], 1, 64);
        if (DYN && has && half == 0) {
            // M / h^2 (implicit dynamics, MatrixVectorMultiply FemPhysicsMatrixFree.cpp:523-534;
            // the residual's -M / h^2 (x - x0), ComputeGradients :553-557)
            const T md = s_md[j];
            if (KIND == 0) {
                const V4 dv = sd4[j];
                acc[0] += md * dv.x; acc[1] += md * dv.y; acc[2] += md * dv.z;
            } else if (KIND == 1) {
                const V4 uv = su4[j];
                acc[0] -= md * uv.x; acc[1] -= md * uv.y; acc[2] -= md * uv.z;
            } else {
                acc[0] += md; acc[4] += md; acc[8] += md;
            }
        }
        if (has && half == 0) {
            const uint32_t fl = s_out[j];
            const uint32_t node = fl & NODE_MASK;     // the partial slot of a shared row
            if (fl & FLAG_SHARED) {
                T* pp = partial + (size_t)W * node;
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
            // d.(K d) is linear in the node sums, so every row -- shared rows with
            // their partial sums too -- adds its share here; fixed rows have d = 0
            if (KIND == 0 && dot_part) {
                const V4 pv = sd4[j];
                dot += (double)pv.x * acc[0] + (double)pv.y * acc[1] + (double)pv.z * acc[2];
            }
        }
    }
    if (KIND == 0 && dot_part) {
        double v[1] = {dot};
        block_partials<1, BE>(v, dot_part);
    }
}

// Shared nodes: sum each node's contiguous partial rows (block order).  With
// DOT, each workgroup also folds a slice of the element kernel's p.q
// partials and its own nodes' p.q into dot_out[blockIdx.x].
template <typename T, int W, bool DOT>
__global__ __launch_bounds__(256) void k_shared_gather(SharedArgs s, const T* __restrict__ partial, T* __restrict__ out,
                                                       const double* __restrict__ ext, int mask_fixed,
                                                       const int* __restrict__ done, const float* __restrict__ pvec,
                                                       const double* __restrict__ dot_in, int n_in,
                                                       double* __restrict__ dot_out)
{
    if (done && *done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double dot = 0.0;
    if (i < s.nshared) {
        const uint32_t fl = s.node[i];
        const uint32_t node = fl & NODE_MASK;
        T acc[W];
#pragma unroll
        for (int q = 0; q < W; q++) acc[q] = T(0);
        // up to 4 partial rows in flight per round trip, summed in slot order
        const int k1 = s.ptr[i + 1];
        for (int k = s.ptr[i]; k < k1; k += 4) {
            T t[4][W];
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int q = 0; q < W; q++) t[j][q] = k + j < k1 ? partial[(size_t)W * (k + j) + q] : T(0);
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int q = 0; q < W; q++) acc[q] += t[j][q];
        }
        T* op = out + (size_t)W * node;
        if ((fl & FLAG_FIXED) && mask_fixed) {
#pragma unroll
            for (int q = 0; q < W; q++) op[q] = (W == 9 && (q == 0 || q == 4 || q == 8)) ? T(1) : T(0);
        } else {
#pragma unroll
            for (int q = 0; q < W; q++) op[q] = acc[q] + (ext ? (T)ext[3 * (size_t)node + q % 3] : T(0));
            if (DOT) dot += (double)pvec[3 * (size_t)node] * acc[0] + (double)pvec[3 * (size_t)node + 1] * acc[1] +
                            (double)pvec[3 * (size_t)node + 2] * acc[2];
        }
    }
    if (DOT) {
        const int chunk = (n_in + gridDim.x - 1) / gridDim.x;
        const int b0 = blockIdx.x * chunk;
        const int b1 = b0 + chunk < n_in ? b0 + chunk : n_in;
        for (int k = b0 + threadIdx.x; k < b1; k += blockDim.x) dot += dot_in[k];
        double v[1] = {dot};
        block_partials<1, 256>(v, dot_out);
    }
}

// ---------------------------------------------------------------- CG kernels
// Per iteration two launches: the element kernel (K p and the p.q partials;
// from the second iteration on it first makes the previous iteration's test
// and direction update, beta and p = z + beta p, PFuse) and k_cg_update
// (alpha, x, r, z, summing the shared nodes' partial rows; r.z and r.r
// partials).  k_cg_p makes the test and the direction update on its own:
// after the last launch of a solve, and in the multigrid PCG, where the
// V-cycle runs between the update and the direction.  Each workgroup reduces
// the partials it needs itself in a fixed order: no atomics, no extra launch.
// Scalars that a kernel reads are never written by the same kernel
// (S_RZ_PREV holds the previous r.z).

// x = 0, r = b, z = M r, p = z (and p2 = z); partials of r.z, r.r
template <int PC>
__global__ __launch_bounds__(CG_BLOCK) void k_cg_init(int64_t n3, const float* __restrict__ b, float* __restrict__ x,
                                                 float* __restrict__ r, float* __restrict__ z, float* __restrict__ p,
                                                 float* __restrict__ p2, const float* __restrict__ minv,
                                                 const uint8_t* __restrict__ fixed, double* __restrict__ part)
{
    const int64_t nn = n3 / 3;
    double v[2] = {0.0, 0.0};
    for (int64_t i = xcd_block(blockIdx.x, gridDim.x) * (int64_t)CG_BLOCK + threadIdx.x; i < nn; i += (int64_t)gridDim.x * CG_BLOCK) {
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
        if (p2)
#pragma unroll
            for (int c = 0; c < 3; c++) p2[3 * i + c] = zv[c];
        v[0] += (double)rv[0] * zv[0] + (double)rv[1] * zv[1] + (double)rv[2] * zv[2];
        v[1] += (double)rv[0] * rv[0] + (double)rv[1] * rv[1] + (double)rv[2] * rv[2];
    }
    block_partials<2, CG_BLOCK>(v, part);
}

__global__ __launch_bounds__(256) void k_cg_init_finish(const double* __restrict__ part, int n, double* scal, int* done)
{
    double tot[2];
    sum_partials<2, 256>(part, n, n, tot);
    if (threadIdx.x == 0) {
        scal[S_RZ] = tot[0];
        scal[S_RZ0] = tot[0];
        scal[S_RR] = tot[1];
        scal[S_RR0] = tot[1];
        scal[S_ITER] = 0;
        scal[S_RR_BEST] = tot[1];
        scal[S_IT_BEST] = 0;
        *done = (tot[0] == 0.0) ? 1 : 0;    // IterativeSolvers.h:156-157
    }
}

// alpha = rz / p.q (every workgroup); x += a p; r -= a q; z = M r; partials.
// Each thread first loads its first node's operands, then reduces the p.q
// partials, so the two latencies overlap.
// q of a node: written by the element kernel for nodes of one block, the sum
// of the node's contiguous partial slots for shared nodes, 0 for fixed nodes
// and nodes of no element (node_q: {first slot, end slot} of a shared node;
// {-1, 0} free node of one block, {-2, 0} fixed or of no element)
// PC 2 (multigrid PCG): z = zscale M r into z and z2 -- the V-cycle's first
// Chebyshev step from x = 0 on the new residual (k_mg_cheb<0>'s direction);
// no partials (the V-cycle's last step makes them)
template <int PC>
__global__ __launch_bounds__(CG_BLOCK) void k_cg_update(int64_t n3, float* __restrict__ x, float* __restrict__ r,
                                                   float* __restrict__ z, const float* __restrict__ p,
                                                   const float* __restrict__ q, const float* __restrict__ minv,
                                                   const int2* __restrict__ node_q, const float* __restrict__ partial,
                                                   double* scal, const double* __restrict__ pq_part, int n_pq,
                                                   double* __restrict__ part, int* done, int guard, float zscale,
                                                   float* __restrict__ z2)
{
    // the done flag is read with the node loads and the p.q reduction and
    // tested before anything is stored (a vector load: as a scalar load its
    // miss would hold back the reduction's loads, which wait for the scalar
    // kernel arguments)
    const int stop = vload(done);
    const int64_t nn = n3 / 3;
    const int64_t stride = (int64_t)gridDim.x * CG_BLOCK;
    // each XCD updates a contiguous range of nodes (its workgroups' first rounds)
    int64_t i = xcd_block(blockIdx.x, gridDim.x) * (int64_t)CG_BLOCK + threadIdx.x;
    float pv[3], qv[3], rv[3], xv[3], M[9];
    int2 nq = make_int2(-2, 0);
    auto load_rows = [&](int64_t n) {
        nq = node_q[n];
        // every row is loaded in the first round (q is selected, not loaded
        // conditionally on node_q); the partial rows of a shared node follow
#pragma unroll
        for (int c = 0; c < 3; c++) {
            pv[c] = p[3 * n + c]; rv[c] = r[3 * n + c]; xv[c] = x[3 * n + c];
            qv[c] = q[3 * n + c];
        }
        if (PC) {
#pragma unroll
            for (int k = 0; k < 9; k++) M[k] = minv[9 * n + k];
        }
    };
    auto load_partial_rows = [&]() {
#pragma unroll
        for (int c = 0; c < 3; c++) qv[c] = nq.x == -1 ? qv[c] : 0.f;
        if (nq.x >= 0) {
            // up to 4 partial rows in flight per round trip, summed in slot order;
            // the loads are unconditional (an index past the node's last slot
            // reads that slot again and is not added): a load under a branch is
            // waited for where the branch joins
            const int k1 = nq.y;
            for (int k = nq.x; k < k1; k += 4) {
                float t[4][3];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const size_t kk = (size_t)(k + j < k1 ? k + j : k1 - 1);
#pragma unroll
                    for (int c = 0; c < 3; c++) t[j][c] = partial[3 * kk + c];
                }
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (k + j < k1)
#pragma unroll
                        for (int c = 0; c < 3; c++) qv[c] += t[j][c];
            }
        }
    };
    // first round: this thread's first node rows and the first p.q partials;
    // then the node's partial rows (they need node_q); then the reduction
    double pq_v[1][16], pq[1];
    if (i < nn) load_rows(i);
    partials_issue<1, CG_BLOCK, 16>(pq_part, n_pq, n_pq, pq_v);
    // r.z of the previous kernel: loaded with the first round, not after the reduction
    const double rz = scal[S_RZ];
    if (i < nn) load_partial_rows();
    partials_finish<1, CG_BLOCK, 16>(pq_part, n_pq, n_pq, pq_v, pq);
    if (stop) return;
    if (guard && !(pq[0] > 0.0)) {          // breakdown: not positive definite
        if (blockIdx.x == 0 && threadIdx.x == 0) *done = 3;
        return;
    }
    const double alpha_d = rz / pq[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[S_PQ] = pq[0];
        scal[S_ALPHA] = alpha_d;
        scal[S_RZ_PREV] = rz;
    }
    const float alpha = (float)alpha_d;
    double v[2] = {0.0, 0.0};
    for (; i < nn; i += stride) {
        float zv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            xv[c] += alpha * pv[c];
            rv[c] -= alpha * qv[c];
            x[3 * i + c] = xv[c];
            r[3 * i + c] = rv[c];
        }
        if (PC == 2) {
            float mr[3];
            block_mul(M, rv, mr);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                zv[c] = cheb_dir(false, 0.f, 0.f, zscale, mr[c]);
                z[3 * i + c] = zv[c];
                z2[3 * i + c] = zv[c];
            }
        } else if (PC) {
#pragma unroll
            for (int c = 0; c < 3; c++) zv[c] = M[3 * c] * rv[0] + M[3 * c + 1] * rv[1] + M[3 * c + 2] * rv[2];
#pragma unroll
            for (int c = 0; c < 3; c++) z[3 * i + c] = zv[c];
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++) zv[c] = rv[c];
        }
        if (PC != 2) {
            v[0] += (double)rv[0] * zv[0] + (double)rv[1] * zv[1] + (double)rv[2] * zv[2];
            v[1] += (double)rv[0] * rv[0] + (double)rv[1] * rv[1] + (double)rv[2] * rv[2];
        }
        if (i + stride < nn) {
            load_rows(i + stride);
            load_partial_rows();
        }
    }
    if (PC != 2) block_partials<2, CG_BLOCK>(v, part);
}

// beta = rz' / rz (every workgroup); test; p = z + beta p, one node per step
// test 0: the reference criterion |r.z| <= eps^2 |r0.z0| (IterativeSolvers.h:180-185)
// test 1: true residual ||r||^2 <= tol^2 ||r0||^2; with rz_guard, r.z <= 0
//   is a breakdown (the preconditioner is not positive definite here), and
//   with stall_window > 0 an r.r that has not halved within that many
//   iterations of its last reference value stops the solve (stagnation)
template <int PC>
__global__ __launch_bounds__(CG_P_BLOCK) void k_cg_p(int64_t n3, const float* __restrict__ z, const float* __restrict__ r,
                                              float* __restrict__ p, double* scal, const double* __restrict__ part, int n_part,
                                              int* done, int test_kind, double tol2, int rz_guard, int stall_window,
                                              CgSlot* slot)
{
    const int stop = vload(done);    // tested after the loads and the reduction are issued
    const int64_t nn = n3 / 3;
    const int64_t stride = (int64_t)gridDim.x * CG_P_BLOCK;
    int64_t i = blockIdx.x * (int64_t)CG_P_BLOCK + threadIdx.x;
    const float* zz = PC ? z : r;
    float zv[3], pv[3];
    if (i < nn) {
#pragma unroll
        for (int c = 0; c < 3; c++) { zv[c] = zz[3 * i + c]; pv[c] = p[3 * i + c]; }
    }
    // the scalars of the previous kernels: loaded with the first round
    const double rz0 = scal[S_RZ0], rr0 = scal[S_RR0], rz_prev = scal[S_RZ_PREV], it_prev = scal[S_ITER];
    double t[2];
    sum_partials<2, CG_P_BLOCK>(part, n_part, n_part, t);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (stop) {
        if (slot && lead) {
            slot->rr_rel = 0.0;
            slot->iter = (int)it_prev;
            slot->done = stop;
        }
        return;
    }
    const double rz1 = t[0], rr = t[1];
    const bool conv = test_kind == 0 ? (fabs(rz1) <= tol2 * fabs(rz0)) : (rr <= tol2 * rr0);
    const bool bad = !(rr == rr) || !(rz1 == rz1);
    const bool indef = rz_guard && !conv && !bad && !(rz1 > 0.0);
    const double beta = rz1 / rz_prev;
    if (lead) cg_test_record(scal, done, rz1, rr, beta, conv, bad, indef, stall_window, slot, rr0);
    if (conv || bad || indef) return;
    const float fb = (float)beta;
    for (; i < nn; i += stride) {
#pragma unroll
        for (int c = 0; c < 3; c++) p[3 * i + c] = zv[c] + fb * pv[c];
        if (i + stride < nn) {
#pragma unroll
            for (int c = 0; c < 3; c++) { zv[c] = zz[3 * (i + stride) + c]; pv[c] = p[3 * (i + stride) + c]; }
        }
    }
}

// ---------------------------------------------------------------- FP64 vectors
// a.a, b.b and d.a for FP64 node vectors a, b and an FP32 direction d
// -> scal[slot], scal[slot+1], scal[slot+2]  (b, d may be null)
__global__ __launch_bounds__(256) void k_dots3(int64_t n3, const double* __restrict__ a, const double* __restrict__ b,
                                               const float* __restrict__ d, double* __restrict__ part)
{
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3 / 3; i += (int64_t)gridDim.x * blockDim.x) {
        v[0] += a[3 * i] * a[3 * i] + a[3 * i + 1] * a[3 * i + 1] + a[3 * i + 2] * a[3 * i + 2];
        if (b) v[1] += b[3 * i] * b[3 * i] + b[3 * i + 1] * b[3 * i + 1] + b[3 * i + 2] * b[3 * i + 2];
        if (d) v[2] += (double)d[3 * i] * a[3 * i] + (double)d[3 * i + 1] * a[3 * i + 1] + (double)d[3 * i + 2] * a[3 * i + 2];
    }
    block_partials<3, 256>(v, part);
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

// explicit substep on free nodes: v += h f / m + h g; u += h v; u32 = (float) u
__global__ __launch_bounds__(256) void k_explicit(int64_t nn, double* __restrict__ u, double* __restrict__ vel,
                                                  const double* __restrict__ f, const double* __restrict__ mass,
                                                  const uint8_t* __restrict__ fixed, double gx, double gy, double gz,
                                                  double h, float* __restrict__ u32)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x) {
        if (fixed[i]) continue;
        const double g[3] = {gx, gy, gz};
        const double hm = h / mass[i];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double v = vel[3 * i + c] + hm * f[3 * i + c] + h * g[c];
            vel[3 * i + c] = v;
            const double x = u[3 * i + c] + h * v;
            u[3 * i + c] = x;
            u32[3 * i + c] = (float)x;
        }
    }
}

// first Piola-Kirchhoff stress per element (FP64), written at the element's
// original index: the coefficients hold V mu, V lambda and P is linear in
// (mu, lambda), so P = stress(F, V mu, V lambda) / V, V = 1 / (6 det Dm^-1)
template <int MAT>
__global__ __launch_bounds__(256) void k_elem_stress(BlockArgs a, const double* __restrict__ u,
                                                     const int* __restrict__ eorig, double* __restrict__ out)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= a.ne) return;
    const size_t st = (size_t)a.ne_pad;
    double g[9];
#pragma unroll
    for (int k = 0; k < 9; k++) g[k] = a.dminv64[k * st + s];
    const uint2 li = a.lidx[s];
    const int n0 = a.bnode_ptr[s / BLOCK_E];
    const int l[4] = {(int)(li.x & 0xffffu), (int)(li.x >> 16), (int)(li.y & 0xffffu), (int)(li.y >> 16)};
    double uu[4][3];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t node = a.bnodes[n0 + l[k]] & NODE_MASK;
#pragma unroll
        for (int c = 0; c < 3; c++) uu[k][c] = u[3 * (size_t)node + c];
    }
    double F[9], P[9];
    defgrad(g, uu, F);
    stress<double, MAT>(F, a.coef64[s], a.coef64[st + s], P);
    const double detg = g[0] * (g[4] * g[8] - g[5] * g[7]) - g[1] * (g[3] * g[8] - g[5] * g[6]) +
                        g[2] * (g[3] * g[7] - g[4] * g[6]);
    const double invV = 6.0 * detg;
    double* o = out + 9 * (size_t)eorig[s];
#pragma unroll
    for (int k = 0; k < 9; k++) o[k] = P[k] * invV;
}

// the element's displacement gradient input: Dm^-1 and the 4 corner
// displacements of sorted element s (FP64)
__device__ __forceinline__ void elem_state64(const BlockArgs& a, int64_t s, const double* __restrict__ u, double* g,
                                             double uu[4][3])
{
    const size_t st = (size_t)a.ne_pad;
#pragma unroll
    for (int k = 0; k < 9; k++) g[k] = a.dminv64[k * st + s];
    const uint2 li = a.lidx[s];
    const int n0 = a.bnode_ptr[s / BLOCK_E];
    const int l[4] = {(int)(li.x & 0xffffu), (int)(li.x >> 16), (int)(li.y & 0xffffu), (int)(li.y >> 16)};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t node = a.bnodes[n0 + l[k]] & NODE_MASK;
#pragma unroll
        for (int c = 0; c < 3; c++) uu[k][c] = u[3 * (size_t)node + c];
    }
}

// per-workgroup partial of sum_e V_e Psi(F_e) (FP64; grid-stride over the
// sorted elements, then a fixed tree: deterministic)
template <int MAT>
__global__ __launch_bounds__(256) void k_elem_energy(BlockArgs a, const double* __restrict__ u, double* __restrict__ part)
{
    double v[1] = {0.0};
    const size_t st = (size_t)a.ne_pad;
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < a.ne; s += (int64_t)gridDim.x * blockDim.x) {
        double g[9], uu[4][3], F[9];
        elem_state64(a, s, u, g, uu);
        defgrad(g, uu, F);
        v[0] += energy<double, MAT>(F, a.coef64[s], a.coef64[st + s]);
    }
    block_partials<1, 256>(v, part);
}

// per-workgroup partial of frac sum_free (m g + f_dead) . u
__global__ __launch_bounds__(256) void k_dead_work(int64_t nn, const double* __restrict__ mass, const double* __restrict__ fdead,
                                                   double gx, double gy, double gz, double frac,
                                                   const uint8_t* __restrict__ fixed, const double* __restrict__ u,
                                                   double* __restrict__ part)
{
    double v[1] = {0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x) {
        if (fixed[i]) continue;
        double f[3] = {mass[i] * gx, mass[i] * gy, mass[i] * gz};
        if (fdead) {
#pragma unroll
            for (int c = 0; c < 3; c++) f[c] += fdead[3 * i + c];
        }
        v[0] += frac * (f[0] * u[3 * i] + f[1] * u[3 * i + 1] + f[2] * u[3 * i + 2]);
    }
    block_partials<1, 256>(v, part);
}

// cable node j at x = X0 + u (and its velocity): the barycentric combination
// of its element's corners (FemPhysicsBase.cpp:222-245)
__device__ __forceinline__ void cable_point(const CableArgs& c, int j, const double* __restrict__ X0,
                                            const double* __restrict__ u, const double* __restrict__ vel, double* p, double* v)
{
    const int4 nd = c.node[j];
    const int ids[4] = {nd.x, nd.y, nd.z, nd.w};
#pragma unroll
    for (int q = 0; q < 3; q++) { p[q] = 0.0; v[q] = 0.0; }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double wk = c.w[4 * j + k];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            p[q] += wk * (X0[3 * (size_t)ids[k] + q] + u[3 * (size_t)ids[k] + q]);
            if (vel) v[q] += wk * vel[3 * (size_t)ids[k] + q];
        }
    }
}

// spring j -> j+1 (FemPhysicsBase.cpp:247-268): tension of the Bern model
// (unilateral, smoothed over |err| <= eps = rest / 100) plus damping along it
__global__ __launch_bounds__(256) void k_cable_springs(CableArgs c, const double* __restrict__ X0, const double* __restrict__ u,
                                                       const double* __restrict__ vel, double* __restrict__ fs)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= c.n) return;
    const double L = c.spring[4 * j], act = c.spring[4 * j + 1], k = c.spring[4 * j + 2], damp = c.spring[4 * j + 3];
    double f[3] = {0.0, 0.0, 0.0};
    if (k != 0.0) {
        double pa[3], va[3], pb[3], vb[3];
        cable_point(c, j, X0, u, vel, pa, va);
        cable_point(c, j + 1, X0, u, vel, pb, vb);
        const double y[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
        const double len = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        const double inv = 1.0 / len;
        const double dir[3] = {inv * y[0], inv * y[1], inv * y[2]};
        const double eps = L * 0.01;
        const double err = len - L * act;
        double tension = -2.0 * k * err;
        if (err < -eps) tension = 0.0;
        else if (err >= -eps && err <= eps) tension = -k * (0.5 * err * err / eps + err + 0.5 * eps);
        const double dv = (vb[0] - va[0]) * dir[0] + (vb[1] - va[1]) * dir[1] + (vb[2] - va[2]) * dir[2];
        const double t = tension - damp * dv;
#pragma unroll
        for (int q = 0; q < 3; q++) f[q] = t * dir[q];
    }
#pragma unroll
    for (int q = 0; q < 3; q++) fs[3 * (size_t)j + q] = f[q];
}

// node forces of the cable nodes (f_s[j-1] - f_s[j]) back-interpolated to the
// touched mesh nodes, in the fixed order of each node's attachment list
// (FemPhysicsBase.cpp:270-300); fixed nodes are skipped
__global__ __launch_bounds__(256) void k_cable_gather(CableArgs c, const double* __restrict__ fs, double frac,
                                                      const uint8_t* __restrict__ fixed, double* __restrict__ out,
                                                      int accumulate)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c.ntouched) return;
    const int node = c.touched[t];
    if (fixed[node]) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int k = c.tptr[t]; k < c.tptr[t + 1]; k++) {
        const int2 en = c.tent[k];
        const int j = en.x;
        const double wk = c.w[en.y];
        double nf[3] = {-fs[3 * (size_t)j], -fs[3 * (size_t)j + 1], -fs[3 * (size_t)j + 2]};
        if (j > 0) {
#pragma unroll
            for (int q = 0; q < 3; q++) nf[q] += fs[3 * (size_t)(j - 1) + q];
        }
#pragma unroll
        for (int q = 0; q < 3; q++) acc[q] += wk * (frac * nf[q]);
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
        double* o = out + 3 * (size_t)node + q;
        *o = accumulate ? *o + acc[q] : acc[q];
    }
}

// spring potentials (FemPhysicsMatrixFree.cpp:1093-1108), per-workgroup partials
__global__ __launch_bounds__(256) void k_cable_energy(CableArgs c, const double* __restrict__ X0, const double* __restrict__ u,
                                                      double* __restrict__ part)
{
    double v[1] = {0.0};
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < c.n; j += gridDim.x * blockDim.x) {
        const double L = c.spring[4 * j], act = c.spring[4 * j + 1], k = c.spring[4 * j + 2];
        if (k == 0.0) continue;
        double pa[3], va[3], pb[3], vb[3];
        cable_point(c, j, X0, u, nullptr, pa, va);
        cable_point(c, j + 1, X0, u, nullptr, pb, vb);
        const double y[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
        const double len = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        const double eps = L * 0.01;
        const double err = len - L * act;
        double pot = k * (err * err + eps * eps / 3.0);
        if (err < -eps) pot = 0.0;
        else if (err >= -eps && err <= eps) pot = 0.5 * k * (err * err * err / 3.0 / eps + err * err + eps * err + eps * eps / 3.0);
        v[0] += pot;
    }
    block_partials<1, 256>(v, part);
}

// implicit dynamics: m / h^2 per free node (0 on fixed), FP32 and FP64
__global__ __launch_bounds__(256) void k_mass_diag(int64_t nn, const double* __restrict__ mass, const uint8_t* __restrict__ fixed,
                                                   double invh2, float* __restrict__ m32, double* __restrict__ m64)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const double v = fixed[i] ? 0.0 : mass[i] * invh2;
    m32[i] = (float)v;
    m64[i] = v;
}

// after an implicit solve (Solve, FemPhysicsMatrixFree.cpp:589-597): v = (x - x0) / h on free nodes
__global__ __launch_bounds__(256) void k_implicit_vel(int64_t nn, const double* __restrict__ u, const double* __restrict__ u0,
                                                      const uint8_t* __restrict__ fixed, double invh, double* __restrict__ vel)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nn || fixed[i]) return;
#pragma unroll
    for (int c = 0; c < 3; c++) vel[3 * i + c] = invh * (u[3 * i + c] - u0[3 * i + c]);
}

// load-step predictor: uprev <- u, u <- u + s (u - uprev), u32 <- (float) u
__global__ __launch_bounds__(256) void k_extrapolate(int64_t n3, double* __restrict__ u, double* __restrict__ uprev,
                                                     double sc, float* __restrict__ u32)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = u[i], b = uprev[i];
        const double v = a + sc * (a - b);
        uprev[i] = a;
        u[i] = v;
        u32[i] = (float)v;
    }
}

// u <- a u + b uprev + c uprev2 (the quadratic load-step predictor); the
// history shifts: uprev2 <- uprev, uprev <- u
__global__ __launch_bounds__(256) void k_extrapolate2(int64_t n3, double* __restrict__ u, double* __restrict__ uprev,
                                                      double* __restrict__ uprev2, double a, double b, double c,
                                                      float* __restrict__ u32)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        const double x2 = u[i], x1 = uprev[i], x0 = uprev2[i];
        const double v = a * x2 + b * x1 + c * x0;
        uprev2[i] = x1;
        uprev[i] = x2;
        u[i] = v;
        u32[i] = (float)v;
    }
}

// u = a x + b y (FP64) and its FP32 copy
__global__ __launch_bounds__(256) void k_lincomb(int64_t n3, const double* __restrict__ x, const double* __restrict__ y,
                                                 double a, double b, double* __restrict__ u, float* __restrict__ u32)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = a * x[i] + b * y[i];
        u[i] = v;
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
                                                  const double* __restrict__ vel, const double* __restrict__ u0,
                                                  double invh, double* __restrict__ ext)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nn) return;
    double e[3] = {0.0, 0.0, 0.0};
    if (!fixed[i] && vel) {
        // implicit dynamics (ComputeGradients :553-557): m/h v + m/h^2 u0; the
        // element kernel subtracts m/h^2 u at the row of the node's first block
#pragma unroll
        for (int c = 0; c < 3; c++) e[c] = mass[i] * invh * (vel[3 * i + c] + invh * u0[3 * i + c]);
    }
    if (!fixed[i]) {
        // ComputeGradients (FemPhysicsMatrixFree.cpp:547-552): m frac g
        const double m = mass[i] * frac;
        e[0] += m * gx; e[1] += m * gy; e[2] += m * gz;
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

template <int KIND, int BE, int SG = 0>
static hipError_t launch_elem_mat(int mat, const BlockArgs& a, const void* u, const void* d, void* out, void* partial,
                                  const double* ext, int mask_fixed, const int* done_in, double* dot, size_t lds, hipStream_t s,
                                  hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr, const PFuse& pf = PFuse{},
                                  const MgStage& mg = MgStage{})
{
    typedef typename KindTraits<KIND>::T T;
    dim3 grid(a.nblk), block(BE);
    const T* md = (const T*)(sizeof(T) == 8 ? (const void*)a.mdiag64 : (const void*)a.mdiag32);
    // only PF launches write the flag (the test of the previous iteration)
    int* done = const_cast<int*>(done_in ? done_in : a.zero_flag);
    // the dynamics instances also stage m / h^2 per row (lds: the static layout)
    const size_t dyn_lds = (size_t)a.nmax * sizeof(T);
    // with events: hipExtLaunchKernelGGL records them at the dispatch's own start
    // and end, so their interval is the kernel's execution time
#define SFEM_ELEM_DYN(M, D)                                                                                          \
    if (ev0)                                                                                                         \
        hipExtLaunchKernelGGL((k_elem_block<KIND, M, BE, D, SG>), grid, block, (uint32_t)(lds + (D ? dyn_lds : 0)), \
                              s, ev0, ev1, 0u, a,                                                                    \
                              (const T*)u, (const T*)d, (T*)out, (T*)partial, ext, mask_fixed, done, dot, md, pf, mg); \
    else                                                                                                             \
        hipLaunchKernelGGL((k_elem_block<KIND, M, BE, D, SG>), grid, block, lds + (D ? dyn_lds : 0), s, a,          \
                           (const T*)u, (const T*)d,                                                                  \
                           (T*)out, (T*)partial, ext, mask_fixed, done, dot, md, pf, mg)
    // the multigrid stages have no dynamics instance (static solves only)
#define SFEM_ELEM_CASE(M)                                                                                            \
    if constexpr (SG >= MG_STAGE_CHEB) {                                                                             \
        if (md) return hipErrorInvalidValue;                                                                         \
        SFEM_ELEM_DYN(M, false);                                                                                     \
    } else if (md) {                                                                                                 \
        SFEM_ELEM_DYN(M, true);                                                                                      \
    } else {                                                                                                         \
        SFEM_ELEM_DYN(M, false);                                                                                     \
    }
    switch (mat) {
    case LINEAR: SFEM_ELEM_CASE(LINEAR); break;
    case STVK: SFEM_ELEM_CASE(STVK); break;
    case COROTATIONAL: SFEM_ELEM_CASE(COROTATIONAL); break;
    case COROTATIONAL_POLAR: SFEM_ELEM_CASE(COROTATIONAL_POLAR); break;
    case NEO_HOOKEAN_COARSE:
        // coarse multigrid levels: operator and diagonal blocks only
        if constexpr (KIND != 1) {
            SFEM_ELEM_CASE(NEO_HOOKEAN_COARSE);
        } else {
            SFEM_ELEM_CASE(NEO_HOOKEAN);
        }
        break;
    default:
        // the static Hessian-vector kernels read the cached state when it is set
        if constexpr (KIND == 0) {
            if (a.lin32 && !md) {
                // the cached state replaces the staged displacement rows
                lds = lds_layout_bytes<0>(BE, a.nmax, true, false);
                SFEM_ELEM_DYN(NEO_HOOKEAN_LIN, false);
                break;
            }
        }
        SFEM_ELEM_CASE(NEO_HOOKEAN);
        break;
    }
#undef SFEM_ELEM_CASE
#undef SFEM_ELEM_DYN
    return hipGetLastError();
}

// the static layout (launch_elem_mat adds the dynamics rows); dyn: with them
template <int KIND>
static size_t lds_bytes(const BlockArgs& a, bool dyn = false)
{
    return lds_layout_bytes<KIND>(BLOCK_E, a.nmax, false, dyn);
}

template <typename T, int W>
static hipError_t launch_shared(const SharedArgs& s, const void* partial, void* out, const double* ext, int mask_fixed,
                                const int* done, hipStream_t st)
{
    if (s.nshared == 0) return hipSuccess;
    hipLaunchKernelGGL((k_shared_gather<T, W, false>), dim3((s.nshared + 255) / 256), dim3(256), 0, st, s, (const T*)partial,
                       (T*)out, ext, mask_fixed, done, nullptr, nullptr, 0, nullptr);
    return hipGetLastError();
}

int gather_grid(const SharedArgs& s) { return s.nshared > 0 ? (s.nshared + 255) / 256 : 1; }

hipError_t launch_hessvec(int mat, const BlockArgs& a, const SharedArgs& s, const float* u32, const float* d32, float* out,
                          float* partial, const int* done, hipStream_t st, double* dot_elem, double* dot_out)
{
    hipError_t e = launch_elem_mat<0, BLOCK_E>(mat, a, u32, d32, out, partial, nullptr, 1, done, dot_out ? dot_elem : nullptr,
                                               lds_bytes<0>(a), st);
    if (e != hipSuccess) return e;
    if (!dot_out) return launch_shared<float, 3>(s, partial, out, nullptr, 1, done, st);
    hipLaunchKernelGGL((k_shared_gather<float, 3, true>), dim3(gather_grid(s)), dim3(256), 0, st, s, (const float*)partial, out,
                       nullptr, 1, done, d32, dot_elem, a.nblk, dot_out);
    return hipGetLastError();
}

hipError_t launch_shared_only(const SharedArgs& s, const float* partial, float* out, hipStream_t st)
{
    return launch_shared<float, 3>(s, partial, out, nullptr, 1, nullptr, st);
}

hipError_t launch_force(int mat, const BlockArgs& a, const SharedArgs& s, const double* u64, double* out, double* partial,
                        const double* ext, int mask_fixed, hipStream_t st)
{
    hipError_t e = launch_elem_mat<1, BLOCK_E>(mat, a, u64, nullptr, out, partial, ext, mask_fixed, nullptr, nullptr,
                                               lds_bytes<1>(a), st);
    if (e != hipSuccess) return e;
    return launch_shared<double, 3>(s, partial, out, ext, mask_fixed, nullptr, st);
}

hipError_t launch_blockdiag(int mat, const BlockArgs& a, const SharedArgs& s, const float* u32, float* out, float* partial,
                            hipStream_t st)
{
    hipError_t e = launch_elem_mat<2, BLOCK_E>(mat, a, u32, nullptr, out, partial, nullptr, 1, nullptr, nullptr, lds_bytes<2>(a), st);
    if (e != hipSuccess) return e;
    return launch_shared<float, 9>(s, partial, out, nullptr, 1, nullptr, st);
}

hipError_t launch_lin_state(const BlockArgs& a, const float* u32, float* lin, hipStream_t st)
{
    hipLaunchKernelGGL((k_elem_block<3, NEO_HOOKEAN, BLOCK_E, false, 0>), dim3(a.nblk), dim3(BLOCK_E), lds_bytes<3>(a), st, a,
                       u32, nullptr, lin, nullptr, nullptr, 1, const_cast<int*>(a.zero_flag), nullptr, nullptr, PFuse{},
                       MgStage{});
    return hipGetLastError();
}

size_t elem_lds_max(const BlockArgs& a)
{
    size_t m = lds_bytes<0>(a, true);
    if (lds_bytes<1>(a, true) > m) m = lds_bytes<1>(a, true);
    if (lds_bytes<2>(a, true) > m) m = lds_bytes<2>(a, true);
    return m;
}

hipError_t set_elem_lds_limits(const BlockArgs& a)
{
    // dynamic LDS above the 64 KiB default (unstructured meshes with many nodes
    // per block): raise each kernel's limit to what it needs, within the
    // device's opt-in maximum less the kernel's static LDS
    struct Fn {
        const void* f;
        int kind;
    };
    const Fn fns[] = {
#define SFEM_FN(K, M) {(const void*)k_elem_block<K, M, BLOCK_E, false, 0>, K}, \
                      {(const void*)k_elem_block<K, M, BLOCK_E, true, 0>, K}
#define SFEM_FN_PF(M) {(const void*)k_elem_block<0, M, BLOCK_E, false, 1>, 0}, \
                      {(const void*)k_elem_block<0, M, BLOCK_E, true, 1>, 0}
#define SFEM_FN_MG(M) {(const void*)k_elem_block<0, M, BLOCK_E, false, MG_STAGE_CHEB>, 0}, \
                      {(const void*)k_elem_block<0, M, BLOCK_E, false, MG_STAGE_PROLONG>, 0}
        SFEM_FN(0, LINEAR), SFEM_FN(0, STVK), SFEM_FN(0, COROTATIONAL), SFEM_FN(0, NEO_HOOKEAN),
        SFEM_FN(0, NEO_HOOKEAN_COARSE), SFEM_FN(0, COROTATIONAL_POLAR),
        SFEM_FN_PF(LINEAR), SFEM_FN_PF(STVK), SFEM_FN_PF(COROTATIONAL), SFEM_FN_PF(NEO_HOOKEAN),
        SFEM_FN_PF(NEO_HOOKEAN_COARSE), SFEM_FN_PF(COROTATIONAL_POLAR),
        SFEM_FN_MG(LINEAR), SFEM_FN_MG(STVK), SFEM_FN_MG(COROTATIONAL), SFEM_FN_MG(NEO_HOOKEAN),
        SFEM_FN_MG(NEO_HOOKEAN_COARSE), SFEM_FN_MG(COROTATIONAL_POLAR),
        SFEM_FN(1, LINEAR), SFEM_FN(1, STVK), SFEM_FN(1, COROTATIONAL), SFEM_FN(1, NEO_HOOKEAN),
        SFEM_FN(1, COROTATIONAL_POLAR),
        SFEM_FN(2, LINEAR), SFEM_FN(2, STVK), SFEM_FN(2, COROTATIONAL), SFEM_FN(2, NEO_HOOKEAN),
        SFEM_FN(2, NEO_HOOKEAN_COARSE), SFEM_FN(2, COROTATIONAL_POLAR),
        {(const void*)k_elem_block<0, NEO_HOOKEAN_LIN, BLOCK_E, false, 0>, 0},
        {(const void*)k_elem_block<0, NEO_HOOKEAN_LIN, BLOCK_E, false, 1>, 0},
        {(const void*)k_elem_block<0, NEO_HOOKEAN_LIN, BLOCK_E, false, MG_STAGE_CHEB>, 0},
        {(const void*)k_elem_block<0, NEO_HOOKEAN_LIN, BLOCK_E, false, MG_STAGE_PROLONG>, 0},
        {(const void*)k_elem_block<3, NEO_HOOKEAN, BLOCK_E, false, 0>, 3},
#undef SFEM_FN_MG
#undef SFEM_FN_PF
#undef SFEM_FN
    };
    const size_t need[] = {lds_bytes<0>(a, true), lds_bytes<1>(a, true), lds_bytes<2>(a, true), lds_bytes<3>(a, true)};
    if (elem_lds_max(a) <= 64 * 1024) return hipSuccess;
    int dev = 0, optin = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev);
    if (e != hipSuccess) return e;
    for (const Fn& fn : fns) {
        hipFuncAttributes at;
        e = hipFuncGetAttributes(&at, fn.f);
        if (e != hipSuccess) return e;
        const size_t n = need[fn.kind];
        if (n <= 64 * 1024) continue;
        if (n + at.sharedSizeBytes > (size_t)optin) return hipErrorInvalidValue;   // mesh too irregular for one block
        e = hipFuncSetAttribute(fn.f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)n);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int cg_grid(int64_t nn) { return grid_for(nn, CG_BLOCK, CG_GRID); }

hipError_t launch_cg_init(int pc, int64_t n3, const float* b, float* x, float* r, float* z, float* p, const float* minv,
                          const uint8_t* fixed, double* scal, double* part, int* done, hipStream_t st, float* p2)
{
    const int g = cg_grid(n3 / 3);
    if (pc) hipLaunchKernelGGL(k_cg_init<1>, dim3(g), dim3(CG_BLOCK), 0, st, n3, b, x, r, z, p, p2, minv, fixed, part);
    else hipLaunchKernelGGL(k_cg_init<0>, dim3(g), dim3(CG_BLOCK), 0, st, n3, b, x, r, z, p, p2, minv, fixed, part);
    hipLaunchKernelGGL(k_cg_init_finish, dim3(1), dim3(256), 0, st, part, g, scal, done);
    return hipGetLastError();
}

hipError_t launch_hessvec_cg(int mat, const BlockArgs& a, const float* u32, const float* d32, float* out, float* partial,
                             const int* done, double* dot_elem, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1)
{
    return launch_elem_mat<0, BLOCK_E>(mat, a, u32, d32, out, partial, nullptr, 1, done, dot_elem, lds_bytes<0>(a), st,
                                       ev0, ev1);
}

hipError_t launch_hessvec_cg_fused(int mat, const BlockArgs& a, const float* u32, const PFuse& pf, float* out,
                                   float* partial, int* done, double* dot_elem, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1)
{
    return launch_elem_mat<0, BLOCK_E, 1>(mat, a, u32, nullptr, out, partial, nullptr, 1, done, dot_elem, lds_bytes<0>(a),
                                          st, ev0, ev1, pf);
}

hipError_t launch_hessvec_mg(int mat, const BlockArgs& a, const float* u32, int stage, const MgStage& mg, float* out,
                             float* partial, const int* done, hipStream_t st)
{
    if (stage == MG_STAGE_CHEB)
        return launch_elem_mat<0, BLOCK_E, MG_STAGE_CHEB>(mat, a, u32, nullptr, out, partial, nullptr, 1, done, nullptr,
                                                          lds_bytes<0>(a), st, nullptr, nullptr, PFuse{}, mg);
    if (stage == MG_STAGE_PROLONG)
        return launch_elem_mat<0, BLOCK_E, MG_STAGE_PROLONG>(mat, a, u32, nullptr, out, partial, nullptr, 1, done, nullptr,
                                                             lds_bytes<0>(a), st, nullptr, nullptr, PFuse{}, mg);
    return hipErrorInvalidValue;
}

hipError_t launch_cg_test(int pc, int64_t n3, const float* z, const float* r, float* p, double* scal, const double* part,
                          int* done, int test_kind, double tol2, int rz_guard, int stall_window, hipStream_t st,
                          CgSlot* slot)
{
    const int gp = grid_for(n3 / 3, CG_P_BLOCK, CG_GRID);
    if (pc) hipLaunchKernelGGL(k_cg_p<1>, dim3(gp), dim3(CG_P_BLOCK), 0, st, n3, z, r, p, scal, part, cg_grid(n3 / 3), done,
                               test_kind, tol2, rz_guard, stall_window, slot);
    else hipLaunchKernelGGL(k_cg_p<0>, dim3(gp), dim3(CG_P_BLOCK), 0, st, n3, z, r, p, scal, part, cg_grid(n3 / 3), done,
                            test_kind, tol2, 0, stall_window, slot);
    return hipGetLastError();
}

hipError_t launch_cg_update(int pc, int64_t n3, float* x, float* r, float* z, const float* p, const float* q, const float* minv,
                            const int2* node_q, const float* partial, double* scal, const double* pq_part, int n_pq,
                            double* part, int* done, int guard, hipStream_t st)
{
    const int g = cg_grid(n3 / 3);
    if (pc) hipLaunchKernelGGL(k_cg_update<1>, dim3(g), dim3(CG_BLOCK), 0, st, n3, x, r, z, p, q, minv, node_q, partial, scal, pq_part, n_pq, part, done, guard, 1.f, nullptr);
    else hipLaunchKernelGGL(k_cg_update<0>, dim3(g), dim3(CG_BLOCK), 0, st, n3, x, r, z, p, q, minv, node_q, partial, scal, pq_part, n_pq, part, done, guard, 1.f, nullptr);
    return hipGetLastError();
}

// the two halves of the CG tail for a preconditioner applied between them
// (multigrid): the update without z, then p from z and r.z partials the
// caller wrote in the update's layout (2 x cg_grid)
hipError_t launch_cg_update_only(int64_t n3, float* x, float* r, const float* p, const float* q, const int2* node_q,
                                 const float* partial, double* scal, const double* pq_part, int n_pq, double* part,
                                 int* done, int guard, hipStream_t st)
{
    const int g = cg_grid(n3 / 3);
    hipLaunchKernelGGL(k_cg_update<0>, dim3(g), dim3(CG_BLOCK), 0, st, n3, x, r, nullptr, p, q, nullptr, node_q, partial,
                       scal, pq_part, n_pq, part, done, guard, 1.f, nullptr);
    return hipGetLastError();
}

hipError_t launch_cg_update_mg(int64_t n3, float* x, float* r, float* z, float* z2, const float* p, const float* q,
                               const float* minv, float c2, const int2* node_q, const float* partial, double* scal,
                               const double* pq_part, int n_pq, int* done, int guard, hipStream_t st)
{
    const int g = cg_grid(n3 / 3);
    hipLaunchKernelGGL(k_cg_update<2>, dim3(g), dim3(CG_BLOCK), 0, st, n3, x, r, z, p, q, minv, node_q, partial, scal,
                       pq_part, n_pq, nullptr, done, guard, c2, z2);
    return hipGetLastError();
}

hipError_t launch_cg_p_only(int64_t n3, const float* z, const float* r, float* p, double* scal, const double* part,
                            int* done, int test_kind, double tol2, int rz_guard, int stall_window, hipStream_t st,
                            CgSlot* slot)
{
    const int gp = grid_for(n3 / 3, CG_P_BLOCK, CG_GRID);
    hipLaunchKernelGGL(k_cg_p<1>, dim3(gp), dim3(CG_P_BLOCK), 0, st, n3, z, r, p, scal, part, cg_grid(n3 / 3), done,
                       test_kind, tol2, rz_guard, stall_window, slot);
    return hipGetLastError();
}

// CG scalars from r.z, r.r partials (2 x n): rz, rz0, rr, rr0, iteration 0
hipError_t launch_cg_scalars(const double* part, int n, double* scal, int* done, hipStream_t st)
{
    hipLaunchKernelGGL(k_cg_init_finish, dim3(1), dim3(256), 0, st, part, n, scal, done);
    return hipGetLastError();
}

hipError_t launch_dots3(int64_t n3, const double* a, const double* b, const float* d, double* scal, int slot, double* part,
                        hipStream_t st)
{
    const int g = grid_for(n3 / 3, 256, RED_GRID);
    hipLaunchKernelGGL(k_dots3, dim3(g), dim3(256), 0, st, n3, a, b, d, part);
    hipLaunchKernelGGL(k_finish<3>, dim3(1), dim3(256), 0, st, part, g, scal, slot);
    return hipGetLastError();
}

hipError_t launch_axpy_state(int64_t n3, const double* u, const float* dx, double alpha, double* ut, float* u32, hipStream_t st)
{
    hipLaunchKernelGGL(k_axpy_state, dim3(grid_for(n3, 256, 8192)), dim3(256), 0, st, n3, u, dx, alpha, ut, u32);
    return hipGetLastError();
}

hipError_t launch_explicit(int64_t nn, double* u, double* vel, const double* f, const double* mass, const uint8_t* fixed,
                           const double* g, double h, float* u32, hipStream_t st)
{
    hipLaunchKernelGGL(k_explicit, dim3(grid_for(nn, 256, 8192)), dim3(256), 0, st, nn, u, vel, f, mass, fixed, g[0], g[1],
                       g[2], h, u32);
    return hipGetLastError();
}

static_assert(MAT_NEO_HOOKEAN_COARSE == NEO_HOOKEAN_COARSE && MAT_COROTATIONAL_POLAR == COROTATIONAL_POLAR,
              "host and device material ids");

hipError_t launch_elem_stress(int mat, const BlockArgs& a, const double* u, const int* eorig, double* out, hipStream_t st)
{
    const dim3 grid((unsigned)((a.ne + 255) / 256));
    switch (mat) {
    case LINEAR: hipLaunchKernelGGL(k_elem_stress<LINEAR>, grid, dim3(256), 0, st, a, u, eorig, out); break;
    case STVK: hipLaunchKernelGGL(k_elem_stress<STVK>, grid, dim3(256), 0, st, a, u, eorig, out); break;
    case COROTATIONAL: hipLaunchKernelGGL(k_elem_stress<COROTATIONAL>, grid, dim3(256), 0, st, a, u, eorig, out); break;
    case COROTATIONAL_POLAR:
        hipLaunchKernelGGL(k_elem_stress<COROTATIONAL_POLAR>, grid, dim3(256), 0, st, a, u, eorig, out);
        break;
    default: hipLaunchKernelGGL(k_elem_stress<NEO_HOOKEAN>, grid, dim3(256), 0, st, a, u, eorig, out); break;
    }
    return hipGetLastError();
}

hipError_t launch_energy(int mat, const BlockArgs& a, const double* u, double* scal, int slot, double* part, hipStream_t st)
{
    const int g = grid_for(a.ne, 256, RED_GRID);
    switch (mat) {
    case LINEAR: hipLaunchKernelGGL(k_elem_energy<LINEAR>, dim3(g), dim3(256), 0, st, a, u, part); break;
    case STVK: hipLaunchKernelGGL(k_elem_energy<STVK>, dim3(g), dim3(256), 0, st, a, u, part); break;
    case COROTATIONAL: hipLaunchKernelGGL(k_elem_energy<COROTATIONAL>, dim3(g), dim3(256), 0, st, a, u, part); break;
    case COROTATIONAL_POLAR:
        hipLaunchKernelGGL(k_elem_energy<COROTATIONAL_POLAR>, dim3(g), dim3(256), 0, st, a, u, part);
        break;
    default: hipLaunchKernelGGL(k_elem_energy<NEO_HOOKEAN>, dim3(g), dim3(256), 0, st, a, u, part); break;
    }
    hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(256), 0, st, part, g, scal, slot);
    return hipGetLastError();
}

hipError_t launch_dead_work(int64_t nn, const double* mass, const double* fdead, const double* g, double frac,
                            const uint8_t* fixed, const double* u, double* scal, int slot, double* part, hipStream_t st)
{
    const int gr = grid_for(nn, 256, RED_GRID);
    hipLaunchKernelGGL(k_dead_work, dim3(gr), dim3(256), 0, st, nn, mass, fdead, g[0], g[1], g[2], frac, fixed, u, part);
    hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(256), 0, st, part, gr, scal, slot);
    return hipGetLastError();
}

hipError_t launch_cable_forces(const CableArgs& c, const double* X0, const double* u, const double* vel, double* fs,
                               double frac, const uint8_t* fixed, double* out, int accumulate, hipStream_t st)
{
    if (c.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_cable_springs, dim3((c.n + 255) / 256), dim3(256), 0, st, c, X0, u, vel, fs);
    if (c.ntouched)
        hipLaunchKernelGGL(k_cable_gather, dim3((c.ntouched + 255) / 256), dim3(256), 0, st, c, (const double*)fs, frac,
                           fixed, out, accumulate);
    return hipGetLastError();
}

hipError_t launch_cable_energy(const CableArgs& c, const double* X0, const double* u, double* scal, int slot, double* part,
                               hipStream_t st)
{
    const int g = grid_for(c.n > 0 ? c.n : 1, 256, RED_GRID);
    hipLaunchKernelGGL(k_cable_energy, dim3(g), dim3(256), 0, st, c, X0, u, part);
    hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(256), 0, st, part, g, scal, slot);
    return hipGetLastError();
}

hipError_t launch_mass_diag(int64_t nn, const double* mass, const uint8_t* fixed, double invh2, float* m32, double* m64,
                            hipStream_t st)
{
    hipLaunchKernelGGL(k_mass_diag, dim3((nn + 255) / 256), dim3(256), 0, st, nn, mass, fixed, invh2, m32, m64);
    return hipGetLastError();
}

hipError_t launch_implicit_vel(int64_t nn, const double* u, const double* u0, const uint8_t* fixed, double invh, double* vel,
                               hipStream_t st)
{
    hipLaunchKernelGGL(k_implicit_vel, dim3((nn + 255) / 256), dim3(256), 0, st, nn, u, u0, fixed, invh, vel);
    return hipGetLastError();
}

hipError_t launch_extrapolate(int64_t n3, double* u, double* uprev, double sc, float* u32, hipStream_t st)
{
    hipLaunchKernelGGL(k_extrapolate, dim3(grid_for(n3, 256, 8192)), dim3(256), 0, st, n3, u, uprev, sc, u32);
    return hipGetLastError();
}

hipError_t launch_lincomb(int64_t n3, const double* x, const double* y, double a, double b, double* u, float* u32,
                          hipStream_t st)
{
    hipLaunchKernelGGL(k_lincomb, dim3(grid_for(n3, 256, 8192)), dim3(256), 0, st, n3, x, y, a, b, u, u32);
    return hipGetLastError();
}

hipError_t launch_extrapolate2(int64_t n3, double* u, double* uprev, double* uprev2, double a, double b, double c, float* u32,
                               hipStream_t st)
{
    hipLaunchKernelGGL(k_extrapolate2, dim3(grid_for(n3, 256, 8192)), dim3(256), 0, st, n3, u, uprev, uprev2, a, b, c, u32);
    return hipGetLastError();
}

hipError_t launch_to_f32(int64_t n, const double* a, float* b, hipStream_t st)
{
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, n, a, b);
    return hipGetLastError();
}

hipError_t launch_external(int64_t nn, const double* mass, const double* fdead, const double* g, double frac, const int* pptr,
                           const int* pidx, const double* ftri, const uint8_t* fixed, const double* vel, const double* u0,
                           double invh, double* ext, hipStream_t st)
{
    hipLaunchKernelGGL(k_external, dim3((nn + 255) / 256), dim3(256), 0, st, nn, mass, fdead, g[0], g[1], g[2], frac, pptr, pidx,
                       ftri, fixed, vel, u0, invh, ext);
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
