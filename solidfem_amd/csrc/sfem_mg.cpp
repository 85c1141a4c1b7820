// sfem_mg.cpp -- the multigrid preconditioner of the converged Newton-PCG:
// hierarchy set-up (sfem_mg_add_level), the level operators at each Newton
// iterate, the V-cycle, and the PCG loop that applies it.
//
// Each coarse level is a full context of its own mesh (partition, element
// kernels, block preconditioner) whose state is the level above's
// displacement at the coinciding nodes, so every level's operator is the
// matrix-free tangent of its own mesh at (an injection of) the current state
// -- a rediscretisation, not a Galerkin product.  Smoothing is a Chebyshev
// polynomial in M A (M the level's inverse 3x3 diagonal blocks, or the
// learned blocks on the finest level) on [lmax / ratio, lmax], with lmax from
// a power iteration at each Newton iterate.  Pre- and post-smoothing use the
// same polynomial, so with a symmetric coarsest solve (Chebyshev again) the
// V-cycle is a symmetric operator, as PCG needs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "sfem_capi.h"
#include "sfem_internal.hpp"
#include "sfem_kernels.hpp"
#include "sfem_mg.hpp"

using namespace sfem;

static_assert(MG_DENSE_MAX == SFEM_MG_DENSE_MAX, "dense coarsest limit of the C-ABI");

MgLevel::~MgLevel() { delete c; }

namespace {

constexpr int MG_DEGREE = 2;            // Chebyshev degree of pre- and post-smoothing
constexpr int MG_COARSE_DEGREE = 16;    // Chebyshev degree of the coarsest-level solve
constexpr double MG_RATIO = 10.0;       // smoothed band [lmax / ratio, lmax]
constexpr double MG_COARSE_RATIO = 100.0;
constexpr double MG_SAFETY = 1.15;      // the power iteration estimates lmax from below
constexpr int MG_POWER_ITERS = 12;
constexpr int MG_POWER_ITERS_WARM = 4;  // from the previous full set-up's last vector
// a multigrid-PCG solve stops (and the Newton driver redoes it with block
// Jacobi) when it breaks down (p.Kp <= 0 or r.z <= 0: the V-cycle is not
// positive definite at this state), when r.r has not halved within
// MG_STALL_WINDOW iterations, or after MG_MAX_ITERS iterations.  Where the
// V-cycle works it needs a few to a few tens of iterations per Newton step
// and reduces r.r by ~10x per iteration.
constexpr int MG_MAX_ITERS = 500;
constexpr int MG_STALL_WINDOW = 150;

struct LevelView {
    sfem_ctx* c;           // mesh, partition, state of the level
    const float* minv;     // smoother blocks
    const float* b;        // right-hand side
    float *x, *r, *d, *q;  // solution, residual, Chebyshev direction, operator image
    double lmax;
    // the folded schedule: partial slots of q (pa), second residual and
    // direction, the image of one more apply (qb rows, pb partial slots)
    float *pa, *r2, *d2, *qb, *pb;
};

LevelView level(sfem_ctx* f, int l)
{
    LevelView L;
    if (l == 0) {
        L.c = f;
        L.minv = f->mg_minv0;
        L.b = f->d_r32.as<float>();
        L.x = f->d_z32.as<float>();
        L.r = f->d_mg_r.as<float>();
        L.d = f->d_mg_d.as<float>();
        L.q = f->d_mg_q.as<float>();
        L.lmax = f->mg_lmax0;
    } else {
        MgLevel& m = *f->mg[l - 1];
        sfem_ctx* c = m.c;
        L.c = c;
        L.minv = c->d_minv.as<float>();
        L.b = c->d_b32.as<float>();
        L.x = c->d_x32.as<float>();
        L.r = c->d_r32.as<float>();
        L.d = c->d_p32.as<float>();
        L.q = c->d_q32.as<float>();
        L.lmax = m.lmax;
    }
    sfem_ctx* c = L.c;
    L.pa = c->d_partial.as<float>();
    L.r2 = c->d_mgw_r2.as<float>();
    L.d2 = c->d_mgw_d2.as<float>();
    L.qb = c->d_mgw_qb.as<float>();
    L.pb = c->d_mgw_pb.as<float>();
    return L;
}

// the folded schedule's work vectors of one level context
void alloc_fused(sfem_ctx* c)
{
    if (c->d_mgw_r2.p) return;
    const size_t n3 = 3 * (size_t)c->nn * sizeof(float), ns = 3 * (size_t)c->nslots * sizeof(float);
    for (DevBuf* b : {&c->d_mgw_r2, &c->d_mgw_d2, &c->d_mgw_qb}) b->alloc(n3);
    c->d_mgw_pb.alloc(ns);
}

// out = K_l v (element rows and partial rows; operator_row assembles them);
// done: the PCG's flag, or a flag that stays 0
void matvec(const LevelView& L, const float* v, float* out, const int* done)
{
    sfem_ctx* c = L.c;
    HIPCK(launch_hessvec_cg(c->material, c->blk, c->d_u32.as<float>(), v, out, c->d_partial.as<float>(), done, nullptr,
                            c->stream, nullptr, nullptr));
}

struct Cheb {
    double theta, delta, sigma;
};

Cheb cheb_band(double lmax, double ratio)
{
    const double hi = MG_SAFETY * lmax, lo = hi / ratio;
    return Cheb{0.5 * (hi + lo), 0.5 * (hi - lo), (hi + lo) / (hi - lo)};
}

// Chebyshev smoothing of K x = b (Saad, Iterative Methods for Sparse Linear
// Systems, 2nd ed., Alg. 12.1, with the preconditioner M): from x = 0, or from
// the current x when r already holds b - K x_old and q = K e for x = x_old + e.
// first_done: the first step from x = 0 was made by the restriction that
// produced b.
void smooth(const LevelView& L, int degree, bool from_zero, const Cheb& ch, const int* done, bool first_done = false)
{
    sfem_ctx* c = L.c;
    const int64_t nn = c->nn;
    const int2* nq = c->d_node_info.as<int2>();
    const float* part = c->d_partial.as<float>();
    if (!first_done)
        HIPCK(launch_mg_cheb(from_zero ? 0 : 2, nn, L.b, L.r, L.d, L.x, L.q, nq, part, L.minv, 0.f,
                             (float)(1.0 / ch.theta), done, c->stream));
    double rho = 1.0 / ch.sigma;
    for (int i = 1; i < degree; i++) {
        matvec(L, L.d, L.q, done);
        const double rho1 = 1.0 / (2.0 * ch.sigma - rho);
        HIPCK(launch_mg_cheb(1, nn, L.b, L.r, L.d, L.x, L.q, nq, part, L.minv, (float)(rho1 * rho),
                             (float)(2.0 * rho1 / ch.delta), done, c->stream));
        rho = rho1;
    }
}

// the Chebyshev band a level smooths with
Cheb level_band(sfem_ctx* f, int l, double lmax)
{
    return cheb_band(lmax, l == (int)f->mg.size() ? MG_COARSE_RATIO : MG_RATIO);
}

// One V-cycle from level l, one kernel per step.  first_done: the restriction
// into this level made its first smoothing step.  done: the PCG's flag, or a
// flag that stays 0.
void vcycle(sfem_ctx* f, int l, const int* done, bool first_done = false)
{
    const LevelView L = level(f, l);
    sfem_ctx* c = L.c;
    const int nlev = (int)f->mg.size();
    if (l == nlev) {
        const MgLevel* m = l > 0 ? f->mg[l - 1].get() : nullptr;
        if (m && m->dense) HIPCK(launch_mg_dense_solve(3 * (int)c->nn, m->d_Ainv.as<float>(), L.b, L.x, done, c->stream));
        else smooth(L, MG_COARSE_DEGREE, true, level_band(f, l, L.lmax), done, first_done);
        return;
    }
    const Cheb ch = level_band(f, l, L.lmax);
    smooth(L, MG_DEGREE, true, ch, done, first_done);
    matvec(L, L.d, L.q, done);
    HIPCK(launch_mg_cheb(3, c->nn, L.b, L.r, L.d, L.x, L.q, c->d_node_info.as<int2>(), c->d_partial.as<float>(), L.minv,
                         0.f, 0.f, done, c->stream));
    MgLevel& m = *f->mg[l];
    const LevelView C = level(f, l + 1);
    // a smoothed coarse level's first step is made by the restriction
    const bool fuse = !(l + 1 == nlev && m.dense);
    const float c2 = fuse ? (float)(1.0 / level_band(f, l + 1, C.lmax).theta) : 0.f;
    HIPCK(launch_mg_restrict(C.c->nn, m.t_ptr.as<int>(), m.t_row.as<int>(), m.t_val.as<float>(), L.r,
                             C.c->d_fixed.as<uint8_t>(), const_cast<float*>(C.b), done, c->stream, fuse ? C.minv : nullptr,
                             C.r, C.d, C.x, c2));
    vcycle(f, l + 1, done, fuse);
    HIPCK(launch_mg_prolong(c->nn, m.p_ptr.as<int>(), m.p_col.as<int>(), m.p_val.as<float>(), C.x,
                            c->d_fixed.as<uint8_t>(), L.d, L.x, done, c->stream));
    matvec(L, L.d, L.q, done);
    smooth(L, MG_DEGREE, false, ch, done);
}

// The same V-cycle with steps folded into neighbouring kernels, bitwise the
// same result (every node value is the same FP32 expression):
//  - on every level the prolongation into it is made by the staging of the
//    level's next operator apply (MG_STAGE_PROLONG: e = P x_c, x += e, K e);
//  - on the coarse levels, where each kernel costs about one workgroup's
//    latency chain, both Chebyshev steps that follow an apply are made by the
//    staging of the next apply (MG_STAGE_CHEB); on the finest level, whose
//    element kernel is VALU-bound, the extra staging work costs more than the
//    launch it saves, and they stay separate kernels;
//  - the finest level's first step is made by the PCG update
//    (launch_cg_update_mg), and its last step also writes the PCG's r.z, r.r
//    partials (rz_part).
// A smoothed level (degree 2; q_k: the image of the k-th apply in rows and
// partial slots), on the coarse levels:
//   [first step: by the restriction into this level]
//   q1 = K d                                              matvec
//   r' = b - q1, d' = c1 d + c2 M r', x += d'; q2 = K d'   CHEB stage
//   r' -= q2                                              k_mg_cheb<3>
//   b_c = P^T r'                                          restriction (+ the coarse first step)
//   e = P x_c, x += e; q3 = K e                           PROLONG stage
//   r'' = r' - q3, d = c2 M r'', x += d; q4 = K d          CHEB stage
//   r = r'' - q4, d = c1 d + c2 M r, x += d               k_mg_cheb<1>
void vcycle_fused(sfem_ctx* f, int l, const int* done, bool first_done, double* rz_part)
{
    static_assert(MG_DEGREE == 2, "the folded schedule is written for degree-2 smoothing");
    const LevelView L = level(f, l);
    sfem_ctx* c = L.c;
    const int nlev = (int)f->mg.size();
    const int2* nq = c->d_node_info.as<int2>();
    if (l == nlev) {
        const MgLevel* m = l > 0 ? f->mg[l - 1].get() : nullptr;
        if (m && m->dense) HIPCK(launch_mg_dense_solve(3 * (int)c->nn, m->d_Ainv.as<float>(), L.b, L.x, done, c->stream));
        else smooth(L, MG_COARSE_DEGREE, true, level_band(f, l, L.lmax), done, first_done);
        return;
    }
    const bool fold = l > 0;
    const Cheb ch = level_band(f, l, L.lmax);
    // the coefficients of smooth(): the first step c2 = 1 / theta, the second from rho
    const double rho = 1.0 / ch.sigma, rho1 = 1.0 / (2.0 * ch.sigma - rho);
    const float c1 = (float)(rho1 * rho), c2 = (float)(2.0 * rho1 / ch.delta), c2first = (float)(1.0 / ch.theta);
    const float* u32 = c->d_u32.as<float>();
    if (!first_done)
        HIPCK(launch_mg_cheb(0, c->nn, L.b, L.r, L.d, L.x, L.q, nq, L.pa, L.minv, 0.f, c2first, done, c->stream));
    matvec(L, L.d, L.q, done);
    // q2: the image of the pre-smoothing's second direction
    const float* q2 = L.q;
    const float* p2 = L.pa;
    if (fold) {
        MgStage st{};
        st.r_old = L.b;
        st.r_new = L.r;
        st.d_old = L.d;
        st.d_new = L.d2;
        st.x = L.x;
        st.minv = L.minv;
        st.qa = L.q;
        st.pa = L.pa;
        st.has_dold = 1;
        st.c1 = c1;
        st.c2 = c2;
        HIPCK(launch_hessvec_mg(c->material, c->blk, u32, MG_STAGE_CHEB, st, L.qb, L.pb, done, c->stream));
        q2 = L.qb;
        p2 = L.pb;
    } else {
        // r from b: the first step may have been the PCG update's, which does
        // not copy b to r
        HIPCK(launch_mg_cheb(4, c->nn, L.b, L.r, L.d, L.x, L.q, nq, L.pa, L.minv, c1, c2, done, c->stream));
        matvec(L, L.d, L.q, done);
    }
    HIPCK(launch_mg_cheb(3, c->nn, L.b, L.r, L.d, L.x, q2, nq, p2, L.minv, 0.f, 0.f, done, c->stream));
    MgLevel& m = *f->mg[l];
    const LevelView C = level(f, l + 1);
    const bool cfirst = !(l + 1 == nlev && m.dense);
    const float cc2 = cfirst ? (float)(1.0 / level_band(f, l + 1, C.lmax).theta) : 0.f;
    HIPCK(launch_mg_restrict(C.c->nn, m.t_ptr.as<int>(), m.t_row.as<int>(), m.t_val.as<float>(), L.r,
                             C.c->d_fixed.as<uint8_t>(), const_cast<float*>(C.b), done, c->stream,
                             cfirst ? C.minv : nullptr, C.r, C.d, C.x, cc2));
    vcycle_fused(f, l + 1, done, cfirst, nullptr);
    MgStage pr{};
    pr.x = L.x;
    pr.p_ptr = m.p_ptr.as<int>();
    pr.p_col = m.p_col.as<int>();
    pr.p_val = m.p_val.as<float>();
    pr.xc = C.x;
    HIPCK(launch_hessvec_mg(c->material, c->blk, u32, MG_STAGE_PROLONG, pr, L.q, L.pa, done, c->stream));
    float* rl = L.r;
    float* q4 = L.q;
    float* p4 = L.pa;
    if (fold) {
        MgStage s1{};
        s1.r_old = L.r;
        s1.r_new = L.r2;
        s1.d_old = L.d2;   // not read (a first step); no block writes it here
        s1.d_new = L.d;
        s1.x = L.x;
        s1.minv = L.minv;
        s1.qa = L.q;
        s1.pa = L.pa;
        s1.has_dold = 0;
        s1.c1 = 0.f;
        s1.c2 = c2first;
        HIPCK(launch_hessvec_mg(c->material, c->blk, u32, MG_STAGE_CHEB, s1, L.qb, L.pb, done, c->stream));
        rl = L.r2;
        q4 = L.qb;
        p4 = L.pb;
    } else {
        HIPCK(launch_mg_cheb(2, c->nn, L.b, L.r, L.d, L.x, L.q, nq, L.pa, L.minv, 0.f, c2first, done, c->stream));
        matvec(L, L.d, L.q, done);
    }
    if (rz_part)
        HIPCK(launch_mg_cheb_rz(c->nn, L.b, rl, L.d, L.x, q4, nq, p4, L.minv, c1, c2, rz_part, done, c->stream));
    else
        HIPCK(launch_mg_cheb(1, c->nn, L.b, rl, L.d, L.x, q4, nq, p4, L.minv, c1, c2, done, c->stream));
}

// element evaluations of one V-cycle (matvecs per level x its elements)
int64_t vcycle_evals(sfem_ctx* f)
{
    const int nlev = (int)f->mg.size();
    int64_t n = 0;
    for (int l = 0; l <= nlev; l++) {
        const LevelView L = level(f, l);
        const bool dense = l == nlev && l > 0 && f->mg[l - 1]->dense;
        const int mv = l == nlev ? (dense ? 0 : MG_COARSE_DEGREE - 1) : 2 * MG_DEGREE;
        n += (int64_t)mv * L.c->ne;
    }
    return n;
}

}  // namespace

namespace sfem {

bool mg_prepare(sfem_ctx* f, const float* minv0, bool full, bool warm_start)
{
    hipStream_t s = f->stream;
    const size_t n3 = 3 * (size_t)f->nn;
    if (!f->d_mg_r.p) {
        f->d_mg_r.alloc(n3 * sizeof(float));
        f->d_mg_d.alloc(n3 * sizeof(float));
        f->d_mg_q.alloc(n3 * sizeof(float));
        f->d_mg_part.alloc(1024 * sizeof(double));
    }
    const int nlev = (int)f->mg.size();
    if (f->mg_fused) {
        alloc_fused(f);
        for (auto& m : f->mg) alloc_fused(m->c);
    }
    if (!f->d_mg_norm.p || f->d_mg_norm.n < 2 * (size_t)(nlev + 1) * sizeof(double))
        f->d_mg_norm.alloc(2 * (size_t)(nlev + 1) * sizeof(double));
    f->mg_minv0 = minv0;
    full = full || !f->mg_ready;
    // coarse states, finest to coarsest; then the coarsest operator's inverse
    // on the side stream; then the other levels' block preconditioners
    const double* ua = f->d_u64.as<double>();
    const bool dense = nlev > 0 && f->mg[nlev - 1]->dense;
    bool side = false;
    for (int l = 0; l < nlev; l++) {
        MgLevel& m = *f->mg[l];
        sfem_ctx* c = m.c;
        HIPCK(launch_mg_inject(c->nn, m.inj.as<int>(), ua, c->d_u64.as<double>(), c->d_u32.as<float>(), s));
        ua = c->d_u64.as<double>();
    }
    if (dense && full) {
        // the coarsest operator, assembled and inverted on the side stream
        // once its state is injected; the main stream goes on with the other
        // levels' block preconditioners and estimates and waits for it before
        // the norms are read back
        MgLevel& m = *f->mg[nlev - 1];
        sfem_ctx* c = m.c;
        if (!f->mg_side) {
            HIPCK(hipStreamCreateWithFlags(&f->mg_side, hipStreamNonBlocking));
            for (hipEvent_t& e : f->mg_ev) HIPCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        hipStream_t ss = f->mg_side;
        HIPCK(hipEventRecord(f->mg_ev[0], s));
        HIPCK(hipStreamWaitEvent(ss, f->mg_ev[0], 0));
        const int n = 3 * (int)c->nn;
        HIPCK(launch_mg_elem_matrix(c->material, c->ne, m.d_tets4.as<int4>(), c->d_u32.as<float>(), m.d_dminv.as<float>(),
                                    m.d_cvol.as<float>(), m.d_Ke.as<float>(), ss));
        HIPCK(launch_mg_dense_assemble((int)c->nn, m.d_inc_ptr.as<int>(), m.d_inc.as<int2>(), m.d_tets4.as<int4>(),
                                       m.d_Ke.as<float>(), c->d_fixed.as<uint8_t>(), m.d_A.as<float>(), ss));
        HIPCK(launch_mg_dense_inverse(n, m.d_A.as<float>(), m.d_Ainv.as<float>(), m.d_W.as<double>(),
                                      m.d_C.as<double>(), m.d_S.as<double>(), ss));
        HIPCK(hipEventRecord(f->mg_ev[1], ss));
        side = true;
    }
    for (int l = 0; l < nlev; l++)
        if (!(l == nlev - 1 && dense)) build_jacobi(f->mg[l]->c);
    // lmax(M A) per level: K unnormalised power steps (the growth of K steps
    // stays far inside FP32 range for lmax of a few units), lmax ~
    // |v_K| / |v_{K-1}|: 12 from a hashed start, or, with warm_start (the
    // Newton solves), 4 from the previous warm set-up's last vector scaled to
    // unit norm (the tangents of consecutive load steps differ little, so it
    // is already close to the top eigenvector; sfem_reset and changes of the
    // hierarchy forget it)
    f->mg_evals = vcycle_evals(f);
    if (!full) return true;
    double* norms = f->d_mg_norm.as<double>();
    const int nest = dense ? nlev : nlev + 1;   // levels that smooth
    for (int l = 0; l < nest; l++) {
        const LevelView L = level(f, l);
        sfem_ctx* c = L.c;
        const int64_t n3c = 3 * c->nn;
        float* v = L.d;
        float* y = L.r;
        const bool warm = warm_start && c->mg_eig_ok;
        if (warm) HIPCK(launch_mg_scale(n3c, c->d_mg_eig.as<float>(), (float)(1.0 / c->mg_eig_norm), v, s));
        else HIPCK(launch_mg_start(c->nn, c->d_fixed.as<uint8_t>(), v, s));
        for (int it = 0; it < (warm ? MG_POWER_ITERS_WARM : MG_POWER_ITERS); it++) {
            matvec(L, v, L.q, c->blk.zero_flag);
            HIPCK(launch_mg_power(c->nn, L.q, c->d_node_info.as<int2>(), c->d_partial.as<float>(), L.minv, y, s));
            std::swap(v, y);
        }
        HIPCK(launch_mg_sqnorm(n3c, v, f->d_mg_part.as<double>(), norms + 2 * l, s));
        HIPCK(launch_mg_sqnorm(n3c, y, f->d_mg_part.as<double>(), norms + 2 * l + 1, s));
        if (warm_start) {
            if (c->d_mg_eig.n < (size_t)n3c * sizeof(float)) c->d_mg_eig.alloc((size_t)n3c * sizeof(float));
            HIPCK(launch_mg_copy(n3c, v, c->d_mg_eig.as<float>(), s));
        }
    }
    if (side) HIPCK(hipStreamWaitEvent(s, f->mg_ev[1], 0));
    std::vector<double> h(2 * nest);
    HIPCK(hipMemcpyAsync(h.data(), norms, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    // a level whose operator is not finite at this state (an element inverted
    // by the injected coarse displacement) gives no estimate: no V-cycle here
    std::vector<double> lm(nest);
    for (int l = 0; l < nest; l++) {
        lm[l] = h[2 * l + 1] > 0 ? std::sqrt(h[2 * l] / h[2 * l + 1]) : 0.0;
        if (!(lm[l] > 0) || !std::isfinite(lm[l])) {
            f->mg_ready = false;
            for (int j = 0; j < nest; j++) level(f, j).c->mg_eig_ok = false;
            return false;
        }
    }
    for (int l = 0; l < nest; l++) {
        if (l == 0) f->mg_lmax0 = lm[l];
        else f->mg[l - 1]->lmax = lm[l];
        if (warm_start) {
            sfem_ctx* c = level(f, l).c;
            c->mg_eig_norm = std::sqrt(h[2 * l]);
            c->mg_eig_ok = c->mg_eig_norm > 0 && std::isfinite(c->mg_eig_norm);
        }
    }
    f->mg_ready = true;
    return true;
}

void mg_vcycle(sfem_ctx* f)
{
    if (f->mg_fused) vcycle_fused(f, 0, f->blk.zero_flag, false, nullptr);
    else vcycle(f, 0, f->blk.zero_flag);
}

// PCG with one V-cycle per iteration: the element kernel (q = K p, p.q
// partials), the CG update without z, z = V(r), r.z partials, then beta,
// the convergence test and p.  The host keeps one iteration enqueued beyond
// the done flag it waits for (a copy of the flag into pinned memory and an
// event after every iteration), so the device never waits for the host; the
// iteration enqueued past the end of a solve finds the flag set and every
// kernel of it returns at once.
CgResult mg_pcg(sfem_ctx* c, int maxit, double tol2, int64_t* hv_evals)
{
    maxit = std::min(maxit, MG_MAX_ITERS);
    const int64_t n3 = 3 * c->nn;
    hipStream_t s = c->stream;
    double* scal = c->d_scal.as<double>();
    double* parts = c->d_partials.as<double>();
    double* pq_elem = parts + 2 * CG_GRID;
    int* done = c->d_done.as<int>();
    const int* zero = c->blk.zero_flag;
    float* x = c->d_x32.as<float>();
    float* r = c->d_r32.as<float>();
    float* z = c->d_z32.as<float>();
    float* p = c->d_p32.as<float>();
    float* q = c->d_q32.as<float>();
    ensure_cg_slots(c);
    HIPCK(launch_cg_init(0, n3, c->d_b32.as<float>(), x, r, z, p, c->d_minv.as<float>(), c->d_fixed.as<uint8_t>(), scal,
                         parts, done, s));
    // the folded schedule: the update makes the V-cycle's first step and the
    // V-cycle's last step the r.z partials.  The first V-cycle runs whatever
    // the flag says (its partials set the flag again).
    const bool fused = c->mg_fused;
    const float c2first = (float)(1.0 / level_band(c, 0, c->mg_lmax0).theta);
    if (fused) {
        vcycle_fused(c, 0, zero, false, parts);
    } else {
        vcycle(c, 0, zero);
        HIPCK(launch_mg_rz(c->nn, r, z, parts, zero, s));
    }
    HIPCK(launch_cg_scalars(parts, cg_grid(c->nn), scal, done, s));
    HIPCK(launch_mg_copy(n3, z, p, s));
    int dflag = 0, launched = 0;   // iterations enqueued
    auto iteration = [&]() {
        HIPCK(launch_hessvec_cg(c->material, c->blk, c->d_u32.as<float>(), p, q, c->d_partial.as<float>(), done, pq_elem, s,
                                nullptr, nullptr));
        if (fused) {
            HIPCK(launch_cg_update_mg(n3, x, r, z, c->d_mg_d.as<float>(), p, q, c->mg_minv0, c2first,
                                      c->d_node_info.as<int2>(), c->d_partial.as<float>(), scal, pq_elem, c->nblk, done, 1,
                                      s));
            vcycle_fused(c, 0, done, true, parts);
        } else {
            HIPCK(launch_cg_update_only(n3, x, r, p, q, c->d_node_info.as<int2>(), c->d_partial.as<float>(), scal, pq_elem,
                                        c->nblk, parts, done, 1, s));
            vcycle(c, 0, done);
            HIPCK(launch_mg_rz(c->nn, r, z, parts, done, s));
        }
        // the test is on the true residual norm (r.r); a V-cycle that is not
        // positive definite at this state (tangents of near-incompressible
        // material at strongly deformed Newton iterates) shows as r.z <= 0 or
        // stagnation, and the Newton driver falls back to block Jacobi
        HIPCK(launch_cg_p_only(n3, z, r, p, scal, parts, done, 1, tol2, 1, MG_STALL_WINDOW, s,
                               c->d_slot + ((launched + 1) & 1)));
    };
    // The host keeps one iteration enqueued beyond the one whose outcome it
    // waits for (the device never waits for the host), except where the last
    // two outcomes predict that the iteration in flight converges: then it
    // waits for that one first, which saves the iteration past the end of the
    // solve (~70 us of early returns) when the prediction holds and costs one
    // host round trip when it does not.  Outcome k (after k iterations) is in
    // slot k & 1: the iteration's direction kernel writes it, or, for k = 0, a
    // copy of the flag; slot k & 1 is reused by outcome k + 2 only after the
    // host has read outcome k.
    auto record = [&](int k) { HIPCK(hipEventRecord(c->flag_ev[k & 1], s)); };
    // outcome 0: the flag and 0 iterations (the flag word after it, which
    // stays 0), copied in stream order (an iteration of the previous solve may
    // still write this slot)
    HIPCK(hipMemcpyAsync(&c->h_slot[0], done, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    record(0);
    auto launch = [&]() {
        iteration();
        launched++;
        record(launched);
    };
    if (maxit > 0) launch();
    double rr_prev = 0.0;
    int iters = 0;
    for (int k = 0;; k++) {
        HIPCK(hipEventSynchronize(c->flag_ev[k & 1]));
        dflag = c->h_slot[k & 1].done;
        iters = c->h_slot[k & 1].iter;
        const double rr = k > 0 ? c->h_slot[k & 1].rr_rel : 0.0;
        if (dflag) break;
        if (launched == k) {    // nothing in flight: the prediction held the next one back, or maxit
            if (launched >= maxit) break;
            launch();
        }
        const bool ends = k > 1 && rr > 0.0 && rr_prev > 0.0 && rr * std::min(1.0, rr / rr_prev) <= tol2;
        if (launched == k + 1 && launched < maxit && !ends) launch();
        rr_prev = rr;
    }
    // no read-back: the outcomes carry the status and the iteration count, and
    // what the caller launches next queues behind an iteration still in flight
    CgResult res;
    res.iters = iters;
    res.status = dflag;
    res.rr0 = res.rr = 0.0;
    if (hv_evals) *hv_evals += (int64_t)res.iters * (c->ne + c->mg_evals) + c->mg_evals;
    return res;
}

}  // namespace sfem

namespace {

// the element data of a small level for its direct solve (dense_level_data)
void dense_setup(MgLevel& m, const sfem_ctx* f, const sfem_mesh* mesh, const double* young_e, const double* poisson_e)
{
    const int64_t ne = mesh->ne, nn = mesh->nn;
    const DenseLevelData d = dense_level_data(mesh->tets, ne, mesh->X0, nn, f->young, f->poisson, young_e, poisson_e);
    m.d_tets4.upload(d.tets4);
    m.d_dminv.upload(d.dminv);
    m.d_cvol.upload(d.cvol);
    m.d_inc_ptr.upload(d.inc_ptr);
    m.d_inc.upload(d.inc);
    m.d_Ke.alloc(144 * (size_t)ne * sizeof(float));
    const size_t n = 3 * (size_t)nn;
    m.d_A.alloc(n * n * sizeof(float));
    m.d_Ainv.alloc(n * n * sizeof(float));
    m.d_W.alloc(n * n * sizeof(double));
    m.d_C.alloc(3 * n * 32 * sizeof(double));
    m.d_S.alloc(32 * 32 * sizeof(double));
    m.dense = true;
}

}  // namespace

// ======================================================================= C-ABI
extern "C" {

int sfem_mg_add_level(sfem_ctx* f, const sfem_mesh* coarse, const int32_t* inject, const int64_t* p_ptr,
                      const int32_t* p_col, const double* p_val, const double* young_e, const double* poisson_e)
{
    if (!f || !coarse || !inject || !p_ptr || !p_col || !p_val) return fail(SFEM_EINVAL, "sfem_mg_add_level: null argument");
    if (f->per_element && (!young_e || !poisson_e))
        return fail(SFEM_EINVAL, "sfem_mg_add_level: this context has per-element moduli; give the level's own");
    sfem_ctx* above = f->mg.empty() ? f : f->mg.back()->c;
    const int64_t na = above->nn, nc = coarse->nn;
    std::string err;
    if (!check_level_transfer(na, nc, inject, p_ptr, p_col, &err)) return fail(SFEM_EINVAL, "sfem_mg_add_level: " + err);
    const int64_t nnz = p_ptr[na];
    sfem_material mat{};
    mat.material = f->material == MAT_COROTATIONAL_POLAR ? SFEM_COROTATIONAL_POLAR : f->material;
    mat.young = f->young;
    mat.poisson = f->poisson;
    mat.density = f->density;
    mat.young_e = young_e;
    mat.poisson_e = poisson_e;
    sfem_ctx* lc = nullptr;
    int rc = sfem_create(coarse, &mat, &lc);
    if (rc != SFEM_OK) return rc;
    return guard_call("sfem_mg_add_level", [&]() -> int {
        std::unique_ptr<MgLevel> m(new MgLevel);
        m->c = lc;
        HIPCK(hipStreamSynchronize(lc->stream));
        HIPCK(hipStreamDestroy(lc->stream));
        lc->stream = f->stream;
        lc->owns_stream = false;
        // coarse Neo-Hookean levels: the safeguarded tangent
        if (lc->material == SFEM_NEO_HOOKEAN) lc->material = MAT_NEO_HOOKEAN_COARSE;
        std::vector<int> inj(inject, inject + nc);
        m->inj.upload(inj);
        std::vector<int> ptr(na + 1), col(nnz);
        std::vector<float> val(nnz);
        for (int64_t i = 0; i <= na; i++) ptr[i] = (int)p_ptr[i];
        for (int64_t k = 0; k < nnz; k++) {
            col[k] = p_col[k];
            val[k] = (float)p_val[k];
        }
        // transpose: for each coarse node its (fine row, weight) entries in
        // increasing fine-row order (the restriction's fixed summation order)
        std::vector<int> tptr, trow;
        std::vector<float> tval;
        transpose_prolongation(p_ptr, p_col, p_val, na, nc, &tptr, &trow, &tval);
        // only the coarsest level is solved directly: a level that now has one
        // below it gives its dense set-up back
        if (!f->mg.empty() && f->mg.back()->dense) {
            MgLevel& prev = *f->mg.back();
            HIPCK(hipStreamSynchronize(f->stream));
            if (f->mg_side) HIPCK(hipStreamSynchronize(f->mg_side));
            for (DevBuf* d : {&prev.d_tets4, &prev.d_dminv, &prev.d_cvol, &prev.d_inc_ptr, &prev.d_inc, &prev.d_Ke,
                              &prev.d_A, &prev.d_Ainv, &prev.d_W, &prev.d_C, &prev.d_S})
                d->release();
            prev.dense = false;
        }
        if (3 * nc <= MG_DENSE_MAX) dense_setup(*m, f, coarse, young_e, poisson_e);
        m->p_ptr.upload(ptr);
        m->p_col.upload(col);
        m->p_val.upload(val);
        m->t_ptr.upload(tptr);
        m->t_row.upload(trow);
        m->t_val.upload(tval);
        f->mg.push_back(std::move(m));
        f->mg_ready = false;
        f->mg_eig_ok = false;
        return SFEM_OK;
    });
}

int sfem_mg_clear(sfem_ctx* f)
{
    if (!f) return fail(SFEM_EINVAL, "sfem_mg_clear: null context");
    return guard_call("sfem_mg_clear", [&]() -> int {
        HIPCK(hipStreamSynchronize(f->stream));
        if (f->mg_side) HIPCK(hipStreamSynchronize(f->mg_side));
        f->mg.clear();
        f->mg_ready = false;
        f->mg_eig_ok = false;
        return SFEM_OK;
    });
}

int sfem_mg_set_schedule(sfem_ctx* f, int fused)
{
    if (!f) return fail(SFEM_EINVAL, "sfem_mg_set_schedule: null context");
    if (fused != 0 && fused != 1) return fail(SFEM_EINVAL, "sfem_mg_set_schedule: fused must be 0 or 1");
    f->mg_fused = fused != 0;
    return SFEM_OK;
}

int sfem_mg_info(sfem_ctx* f, int32_t* nlevels, double* lmax)
{
    if (!f || !nlevels) return fail(SFEM_EINVAL, "sfem_mg_info: null argument");
    *nlevels = (int32_t)f->mg.size() + 1;
    if (lmax) {
        lmax[0] = f->mg_lmax0;
        for (size_t l = 0; l < f->mg.size(); l++) lmax[l + 1] = f->mg[l]->lmax;
    }
    return SFEM_OK;
}

int sfem_mg_coarse_dense(sfem_ctx* f, int64_t capacity, int32_t* n, double* a, double* ainv)
{
    if (!f || !n) return fail(SFEM_EINVAL, "sfem_mg_coarse_dense: null argument");
    if (f->mg.empty() || !f->mg.back()->dense)
        return fail(SFEM_ESTATE, "sfem_mg_coarse_dense: the coarsest level is not solved directly");
    MgLevel& m = *f->mg.back();
    const int nd = 3 * (int)m.c->nn;
    *n = nd;
    if (!a && !ainv) return SFEM_OK;
    if (capacity < (int64_t)nd * nd) return fail(SFEM_EINVAL, "sfem_mg_coarse_dense: buffer smaller than n x n");
    return guard_call("sfem_mg_coarse_dense", [&]() -> int {
        build_jacobi(f);
        if (!mg_prepare(f, f->d_minv.as<float>(), true))
            return fail(SFEM_ESTATE, "multigrid: a level's operator is not finite at this state");
        HIPCK(hipStreamSynchronize(f->stream));
        std::vector<float> h((size_t)nd * nd);
        for (int k = 0; k < 2; k++) {
            double* out = k == 0 ? a : ainv;
            if (!out) continue;
            HIPCK(hipMemcpy(h.data(), (k == 0 ? m.d_A : m.d_Ainv).p, h.size() * sizeof(float), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < h.size(); i++) out[i] = h[i];
        }
        return SFEM_OK;
    });
}

int sfem_mg_apply(sfem_ctx* f, const double* r, double* z)
{
    if (!f || !r || !z) return fail(SFEM_EINVAL, "sfem_mg_apply: null argument");
    if (f->mg.empty()) return fail(SFEM_ESTATE, "sfem_mg_apply: no coarse levels (sfem_mg_add_level)");
    return guard_call("sfem_mg_apply", [&]() -> int {
        const int64_t n3 = 3 * f->nn;
        std::vector<float> h(n3);
        for (int64_t i = 0; i < n3; i++) h[i] = f->fixed_h[i / 3] ? 0.f : (float)r[i];
        build_jacobi(f);
        if (!mg_prepare(f, f->d_minv.as<float>(), true))
            return fail(SFEM_ESTATE, "multigrid: a level's operator is not finite at this state");
        HIPCK(hipMemcpyAsync(f->d_r32.p, h.data(), n3 * sizeof(float), hipMemcpyHostToDevice, f->stream));
        mg_vcycle(f);
        HIPCK(hipMemcpyAsync(h.data(), f->d_z32.p, n3 * sizeof(float), hipMemcpyDeviceToHost, f->stream));
        HIPCK(hipStreamSynchronize(f->stream));
        for (int64_t i = 0; i < n3; i++) z[i] = h[i];
        return SFEM_OK;
    });
}

int sfem_mg_pcg(sfem_ctx* f, const double* b, double* x, double rtol, int32_t maxit, int32_t* iters, int32_t* status)
{
    if (!f || !b || !x || !iters || !status || maxit <= 0) return fail(SFEM_EINVAL, "sfem_mg_pcg: bad argument");
    if (f->mg.empty()) return fail(SFEM_ESTATE, "sfem_mg_pcg: no coarse levels (sfem_mg_add_level)");
    return guard_call("sfem_mg_pcg", [&]() -> int {
        const int64_t n3 = 3 * f->nn;
        std::vector<float> h(n3);
        for (int64_t i = 0; i < n3; i++) h[i] = f->fixed_h[i / 3] ? 0.f : (float)b[i];
        build_jacobi(f);
        if (!mg_prepare(f, f->d_minv.as<float>(), true))
            return fail(SFEM_ESTATE, "multigrid: a level's operator is not finite at this state");
        HIPCK(hipMemcpyAsync(f->d_b32.p, h.data(), n3 * sizeof(float), hipMemcpyHostToDevice, f->stream));
        CgResult cg = mg_pcg(f, maxit, rtol * rtol, nullptr);
        HIPCK(hipMemcpyAsync(h.data(), f->d_x32.p, n3 * sizeof(float), hipMemcpyDeviceToHost, f->stream));
        HIPCK(hipStreamSynchronize(f->stream));
        for (int64_t i = 0; i < n3; i++) x[i] = h[i];
        *iters = cg.iters;
        *status = cg.status;
        return SFEM_OK;
    });
}

}  // extern "C"
