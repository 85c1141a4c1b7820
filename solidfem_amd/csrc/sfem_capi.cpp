// sfem_capi.cpp -- host side of libsfem_hip.so: the C-ABI declared in
// include/sfem_capi.h, the element partition, and the Newton-CG driver.
//
// Host <-> device data layout (HBM):
//   element SoA, block order:  Dm^-1 (9 x FP32 and 9 x FP64), V mu, V lambda
//                              (2 x FP32, 2 x FP64), 4 x uint16 block-local ids
//   block tables:              node list (uint32 id | flags), incidence lists
//                              (uint16), shared-node partial-row CSR
//   node vectors, x,y,z interleaved, original node order:
//       FP64: X0, u, trial u, residual, external load, mass (scalar)
//       FP32: u, CG x r z p q, preconditioner blocks (9 per node)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "sfem_capi.h"
#include "sfem_kernels.hpp"

#include "sfem_internal.hpp"

using namespace sfem;

namespace sfem {
thread_local std::string g_err;
int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}
}  // namespace sfem

namespace {

void build_partition(sfem_ctx* c, const sfem_mesh* m, const std::vector<double>& mu, const std::vector<double>& lam,
                     const std::vector<double>& dminv, const std::vector<double>& vol)
{
    const int64_t ne = m->ne, nn = m->nn;
    const Partition P = make_partition(m->tets, ne, m->X0, nn, m->fixed, mu, lam, dminv, vol, BLOCK_E);
    const int nblk = P.nblk;
    const int64_t ne_pad = P.ne_pad;
    const int nmax = P.nmax;
    const bool uniform = P.uniform;
    c->nblk = nblk;
    c->d_rel32.upload(P.rel32);
    if (!uniform) c->d_mat32.upload(P.mat32);
    c->d_eorig.upload(P.eorig);
    c->d_bnode_ptr.upload(P.bnode_ptr);
    c->d_bnodes.upload(P.bnodes);
    c->d_lptr.upload(P.lptr);
    c->d_lperm.upload(P.lperm);
    c->d_lidx.upload(P.lidx);
    c->d_dminv32.upload(P.dm32);
    c->d_coef32.upload(P.cf32);
    c->d_dminv64.upload(P.dm64);
    c->d_coef64.upload(P.cf64);
    c->d_sh_node.upload(P.sh_node);
    c->d_sh_ptr.upload(P.sh_ptr);
    c->d_bslot.upload(P.bslot);
    c->d_node_info.upload(P.node_q);
    c->nrows_partial = (int64_t)P.bnodes.size();
    c->nslots = P.sh_ptr.back();
    c->d_partial.alloc((size_t)c->nrows_partial * 9 * sizeof(float) + (size_t)c->nrows_partial * 3 * sizeof(double));

    BlockArgs& a = c->blk;
    a.ne = ne;
    a.ne_pad = ne_pad;
    a.nblk = nblk;
    a.nmax = nmax;
    a.bnode_ptr = c->d_bnode_ptr.as<int>();
    a.bnodes = c->d_bnodes.as<uint32_t>();
    a.bslot = c->d_bslot.as<int>();
    a.lptr = c->d_lptr.as<uint16_t>();
    a.lperm = c->d_lperm.as<uint2>();
    a.lidx = c->d_lidx.as<uint2>();
    a.dminv32 = c->d_dminv32.as<float>();
    a.coef32 = c->d_coef32.as<float>();
    a.dminv64 = c->d_dminv64.as<double>();
    a.coef64 = c->d_coef64.as<double>();
    a.rel32 = c->d_rel32.as<float>();
    a.mdiag32 = nullptr;
    a.mdiag64 = nullptr;
    a.node_q = c->d_node_info.as<int2>();
    a.zero_flag = c->d_node_info.as<int>() + 2 * nn;
    a.mat32 = uniform ? nullptr : c->d_mat32.as<float>();
    a.lin32 = nullptr;
    a.mu0 = (float)mu[0];
    a.lam0 = (float)lam[0];
    SharedArgs& s = c->shr;
    s.nshared = (int)P.sh_node.size();
    s.node = c->d_sh_node.as<uint32_t>();
    s.ptr = c->d_sh_ptr.as<int>();
}

template <typename T> void zero(DevBuf& b, size_t n, hipStream_t s)
{
    b.alloc(n * sizeof(T));
    HIPCK(hipMemsetAsync(b.p, 0, n * sizeof(T), s));
}

// Implicit dynamics time step h (0: static): m / h^2 per node for the element
// kernels (operator diagonal, residual inertia) and the load kernel
void set_time_step(sfem_ctx* c, double h)
{
    c->tstep = h;
    if (h > 0) {
        HIPCK(launch_mass_diag(c->nn, c->d_mass.as<double>(), c->d_fixed.as<uint8_t>(), 1.0 / (h * h), c->d_mdiag32.as<float>(),
                               c->d_mdiag64.as<double>(), c->stream));
        c->blk.mdiag32 = c->d_mdiag32.as<float>();
        c->blk.mdiag64 = c->d_mdiag64.as<double>();
    } else {
        c->blk.mdiag32 = nullptr;
        c->blk.mdiag64 = nullptr;
    }
}

// The Neo-Hookean tangent's state per element (E = Dm^-1 F^-1, V c) for the
// linear solves of one Newton step: built from u32 when they begin, read by
// every Hessian-vector product they make, dropped when they end, before u32
// changes, and by the destructor on every other way out.  The CG loops use it
// while it stays small beside the 256 MB Infinity Cache (their fused
// direction-update-and-product kernel: 17.4 instead of 18.1 us per launch on
// the 786,432-tet beam, 130 instead of 116 us on the 6.3M-tet one, where the
// 40 B per element come from HBM); the multigrid PCG, the stand-alone
// products and all other kernels read the state itself (a plain product is
// latency-bound, and the 40 B per element cost it more than the arithmetic
// they save: 15.2 against 13.7 us on the 786,432-tet beam).
constexpr size_t LIN_STATE_MAX_BYTES = size_t(64) << 20;

class LinState {
public:
    explicit LinState(sfem_ctx* c) : c_(c) {}
    ~LinState() { c_->blk.lin32 = nullptr; }
    LinState(const LinState&) = delete;
    LinState& operator=(const LinState&) = delete;
    void begin()
    {
        const size_t bytes = 10 * (size_t)c_->blk.ne_pad * sizeof(float);
        if (!c_->lin_cache || c_->material != SFEM_NEO_HOOKEAN || c_->blk.mdiag32 || bytes > LIN_STATE_MAX_BYTES) return;
        if (c_->d_lin.n < bytes) c_->d_lin.alloc(bytes);
        HIPCK(launch_lin_state(c_->blk, c_->d_u32.as<float>(), c_->d_lin.as<float>(), c_->stream));
        c_->blk.lin32 = c_->d_lin.as<float>();
    }
    void end() { c_->blk.lin32 = nullptr; }

private:
    sfem_ctx* c_;
};

// Newton steps a predicted load step gets to fall below the residual of the
// unpredicted start before it is abandoned
constexpr int PRED_PATIENCE = 4;

// SFEM_TRACE=1 in the environment: one line per Newton step of the converged
// solve on stderr (residual, forcing, inner solves and their CG status)
bool trace_on()
{
    static const bool on = [] {
        const char* v = std::getenv("SFEM_TRACE");
        return v && v[0] == '1';
    }();
    return on;
}

// ---------------------------------------------------------------- operations

// residual r = f_int + ext on free rows (0 on fixed) at displacement u (device)
void residual(sfem_ctx* c, const double* u, double* r)
{
    const int64_t nn = c->nn;
    if (c->ntri)
        HIPCK(launch_pressure(c->ntri, c->d_tris.as<int>(), c->d_X0.as<double>(), u, c->pressure, c->frac,
                              c->d_ftri.as<double>(), c->stream));
    const bool dyn = c->tstep > 0;
    HIPCK(launch_external(nn, c->d_mass.as<double>(), c->has_fdead ? c->d_fdead.as<double>() : nullptr, c->gravity, c->frac,
                          c->ntri ? c->d_pptr.as<int>() : nullptr, c->ntri ? c->d_pidx.as<int>() : nullptr,
                          c->ntri ? c->d_ftri.as<double>() : nullptr, c->d_fixed.as<uint8_t>(),
                          dyn ? c->d_vel.as<double>() : nullptr, dyn ? c->d_u0step.as<double>() : nullptr,
                          dyn ? 1.0 / c->tstep : 0.0, c->d_ext64.as<double>(), c->stream));
    // cable spring forces (ComputeGradients adds them, FemPhysicsMatrixFree.cpp:544)
    if (c->cab.n)
        HIPCK(launch_cable_forces(c->cab, c->d_X0.as<double>(), u, c->d_vel.as<double>(), c->d_cab_fs.as<double>(), c->frac,
                                  c->d_fixed.as<uint8_t>(), c->d_ext64.as<double>(), 1, c->stream));
    double* part = reinterpret_cast<double*>(c->d_partial.as<char>() + (size_t)c->nrows_partial * 9 * sizeof(float));
    HIPCK(launch_force(c->material, c->blk, c->shr, u, r, part, c->d_ext64.as<double>(), 1, c->stream));
}

// The cables' device arrays (CableArgs) from c->cables: per cable node its
// element's 4 nodes and 4 weights (w3 = 1 - w0 - w1 - w2, FemPhysicsBase.cpp:
// 233-238); per spring j -> j+1 of one cable (rest length, actuation,
// stiffness, damping), stiffness 0 after a cable's last node and for a cable
// with zero actuation (no force, :216-217); the touched mesh nodes with their
// attachments in cable-node order (the deterministic back-interpolation)
void rebuild_cables(sfem_ctx* c)
{
    std::vector<int> node4;
    std::vector<double> w4, spring;
    for (const sfem_ctx::Cable& cb : c->cables) {
        const size_t n = cb.elem.size();
        for (size_t i = 0; i < n; i++) {
            const int e = cb.elem[i];
            for (int k = 0; k < 4; k++) node4.push_back(c->tets_h[4 * (size_t)e + k]);
            const double w0 = cb.bary[3 * i], w1 = cb.bary[3 * i + 1], w2 = cb.bary[3 * i + 2];
            w4.insert(w4.end(), {w0, w1, w2, 1.0 - w0 - w1 - w2});
            const bool live = i + 1 < n && cb.actuation != 0.0;
            spring.insert(spring.end(), {cb.rest, cb.actuation, live ? cb.stiffness : 0.0, cb.damping});
        }
    }
    const int n = (int)(node4.size() / 4);
    std::vector<int> touched, tptr;
    std::vector<std::pair<int, int>> att;
    cable_attachments(node4, c->nn, &touched, &tptr, &att);
    std::vector<int2> tent(att.size());
    for (size_t k = 0; k < att.size(); k++) tent[k] = make_int2(att[k].first, att[k].second);
    c->d_cab_node.upload(node4);
    c->d_cab_w.upload(w4);
    c->d_cab_spring.upload(spring);
    c->d_cab_touched.upload(touched);
    c->d_cab_tptr.upload(tptr);
    c->d_cab_tent.upload(tent);
    c->d_cab_fs.alloc(3 * (size_t)std::max(n, 1) * sizeof(double));
    zero<double>(c->d_fcable, 3 * (size_t)c->nn, c->stream);
    HIPCK(hipStreamSynchronize(c->stream));
    c->cab.n = n;
    c->cab.node = c->d_cab_node.as<int4>();
    c->cab.w = c->d_cab_w.as<double>();
    c->cab.spring = c->d_cab_spring.as<double>();
    c->cab.ntouched = (int)touched.size();
    c->cab.touched = c->d_cab_touched.as<int>();
    c->cab.tptr = c->d_cab_tptr.as<int>();
    c->cab.tent = c->d_cab_tent.as<int2>();
}

}  // namespace

namespace sfem {

void hessvec(sfem_ctx* c, const float* d, float* out, const int* done)
{
    HIPCK(launch_hessvec(c->material, c->blk, c->shr, c->d_u32.as<float>(), d, out, c->d_partial.as<float>(), done, c->stream));
}

// a.a, b.b, d.a (FP64) -> host
Dots dots3(sfem_ctx* c, const double* a, const double* b, const float* d)
{
    HIPCK(launch_dots3(3 * c->nn, a, b, d, c->d_scal.as<double>(), S_NT, c->d_partials.as<double>(), c->stream));
    double h[3];
    HIPCK(hipMemcpyAsync(h, c->d_scal.as<double>() + S_NT, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return Dots{h[0], h[1], h[2]};
}

void build_jacobi(sfem_ctx* c)
{
    HIPCK(launch_blockdiag(c->material, c->blk, c->shr, c->d_u32.as<float>(), c->d_bd32.as<float>(), c->d_partial.as<float>(),
                           c->stream));
    HIPCK(launch_invert_blocks(c->nn, c->d_bd32.as<float>(), c->d_fixed.as<uint8_t>(), c->d_minv.as<float>(), c->stream));
}

void ensure_cg_slots(sfem_ctx* c)
{
    if (c->h_slot) return;
    HIPCK(hipHostMalloc((void**)&c->h_slot, 2 * sizeof(CgSlot), hipHostMallocMapped | hipHostMallocCoherent));
    HIPCK(hipHostGetDevicePointer((void**)&c->d_slot, c->h_slot, 0));
    for (hipEvent_t& e : c->flag_ev) HIPCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

// PCG on K x = b (b already in d_b32); x in d_x32.
// Two launches per iteration: the element kernel, which from the second
// iteration on also makes the previous iteration's convergence test and
// direction update (p = z + beta p, into the other buffer of d_p32 / d_p32b),
// then the update of x, r and z.  After the last launch, one direction kernel
// makes the test of the last update.
CgResult pcg(sfem_ctx* c, const float* minv, int pc, int maxit, int test_kind, double tol2, int guard, int check_every,
             int64_t* hv_evals, CgHook hook, void* user)
{
    const int64_t n3 = 3 * c->nn;
    double* scal = c->d_scal.as<double>();
    double* parts = c->d_partials.as<double>();          // CG update partials (2 x CG_GRID)
    double* pq_elem = parts + 2 * CG_GRID;               // p.q per element block (nblk)
    int* done = c->d_done.as<int>();
    float* pbuf[2] = {c->d_p32.as<float>(), c->d_p32b.as<float>()};
    int cur = 0;
    HIPCK(launch_cg_init(pc, n3, c->d_b32.as<float>(), c->d_x32.as<float>(), c->d_r32.as<float>(), c->d_z32.as<float>(),
                         pbuf[0], minv, c->d_fixed.as<uint8_t>(), scal, parts, done, c->stream, pbuf[1]));
    // guard 2: also stop at r.z <= 0 (a learned preconditioner that is not
    // positive definite at this state; its solve is redone with block Jacobi)
    const int rz_guard = guard >= 2 && pc;
    PFuse pf{};
    pf.z = pc ? c->d_z32.as<float>() : c->d_r32.as<float>();
    pf.rz_part = parts;
    pf.n_rz = cg_grid(c->nn);
    pf.scal = scal;
    pf.test_kind = test_kind;
    pf.tol2 = tol2;
    pf.rz_guard = rz_guard;
    pf.stall_window = 0;
    if (check_every < 1) check_every = 1;
    int launched = 0;
    int dflag = 0;
    double hs[S_COUNT];
    // one batch of up to check_every iterations; its tests write `slot`
    auto batch = [&](CgSlot* slot) {
        const int k = std::min(check_every, maxit - launched);
        pf.slot = slot;
        for (int i = 0; i < k; i++) {
            const bool timed = c->tm_every > 0 && c->tm_counter++ % c->tm_every == 0 && c->tm_pending < sfem_ctx::TM_POOL;
            hipEvent_t ev0 = timed ? c->tm_ev[2 * c->tm_pending] : nullptr;
            hipEvent_t ev1 = timed ? c->tm_ev[2 * c->tm_pending + 1] : nullptr;
            if (launched + i == 0) {
                HIPCK(launch_hessvec_cg(c->material, c->blk, c->d_u32.as<float>(), pbuf[cur], c->d_q32.as<float>(),
                                        c->d_partial.as<float>(), done, pq_elem, c->stream, ev0, ev1));
            } else {
                pf.p_old = pbuf[cur];
                pf.p_new = pbuf[cur ^ 1];
                cur ^= 1;
                HIPCK(launch_hessvec_cg_fused(c->material, c->blk, c->d_u32.as<float>(), pf, c->d_q32.as<float>(),
                                              c->d_partial.as<float>(), done, pq_elem, c->stream, ev0, ev1));
            }
            if (timed) c->tm_iter[c->tm_pending++] = launched + i;
            HIPCK(launch_cg_update(pc, n3, c->d_x32.as<float>(), c->d_r32.as<float>(), c->d_z32.as<float>(), pbuf[cur],
                                   c->d_q32.as<float>(), minv, c->d_node_info.as<int2>(), c->d_partial.as<float>(), scal,
                                   pq_elem, c->nblk, parts, done, guard, c->stream));
        }
        launched += k;
        if (launched >= maxit)
            HIPCK(launch_cg_test(pc, n3, pf.z, c->d_r32.as<float>(), pbuf[cur], scal, parts, done, test_kind, tol2, rz_guard, 0,
                                 c->stream, slot));
    };
    // the sampled launch timings of the launches before iteration `end`
    // (all of them complete), those at or after `upto` dropped (launches
    // past the converged iteration return at once)
    auto take_timings = [&](int end, double upto) {
        int kept = 0;
        for (int k = 0; k < c->tm_pending; k++) {
            if (c->tm_iter[k] < end) {
                float ms = 0.f;
                HIPCK(hipEventElapsedTime(&ms, c->tm_ev[2 * k], c->tm_ev[2 * k + 1]));
                if ((double)c->tm_iter[k] < upto) {
                    c->tm_ms += ms;
                    c->tm_samples++;
                }
            } else {
                std::swap(c->tm_ev[2 * kept], c->tm_ev[2 * k]);
                std::swap(c->tm_ev[2 * kept + 1], c->tm_ev[2 * k + 1]);
                c->tm_iter[kept++] = c->tm_iter[k];
            }
        }
        c->tm_pending = kept;
    };
    if (hook) {
        // the caller acts between batches: one batch at a time
        while (launched < maxit) {
            batch(nullptr);
            HIPCK(hipMemcpyAsync(&dflag, done, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPCK(hipStreamSynchronize(c->stream));
            if (c->tm_pending) {
                double it_done = 1e300;
                if (dflag) HIPCK(hipMemcpy(&it_done, scal + S_ITER, sizeof(double), hipMemcpyDeviceToHost));
                take_timings(launched, it_done);
            }
            if (dflag) break;
            hook(c, launched, user);
        }
    } else if (maxit > 0) {
        // One batch stays enqueued beyond the one whose outcome the host waits
        // for (the device never waits for the host), except where the last two
        // outcomes predict that the batch in flight converges: then the host
        // waits for it first, which saves a batch of early returns at the end
        // of the solve when the prediction holds and costs one host round trip
        // when it does not.  Batch j's tests write slot j & 1 (reset by the
        // host before the batch is enqueued: batch j - 2, its last writer, has
        // completed), and an event follows it.
        ensure_cg_slots(c);
        int nenq = 0, bend[2] = {0, 0};
        auto enqueue = [&]() {
            CgSlot& hs_ = c->h_slot[nenq & 1];
            hs_.done = 0;
            hs_.iter = 0;
            hs_.rr_rel = 0.0;
            batch(c->d_slot + (nenq & 1));
            HIPCK(hipEventRecord(c->flag_ev[nenq & 1], c->stream));
            bend[nenq & 1] = launched;
            nenq++;
        };
        enqueue();
        double rr_prev = 0.0;
        bool hold = false;
        for (int j = 0;; j++) {
            if (nenq == j + 1 && launched < maxit && !hold) enqueue();
            HIPCK(hipEventSynchronize(c->flag_ev[j & 1]));
            const CgSlot out = c->h_slot[j & 1];
            dflag = out.done;
            if (c->tm_pending) take_timings(bend[j & 1], dflag ? (double)out.iter : 1e300);
            if (dflag) break;
            if (nenq == j + 1) {      // nothing in flight: held back, or the iteration limit
                if (launched >= maxit) break;
                enqueue();
            }
            const double rr = out.rr_rel;
            hold = test_kind == 1 && rr > 0.0 && rr_prev > 0.0 && rr * std::min(1.0, rr / rr_prev) <= tol2;
            rr_prev = rr;
        }
        c->tm_pending = 0;   // launches past the end of the solve
    }
    HIPCK(hipMemcpy(hs, scal, sizeof hs, hipMemcpyDeviceToHost));
    CgResult res;
    res.iters = (int)hs[S_ITER];
    res.status = dflag;
    res.rr0 = hs[S_RR0];
    res.rr = hs[S_RR];
    if (hv_evals) *hv_evals += (int64_t)res.iters * c->ne;
    return res;
}

}  // namespace sfem

namespace {

// ----------------------------------------------------------- Newton drivers

// SolveNewtonCG (FemPhysicsMatrixFree.cpp:604-628): fixed Newton steps, CG
// from zero each step, eps 1e-3f on r.r, no residual test, no line search.
int solve_reference(sfem_ctx* c, const sfem_solve_params* p, sfem_solve_stats* st)
{
    const int64_t n3 = 3 * c->nn;
    HIPCK(launch_identity_blocks(c->nn, c->d_fixed.as<uint8_t>(), c->d_minv.as<float>(), c->stream));
    const float eps = 1e-3f;
    const float e2 = eps * eps;   // float product, as the reference evaluates eps * eps
    LinState lin(c);
    for (int it = 0; it < p->outer_iterations; it++) {
        residual(c, c->d_u64.as<double>(), c->d_r64.as<double>());
        st->force_evals += c->ne;
        if (it == 0) st->res0 = std::sqrt(dots3(c, c->d_r64.as<double>(), nullptr, nullptr).aa);
        HIPCK(launch_to_f32(n3, c->d_r64.as<double>(), c->d_b32.as<float>(), c->stream));
        lin.begin();
        CgResult cg = pcg(c, c->d_minv.as<float>(), 0, p->inner_iterations, 0, (double)e2, 0, p->check_every,
                          &st->hessvec_evals);
        lin.end();
        st->cg_iters += cg.iters;
        HIPCK(launch_axpy_state(n3, c->d_u64.as<double>(), c->d_x32.as<float>(), 1.0, c->d_ut64.as<double>(),
                                c->d_u32.as<float>(), c->stream));
        std::swap(c->d_u64.p, c->d_ut64.p);
        st->newton_iters++;
    }
    residual(c, c->d_u64.as<double>(), c->d_r64.as<double>());
    st->force_evals += c->ne;
    Dots dn = dots3(c, c->d_r64.as<double>(), c->d_ext64.as<double>(), nullptr);
    st->res = std::sqrt(dn.aa);
    st->ext_norm = std::sqrt(dn.bb);
    st->converged = 0;
    return SFEM_OK;
}

// Converged Newton-PCG: FP64 residual test and line search, FP32 inner PCG.
//  - forcing: Eisenstat-Walker choice 2 (gamma 0.9, exponent 2) capped by
//    cg_rtol, floored where the outer target no longer needs more;
//  - line search on the slope R(a) = dx . r(u + a dx) (Bonet & Wood): accept
//    when |R(a)| <= 0.5 R(0) or R(a) > 0 (still descending), otherwise move to
//    the secant root of R, at most 8 trials (the reference's back-tracking
//    Newton, Include/NewtonSolver.h:229-293, bisects on the same quantity).
int solve_converged(sfem_ctx* c, const sfem_solve_params* p, sfem_solve_stats* st)
{
    const int64_t n3 = 3 * c->nn;
    int pc_kind = p->precond;
    if ((pc_kind == SFEM_PC_LEARNED || pc_kind == SFEM_PC_MULTIGRID_LEARNED) && !c->has_learned && !c->gnn)
        return fail(SFEM_ESTATE, "learned preconditioner not installed");
    const bool multigrid = pc_kind == SFEM_PC_MULTIGRID || pc_kind == SFEM_PC_MULTIGRID_LEARNED;
    if (multigrid && c->mg.empty()) return fail(SFEM_ESTATE, "multigrid preconditioner without coarse levels");
    if (multigrid && c->tstep > 0)
        return fail(SFEM_ESTATE, "multigrid preconditioner: static and quasi-static solves only");
    if (pc_kind == SFEM_PC_NONE)
        HIPCK(launch_identity_blocks(c->nn, c->d_fixed.as<uint8_t>(), c->d_minv.as<float>(), c->stream));
    residual(c, c->d_u64.as<double>(), c->d_r64.as<double>());
    st->force_evals += c->ne;
    Dots d0 = dots3(c, c->d_r64.as<double>(), c->d_ext64.as<double>(), nullptr);
    double rn = std::sqrt(d0.aa), en = std::sqrt(d0.bb);
    st->res0 = rn;
    const bool adaptive = pc_kind == SFEM_PC_LEARNED || multigrid;
    int maxn = p->outer_iterations > 0 ? p->outer_iterations : 50;
    if (adaptive && p->pc_newton_cap > 0) maxn = std::min(maxn, p->pc_newton_cap);
    // the forcing cap: 0.3 for the multigrid PCG, whose few iterations per Newton
    // step make a looser cap pay, 0.1 for the Jacobi-type PCGs (with 0.3 the
    // nu = 0.495 beam's first load step no longer converged in 50 Newton steps)
    const double eta_max = p->cg_rtol > 0 ? p->cg_rtol : (multigrid ? 0.3 : 0.1);
    double eta = eta_max, rn_prev = 0;
    const double abs_tol = p->abs_tol > 0 ? p->abs_tol : 0.0;
    auto target = [&](double e) { return std::max(p->tol * e, abs_tol); };
    LinState lin(c);
    for (int it = 0; it < maxn; it++) {
        if (rn <= target(en)) break;
        if (it > 0) {
            const double ew = 0.9 * (rn / rn_prev) * (rn / rn_prev);
            const double safe = 0.9 * eta * eta;
            eta = std::min(eta_max, safe > 0.1 ? std::max(ew, safe) : ew);
        }
        eta = std::max(eta, std::max(0.5 * target(en) / rn, 1e-6));
        const bool learned = pc_kind == SFEM_PC_LEARNED || pc_kind == SFEM_PC_MULTIGRID_LEARNED;
        // the cached state pays in the CG loop's fused kernel (direction update
        // and product); the V-cycle's plain products are faster without it
        if (!multigrid) lin.begin();
        if (pc_kind == SFEM_PC_JACOBI || multigrid) build_jacobi(c);
        if (learned && c->gnn) {
            if (!multigrid) build_jacobi(c);   // network input and the blocks it corrects
            gnn_install(c);
        }
        HIPCK(launch_to_f32(n3, c->d_r64.as<double>(), c->d_b32.as<float>(), c->stream));
        const float* minv = learned ? c->d_minv_learned.as<float>() : c->d_minv.as<float>();
        CgResult cg{0, 0, 0.0, 0.0};
        if (multigrid) {
            if (mg_prepare(c, minv, it == 0, true)) cg = mg_pcg(c, p->inner_iterations, eta * eta, &st->hessvec_evals);
        } else {
            // learned blocks: also stop at r.z <= 0 (blocks not positive definite)
            cg = pcg(c, minv, pc_kind == SFEM_PC_NONE ? 0 : 1, p->inner_iterations, 1, eta * eta, learned ? 2 : 1,
                     p->check_every, &st->hessvec_evals);
        }
        st->cg_iters += cg.iters;
        if (c->abort_after > 0 && cg.status == CG_BREAKDOWN) {
            // a predicted start where the tangent (or its preconditioner) is not
            // positive definite: abandon it at once, the step is redone from the
            // last state (sfem_step)
            if (trace_on()) std::fprintf(stderr, "sfem newton %d: predicted start rejected (breakdown)\n", it);
            c->pred_aborted = true;
            break;
        }
        if (trace_on())
            std::fprintf(stderr, "sfem newton %d: |r| %.3e target %.3e eta %.2e pc %d: cg %d status %d\n", it, rn,
                         target(en), eta, pc_kind, cg.iters, cg.status);
        if ((learned || multigrid) && cg.status != CG_CONVERGED) {
            // the learned / multigrid PCG did not reach its forcing (iteration
            // limit, breakdown, stagnation or a non-finite value): redo this
            // Newton step's linear solve with block Jacobi
            st->fallbacks++;
            build_jacobi(c);
            cg = pcg(c, c->d_minv.as<float>(), 1, p->inner_iterations, 1, eta * eta, 1, p->check_every, &st->hessvec_evals);
            st->cg_iters += cg.iters;
            if (trace_on()) std::fprintf(stderr, "sfem   fallback jacobi: cg %d status %d\n", cg.iters, cg.status);
        }
        lin.end();
        if (cg.status == CG_BREAKDOWN && cg.iters == 0) {
            // non-positive curvature along the first direction: no direction
            // at all, and every further Newton step would repeat this one
            if (trace_on()) std::fprintf(stderr, "sfem newton %d: no direction (breakdown at the first CG step)\n", it);
            break;
        }
        // the slope R(0) = dx . r(u): its dots are read with the first trial's
        // (one host round trip fewer per Newton step)
        HIPCK(launch_dots3(3 * c->nn, c->d_r64.as<double>(), nullptr, c->d_x32.as<float>(), c->d_scal.as<double>(), S_SL,
                           c->d_partials.as<double>(), c->stream));
        double R0 = 0.0;
        bool have_r0 = false;
        double alpha = 1.0;
        Dots dt{0, 0, 0};
        bool finite = false;
        const int ntrial = p->line_search ? 8 : 1;
        auto trial = [&]() {
            HIPCK(launch_axpy_state(n3, c->d_u64.as<double>(), c->d_x32.as<float>(), alpha, c->d_ut64.as<double>(),
                                    c->d_u32.as<float>(), c->stream));
            residual(c, c->d_ut64.as<double>(), c->d_f64tmp.as<double>());
            st->force_evals += c->ne;
            HIPCK(launch_dots3(3 * c->nn, c->d_f64tmp.as<double>(), c->d_ext64.as<double>(), c->d_x32.as<float>(),
                               c->d_scal.as<double>(), S_NT, c->d_partials.as<double>(), c->stream));
            double h[S_COUNT];
            HIPCK(hipMemcpyAsync(h, c->d_scal.p, sizeof h, hipMemcpyDeviceToHost, c->stream));
            HIPCK(hipStreamSynchronize(c->stream));
            dt = Dots{h[S_NT], h[S_NT + 1], h[S_NT + 2]};
            if (!have_r0) {
                R0 = h[S_SL + 2];
                have_r0 = true;
            }
            finite = std::isfinite(dt.da) && std::isfinite(dt.aa);
        };
        for (int ls = 0; ls < ntrial; ls++) {
            trial();
            const double Ra = dt.da;
            if (!p->line_search || ls == ntrial - 1) break;
            if (finite && (std::fabs(Ra) <= 0.5 * std::fabs(R0) || Ra * R0 > 0)) break;
            alpha = finite ? std::min(std::max(alpha * R0 / (R0 - Ra), 0.1 * alpha), 0.9 * alpha) : 0.5 * alpha;
        }
        // never accept a state with inverted elements (non-finite residual):
        // shorten the step until it is finite, or stop at the current state
        for (int k = 0; !finite && k < 40; k++) {
            alpha *= 0.5;
            trial();
        }
        if (!finite) {
            HIPCK(launch_to_f32(n3, c->d_u64.as<double>(), c->d_u32.as<float>(), c->stream));
            break;
        }
        std::swap(c->d_u64.p, c->d_ut64.p);
        std::swap(c->d_r64.p, c->d_f64tmp.p);
        rn_prev = rn;
        rn = std::sqrt(dt.aa);
        en = std::sqrt(dt.bb);
        st->newton_iters++;
        if (!std::isfinite(rn)) break;
        // a predicted start that is not paying off (see sfem_step)
        if (c->abort_after > 0 && st->newton_iters >= c->abort_after && rn > c->abort_rn && rn > target(en)) {
            c->pred_aborted = true;
            break;
        }
    }
    // u32 must match the accepted state
    HIPCK(launch_to_f32(n3, c->d_u64.as<double>(), c->d_u32.as<float>(), c->stream));
    st->res = rn;
    st->ext_norm = en;
    st->converged = rn <= target(en);
    return st->converged ? SFEM_OK : SFEM_ENOTCONV;
}

}  // namespace

// ======================================================================= C-ABI
extern "C" {

const char* sfem_last_error(void) { return g_err.c_str(); }

int sfem_create(const sfem_mesh* m, const sfem_material* mat, sfem_ctx** out)
{
    if (!m || !mat || !out || !m->tets || !m->X0 || !m->fixed || m->ne <= 0 || m->nn <= 0)
        return fail(SFEM_EINVAL, "sfem_create: null or empty mesh");
    if (m->nn >= (1 << 29)) return fail(SFEM_EINVAL, "sfem_create: more than 2^29 nodes");
    // partial slots (at most 4 per element) share the node ids' 29 bits in the element kernel's LDS
    if (m->ne >= (1 << 27)) return fail(SFEM_EINVAL, "sfem_create: more than 2^27 elements");
    if ((mat->material < SFEM_LINEAR || mat->material > SFEM_NEO_HOOKEAN) && mat->material != SFEM_COROTATIONAL_POLAR)
        return fail(SFEM_EINVAL, "sfem_create: unknown material");
    for (int64_t i = 0; i < 4 * m->ne; i++)
        if (m->tets[i] < 0 || m->tets[i] >= m->nn) return fail(SFEM_EINVAL, "sfem_create: node index out of range");
    *out = nullptr;
    sfem_ctx* c = new sfem_ctx;
    int rc = guard_call("sfem_create", [&]() -> int {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw HipError{hipErrorNoDevice, "no HIP device"};
        HIPCK(hipGetDevice(&c->device));
        // the allocation-free entry point's scratch array on this device
        if (noalloc_init(c->device) != SFEM_OK) throw std::runtime_error(g_err);
        HIPCK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->ne = m->ne;
        c->nn = m->nn;
        c->material = mat->material == SFEM_COROTATIONAL_POLAR ? MAT_COROTATIONAL_POLAR : mat->material;
        c->young = mat->young;
        c->poisson = mat->poisson;
        c->density = mat->density;
        c->per_element = mat->young_e != nullptr || mat->poisson_e != nullptr;
        for (int k = 0; k < 3; k++) c->gravity[k] = mat->gravity[k];
        const int64_t ne = m->ne, nn = m->nn;
        // --- geometry (FemPhysicsMatrixFree.cpp:126-146): Dm^-1, V, lumped masses
        ElementGeometry g;
        std::string err;
        if (!element_geometry(m->tets, ne, m->X0, nn, m->fixed, mat->young, mat->poisson, mat->young_e, mat->poisson_e,
                              mat->density, &g, &err))
            throw std::runtime_error(err);
        c->fixed_h = g.held;
        c->nfixed = g.nheld;
        sfem_mesh mf = *m;
        mf.fixed = c->fixed_h.data();
        c->tets_h.assign(m->tets, m->tets + 4 * ne);
        const std::vector<double>& mass = g.mass;
        const std::vector<double>&dminv = g.dminv, &vol = g.vol, &mu = g.mu, &lam = g.lam;
        build_partition(c, &mf, mu, lam, dminv, vol);
        HIPCK(set_elem_lds_limits(c->blk));
        if (elem_lds_max(c->blk) > 160 * 1024) throw std::runtime_error("element block exceeds LDS");
        std::vector<double> X0(m->X0, m->X0 + 3 * nn);
        c->d_X0.upload(X0);
        c->d_mass.upload(mass);
        c->d_fixed.upload(c->fixed_h);
        const size_t n3 = 3 * (size_t)nn;
        hipStream_t s = c->stream;
        zero<double>(c->d_u64, n3, s);
        zero<double>(c->d_uprev, n3, s);
        zero<double>(c->d_uprev2, n3, s);
        zero<double>(c->d_ustart, n3, s);
        zero<double>(c->d_vel, n3, s);
        zero<double>(c->d_u0step, n3, s);
        zero<float>(c->d_mdiag32, (size_t)nn, s);
        zero<double>(c->d_mdiag64, (size_t)nn, s);
        zero<double>(c->d_ut64, n3, s);
        zero<float>(c->d_u32, n3, s);
        zero<double>(c->d_r64, n3, s);
        zero<double>(c->d_ext64, n3, s);
        zero<double>(c->d_f64tmp, n3, s);
        for (DevBuf* b : {&c->d_b32, &c->d_x32, &c->d_r32, &c->d_z32, &c->d_p32, &c->d_p32b, &c->d_q32, &c->d_d32})
            zero<float>(*b, n3, s);
        zero<float>(c->d_minv, 9 * (size_t)nn, s);
        zero<float>(c->d_bd32, 9 * (size_t)nn, s);
        zero<double>(c->d_scal, S_COUNT, s);
        zero<double>(c->d_partials, std::max<size_t>(3 * (size_t)RED_GRID, 2 * (size_t)CG_GRID + c->nblk), s);
        zero<int>(c->d_done, 4, s);   // [0] the CG flag; [1] stays 0 (the multigrid PCG copies [0..1])
        HIPCK(hipStreamSynchronize(s));
        return SFEM_OK;
    });
    if (rc != SFEM_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return SFEM_OK;
}

void sfem_destroy(sfem_ctx* c) { delete c; }

int sfem_time_hessvec(sfem_ctx* c, int every, double* ms_avg, int64_t* samples)
{
    if (!c || every < 0) return fail(SFEM_EINVAL, "sfem_time_hessvec: bad argument");
    return guard_call("sfem_time_hessvec", [&]() -> int {
        if (ms_avg) *ms_avg = c->tm_samples ? c->tm_ms / (double)c->tm_samples : 0.0;
        if (samples) *samples = c->tm_samples;
        if (every > 0 && !c->tm_ev[0])
            for (hipEvent_t& e : c->tm_ev) HIPCK(hipEventCreate(&e));
        if (every != c->tm_every || every > 0) {
            c->tm_every = every;
            c->tm_counter = 0;
            c->tm_samples = 0;
            c->tm_ms = 0.0;
        }
        return SFEM_OK;
    });
}

int sfem_info(sfem_ctx* c, int64_t* ne, int64_t* nn, int64_t* nfixed, int64_t* nblocks)
{
    if (!c) return fail(SFEM_EINVAL, "sfem_info: null context");
    if (ne) *ne = c->ne;
    if (nn) *nn = c->nn;
    if (nfixed) *nfixed = c->nfixed;
    if (nblocks) *nblocks = c->nblk;
    return SFEM_OK;
}

int sfem_set_positions(sfem_ctx* c, const double* x)
{
    if (!c || !x) return fail(SFEM_EINVAL, "sfem_set_positions: null argument");
    return guard_call("sfem_set_positions", [&]() -> int {
        const size_t n3 = 3 * (size_t)c->nn;
        std::vector<double> X0(n3), u(n3);
        HIPCK(hipMemcpy(X0.data(), c->d_X0.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n3; i++) u[i] = x[i] - X0[i];
        HIPCK(hipMemcpy(c->d_u64.p, u.data(), n3 * sizeof(double), hipMemcpyHostToDevice));
        HIPCK(launch_to_f32((int64_t)n3, c->d_u64.as<double>(), c->d_u32.as<float>(), c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        c->has_prev = c->has_prev2 = false;
        return SFEM_OK;
    });
}

int sfem_get_positions(sfem_ctx* c, double* x)
{
    if (!c || !x) return fail(SFEM_EINVAL, "sfem_get_positions: null argument");
    return guard_call("sfem_get_positions", [&]() -> int {
        const size_t n3 = 3 * (size_t)c->nn;
        std::vector<double> X0(n3);
        HIPCK(hipStreamSynchronize(c->stream));
        HIPCK(hipMemcpy(X0.data(), c->d_X0.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
        HIPCK(hipMemcpy(x, c->d_u64.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n3; i++) x[i] += X0[i];
        return SFEM_OK;
    });
}

int sfem_set_load_fraction(sfem_ctx* c, double frac)
{
    if (!c) return fail(SFEM_EINVAL, "sfem_set_load_fraction: null context");
    c->frac = frac;
    c->has_prev = c->has_prev2 = false;
    return SFEM_OK;
}

int sfem_get_load_fraction(sfem_ctx* c, double* frac)
{
    if (!c || !frac) return fail(SFEM_EINVAL, "sfem_get_load_fraction: null argument");
    *frac = c->frac;
    return SFEM_OK;
}

int sfem_set_pressure_surface(sfem_ctx* c, const int32_t* tris, int64_t ntri, double pressure)
{
    if (!c || (ntri > 0 && !tris) || ntri < 0) return fail(SFEM_EINVAL, "sfem_set_pressure_surface: bad argument");
    for (int64_t i = 0; i < 3 * ntri; i++)
        if (tris[i] < 0 || tris[i] >= c->nn) return fail(SFEM_EINVAL, "sfem_set_pressure_surface: node index out of range");
    return guard_call("sfem_set_pressure_surface", [&]() -> int {
        c->ntri = ntri;
        c->pressure = pressure;
        if (ntri == 0) return SFEM_OK;
        std::vector<int> t(tris, tris + 3 * ntri);
        // node -> incident triangles (CSR, triangle order)
        std::vector<int> ptr, idx;
        node_triangle_csr(tris, ntri, c->nn, &ptr, &idx);
        c->d_tris.upload(t);
        c->d_pptr.upload(ptr);
        c->d_pidx.upload(idx);
        c->d_ftri.alloc(3 * ntri * sizeof(double));
        return SFEM_OK;
    });
}

int sfem_set_external_forces(sfem_ctx* c, const double* f)
{
    if (!c) return fail(SFEM_EINVAL, "sfem_set_external_forces: null context");
    return guard_call("sfem_set_external_forces", [&]() -> int {
        if (!f) { c->has_fdead = false; return SFEM_OK; }
        std::vector<double> v(f, f + 3 * c->nn);
        c->d_fdead.upload(v);
        c->has_fdead = true;
        return SFEM_OK;
    });
}

int sfem_internal_forces(sfem_ctx* c, double* f)
{
    if (!c || !f) return fail(SFEM_EINVAL, "sfem_internal_forces: null argument");
    return guard_call("sfem_internal_forces", [&]() -> int {
        double* part = reinterpret_cast<double*>(c->d_partial.as<char>() + (size_t)c->nrows_partial * 9 * sizeof(float));
        BlockArgs a = c->blk;   // elastic forces only, whatever the time step
        a.mdiag32 = nullptr;
        a.mdiag64 = nullptr;
        HIPCK(launch_force(c->material, a, c->shr, c->d_u64.as<double>(), c->d_f64tmp.as<double>(), part, nullptr, 0,
                           c->stream));
        HIPCK(hipMemcpyAsync(f, c->d_f64tmp.p, 3 * c->nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        return SFEM_OK;
    });
}

int sfem_element_stress(sfem_ctx* c, double* P)
{
    if (!c || !P) return fail(SFEM_EINVAL, "sfem_element_stress: null argument");
    return guard_call("sfem_element_stress", [&]() -> int {
        DevBuf out;
        out.alloc((size_t)9 * c->ne * sizeof(double));
        HIPCK(launch_elem_stress(c->material, c->blk, c->d_u64.as<double>(), c->d_eorig.as<int>(), out.as<double>(), c->stream));
        HIPCK(hipMemcpyAsync(P, out.p, (size_t)9 * c->ne * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        return SFEM_OK;
    });
}

int sfem_get_velocities(sfem_ctx* c, double* v)
{
    if (!c || !v) return fail(SFEM_EINVAL, "sfem_get_velocities: null argument");
    return guard_call("sfem_get_velocities", [&]() -> int {
        HIPCK(hipMemcpy(v, c->d_vel.p, 3 * c->nn * sizeof(double), hipMemcpyDeviceToHost));
        return SFEM_OK;
    });
}

int sfem_gradients(sfem_ctx* c, double* r)
{
    if (!c || !r) return fail(SFEM_EINVAL, "sfem_gradients: null argument");
    return guard_call("sfem_gradients", [&]() -> int {
        residual(c, c->d_u64.as<double>(), c->d_f64tmp.as<double>());
        HIPCK(hipMemcpyAsync(r, c->d_f64tmp.p, 3 * c->nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        return SFEM_OK;
    });
}

int sfem_matvec(sfem_ctx* c, const double* d, double* Kd)
{
    if (!c || !d || !Kd) return fail(SFEM_EINVAL, "sfem_matvec: null argument");
    return guard_call("sfem_matvec", [&]() -> int {
        const size_t n3 = 3 * (size_t)c->nn;
        std::vector<float> d32(n3), o(n3);
        for (int64_t i = 0; i < c->nn; i++)
            for (int k = 0; k < 3; k++) d32[3 * i + k] = c->fixed_h[i] ? 0.f : (float)d[3 * i + k];
        HIPCK(hipMemcpyAsync(c->d_d32.p, d32.data(), n3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
        LinState lin(c);
        if (c->lin_cache == 2) lin.begin();
        hessvec(c, c->d_d32.as<float>(), c->d_q32.as<float>(), nullptr);
        HIPCK(hipMemcpyAsync(o.data(), c->d_q32.p, n3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < n3; i++) Kd[i] = o[i];
        return SFEM_OK;
    });
}

int sfem_energy(sfem_ctx* c, double* out)
{
    if (!c || !out) return fail(SFEM_EINVAL, "sfem_energy: null argument");
    return guard_call("sfem_energy", [&]() -> int {
        double* scal = c->d_scal.as<double>();
        double* part = c->d_partials.as<double>();
        HIPCK(launch_energy(c->material, c->blk, c->d_u64.as<double>(), scal, S_NT, part, c->stream));
        if (c->cab.n) HIPCK(launch_cable_energy(c->cab, c->d_X0.as<double>(), c->d_u64.as<double>(), scal, S_NT2, part, c->stream));
        else HIPCK(hipMemsetAsync(scal + S_NT2, 0, sizeof(double), c->stream));
        HIPCK(launch_dead_work(c->nn, c->d_mass.as<double>(), c->has_fdead ? c->d_fdead.as<double>() : nullptr, c->gravity,
                               c->frac, c->d_fixed.as<uint8_t>(), c->d_u64.as<double>(), scal, S_NT3, part, c->stream));
        HIPCK(hipMemcpyAsync(out, scal + S_NT, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        return SFEM_OK;
    });
}

int sfem_force_param_grads(sfem_ctx* c, double* fmu, double* flam, double* frho)
{
    if (!c || !fmu || !flam || !frho) return fail(SFEM_EINVAL, "sfem_force_param_grads: null argument");
    return guard_call("sfem_force_param_grads", [&]() -> int {
        const size_t n3 = 3 * (size_t)c->nn;
        double* part = reinterpret_cast<double*>(c->d_partial.as<char>() + (size_t)c->nrows_partial * 9 * sizeof(float));
        BlockArgs a = c->blk;   // elastic forces only, whatever the time step
        a.mdiag32 = nullptr;
        a.mdiag64 = nullptr;
        for (int mode = 1; mode <= 2; mode++) {
            a.coef_mode = mode;
            HIPCK(launch_force(c->material, a, c->shr, c->d_u64.as<double>(), c->d_f64tmp.as<double>(), part, nullptr, 0,
                               c->stream));
            HIPCK(hipMemcpyAsync(mode == 1 ? fmu : flam, c->d_f64tmp.p, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        }
        // d(m g)/d rho = (sum of V / 4 over the node's elements) g on free nodes
        // (FemPhysicsBase.cpp:171-186)
        std::vector<double> X0(n3);
        HIPCK(hipMemcpyAsync(X0.data(), c->d_X0.p, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        std::vector<double> nv(c->nn, 0.0);
        for (int64_t e = 0; e < c->ne; e++) {
            const int32_t* t = &c->tets_h[4 * e];
            double D[9];
            for (int k = 1; k < 4; k++)
                for (int r = 0; r < 3; r++) D[3 * r + (k - 1)] = X0[3 * (size_t)t[k] + r] - X0[3 * (size_t)t[0] + r];
            const double det = D[0] * (D[4] * D[8] - D[5] * D[7]) - D[1] * (D[3] * D[8] - D[5] * D[6]) +
                               D[2] * (D[3] * D[7] - D[4] * D[6]);
            for (int k = 0; k < 4; k++) nv[t[k]] += 0.25 * (det / 6.0);
        }
        for (int64_t i = 0; i < c->nn; i++)
            for (int k = 0; k < 3; k++) frho[3 * i + k] = c->fixed_h[i] ? 0.0 : nv[i] * c->gravity[k];
        return SFEM_OK;
    });
}

int sfem_add_cable(sfem_ctx* c, const int32_t* elem, const double* bary, int32_t n, double rest_length, double stiffness,
                   double damping, int32_t* id)
{
    if (!c || !elem || !bary || n < 2 || !(rest_length > 0)) return fail(SFEM_EINVAL, "sfem_add_cable: bad argument");
    for (int32_t i = 0; i < n; i++) {
        if (elem[i] < 0)
            return fail(SFEM_EINVAL, "sfem_add_cable: free cable nodes (outside the mesh, elem < 0) are not supported: "
                                     "they would add degrees of freedom with no elastic stiffness");
        if (elem[i] >= c->ne) return fail(SFEM_EINVAL, "sfem_add_cable: element index out of range");
    }
    return guard_call("sfem_add_cable", [&]() -> int {
        sfem_ctx::Cable cb;
        cb.elem.assign(elem, elem + n);
        cb.bary.assign(bary, bary + 3 * (size_t)n);
        cb.rest = rest_length;
        cb.stiffness = stiffness;
        cb.damping = damping;
        cb.actuation = 1.0;
        c->cables.push_back(cb);
        if (id) *id = (int32_t)c->cables.size() - 1;
        rebuild_cables(c);
        return SFEM_OK;
    });
}

int sfem_set_cable_actuation(sfem_ctx* c, int32_t cable, double actuation)
{
    if (!c) return fail(SFEM_EINVAL, "sfem_set_cable_actuation: null context");
    // SetCableActuation (FemPhysicsBase.h:134) ignores an index out of range
    if (cable < 0 || (size_t)cable >= c->cables.size()) return SFEM_OK;
    return guard_call("sfem_set_cable_actuation", [&]() -> int {
        c->cables[cable].actuation = actuation;
        rebuild_cables(c);
        return SFEM_OK;
    });
}

int sfem_num_cables(sfem_ctx* c, int32_t* n)
{
    if (!c || !n) return fail(SFEM_EINVAL, "sfem_num_cables: null argument");
    *n = (int32_t)c->cables.size();
    return SFEM_OK;
}

int sfem_get_cable_positions(sfem_ctx* c, int32_t cable, double* pos)
{
    if (!c || !pos || cable < 0 || (size_t)cable >= c->cables.size())
        return fail(SFEM_EINVAL, "sfem_get_cable_positions: bad argument");
    return guard_call("sfem_get_cable_positions", [&]() -> int {
        const sfem_ctx::Cable& cb = c->cables[cable];
        const size_t n3 = 3 * (size_t)c->nn;
        std::vector<double> X0(n3), u(n3);
        HIPCK(hipStreamSynchronize(c->stream));
        HIPCK(hipMemcpy(X0.data(), c->d_X0.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
        HIPCK(hipMemcpy(u.data(), c->d_u64.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < cb.elem.size(); i++) {
            const double w[4] = {cb.bary[3 * i], cb.bary[3 * i + 1], cb.bary[3 * i + 2],
                                 1.0 - cb.bary[3 * i] - cb.bary[3 * i + 1] - cb.bary[3 * i + 2]};
            for (int q = 0; q < 3; q++) {
                double v = 0.0;
                for (int k = 0; k < 4; k++) {
                    const size_t nd = (size_t)c->tets_h[4 * (size_t)cb.elem[i] + k];
                    v += w[k] * (X0[3 * nd + q] + u[3 * nd + q]);
                }
                pos[3 * i + q] = v;
            }
        }
        return SFEM_OK;
    });
}

int sfem_clear_cables(sfem_ctx* c)
{
    if (!c) return fail(SFEM_EINVAL, "sfem_clear_cables: null context");
    return guard_call("sfem_clear_cables", [&]() -> int {
        c->cables.clear();
        rebuild_cables(c);
        return SFEM_OK;
    });
}

int sfem_block_diagonal(sfem_ctx* c, double* B)
{
    if (!c || !B) return fail(SFEM_EINVAL, "sfem_block_diagonal: null argument");
    return guard_call("sfem_block_diagonal", [&]() -> int {
        const size_t n9 = 9 * (size_t)c->nn;
        std::vector<float> o(n9);
        HIPCK(launch_blockdiag(c->material, c->blk, c->shr, c->d_u32.as<float>(), c->d_bd32.as<float>(),
                               c->d_partial.as<float>(), c->stream));
        HIPCK(hipMemcpyAsync(o.data(), c->d_bd32.p, n9 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < n9; i++) B[i] = o[i];
        return SFEM_OK;
    });
}

int sfem_reset(sfem_ctx* c)
{
    if (!c) return fail(SFEM_EINVAL, "sfem_reset: null context");
    return guard_call("sfem_reset", [&]() -> int {
        HIPCK(hipMemsetAsync(c->d_u64.p, 0, 3 * c->nn * sizeof(double), c->stream));
        HIPCK(hipMemsetAsync(c->d_u32.p, 0, 3 * c->nn * sizeof(float), c->stream));
        HIPCK(hipMemsetAsync(c->d_vel.p, 0, 3 * c->nn * sizeof(double), c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        c->frac = 0.0;
        c->has_prev = c->has_prev2 = false;
        set_time_step(c, 0.0);
        // the multigrid set-up starts again as on a new context
        c->mg_eig_ok = false;
        for (auto& m : c->mg) m->c->mg_eig_ok = false;
        return SFEM_OK;
    });
}

int sfem_profile_hessvec(sfem_ctx* c, int reps, double* ms_elem, double* ms_gather)
{
    if (!c || reps <= 0 || !ms_elem || !ms_gather) return fail(SFEM_EINVAL, "sfem_profile_hessvec: bad argument");
    return guard_call("sfem_profile_hessvec", [&]() -> int {
        // direction: the current displacement (a smooth, realistic field)
        HIPCK(hipMemcpyAsync(c->d_d32.p, c->d_u32.p, 3 * c->nn * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        hipEvent_t e0, e1, e2;
        HIPCK(hipEventCreate(&e0));
        HIPCK(hipEventCreate(&e1));
        HIPCK(hipEventCreate(&e2));
        SharedArgs none = c->shr;
        none.nshared = 0;
        // warm
        HIPCK(launch_hessvec(c->material, c->blk, c->shr, c->d_u32.as<float>(), c->d_d32.as<float>(), c->d_q32.as<float>(),
                             c->d_partial.as<float>(), nullptr, c->stream));
        HIPCK(hipEventRecord(e0, c->stream));
        for (int i = 0; i < reps; i++)
            HIPCK(launch_hessvec(c->material, c->blk, none, c->d_u32.as<float>(), c->d_d32.as<float>(),
                                 c->d_q32.as<float>(), c->d_partial.as<float>(), nullptr, c->stream));
        HIPCK(hipEventRecord(e1, c->stream));
        for (int i = 0; i < reps; i++)
            HIPCK(launch_shared_only(c->shr, c->d_partial.as<float>(), c->d_q32.as<float>(), c->stream));
        HIPCK(hipEventRecord(e2, c->stream));
        HIPCK(hipEventSynchronize(e2));
        float t1 = 0, t2 = 0;
        HIPCK(hipEventElapsedTime(&t1, e0, e1));
        HIPCK(hipEventElapsedTime(&t2, e1, e2));
        *ms_elem = t1 / reps;
        *ms_gather = t2 / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(e2);
        return SFEM_OK;
    });
}

int sfem_set_tangent_cache(sfem_ctx* c, int mode)
{
    if (!c) return fail(SFEM_EINVAL, "sfem_set_tangent_cache: null context");
    if (mode < 0 || mode > 2) return fail(SFEM_EINVAL, "sfem_set_tangent_cache: mode must be 0, 1 or 2");
    c->lin_cache = mode;
    return SFEM_OK;
}

int sfem_set_precond_blocks(sfem_ctx* c, const float* minv9)
{
    if (!c) return fail(SFEM_EINVAL, "sfem_set_precond_blocks: null context");
    return guard_call("sfem_set_precond_blocks", [&]() -> int {
        if (!minv9) { c->has_learned = false; return SFEM_OK; }
        std::vector<float> v(minv9, minv9 + 9 * c->nn);
        for (int64_t i = 0; i < c->nn; i++)
            if (c->fixed_h[i]) std::fill(v.begin() + 9 * i, v.begin() + 9 * i + 9, 0.f);
        c->d_minv_learned.upload(v);
        c->has_learned = true;
        return SFEM_OK;
    });
}

int sfem_solve(sfem_ctx* c, const sfem_solve_params* p, sfem_solve_stats* st)
{
    if (!c || !p || !st) return fail(SFEM_EINVAL, "sfem_solve: null argument");
    if (p->inner_iterations <= 0) return fail(SFEM_EINVAL, "sfem_solve: inner_iterations must be positive");
    std::memset(st, 0, sizeof *st);
    return guard_call("sfem_solve", [&]() -> int {
        HIPCK(hipStreamSynchronize(c->stream));
        auto t0 = std::chrono::steady_clock::now();
        int rc;
        if (p->mode == SFEM_SOLVE_REFERENCE) {
            rc = solve_reference(c, p, st);
        } else {
            const bool adaptive = p->precond == SFEM_PC_LEARNED || p->precond == SFEM_PC_MULTIGRID ||
                                  p->precond == SFEM_PC_MULTIGRID_LEARNED;
            const size_t n3b = 3 * (size_t)c->nn * sizeof(double);
            if (adaptive) HIPCK(hipMemcpyAsync(c->d_ustart.p, c->d_u64.p, n3b, hipMemcpyDeviceToDevice, c->stream));
            c->pred_aborted = false;
            rc = solve_converged(c, p, st);
            if (adaptive && rc == SFEM_ENOTCONV && !c->pred_aborted) {
                // the judge: the learned / multigrid solve's true FP64 residual missed
                // the target; the solve reruns from its start state with Jacobi-PCG
                const sfem_solve_stats first = *st;
                HIPCK(hipMemcpyAsync(c->d_u64.p, c->d_ustart.p, n3b, hipMemcpyDeviceToDevice, c->stream));
                HIPCK(launch_to_f32(3 * c->nn, c->d_u64.as<double>(), c->d_u32.as<float>(), c->stream));
                sfem_solve_params pj = *p;
                pj.precond = SFEM_PC_JACOBI;
                pj.pc_newton_cap = 0;
                std::memset(st, 0, sizeof *st);
                rc = solve_converged(c, &pj, st);
                st->newton_iters += first.newton_iters;
                st->cg_iters += first.cg_iters;
                st->fallbacks += first.fallbacks + 1;
                st->reruns += first.reruns + 1;
                st->hessvec_evals += first.hessvec_evals;
                st->force_evals += first.force_evals;
                st->res0 = first.res0;
            }
        }
        HIPCK(hipStreamSynchronize(c->stream));
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rc == SFEM_ENOTCONV)
            fail(rc, "sfem_solve: residual " + std::to_string(st->res) + " above tol * " + std::to_string(st->ext_norm));
        return rc;
    });
}

int sfem_step(sfem_ctx* c, const sfem_solve_params* p, sfem_solve_stats* st)
{
    if (!c || !p || !st) return fail(SFEM_EINVAL, "sfem_step: null argument");
    // Step (FemPhysicsMatrixFree.cpp:325-340)
    if (p->sim_type == SFEM_ST_STATIC) {
        if (c->frac == 0) {
            c->frac = 1;
            return sfem_solve(c, p, st);
        }
        std::memset(st, 0, sizeof *st);
        return SFEM_OK;
    }
    if (p->sim_type == SFEM_ST_QUASI_STATIC) {
        if (c->frac < 1) {
            const double f_old = c->frac, f_new = std::min(1.0, c->frac + p->force_step);
            int extra = 0;
            bool predicted = false, quadratic = false;
            int rc = guard_call("sfem_step", [&]() -> int {
                const int64_t n3 = 3 * c->nn;
                double* u = c->d_u64.as<double>();
                double* r = c->d_r64.as<double>();
                const int order = p->mode == SFEM_SOLVE_CONVERGED ? p->predictor : 0;
                const bool pred = order > 0 && c->has_prev && f_old > c->frac_prev;
                const bool quad = order >= 2 && pred && c->has_prev2 && c->frac_prev > c->frac_prev2;
                c->frac = f_new;
                if (pred) {
                    // predictor: the secant u + s (u - u_prev) of the last two
                    // states, or the quadratic through the last three (Lagrange
                    // in the load fraction), dropped if it inverts an element
                    // (non-finite residual).  Its residual norm is usually
                    // larger than at u (the stiff modes it misses dominate ||r||)
                    // while its error in the soft modes that Newton-CG converges
                    // slowest is much smaller, so ||r|| is no test of it at the
                    // start; Newton from it must get below ||r(u)|| within
                    // PRED_PATIENCE steps, else the step is redone from u.
                    residual(c, u, r);
                    c->abort_rn = std::sqrt(dots3(c, r, nullptr, nullptr).aa);
                    if (quad) {
                        const double f0 = c->frac_prev2, f1 = c->frac_prev, f2 = f_old, f = f_new;
                        const double l0 = (f - f1) * (f - f2) / ((f0 - f1) * (f0 - f2));
                        const double l1 = (f - f0) * (f - f2) / ((f1 - f0) * (f1 - f2));
                        const double l2 = (f - f0) * (f - f1) / ((f2 - f0) * (f2 - f1));
                        HIPCK(launch_extrapolate2(n3, u, c->d_uprev.as<double>(), c->d_uprev2.as<double>(), l2, l1, l0,
                                                  c->d_u32.as<float>(), c->stream));
                    } else {
                        HIPCK(hipMemcpyAsync(c->d_uprev2.p, c->d_uprev.p, n3 * sizeof(double), hipMemcpyDeviceToDevice,
                                             c->stream));
                        HIPCK(launch_extrapolate(n3, u, c->d_uprev.as<double>(), (f_new - f_old) / (f_old - c->frac_prev),
                                                 c->d_u32.as<float>(), c->stream));
                    }
                    residual(c, u, r);
                    const double r1 = dots3(c, r, nullptr, nullptr).aa;
                    extra = 2;
                    predicted = std::isfinite(r1);
                    quadratic = quad && predicted;
                    c->abort_after = predicted ? PRED_PATIENCE : 0;
                    if (!predicted) {
                        HIPCK(hipMemcpyAsync(u, c->d_uprev.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
                        HIPCK(launch_to_f32(n3, u, c->d_u32.as<float>(), c->stream));
                    }
                } else {
                    HIPCK(hipMemcpyAsync(c->d_uprev2.p, c->d_uprev.p, n3 * sizeof(double), hipMemcpyDeviceToDevice,
                                         c->stream));
                    HIPCK(hipMemcpyAsync(c->d_uprev.p, u, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
                }
                c->frac_prev2 = c->frac_prev;
                c->has_prev2 = c->has_prev;
                c->frac_prev = f_old;
                c->has_prev = true;
                return SFEM_OK;
            });
            if (rc != SFEM_OK) return rc;
            rc = sfem_solve(c, p, st);
            c->abort_after = 0;
            // adds the work of an abandoned attempt to the stats of the next one
            auto add_attempt = [](sfem_solve_stats* st, const sfem_solve_stats& first) {
                st->newton_iters += first.newton_iters;
                st->cg_iters += first.cg_iters;
                st->fallbacks += first.fallbacks;
                st->rejects += first.rejects + 1;
                st->reruns += first.reruns;
                st->seconds += first.seconds;
                st->hessvec_evals += first.hessvec_evals;
                st->force_evals += first.force_evals;
                st->res0 = first.res0;
            };
            if (quadratic && rc == SFEM_ENOTCONV && c->pred_aborted) {
                // the quadratic prediction was abandoned (it can overshoot into a
                // state where the tangent is not positive definite): try the secant
                // one u + s (u - u_prev) from the same history before the last
                // state.  An attempt that used up its Newton steps is not retried
                // from another prediction (Newton stalls there from any start).
                const sfem_solve_stats first = *st;
                bool finite = false;
                rc = guard_call("sfem_step", [&]() -> int {
                    const int64_t n3 = 3 * c->nn;
                    const double sc = (f_new - c->frac_prev) / (c->frac_prev - c->frac_prev2);
                    HIPCK(launch_lincomb(n3, c->d_uprev.as<double>(), c->d_uprev2.as<double>(), 1.0 + sc, -sc,
                                         c->d_u64.as<double>(), c->d_u32.as<float>(), c->stream));
                    residual(c, c->d_u64.as<double>(), c->d_r64.as<double>());
                    finite = std::isfinite(dots3(c, c->d_r64.as<double>(), nullptr, nullptr).aa);
                    return SFEM_OK;
                });
                if (rc != SFEM_OK) return rc;
                extra++;
                if (finite) {
                    c->abort_after = PRED_PATIENCE;
                    rc = sfem_solve(c, p, st);
                    c->abort_after = 0;
                    add_attempt(st, first);
                } else {
                    *st = first;
                    rc = SFEM_ENOTCONV;
                }
            }
            if (predicted && rc == SFEM_ENOTCONV) {
                // the solve from the predicted state failed: redo the step from the
                // last state (what Step does), adding up the work of both attempts
                const sfem_solve_stats first = *st;
                rc = guard_call("sfem_step", [&]() -> int {
                    HIPCK(hipMemcpyAsync(c->d_u64.p, c->d_uprev.p, 3 * c->nn * sizeof(double), hipMemcpyDeviceToDevice,
                                         c->stream));
                    HIPCK(launch_to_f32(3 * c->nn, c->d_u64.as<double>(), c->d_u32.as<float>(), c->stream));
                    return SFEM_OK;
                });
                if (rc != SFEM_OK) return rc;
                rc = sfem_solve(c, p, st);
                add_attempt(st, first);
            }
            st->force_evals += (int64_t)extra * c->ne;
            return rc;
        }
        std::memset(st, 0, sizeof *st);
        return SFEM_OK;
    }
    if (p->sim_type == SFEM_ST_EXPLICIT) {
        // Step for ST_EXPLICIT (FemPhysicsMatrixFree.cpp:316-323) and SubStep (:371-387)
        if (!(p->time_step > 0) || p->substeps < 1) return fail(SFEM_EINVAL, "sfem_step: explicit needs time_step > 0, substeps >= 1");
        std::memset(st, 0, sizeof *st);
        return guard_call("sfem_step", [&]() -> int {
            HIPCK(hipStreamSynchronize(c->stream));
            auto t0 = std::chrono::steady_clock::now();
            c->frac = 1.0;
            c->has_prev = c->has_prev2 = false;
            const double h = p->time_step / p->substeps;
            double* part = reinterpret_cast<double*>(c->d_partial.as<char>() + (size_t)c->nrows_partial * 9 * sizeof(float));
            BlockArgs a = c->blk;   // elastic forces only (ElasticEnergy::ComputeForces)
            a.mdiag32 = nullptr;
            a.mdiag64 = nullptr;
            for (int k = 0; k < p->substeps; k++) {
                // SubStep (FemPhysicsMatrixFree.cpp:371-387): elastic plus cable spring forces
                if (c->cab.n)
                    HIPCK(launch_cable_forces(c->cab, c->d_X0.as<double>(), c->d_u64.as<double>(), c->d_vel.as<double>(),
                                              c->d_cab_fs.as<double>(), c->frac, c->d_fixed.as<uint8_t>(),
                                              c->d_fcable.as<double>(), 0, c->stream));
                HIPCK(launch_force(c->material, a, c->shr, c->d_u64.as<double>(), c->d_f64tmp.as<double>(), part,
                                   c->cab.n ? c->d_fcable.as<double>() : nullptr, 0, c->stream));
                HIPCK(launch_explicit(c->nn, c->d_u64.as<double>(), c->d_vel.as<double>(), c->d_f64tmp.as<double>(),
                                      c->d_mass.as<double>(), c->d_fixed.as<uint8_t>(), c->gravity, h, c->d_u32.as<float>(),
                                      c->stream));
            }
            HIPCK(hipStreamSynchronize(c->stream));
            st->force_evals = (int64_t)p->substeps * c->ne;
            st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            return SFEM_OK;
        });
    }
    if (p->sim_type == SFEM_ST_IMPLICIT) {
        // Step for ST_IMPLICIT (FemPhysicsMatrixFree.cpp:299-314): x0 = x at the start
        // of the step; `substeps` solves with M / h^2 in the operator and
        // m/h v - m/h^2 (x - x0) in the residual; v = (x - x0) / h after each
        if (!(p->time_step > 0) || p->substeps < 1) return fail(SFEM_EINVAL, "sfem_step: implicit needs time_step > 0, substeps >= 1");
        const double h = p->time_step / p->substeps;
        int rc = guard_call("sfem_step", [&]() -> int {
            const int64_t n3 = 3 * c->nn;
            HIPCK(hipMemcpyAsync(c->d_u0step.p, c->d_u64.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            set_time_step(c, h);
            c->frac = 1.0;
            c->has_prev = c->has_prev2 = false;
            return SFEM_OK;
        });
        if (rc != SFEM_OK) return rc;
        sfem_solve_stats tot{};
        for (int k = 0; k < p->substeps; k++) {
            rc = sfem_solve(c, p, st);
            if (rc != SFEM_OK && rc != SFEM_ENOTCONV) return rc;
            const int rc2 = guard_call("sfem_step", [&]() -> int {
                HIPCK(launch_implicit_vel(c->nn, c->d_u64.as<double>(), c->d_u0step.as<double>(), c->d_fixed.as<uint8_t>(),
                                          1.0 / h, c->d_vel.as<double>(), c->stream));
                HIPCK(hipStreamSynchronize(c->stream));
                return SFEM_OK;
            });
            if (rc2 != SFEM_OK) return rc2;
            tot.newton_iters += st->newton_iters;
            tot.cg_iters += st->cg_iters;
            tot.fallbacks += st->fallbacks;
            tot.reruns += st->reruns;
            tot.rejects += st->rejects;
            tot.seconds += st->seconds;
            tot.hessvec_evals += st->hessvec_evals;
            tot.force_evals += st->force_evals;
            if (k == 0) tot.res0 = st->res0;
            tot.res = st->res;
            tot.ext_norm = st->ext_norm;
            tot.converged = st->converged;
        }
        *st = tot;
        return rc;
    }
    return fail(SFEM_EINVAL, "sfem_step: unknown simulation type");
}

}  // extern "C"
