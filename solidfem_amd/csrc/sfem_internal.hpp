// sfem_internal.hpp -- context layout and helpers shared by the host sources
// of libsfem_hip.so (not part of the public C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "sfem_capi.h"
#include "sfem_host.hpp"
#include "sfem_kernels.hpp"


namespace sfem {

// records the message returned by sfem_last_error() on this thread
int fail(int code, const std::string& msg);

struct HipError {
    hipError_t e;
    const char* what;
};

#define HIPCK(x)                                            \
    do {                                                    \
        hipError_t e_ = (x);                                \
        if (e_ != hipSuccess) throw HipError{e_, #x};       \
    } while (0)

struct DevBuf {   // owning device allocation
    void* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t bytes)
    {
        release();
        if (bytes == 0) bytes = 16;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            throw HipError{hipErrorOutOfMemory, "hipMalloc"};
        }
        n = bytes;
    }
    template <typename T> void upload(const std::vector<T>& v)
    {
        alloc(v.size() * sizeof(T));
        if (!v.empty()) HIPCK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace sfem

struct sfem_gnn;

using sfem::BlockArgs;
using sfem::DevBuf;
using sfem::SharedArgs;

// residual/solution training pairs collected by the solver at one state
struct SfemPairs {
    int64_t n = 0;                                  // number of pairs
    std::vector<std::unique_ptr<sfem::DevBuf>> r;   // FP32 residual vectors (nn*3)
    std::vector<double> er;                         // r . e per pair (FP64)
    sfem::DevBuf d_part;                            // the recorder's partial sums
};

// a coarse level of the multigrid preconditioner: its own context (mesh,
// partition, state, block preconditioner), which runs on the stream of the
// finest level's context, and the transfer operators to the level above
struct MgLevel {
    sfem_ctx* c = nullptr;
    sfem::DevBuf inj;                 // per node: the node of the level above at its position
    sfem::DevBuf p_ptr, p_col, p_val; // prolongation rows (nodes above) -> nodes here, FP32 weights
    sfem::DevBuf t_ptr, t_row, t_val; // its transpose (restriction), rows = nodes here
    double lmax = 0.0;                // largest eigenvalue estimate of M A on this level
    // direct solve when this level is the coarsest and 3 nn <= MG_DENSE_MAX:
    // element table, Dm^-1 (9 per element) and V mu, V lambda (2 per element)
    // in element order, node -> (element, corner) incidences, element
    // matrices, the dense operator and its inverse
    bool dense = false;
    // dense coarsest level: operator A, inverse Ainv (FP32), and the FP64
    // working matrix W and scratch C, S of its blocked inverse
    sfem::DevBuf d_tets4, d_dminv, d_cvol, d_inc_ptr, d_inc, d_Ke, d_A, d_Ainv, d_W, d_C, d_S;
    ~MgLevel();
};

struct sfem_ctx {
    int64_t ne = 0, nn = 0, nfixed = 0;
    int device = 0;          // the HIP device current at creation (the context's device)
    int64_t nslots = 0;      // partial-row slots of the nodes shared by several blocks
    int material = 0;
    // the material as given (coarse multigrid levels inherit a uniform one)
    double young = 0.0, poisson = 0.0, density = 0.0;
    bool per_element = false;
    double gravity[3] = {0, 0, 0};
    double frac = 0.0;
    double tstep = 0.0;   // implicit dynamics substep h; 0 = static (FemPhysicsMatrixFree::mTimeStep)
    // load-step predictor history: the converged displacement of the previous
    // quasi-static step (d_uprev) and its load fraction,
    double frac_prev = 0.0;
    bool has_prev = false;
    // the state before it (d_uprev2, frac_prev2): the quadratic predictor
    double frac_prev2 = 0.0;
    bool has_prev2 = false;
    // a predicted solve stops after abort_after Newton steps if ||r|| > abort_rn
    // (pred_aborted records that it did: the judge does not rerun such a solve,
    // the step is redone from the last state instead)
    int abort_after = 0;
    double abort_rn = 0.0;
    bool pred_aborted = false;
    double pressure = 0.0;
    int64_t ntri = 0;
    bool has_fdead = false;
    bool has_learned = false;
    hipStream_t stream = nullptr;
    bool owns_stream = true;   // false for multigrid levels (the finest level's stream)

    // host copies kept for preprocessing
    std::vector<uint8_t> fixed_h;
    std::vector<int32_t> tets_h;   // element table (node graph of the learned preconditioner)

    // partition
    BlockArgs blk{};
    SharedArgs shr{};
    int nblk = 0;
    int64_t nrows_partial = 0;
    DevBuf d_rel32, d_mat32, d_bnode_ptr, d_bnodes, d_lptr, d_lperm, d_lidx, d_dminv32, d_coef32, d_dminv64, d_coef64;
    DevBuf d_sh_node, d_sh_ptr, d_bslot, d_node_info;
    DevBuf d_partial;        // nrows_partial * 9 floats (also used as FP64 x3)

    // node data
    DevBuf d_ustart, d_uprev, d_uprev2, d_vel, d_u0step, d_mdiag32, d_mdiag64, d_eorig, d_X0, d_mass, d_fixed, d_fdead;
    DevBuf d_u64, d_ut64, d_u32;
    DevBuf d_r64, d_ext64, d_f64tmp;
    DevBuf d_b32, d_x32, d_r32, d_z32, d_p32, d_q32, d_d32, d_minv, d_minv_learned, d_bd32;
    DevBuf d_p32b;           // the second CG direction buffer (the direction update writes the other one)
    DevBuf d_tris, d_pptr, d_pidx, d_ftri;

    // reductions
    DevBuf d_scal, d_partials, d_done;

    // cable springs (FemPhysicsBase::mCables): the cables as added, and their
    // concatenation on the device (CableArgs); d_cab_fs = spring forces,
    // d_fcable = cable node forces on the mesh for the explicit integrator
    struct Cable {
        std::vector<int> elem;       // element (original index) of each cable node
        std::vector<double> bary;    // 3 barycentric weights per cable node
        double rest = 0.0, stiffness = 0.0, damping = 0.0, actuation = 1.0;
    };
    std::vector<Cable> cables;
    sfem::CableArgs cab{};
    DevBuf d_cab_node, d_cab_w, d_cab_spring, d_cab_touched, d_cab_tptr, d_cab_tent, d_cab_fs, d_fcable;

    // learned preconditioner: node graph (CSR, no self loops), network, pairs
    DevBuf d_adj_ptr, d_adj_idx;
    int64_t nadj = 0;
    sfem_gnn* gnn = nullptr;
    SfemPairs pairs;

    // multigrid preconditioner: coarse levels (finest first), the finest
    // level's smoother blocks and work vectors, power-iteration norms
    std::vector<std::unique_ptr<MgLevel>> mg;
    const float* mg_minv0 = nullptr;
    double mg_lmax0 = 0.0;
    bool mg_ready = false;     // eigenvalue estimates and coarsest inverse exist
    int64_t mg_evals = 0;      // element evaluations of one V-cycle
    DevBuf d_mg_r, d_mg_d, d_mg_q, d_mg_part, d_mg_norm;
    // the V-cycle's launch schedule: the Chebyshev steps and the prolongation
    // folded into the element kernels (true), or one kernel per step; the
    // results are bitwise the same
    bool mg_fused = true;
    // Neo-Hookean static solves: the tangent's state per element at each
    // Newton step's displacements (BlockArgs::lin32), built when its linear
    // solves begin; sfem_set_tangent_cache: 0 off, 1 in solves, 2 also in
    // sfem_matvec (built at each call)
    int lin_cache = 1;
    DevBuf d_lin;
    // this context as a smoothed multigrid level: the last vector of its
    // power iteration and that vector's norm, the next full set-up's start
    // (mg_eig_ok false: the hashed start)
    DevBuf d_mg_eig;
    double mg_eig_norm = 0.0;
    bool mg_eig_ok = false;
    // this context as a multigrid level under the folded schedule: a second
    // residual and direction, and a second operator image (rows and partial
    // slots)
    DevBuf d_mgw_r2, d_mgw_d2, d_mgw_qb, d_mgw_pb;
    // the multigrid PCG's test outcomes (pinned host memory mapped into the
    // device, two slots: h_slot the host's address, d_slot the device's) and
    // the events after them
    sfem::CgSlot* h_slot = nullptr;
    sfem::CgSlot* d_slot = nullptr;
    hipEvent_t flag_ev[2] = {};
    // the coarsest level's inverse is computed on a second stream (default
    // priority) beside the other levels' set-up: [0] coarse state ready,
    // [1] inverse ready
    hipStream_t mg_side = nullptr;
    hipEvent_t mg_ev[2] = {};

    // sampled timing of the CG loops' Hessian-vector launches (sfem_time_hessvec)
    static constexpr int TM_POOL = 64;
    hipEvent_t tm_ev[2 * TM_POOL] = {};
    int tm_iter[TM_POOL] = {};
    int tm_every = 0, tm_pending = 0;
    int64_t tm_counter = 0, tm_samples = 0;
    double tm_ms = 0.0;

    ~sfem_ctx()
    {
        for (hipEvent_t e : tm_ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : flag_ev)
            if (e) (void)hipEventDestroy(e);
        if (h_slot) (void)hipHostFree(h_slot);
        if (mg_side) (void)hipStreamSynchronize(mg_side);
        mg.clear();
        for (hipEvent_t e : mg_ev)
            if (e) (void)hipEventDestroy(e);
        if (mg_side) (void)hipStreamDestroy(mg_side);
        if (stream && owns_stream) (void)hipStreamDestroy(stream);
    }
};


namespace sfem {

template <typename F> int guard_call(const char* where, F&& f)
{
    try {
        return f();
    } catch (const HipError& e) {
        return fail(e.e == hipErrorOutOfMemory ? SFEM_ENOMEM : SFEM_EDEVICE,
                    std::string(where) + ": " + e.what + ": " + hipGetErrorString(e.e));
    } catch (const std::exception& e) {
        return fail(SFEM_EDEVICE, std::string(where) + ": " + e.what());
    }
}

// sfem_matvec_noalloc (sfem_noalloc.hip): the per-device scratch array's
// size, the bytes one call needs for context c, and the once-per-device
// lookup of the array's address (at context creation)
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(16) << 20;
size_t noalloc_matvec_bytes(const sfem_ctx* c);
int noalloc_init(int device);

// operations on a context (sfem_capi.cpp)
void hessvec(sfem_ctx* c, const float* d, float* out, const int* done);
void build_jacobi(sfem_ctx* c);
struct Dots {
    double aa, bb, da;
};
Dots dots3(sfem_ctx* c, const double* a, const double* b, const float* d);
struct CgResult {
    int iters;
    int status;   // 1 converged, 0 max iterations, 3 breakdown, 4 NaN
    double rr0, rr;   // r0.r0 and the last r.r (pcg; the multigrid PCG leaves them 0)
};
// on_check(iters) runs after every check_every iterations while not done
typedef void (*CgHook)(sfem_ctx* c, int iters, void* user);
CgResult pcg(sfem_ctx* c, const float* minv, int pc, int maxit, int test_kind, double tol2, int guard, int check_every,
             int64_t* hv_evals, CgHook hook = nullptr, void* user = nullptr);

// learned preconditioner (sfem_gnn_host.cpp): minv_learned from the network at
// the current state (needs the Jacobi blocks in d_bd32)
void gnn_install(sfem_ctx* c);
void ensure_graph(sfem_ctx* c);

// multigrid preconditioner (sfem_mg.cpp): level operators at the current
// state with minv0 as the finest level's smoother blocks, then PCG on
// K x = b (b in d_b32, x in d_x32) with one V-cycle per iteration.  Without
// `full`, the Chebyshev bands and the coarsest level's inverse of the last
// full set-up are kept (the Newton steps of one load step share them); the
// states and block preconditioners are always the current ones.  `warm`:
// the eigenvalue estimates start from the previous warm set-up's vectors
// (Newton solves; the test entry points start from the hashed vector).
// False when a level's operator is not finite at this state (no V-cycle
// possible).
bool mg_prepare(sfem_ctx* c, const float* minv0, bool full, bool warm = false);
CgResult mg_pcg(sfem_ctx* c, int maxit, double tol2, int64_t* hv_evals);
// the CG test-outcome slots (mapped pinned memory) and their events, once per context
void ensure_cg_slots(sfem_ctx* c);
// z = V(r) on the finest level (r in d_r32, z in d_z32), after mg_prepare
void mg_vcycle(sfem_ctx* c);

}  // namespace sfem
