// sfem_kernels.hpp -- device data layout shared by the kernels and the host.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfem {

// elements per workgroup of the element-block kernels
#ifndef SFEM_BLOCK_E
#define SFEM_BLOCK_E 256
#endif
constexpr int BLOCK_E = SFEM_BLOCK_E;
// cap on workgroups of the reduction kernels (their partials buffer size)
constexpr int RED_GRID = 2048;   // reduction partials buffer holds 3 * RED_GRID doubles
// workgroups of the CG vector kernels; each of them re-reduces the previous
// kernel's partials, so this stays small
#ifndef SFEM_CG_GRID
#define SFEM_CG_GRID 1024
#endif
constexpr int CG_GRID = SFEM_CG_GRID;
// threads per workgroup of the CG vector kernels (one node per thread)
#ifndef SFEM_CG_BLOCK
#define SFEM_CG_BLOCK 256
#endif
constexpr int CG_BLOCK = SFEM_CG_BLOCK;
#ifndef SFEM_CG_P_BLOCK
#define SFEM_CG_P_BLOCK 256
#endif
constexpr int CG_P_BLOCK = SFEM_CG_P_BLOCK;

// FP64 scalars kept on the device during a CG solve
// S_RR_BEST / S_IT_BEST: the stagnation test's reference r.r and the iteration
// it was reached at (read and written by one thread of k_cg_p only); S_SL..+2:
// the Newton line search's slope at its start (three dots3 slots)
enum Scal {
    S_RZ = 0, S_RZ0, S_RR, S_RR0, S_PQ, S_ALPHA, S_BETA, S_ITER, S_NT, S_NT2, S_NT3, S_RZ_PREV, S_RR_BEST, S_IT_BEST,
    S_SL, S_COUNT = 20
};
// CG status in the done flag: 0 running / iteration limit, 1 converged,
// 3 breakdown (p.Kp <= 0, or r.z <= 0: operator or preconditioner not
// positive definite), 4 not finite, 5 stagnation (r.r not halved within the
// stagnation window)
enum CgStatus { CG_RUNNING = 0, CG_CONVERGED = 1, CG_BREAKDOWN = 3, CG_NOT_FINITE = 4, CG_STAGNATED = 5 };

// Element-block partition, all device pointers.
// the material id of coarse multigrid levels of a Neo-Hookean context (the
// safeguarded tangent, sfem_device.hpp); the C-ABI ids 0 .. 3 are the kernels'
// ids, SFEM_COROTATIONAL_POLAR (16) is MAT_COROTATIONAL_POLAR
constexpr int MAT_NEO_HOOKEAN_COARSE = 4;
constexpr int MAT_COROTATIONAL_POLAR = 5;

struct BlockArgs {
    int64_t ne;            // elements
    int64_t ne_pad;        // SoA stride (nblk * BLOCK_E)
    int nblk;              // workgroups
    int nmax;              // max nodes of any block (LDS sizing)
    const int* bnode_ptr;  // nblk+1 offsets into bnodes
    const uint32_t* bnodes;// block node list: node id | FLAG_SHARED | FLAG_FIXED
    const int* bslot;      // per block node row: its partial slot (shared rows only)
    const uint16_t* lptr;  // per block nb+1 offsets into its incidence list
    const uint2* lidx;     // per element 4 block-local node ids (16 bit each)
    const uint2* lperm;    // per element, per corner its slot in the block's node-sorted incidences (16 bit each)
    const float* dminv32;  // 9 x ne_pad, Dm^-1 row-major, FP32
    const float* coef32;   // 2 x ne_pad: V mu, V lambda, FP32
    const double* dminv64; // 9 x ne_pad, FP64
    const double* coef64;  // 2 x ne_pad, FP64
    // Hessian-vector kernel: Dm^-1 and V are recomputed in FP32 from the rest
    // coordinates of the block rows, relative to the block's centre
    const float* rel32;    // 3 per block row
    const float* mat32;    // 2 x ne_pad: mu, lambda per element; null when uniform
    // Neo-Hookean Hessian-vector kernel: 10 x ne_pad floats, the tangent's
    // state at the current displacements (E = Dm^-1 F^-1 row-major, then
    // V (mu - lam ln J)), value k of element b * BLOCK_E + t at
    // (b * 10 + k) * BLOCK_E + t; set only while the linear solves of one
    // Newton step run, null otherwise
    const float* lin32;
    float mu0, lam0;       // uniform mu, lambda
    // implicit dynamics: m / h^2 per node (null when static) and each node's
    // partial-slot range (the first slot's block adds the mass term)
    const float* mdiag32;
    const double* mdiag64;
    const int2* node_q;
    const int* zero_flag;  // a device int that stays 0: the done flag of launches outside CG
    // internal-force kernel only: 0 = the element's (V mu, V lambda); 1 = (V, 0)
    // and 2 = (0, V), the force's derivatives in mu and in lambda (P is linear in both)
    int coef_mode;
};

// Cable springs (FemPhysicsBase::ComputeSpringForces, FemPhysicsBase.cpp:214-309):
// cable nodes embedded in elements by barycentric weights, one spring between
// consecutive nodes of a cable; all cables concatenated
struct CableArgs {
    int n;                  // cable nodes
    const int4* node;       // per cable node: the 4 mesh nodes of its element
    const double* w;        // per cable node: 4 barycentric weights
    const double* spring;   // per cable node j: spring j -> j+1 as (rest length, actuation, stiffness, damping);
                            // stiffness 0 where j is the last node of its cable or the cable is not actuated
    int ntouched;           // mesh nodes that carry cable forces
    const int* touched;     // their ids
    const int* tptr;        // ntouched+1 offsets into tent
    const int2* tent;       // (cable node, index into w) per attachment, cable-node order
};

// nodes shared by several blocks: partial rows to sum
struct SharedArgs {
    int nshared;
    const uint32_t* node;  // node id | flags
    const int* ptr;        // nshared+1: node i owns partial slots [ptr[i], ptr[i+1]), block order
};

// q = K d; with dot_out, also p.q partials: dot_elem (nblk doubles) then
// dot_out (gather_grid(s) doubles)
hipError_t launch_hessvec(int mat, const BlockArgs& a, const SharedArgs& s, const float* u32, const float* d32, float* out,
                          float* partial, const int* done, hipStream_t st, double* dot_elem = nullptr,
                          double* dot_out = nullptr);
int gather_grid(const SharedArgs& s);
int cg_grid(int64_t nn);
hipError_t launch_shared_only(const SharedArgs& s, const float* partial, float* out, hipStream_t st);
hipError_t launch_force(int mat, const BlockArgs& a, const SharedArgs& s, const double* u64, double* out, double* partial,
                        const double* ext, int mask_fixed, hipStream_t st);
hipError_t launch_blockdiag(int mat, const BlockArgs& a, const SharedArgs& s, const float* u32, float* out, float* partial,
                            hipStream_t st);
// the Neo-Hookean tangent's state at the displacements u32 into lin
// (10 x ne_pad floats, BlockArgs::lin32)
hipError_t launch_lin_state(const BlockArgs& a, const float* u32, float* lin, hipStream_t st);
size_t elem_lds_max(const BlockArgs& a);
hipError_t set_elem_lds_limits(const BlockArgs& a);
// p2 (may be null): a second direction buffer that also gets p = z
hipError_t launch_cg_init(int pc, int64_t n3, const float* b, float* x, float* r, float* z, float* p, const float* minv,
                          const uint8_t* fixed, double* scal, double* part, int* done, hipStream_t st, float* p2 = nullptr);
// CG operator apply: element blocks only (single-block rows -> out, shared
// rows -> partial slots) plus one d.(K d) partial per element block
// with events ev0/ev1: they record the dispatch's start and end (timing)
hipError_t launch_hessvec_cg(int mat, const BlockArgs& a, const float* u32, const float* d32, float* out, float* partial,
                             const int* done, double* dot_elem, hipStream_t st, hipEvent_t ev0 = nullptr,
                             hipEvent_t ev1 = nullptr);
// The CG direction update folded into the next Hessian-vector apply: every
// block reduces the update's r.z, r.r partials (the fixed order of the
// separate direction kernel), makes the convergence test, forms
// p = z + beta p_old for its staged rows and applies K to that p; each row's
// owner (its only block, or the first block of a shared node) stores p into
// p_new, the other buffer of a pair, while other blocks still read p_old.
// The outcome of a CG test, written by the kernel that makes it (the
// direction kernel, or the element kernel with the direction update) into
// pinned host memory (a device pointer to it), so the host reads it after the
// kernel's completion event without a copy: the flag after the test, the
// iterations counted (S_ITER) and r.r / r0.r0 (0 once the flag was already set)
struct CgSlot {
    int done;
    int iter;
    double rr_rel;
};
struct PFuse {
    const float* z;          // z, or r without a preconditioner
    const float* p_old;
    float* p_new;
    const double* rz_part;   // 2 x n_rz partials (r.z, then r.r) of the update
    int n_rz;
    double* scal;
    int test_kind;           // as launch_cg_p_only
    double tol2;
    int rz_guard, stall_window;
    CgSlot* slot;            // the test's outcome (may be null)
};
hipError_t launch_hessvec_cg_fused(int mat, const BlockArgs& a, const float* u32, const PFuse& pf, float* out,
                                   float* partial, int* done, double* dot_elem, hipStream_t st, hipEvent_t ev0 = nullptr,
                                   hipEvent_t ev1 = nullptr);
// A multigrid smoothing step folded into the level's next operator apply:
// every block forms the direction d of each staged row itself and applies K
// to it; the row's owner (its only block, or the first block of a shared
// node) stores the node's new vectors.  Reads and writes go to different
// buffers wherever another block may still read the old value.
//   MG_STAGE_CHEB: q = the node's complete image of the previous apply (qa
//     rows and pa partial slots by node_q); r = r_old - q; d = c1 d_old +
//     c2 M r (c2 M r without has_dold); the owner stores r_new, d_new and
//     x += d.
//   MG_STAGE_PROLONG: d = e = P xc at the row (0 at fixed rows); the owner
//     stores x += e.
enum MgStageKind { MG_STAGE_CHEB = 2, MG_STAGE_PROLONG = 3 };
struct MgStage {
    const float* r_old;
    float* r_new;
    const float* d_old;
    float* d_new;
    float* x;
    const float* minv;
    const float* qa;
    const float* pa;
    float c1, c2;
    int has_dold;
    const int* p_ptr;
    const int* p_col;
    const float* p_val;
    const float* xc;
};
// done: the PCG's flag (the block returns at once when it is set), or null
hipError_t launch_hessvec_mg(int mat, const BlockArgs& a, const float* u32, int stage, const MgStage& mg, float* out,
                             float* partial, const int* done, hipStream_t st);
// alpha, x, r and z (pc: z = M r), r.z and r.r partials (2 x cg_grid); guard
// 1: stop at p.Kp <= 0
hipError_t launch_cg_update(int pc, int64_t n3, float* x, float* r, float* z, const float* p, const float* q, const float* minv,
                            const int2* node_q, const float* partial, double* scal, const double* pq_part, int n_pq,
                            double* part, int* done, int guard, hipStream_t st);
// beta, the convergence test and p from r.z, r.r partials (2 x cg_grid)
hipError_t launch_cg_test(int pc, int64_t n3, const float* z, const float* r, float* p, double* scal, const double* part,
                          int* done, int test_kind, double tol2, int rz_guard, int stall_window, hipStream_t st,
                          CgSlot* slot = nullptr);
hipError_t launch_cg_update_only(int64_t n3, float* x, float* r, const float* p, const float* q, const int2* node_q,
                                 const float* partial, double* scal, const double* pq_part, int n_pq, double* part,
                                 int* done, int guard, hipStream_t st);
// the update of the multigrid PCG with the V-cycle's first Chebyshev step on
// the new residual: alpha, x, r, then z = c2 M r into z and z2 (no partials)
hipError_t launch_cg_update_mg(int64_t n3, float* x, float* r, float* z, float* z2, const float* p, const float* q,
                               const float* minv, float c2, const int2* node_q, const float* partial, double* scal,
                               const double* pq_part, int n_pq, int* done, int guard, hipStream_t st);
// rz_guard: stop with CG_BREAKDOWN at r.z <= 0; stall_window > 0: stop with
// CG_STAGNATED when r.r has not fallen below half its last reference value
// within that many iterations; slot (may be null): the test's outcome
hipError_t launch_cg_p_only(int64_t n3, const float* z, const float* r, float* p, double* scal, const double* part,
                            int* done, int test_kind, double tol2, int rz_guard, int stall_window, hipStream_t st,
                            CgSlot* slot = nullptr);
hipError_t launch_cg_scalars(const double* part, int n, double* scal, int* done, hipStream_t st);
hipError_t launch_dots3(int64_t n3, const double* a, const double* b, const float* d, double* scal, int slot, double* part,
                        hipStream_t st);
hipError_t launch_axpy_state(int64_t n3, const double* u, const float* dx, double alpha, double* ut, float* u32, hipStream_t st);
hipError_t launch_explicit(int64_t nn, double* u, double* vel, const double* f, const double* mass, const uint8_t* fixed,
                           const double* g, double h, float* u32, hipStream_t st);
// P per element at displacement u, 9 doubles at the element's original index
hipError_t launch_elem_stress(int mat, const BlockArgs& a, const double* u, const int* eorig, double* out, hipStream_t st);
hipError_t launch_extrapolate(int64_t n3, double* u, double* uprev, double sc, float* u32, hipStream_t st);
// u = a x + b y and its FP32 copy u32
hipError_t launch_lincomb(int64_t n3, const double* x, const double* y, double a, double b, double* u, float* u32,
                          hipStream_t st);
// u <- a u + b uprev + c uprev2, then uprev2 <- uprev, uprev <- u (old)
hipError_t launch_extrapolate2(int64_t n3, double* u, double* uprev, double* uprev2, double a, double b, double c, float* u32,
                               hipStream_t st);
hipError_t launch_to_f32(int64_t n, const double* a, float* b, hipStream_t st);
// ext = frac (m g + f_dead) + pressure (+ m/h v + m/h^2 u0 with vel, u0: implicit dynamics)
hipError_t launch_external(int64_t nn, const double* mass, const double* fdead, const double* g, double frac, const int* pptr,
                           const int* pidx, const double* ftri, const uint8_t* fixed, const double* vel, const double* u0,
                           double invh, double* ext, hipStream_t st);
hipError_t launch_mass_diag(int64_t nn, const double* mass, const uint8_t* fixed, double invh2, float* m32, double* m64,
                            hipStream_t st);
hipError_t launch_implicit_vel(int64_t nn, const double* u, const double* u0, const uint8_t* fixed, double invh, double* vel,
                               hipStream_t st);
hipError_t launch_pressure(int64_t ntri, const int* tris, const double* X0, const double* u, double pressure, double frac,
                           double* ftri, hipStream_t st);
// sum_e V_e Psi(F_e) at displacement u (FP64) -> scal[slot]
hipError_t launch_energy(int mat, const BlockArgs& a, const double* u, double* scal, int slot, double* part, hipStream_t st);
// frac sum over free nodes of (m g + f_dead) . u -> scal[slot] (fdead may be null)
hipError_t launch_dead_work(int64_t nn, const double* mass, const double* fdead, const double* g, double frac,
                            const uint8_t* fixed, const double* u, double* scal, int slot, double* part, hipStream_t st);
// spring forces f_s (3 per cable node, FP64) at x = X0 + u, velocities vel;
// then out[node] (+)= frac * sum_j w_j (f_s[j-1] - f_s[j]) over the touched
// free mesh nodes (accumulate: add to out, else overwrite those rows)
hipError_t launch_cable_forces(const CableArgs& c, const double* X0, const double* u, const double* vel, double* fs,
                               double frac, const uint8_t* fixed, double* out, int accumulate, hipStream_t st);
// spring potentials (FemPhysicsMatrixFree::ComputeSpringEnergy, FemPhysicsMatrixFree.cpp:1062-1111) -> scal[slot]
hipError_t launch_cable_energy(const CableArgs& c, const double* X0, const double* u, double* scal, int slot, double* part,
                               hipStream_t st);
hipError_t launch_invert_blocks(int64_t nn, const float* B, const uint8_t* fixed, float* Minv, hipStream_t st);
hipError_t launch_identity_blocks(int64_t nn, const uint8_t* fixed, float* Minv, hipStream_t st);

}  // namespace sfem
