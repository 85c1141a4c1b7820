/*
 * sfem_capi.h -- C-ABI of the MI355X matrix-free FEM elasticity library
 * (libsfem_hip.so).  Drop-in for SolidFEM's FemPhysicsMatrixFree Newton-CG path.
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative SFEM_E* code on
 *     failure; sfem_last_error() gives the message (thread-local).  Nothing
 *     throws across the ABI.
 *   - Node arrays are nn*3 doubles (x,y,z interleaved) in the caller's ORIGINAL
 *     node order.  The reference works on free nodes in its own shuffled order
 *     (FemPhysicsMatrixFree.cpp:229-263); here rows of fixed nodes are present
 *     and zero wherever the reference drops them.
 *   - Host buffers are borrowed for the duration of a call.  The context owns
 *     all device memory; calls are synchronous at return.  One context is not
 *     thread-safe; distinct contexts may be used from distinct threads.
 *   - Element kernels and the CG operator run in FP32; residuals, dot products,
 *     norms, line search and the Newton update run in FP64 on the device.
 */
#ifndef SFEM_CAPI_H
#define SFEM_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* MaterialModelType values (Include/FemDataStructures.h:84-88).
   SFEM_COROTATIONAL is the reference's model: the Gram-Schmidt rotation of
   QRPolarDecomposition (PolarDecomposition.cpp:269-280) and its tangent
   (FemPhysicsMatrixFree.cpp:441-456), which is not the derivative of its force.
   SFEM_COROTATIONAL_POLAR (no reference counterpart; a value past the
   reference's MaterialModelType range) is the corotational model with the
   rotation of the polar decomposition F = R S, an energy
   Psi = mu |F - R|^2 + lam/2 tr(S - I)^2 and its consistent, symmetric tangent. */
enum { SFEM_LINEAR = 0, SFEM_STVK = 1, SFEM_COROTATIONAL = 2, SFEM_NEO_HOOKEAN = 3, SFEM_COROTATIONAL_POLAR = 16 };

/* SimulationType values (Include/FemDataStructures.h:71-77) */
enum { SFEM_ST_STATIC = 0, SFEM_ST_QUASI_STATIC = 1, SFEM_ST_EXPLICIT = 2, SFEM_ST_IMPLICIT = 3 };

/* Solve modes.
   REFERENCE: SolveNewtonCG exactly (FemPhysicsMatrixFree.cpp:604-628): outer
     Newton steps, each an unpreconditioned CG from zero to |r.r| <= (1e-3f)^2
     |r0.r0| or inner_iterations; no residual test, no line search.
   CONVERGED: Newton to ||r||_2 <= tol * ||r_ext||_2 (FP64), PCG inner solves
     to cg_rtol, FP64 back-tracking line search, block-Jacobi fallback; with
     a learned or multigrid preconditioner Jacobi-PCG is also the judge: a
     solve whose final FP64 residual misses the target is rerun from its start
     state with Jacobi-PCG.                                                    */
enum { SFEM_SOLVE_REFERENCE = 0, SFEM_SOLVE_CONVERGED = 1 };

/* Preconditioners for the inner CG.  MULTIGRID: one V-cycle over the levels
   added by sfem_mg_add_level per CG iteration, Chebyshev smoothing on block
   Jacobi; MULTIGRID_LEARNED: the same with the learned blocks in the finest
   level's smoother.  A learned or multigrid PCG that breaks down (p.Kp <= 0,
   r.z <= 0), stagnates, hits its iteration limit or meets a non-finite value
   is redone with block-Jacobi PCG for that Newton step. */
enum { SFEM_PC_NONE = 0, SFEM_PC_JACOBI = 1, SFEM_PC_LEARNED = 2, SFEM_PC_MULTIGRID = 3, SFEM_PC_MULTIGRID_LEARNED = 4 };

enum {
    SFEM_OK = 0,
    SFEM_EINVAL = -1,     /* bad argument */
    SFEM_EDEVICE = -2,    /* HIP runtime error */
    SFEM_ENOMEM = -3,     /* device allocation failed */
    SFEM_ENOTCONV = -4,   /* solve did not reach its tolerance */
    SFEM_ESTATE = -5      /* call not valid in the current state */
};

typedef struct sfem_ctx sfem_ctx;

typedef struct {
    const int32_t* tets;    /* ne*4 node ids (Tet::idx, FemDataStructures.h:167-170) */
    int64_t ne;
    const double* X0;       /* nn*3 rest positions (Node::pos0) */
    int64_t nn;
    const uint8_t* fixed;   /* nn; 1 = Dirichlet node (Node::invMass == 0); a node
                               that no element references is held fixed as well */
} sfem_mesh;

typedef struct {
    int32_t material;       /* SFEM_LINEAR .. SFEM_NEO_HOOKEAN (FemConfig::mMaterial) */
    double young;           /* FemConfig::mYoungsModulus */
    double poisson;         /* FemConfig::mPoissonRatio */
    const double* young_e;  /* optional per-element E (ne), overrides young */
    const double* poisson_e;/* optional per-element nu (ne), overrides poisson */
    double density;         /* FemConfig::mDensity */
    double gravity[3];      /* FemPhysicsBase::mGravity (FemPhysicsBase.cpp:30-33) */
} sfem_material;

typedef struct {
    int32_t mode;              /* SFEM_SOLVE_REFERENCE | SFEM_SOLVE_CONVERGED */
    int32_t precond;           /* SFEM_PC_* (REFERENCE mode ignores it: none) */
    int32_t outer_iterations;  /* FemConfig::mOuterIterations; CONVERGED: max Newton steps */
    int32_t inner_iterations;  /* FemConfig::mInnerIterations; max CG iterations per step */
    double tol;                /* CONVERGED: relative FP64 residual target (e.g. 1e-8) */
    double cg_rtol;            /* CONVERGED: cap of the Eisenstat-Walker forcing ||r_cg|| <= eta ||b||;
                                  <= 0: 0.3 with the multigrid preconditioners, else 0.1 */
    int32_t line_search;       /* CONVERGED: 1 = FP64 back-tracking on ||r|| */
    int32_t check_every;       /* CG iterations between host convergence checks (>=1) */
    int32_t sim_type;          /* sfem_step only: SFEM_ST_STATIC | SFEM_ST_QUASI_STATIC */
    double force_step;         /* sfem_step only: FemConfig::mForceApplicationStep */
    double abs_tol;            /* CONVERGED: also stop when ||r||_2 <= abs_tol
                                  (FemConfig::mAbsNewtonRsidualThreshold); 0 = off */
    int32_t predictor;         /* sfem_step, ST_QUASI_STATIC, CONVERGED only: 1 = start each
                                  load step from the secant extrapolation of the last two
                                  states, 2 = from the quadratic one through the last three
                                  (the secant while there are two); dropped if it inverts an
                                  element; the step is redone from the last state if an inner
                                  solve breaks down there or that solve fails; 0 = from the
                                  last state, as Step does */
    double time_step;          /* sfem_step, ST_EXPLICIT / ST_IMPLICIT: Step(dt)'s dt */
    int32_t substeps;          /* sfem_step, ST_EXPLICIT / ST_IMPLICIT: FemConfig::mNumSubsteps (>= 1) */
    int32_t pc_newton_cap;     /* CONVERGED, learned / multigrid preconditioners: Newton steps the
                                  preconditioned attempt gets before the judge reruns the solve
                                  with Jacobi-PCG (0 = outer_iterations) */
} sfem_solve_params;

typedef struct {
    int32_t newton_iters;      /* Newton steps taken */
    int32_t cg_iters;          /* total CG iterations */
    int32_t converged;         /* CONVERGED mode: 1 if ||r|| <= max(tol ||r_ext||, abs_tol) */
    int32_t fallbacks;         /* linear solves of the learned / multigrid preconditioner redone
                                  with block Jacobi (breakdown, stagnation, iteration limit or a
                                  non-finite value), and solves rerun by the judge */
    double res0;               /* ||r||_2 at the start (FP64) */
    double res;                /* ||r||_2 at the end (FP64) */
    double ext_norm;           /* ||r_ext||_2 at the end (FP64) */
    double seconds;            /* device wall time of the solve */
    int64_t hessvec_evals;     /* element Hessian-vector evaluations (ne per apply) */
    int64_t force_evals;       /* element internal-force evaluations (ne per residual) */
    int32_t reruns;            /* solves rerun by the judge: a learned / multigrid solve whose FP64
                                  residual missed the target is redone from its start state with
                                  Jacobi-PCG (counted in fallbacks too) */
    int32_t rejects;           /* quasi-static load steps whose predicted start was abandoned (an
                                  inner solve broke down there, or Newton did not improve on the
                                  last state within 4 steps) and that were redone from the last
                                  state */
} sfem_solve_stats;

/* FemPhysicsMatrixFree(tets, nodes, config) (Include/FemPhysicsMatrixFree.h:61-62,
   Source/FemPhysicsMatrixFree.cpp:71-213): copies the mesh to the device,
   computes Dm^-1, volumes and lumped masses, partitions elements. */
int sfem_create(const sfem_mesh* mesh, const sfem_material* mat, sfem_ctx** out);

/* releases every device buffer of the context */
void sfem_destroy(sfem_ctx* ctx);

/* message of the last failed call on this thread */
const char* sfem_last_error(void);

/* SetDeformedPosition for all nodes (Include/FemPhysicsMatrixFree.h:72) */
int sfem_set_positions(sfem_ctx* ctx, const double* x);

/* GetDeformedPosition for all nodes, original order (Include/FemPhysicsMatrixFree.h:70) */
int sfem_get_positions(sfem_ctx* ctx, double* x);

/* the load fraction mForceFraction (FemPhysicsBase.h:165; GetForceFraction :108) */
int sfem_set_load_fraction(sfem_ctx* ctx, double frac);
int sfem_get_load_fraction(sfem_ctx* ctx, double* frac);

/* SetBoundaryConditionsSurface (FemPhysicsMatrixFree.cpp:1051-1060): follower
   pressure on a triangle list (original node ids) */
int sfem_set_pressure_surface(sfem_ctx* ctx, const int32_t* tris, int64_t ntri, double pressure);

/* SetExternalForces (FemPhysicsBase.h:68; asserts for the matrix-free class in
   the reference): dead nodal loads scaled by the load fraction.  NULL clears. */
int sfem_set_external_forces(sfem_ctx* ctx, const double* f);

/* ElasticEnergy::ComputeForces (Source/ElasticEnergy.cpp:259-282), FP64, all rows */
int sfem_internal_forces(sfem_ctx* ctx, double* f);

/* ElasticEnergy::ComputeElementStress (ElasticEnergy.cpp:114-257): first
   Piola-Kirchhoff stress of every element at the current positions, ne*9
   doubles row-major, original element order (FP64; the VTK heat map's cell
   data, FemIO.cpp:317-349) */
int sfem_element_stress(sfem_ctx* ctx, double* P);

/* Energies at the current state (FP64, device reductions in a fixed order):
   out[0] = sum_e V_e Psi(F_e), FemPhysicsBase::ComputeElasticEnergy
            (FemPhysicsBase.cpp:203-212, ElasticEnergy::ComputeElementEnergy
            ElasticEnergy.cpp:40-112);
   out[1] = the cable springs' potential, FemPhysicsMatrixFree::ComputeSpringEnergy
            (FemPhysicsMatrixFree.cpp:1062-1121; not scaled by the load fraction,
            as there);
   out[2] = frac sum over free nodes of (m g + f_dead) . u, the work of the dead
            loads (gravity, sfem_set_external_forces).
   out[0] - out[2] is the total potential energy of a problem with dead loads
   only, whose negative gradient is sfem_gradients' residual; the C++ class
   builds FemPhysicsMatrixFree::ComputeEnergy (:882-916) from these. */
int sfem_energy(sfem_ctx* ctx, double* out);

/* ElasticEnergy::ComputeForceParamGrads (ElasticEnergy.cpp:284-315) and the
   density term of FemPhysicsBase::GetForceParamGrads (FemPhysicsBase.cpp:162-187)
   at the current state: fmu, flam = d f_int / d mu and d f_int / d lambda per node
   (the internal force with (mu, lambda) = (1, 0) and (0, 1): P is linear in both),
   frho = (sum of V/4 over the node's elements) g on free nodes; nn*3 doubles each,
   all rows, original order.  The reference fills these for NEO_HOOKEAN only
   (ComputeElementStress sets Pmu / Plambda for that material alone, zeros
   otherwise); here every material gets its derivative. */
int sfem_force_param_grads(sfem_ctx* ctx, double* fmu, double* flam, double* frho);

/* Cable springs: AddCable (FemPhysicsBase.cpp:152-155) with a cable as
   CreateCable builds it (FemBody.cpp:31-104): n >= 2 cable nodes, node k inside
   element elem[k] (original element index) at barycentric weights bary[3k ..
   3k+2] (the fourth is 1 - their sum), n - 1 springs of rest length
   rest_length, stiffness and damping, actuation 1.  Their forces (the Bern
   tension model of ComputeSpringForces, FemPhysicsBase.cpp:214-309, scaled by
   the load fraction) enter the residual (ComputeGradients :544) and the
   explicit substep (:377), and are counted in ||r_ext||; as in the reference
   they add no stiffness to the matrix-free operator.  Free cable nodes
   (elem < 0: points outside the mesh, which the reference turns into extra
   massless degrees of freedom, FemPhysicsMatrixFree.cpp:265-294) are refused
   with SFEM_EINVAL.  *id (nullable) receives the cable's index. */
int sfem_add_cable(sfem_ctx* ctx, const int32_t* elem, const double* bary, int32_t n, double rest_length, double stiffness,
                   double damping, int32_t* id);
/* SetCableActuation (FemPhysicsBase.h:134): the rest length scale of one
   cable; 0 switches its forces off; an index out of range is ignored */
int sfem_set_cable_actuation(sfem_ctx* ctx, int32_t cable, double actuation);
int sfem_num_cables(sfem_ctx* ctx, int32_t* n);
/* GetCablePosition (FemPhysicsBase.h:133) for all nodes of one cable, n*3 */
int sfem_get_cable_positions(sfem_ctx* ctx, int32_t cable, double* pos);
int sfem_clear_cables(sfem_ctx* ctx);

/* Node velocities (Node::vel) of the explicit integrator, nn*3 doubles */
int sfem_get_velocities(sfem_ctx* ctx, double* v);

/* ComputeGradients (FemPhysicsMatrixFree.cpp:536-560), FP64; fixed rows 0 */
int sfem_gradients(sfem_ctx* ctx, double* r);

/* MatrixVectorMultiply (FemPhysicsMatrixFree.cpp:523-534), static: K d on free
   rows, FP32 element kernel, fixed rows of d ignored and of Kd zero */
int sfem_matvec(sfem_ctx* ctx, const double* d, double* Kd);

/* sfem_matvec without allocating: no reference counterpart; for callers that
   cannot allocate (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc).  Same kernels and result as sfem_matvec,
   but the host staging, the FP32 direction and result and the shared-node
   partial rows live in a static per-device array of 16 MiB (held by each
   device the library runs on until the process ends); no hipMalloc, hipFree
   or host allocation.  Meshes needing more (about 24 B per node plus 12 B per
   partial-row slot, i.e. up to ~1M elements of the beam meshes) get SFEM_EINVAL.
   All calls on one device share the array: calls that overlap on one device
   race. */
int sfem_matvec_noalloc(sfem_ctx* ctx, const double* d, double* Kd);

/* 3x3 diagonal blocks of K per node (nn*9, row-major), FP32 kernel; fixed
   nodes get the identity.  The block-Jacobi preconditioner is their inverse. */
int sfem_block_diagonal(sfem_ctx* ctx, double* B);

/* Solve() with NST_NEWTON_CG (FemPhysicsMatrixFree.cpp:570-628) at the current
   load fraction, from the current positions */
int sfem_solve(sfem_ctx* ctx, const sfem_solve_params* p, sfem_solve_stats* st);

/* Step(dt) (FemPhysicsMatrixFree.cpp:297-342): ST_STATIC solves once at full
   load; ST_QUASI_STATIC raises the load fraction by force_step and solves;
   ST_EXPLICIT runs `substeps` symplectic Euler substeps of time_step/substeps
   (SubStep, FemPhysicsMatrixFree.cpp:371-387: v += h f_int / m + h g,
   x += h v on free nodes, elastic forces only); ST_IMPLICIT (:299-314) saves
   x0 = x, then runs `substeps` solves with h = time_step/substeps: operator
   M/h^2 + K (MatrixVectorMultiply :523-534), residual + m/h v - m/h^2 (x - x0)
   (ComputeGradients :536-560), v = (x - x0)/h after each (:589-597).  After an
   implicit step h stays set, as mTimeStep does (sfem_matvec and
   sfem_gradients include the dynamic terms); sfem_reset clears it. */
int sfem_step(sfem_ctx* ctx, const sfem_solve_params* p, sfem_solve_stats* st);

/* Installs per-node 3x3 inverse preconditioner blocks (nn*9 FP32 row-major,
   original order) used by SFEM_PC_LEARNED; NULL uninstalls.  No reference
   counterpart: the reference CG's only preconditioner is a probed scalar
   diagonal (Math/IterativeSolvers.h:133-152). */
int sfem_set_precond_blocks(sfem_ctx* ctx, const float* minv9);

/* Returns every node to its rest position and the load fraction to 0 without
   a host round trip (the state FemPhysicsMatrixFree's constructor leaves). */
int sfem_reset(sfem_ctx* ctx);

/* Neo-Hookean static solves (no reference counterpart; the reference rebuilds
   F, F^-1 and ln J of every element in every product,
   FemPhysicsMatrixFree.cpp:417-470): mode 1 (the default) builds the
   tangent's state per element (Dm^-1 F^-1 and V (mu - lam ln J), 40 B) once
   per Newton step and makes every Hessian-vector product of that step's
   block-Jacobi, learned-block or unpreconditioned PCG (and of the reference
   semantics' CG) from it, on meshes of up to 1.6M elements (a state of up to
   64 MiB; beyond the 256 MB Infinity Cache its reads cost more than the
   arithmetic they save); mode 0 rebuilds it in every product; mode 2 as 1,
   and sfem_matvec builds it at each call too (for tests).  They agree to
   FP32 rounding.  The multigrid PCG and the other stand-alone calls always
   use the state itself. */
int sfem_set_tangent_cache(sfem_ctx* ctx, int mode);

/* Instrumentation: average device time (ms) of one launch of the Hessian-vector
   element-block kernel and of its shared-node gather kernel over `reps`
   back-to-back launches at the current state, timed with HIP events on the
   context's stream.  Corresponds to the reference's PROFILE_SCOPE("Differential")
   (FemPhysicsMatrixFree.cpp:408). */
int sfem_profile_hessvec(sfem_ctx* ctx, int reps, double* ms_elem, double* ms_gather);

/* Instrumentation inside solves: every > 0 starts timing every `every`-th
   Hessian-vector element-block launch of the CG loops with a pair of HIP
   events that the dispatch itself records at its start and end, on the
   context's stream (accumulators reset); every == 0 stops.
   Returns the average device time (ms) and the number of timed launches since
   the last start.  Same counterpart as sfem_profile_hessvec. */
int sfem_time_hessvec(sfem_ctx* ctx, int every, double* ms_avg, int64_t* samples);

/* ---------------------------------------------------------------------------
 * Multigrid preconditioner (no reference counterpart: the reference CG is
 * unpreconditioned, Math/IterativeSolvers.h:115-195).  A hierarchy of nested
 * tetrahedral meshes, finest (the context's) first; each coarse level runs the
 * same element kernels on its own mesh at the state injected from the level
 * above, so its operator is the tangent of its mesh (rediscretisation).
 * ------------------------------------------------------------------------- */

/* Adds a level below the current coarsest one (the context's mesh, before the
   first call).  coarse: the level's mesh (positive volumes; fixed flags of its
   Dirichlet nodes).  inject[k] (coarse nn): the node of the level above at
   coarse node k's position; the coarse displacement is sampled there.
   Prolongation from this level to the one above, CSR over the nodes of the
   level above: row i holds entries p_ptr[i] .. p_ptr[i+1]-1 of
   (p_col = coarse node, p_val = weight); node values, the same for x, y, z.
   The level has the context's material model; its moduli are the context's
   uniform ones, or per coarse element young_e / poisson_e (required when the
   context has per-element moduli, e.g. averages over each coarse element).
   The coarsest level is solved directly (a dense inverse rebuilt once per load
   step) when its dofs, 3 x nodes, are at most SFEM_MG_DENSE_MAX, else by
   Chebyshev smoothing. */
enum { SFEM_MG_DENSE_MAX = 1536 };
int sfem_mg_add_level(sfem_ctx* ctx, const sfem_mesh* coarse, const int32_t* inject, const int64_t* p_ptr,
                      const int32_t* p_col, const double* p_val, const double* young_e, const double* poisson_e);
int sfem_mg_clear(sfem_ctx* ctx);
/* levels (1 + coarse levels) and, when lmax is non-null, the largest-eigenvalue
   estimates of M A per level from the last set-up */
int sfem_mg_info(sfem_ctx* ctx, int32_t* nlevels, double* lmax);
/* The directly solved coarsest level at the current state (after a full
   set-up): n = 3 x its nodes; when a / ainv are non-null (n x n each,
   capacity >= n * n), the assembled operator A (FP32 element matrices; fixed
   rows and columns identity) and the inverse the V-cycle applies, row-major,
   dof 3 k + c of coarse node k.  SFEM_ESTATE when the coarsest level is not
   solved directly. */
int sfem_mg_coarse_dense(sfem_ctx* ctx, int64_t capacity, int32_t* n, double* a, double* ainv);
/* z = V(r): one V-cycle (block-Jacobi smoothers) at the current state;
   r, z nn*3 in original order, fixed rows ignored / zero */
int sfem_mg_apply(sfem_ctx* ctx, const double* r, double* z);
/* One linear solve K x = b at the current state by PCG with the V-cycle
   (block-Jacobi smoothers) to ||r|| <= rtol ||b|| in at most maxit iterations;
   b, x nn*3 in original order; status 1 converged, 0 iteration limit,
   3 breakdown (p.Kp <= 0 or r.z <= 0), 4 not finite, 5 stagnation (r.r not
   halved within 150 iterations); at most 500 iterations */
int sfem_mg_pcg(sfem_ctx* ctx, const double* b, double* x, double rtol, int32_t maxit, int32_t* iters, int32_t* status);
/* The V-cycle's launch schedule (no reference counterpart; for tests and
   measurement): fused = 1 (the default) folds each level's Chebyshev steps
   and its prolongation into the element kernels' row staging and the
   residual into the restriction (6 launches per smoothed level instead of
   10); fused = 0 launches one kernel per step.  Both give bitwise the same
   V-cycle. */
int sfem_mg_set_schedule(sfem_ctx* ctx, int fused);

/* ---------------------------------------------------------------------------
 * Learned block preconditioner (no reference counterpart).  A message-passing
 * network over the node graph (3 rounds, width 32, BF16 MFMA GEMMs with FP32
 * accumulation) maps per-node features of the block-Jacobi blocks D_i to SPD
 * blocks Minv_i = C_i (L_i L_i^T + 0.01 I) C_i^T with C_i C_i^T = D_i^-1; the
 * untrained network is block Jacobi.  Trained on the device on residual /
 * solution pairs (r, e), K e = r, that the solver collects.
 * ------------------------------------------------------------------------- */
typedef struct sfem_gnn sfem_gnn;

int sfem_gnn_create(uint64_t seed, sfem_gnn** out);
void sfem_gnn_destroy(sfem_gnn* gnn);
int sfem_gnn_num_params(sfem_gnn* gnn, int64_t* n);
int sfem_gnn_get_params(sfem_gnn* gnn, float* w);
int sfem_gnn_set_params(sfem_gnn* gnn, const float* w);

/* Attaches the network to a context: SFEM_PC_LEARNED solves then rebuild the
   blocks from it at every Newton step (block-Jacobi fallback unchanged).
   NULL detaches.  The network must outlive its attachment. */
int sfem_set_learned_preconditioner(sfem_ctx* ctx, sfem_gnn* gnn);

/* Solves K e = load at the current state (block-Jacobi PCG to rtol) and keeps
   nsnap pairs (r_k, r_k . e_k): k = 0, snap_every, 2 snap_every, ... of a
   second identical PCG run, with e_k = e - x_k so that K e_k = r_k.  The
   pairs belong to the current state: change it only after training. */
int sfem_collect_pairs(sfem_ctx* ctx, const double* load, int nsnap, int snap_every, double rtol, int* collected);
/* Keeps `count` pairs (K e, e . K e) at the current state with e white noise
   on the free dofs (seeded): the high-frequency errors a multigrid smoother
   has to remove, for training the network as a smoother. */
int sfem_collect_rough_pairs(sfem_ctx* ctx, int count, uint64_t seed, int* collected);
int sfem_num_pairs(sfem_ctx* ctx, int64_t* n);
int sfem_get_pair(sfem_ctx* ctx, int64_t k, float* r, double* er);
int sfem_clear_pairs(sfem_ctx* ctx);

/* Mean loss 1 - (z.r)^2 / ((z.Kz)(r.e)), z = Minv r, over pairs
   [first, first+count) (count <= 16), and its parameter gradient when grad is
   non-null (nparams floats). */
int sfem_gnn_loss(sfem_gnn* gnn, sfem_ctx* ctx, int64_t first, int count, double* loss, float* grad);

/* Mean loss over all pairs of the context for the network and for block Jacobi */
int sfem_gnn_eval(sfem_gnn* gnn, sfem_ctx* ctx, double* loss_learned, double* loss_jacobi);

/* `steps` Adam steps (lr, beta 0.9 / 0.999) on minibatches of `batch` (<= 16)
   pairs drawn with `seed`; losses[steps] receives each step's loss (nullable) */
int sfem_gnn_train(sfem_gnn* gnn, sfem_ctx* ctx, int steps, int batch, float lr, uint64_t seed, double* losses);

/* The network's inputs and outputs at the current state: features (nn*16),
   Cholesky factors of the inverse Jacobi blocks (nn*6), blocks (nn*9) */
int sfem_gnn_inputs(sfem_gnn* gnn, sfem_ctx* ctx, float* feat, float* cbase, float* minv);

/* Y = X W + b through the BF16 MFMA GEMM of layer `layer` (N rows) */
int sfem_gnn_gemm_check(sfem_gnn* gnn, int layer, int64_t N, const float* X, float* Y);

/* Number of elements, nodes, fixed nodes, and element blocks of the partition */
int sfem_info(sfem_ctx* ctx, int64_t* ne, int64_t* nn, int64_t* nfixed, int64_t* nblocks);

#ifdef __cplusplus
}
#endif
#endif /* SFEM_CAPI_H */
