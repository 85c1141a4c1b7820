// sfem_mg.hpp -- launchers of the multigrid preconditioner kernels (sfem_mg.hip)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfem {

// the coarsest level is solved directly when its dofs (3 x nodes) are at most
// this many (a dense inverse of up to 9.4 MB, rebuilt once per load step)
constexpr int MG_DENSE_MAX = 1536;

// Chebyshev step: mode 0 first step from x = 0, 1 next step, 2 first step from
// x after a coarse correction, 3 residual only (see k_mg_cheb)
// done: the PCG's flag, or a flag that stays 0; every kernel of the V-cycle
// returns at once when it is set (V-cycles enqueued past the end of a solve)
hipError_t launch_mg_cheb(int mode, int64_t nn, const float* b, float* r, float* d, float* x, const float* q,
                          const int2* node_q, const float* partial, const float* minv, float c1, float c2, const int* done,
                          hipStream_t st);
hipError_t launch_mg_prolong(int64_t nn, const int* ptr, const int* col, const float* val, const float* xc,
                             const uint8_t* fixed, float* e, float* x, const int* done, hipStream_t st);
// minv non-null: also the coarse level's first Chebyshev step from x = 0
// (rc = bc, dc = c2 M bc, xc = dc)
hipError_t launch_mg_restrict(int64_t nnc, const int* ptr, const int* row, const float* val, const float* r,
                              const uint8_t* fixed, float* bc, const int* done, hipStream_t st, const float* minv = nullptr,
                              float* rc = nullptr, float* dc = nullptr, float* xc = nullptr, float c2 = 0.f);
// the finest level's last Chebyshev step (k_mg_cheb mode 1) and the PCG's
// r.z, r.r partials of z = x against b in the CG update's layout (2 x cg_grid)
hipError_t launch_mg_cheb_rz(int64_t nn, const float* b, float* r, float* d, float* x, const float* q, const int2* node_q,
                             const float* partial, const float* minv, float c1, float c2, double* part, const int* done,
                             hipStream_t st);
hipError_t launch_mg_inject(int64_t nnc, const int* inj, const double* ua, double* u64, float* u32, hipStream_t st);
hipError_t launch_mg_power(int64_t nn, const float* q, const int2* node_q, const float* partial, const float* minv, float* y,
                           hipStream_t st);
hipError_t launch_mg_start(int64_t nn, const uint8_t* fixed, float* v, hipStream_t st);
// out[0] = a.a (FP64, fixed order); part holds mg_sqnorm_grid(n3) doubles
int mg_sqnorm_grid(int64_t n3);
hipError_t launch_mg_sqnorm(int64_t n3, const float* a, double* part, double* out, hipStream_t st);
// r.z, r.r partials in the CG update's layout (2 x cg_grid(nn))
hipError_t launch_mg_rz(int64_t nn, const float* r, const float* z, double* part, const int* done, hipStream_t st);
hipError_t launch_mg_copy(int64_t n, const float* a, float* b, hipStream_t st);
// b = sc a
hipError_t launch_mg_scale(int64_t n, const float* a, float sc, float* b, hipStream_t st);
// dense direct solve of the coarsest level: element matrices (12 x 12 per
// element), the dense operator (n = 3 nn), its inverse, x = Ainv b
hipError_t launch_mg_elem_matrix(int mat, int64_t ne, const int4* tets, const float* u, const float* dminv,
                                 const float* cvol, float* Ke, hipStream_t st);
hipError_t launch_mg_dense_assemble(int nn, const int* inc_ptr, const int2* inc, const int4* tets, const float* Ke,
                                    const uint8_t* fixed, float* A, hipStream_t st);
// Ainv = A^-1 (symmetric A) in FP32, computed in FP64 in W (n x n), C (3 x n x 32:
// two panel-column buffers and G) and S (32 x 32) of scratch
hipError_t launch_mg_dense_inverse(int n, const float* A, float* Ainv, double* W, double* C, double* S, hipStream_t st);
// x = Ainv b
hipError_t launch_mg_dense_solve(int n, const float* Ainv, const float* b, float* x, const int* done, hipStream_t st);

}  // namespace sfem
