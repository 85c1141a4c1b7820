// sfem_gnn.hpp -- launchers of the learned-preconditioner kernels (sfem_gnn.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfem {

// Y = [X1 | X2] W + b (+ R) (W: K x NO bf16 row-major), Y and its relu copy
// Yact each optional; K1 a multiple of 4
hipError_t gnn_gemm(int K, int NO, int64_t N, const float* X1, int K1, const float* X2, const uint16_t* W, const float* bias,
                    float* Y, float* Yact, hipStream_t s, const float* R = nullptr);
// out (K x NO) = [X1 | X2]^T dY via nparts partials (partial: nparts*K*NO floats)
hipError_t gnn_wgrad(int K, int NO, int64_t N, const float* X1, int K1, const float* X2, const float* dY, float* partial,
                     int nparts, float* out, hipStream_t s);
// out (NO) = column sums of dY
hipError_t gnn_bgrad(int NO, int64_t N, const float* dY, float* partial, int nparts, float* out, hipStream_t s);
hipError_t gnn_relu_bwd(int64_t n, const float* pre, const float* dpost, float* dpre, hipStream_t s);
hipError_t gnn_split_concat(int64_t N, const float* dU, float* dh, float* dm, hipStream_t s);
hipError_t gnn_mean_fwd(int64_t N, const int* ptr, const int* idx, const float* h, float* m, hipStream_t s);
hipError_t gnn_mean_bwd(int64_t N, const int* ptr, const int* idx, const float* dm, float* dh, hipStream_t s);
hipError_t gnn_features(int64_t N, const float* B, const uint8_t* fixed, const int* ptr, const int* idx, float* feat,
                        float* cbase, hipStream_t s);
hipError_t gnn_blocks_fwd(int64_t N, const float* O, const float* cbase, float eps, float* minv, hipStream_t s);
hipError_t gnn_blocks_bwd(int64_t N, const float* O, const float* cbase, const float* dM, float* dO, hipStream_t s);
hipError_t gnn_apply_blocks(int64_t N, int S, const float* minv, const float* r, float* z, hipStream_t s);
hipError_t gnn_loss_grad_blocks(int64_t N, int S, const float* r, const float* q, const float* c1, const float* c2, float* dM,
                                hipStream_t s);
hipError_t gnn_pair_dots(int64_t n3, int S, const float* z, const float* r, const float* q, double* partial, int nparts,
                         hipStream_t s);
hipError_t gnn_err_dot(int64_t n3, const float* e, const float* x, const float* r, double* partial, int nparts, hipStream_t s);
hipError_t gnn_adam(int n, float* w, const float* g, float* m, float* v, float lr_t, float b1, float b2, float eps, float clip,
                    hipStream_t s);
hipError_t gnn_to_bf16_pair(int K, int NO, const float* w, uint16_t* wb, uint16_t* wtb, hipStream_t s);

}  // namespace sfem
