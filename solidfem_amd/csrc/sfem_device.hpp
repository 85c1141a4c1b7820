// sfem_device.hpp -- per-tetrahedron constitutive math for the MI355X kernels.
//
// All element quantities are computed from DISPLACEMENTS: with the rest shape
// matrix Dm and G = Dm^-1 (row-major g[3k+j] = G(k,j)),
//     F  = I + [u1-u0 | u2-u0 | u3-u0] G          (= Ds G, FemPhysicsMatrixFree.cpp:354-368)
//     dF = [d1-d0 | d2-d0 | d3-d0] G              (:426-430)
// so FP32 evaluation does not lose the small strains to cancellation against
// the rest positions.  Material coefficients arrive pre-multiplied by the rest
// volume (cmu = V mu, clam = V lambda): every stress and stress differential
// below is linear in (mu, lambda), so the result is already V P / V dP.
//
// Corner outputs: column k of (V dP) G^T is the contribution to corner k+1,
// corner 0 gets minus their sum (FemPhysicsMatrixFree.cpp:479-495).
#pragma once
#include <hip/hip_runtime.h>

namespace sfem {

enum { LINEAR = 0, STVK = 1, COROTATIONAL = 2, NEO_HOOKEAN = 3,
       NEO_HOOKEAN_COARSE = 4,     // internal: coarse multigrid levels (MAT_NEO_HOOKEAN_COARSE)
       COROTATIONAL_POLAR = 5,     // internal kind of SFEM_COROTATIONAL_POLAR
       NEO_HOOKEAN_LIN = 6 };      // internal: Hessian-vector kernel on the cached Neo-Hookean state (nh_lin_corners)

template <typename T> __device__ __forceinline__ T dsqrt(T x) { return sqrt(x); }
template <typename T> __device__ __forceinline__ T dlog(T x) { return log(x); }
// 1/x: FP32 kernels (the CG operator) take the hardware reciprocal (1 ulp);
// the FP64 kernels divide exactly
__device__ __forceinline__ float drcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double drcp(double x) { return 1.0 / x; }

// C = A B (3x3 row-major)
template <typename T>
__device__ __forceinline__ void mul3(const T* A, const T* B, T* C)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

// adjugate inverse (Math/matrix3.h:234-249); returns the determinant
template <typename T>
__device__ __forceinline__ T inv3(const T* m, T* v)
{
    v[0] = m[4] * m[8] - m[5] * m[7];
    v[3] = -m[3] * m[8] + m[6] * m[5];
    v[6] = m[3] * m[7] - m[6] * m[4];
    v[1] = m[2] * m[7] - m[1] * m[8];
    v[4] = m[0] * m[8] - m[2] * m[6];
    v[7] = m[1] * m[6] - m[0] * m[7];
    v[2] = m[1] * m[5] - m[2] * m[4];
    v[5] = m[2] * m[3] - m[0] * m[5];
    v[8] = m[0] * m[4] - m[1] * m[3];
    T det = m[0] * v[0] + m[1] * v[3] + m[2] * v[6];
    T s = drcp(det);
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] *= s;
    return det;
}

// [a1|a2|a3] G  with a_k = w_k - w_0
template <typename T>
__device__ __forceinline__ void grad4(const T* g, const T w[4][3], T* M)
{
    T a[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int c = 0; c < 3; c++) a[k][c] = w[k + 1][c] - w[0][c];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            M[3 * i + j] = a[0][i] * g[j] + a[1][i] * g[3 + j] + a[2][i] * g[6 + j];
}

// corner outputs from a (volume-scaled) stress-like matrix S: col k of S G^T
template <typename T>
__device__ __forceinline__ void corners(const T* S, const T* g, T sign, T out[4][3])
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        T h0 = sign * (S[3 * c] * g[0] + S[3 * c + 1] * g[1] + S[3 * c + 2] * g[2]);
        T h1 = sign * (S[3 * c] * g[3] + S[3 * c + 1] * g[4] + S[3 * c + 2] * g[5]);
        T h2 = sign * (S[3 * c] * g[6] + S[3 * c + 1] * g[7] + S[3 * c + 2] * g[8]);
        out[1][c] = h0;
        out[2][c] = h1;
        out[3][c] = h2;
        out[0][c] = -(h0 + h1 + h2);
    }
}

// Gram-Schmidt rotation of PolarDecomposition::QRPolarDecomposition
// (Source/PolarDecomposition.cpp:269-280) and S = sym(R^T F) (:37-47)
template <typename T>
__device__ __forceinline__ void qr_rot(const T* F, T* R, T* S)
{
    T r0[3] = {F[0], F[3], F[6]};
    T s = T(1) / dsqrt(r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2]);
    r0[0] *= s; r0[1] *= s; r0[2] *= s;
    T a1[3] = {F[1], F[4], F[7]};
    T dd = a1[0] * r0[0] + a1[1] * r0[1] + a1[2] * r0[2];
    T r1[3] = {a1[0] - dd * r0[0], a1[1] - dd * r0[1], a1[2] - dd * r0[2]};
    s = T(1) / dsqrt(r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]);
    r1[0] *= s; r1[1] *= s; r1[2] *= s;
    T r2[3] = {r0[1] * r1[2] - r1[1] * r0[2], -r0[0] * r1[2] + r0[2] * r1[0], r0[0] * r1[1] - r1[0] * r0[1]};
    s = T(1) / dsqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
    r2[0] *= s; r2[1] *= s; r2[2] *= s;
#pragma unroll
    for (int i = 0; i < 3; i++) { R[3 * i] = r0[i]; R[3 * i + 1] = r1[i]; R[3 * i + 2] = r2[i]; }
    T V[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            V[3 * i + j] = R[i] * F[j] + R[3 + i] * F[3 + j] + R[6 + i] * F[6 + j];   // (R^T F)_ij
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S[3 * i + j] = T(0.5) * (V[3 * i + j] + V[3 * j + i]);
}

// Rotation of the polar decomposition F = R U (U symmetric, positive definite
// for det F > 0) by Newton's iteration X <- (g X + X^-T / g) / 2 with the
// Frobenius-norm scaling g = (|X^-1|_F / |X|_F)^(1/2) (N. J. Higham, SIAM J.
// Sci. Stat. Comput. 7 (1986) 1160-1174), until the update is below
// 2^-14 (FP32) / 2^-45 (FP64) of |X|_F.  Near the solution the iteration
// converges quadratically: the error of the new iterate is about
// |update|^2 / (2 sigma_min(F)), so an FP32 update below 2^-14 |X|_F already
// leaves the returned R exact to FP32 rounding (for sigma_min >= 0.05), one
// iteration earlier than a test at the rounding level would stop.
// S = sym(R^T F).  An inverted F (det < 0) converges to the orthogonal factor
// with det R = -1.
template <typename T>
__device__ __forceinline__ void polar_rot(const T* F, T* R, T* S)
{
    constexpr int maxit = sizeof(T) == 4 ? 12 : 20;
    const T tol2 = sizeof(T) == 4 ? T(3.0 * 0x1p-28) : T(3.0 * 0x1p-90);
    T X[9];
#pragma unroll
    for (int k = 0; k < 9; k++) X[k] = F[k];
    for (int it = 0; it < maxit; it++) {
        T Xi[9];
        inv3(X, Xi);
        T nx = 0, ni = 0;
#pragma unroll
        for (int k = 0; k < 9; k++) { nx += X[k] * X[k]; ni += Xi[k] * Xi[k]; }
        const T g = dsqrt(dsqrt(ni / nx)), h = T(0.5) / g, g2 = T(0.5) * g;
        T du = 0;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const T xn = g2 * X[3 * i + j] + h * Xi[3 * j + i];
                du += (xn - X[3 * i + j]) * (xn - X[3 * i + j]);
                X[3 * i + j] = xn;
            }
        if (!(du > tol2)) break;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) R[k] = X[k];
    T V[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            V[3 * i + j] = R[i] * F[j] + R[3 + i] * F[3 + j] + R[6 + i] * F[6 + j];   // (R^T F)_ij
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S[3 * i + j] = T(0.5) * (V[3 * i + j] + V[3 * j + i]);
}

// ---------------------------------------------------------------------------
// State-dependent part of the tangent, computed once per element per apply.
template <typename T, int MAT> struct Tangent;

// LINEAR: dP = mu (dF + dF^T) + lam tr(dF) I   (FemPhysicsMatrixFree.cpp:436-440)
template <typename T> struct Tangent<T, LINEAR> {
    T cmu, clam;
    __device__ __forceinline__ void init(const T*, T m, T l) { cmu = m; clam = l; }
    __device__ __forceinline__ void apply(const T* dF, T* dP) const
    {
        T tr = dF[0] + dF[4] + dF[8];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) dP[3 * i + j] = cmu * (dF[3 * i + j] + dF[3 * j + i]);
        dP[0] += clam * tr; dP[4] += clam * tr; dP[8] += clam * tr;
    }
};

// STVK: dP = dF (2mu E + lam trE I) + F (2mu dE + lam tr dE I), dE = sym(F^T dF)
// (:457-463)
template <typename T> struct Tangent<T, STVK> {
    T F[9], S2[9], cmu, clam;
    __device__ __forceinline__ void init(const T* Fin, T m, T l)
    {
        cmu = m; clam = l;
#pragma unroll
        for (int k = 0; k < 9; k++) F[k] = Fin[k];
        T E[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                E[3 * i + j] = T(0.5) * (F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j] - (i == j ? T(1) : T(0)));
        T tr = E[0] + E[4] + E[8];
#pragma unroll
        for (int k = 0; k < 9; k++) S2[k] = 2 * cmu * E[k];
        S2[0] += clam * tr; S2[4] += clam * tr; S2[8] += clam * tr;
    }
    __device__ __forceinline__ void apply(const T* dF, T* dP) const
    {
        T FtdF[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                FtdF[3 * i + j] = F[i] * dF[j] + F[3 + i] * dF[3 + j] + F[6 + i] * dF[6 + j];
        T trdE = FtdF[0] + FtdF[4] + FtdF[8];
        T B[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) B[3 * i + j] = cmu * (FtdF[3 * i + j] + FtdF[3 * j + i]);
        B[0] += clam * trdE; B[4] += clam * trdE; B[8] += clam * trdE;
        T t1[9], t2[9];
        mul3(dF, S2, t1);
        mul3(F, B, t2);
#pragma unroll
        for (int k = 0; k < 9; k++) dP[k] = t1[k] + t2[k];
    }
};

// COROTATIONAL with the reference's QR rotation and its (non-symmetric)
// tangent (:441-456): dS = R^T dF; dP = 2mu dF + lam tr(dS) R
//   + (lam (trS-3) - 2mu) R skew((S - trS I)^-1 w(dS))
template <typename T> struct Tangent<T, COROTATIONAL> {
    T R[9], Ai[9], cmu, clam, cc;
    __device__ __forceinline__ void init(const T* F, T m, T l)
    {
        cmu = m; clam = l;
        T S[9];
        qr_rot(F, R, S);
        T trS = S[0] + S[4] + S[8];
        T A[9];
#pragma unroll
        for (int k = 0; k < 9; k++) A[k] = S[k];
        A[0] -= trS; A[4] -= trS; A[8] -= trS;
        inv3(A, Ai);
        cc = clam * (trS - 3) - 2 * cmu;
    }
    __device__ __forceinline__ void apply(const T* dF, T* dP) const
    {
        T dS[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                dS[3 * i + j] = R[i] * dF[j] + R[3 + i] * dF[3 + j] + R[6 + i] * dF[6 + j];
        T tr = dS[0] + dS[4] + dS[8];
        T w0 = dS[5] - dS[7], w1 = dS[2] - dS[6], w2 = dS[1] - dS[3];
        T r0 = Ai[0] * w0 + Ai[1] * w1 + Ai[2] * w2;
        T r1 = Ai[3] * w0 + Ai[4] * w1 + Ai[5] * w2;
        T r2 = Ai[6] * w0 + Ai[7] * w1 + Ai[8] * w2;
        // dR = R skew(r), skew(r) = [[0,-r2,r1],[r2,0,-r0],[-r1,r0,0]] (matrix3.h:264-269)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            T d0 = R[3 * i + 1] * r2 - R[3 * i + 2] * r1;
            T d1 = -R[3 * i] * r2 + R[3 * i + 2] * r0;
            T d2 = R[3 * i] * r1 - R[3 * i + 1] * r0;
            dP[3 * i + 0] = 2 * cmu * dF[3 * i + 0] + clam * tr * R[3 * i + 0] + cc * d0;
            dP[3 * i + 1] = 2 * cmu * dF[3 * i + 1] + clam * tr * R[3 * i + 1] + cc * d1;
            dP[3 * i + 2] = 2 * cmu * dF[3 * i + 2] + clam * tr * R[3 * i + 2] + cc * d2;
        }
    }
};

// COROTATIONAL_POLAR (no reference counterpart): the corotational model with
// the rotation of the polar decomposition, P = 2mu (F - R) + lam tr(S - I) R =
// dPsi/dF for Psi = mu |F - R|^2 + lam/2 tr(S - I)^2, and its consistent
// tangent.  With K = R^T dF, the rotation rate dR = R [w]x solves
// (tr S I - S) w = axial(K - K^T) (F = R S, S symmetric), so
//   dP = 2mu dF + lam tr(K) R + (lam (trS - 3) - 2mu) dR,
// the COROTATIONAL formula with the polar R and the axial vector
// (K21 - K12, K02 - K20, K10 - K01).  An inverted element (det F <= 0) takes
// the tangent at F = I.
template <typename T> struct Tangent<T, COROTATIONAL_POLAR> {
    T R[9], Ai[9], cmu, clam, cc;
    __device__ __forceinline__ void init(const T* F, T m, T l)
    {
        cmu = m; clam = l;
        const T J = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                    F[2] * (F[3] * F[7] - F[4] * F[6]);
        T S[9];
        if (J > T(0)) {
            polar_rot(F, R, S);
        } else {
#pragma unroll
            for (int k = 0; k < 9; k++) R[k] = S[k] = (k % 4 == 0) ? T(1) : T(0);
        }
        const T trS = S[0] + S[4] + S[8];
        T A[9];
#pragma unroll
        for (int k = 0; k < 9; k++) A[k] = -S[k];
        A[0] += trS; A[4] += trS; A[8] += trS;
        inv3(A, Ai);
        cc = clam * (trS - 3) - 2 * cmu;
    }
    __device__ __forceinline__ void apply(const T* dF, T* dP) const
    {
        T K[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                K[3 * i + j] = R[i] * dF[j] + R[3 + i] * dF[3 + j] + R[6 + i] * dF[6 + j];
        const T tr = K[0] + K[4] + K[8];
        const T a0 = K[7] - K[5], a1 = K[2] - K[6], a2 = K[3] - K[1];
        const T w0 = Ai[0] * a0 + Ai[1] * a1 + Ai[2] * a2;
        const T w1 = Ai[3] * a0 + Ai[4] * a1 + Ai[5] * a2;
        const T w2 = Ai[6] * a0 + Ai[7] * a1 + Ai[8] * a2;
        // dR = R [w]x, [w]x = [[0,-w2,w1],[w2,0,-w0],[-w1,w0,0]]
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const T d0 = R[3 * i + 1] * w2 - R[3 * i + 2] * w1;
            const T d1 = -R[3 * i] * w2 + R[3 * i + 2] * w0;
            const T d2 = R[3 * i] * w1 - R[3 * i + 1] * w0;
            dP[3 * i + 0] = 2 * cmu * dF[3 * i + 0] + clam * tr * R[3 * i + 0] + cc * d0;
            dP[3 * i + 1] = 2 * cmu * dF[3 * i + 1] + clam * tr * R[3 * i + 1] + cc * d1;
            dP[3 * i + 2] = 2 * cmu * dF[3 * i + 2] + clam * tr * R[3 * i + 2] + cc * d2;
        }
    }
};

// NEO_HOOKEAN: dP = mu dF + (mu - lam ln J) F^-T dF^T F^-T + lam tr(F^-1 dF) F^-T
// (:464-470); F^-T dF^T F^-T = (F^-1 dF F^-1)^T
template <typename T> struct Tangent<T, NEO_HOOKEAN> {
    T Fi[9], cmu, clam, c;
    __device__ __forceinline__ void init(const T* F, T m, T l)
    {
        cmu = m; clam = l;
        T J = inv3(F, Fi);
        c = cmu - clam * dlog(J);
    }
    __device__ __forceinline__ void apply(const T* dF, T* dP) const
    {
        T G[9], H[9];
        mul3(Fi, dF, G);
        T tr = G[0] + G[4] + G[8];
        mul3(G, Fi, H);
        T lt = clam * tr;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                dP[3 * i + j] = cmu * dF[3 * i + j] + c * H[3 * j + i] + lt * Fi[3 * j + i];
    }
};

// The Neo-Hookean tangent on the coarse levels of the multigrid
// preconditioner, whose states are injected from the level above and need not
// be valid deformations of the coarse mesh: an element with J <= 0.1 (or not
// finite) takes the tangent at F = I, and the coefficient c = mu - lam ln J of
// the F^-T dF^T F^-T term is kept in [mu / 10, mu].  Below: c turns negative
// for J > exp(mu / lam) (J > 1.01 at nu = 0.495).  Above: for a compressed
// element c grows like lam |ln J| (6 mu at J = 0.95, nu = 0.495), and since
// c (F^-T dF^T F^-T) : dF = c tr((F^-1 dF)^2) is -c |F^-1 dF|^2 for a
// rotation-like dF, the tangent is indefinite once c exceeds ~mu.  Both make
// the coarse operator, solved exactly on the coarsest level, indefinite.  A
// preconditioner operator, not the element's tangent.
template <typename T> struct Tangent<T, NEO_HOOKEAN_COARSE> : Tangent<T, NEO_HOOKEAN> {
    __device__ __forceinline__ void init(const T* F, T m, T l)
    {
        this->cmu = m;
        this->clam = l;
        T J = inv3(F, this->Fi);
        if (!(J > T(0.1))) {
#pragma unroll
            for (int k = 0; k < 9; k++) this->Fi[k] = (k % 4 == 0) ? T(1) : T(0);
            J = T(1);
        }
        const T c = m - l * dlog(J);
        this->c = c > T(0.1) * m ? (c < m ? c : m) : T(0.1) * m;
    }
};

// First Piola-Kirchhoff stress, volume-scaled (ElasticEnergy.cpp:114-187)
template <typename T, int MAT>
__device__ __forceinline__ void stress(const T* F, T cmu, T clam, T* P)
{
    if (MAT == NEO_HOOKEAN || MAT == NEO_HOOKEAN_COARSE) {
        T Fi[9];
        T J = inv3(F, Fi);
        T lJ = clam * dlog(J);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) P[3 * i + j] = cmu * (F[3 * i + j] - Fi[3 * j + i]) + lJ * Fi[3 * j + i];
    } else if (MAT == STVK) {
        Tangent<T, STVK> t;
        t.init(F, cmu, clam);
        mul3(F, t.S2, P);
    } else {
        T R[9], S[9];
        const T* U = F;
        if (MAT == COROTATIONAL) { qr_rot(F, R, S); U = S; }
        if (MAT == COROTATIONAL_POLAR) { polar_rot(F, R, S); U = S; }
        T e[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) e[3 * i + j] = T(0.5) * (U[3 * i + j] + U[3 * j + i]) - (i == j ? T(1) : T(0));
        T tr = e[0] + e[4] + e[8];
        T Q[9];
#pragma unroll
        for (int k = 0; k < 9; k++) Q[k] = 2 * cmu * e[k];
        Q[0] += clam * tr; Q[4] += clam * tr; Q[8] += clam * tr;
        if (MAT == COROTATIONAL || MAT == COROTATIONAL_POLAR) mul3(R, Q, P);
        else {
#pragma unroll
            for (int k = 0; k < 9; k++) P[k] = Q[k];
        }
    }
}

// Strain energy of one element, volume-scaled like the stresses (cmu = V mu,
// clam = V lambda; Psi is linear in mu, lambda), ElasticEnergy::
// ComputeElementEnergy (ElasticEnergy.cpp:40-112): LINEAR, COROTATIONAL (on
// the stretch S of the reference's QR rotation) and STVK mu e:e + lam/2 tr(e)^2
// with the Cauchy (LINEAR, COROTATIONAL) or Green (STVK) strain; NEO_HOOKEAN
// mu/2 (I1 - 3) - mu log J + lam/2 (log J)^2; COROTATIONAL_POLAR the
// corotational form on the polar stretch, mu |F - R|^2 + lam/2 tr(S - I)^2.
template <typename T, int MAT>
__device__ __forceinline__ T energy(const T* F, T cmu, T clam)
{
    if (MAT == NEO_HOOKEAN || MAT == NEO_HOOKEAN_COARSE) {
        T I1 = T(0);
#pragma unroll
        for (int k = 0; k < 9; k++) I1 += F[k] * F[k];
        const T J = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                    F[2] * (F[3] * F[7] - F[4] * F[6]);
        const T lJ = dlog(J);
        return T(0.5) * cmu * (I1 - T(3)) - cmu * lJ + T(0.5) * clam * lJ * lJ;
    }
    T e[9];
    if (MAT == STVK) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                e[3 * i + j] = T(0.5) * (F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j] - (i == j ? T(1) : T(0)));
    } else {
        T R[9], S[9];
        const T* U = F;
        if (MAT == COROTATIONAL) { qr_rot(F, R, S); U = S; }
        if (MAT == COROTATIONAL_POLAR) { polar_rot(F, R, S); U = S; }
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) e[3 * i + j] = T(0.5) * (U[3 * i + j] + U[3 * j + i]) - (i == j ? T(1) : T(0));
    }
    T ee = T(0);
#pragma unroll
    for (int k = 0; k < 9; k++) ee += e[k] * e[k];
    const T tr = e[0] + e[4] + e[8];
    return cmu * ee + T(0.5) * clam * tr * tr;
}

// F = I + grad(u)
template <typename T>
__device__ __forceinline__ void defgrad(const T* g, const T u[4][3], T* F)
{
    grad4(g, u, F);
    F[0] += T(1); F[4] += T(1); F[8] += T(1);
}

// Neo-Hookean Hessian-vector product on a state cached per element at the
// linearisation point.  With A = [d1-d0 | d2-d0 | d3-d0] (dF = A G) and
// E = G F^-1: F^-1 dF F^-1 = F^-1 A E, G F^-T = E^T and tr(F^-1 dF) = tr(E A),
// so the corner matrix (V dP) G^T of Tangent<NEO_HOOKEAN> is
//     cmu A (G G^T) + cc (E A E)^T + clam tr(E A) E^T
// (cmu = V mu, clam = V lam, cc = V (mu - lam ln J)): about 150 operations per
// product instead of about 230 that rebuild F, F^-1 and ln J from the
// displacements and apply the tangent to dF.
template <typename T>
__device__ __forceinline__ void nh_lin_state(const T* g, const T* F, T cmu, T clam, T* E, T& cc)
{
    T Fi[9];
    const T J = inv3(F, Fi);
    mul3(g, Fi, E);
    cc = cmu - clam * dlog(J);
}

template <typename T>
__device__ __forceinline__ void nh_lin_corners(const T* g, const T* E, T cmu, T clam, T cc, const T w[4][3], T out[4][3])
{
    T a[3][3];   // a[k] = column k of A
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int c = 0; c < 3; c++) a[k][c] = w[k + 1][c] - w[0][c];
    T S[3][3];   // G G^T
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i; j < 3; j++) {
            S[i][j] = g[3 * i] * g[3 * j] + g[3 * i + 1] * g[3 * j + 1] + g[3 * i + 2] * g[3 * j + 2];
            S[j][i] = S[i][j];
        }
    T Y[3][3];   // E A
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) Y[i][k] = E[3 * i] * a[k][0] + E[3 * i + 1] * a[k][1] + E[3 * i + 2] * a[k][2];
    const T lt = clam * (Y[0][0] + Y[1][1] + Y[2][2]);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        // column j of the corner matrix: cmu (A S)(:,j) + cc (Y E)(j,:)^T + lt E(j,:)^T
        T z[3];
#pragma unroll
        for (int c = 0; c < 3; c++) z[c] = Y[j][0] * E[c] + Y[j][1] * E[3 + c] + Y[j][2] * E[6 + c];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T m1 = a[0][c] * S[0][j] + a[1][c] * S[1][j] + a[2][c] * S[2][j];
            out[j + 1][c] = cmu * m1 + cc * z[c] + lt * E[3 * j + c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) out[0][c] = -(out[1][c] + out[2][c] + out[3][c]);
}

}  // namespace sfem
