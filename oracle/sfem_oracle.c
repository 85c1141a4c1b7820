/*
 * sfem_oracle.c -- CPU FP64 restatement of SolidFEM's matrix-free Newton-CG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the FP64 reference that the HIP product
 * path is compared against; it is never linked into, loaded by, or called from the
 * product library (solidfem_amd/).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it.
 *
 * Every function names the reference lines it restates.  Matrices are 3x3
 * row-major (m[3*i+j] = M(i,j)); "column k" of a matrix is what the reference's
 * Matrix3R::operator()(k) returns (Math/matrix3.h:173-179).
 *
 * Node vectors are full length (nn*3) in the caller's ORIGINAL node order; rows
 * of fixed nodes are zero wherever the reference drops them.  The reference
 * reorders nodes (fixed first, FemPhysicsMatrixFree.cpp:229-263); that
 * permutation changes nothing but the order of floating-point sums.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   geometry / mass     : box2.tet fixture, sum of volumes, sum of masses = rho V
 *   stress P(F)         : -dPsi/dx by central differences on orc_energy, and the
 *                         compiled reference (oracle/_ref) internal forces
 *   dstress dP(F,dF)    : central differences of P for LINEAR/STVK/NH, compiled
 *                         reference matvec for all four (COROT's tangent is not
 *                         the derivative of its force: only the reference pins it)
 *   cg / newton_cg_ref  : dense direct solve on small meshes, n-step termination,
 *                         compiled reference trajectories (iteration counts, x)
 *   loads               : closed-form totals (rho V g, p * sum(area n), torque)
 *   block_diag          : brute-force K e_j probes
 *   COROTATIONAL_POLAR  : no reference counterpart (the corotational model
 *                         with the polar rotation).  The rotation against
 *                         scipy.linalg.polar (SVD), P against -dPsi/dx by
 *                         central differences, dP against central differences
 *                         of P and K against its transpose.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { MAT_LINEAR = 0, MAT_STVK = 1, MAT_COROTATIONAL = 2, MAT_NEO_HOOKEAN = 3, MAT_COROTATIONAL_POLAR = 16 };

typedef struct {
    int64_t ne, nn;
    const int32_t* tets;   /* ne*4 node ids, original numbering              */
    const double* X0;      /* nn*3 rest positions                             */
    const uint8_t* fixed;  /* nn, 1 = Dirichlet node (invMass == 0)           */
    const double* DmInv;   /* ne*9 Dm^-1, row-major                           */
    const double* vol;     /* ne signed rest volume                           */
    const double* mu;      /* ne shear modulus per element                    */
    const double* lam;     /* ne Lame lambda per element                      */
    const double* mass;    /* nn lumped mass as the reference uses it         */
    int32_t material;
    int32_t nthreads;      /* <=0: OpenMP default                             */
    double grav[3];
    int64_t ntri;          /* follower-pressure triangles (original ids)      */
    const int32_t* tris;
    double pressure;
    const double* fext;    /* nn*3 dead external nodal forces, or NULL        */
    /* node -> (element*4 + corner) incidence (CSR), optional; enables the
       parallel deterministic gather used by the large-mesh solve            */
    const int64_t* inc_ptr;
    const int32_t* inc_idx;
} orc_problem;

/* ---------------------------------------------------------------- 3x3 math */
/* Math/matrix3.h conventions: column constructor :101-109, GetInverse
   (adjugate / det of row-0 cofactor expansion) :234-249, Determinant :206-212. */

static inline void m3_cols(const double* a, const double* b, const double* c, double* M)
{
    for (int i = 0; i < 3; i++) { M[3*i+0] = a[i]; M[3*i+1] = b[i]; M[3*i+2] = c[i]; }
}
static inline void m3_mul(const double* A, const double* B, double* C)
{
    double T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            T[3*i+j] = A[3*i+0]*B[0*3+j] + A[3*i+1]*B[1*3+j] + A[3*i+2]*B[2*3+j];
    memcpy(C, T, sizeof T);
}
static inline void m3_tr(const double* A, double* T)
{
    double U[9];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) U[3*i+j] = A[3*j+i];
    memcpy(T, U, sizeof U);
}
static inline double m3_trace(const double* A) { return A[0] + A[4] + A[8]; }
static inline double m3_det(const double* m)
{
    double a00 = m[4]*m[8] - m[5]*m[7];
    double a01 = -m[3]*m[8] + m[6]*m[5];
    double a02 = m[3]*m[7] - m[6]*m[4];
    return m[0]*a00 + m[1]*a01 + m[2]*a02;
}
static inline void m3_inv(const double* m, double* inv)
{
    double v[9];
    v[0] = m[4]*m[8] - m[5]*m[7];
    v[3] = -m[3]*m[8] + m[6]*m[5];
    v[6] = m[3]*m[7] - m[6]*m[4];
    v[1] = m[2]*m[7] - m[1]*m[8];
    v[4] = m[0]*m[8] - m[2]*m[6];
    v[7] = m[1]*m[6] - m[0]*m[7];
    v[2] = m[1]*m[5] - m[2]*m[4];
    v[5] = m[2]*m[3] - m[0]*m[5];
    v[8] = m[0]*m[4] - m[1]*m[3];
    double det = m[0]*v[0] + m[1]*v[3] + m[2]*v[6];
    double s = 1.0 / det;
    for (int k = 0; k < 9; k++) inv[k] = v[k] * s;
}
static inline void v3_normalize(double* v)
{
    double len = sqrt(v[0]*v[0] + v[1]*v[1] + v[2]*v[2]);
    double s = 1.0 / len;   /* Vector3T::Normalize scales by 1/len, vector3.h:109-114 */
    v[0] *= s; v[1] *= s; v[2] *= s;
}

/* PolarDecomposition::QRPolarDecomposition (Source/PolarDecomposition.cpp:269-280)
   followed by the symmetric part U = (R^T F + F^T R)/2 (:37-47).  Gram-Schmidt
   on columns 0 and 1 of F; r2 = r0 x r1.  Not a polar decomposition.          */
static void qr_rotation(const double* F, double* R, double* S)
{
    double r0[3] = { F[0], F[3], F[6] };
    v3_normalize(r0);
    double a1[3] = { F[1], F[4], F[7] };
    double d = a1[0]*r0[0] + a1[1]*r0[1] + a1[2]*r0[2];
    double r1[3] = { a1[0] - d*r0[0], a1[1] - d*r0[1], a1[2] - d*r0[2] };
    v3_normalize(r1);
    double r2[3] = { r0[1]*r1[2] - r1[1]*r0[2], -r0[0]*r1[2] + r0[2]*r1[0], r0[0]*r1[1] - r1[0]*r0[1] };
    v3_normalize(r2);
    m3_cols(r0, r1, r2, R);
    if (S) {
        double Rt[9], V[9], Vt[9];
        m3_tr(R, Rt);
        m3_mul(Rt, F, V);
        m3_tr(V, Vt);
        for (int k = 0; k < 9; k++) S[k] = 0.5 * (V[k] + Vt[k]);
    }
}

/* Orthogonal factor of the polar decomposition F = R U by Newton's iteration
   X <- (X + X^-T)/2 (Higham 1986), scaled by g = sqrt(|X^-1|_F / |X|_F)
   while the update is large, run until the update stops shrinking; then
   S = (R^T F + F^T R)/2.  COROTATIONAL_POLAR only (no reference counterpart). */
static void polar_rotation(const double* F, double* R, double* S)
{
    double X[9], prev = INFINITY;
    memcpy(X, F, sizeof X);
    for (int it = 0; it < 100; it++) {
        double Xi[9], nx = 0, ni = 0, du = 0;
        m3_inv(X, Xi);
        for (int k = 0; k < 9; k++) { nx += X[k] * X[k]; ni += Xi[k] * Xi[k]; }
        const double g = prev > 1e-2 ? sqrt(sqrt(ni / nx)) : 1.0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const double xn = 0.5 * (g * X[3*i+j] + Xi[3*j+i] / g);
                du += (xn - X[3*i+j]) * (xn - X[3*i+j]);
                X[3*i+j] = xn;
            }
        du = sqrt(du);
        if (du >= prev || du < 1e-15) break;
        prev = du;
    }
    memcpy(R, X, sizeof X);
    double Rt[9], V[9], Vt[9];
    m3_tr(R, Rt);
    m3_mul(Rt, F, V);
    m3_tr(V, Vt);
    for (int k = 0; k < 9; k++) S[k] = 0.5 * (V[k] + Vt[k]);
}

/* ------------------------------------------------------ constitutive laws */

/* First Piola-Kirchhoff stress P(F): ElasticEnergy::ComputeElementStress
   (Source/ElasticEnergy.cpp:114-187) for LINEAR (:141-146,159-163), STVK
   (:148-158), COROTATIONAL (:132-137,164-167) and NEO_HOOKEAN (:168-187).      */
void orc_stress(int mat, double mu, double lam, const double* F_in, double* P)
{
    double F[9], R[9];
    memcpy(F, F_in, sizeof F);
    if (mat == MAT_COROTATIONAL || mat == MAT_COROTATIONAL_POLAR) {
        double S[9];
        if (mat == MAT_COROTATIONAL) qr_rotation(F_in, R, S);
        else polar_rotation(F_in, R, S);
        memcpy(F, S, sizeof F);     /* F = U (:133-136) */
    }
    if (mat == MAT_LINEAR || mat == MAT_COROTATIONAL || mat == MAT_COROTATIONAL_POLAR) {
        double e[9], Ft[9];
        m3_tr(F, Ft);
        for (int k = 0; k < 9; k++) e[k] = 0.5 * (F[k] + Ft[k]);
        e[0] -= 1; e[4] -= 1; e[8] -= 1;
        double tr = m3_trace(e), T[9];
        for (int k = 0; k < 9; k++) T[k] = 2 * mu * e[k];
        T[0] += lam * tr; T[4] += lam * tr; T[8] += lam * tr;
        if (mat == MAT_LINEAR) memcpy(P, T, sizeof T);
        else m3_mul(R, T, P);
    } else if (mat == MAT_STVK) {
        double Ft[9], C[9], T[9];
        m3_tr(F, Ft);
        m3_mul(Ft, F, C);
        C[0] -= 1; C[4] -= 1; C[8] -= 1;
        for (int k = 0; k < 9; k++) C[k] *= 0.5;       /* Green strain E */
        double tr = m3_trace(C);
        for (int k = 0; k < 9; k++) T[k] = 2 * mu * C[k];
        T[0] += lam * tr; T[4] += lam * tr; T[8] += lam * tr;
        m3_mul(F, T, P);
    } else { /* NEO_HOOKEAN */
        double Fi[9], Fit[9];
        m3_inv(F, Fi);
        m3_tr(Fi, Fit);
        double J = m3_det(F), lJ = log(J);
        for (int k = 0; k < 9; k++) P[k] = mu * (F[k] - Fit[k]) + lam * (lJ * Fit[k]);
    }
}

/* Stress differential dP(F; dF): FemPhysicsMatrixFree::ComputeForceDifferential
   material branches (Source/FemPhysicsMatrixFree.cpp:436-470).               */
void orc_dstress(int mat, double mu, double lam, const double* F, const double* dF, double* dP)
{
    if (mat == MAT_LINEAR) {                                   /* :436-440 */
        double tr = m3_trace(dF);
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++)
            dP[3*i+j] = mu * (dF[3*i+j] + dF[3*j+i]);
        dP[0] += lam * tr; dP[4] += lam * tr; dP[8] += lam * tr;
    } else if (mat == MAT_COROTATIONAL) {                      /* :441-456 */
        double R[9], S[9], Rt[9], dS[9];
        qr_rotation(F, R, S);
        m3_tr(R, Rt);
        m3_mul(Rt, dF, dS);
        double trdS = m3_trace(dS);
        for (int k = 0; k < 9; k++) dP[k] = 2 * mu * dF[k] + lam * trdS * R[k];
        double trS = m3_trace(S);
        double A[9], Ai[9];
        memcpy(A, S, sizeof A);
        A[0] -= trS; A[4] -= trS; A[8] -= trS;
        double w[3] = { dS[5] - dS[7], dS[2] - dS[6], dS[1] - dS[3] };
        m3_inv(A, Ai);
        double r[3];
        for (int i = 0; i < 3; i++) r[i] = Ai[3*i+0]*w[0] + Ai[3*i+1]*w[1] + Ai[3*i+2]*w[2];
        double X[9] = { 0, -r[2], r[1],  r[2], 0, -r[0],  -r[1], r[0], 0 };  /* Skew, matrix3.h:264-269 */
        double dR[9];
        m3_mul(R, X, dR);
        double c = lam * (trS - 3) - 2 * mu;
        for (int k = 0; k < 9; k++) dP[k] += c * dR[k];
    } else if (mat == MAT_COROTATIONAL_POLAR) {
        /* d/dF of R (2 mu (S - I) + lam tr(S - I) I) with F = R S: R^T dR = [w]x
           where (S - tr(S) I) w = -axial(R^T dF - dF^T R), so that
           dP = 2 mu dF + lam tr(R^T dF) R + (lam (tr S - 3) - 2 mu) dR.  An
           element with det F <= 0 takes the tangent at F = I.               */
        double R[9], S[9], Rt[9], K[9];
        if (m3_det(F) > 0) polar_rotation(F, R, S);
        else {
            for (int k = 0; k < 9; k++) R[k] = S[k] = (k % 4 == 0) ? 1.0 : 0.0;
        }
        m3_tr(R, Rt);
        m3_mul(Rt, dF, K);
        const double trS = m3_trace(S), trK = m3_trace(K);
        double A[9], Ai[9];
        memcpy(A, S, sizeof A);
        A[0] -= trS; A[4] -= trS; A[8] -= trS;
        m3_inv(A, Ai);
        const double ax[3] = { K[7] - K[5], K[2] - K[6], K[3] - K[1] };
        double w[3];
        for (int i = 0; i < 3; i++) w[i] = -(Ai[3*i+0]*ax[0] + Ai[3*i+1]*ax[1] + Ai[3*i+2]*ax[2]);
        double W[9] = { 0, -w[2], w[1],  w[2], 0, -w[0],  -w[1], w[0], 0 };
        double dR[9];
        m3_mul(R, W, dR);
        const double c = lam * (trS - 3) - 2 * mu;
        for (int k = 0; k < 9; k++) dP[k] = 2 * mu * dF[k] + lam * trK * R[k] + c * dR[k];
    } else if (mat == MAT_STVK) {                              /* :457-463 */
        double Ft[9], dFt[9], E[9], dE[9], a[9], b[9], T1[9], T2[9];
        m3_tr(F, Ft); m3_tr(dF, dFt);
        m3_mul(Ft, F, E);
        E[0] -= 1; E[4] -= 1; E[8] -= 1;
        for (int k = 0; k < 9; k++) E[k] *= 0.5;
        m3_mul(dFt, F, a);
        m3_mul(Ft, dF, b);
        for (int k = 0; k < 9; k++) dE[k] = 0.5 * (a[k] + b[k]);
        double trE = m3_trace(E), trdE = m3_trace(dE);
        for (int k = 0; k < 9; k++) { a[k] = 2 * mu * E[k]; b[k] = 2 * mu * dE[k]; }
        a[0] += lam * trE; a[4] += lam * trE; a[8] += lam * trE;
        b[0] += lam * trdE; b[4] += lam * trdE; b[8] += lam * trdE;
        m3_mul(dF, a, T1);
        m3_mul(F, b, T2);
        for (int k = 0; k < 9; k++) dP[k] = T1[k] + T2[k];
    } else {                                                   /* :464-470 */
        double Fi[9], Fit[9], dFt[9], T[9], FidF[9];
        m3_inv(F, Fi);
        m3_tr(Fi, Fit);
        double J = m3_det(F);
        double c = mu - lam * log(J), cF[9];
        for (int k = 0; k < 9; k++) cF[k] = c * Fit[k];
        m3_tr(dF, dFt);
        m3_mul(cF, dFt, T);
        m3_mul(T, Fit, T);
        m3_mul(Fi, dF, FidF);
        double tr = m3_trace(FidF);
        for (int k = 0; k < 9; k++) dP[k] = mu * dF[k] + T[k] + lam * tr * Fit[k];
    }
}

/* Strain energy density Psi(F) (ElasticEnergy::ComputeElementEnergy,
   Source/ElasticEnergy.cpp:40-87): LINEAR/COROT/STVK mu e:e + lam/2 tr(e)^2,
   NH mu/2 (I1-3) - mu ln J + lam/2 (ln J)^2.                                   */
double orc_energy_density(int mat, double mu, double lam, const double* F_in)
{
    double F[9];
    memcpy(F, F_in, sizeof F);
    if (mat == MAT_COROTATIONAL || mat == MAT_COROTATIONAL_POLAR) {
        double R[9], S[9];
        if (mat == MAT_COROTATIONAL) qr_rotation(F_in, R, S);
        else polar_rotation(F_in, R, S);
        memcpy(F, S, sizeof F);
    }
    if (mat == MAT_NEO_HOOKEAN) {
        double Ft[9], C[9];
        m3_tr(F, Ft); m3_mul(Ft, F, C);
        double J = m3_det(F), lJ = log(J);
        return 0.5 * mu * (m3_trace(C) - 3) - mu * lJ + 0.5 * lam * lJ * lJ;
    }
    double e[9];
    if (mat == MAT_STVK) {
        double Ft[9];
        m3_tr(F, Ft); m3_mul(Ft, F, e);
        e[0] -= 1; e[4] -= 1; e[8] -= 1;
        for (int k = 0; k < 9; k++) e[k] *= 0.5;
    } else {
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) e[3*i+j] = 0.5 * (F[3*i+j] + F[3*j+i]);
        e[0] -= 1; e[4] -= 1; e[8] -= 1;
    }
    double ee = 0;
    for (int k = 0; k < 9; k++) ee += e[k] * e[k];
    double tr = m3_trace(e);
    return mu * ee + 0.5 * lam * tr * tr;
}

/* --------------------------------------------------------------- geometry */

/* Rest shape matrix Dm = [X1-X0 | X2-X0 | X3-X0], Dm^-1, signed volume
   det(Dm)/6 (Source/FemPhysicsMatrixFree.cpp:126-139).                         */
void orc_geometry(const orc_problem* p, double* DmInv, double* vol)
{
    for (int64_t e = 0; e < p->ne; e++) {
        const int32_t* t = p->tets + 4 * e;
        const double* x0 = p->X0 + 3 * (int64_t)t[0];
        double d[3][3];
        for (int k = 1; k < 4; k++)
            for (int c = 0; c < 3; c++) d[k-1][c] = p->X0[3 * (int64_t)t[k] + c] - x0[c];
        double Dm[9];
        m3_cols(d[0], d[1], d[2], Dm);
        m3_inv(Dm, DmInv + 9 * e);
        vol[e] = m3_det(Dm) / 6.0;
    }
}

/* Lumped masses m_i = sum_e rho V_e / 4, and the value the reference then uses,
   1/invMass = 1/(1/m) (FemPhysicsMatrixFree.cpp:122,141-146,200-205,530,547).
   Free nodes touched by no element keep invMass = 1 (Node ctor).               */
void orc_lumped_mass(const orc_problem* p, double density, double* mass)
{
    double* m = (double*)calloc((size_t)p->nn, sizeof(double));
    for (int64_t e = 0; e < p->ne; e++) {
        double lm = 0.25 * p->vol[e] * density;
        for (int k = 0; k < 4; k++) m[p->tets[4*e+k]] += lm;
    }
    for (int64_t i = 0; i < p->nn; i++) {
        if (p->fixed[i]) { mass[i] = 0; continue; }
        double inv = (m[i] != 0) ? 1.0 / m[i] : 1.0;
        mass[i] = 1.0 / inv;
    }
    free(m);
}

static inline void def_grad(const orc_problem* p, int64_t e, const double* x, double* F)
{
    /* FemPhysicsMatrixFree::ComputeDeformationGradient (:354-368): F = Ds Dm^-1 */
    const int32_t* t = p->tets + 4 * e;
    const double* x0 = x + 3 * (int64_t)t[0];
    double d[3][3], Ds[9];
    for (int k = 1; k < 4; k++)
        for (int c = 0; c < 3; c++) d[k-1][c] = x[3 * (int64_t)t[k] + c] - x0[c];
    m3_cols(d[0], d[1], d[2], Ds);
    m3_mul(Ds, p->DmInv + 9 * e, F);
}

static void set_threads(const orc_problem* p)
{
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#else
    (void)p;
#endif
}

/* H = -V P Dm^-T: columns 0..2 are the forces on corners 1..3, corner 0 gets
   minus their sum (ElasticEnergy.cpp:265-280, FemPhysicsMatrixFree.cpp:479-495) */
static inline void element_H(const double* P, const double* DmInv, double vol, double* H)
{
    double S[9];
    for (int k = 0; k < 9; k++) S[k] = -vol * P[k];      /* (-V P) first, then * Dm^-T */
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++)
            H[3*i+k] = S[3*i+0]*DmInv[3*k+0] + S[3*i+1]*DmInv[3*k+1] + S[3*i+2]*DmInv[3*k+2];
}

/* Internal elastic forces f (all nodes, original numbering): ElasticEnergy::
   ComputeForces (Source/ElasticEnergy.cpp:259-282) with F, F^-1, det cached by
   FemPhysicsBase::ComputeDeformationGradients (Source/FemPhysicsBase.cpp:189-201). */
static void accumulate_internal_forces(const orc_problem* p, const double* x, double* f)
{
    for (int64_t e = 0; e < p->ne; e++) {
        double F[9], P[9], H[9];
        def_grad(p, e, x, F);
        orc_stress(p->material, p->mu[e], p->lam[e], F, P);
        element_H(P, p->DmInv + 9 * e, p->vol[e], H);
        const int32_t* t = p->tets + 4 * e;
        for (int j = 1; j < 4; j++)
            for (int c = 0; c < 3; c++) {
                f[3 * (int64_t)t[j] + c] += H[3*c + (j-1)];
                f[3 * (int64_t)t[0] + c] -= H[3*c + (j-1)];
            }
    }
}

void orc_internal_forces(const orc_problem* p, const double* x, double* f)
{
    memset(f, 0, sizeof(double) * 3 * (size_t)p->nn);
    accumulate_internal_forces(p, x, f);
}

/* Total elastic energy sum_e V_e Psi(F_e) (FemPhysicsBase::ComputeElasticEnergy,
   Source/FemPhysicsBase.cpp:203-212).                                          */
double orc_energy(const orc_problem* p, const double* x)
{
    double s = 0;
    for (int64_t e = 0; e < p->ne; e++) {
        double F[9];
        def_grad(p, e, x, F);
        s += p->vol[e] * orc_energy_density(p->material, p->mu[e], p->lam[e], F);
    }
    return s;
}

/* Per-element dH = -V dP Dm^-T for direction dx; fixed nodes contribute dx = 0
   (FemPhysicsMatrixFree.cpp:417-480).                                          */
static inline void element_dH(const orc_problem* p, int64_t e, const double* x, const double* dx, double* dH)
{
    const int32_t* t = p->tets + 4 * e;
    double u[4][3];
    for (int k = 0; k < 4; k++)
        for (int c = 0; c < 3; c++)
            u[k][c] = p->fixed[t[k]] ? 0.0 : dx[3 * (int64_t)t[k] + c];
    double d[3][3], M[9], dF[9], F[9], dP[9];
    for (int k = 1; k < 4; k++)
        for (int c = 0; c < 3; c++) d[k-1][c] = u[k][c] - u[0][c];
    m3_cols(d[0], d[1], d[2], M);
    m3_mul(M, p->DmInv + 9 * e, dF);
    def_grad(p, e, x, F);
    orc_dstress(p->material, p->mu[e], p->lam[e], F, dF, dP);
    element_H(dP, p->DmInv + 9 * e, p->vol[e], dH);
}

/* Static MatrixVectorMultiply: q = -df = K dx on free rows, 0 on fixed rows
   (FemPhysicsMatrixFree.cpp:406-496 and :523-534 with mTimeStep = 0).
   Loop structure follows the reference: an OpenMP element loop writes dH per
   element, then one serial scatter pass.  With p->inc_ptr set, the scatter is
   replaced by a parallel fixed-order gather over the node incidence lists
   (same sums, deterministic, used for large meshes).                          */
void orc_matvec(const orc_problem* p, const double* x, const double* dx, double* q)
{
    set_threads(p);
    double* dH = (double*)malloc(sizeof(double) * 9 * (size_t)p->ne);
    int64_t ne = p->ne;
    #pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; e++) element_dH(p, e, x, dx, dH + 9 * e);

    if (!p->inc_ptr) {
        memset(q, 0, sizeof(double) * 3 * (size_t)p->nn);
        for (int64_t e = 0; e < ne; e++) {
            const int32_t* t = p->tets + 4 * e;
            double f3[3] = { 0, 0, 0 };
            for (int j = 1; j < 4; j++)
                for (int c = 0; c < 3; c++) {
                    double v = dH[9*e + 3*c + (j-1)];
                    if (!p->fixed[t[j]]) q[3 * (int64_t)t[j] + c] += v;
                    f3[c] -= v;
                }
            if (!p->fixed[t[0]])
                for (int c = 0; c < 3; c++) q[3 * (int64_t)t[0] + c] += f3[c];
        }
        for (int64_t i = 0; i < 3 * p->nn; i++) q[i] = -q[i];
    } else {
        int64_t nn = p->nn;
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nn; i++) {
            double s[3] = { 0, 0, 0 };
            if (!p->fixed[i]) {
                for (int64_t k = p->inc_ptr[i]; k < p->inc_ptr[i+1]; k++) {
                    int64_t e = p->inc_idx[k] >> 2;
                    int corner = p->inc_idx[k] & 3;
                    for (int c = 0; c < 3; c++) {
                        const double* h = dH + 9*e + 3*c;
                        s[c] += corner ? h[corner-1] : -(h[0] + h[1] + h[2]);
                    }
                }
            }
            for (int c = 0; c < 3; c++) q[3*i + c] = -s[c];
        }
    }
    free(dH);
}

/* Follower pressure loads f_i += frac p (A/3) n_hat per triangle vertex
   (FemPhysicsMatrixFree::ComputePressureForces, :967-1013).                    */
void orc_pressure_forces(const orc_problem* p, const double* x, double frac, double* f)
{
    for (int64_t t = 0; t < p->ntri; t++) {
        const int32_t* tr = p->tris + 3 * t;
        const double* p1 = x + 3 * (int64_t)tr[0];
        const double* p2 = x + 3 * (int64_t)tr[1];
        const double* p3 = x + 3 * (int64_t)tr[2];
        double a[3], b[3];
        for (int c = 0; c < 3; c++) { a[c] = p2[c] - p1[c]; b[c] = p3[c] - p1[c]; }
        double n[3] = { a[1]*b[2] - b[1]*a[2], -a[0]*b[2] + a[2]*b[0], a[0]*b[1] - b[0]*a[1] };
        double area = 0.5 * sqrt(n[0]*n[0] + n[1]*n[1] + n[2]*n[2]);
        v3_normalize(n);
        for (int c = 0; c < 3; c++) {
            double force = frac * ((area / 3.0) * (p->pressure * n[c]));
            for (int k = 0; k < 3; k++) f[3 * (int64_t)tr[k] + c] += force;
        }
    }
}

/* External load r_ext = f_press(x) + frac (m g + f_dead) on free rows.  The
   dead-load term is the SetExternalForces extension (FemPhysicsBase.h:68).    */
void orc_external(const orc_problem* p, const double* x, double frac, double* r)
{
    memset(r, 0, sizeof(double) * 3 * (size_t)p->nn);
    orc_pressure_forces(p, x, frac, r);
    for (int64_t i = 0; i < p->nn; i++) {
        if (p->fixed[i]) { r[3*i] = r[3*i+1] = r[3*i+2] = 0; continue; }
        for (int c = 0; c < 3; c++) {
            r[3*i+c] += p->mass[i] * frac * p->grav[c];
            if (p->fext) r[3*i+c] += frac * p->fext[3*i+c];
        }
    }
}

/* Negative energy gradient r = f_int + f_press + frac (m g + f_dead) on free
   nodes, 0 on fixed ones (FemPhysicsMatrixFree::ComputeGradients, :536-560). */
void orc_gradients(const orc_problem* p, const double* x, double frac, double* r)
{
    /* same accumulation order as the reference: pressure, then element
       forces into the same array, then the body/dead load per free node      */
    memset(r, 0, sizeof(double) * 3 * (size_t)p->nn);
    orc_pressure_forces(p, x, frac, r);
    accumulate_internal_forces(p, x, r);
    for (int64_t i = 0; i < p->nn; i++) {
        if (p->fixed[i]) { r[3*i] = r[3*i+1] = r[3*i+2] = 0; continue; }
        for (int c = 0; c < 3; c++) {
            r[3*i+c] += p->mass[i] * frac * p->grav[c];
            if (p->fext) r[3*i+c] += frac * p->fext[3*i+c];
        }
    }
}

/* 3x3 diagonal blocks of K per node, K_ii = sum over incident elements of the
   corner block, probed with unit directions the way ElasticEnergy.cpp:493-517
   builds local stiffness.  Fixed nodes get the identity.  B is nn*9 row-major. */
void orc_block_diag(const orc_problem* p, const double* x, double* B)
{
    memset(B, 0, sizeof(double) * 9 * (size_t)p->nn);
    for (int64_t e = 0; e < p->ne; e++) {
        const int32_t* t = p->tets + 4 * e;
        const double* Xi = p->DmInv + 9 * e;
        double g[4][3];   /* gradients of the shape functions: rows of Dm^-1 */
        for (int c = 0; c < 3; c++) {
            g[1][c] = Xi[0*3+c]; g[2][c] = Xi[1*3+c]; g[3][c] = Xi[2*3+c];
            g[0][c] = -(g[1][c] + g[2][c] + g[3][c]);
        }
        double F[9];
        def_grad(p, e, x, F);
        for (int a = 0; a < 4; a++) {
            if (p->fixed[t[a]]) continue;
            for (int k = 0; k < 3; k++) {
                double dF[9] = { 0 }, dP[9], dH[9];
                for (int c = 0; c < 3; c++) dF[3*k + c] = g[a][c];
                orc_dstress(p->material, p->mu[e], p->lam[e], F, dF, dP);
                element_H(dP, Xi, p->vol[e], dH);
                for (int c = 0; c < 3; c++) {
                    double fa = a ? dH[3*c + a-1] : -(dH[3*c] + dH[3*c+1] + dH[3*c+2]);
                    B[9 * (int64_t)t[a] + 3*c + k] -= fa;   /* K = -df/dx */
                }
            }
        }
    }
    for (int64_t i = 0; i < p->nn; i++)
        if (p->fixed[i]) { double* b = B + 9*i; memset(b, 0, 9*sizeof(double)); b[0] = b[4] = b[8] = 1; }
}

static double dotv(int64_t n, const double* a, const double* b)
{
    /* InnerProduct (IterativeSolvers.h:94-101): per-node Vector3 dot, summed
       in node order (fixed rows are zero and add nothing)                     */
    double s = 0;
    for (int64_t i = 0; i + 2 < n; i += 3) s += a[i]*b[i] + a[i+1]*b[i+1] + a[i+2]*b[i+2];
    return s;
}

static void apply_blocks(int64_t nn, const double* Binv, const double* r, double* z)
{
    for (int64_t i = 0; i < nn; i++) {
        const double* M = Binv + 9 * i;
        const double* v = r + 3 * i;
        for (int c = 0; c < 3; c++) z[3*i+c] = M[3*c]*v[0] + M[3*c+1]*v[1] + M[3*c+2]*v[2];
    }
}

/* SolveConjugateGradientMF (Math/IterativeSolvers.h:115-195), Shewchuk CG.
   x is in/out.  eps is the reference's float tolerance: stop when
   |delta| <= eps*eps*|delta0| with eps*eps rounded in float.  Binv (nn*9)
   selects the block-preconditioned variant (z = Binv r, delta = r.z); NULL is
   the reference default (usePrecond = false).  Returns the iteration count.   */
static int cg_impl(const orc_problem* p, const double* xpos, const double* b, double* x,
                   int maxit, double tol2, int true_res, const double* Binv);

int orc_cg(const orc_problem* p, const double* xpos, const double* b, double* x,
           int maxit, float eps, const double* Binv)
{
    float e2 = eps * eps;     /* float product, as the reference evaluates it */
    return cg_impl(p, xpos, b, x, maxit, (double)e2, 0, Binv);
}

/* true_res = 0: the reference test |r.z| <= tol2 |r0.z0|;
   true_res = 1: ||r||^2 <= tol2 ||r0||^2 (the converged solver's test)       */
static int cg_impl(const orc_problem* p, const double* xpos, const double* b, double* x,
                   int maxit, double tol2, int true_res, const double* Binv)
{
    int64_t n = 3 * p->nn;
    double* r = (double*)malloc(sizeof(double) * n);
    double* z = (double*)malloc(sizeof(double) * n);
    double* d = (double*)malloc(sizeof(double) * n);
    double* q = (double*)malloc(sizeof(double) * n);
    memcpy(r, b, sizeof(double) * n);
    if (dotv(n, x, x) != 0) {
        orc_matvec(p, xpos, x, q);
        for (int64_t i = 0; i < n; i++) r[i] = b[i] - q[i];
    }
    if (Binv) apply_blocks(p->nn, Binv, r, z); else memcpy(z, r, sizeof(double) * n);
    memcpy(d, z, sizeof(double) * n);
    double delta = dotv(n, r, z);
    double rr0 = dotv(n, r, r);
    int it = 0;
    if (delta != 0) {
        double delta0 = delta;
        it = maxit;
        for (int iter = 0; iter < maxit; iter++) {
            orc_matvec(p, xpos, d, q);
            double alpha = delta / dotv(n, d, q);
            for (int64_t i = 0; i < n; i++) x[i] += alpha * d[i];
            for (int64_t i = 0; i < n; i++) r[i] -= alpha * q[i];
            if (Binv) apply_blocks(p->nn, Binv, r, z); else memcpy(z, r, sizeof(double) * n);
            double delta1 = dotv(n, r, z);
            double beta = delta1 / delta;
            delta = delta1;
            int conv = true_res ? (dotv(n, r, r) <= tol2 * rr0) : (fabs(delta) <= tol2 * fabs(delta0));
            if (conv) { it = iter + 1; break; }
            for (int64_t i = 0; i < n; i++) d[i] = z[i] + beta * d[i];
        }
    }
    free(r); free(z); free(d); free(q);
    return it;
}

/* FemPhysicsMatrixFree::SolveNewtonCG (:604-628): `outer` full Newton steps,
   each solving K delta = r from delta = 0 with unpreconditioned CG (eps 1e-3f,
   `inner` iterations at most); no residual test, no line search.
   cg_iters (length outer, may be NULL) receives the CG counts.                */
void orc_newton_cg_ref(const orc_problem* p, double* x, double frac, int outer, int inner, int* cg_iters)
{
    int64_t n = 3 * p->nn;
    double* b = (double*)malloc(sizeof(double) * n);
    double* dl = (double*)malloc(sizeof(double) * n);
    for (int it = 0; it < outer; it++) {
        orc_gradients(p, x, frac, b);
        memset(dl, 0, sizeof(double) * n);
        int k = orc_cg(p, x, b, dl, inner, 1e-3f, NULL);
        if (cg_iters) cg_iters[it] = k;
        for (int64_t i = 0; i < p->nn; i++)
            if (!p->fixed[i]) for (int c = 0; c < 3; c++) x[3*i+c] += dl[3*i+c];
    }
    free(b); free(dl);
}

static void invert_blocks(int64_t nn, const double* B, double* Binv)
{
    for (int64_t i = 0; i < nn; i++) m3_inv(B + 9 * i, Binv + 9 * i);
}

/* Converged Newton solve used as the FP64 oracle at scale (an extension: the
   reference has no residual test), the same algorithm as the device driver:
   block-Jacobi PCG inner solves with Eisenstat-Walker forcing (choice 2,
   gamma 0.9, capped at cg_rtol), a line search on the slope
   R(a) = dx . r(x + a dx) (accept |R(a)| <= R(0)/2 or R(a) R(0) > 0, else the
   secant root, 8 trials; cf. the bisection on the same quantity in
   NewtonSolverBackTrack, Include/NewtonSolver.h:229-293), stop when
   ||r|| <= tol ||r_ext||.  stats[0..3] = newton iters, total CG iters,
   final ||r||, ||r_ext||.  Returns 0 on convergence, 1 otherwise.            */
int orc_newton_solve(const orc_problem* p, double* x, double frac, double tol, int max_newton,
                     double cg_rtol, int max_cg, double* stats)
{
    int64_t n = 3 * p->nn;
    double* r = (double*)malloc(sizeof(double) * n);
    double* rt = (double*)malloc(sizeof(double) * n);
    double* ext = (double*)malloc(sizeof(double) * n);
    double* dl = (double*)malloc(sizeof(double) * n);
    double* xt = (double*)malloc(sizeof(double) * n);
    double* B = (double*)malloc(sizeof(double) * 9 * p->nn);
    double* Binv = (double*)malloc(sizeof(double) * 9 * p->nn);
    int total_cg = 0, it;
    orc_gradients(p, x, frac, r);
    orc_external(p, x, frac, ext);
    double rn = sqrt(dotv(n, r, r)), en = sqrt(dotv(n, ext, ext));
    double eta = cg_rtol, rn_prev = 0;
    for (it = 0; it < max_newton; it++) {
        if (rn <= tol * en) break;
        if (it > 0) {
            double ew = 0.9 * (rn / rn_prev) * (rn / rn_prev), safe = 0.9 * eta * eta;
            eta = safe > 0.1 ? (ew > safe ? ew : safe) : ew;
            if (eta > cg_rtol) eta = cg_rtol;
        }
        double floor = 0.5 * tol * en / rn;
        if (eta < floor) eta = floor;
        if (eta < 1e-12) eta = 1e-12;
        orc_block_diag(p, x, B);
        invert_blocks(p->nn, B, Binv);
        memset(dl, 0, sizeof(double) * n);
        total_cg += cg_impl(p, x, r, dl, max_cg, eta * eta, 1, Binv);
        double R0 = dotv(n, dl, r), alpha = 1.0, Ra = 0;
        int finite = 0;
        for (int ls = 0; ls < 8; ls++) {
            for (int64_t i = 0; i < n; i++) xt[i] = x[i] + alpha * dl[i];
            orc_gradients(p, xt, frac, rt);
            Ra = dotv(n, dl, rt);
            double rr = dotv(n, rt, rt);
            finite = isfinite(Ra) && isfinite(rr);
            if (ls == 7) break;
            if (finite && (fabs(Ra) <= 0.5 * fabs(R0) || Ra * R0 > 0)) break;
            if (finite) {
                double a = alpha * R0 / (R0 - Ra);
                if (a < 0.1 * alpha) a = 0.1 * alpha;
                if (a > 0.9 * alpha) a = 0.9 * alpha;
                alpha = a;
            } else alpha *= 0.5;
        }
        /* never accept a state with inverted elements: shorten the step until
           the residual is finite, or stop at the current state */
        for (int k = 0; !finite && k < 40; k++) {
            alpha *= 0.5;
            for (int64_t i = 0; i < n; i++) xt[i] = x[i] + alpha * dl[i];
            orc_gradients(p, xt, frac, rt);
            finite = isfinite(dotv(n, dl, rt)) && isfinite(dotv(n, rt, rt));
        }
        if (!finite) break;
        memcpy(x, xt, sizeof(double) * n);
        memcpy(r, rt, sizeof(double) * n);
        orc_external(p, x, frac, ext);
        rn_prev = rn;
        rn = sqrt(dotv(n, r, r));
        en = sqrt(dotv(n, ext, ext));
        if (!isfinite(rn)) break;
    }
    int ok = rn <= tol * en;
    if (stats) { stats[0] = it; stats[1] = total_cg; stats[2] = rn; stats[3] = en; }
    free(r); free(rt); free(ext); free(dl); free(xt); free(B); free(Binv);
    return ok ? 0 : 1;
}

/* ------------------------------------------------------------ cable springs */
/* One cable of n nodes, node j at sum_k w4[4j+k] x[node4[4j+k]] (the
   barycentric interpolation of FemPhysicsBase::ComputeSpringForces,
   FemPhysicsBase.cpp:222-245), springs j -> j+1 of rest length `rest`
   scaled by `act`.  Adds frac * the back-interpolated node forces to f (nn*3,
   all rows) following :247-300: Bern tension (unilateral, smoothed over
   |err| <= eps = rest / 100) plus damping along the spring from the
   interpolated velocities v (may be NULL: zero).  No forces when act == 0
   (:216-217).  Restated, not copied; pinned by the compiled reference
   (FemPhysicsMatrixFree::ComputeGradients with AddCable) and by
   f = -frac dE/dx of orc_cable_energy. */
static void cable_point(const int32_t* node4, const double* w4, int64_t j, const double* x, double* p)
{
    for (int q = 0; q < 3; q++) p[q] = 0.0;
    for (int k = 0; k < 4; k++) {
        const double wk = w4[4 * j + k];
        for (int q = 0; q < 3; q++) p[q] += wk * x[3 * (int64_t)node4[4 * j + k] + q];
    }
}

void orc_cable_forces(int64_t n, const int32_t* node4, const double* w4, double rest, double act, double k,
                      double damp, const double* x, const double* v, double frac, double* f)
{
    if (n < 2 || act == 0.0) return;
    double* fs = (double*)calloc((size_t)(3 * (n - 1)), sizeof(double));
    const double eps = rest * 0.01;
    for (int64_t j = 0; j + 1 < n; j++) {
        double pa[3], pb[3], va[3] = {0, 0, 0}, vb[3] = {0, 0, 0};
        cable_point(node4, w4, j, x, pa);
        cable_point(node4, w4, j + 1, x, pb);
        if (v) { cable_point(node4, w4, j, v, va); cable_point(node4, w4, j + 1, v, vb); }
        double y[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
        double len = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        double dir[3] = {y[0] / len, y[1] / len, y[2] / len};
        double err = len - rest * act;
        double tension = -2.0 * k * err;
        if (err < -eps) tension = 0.0;
        else if (err >= -eps && err <= eps) tension = -k * (0.5 * err * err / eps + err + 0.5 * eps);
        double dv = (vb[0] - va[0]) * dir[0] + (vb[1] - va[1]) * dir[1] + (vb[2] - va[2]) * dir[2];
        for (int q = 0; q < 3; q++) fs[3 * j + q] = (tension - damp * dv) * dir[q];
    }
    for (int64_t j = 0; j < n; j++) {
        double nf[3] = {0, 0, 0};
        for (int q = 0; q < 3; q++) {
            if (j + 1 < n) nf[q] -= fs[3 * j + q];
            if (j > 0) nf[q] += fs[3 * (j - 1) + q];
        }
        for (int kk = 0; kk < 4; kk++) {
            const double wk = w4[4 * j + kk];
            for (int q = 0; q < 3; q++) f[3 * (int64_t)node4[4 * j + kk] + q] += wk * (frac * nf[q]);
        }
    }
    free(fs);
}

/* FemPhysicsMatrixFree::ComputeSpringEnergy (FemPhysicsMatrixFree.cpp:1062-1111)
   for one cable, restated: sum over springs of k (err^2 + eps^2 / 3), 0 below
   -eps, the smoothed cubic k/2 (err^3 / (3 eps) + err^2 + eps err + eps^2 / 3)
   on |err| <= eps; 0 when act == 0 */
double orc_cable_energy(int64_t n, const int32_t* node4, const double* w4, double rest, double act, double k,
                        const double* x)
{
    if (n < 2 || act == 0.0) return 0.0;
    const double eps = rest * 0.01;
    double e = 0.0;
    for (int64_t j = 0; j + 1 < n; j++) {
        double pa[3], pb[3];
        cable_point(node4, w4, j, x, pa);
        cable_point(node4, w4, j + 1, x, pb);
        double y[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
        double len = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        double err = len - rest * act;
        double pot = k * (err * err + eps * eps / 3.0);
        if (err < -eps) pot = 0.0;
        else if (err >= -eps && err <= eps) pot = 0.5 * k * (err * err * err / 3.0 / eps + err * err + eps * err + eps * eps / 3.0);
        e += pot;
    }
    return e;
}
