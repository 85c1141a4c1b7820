// FemPhysicsMatrixFree.h -- C++ host API of the MI355X matrix-free solver,
// mirroring SolidFEM's FemPhysicsBase / FemPhysicsMatrixFree interface
// (Include/FemPhysicsBase.h:41-207, Include/FemPhysicsMatrixFree.h:40-155) and
// the data types it takes (Include/FemDataStructures.h:50-194).  The bodies
// call the C-ABI of libsfem_hip.so (include/sfem_capi.h); no HIP types appear
// here, so code written against the reference compiles against this header.
//
// Vectors passed to MatrixVectorMultiply / ComputeGradients /
// UpdatePosAndComputeGradients have GetNumFreeNodes() entries in the
// reference's shuffled order: fixed nodes first, then free nodes, each in
// original order (FemPhysicsMatrixFree.cpp:229-263).
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "sfem_capi.h"

// The reference keeps `real`, `uint32` in the global namespace and the math
// types in namespace Math, pulled into the global one by `using namespace
// Math` (Include/FemDataStructures.h:52-57, Engine/Types.h, Math/matrix3.h:38);
// the same arrangement here, so that templates written against it (the
// global std::vector operators of Math/IterativeSolvers.h found from inside
// FEM_SYSTEM) resolve the same way.
typedef double real;
typedef uint32_t uint32;

namespace Math {

// 3-vector with the reference's layout and operators (Math/vector3.h:40-275):
// x, y, z aliasing v[3]; Vector3T(a) fills all three; u * v is the dot
// product; scalars multiply from either side.  Written here so that code
// templated on the reference's vector type (Math/IterativeSolvers.h:115-195,
// Include/NonlinearCG.h:51-114) compiles against this header.
template <typename Real> struct Vector3T {
    union {
        struct { Real x, y, z; };
        Real v[3];
    };
    Vector3T() : v{0, 0, 0} {}
    Vector3T(Real a, Real b, Real c) : v{a, b, c} {}
    explicit Vector3T(Real a) : v{a, a, a} {}
    Vector3T(const Vector3T& o) : v{o.v[0], o.v[1], o.v[2]} {}
    Vector3T& operator=(const Vector3T& o) { v[0] = o.v[0]; v[1] = o.v[1]; v[2] = o.v[2]; return *this; }
    Real& operator[](int i) { return v[i]; }
    Real operator[](int i) const { return v[i]; }
    Real& operator()(int i) { return v[i]; }
    Real operator()(int i) const { return v[i]; }
    Real X() const { return v[0]; }
    Real Y() const { return v[1]; }
    Real Z() const { return v[2]; }
    void SetZero() { v[0] = v[1] = v[2] = 0; }
    void Set(Real a, Real b, Real c) { v[0] = a; v[1] = b; v[2] = c; }
    Real Dot(const Vector3T& o) const { return v[0] * o.v[0] + v[1] * o.v[1] + v[2] * o.v[2]; }
    Real LengthSquared() const { return Dot(*this); }
    Real Length() const { return std::sqrt(Dot(*this)); }
    void Scale(Real s) { v[0] *= s; v[1] *= s; v[2] *= s; }
    void Normalize() { Scale(Real(1) / Length()); }
    void Flip() { Scale(Real(-1)); }
    Vector3T Cross(const Vector3T& o) const
    {
        return Vector3T(v[1] * o.v[2] - v[2] * o.v[1], v[2] * o.v[0] - v[0] * o.v[2], v[0] * o.v[1] - v[1] * o.v[0]);
    }
    Vector3T& operator+=(const Vector3T& o) { v[0] += o.v[0]; v[1] += o.v[1]; v[2] += o.v[2]; return *this; }
    Vector3T& operator-=(const Vector3T& o) { v[0] -= o.v[0]; v[1] -= o.v[1]; v[2] -= o.v[2]; return *this; }
    Vector3T& operator*=(Real s) { Scale(s); return *this; }
};
template <typename R> Vector3T<R> operator+(Vector3T<R> a, const Vector3T<R>& b) { return a += b; }
template <typename R> Vector3T<R> operator-(Vector3T<R> a, const Vector3T<R>& b) { return a -= b; }
template <typename R> Vector3T<R> operator-(const Vector3T<R>& a) { return Vector3T<R>(-a[0], -a[1], -a[2]); }
template <typename R> Vector3T<R> operator*(double s, const Vector3T<R>& a) { return Vector3T<R>(R(s * a[0]), R(s * a[1]), R(s * a[2])); }
template <typename R> Vector3T<R> operator*(float s, const Vector3T<R>& a) { return Vector3T<R>(R(s * a[0]), R(s * a[1]), R(s * a[2])); }
template <typename R> Vector3T<R> operator*(const Vector3T<R>& a, double s) { return s * a; }
template <typename R> Vector3T<R> operator*(const Vector3T<R>& a, float s) { return s * a; }
template <typename R> R operator*(const Vector3T<R>& a, const Vector3T<R>& b) { return a.Dot(b); }   // vector3.h:243-247
template <typename R> Vector3T<R> cross(const Vector3T<R>& a, const Vector3T<R>& b) { return a.Cross(b); }
template <typename R> R triple(const Vector3T<R>& a, const Vector3T<R>& b, const Vector3T<R>& c) { return c.Dot(a.Cross(b)); }

// 3x3 with the reference's conventions (Math/matrix3.h:42-359): row storage
// m[row][col]; the default is the identity; Matrix3T(a, b, c) takes COLUMNS;
// Matrix3T(v) = diag(v); M(i) = column i, M[i] = row i, M(i, j) = element;
// !M = transpose; GetInverse = adjugate / det (row-0 cofactor expansion).
template <typename Real> struct Matrix3T {
    Real m[3][3];
    Matrix3T() : m{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}} {}
    Matrix3T(const Vector3T<Real>& c0, const Vector3T<Real>& c1, const Vector3T<Real>& c2)
    {
        for (int i = 0; i < 3; i++) { m[i][0] = c0[i]; m[i][1] = c1[i]; m[i][2] = c2[i]; }
    }
    Matrix3T(Real a00, Real a01, Real a02, Real a10, Real a11, Real a12, Real a20, Real a21, Real a22)
        : m{{a00, a01, a02}, {a10, a11, a12}, {a20, a21, a22}} {}
    explicit Matrix3T(const Vector3T<Real>& d) : m{{d[0], 0, 0}, {0, d[1], 0}, {0, 0, d[2]}} {}
    explicit Matrix3T(Real d) : m{{d, 0, 0}, {0, d, 0}, {0, 0, d}} {}
    static Matrix3T Identity() { return Matrix3T(); }
    static Matrix3T Zero() { return Matrix3T(Real(0)); }
    Real operator()(int i, int j) const { return m[i][j]; }
    Real& operator()(int i, int j) { return m[i][j]; }
    Vector3T<Real> operator()(int j) const { return Vector3T<Real>(m[0][j], m[1][j], m[2][j]); }
    Vector3T<Real> operator[](int i) const { return Vector3T<Real>(m[i][0], m[i][1], m[i][2]); }
    Vector3T<Real> GetDiagonal() const { return Vector3T<Real>(m[0][0], m[1][1], m[2][2]); }
    Real Trace() const { return m[0][0] + m[1][1] + m[2][2]; }
    Real Determinant() const
    {
        return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    }
    Matrix3T GetTranspose() const
    {
        Matrix3T t;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) t.m[i][j] = m[j][i];
        return t;
    }
    Matrix3T GetInverse() const
    {
        Matrix3T a(m[1][1] * m[2][2] - m[1][2] * m[2][1], m[0][2] * m[2][1] - m[0][1] * m[2][2], m[0][1] * m[1][2] - m[0][2] * m[1][1],
                   m[1][2] * m[2][0] - m[1][0] * m[2][2], m[0][0] * m[2][2] - m[0][2] * m[2][0], m[0][2] * m[1][0] - m[0][0] * m[1][2],
                   m[1][0] * m[2][1] - m[1][1] * m[2][0], m[0][1] * m[2][0] - m[0][0] * m[2][1], m[0][0] * m[1][1] - m[0][1] * m[1][0]);
        const Real s = Real(1) / Determinant();
        a.Scale(s);
        return a;
    }
    void Scale(Real s)
    {
        for (auto& r : m)
            for (Real& e : r) e *= s;
    }
    Matrix3T operator-() const
    {
        Matrix3T r = *this;
        r.Scale(Real(-1));
        return r;
    }
    static Matrix3T TensorProduct(const Vector3T<Real>& a, const Vector3T<Real>& b)
    {
        Matrix3T r;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) r.m[i][j] = a[i] * b[j];
        return r;
    }
    static Matrix3T Skew(const Vector3T<Real>& v) { return Matrix3T(0, -v[2], v[1], v[2], 0, -v[0], -v[1], v[0], 0); }
    static Real DoubleContraction(const Matrix3T& A, const Matrix3T& B)
    {
        Real s = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) s += A.m[i][j] * B.m[i][j];
        return s;
    }
};
template <typename R> Matrix3T<R> operator!(const Matrix3T<R>& A) { return A.GetTranspose(); }
template <typename R> Matrix3T<R> operator+(Matrix3T<R> A, const Matrix3T<R>& B)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A.m[i][j] += B.m[i][j];
    return A;
}
template <typename R> Matrix3T<R> operator-(Matrix3T<R> A, const Matrix3T<R>& B)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A.m[i][j] -= B.m[i][j];
    return A;
}
template <typename R> Matrix3T<R> operator*(const Matrix3T<R>& A, const Matrix3T<R>& B)
{
    Matrix3T<R> C(R(0));
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) C.m[i][j] += A.m[i][k] * B.m[k][j];
    return C;
}
template <typename R> Vector3T<R> operator*(const Matrix3T<R>& M, const Vector3T<R>& v)
{
    return Vector3T<R>(M.m[0][0] * v[0] + M.m[0][1] * v[1] + M.m[0][2] * v[2], M.m[1][0] * v[0] + M.m[1][1] * v[1] + M.m[1][2] * v[2],
                       M.m[2][0] * v[0] + M.m[2][1] * v[1] + M.m[2][2] * v[2]);
}
template <typename R> Matrix3T<R> operator*(double s, Matrix3T<R> M) { M.Scale(R(s)); return M; }
template <typename R> Matrix3T<R> operator*(float s, Matrix3T<R> M) { M.Scale(R(s)); return M; }

}  // namespace Math

using namespace Math;

namespace FEM_SYSTEM {

using ::real;
using ::uint32;
typedef Vector3T<real> Vector3R;
typedef Matrix3T<real> Matrix3R;

enum SimulationType { ST_STATIC, ST_QUASI_STATIC, ST_EXPLICIT, ST_IMPLICIT };
enum MaterialModelType {
    MMT_LINEAR, MMT_STVK, MMT_COROTATIONAL, MMT_NEO_HOOKEAN, MMT_NEO_HOOKEAN_OGDEN, MMT_DISTORTIONAL_LINEAR,
    MMT_DISTORTIONAL_MOONEY, MMT_DISTORTIONAL_NH2, MMT_DISTORTIONAL_NH3, MMT_DISTORTIONAL_BW, MMT_DISTORTIONAL_OGDEN,
    MMT_DISTORTIONAL_NH6, MMT_DISTORTIONAL_NH7,
    // no reference counterpart: corotational with the polar rotation and its
    // consistent tangent (SFEM_COROTATIONAL_POLAR)
    MMT_COROTATIONAL_POLAR = SFEM_COROTATIONAL_POLAR,
};
enum NonlinearSolverType {
    NST_NEWTON, NST_NEWTON_LS, NST_NEWTON_CG, NST_NONLINEAR_CG, NST_GRADIENT_DESCENT, NST_STEEPEST_DESCENT, NST_NLOPT,
    NST_GSL, NST_MCL, NST_MIXED_NEWTON, NST_MIXED_DUAL_ASCENT,
};
enum MethodType { MT_LINEAR_ELASTICITY, MT_COROTATIONAL_ELASTICITY, MT_NONLINEAR_ELASTICITY = 4 };

// FemConfig (FemDataStructures.h:131-156): same fields and defaults
struct FemConfig {
    real mYoungsModulus = 66000;
    real mPoissonRatio = 0.4f;
    real mDampingYoungsModulus = 0;
    real mDampingPoissonRatio = 0;
    int mType = MT_LINEAR_ELASTICITY;
    int mOrder = 1;
    real mDensity = 1070.f;
    real mGravity = -9.8f;
    SimulationType mSimType = ST_STATIC;
    int mNumSubsteps = 1;
    real mForceApplicationStep = 0.1;
    void* mCustomConfig = nullptr;
    MaterialModelType mMaterial = MMT_LINEAR;
    uint32 mOuterIterations = 10;
    uint32 mInnerIterations = 100;
    bool mHasCollisions = false;
    real mInvBulkModulus = -1;
    real mContactStiffness = 20000;
    uint32 mNumFixed = 0;
    real mAbsNewtonRsidualThreshold = 0.1;
    bool mVerbose = true;
    real mAppliedPressure = 0;
    bool mZUpAxis = false;
};

struct Node {
    Node() : invMass(1.f) {}
    Vector3R pos, vel, force;
    Vector3R pos0;
    real invMass;
};

struct Tet {
    uint32 idx[4];
};

// SpringNode, Cable (FemDataStructures.h:206-220): a cable node inside element
// `elem` at barycentric weights bary (the fourth is 1 - their sum)
struct SpringNode {
    Vector3R bary;
    int elem;
};

struct Cable {
    std::vector<SpringNode> mCableNodes;
    std::vector<Vector3R> mCablePositions;
    real mActuation = 1.0;
    real mCableRestLength;
    real mCableStiffness = 1000;
    real mCableDamping = 0;
};

class FemPhysicsBase {
public:
    explicit FemPhysicsBase(const FemConfig& config);
    virtual ~FemPhysicsBase() {}
    virtual void Step(real dt) = 0;
    virtual Vector3R GetDeformedPosition(uint32 idx, bool shuffle = true) const = 0;
    virtual uint32 GetNumNodes() const = 0;
    virtual uint32 GetNumFreeNodes() const = 0;
    virtual uint32 GetNumElements() const = 0;
    virtual bool IsNodeFixed(uint32 i) const = 0;
    virtual real GetTotalVolume() const = 0;
    virtual void SetExternalForces(const std::vector<Vector3R>& forces) = 0;
    virtual void SetBoundaryConditionsSurface(const std::vector<uint32>& triangleList, real pressure) = 0;

    MaterialModelType GetMaterial() const { return mMaterial; }
    real GetYoungsModulus() const { return mYoungsModulus; }
    real GetLameFirstParam() const { return mLameLambda; }
    real GetShearModulus() const { return mShearModulus; }
    real GetForceFraction() const { return mForceFraction; }

protected:
    real mYoungsModulus, mPoissonRatio, mLameLambda, mShearModulus;
    Vector3R mGravity;
    real mDensity;
    SimulationType mSimType;
    int mNumSteps = 1;   // FemConfig::mNumSubsteps
    MaterialModelType mMaterial;
    uint32 mOuterIterations, mInnerIterations;
    real mForceFraction = 0;
    real mForceStep;
    real mAbsNewtonResidualThreshold;
    bool mVerbose;
    uint32 mNumBCs = 0;
    std::vector<uint32> mReshuffleMap;   // original -> shuffled
};

class FemPhysicsMatrixFree : public FemPhysicsBase {
public:
    enum NodalForcesType { NFT_DIRECT, NFT_FINITE_VOLUME };
    enum Preconditioner {
        PC_NONE = SFEM_PC_NONE,
        PC_JACOBI = SFEM_PC_JACOBI,
        PC_LEARNED = SFEM_PC_LEARNED,
        PC_MULTIGRID = SFEM_PC_MULTIGRID,                  // needs AddMultigridLevel
        PC_MULTIGRID_LEARNED = SFEM_PC_MULTIGRID_LEARNED   // + SetPreconditionerBlocks
    };

    // One coarser nested mesh of the multigrid hierarchy (no reference
    // counterpart; sfem_mg_add_level): coarse rest positions, tets and fixed
    // flags, the fine node each coarse node injects from, the prolongation
    // (CSR, one row per node of the level above, columns coarse nodes) and,
    // when the fine mesh has per-element moduli, the coarse cells' moduli.
    struct MultigridLevel {
        std::vector<Vector3R> mNodes;
        std::vector<int32_t> mTets;      // 4 per element
        std::vector<uint8_t> mFixed;
        std::vector<int32_t> mInject;
        std::vector<int64_t> mProlongPtr;
        std::vector<int32_t> mProlongCol;
        std::vector<double> mProlongVal;
        std::vector<real> mYoung, mPoisson;
    };

    // FemPhysicsMatrixFree::Config (FemPhysicsMatrixFree.h:52-58) plus the
    // solver controls of this implementation (no reference counterpart):
    //   mTolerance > 0 turns NST_NEWTON_CG into the converged Newton-PCG
    //   (||r|| <= mTolerance ||r_ext||); 0 keeps the reference's fixed-step
    //   SolveNewtonCG.  NST_NEWTON / NST_NEWTON_LS (assembled direct solves in
    //   the reference) run the converged matrix-free Newton-PCG without / with
    //   line search.  mPredictor (converged modes, quasi-static steps) starts
    //   each load step from an extrapolation of the last states: 2 quadratic
    //   through the last three (the secant one if it fails), 1 the secant of
    //   the last two, 0 none (the last state, as Step does).
    struct Config {
        NodalForcesType mNodalForcesComputation = NFT_DIRECT;
        NonlinearSolverType mSolver = NST_NEWTON;
        float mDescentRate = 1e-5f;
        bool mOptimizer = false;
        real mTolerance = 0;
        real mCGTolerance = 0.1;
        Preconditioner mPreconditioner = PC_JACOBI;
        int mCheckEvery = 16;
        int mPredictor = 2;
    };

    FemPhysicsMatrixFree(const std::vector<Tet>& tetrahedra, const std::vector<Node>& allNodes, const FemConfig& config);
    ~FemPhysicsMatrixFree() override;
    FemPhysicsMatrixFree(const FemPhysicsMatrixFree&) = delete;
    FemPhysicsMatrixFree& operator=(const FemPhysicsMatrixFree&) = delete;

    void Step(real dt) override;

    Vector3R GetDeformedPosition(uint32 idx, bool shuffle = true) const override;
    Vector3R GetInitialPosition(uint32 idx) const;
    void SetDeformedPosition(uint32 idx, Vector3R val, bool shuffle = true);
    uint32 GetNumNodes() const override { return (uint32)mNodes.size(); }
    uint32 GetNumLocalNodes() const { return 4; }
    uint32 GetNumFreeNodes() const override { return GetNumNodes() - mNumBCs; }
    uint32 GetNumElements() const override { return (uint32)mTets.size(); }
    bool IsNodeFixed(uint32 i) const override { return mFixed[i] != 0; }
    Matrix3R GetBarycentricJacobianMatrix(uint32 e) const;   // Dm^-T
    real GetElementInitialVolume(uint32 e) const { return mVol[e]; }
    uint32 GetGlobalIndex(uint32 e, uint32 l) const { return mReshuffleMap[mTets[e].idx[l]]; }
    uint32 GetGlobalOriginalIndex(uint32 e, uint32 l) const { return mTets[e].idx[l]; }
    real GetTotalVolume() const override;
    bool CheckForInversion(int verbose = 0) const;
    // first Piola-Kirchhoff stress per element (ElasticEnergy::ComputeElementStress,
    // ElasticEnergy.cpp:114-257), original element order, computed on the device
    void GetElementStresses(std::vector<Matrix3R>& P) const;

    Vector3R GetVelocity(uint32 idx) const;   // original index (FemPhysicsMatrixFree.h:73)
    // pos - pos0 of free node i in shuffled order (FemPhysicsMatrixFree.h:100)
    Vector3R GetTotalDisplacement(uint32 i) const;

    // Energies, FP64 on the device: ComputeElasticEnergy (FemPhysicsBase.cpp:203-212),
    // ComputeSpringEnergy (FemPhysicsMatrixFree.cpp:1113-1121) and ComputeEnergy
    // (:882-916) with the reference's terms as written there: the elastic and
    // spring energies, minus m frac g_y y over free nodes (the y components
    // whatever the gravity axis), minus 0.5 m |x - x0|^2 / h^2 after implicit
    // steps, plus |det[p1, p2, p3]| (float) / 6 times the pressure for every
    // third triangle of the pressure surface (its loop steps by 3 over the
    // triangle count); x0 here is the rest position (the reference's pos0
    // becomes a step's start position in implicit dynamics)
    real ComputeElasticEnergy();
    real ComputeSpringEnergy();
    real ComputeEnergy();

    // GetForceParamGrads (FemPhysicsBase.cpp:162-187): d f / d mu, d f / d lambda
    // and d f / d rho on the free nodes in shuffled order, 3 per node (the
    // reference's EigenVectors); every material has its derivatives here
    // (the reference fills them for Neo-Hookean only)
    void GetForceParamGrads(std::vector<real>& gradMu, std::vector<real>& gradLambda, std::vector<real>& gradRho);

    // The matrix the matrix-free operator applies (K, + M / h^2 after implicit
    // steps) on the free dofs in shuffled order, dense row-major (3 n_free)^2,
    // recovered from a few operator applies on distance-2-colored probe vectors;
    // the reference's GetHessian returns the system matrix of its last
    // assembled Newton solve (FemPhysicsBase.h:126, FemPhysicsMatrixFree.cpp:778)
    void GetHessian(std::vector<real>& K) const;

    // Cables (FemPhysicsBase.h:130-135): the cable's nodes must lie in elements
    // (free cable nodes, elem < 0, are refused); their spring forces enter the
    // residual and the explicit substep, not the operator (as in the reference)
    void AddCable(const Cable& cable);
    uint32 GetNumCables() const { return (uint32)mCables.size(); }
    const Cable& GetCable(uint32 i) const { return mCables[i]; }
    Vector3R GetCablePosition(uint32 cable, uint32 node) const;
    void SetCableActuation(uint32 cable, real actuation);

    // Krylov solver helpers (FemPhysicsMatrixFree.h:85-89)
    void MatrixVectorMultiply(const std::vector<Vector3R>& in, std::vector<Vector3R>& out) const;
    void ComputeGradients(std::vector<Vector3R>& r);
    void UpdatePosAndComputeGradients(const std::vector<Vector3R>& pos, std::vector<Vector3R>& r);
    real DotProduct(const std::vector<Vector3R>& a, const std::vector<Vector3R>& b) const;

    void SetBoundaryConditionsSurface(const std::vector<uint32>& triangleList, real pressure) override;
    void SetExternalForces(const std::vector<Vector3R>& forces) override;
    void SetLameParams(real mu, real lambda);
    void SetPerElementMaterial(const std::vector<real>& young, const std::vector<real>& poisson);
    void SetPreconditionerBlocks(const std::vector<float>& minv9);   // original node order, nn*9
    // appends the next coarser level (finest first); SetLameParams keeps the
    // levels (with the new uniform moduli), SetPerElementMaterial drops them:
    // add them again with the coarse cells' moduli.
    void AddMultigridLevel(const MultigridLevel& level);
    void ClearMultigrid();
    uint32 GetNumMultigridLevels() const { return 1 + (uint32)mMultigrid.size(); }

    const sfem_solve_stats& GetLastSolveStats() const { return mStats; }
    const Config& GetConfig() const { return mConfig; }
    Config& GetConfig() { return mConfig; }
    const std::vector<Node>& GetNodes() const;

private:
    bool Solve();
    void SolveNewtonCG();
    void Create(const double* youngE, const double* poissonE);
    void PushPositions() const;
    void PullPositions() const;
    sfem_solve_params Params() const;
    void Check(int rc, const char* what) const;
    void PushMultigridLevel(size_t k);
    void PushCable(const Cable& cable);

    // a driver's access to the private solve entry points and the load
    // fraction, as the reference grants (FemPhysicsMatrixFree.h:152-153)
    friend class FemSystem;
    friend class ElasticProblem;

    sfem_ctx* mCtx = nullptr;
    Config mConfig;
    std::vector<Tet> mTets;
    mutable std::vector<Node> mNodes;   // original order; pos mirrors the device state
    std::vector<uint8_t> mFixed;
    std::vector<real> mVol, mDmInv;
    mutable bool mHostDirty = false;    // host positions newer than the device
    mutable bool mDeviceDirty = false;  // device positions newer than the host
    std::vector<MultigridLevel> mMultigrid;
    std::vector<Cable> mCables;
    std::vector<double> mMass;              // lumped masses, original order (:126-146)
    std::vector<uint32> mShuffledToOriginal;
    std::vector<uint32> mPressureTris;      // SetBoundaryConditionsSurface, original ids
    real mAppliedPressure = 0;
    std::vector<double> mExternalForces;    // SetExternalForces, nn * 3 (empty: none)
    real mTimeStep = 0;                     // the implicit substep of the last Step (mTimeStep)
    sfem_solve_stats mStats{};
};

}  // namespace FEM_SYSTEM
