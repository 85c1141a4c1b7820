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

namespace FEM_SYSTEM {

typedef double real;
typedef uint32_t uint32;

template <typename Real> struct Vector3T {
    Real v[3];
    Vector3T() : v{0, 0, 0} {}
    Vector3T(Real x, Real y, Real z) : v{x, y, z} {}
    Real& operator[](int i) { return v[i]; }
    Real operator[](int i) const { return v[i]; }
    Real X() const { return v[0]; }
    Real Y() const { return v[1]; }
    Real Z() const { return v[2]; }
    void SetZero() { v[0] = v[1] = v[2] = 0; }
    Real Dot(const Vector3T& o) const { return v[0] * o.v[0] + v[1] * o.v[1] + v[2] * o.v[2]; }
    Real Length() const { return std::sqrt(Dot(*this)); }
    Vector3T& operator+=(const Vector3T& o) { v[0] += o.v[0]; v[1] += o.v[1]; v[2] += o.v[2]; return *this; }
    Vector3T& operator-=(const Vector3T& o) { v[0] -= o.v[0]; v[1] -= o.v[1]; v[2] -= o.v[2]; return *this; }
};
template <typename R> Vector3T<R> operator+(Vector3T<R> a, const Vector3T<R>& b) { return a += b; }
template <typename R> Vector3T<R> operator-(Vector3T<R> a, const Vector3T<R>& b) { return a -= b; }
template <typename R> Vector3T<R> operator*(double s, const Vector3T<R>& a) { return Vector3T<R>(s * a[0], s * a[1], s * a[2]); }
template <typename R> Vector3T<R> operator*(float s, const Vector3T<R>& a) { return Vector3T<R>(s * a[0], s * a[1], s * a[2]); }
template <typename R> R operator*(const Vector3T<R>& a, const Vector3T<R>& b) { return a.Dot(b); }   // vector3.h:203-207

// 3x3 with the reference's conventions: row storage, default = identity
struct Matrix3R {
    real m[3][3];
    Matrix3R() : m{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}} {}
    real operator()(int i, int j) const { return m[i][j]; }
    Vector3T<real> operator()(int j) const { return Vector3T<real>(m[0][j], m[1][j], m[2][j]); }
};

typedef Vector3T<real> Vector3R;

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
    //   each load step from the secant extrapolation of the last two states.
    struct Config {
        NodalForcesType mNodalForcesComputation = NFT_DIRECT;
        NonlinearSolverType mSolver = NST_NEWTON;
        float mDescentRate = 1e-5f;
        bool mOptimizer = false;
        real mTolerance = 0;
        real mCGTolerance = 0.1;
        Preconditioner mPreconditioner = PC_JACOBI;
        int mCheckEvery = 16;
        bool mPredictor = true;
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

    sfem_ctx* mCtx = nullptr;
    Config mConfig;
    std::vector<Tet> mTets;
    mutable std::vector<Node> mNodes;   // original order; pos mirrors the device state
    std::vector<uint8_t> mFixed;
    std::vector<real> mVol, mDmInv;
    mutable bool mHostDirty = false;    // host positions newer than the device
    mutable bool mDeviceDirty = false;  // device positions newer than the host
    std::vector<MultigridLevel> mMultigrid;
    sfem_solve_stats mStats{};
};

}  // namespace FEM_SYSTEM
