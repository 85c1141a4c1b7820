// FemPhysicsMatrixFree.cpp -- host implementation of the reference-shaped C++
// API on top of the C-ABI (include/sfem_capi.h).
//
// Error behaviour: the reference reports through Printf and never throws;
// here a failing device call (out of memory, HIP error) throws
// std::runtime_error with the library's message, because continuing after a
// device failure would return garbage.  A solve that misses its tolerance is
// not an error (as in the reference, which has no residual test): it is
// recorded in GetLastSolveStats() and printed when mVerbose is set.
#include "FemPhysicsMatrixFree.h"

#include <algorithm>
#include <cstdio>
#include <stdexcept>

namespace FEM_SYSTEM {

// FemPhysicsBase.cpp:12-57: Lame parameters and gravity axis
FemPhysicsBase::FemPhysicsBase(const FemConfig& config)
    : mYoungsModulus(config.mYoungsModulus), mPoissonRatio(config.mPoissonRatio), mDensity(config.mDensity),
      mSimType(config.mSimType), mMaterial(config.mMaterial), mOuterIterations(config.mOuterIterations),
      mInnerIterations(config.mInnerIterations), mForceStep(config.mForceApplicationStep),
      mAbsNewtonResidualThreshold(config.mAbsNewtonRsidualThreshold), mVerbose(config.mVerbose),
      mNumSteps(config.mNumSubsteps)
{
    mGravity = config.mZUpAxis ? Vector3R(0, 0, config.mGravity) : Vector3R(0, config.mGravity, 0);
    mLameLambda = real(mYoungsModulus * mPoissonRatio / (1.0 + mPoissonRatio) / (1.0 - 2.0 * mPoissonRatio));
    mShearModulus = 0.5 * mYoungsModulus / (1.0 + mPoissonRatio);
}

FemPhysicsMatrixFree::FemPhysicsMatrixFree(const std::vector<Tet>& tetrahedra, const std::vector<Node>& allNodes,
                                           const FemConfig& config)
    : FemPhysicsBase(config), mTets(tetrahedra), mNodes(allNodes)
{
    if (config.mCustomConfig) mConfig = *static_cast<const Config*>(config.mCustomConfig);
    if (mMaterial > MMT_NEO_HOOKEAN && mMaterial != MMT_COROTATIONAL_POLAR)
        throw std::runtime_error("FemPhysicsMatrixFree: only LINEAR, STVK, COROTATIONAL, NEO_HOOKEAN and "
                                 "COROTATIONAL_POLAR are on the matrix-free Newton-CG path");
    const size_t nn = mNodes.size();
    // ReshuffleFixedNodes (FemPhysicsMatrixFree.cpp:229-263): fixed nodes first
    mFixed.resize(nn);
    mReshuffleMap.resize(nn);
    uint32 k = 0;
    for (size_t i = 0; i < nn; i++) {
        mFixed[i] = mNodes[i].invMass == 0 ? 1 : 0;
        if (mFixed[i]) mReshuffleMap[i] = k++;
    }
    mNumBCs = k;
    for (size_t i = 0; i < nn; i++)
        if (!mFixed[i]) mReshuffleMap[i] = k++;
    mShuffledToOriginal.resize(nn);
    for (size_t i = 0; i < nn; i++) mShuffledToOriginal[mReshuffleMap[i]] = (uint32)i;
    // rest geometry kept on the host for the geometric getters (:126-139)
    const size_t ne = mTets.size();
    mVol.resize(ne);
    mDmInv.resize(9 * ne);
    for (size_t e = 0; e < ne; e++) {
        const Vector3R& x0 = mNodes[mTets[e].idx[0]].pos0;
        real D[9];
        for (int c = 1; c < 4; c++) {
            const Vector3R& xc = mNodes[mTets[e].idx[c]].pos0;
            for (int r = 0; r < 3; r++) D[3 * r + c - 1] = xc[r] - x0[r];
        }
        real v[9] = {D[4] * D[8] - D[5] * D[7], D[2] * D[7] - D[1] * D[8], D[1] * D[5] - D[2] * D[4],
                     -D[3] * D[8] + D[6] * D[5], D[0] * D[8] - D[2] * D[6], D[2] * D[3] - D[0] * D[5],
                     D[3] * D[7] - D[6] * D[4], D[1] * D[6] - D[0] * D[7], D[0] * D[4] - D[1] * D[3]};
        const real det = D[0] * v[0] + D[1] * v[3] + D[2] * v[6];
        for (int q = 0; q < 9; q++) mDmInv[9 * e + q] = v[q] / det;
        mVol[e] = det / 6.0;
    }
    // lumped masses rho V / 4 per corner (:141-146)
    mMass.assign(nn, 0.0);
    for (size_t e = 0; e < ne; e++)
        for (int c = 0; c < 4; c++) mMass[mTets[e].idx[c]] += 0.25 * mDensity * mVol[e];
    Create(nullptr, nullptr);
}

FemPhysicsMatrixFree::~FemPhysicsMatrixFree()
{
    if (mCtx) sfem_destroy(mCtx);
}

void FemPhysicsMatrixFree::Check(int rc, const char* what) const
{
    if (rc != SFEM_OK) throw std::runtime_error(std::string(what) + ": " + sfem_last_error());
}

void FemPhysicsMatrixFree::Create(const double* youngE, const double* poissonE)
{
    const size_t nn = mNodes.size(), ne = mTets.size();
    std::vector<int32_t> tets(4 * ne);
    for (size_t e = 0; e < ne; e++)
        for (int k = 0; k < 4; k++) tets[4 * e + k] = (int32_t)mTets[e].idx[k];
    std::vector<double> X0(3 * nn);
    for (size_t i = 0; i < nn; i++)
        for (int c = 0; c < 3; c++) X0[3 * i + c] = mNodes[i].pos0[c];
    sfem_mesh m{tets.data(), (int64_t)ne, X0.data(), (int64_t)nn, mFixed.data()};
    sfem_material mat{};
    mat.material = (int32_t)mMaterial;
    mat.young = mYoungsModulus;
    mat.poisson = mPoissonRatio;
    mat.young_e = youngE;
    mat.poisson_e = poissonE;
    mat.density = mDensity;
    for (int c = 0; c < 3; c++) mat.gravity[c] = mGravity[c];
    sfem_ctx* ctx = nullptr;
    Check(sfem_create(&m, &mat, &ctx), "sfem_create");
    if (mCtx) sfem_destroy(mCtx);
    mCtx = ctx;
    // the constructor's node copy may already be deformed (pos != pos0)
    mHostDirty = true;
    PushPositions();
    Check(sfem_set_load_fraction(mCtx, mForceFraction), "sfem_set_load_fraction");
    // loads and cables carry over to the new context
    if (!mPressureTris.empty()) {
        std::vector<int32_t> t(mPressureTris.begin(), mPressureTris.end());
        Check(sfem_set_pressure_surface(mCtx, t.data(), (int64_t)t.size() / 3, mAppliedPressure), "sfem_set_pressure_surface");
    }
    if (!mExternalForces.empty()) Check(sfem_set_external_forces(mCtx, mExternalForces.data()), "sfem_set_external_forces");
    for (const Cable& cb : mCables) PushCable(cb);
    // the multigrid levels follow the new material: uniform moduli re-add them
    // without coarse moduli; per-element moduli need the caller's coarse means
    if (youngE) mMultigrid.clear();
    for (size_t k = 0; k < mMultigrid.size(); k++) {
        mMultigrid[k].mYoung.clear();
        mMultigrid[k].mPoisson.clear();
        PushMultigridLevel(k);
    }
}

// FemPhysicsBase::SetLameParams (FemPhysicsBase.h:98-106)
void FemPhysicsMatrixFree::SetLameParams(real mu, real lambda)
{
    mLameLambda = lambda;
    mShearModulus = mu;
    mYoungsModulus = mu * (3 * lambda + 2 * mu) / (mu + lambda);
    mPoissonRatio = 0.5 * lambda / (mu + lambda);
    PullPositions();
    Create(nullptr, nullptr);
}

void FemPhysicsMatrixFree::SetPerElementMaterial(const std::vector<real>& young, const std::vector<real>& poisson)
{
    if (young.size() != mTets.size() || poisson.size() != mTets.size())
        throw std::runtime_error("SetPerElementMaterial: one value per element expected");
    PullPositions();
    Create(young.data(), poisson.data());
}

void FemPhysicsMatrixFree::PushPositions() const
{
    if (!mHostDirty) return;
    std::vector<double> x(3 * mNodes.size());
    for (size_t i = 0; i < mNodes.size(); i++)
        for (int c = 0; c < 3; c++) x[3 * i + c] = mNodes[i].pos[c];
    Check(sfem_set_positions(mCtx, x.data()), "sfem_set_positions");
    mHostDirty = false;
}

void FemPhysicsMatrixFree::PullPositions() const
{
    if (!mDeviceDirty) return;
    std::vector<double> x(3 * mNodes.size());
    Check(sfem_get_positions(mCtx, x.data()), "sfem_get_positions");
    for (size_t i = 0; i < mNodes.size(); i++) mNodes[i].pos = Vector3R(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    mDeviceDirty = false;
}

const std::vector<Node>& FemPhysicsMatrixFree::GetNodes() const
{
    PullPositions();
    return mNodes;
}

// GetDeformedPosition (FemPhysicsMatrixFree.h:70): with shuffle the index is an
// original node id; without, a shuffled one.
Vector3R FemPhysicsMatrixFree::GetDeformedPosition(uint32 idx, bool shuffle) const
{
    PullPositions();
    if (shuffle) return mNodes[idx].pos;
    // shuffled index -> original
    for (size_t i = 0; i < mReshuffleMap.size(); i++)
        if (mReshuffleMap[i] == idx) return mNodes[i].pos;
    return Vector3R();
}

Vector3R FemPhysicsMatrixFree::GetInitialPosition(uint32 idx) const { return mNodes[idx].pos0; }

void FemPhysicsMatrixFree::GetElementStresses(std::vector<Matrix3R>& P) const
{
    PushPositions();
    std::vector<double> buf(9 * mTets.size());
    Check(sfem_element_stress(mCtx, buf.data()), "sfem_element_stress");
    P.resize(mTets.size());
    for (size_t e = 0; e < mTets.size(); e++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) P[e].m[i][j] = buf[9 * e + 3 * i + j];
}

void FemPhysicsMatrixFree::SetDeformedPosition(uint32 idx, Vector3R val, bool shuffle)
{
    PullPositions();
    if (!shuffle) {
        for (size_t i = 0; i < mReshuffleMap.size(); i++)
            if (mReshuffleMap[i] == idx) { idx = (uint32)i; break; }
    }
    mNodes[idx].pos = val;
    mHostDirty = true;
}

Matrix3R FemPhysicsMatrixFree::GetBarycentricJacobianMatrix(uint32 e) const
{
    Matrix3R M;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M.m[i][j] = mDmInv[9 * e + 3 * j + i];
    return M;
}

real FemPhysicsMatrixFree::GetTotalVolume() const   // :498-521, current volume
{
    PullPositions();
    real tot = 0;
    for (const Tet& t : mTets) {
        const Vector3R& x0 = mNodes[t.idx[0]].pos;
        Vector3R a = mNodes[t.idx[1]].pos - x0, b = mNodes[t.idx[2]].pos - x0, c = mNodes[t.idx[3]].pos - x0;
        tot += (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0])) / 6.0;
    }
    return tot;
}

// FemPhysicsBase::CheckForInversion (FemPhysicsBase.cpp:93-117): det F <= 0
bool FemPhysicsMatrixFree::CheckForInversion(int verbose) const
{
    PullPositions();
    bool inv = false;
    for (size_t e = 0; e < mTets.size(); e++) {
        const Tet& t = mTets[e];
        const Vector3R& x0 = mNodes[t.idx[0]].pos;
        Vector3R a = mNodes[t.idx[1]].pos - x0, b = mNodes[t.idx[2]].pos - x0, c = mNodes[t.idx[3]].pos - x0;
        const real ds = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
        if (ds / (6.0 * mVol[e]) <= 0) {
            inv = true;
            if (verbose > 1) std::printf("Inverted element %zu\n", e);
        }
    }
    if (inv && verbose) std::printf("Inversion detected\n");
    return inv;
}

void FemPhysicsMatrixFree::MatrixVectorMultiply(const std::vector<Vector3R>& in, std::vector<Vector3R>& out) const
{
    PushPositions();
    const size_t nn = mNodes.size();
    std::vector<double> d(3 * nn, 0.0), q(3 * nn);
    for (size_t i = 0; i < nn; i++)
        if (!mFixed[i])
            for (int c = 0; c < 3; c++) d[3 * i + c] = in[mReshuffleMap[i] - mNumBCs][c];
    Check(sfem_matvec(mCtx, d.data(), q.data()), "sfem_matvec");
    out.resize(GetNumFreeNodes());
    for (size_t i = 0; i < nn; i++)
        if (!mFixed[i]) out[mReshuffleMap[i] - mNumBCs] = Vector3R(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
}

void FemPhysicsMatrixFree::ComputeGradients(std::vector<Vector3R>& r)
{
    PushPositions();
    Check(sfem_set_load_fraction(mCtx, mForceFraction), "sfem_set_load_fraction");
    const size_t nn = mNodes.size();
    std::vector<double> g(3 * nn);
    Check(sfem_gradients(mCtx, g.data()), "sfem_gradients");
    r.resize(GetNumFreeNodes());
    for (size_t i = 0; i < nn; i++)
        if (!mFixed[i]) r[mReshuffleMap[i] - mNumBCs] = Vector3R(g[3 * i], g[3 * i + 1], g[3 * i + 2]);
}

void FemPhysicsMatrixFree::UpdatePosAndComputeGradients(const std::vector<Vector3R>& pos, std::vector<Vector3R>& r)
{
    PullPositions();
    for (size_t i = 0; i < mNodes.size(); i++)
        if (!mFixed[i]) mNodes[i].pos = pos[mReshuffleMap[i] - mNumBCs];
    mHostDirty = true;
    ComputeGradients(r);
}

real FemPhysicsMatrixFree::DotProduct(const std::vector<Vector3R>& a, const std::vector<Vector3R>& b) const
{
    real s = 0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
}

void FemPhysicsMatrixFree::SetBoundaryConditionsSurface(const std::vector<uint32>& triangleList, real pressure)
{
    std::vector<int32_t> t(triangleList.begin(), triangleList.end());
    Check(sfem_set_pressure_surface(mCtx, t.data(), (int64_t)t.size() / 3, pressure), "sfem_set_pressure_surface");
    mPressureTris = triangleList;
    mAppliedPressure = pressure;
}

void FemPhysicsMatrixFree::SetExternalForces(const std::vector<Vector3R>& forces)
{
    if (forces.empty()) {
        Check(sfem_set_external_forces(mCtx, nullptr), "sfem_set_external_forces");
        mExternalForces.clear();
        return;
    }
    if (forces.size() != mNodes.size()) throw std::runtime_error("SetExternalForces: one force per node expected");
    std::vector<double> f(3 * forces.size());
    for (size_t i = 0; i < forces.size(); i++)
        for (int c = 0; c < 3; c++) f[3 * i + c] = forces[i][c];
    Check(sfem_set_external_forces(mCtx, f.data()), "sfem_set_external_forces");
    mExternalForces = f;
}

Vector3R FemPhysicsMatrixFree::GetVelocity(uint32 idx) const
{
    std::vector<double> v(3 * mNodes.size());
    Check(sfem_get_velocities(mCtx, v.data()), "sfem_get_velocities");
    return Vector3R(v[3 * idx], v[3 * idx + 1], v[3 * idx + 2]);
}

Vector3R FemPhysicsMatrixFree::GetTotalDisplacement(uint32 i) const
{
    PullPositions();
    const Node& n = mNodes[mShuffledToOriginal[i + mNumBCs]];
    return n.pos - n.pos0;
}

real FemPhysicsMatrixFree::ComputeElasticEnergy()
{
    PushPositions();
    double e[3];
    Check(sfem_energy(mCtx, e), "sfem_energy");
    return e[0];
}

real FemPhysicsMatrixFree::ComputeSpringEnergy()
{
    PushPositions();
    double e[3];
    Check(sfem_energy(mCtx, e), "sfem_energy");
    return e[1];
}

real FemPhysicsMatrixFree::ComputeEnergy()
{
    PushPositions();
    Check(sfem_set_load_fraction(mCtx, mForceFraction), "sfem_set_load_fraction");
    double e[3];
    Check(sfem_energy(mCtx, e), "sfem_energy");
    PullPositions();
    real val = e[0];
    const real invHSqr = mTimeStep != 0 ? 0.5f / mTimeStep / mTimeStep : 0;
    for (uint32 s = mNumBCs; s < GetNumNodes(); s++) {
        const Node& n = mNodes[mShuffledToOriginal[s]];
        const real mass = mMass[mShuffledToOriginal[s]];
        val -= mass * mForceFraction * mGravity.y * n.pos.y;
        val -= invHSqr * mass * (n.pos - n.pos0).LengthSquared();
    }
    for (size_t i = 0; i < mPressureTris.size() / 3; i += 3) {
        const Vector3R& p1 = mNodes[mPressureTris[3 * i]].pos;
        const Vector3R& p2 = mNodes[mPressureTris[3 * i + 1]].pos;
        const Vector3R& p3 = mNodes[mPressureTris[3 * i + 2]].pos;
        const real vol = (1.f / 6.f) * triple(p1, p2, p3);
        val += std::fabs(vol) * mAppliedPressure;
    }
    return val + e[1];
}

void FemPhysicsMatrixFree::GetForceParamGrads(std::vector<real>& gradMu, std::vector<real>& gradLambda,
                                              std::vector<real>& gradRho)
{
    PushPositions();
    const size_t nn = mNodes.size();
    std::vector<double> fm(3 * nn), fl(3 * nn), fr(3 * nn);
    Check(sfem_force_param_grads(mCtx, fm.data(), fl.data(), fr.data()), "sfem_force_param_grads");
    const size_t nf = GetNumFreeNodes();
    gradMu.assign(3 * nf, 0);
    gradLambda.assign(3 * nf, 0);
    gradRho.assign(3 * nf, 0);
    for (size_t i = 0; i < nn; i++) {
        if (mFixed[i]) continue;
        const size_t s = mReshuffleMap[i] - mNumBCs;
        for (int c = 0; c < 3; c++) {
            gradMu[3 * s + c] = fm[3 * i + c];
            gradLambda[3 * s + c] = fl[3 * i + c];
            gradRho[3 * s + c] = fr[3 * i + c];
        }
    }
}

void FemPhysicsMatrixFree::GetHessian(std::vector<real>& K) const
{
    PushPositions();
    const size_t nn = mNodes.size(), nf = GetNumFreeNodes(), n = 3 * nf;
    // node adjacency (nodes sharing an element), then a greedy distance-2
    // coloring: two nodes of one color share no neighbour, so one operator
    // apply per color and component recovers every column of that color
    std::vector<std::vector<uint32>> adj(nn);
    for (const Tet& t : mTets)
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) adj[t.idx[a]].push_back(t.idx[b]);
    for (auto& a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    std::vector<int> color(nn, -1);
    int ncolors = 0;
    for (size_t i = 0; i < nn; i++) {
        if (mFixed[i]) continue;
        std::vector<char> used(ncolors + 1, 0);
        for (uint32 j : adj[i])
            for (uint32 k : adj[j])
                if (color[k] >= 0) used[color[k]] = 1;
        int c = 0;
        while (c < ncolors && used[c]) c++;
        color[i] = c;
        ncolors = std::max(ncolors, c + 1);
    }
    K.assign(n * n, 0);
    std::vector<double> d(3 * nn), q(3 * nn);
    for (int c = 0; c < ncolors; c++)
        for (int comp = 0; comp < 3; comp++) {
            std::fill(d.begin(), d.end(), 0.0);
            for (size_t j = 0; j < nn; j++)
                if (color[j] == c) d[3 * j + comp] = 1.0;
            Check(sfem_matvec(mCtx, d.data(), q.data()), "sfem_matvec");
            // column (j, comp) of K: rows of the free nodes adjacent to j
            for (size_t j = 0; j < nn; j++) {
                if (color[j] != c) continue;
                const size_t col = 3 * (mReshuffleMap[j] - mNumBCs) + comp;
                for (uint32 i : adj[j]) {
                    if (mFixed[i]) continue;
                    const size_t row = 3 * (mReshuffleMap[i] - mNumBCs);
                    for (int r = 0; r < 3; r++) K[(row + r) * n + col] = q[3 * (size_t)i + r];
                }
            }
        }
}

void FemPhysicsMatrixFree::PushCable(const Cable& cable)
{
    const size_t n = cable.mCableNodes.size();
    std::vector<int32_t> elem(n);
    std::vector<double> bary(3 * n);
    for (size_t i = 0; i < n; i++) {
        elem[i] = cable.mCableNodes[i].elem;
        for (int c = 0; c < 3; c++) bary[3 * i + c] = cable.mCableNodes[i].bary[c];
    }
    int32_t id = 0;
    Check(sfem_add_cable(mCtx, elem.data(), bary.data(), (int32_t)n, cable.mCableRestLength, cable.mCableStiffness,
                         cable.mCableDamping, &id),
          "sfem_add_cable");
    if (cable.mActuation != 1.0) Check(sfem_set_cable_actuation(mCtx, id, cable.mActuation), "sfem_set_cable_actuation");
}

void FemPhysicsMatrixFree::AddCable(const Cable& cable)
{
    PushCable(cable);
    mCables.push_back(cable);
}

Vector3R FemPhysicsMatrixFree::GetCablePosition(uint32 cable, uint32 node) const
{
    PushPositions();
    std::vector<double> pos(3 * mCables.at(cable).mCableNodes.size());
    Check(sfem_get_cable_positions(mCtx, (int32_t)cable, pos.data()), "sfem_get_cable_positions");
    return Vector3R(pos[3 * node], pos[3 * node + 1], pos[3 * node + 2]);
}

void FemPhysicsMatrixFree::SetCableActuation(uint32 cable, real actuation)
{
    if (cable >= mCables.size()) return;   // FemPhysicsBase.h:134
    mCables[cable].mActuation = actuation;
    Check(sfem_set_cable_actuation(mCtx, (int32_t)cable, actuation), "sfem_set_cable_actuation");
}

void FemPhysicsMatrixFree::SetPreconditionerBlocks(const std::vector<float>& minv9)
{
    if (minv9.empty()) {
        Check(sfem_set_precond_blocks(mCtx, nullptr), "sfem_set_precond_blocks");
        return;
    }
    if (minv9.size() != 9 * mNodes.size()) throw std::runtime_error("SetPreconditionerBlocks: nn*9 values expected");
    Check(sfem_set_precond_blocks(mCtx, minv9.data()), "sfem_set_precond_blocks");
}

void FemPhysicsMatrixFree::PushMultigridLevel(size_t k)
{
    const MultigridLevel& l = mMultigrid[k];
    const size_t nc = l.mNodes.size(), nec = l.mTets.size() / 4;
    if (l.mTets.size() != 4 * nec || l.mFixed.size() != nc || l.mInject.size() != nc)
        throw std::runtime_error("AddMultigridLevel: 4 node ids per tet and one fixed flag and injection index per node expected");
    const size_t nf = k == 0 ? mNodes.size() : mMultigrid[k - 1].mNodes.size();
    if (l.mProlongPtr.size() != nf + 1 || l.mProlongCol.size() != l.mProlongVal.size() ||
        (int64_t)l.mProlongCol.size() != l.mProlongPtr.back())
        throw std::runtime_error("AddMultigridLevel: prolongation rows must cover the level above");
    const bool moduli = !l.mYoung.empty() || !l.mPoisson.empty();
    if (moduli && (l.mYoung.size() != nec || l.mPoisson.size() != nec))
        throw std::runtime_error("AddMultigridLevel: one Young's modulus and Poisson ratio per coarse element expected");
    std::vector<double> X(3 * nc);
    for (size_t i = 0; i < nc; i++)
        for (int c = 0; c < 3; c++) X[3 * i + c] = l.mNodes[i][c];
    sfem_mesh m{l.mTets.data(), (int64_t)nec, X.data(), (int64_t)nc, l.mFixed.data()};
    Check(sfem_mg_add_level(mCtx, &m, l.mInject.data(), l.mProlongPtr.data(), l.mProlongCol.data(),
                            l.mProlongVal.data(), moduli ? l.mYoung.data() : nullptr,
                            moduli ? l.mPoisson.data() : nullptr),
          "sfem_mg_add_level");
}

void FemPhysicsMatrixFree::AddMultigridLevel(const MultigridLevel& level)
{
    mMultigrid.push_back(level);
    try {
        PushMultigridLevel(mMultigrid.size() - 1);
    } catch (...) {
        mMultigrid.pop_back();
        throw;
    }
}

void FemPhysicsMatrixFree::ClearMultigrid()
{
    Check(sfem_mg_clear(mCtx), "sfem_mg_clear");
    mMultigrid.clear();
}

sfem_solve_params FemPhysicsMatrixFree::Params() const
{
    sfem_solve_params p{};
    const bool reference = mConfig.mSolver == NST_NEWTON_CG && mConfig.mTolerance <= 0;
    p.mode = reference ? SFEM_SOLVE_REFERENCE : SFEM_SOLVE_CONVERGED;
    p.precond = reference ? SFEM_PC_NONE : (int32_t)mConfig.mPreconditioner;
    p.outer_iterations = (int32_t)mOuterIterations;
    p.inner_iterations = (int32_t)mInnerIterations;
    // NST_NEWTON / NST_NEWTON_LS keep the reference's absolute residual
    // threshold (FemConfig::mAbsNewtonRsidualThreshold); mTolerance adds a
    // relative one; with neither, 1e-8 relative
    const bool direct_kind = mConfig.mSolver == NST_NEWTON || mConfig.mSolver == NST_NEWTON_LS;
    p.abs_tol = direct_kind ? mAbsNewtonResidualThreshold : 0.0;
    p.tol = mConfig.mTolerance > 0 ? mConfig.mTolerance : (p.abs_tol > 0 ? 0.0 : 1e-8);
    p.cg_rtol = mConfig.mCGTolerance;
    p.line_search = mConfig.mSolver != NST_NEWTON;
    p.check_every = mConfig.mCheckEvery;
    p.sim_type = mSimType;
    p.force_step = mForceStep;
    p.predictor = reference ? 0 : mConfig.mPredictor;
    return p;
}

// Step (FemPhysicsMatrixFree.cpp:297-342), static and quasi-static branches
void FemPhysicsMatrixFree::Step(real dt)
{
    if (mSimType == ST_IMPLICIT) {
        // mNumSteps backward-Euler solves on the device (Step :299-314, Solve :589-597).
        // Unlike the reference, the nodes' pos0 keep the rest positions (the
        // reference overwrites them with the step's start positions).
        PushPositions();
        sfem_solve_params p = Params();
        p.time_step = dt;
        p.substeps = std::max(1, mNumSteps);
        mForceFraction = 1;
        mTimeStep = dt / p.substeps;
        const int rc = sfem_step(mCtx, &p, &mStats);
        if (rc != SFEM_OK && rc != SFEM_ENOTCONV) Check(rc, "sfem_step");
        mDeviceDirty = true;
    } else if (mSimType == ST_EXPLICIT) {
        // mNumSteps substeps of symplectic Euler (SubStep, FemPhysicsMatrixFree.cpp:371-387), on the device
        PushPositions();
        sfem_solve_params p = Params();
        p.time_step = dt;
        p.substeps = std::max(1, mNumSteps);
        mForceFraction = 1;
        Check(sfem_step(mCtx, &p, &mStats), "sfem_step");
        mDeviceDirty = true;
    } else if (mSimType == ST_STATIC) {
        if (mForceFraction == 0) {
            mForceFraction = 1;
            Solve();
        }
    } else if (mSimType == ST_QUASI_STATIC) {
        if (mForceFraction < 1) {
            mForceFraction = std::min(real(1), mForceFraction + mForceStep);
            Solve();
        }
    }
}

// Solve (:570-602): every supported solver type runs on the device
bool FemPhysicsMatrixFree::Solve()
{
    switch (mConfig.mSolver) {
    case NST_NEWTON:
    case NST_NEWTON_LS:
    case NST_NEWTON_CG:
        SolveNewtonCG();
        return true;
    default:
        throw std::runtime_error("FemPhysicsMatrixFree: only the Newton solvers run on the matrix-free device path");
    }
}

void FemPhysicsMatrixFree::SolveNewtonCG()
{
    PushPositions();
    Check(sfem_set_load_fraction(mCtx, mForceFraction), "sfem_set_load_fraction");
    sfem_solve_params p = Params();
    const int rc = sfem_solve(mCtx, &p, &mStats);
    if (rc != SFEM_OK && rc != SFEM_ENOTCONV) Check(rc, "sfem_solve");
    mDeviceDirty = true;
    if (mVerbose && rc == SFEM_ENOTCONV)
        std::printf("Newton-CG: residual %g not below %g * %g after %d steps\n", mStats.res, p.tol, mStats.ext_norm,
                    mStats.newton_iters);
}

}  // namespace FEM_SYSTEM
