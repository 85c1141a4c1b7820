// ref_driver.cpp -- a thin C entry layer over the COMPILED REFERENCE
// (FemPhysicsMatrixFree and friends, built from the sources under the
// reference checkout by oracle/Makefile into oracle/_ref/).
//
// TEST INFRASTRUCTURE ONLY: used to validate the FP64 restatement
// (sfem_oracle.c) and to write golden vectors; never linked into the product.
//
// All node arrays are in the caller's ORIGINAL node order; the driver maps
// through mReshuffleMap (FemPhysicsBase.h:162) both ways.
#include "FemPhysicsMatrixFree.h"
#include <cstdint>
#include <cstring>

using namespace FEM_SYSTEM;

// The reference grants `friend class FemSystem` (FemPhysicsMatrixFree.h:152) for
// drivers; this one uses it to set the load fraction and to call SolveNewtonCG.
namespace FEM_SYSTEM {
class FemSystem {
public:
    static void SetFraction(FemPhysicsMatrixFree* p, real f) { p->mForceFraction = f; }
    static void NewtonCG(FemPhysicsMatrixFree* p) { p->SolveNewtonCG(); }
};
}

namespace {
struct RefCtx {
    FemPhysicsMatrixFree* phys = nullptr;
    FemPhysicsMatrixFree::Config cfg;
    std::vector<uint32> perm;   // original -> shuffled
    uint32 nbc = 0;
};
}

extern "C" {

// Builds FemPhysicsMatrixFree(tets, nodes, FemConfig) (FemPhysicsMatrixFree.cpp:71)
// with solver NST_NEWTON_CG selected through mCustomConfig (:104-108).
void* ref_create(const int32_t* tets, int64_t ne, const double* X, int64_t nn, const uint8_t* fixed,
                 int material, double young, double poisson, double density, double gravity,
                 int zup, int simtype, double force_step, int outer, int inner)
{
    std::vector<Tet> t(ne);
    for (int64_t e = 0; e < ne; e++)
        for (int k = 0; k < 4; k++) t[e].idx[k] = (uint32)tets[4 * e + k];
    std::vector<Node> nodes(nn);
    for (int64_t i = 0; i < nn; i++) {
        nodes[i].pos = Vector3R(X[3 * i], X[3 * i + 1], X[3 * i + 2]);
        nodes[i].pos0 = nodes[i].pos;
        nodes[i].invMass = fixed[i] ? 0 : 1;
    }
    RefCtx* c = new RefCtx;
    c->cfg.mSolver = NST_NEWTON_CG;
    FemConfig fc;
    fc.mYoungsModulus = young;
    fc.mPoissonRatio = poisson;
    fc.mDensity = density;
    fc.mGravity = gravity;
    fc.mZUpAxis = zup != 0;
    fc.mMaterial = (MaterialModelType)material;
    fc.mSimType = (SimulationType)simtype;
    fc.mForceApplicationStep = force_step;
    fc.mOuterIterations = outer;
    fc.mInnerIterations = inner;
    fc.mVerbose = false;
    fc.mCustomConfig = &c->cfg;
    c->phys = new FemPhysicsMatrixFree(t, nodes, fc);
    // recover the permutation: the reference renumbers fixed nodes first
    c->perm.resize(nn);
    uint32 k = 0;
    for (int64_t i = 0; i < nn; i++) if (fixed[i]) c->perm[i] = k++;
    c->nbc = k;
    for (int64_t i = 0; i < nn; i++) if (!fixed[i]) c->perm[i] = k++;
    return c;
}

void ref_destroy(void* h)
{
    RefCtx* c = (RefCtx*)h;
    delete c->phys;
    delete c;
}

void ref_set_positions(void* h, const double* x)
{
    RefCtx* c = (RefCtx*)h;
    for (size_t i = 0; i < c->perm.size(); i++)
        c->phys->SetDeformedPosition((uint32)i, Vector3R(x[3 * i], x[3 * i + 1], x[3 * i + 2]), true);
}

void ref_get_positions(void* h, double* x)
{
    RefCtx* c = (RefCtx*)h;
    for (size_t i = 0; i < c->perm.size(); i++) {
        Vector3R p = c->phys->GetDeformedPosition((uint32)i, true);
        x[3 * i] = p.X(); x[3 * i + 1] = p.Y(); x[3 * i + 2] = p.Z();
    }
}

void ref_set_pressure_surface(void* h, const int32_t* tris, int64_t ntri, double pressure)
{
    RefCtx* c = (RefCtx*)h;
    std::vector<uint32> tl(3 * ntri);
    for (int64_t i = 0; i < 3 * ntri; i++) tl[i] = (uint32)tris[i];
    c->phys->SetBoundaryConditionsSurface(tl, pressure);
}

// free-row vectors (shuffled order) <-> full original-order arrays
static void to_free(const RefCtx* c, const double* full, std::vector<Vector3R>& v)
{
    v.assign(c->perm.size() - c->nbc, Vector3R());
    for (size_t i = 0; i < c->perm.size(); i++)
        if (c->perm[i] >= c->nbc)
            v[c->perm[i] - c->nbc] = Vector3R(full[3 * i], full[3 * i + 1], full[3 * i + 2]);
}
static void from_free(const RefCtx* c, const std::vector<Vector3R>& v, double* full)
{
    for (size_t i = 0; i < c->perm.size(); i++) {
        if (c->perm[i] >= c->nbc) {
            const Vector3R& a = v[c->perm[i] - c->nbc];
            full[3 * i] = a.X(); full[3 * i + 1] = a.Y(); full[3 * i + 2] = a.Z();
        } else {
            full[3 * i] = full[3 * i + 1] = full[3 * i + 2] = 0;
        }
    }
}

// ComputeGradients (FemPhysicsMatrixFree.cpp:536-560) at the current positions and
// load fraction.  The fraction is set through Step's own bookkeeping: a caller
// wanting frac must have stepped there; ref_set_force_fraction below pokes it.
void ref_gradients(void* h, double* r)
{
    RefCtx* c = (RefCtx*)h;
    std::vector<Vector3R> v(c->perm.size() - c->nbc);
    c->phys->ComputeGradients(v);
    from_free(c, v, r);
}

// MatrixVectorMultiply (:523-534)
void ref_matvec(void* h, const double* d, double* q)
{
    RefCtx* c = (RefCtx*)h;
    std::vector<Vector3R> in, out(c->perm.size() - c->nbc);
    to_free(c, d, in);
    c->phys->MatrixVectorMultiply(in, out);
    from_free(c, out, q);
}

// SolveConjugateGradientMF (Math/IterativeSolvers.h:115-195) from x = 0.
int ref_cg(void* h, const double* b, double* x, int iters)
{
    RefCtx* c = (RefCtx*)h;
    std::vector<Vector3R> bb, xx(c->perm.size() - c->nbc);
    to_free(c, b, bb);
    int k = SolveConjugateGradientMF<real, Vector3R, FemPhysicsMatrixFree>(*c->phys, bb, xx, iters);
    from_free(c, xx, x);
    return k;
}

// Step(0) (:297-342): ST_STATIC solves once at full load, ST_QUASI_STATIC adds
// mForceStep to the load fraction and solves.
void ref_step(void* h)
{
    RefCtx* c = (RefCtx*)h;
    c->phys->Step(0);
}

// Step(dt): for ST_EXPLICIT, mNumSubsteps symplectic Euler substeps (SubStep, :371-387)
void ref_step_dt(void* h, double dt)
{
    RefCtx* c = (RefCtx*)h;
    c->phys->Step(dt);
}

// SolveNewtonCG (:604-628) at load fraction frac
void ref_newton_cg(void* h, double frac)
{
    RefCtx* c = (RefCtx*)h;
    FemSystem::SetFraction(c->phys, frac);
    FemSystem::NewtonCG(c->phys);
}

void ref_set_force_fraction(void* h, double frac)
{
    FemSystem::SetFraction(((RefCtx*)h)->phys, frac);
}

double ref_force_fraction(void* h)
{
    return ((RefCtx*)h)->phys->GetForceFraction();
}

// FemPhysicsMatrixFree::ComputeEnergy (:882-916), FemPhysicsBase::ComputeElasticEnergy
// (FemPhysicsBase.cpp:203-212) and ComputeSpringEnergy (:1113-1121)
double ref_energy(void* h) { return ((RefCtx*)h)->phys->ComputeEnergy(); }
double ref_elastic_energy(void* h) { return ((RefCtx*)h)->phys->ComputeElasticEnergy(); }
double ref_spring_energy(void* h) { return ((RefCtx*)h)->phys->ComputeSpringEnergy(); }

// AddCable (FemPhysicsBase.cpp:152-155, through the base class's public virtual)
// with n cable nodes inside elements elem[] at barycentric weights bary[3 n]
void ref_add_cable(void* h, const int32_t* elem, const double* bary, int n, double rest, double stiffness, double damping,
                   double actuation)
{
    RefCtx* c = (RefCtx*)h;
    Cable cb;
    cb.mCableNodes.resize(n);
    cb.mCablePositions.resize(n);
    for (int i = 0; i < n; i++) {
        cb.mCableNodes[i].elem = elem[i];
        cb.mCableNodes[i].bary = Vector3R(bary[3 * i], bary[3 * i + 1], bary[3 * i + 2]);
    }
    cb.mActuation = actuation;
    cb.mCableRestLength = rest;
    cb.mCableStiffness = stiffness;
    cb.mCableDamping = damping;
    static_cast<FemPhysicsBase*>(c->phys)->AddCable(cb);
}

void ref_set_cable_actuation(void* h, int cable, double actuation)
{
    ((RefCtx*)h)->phys->SetCableActuation((uint32)cable, actuation);
}

// FemPhysicsBase::GetForceParamGrads (FemPhysicsBase.cpp:162-187): free-node
// vectors in shuffled order, returned here as full original-order arrays
void ref_force_param_grads(void* h, double* gmu, double* glam, double* grho)
{
    RefCtx* c = (RefCtx*)h;
    EigenVector gm, gl, gr;
    c->phys->GetForceParamGrads(gm, gl, gr);
    const EigenVector* src[3] = {&gm, &gl, &gr};
    double* dst[3] = {gmu, glam, grho};
    for (int k = 0; k < 3; k++)
        for (size_t i = 0; i < c->perm.size(); i++)
            for (int q = 0; q < 3; q++) {
                const uint32 sidx = c->perm[i];
                dst[k][3 * i + q] = sidx >= c->nbc ? (*src[k])(3 * (sidx - c->nbc) + q) : 0.0;
            }
}

}  // extern "C"
