// ref_templates.cpp -- drives the C++ class of this library (the device-backed
// FEM_SYSTEM::FemPhysicsMatrixFree of solidfem_amd/csrc/host) through the
// reference's own solver templates, unchanged, from the reference checkout:
//   SolveConjugateGradientMF<real, Vector3R, FemPhysicsMatrixFree>
//       (Math/IterativeSolvers.h:115-195; needs MatrixVectorMultiply), and
//   NonlinearConjugateGradientMinimizer<FemPhysicsMatrixFree, std::vector<Vector3R>>
//       (Include/NonlinearCG.h:51-114; needs ComputeGradients, DotProduct,
//       MatrixVectorMultiply, UpdatePosAndComputeGradients),
// plus the Krylov hooks on their own.  TEST INFRASTRUCTURE ONLY: built by
// oracle/Makefile (target `templates`) into oracle/_ref/ where the reference
// checkout exists; run by tests/test_gpu_krylov_hooks.py, which compares its
// output with the FP64 oracle.
//
// usage: ref_templates IN OUT
//   IN  (binary): int64 nn, ne; double X[3 nn]; int32 tets[4 ne]; uint8 fixed[nn];
//       double x[3 nn] (state, original order); double frac; double d[3 nf],
//       b[3 nf] (free nodes, shuffled order); int32 cg_iters, ncg_outer
//   OUT (binary): double Kd[3 nf], r[3 nf], dot; int32 cg_its; double xcg[3 nf];
//       double ncg_pos[3 nf]; double ncg_r[3 nf]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "FemPhysicsMatrixFree.h"

// what the reference's templates take from its engine headers (Engine/Utils.h)
#define ASSERT(x)
#define Printf printf
using std::abs;

#include "IterativeSolvers.h"
#include "NonlinearCG.h"

using namespace FEM_SYSTEM;

namespace FEM_SYSTEM {
class FemSystem {
public:
    static void SetFraction(FemPhysicsMatrixFree& p, real f) { p.mForceFraction = f; }
};
}  // namespace FEM_SYSTEM

template <typename T> static bool rd(FILE* f, T* p, size_t n) { return std::fread(p, sizeof(T), n, f) == n; }
template <typename T> static void wr(FILE* f, const T* p, size_t n) { std::fwrite(p, sizeof(T), n, f); }

static void put(FILE* f, const std::vector<Vector3R>& v)
{
    for (const Vector3R& a : v) wr(f, a.v, 3);
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    int64_t nn = 0, ne = 0;
    if (!rd(in, &nn, 1) || !rd(in, &ne, 1) || nn <= 0 || ne <= 0) return 2;
    std::vector<double> X(3 * nn), x(3 * nn);
    std::vector<int32_t> tets(4 * ne);
    std::vector<uint8_t> fixed(nn);
    double frac = 0;
    if (!rd(in, X.data(), X.size()) || !rd(in, tets.data(), tets.size()) || !rd(in, fixed.data(), fixed.size()) ||
        !rd(in, x.data(), x.size()) || !rd(in, &frac, 1))
        return 2;
    int64_t nf = 0;
    for (uint8_t q : fixed) nf += q ? 0 : 1;
    std::vector<double> d(3 * nf), b(3 * nf);
    int32_t cg_iters = 0, ncg_outer = 0;
    if (!rd(in, d.data(), d.size()) || !rd(in, b.data(), b.size()) || !rd(in, &cg_iters, 1) || !rd(in, &ncg_outer, 1))
        return 2;
    std::fclose(in);

    std::vector<Node> nodes(nn);
    for (int64_t i = 0; i < nn; i++) {
        nodes[i].pos = nodes[i].pos0 = Vector3R(X[3 * i], X[3 * i + 1], X[3 * i + 2]);
        nodes[i].invMass = fixed[i] ? 0 : 1;
    }
    std::vector<Tet> t(ne);
    for (int64_t e = 0; e < ne; e++)
        for (int k = 0; k < 4; k++) t[e].idx[k] = (uint32)tets[4 * e + k];
    FemConfig cfg;
    cfg.mMaterial = MMT_NEO_HOOKEAN;
    cfg.mYoungsModulus = 1e6;
    cfg.mPoissonRatio = 0.45;
    cfg.mDensity = 1070;
    cfg.mGravity = -9.8;
    cfg.mZUpAxis = true;
    cfg.mVerbose = false;
    FemPhysicsMatrixFree phys(t, nodes, cfg);
    for (int64_t i = 0; i < nn; i++) phys.SetDeformedPosition((uint32)i, Vector3R(x[3 * i], x[3 * i + 1], x[3 * i + 2]));
    FemSystem::SetFraction(phys, frac);

    std::vector<Vector3R> dv(nf), bv(nf), Kd, r;
    for (int64_t i = 0; i < nf; i++) {
        dv[i] = Vector3R(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        bv[i] = Vector3R(b[3 * i], b[3 * i + 1], b[3 * i + 2]);
    }
    // the Krylov hooks (FemPhysicsMatrixFree.h:85-89)
    phys.MatrixVectorMultiply(dv, Kd);
    phys.ComputeGradients(r);
    const real dot = phys.DotProduct(r, dv);
    // the reference's CG template on this class
    std::vector<Vector3R> xcg(nf);
    const int32_t its = SolveConjugateGradientMF<real, Vector3R, FemPhysicsMatrixFree>(phys, bv, xcg, cg_iters);
    // the reference's nonlinear CG template: minimises from the state x over
    // the free nodes' positions (shuffled order)
    std::vector<Vector3R> pos(nf);
    for (uint32 s = 0; s < (uint32)nf; s++) pos[s] = phys.GetDeformedPosition(s + phys.GetNumNodes() - (uint32)nf, false);
    NonlinearConjugateGradientMinimizer<FemPhysicsMatrixFree, std::vector<Vector3R>> ncg;
    ncg.mOuterIterations = ncg_outer;
    ncg.mInnerIterations = 10;
    ncg.mAbsResidualThreshold = 1e-4;
    ncg.Solve(phys, (uint32)nf, pos);
    std::vector<Vector3R> rn;
    phys.ComputeGradients(rn);

    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    put(out, Kd);
    put(out, r);
    wr(out, &dot, 1);
    wr(out, &its, 1);
    put(out, xcg);
    put(out, pos);
    put(out, rn);
    std::fclose(out);
    return 0;
}
