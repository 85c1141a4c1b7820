// pysolidfem.cpp -- the `pysolidfem` Python module (PySolidFEM/Module.cpp:90-111,
// PyNonlinearFEM.cpp) on top of the MI355X FemPhysicsMatrixFree.
//
// NonlinearFEM(tets, nodes, fixed_nodes, config) keeps the reference's
// constructor, `step`, `simulate_static`, `get_nodes`, `get_tets`,
// `set_lame_params`, `get_shear_modulus`, `get_lame_lambda` and its config keys
// (young, poisson, density, simtype, substeps, maxiter, tol, mixed, optimizer;
// PyNonlinearFEM.cpp:240-292).  New keys select what the reference hard-wires
// (NH + NST_NEWTON_LS, :280-291): solver ("newton", "newton_ls", "newton_cg"),
// material ("linear", "stvk", "corotational", "neo-hookean", "corotational_polar"
// or the enum value), gravity, zup, inner_iters, rel_tol, cg_tol, precond
// ("none", "jacobi", "learned", "multigrid", "multigrid_learned"), force_step,
// young_e / poisson_e (per element arrays).
// The rest of the reference's surface: module-level load_from_xml, and
// save_to_obj, get_boundary_mesh, get_visual_mesh, get_hessian,
// compute_force_param_grads with get_force_mu/lambda/rho_grad, and
// set_cable_actuation (PySolidFEM/Module.cpp:43-111); added: add_cable,
// get_cable_positions, compute_energy, build_boundary_mesh.  Unlike the
// reference's FemBody::BuildBoundaryMesh, no uint16 node indices are involved
// anywhere (meshes above 65,535 nodes keep their boundary).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>

#include <fstream>

#include "FemBody.h"
#include "FemIO.h"
#include "FemPhysicsMatrixFree.h"

namespace py = pybind11;
using namespace FEM_SYSTEM;

namespace {

const real DT = 0.016;   // PyNonlinearFEM.h:51

MaterialModelType parse_material(py::handle h)
{
    if (py::isinstance<py::int_>(h)) return (MaterialModelType)py::cast<int>(h);
    std::string s = py::cast<std::string>(h);
    if (s == "linear") return MMT_LINEAR;
    if (s == "stvk") return MMT_STVK;
    if (s == "corotational" || s == "corot") return MMT_COROTATIONAL;
    if (s == "corotational_polar" || s == "corot_polar") return MMT_COROTATIONAL_POLAR;
    if (s == "neo-hookean" || s == "neohookean" || s == "nh") return MMT_NEO_HOOKEAN;
    throw std::invalid_argument("unknown material '" + s + "'");
}

NonlinearSolverType parse_solver(py::handle h)
{
    if (py::isinstance<py::int_>(h)) return (NonlinearSolverType)py::cast<int>(h);
    std::string s = py::cast<std::string>(h);
    if (s == "newton") return NST_NEWTON;
    if (s == "newton_ls") return NST_NEWTON_LS;
    if (s == "newton_cg") return NST_NEWTON_CG;
    throw std::invalid_argument("unknown solver '" + s + "'");
}

// FemBody::Prepare (FemBody.cpp:127-156), which the reference's bindings run
// on every mesh: rest positions, and tetrahedra with negative signed volume get
// their first two indices swapped
void Prepare(std::vector<Node>& nodes, std::vector<Tet>& tets)
{
    for (Node& n : nodes) n.pos0 = n.pos;
    for (Tet& t : tets) {
        const Vector3R x0 = nodes[t.idx[0]].pos0;
        const Vector3R d1 = nodes[t.idx[1]].pos0 - x0, d2 = nodes[t.idx[2]].pos0 - x0, d3 = nodes[t.idx[3]].pos0 - x0;
        const real det = d1[0] * (d2[1] * d3[2] - d2[2] * d3[1]) - d2[0] * (d1[1] * d3[2] - d1[2] * d3[1]) +
                         d3[0] * (d1[1] * d2[2] - d1[2] * d2[1]);
        if (det < 0) std::swap(t.idx[0], t.idx[1]);
    }
}

// IO::LoadFromXmlFile (FemIO.cpp:583-847), the parts the matrix-free path
// uses: a tet1 model (scale, swap, visual mesh path), its Dirichlet nodes,
// material, simulation type / steps / gravity / pressure surface, solver
// iterations and tolerance, and the first <cable>
struct Scene {
    std::vector<Node> nodes;
    std::vector<Tet> tets;
    std::vector<int> fixed, tris;
    FemConfig config;
    real pressure = 0, scale = 1;
    std::string visualPath;
    std::vector<CableDescriptor> cables;
};

std::vector<int> Ints(py::handle text)
{
    std::vector<int> v;
    if (text.is_none()) return v;
    std::string s = py::cast<std::string>(text);
    for (char& ch : s)
        if (ch == ',') ch = ' ';
    std::istringstream in(s);
    int k;
    while (in >> k) v.push_back(k);
    return v;
}

Scene LoadScene(const std::string& path)
{
    Scene sc;
    py::object ET = py::module_::import("xml.etree.ElementTree");
    py::object root = ET.attr("parse")(path).attr("getroot")();
    if (py::cast<std::string>(root.attr("tag")) != "solidfem") throw std::invalid_argument(path + ": not a <solidfem> scene");
    auto attr = [](py::handle el, const char* name) -> py::object { return el.attr("get")(name); };
    auto num = [&](py::handle el, const char* name) { return std::stod(py::cast<std::string>(attr(el, name))); };
    const std::string dir = path.substr(0, path.find_last_of("/\\") + 1);
    FemConfig& c = sc.config;
    c.mVerbose = false;
    py::object model = root.attr("find")("model");
    if (model.is_none()) throw std::invalid_argument(path + ": no <model>");
    py::object mpath = attr(model, "path"), type = attr(model, "type");
    if (mpath.is_none() || type.is_none() || py::cast<std::string>(type) != "tet1")
        throw std::invalid_argument(path + ": only tet1 models are supported");
    sc.scale = attr(model, "scale").is_none() ? 1 : num(model, "scale");
    if (!attr(model, "visual").is_none()) sc.visualPath = dir + py::cast<std::string>(attr(model, "visual"));
    if (!IO::LoadFromTet1File((dir + py::cast<std::string>(mpath)).c_str(), sc.scale, Vector3R(), sc.nodes, sc.tets))
        throw std::invalid_argument(path + ": cannot read the model");
    if (!attr(model, "swap").is_none()) {
        const std::string sw = py::cast<std::string>(attr(model, "swap"));
        const int a = sw == "yz" ? 1 : 0, b = sw == "xy" ? 1 : 2;
        if (sw == "xy" || sw == "xz" || sw == "yz")
            for (Node& n : sc.nodes) {
                std::swap(n.pos[a], n.pos[b]);
                n.pos0 = n.pos;
            }
    }
    py::object dir_el = model.attr("find")("dirichlet");
    if (!dir_el.is_none()) sc.fixed = Ints(dir_el.attr("text"));
    py::object mat = root.attr("find")("material");
    if (!mat.is_none()) {
        if (!attr(mat, "young").is_none()) c.mYoungsModulus = num(mat, "young");
        if (!attr(mat, "poisson").is_none()) c.mPoissonRatio = num(mat, "poisson");
        if (!attr(mat, "density").is_none()) c.mDensity = num(mat, "density");
    }
    py::object sim = root.attr("find")("simulation");
    if (!sim.is_none()) {
        const std::string st = attr(sim, "type").is_none() ? "" : py::cast<std::string>(attr(sim, "type"));
        if (st == "quasi-static") c.mSimType = ST_QUASI_STATIC;
        else if (st == "explicit") c.mSimType = ST_EXPLICIT;
        if (!attr(sim, "method").is_none() && py::cast<std::string>(attr(sim, "method")) == "mixed")
            throw std::invalid_argument("the mixed (pressure) formulation is not on the matrix-free path");
        if (!attr(sim, "steps").is_none()) {
            const int n = std::stoi(py::cast<std::string>(attr(sim, "steps")));
            if (c.mSimType < ST_EXPLICIT) c.mForceApplicationStep = real(1) / n;
            else c.mNumSubsteps = n;
        }
        if (!attr(sim, "gravity").is_none()) c.mGravity = num(sim, "gravity");
        py::object pr = sim.attr("find")("pressure");
        if (!pr.is_none()) {
            if (!attr(pr, "p").is_none()) sc.pressure = num(pr, "p");
            sc.tris = Ints(pr.attr("text"));
        }
    }
    py::object sol = root.attr("find")("solver");
    if (!sol.is_none()) {
        if (!attr(sol, "iterations").is_none()) c.mOuterIterations = std::stoi(py::cast<std::string>(attr(sol, "iterations")));
        if (!attr(sol, "tolerance").is_none()) c.mAbsNewtonRsidualThreshold = num(sol, "tolerance");
    }
    py::object cab = root.attr("find")("cable");
    if (!cab.is_none()) {
        // FemIO.cpp:801-845
        CableDescriptor d;
        if (!attr(cab, "divs").is_none()) d.divs = std::stoi(py::cast<std::string>(attr(cab, "divs")));
        if (!attr(cab, "len").is_none()) d.length = num(cab, "len");
        if (!attr(cab, "offset").is_none()) {
            std::string o = py::cast<std::string>(attr(cab, "offset"));
            for (char& ch : o)
                if (ch == ',') ch = ' ';
            std::istringstream in(o);
            real x = 0, y = 0, z = 0;
            in >> x >> y >> z;
            d.offset = Vector3R(x, y, z);
        }
        if (!attr(cab, "axis").is_none()) {
            const std::string ax = py::cast<std::string>(attr(cab, "axis"));
            if (!ax.empty() && ax[0] == 'y') d.dir = Vector3R(0, 1, 0);
            else if (!ax.empty() && ax[0] == 'z') d.dir = Vector3R(0, 0, 1);
        }
        if (!attr(cab, "stiffness").is_none()) d.stiffness = num(cab, "stiffness");
        sc.cables.push_back(d);
    }
    for (int i : sc.fixed) {
        if (i < 0 || (size_t)i >= sc.nodes.size()) throw std::invalid_argument("fixed node index out of range");
        sc.nodes[i].invMass = 0;
    }
    return sc;
}

// PyLoadFromXml (PySolidFEM/Module.cpp:43-87): (nodes (nn, 3), tets (ne, 4), fixed node list)
py::tuple PyLoadFromXml(const std::string& path)
{
    Scene sc = LoadScene(path);
    py::array_t<double> nodes({(py::ssize_t)sc.nodes.size(), (py::ssize_t)3});
    double* pn = nodes.mutable_data();
    for (size_t i = 0; i < sc.nodes.size(); i++)
        for (int c = 0; c < 3; c++) pn[3 * i + c] = sc.nodes[i].pos[c];
    py::array_t<int> tets({(py::ssize_t)sc.tets.size(), (py::ssize_t)4});
    int* pt = tets.mutable_data();
    for (size_t e = 0; e < sc.tets.size(); e++)
        for (int k = 0; k < 4; k++) pt[4 * e + k] = (int)sc.tets[e].idx[k];
    return py::make_tuple(nodes, tets, py::cast(sc.fixed));
}

// a Mesh as the reference's bindings return it (PyNonlinearFEM.cpp:196-238):
// (vertices (n, 3), flat triangle index list, boundary node list), 32-bit
// indices throughout (the reference's uint16 truncation, FemBody.cpp:263,
// 310-313, is not reproduced)
py::tuple MeshTuple(const Mesh& mesh, const std::vector<int>& boundaryNodes)
{
    py::array_t<double> verts({(py::ssize_t)mesh.vertices.size(), (py::ssize_t)3});
    double* pv = verts.mutable_data();
    for (size_t i = 0; i < mesh.vertices.size(); i++)
        for (int c = 0; c < 3; c++) pv[3 * i + c] = mesh.vertices[i][c];
    return py::make_tuple(verts, py::cast(mesh.indices), py::cast(boundaryNodes));
}

// FemBody::BuildBoundaryMesh on a mesh (tets (ne, 4), nodes (nn, 3)), without
// a physics context: the mesh tuple above at the given positions
py::tuple PyBuildBoundaryMesh(py::array_t<int, py::array::c_style | py::array::forcecast> tets,
                              py::array_t<double, py::array::c_style | py::array::forcecast> nodes)
{
    if (nodes.ndim() != 2 || nodes.shape(1) != 3 || tets.size() % 4) throw std::invalid_argument("tets (ne, 4), nodes (nn, 3)");
    std::vector<Node> n(nodes.shape(0));
    for (size_t i = 0; i < n.size(); i++) n[i].pos = n[i].pos0 = Vector3R(nodes.data()[3 * i], nodes.data()[3 * i + 1], nodes.data()[3 * i + 2]);
    std::vector<Tet> t(tets.size() / 4);
    for (size_t e = 0; e < t.size(); e++)
        for (int k = 0; k < 4; k++) {
            const int v = tets.data()[4 * e + k];
            if (v < 0 || (size_t)v >= n.size()) throw std::invalid_argument("tet index out of range");
            t[e].idx[k] = (uint32)v;
        }
    FemBody body;
    body.BuildBoundaryMesh(t, n);
    return MeshTuple(body.GetBoundaryMesh(), body.GetBoundaryNodes());
}

class PyNonlinearFEM {
public:
    PyNonlinearFEM(py::array_t<int, py::array::c_style | py::array::forcecast> tets,
                   py::array_t<double, py::array::c_style | py::array::forcecast> nodes,
                   py::array_t<int, py::array::c_style | py::array::forcecast> fixed, py::dict config)
    {
        if (nodes.ndim() != 2 || nodes.shape(1) != 3) throw std::invalid_argument("nodes should have 3 columns");
        if (tets.size() % 4) throw std::invalid_argument("tets should have 4 indices per element");
        const size_t nn = nodes.shape(0), ne = tets.size() / 4;
        std::vector<Node> n(nn);
        const double* pn = nodes.data();
        for (size_t i = 0; i < nn; i++) {
            n[i].pos = Vector3R(pn[3 * i], pn[3 * i + 1], pn[3 * i + 2]);
            n[i].pos0 = n[i].pos;
        }
        const int* pf = fixed.data();
        for (ssize_t i = 0; i < fixed.size(); i++) {
            if (pf[i] < 0 || (size_t)pf[i] >= nn) throw std::invalid_argument("fixed node index out of range");
            n[pf[i]].invMass = 0;
        }
        std::vector<Tet> t(ne);
        const int* pt = tets.data();
        for (size_t e = 0; e < ne; e++)
            for (int k = 0; k < 4; k++) {
                if (pt[4 * e + k] < 0 || (size_t)pt[4 * e + k] >= nn) throw std::invalid_argument("tet index out of range");
                t[e].idx[k] = (uint32)pt[4 * e + k];
            }
        Prepare(n, t);
        mBody.BuildBoundaryMesh(t, n);
        FemConfig cfg = ParseConfig(config);
        mPhys.reset(new FemPhysicsMatrixFree(t, n, cfg));
        if (config.contains("young_e") || config.contains("poisson_e")) {
            auto E = py::cast<std::vector<real>>(config.contains("young_e") ? config["young_e"] : py::cast(std::vector<real>(ne, cfg.mYoungsModulus)));
            auto nu = py::cast<std::vector<real>>(config.contains("poisson_e") ? config["poisson_e"] : py::cast(std::vector<real>(ne, cfg.mPoissonRatio)));
            mPhys->SetPerElementMaterial(E, nu);
        }
    }

    // PyNonlinearFEM(path) (PyNonlinearFEM.cpp:68-92): the <solidfem> scene's
    // mesh, Dirichlet nodes and FemConfig (IO::LoadFromXmlFile, FemIO.cpp:583-847),
    // then Neo-Hookean + NST_NEWTON_LS as the reference sets them
    explicit PyNonlinearFEM(const std::string& path) { InitFromXml(path, nullptr); }

    // PyNonlinearFEM(path, config) (:94-121): the scene's mesh and Dirichlet
    // nodes; the config dict replaces the scene's FemConfig
    PyNonlinearFEM(const std::string& path, py::dict config) { InitFromXml(path, &config); }

    // SaveToVTK (:176-184): IO::ExportToVTKHeatMap of the current state
    void SaveToVTK(const std::string& path) const
    {
        std::ofstream out(path);
        if (!out) throw std::runtime_error("cannot write " + path);
        IO::ExportToVTKHeatMap(out, *mPhys);
    }

    // Step (PyNonlinearFEM.cpp:126-132): the physics step, then the boundary
    // and visual meshes follow (SimulateStatic below does not update them, as
    // in the reference)
    void Step(real dt)
    {
        mPhys->Step(dt);
        mBody.UpdateBoundaryMesh(*mPhys);
        if (mBody.HasVisualMesh()) mBody.UpdateVisualMesh();
    }

    // SaveToOBJ (:186-194): the visual mesh when there is one, else the boundary mesh
    void SaveToOBJ(const std::string& path) const
    {
        FILE* f = std::fopen(path.c_str(), "wt");
        if (!f) throw std::runtime_error("cannot write " + path);
        ExportMeshToOBJ(f, mBody.HasVisualMesh() ? mBody.GetVisualMesh() : mBody.GetBoundaryMesh());
        std::fclose(f);
    }

    // GetBoundaryMesh, GetVisualMesh (:196-238)
    py::tuple GetBoundaryMesh() const { return MeshTuple(mBody.GetBoundaryMesh(), mBody.GetBoundaryNodes()); }
    py::tuple GetVisualMesh() const { return MeshTuple(mBody.GetVisualMesh(), mBody.GetBoundaryNodes()); }

    // GetHessian (PyNonlinearFEM.h:67): dense, free dofs in shuffled order
    py::array_t<double> GetHessian() const
    {
        std::vector<real> K;
        mPhys->GetHessian(K);
        const py::ssize_t n = 3 * (py::ssize_t)mPhys->GetNumFreeNodes();
        py::array_t<double> out({n, n});
        std::copy(K.begin(), K.end(), out.mutable_data());
        return out;
    }

    // ComputeForceParamGrads and its getters (PyNonlinearFEM.h:68-71); the
    // reference's get_force_rho_grad returns the lambda gradient, here the
    // density gradient
    void ComputeForceParamGrads() { mPhys->GetForceParamGrads(mGradMu, mGradLambda, mGradRho); }
    py::array_t<double> GetForceMuGrad() const { return py::array_t<double>(mGradMu.size(), mGradMu.data()); }
    py::array_t<double> GetForceLambdaGrad() const { return py::array_t<double>(mGradLambda.size(), mGradLambda.data()); }
    py::array_t<double> GetForceRhoGrad() const { return py::array_t<double>(mGradRho.size(), mGradRho.data()); }

    // SetCableActuation (PyNonlinearFEM.h:74); AddCable takes a cable of n
    // nodes in elements elem at barycentric weights bary (n, 3)
    void SetCableActuation(uint32 cable, real actuation) { mPhys->SetCableActuation(cable, actuation); }
    int AddCable(py::array_t<int, py::array::c_style | py::array::forcecast> elem,
                 py::array_t<double, py::array::c_style | py::array::forcecast> bary, real rest, real stiffness, real damping)
    {
        if (bary.size() != 3 * elem.size()) throw std::invalid_argument("add_cable: bary must be n x 3");
        Cable cb;
        cb.mCableNodes.resize(elem.size());
        for (py::ssize_t i = 0; i < elem.size(); i++) {
            cb.mCableNodes[i].elem = elem.data()[i];
            cb.mCableNodes[i].bary = Vector3R(bary.data()[3 * i], bary.data()[3 * i + 1], bary.data()[3 * i + 2]);
        }
        cb.mCablePositions.resize(elem.size());
        cb.mCableRestLength = rest;
        cb.mCableStiffness = stiffness;
        cb.mCableDamping = damping;
        AddCreatedCable(cb);
        return (int)mPhys->GetNumCables() - 1;
    }
    int NumCables() const { return (int)mPhys->GetNumCables(); }
    py::array_t<double> GetCablePositions(uint32 cable) const
    {
        if (cable >= mPhys->GetNumCables()) throw std::out_of_range("no such cable");
        const size_t n = mPhys->GetCable(cable).mCableNodes.size();
        py::array_t<double> out({(py::ssize_t)n, (py::ssize_t)3});
        double* p = out.mutable_data();
        for (size_t i = 0; i < n; i++) {
            const Vector3R v = mPhys->GetCablePosition(cable, (uint32)i);
            for (int c = 0; c < 3; c++) p[3 * i + c] = v[c];
        }
        return out;
    }

    // FemPhysicsMatrixFree::ComputeEnergy (FemPhysicsMatrixFree.cpp:882-916)
    real ComputeEnergy() { return mPhys->ComputeEnergy(); }
    real ComputeElasticEnergy() { return mPhys->ComputeElasticEnergy(); }

    // SimulateStatic (:134-143): step until the load fraction reaches 1
    void SimulateStatic()
    {
        int guard = 0;
        while (mPhys->GetForceFraction() < 1 && guard++ < 1000000) mPhys->Step(0);
    }

    // GetNodes (:145-159): positions in original order, (nn, 3)
    py::array_t<double> GetNodes() const
    {
        const auto& nodes = mPhys->GetNodes();
        py::array_t<double> out({(py::ssize_t)nodes.size(), (py::ssize_t)3});
        double* p = out.mutable_data();
        for (size_t i = 0; i < nodes.size(); i++)
            for (int c = 0; c < 3; c++) p[3 * i + c] = nodes[i].pos[c];
        return out;
    }

    py::array_t<int> GetTets() const
    {
        const uint32 ne = mPhys->GetNumElements();
        py::array_t<int> out({(py::ssize_t)ne, (py::ssize_t)4});
        int* p = out.mutable_data();
        for (uint32 e = 0; e < ne; e++)
            for (int k = 0; k < 4; k++) p[4 * e + k] = (int)mPhys->GetGlobalOriginalIndex(e, k);
        return out;
    }

    void SetLameParams(real mu, real lambda) { mPhys->SetLameParams(mu, lambda); }
    real GetShearModulus() const { return mPhys->GetShearModulus(); }
    real GetLameLambda() const { return mPhys->GetLameFirstParam(); }

    void SetPressure(py::array_t<int, py::array::c_style | py::array::forcecast> tris, real p)
    {
        std::vector<uint32> t(tris.data(), tris.data() + tris.size());
        mPhys->SetBoundaryConditionsSurface(t, p);
    }

    void SetExternalForces(py::array_t<double, py::array::c_style | py::array::forcecast> f)
    {
        std::vector<Vector3R> v(f.size() / 3);
        const double* p = f.data();
        for (size_t i = 0; i < v.size(); i++) v[i] = Vector3R(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        mPhys->SetExternalForces(v);
    }

    void SetPreconditionerBlocks(py::array_t<float, py::array::c_style | py::array::forcecast> m)
    {
        mPhys->SetPreconditionerBlocks(std::vector<float>(m.data(), m.data() + m.size()));
    }

    // one coarser nested level of the multigrid hierarchy, finest first (the
    // tuples solidfem_amd.mesh.beam_hierarchy returns unpack into it)
    void AddMultigridLevel(py::array_t<double, py::array::c_style | py::array::forcecast> nodes,
                           py::array_t<int, py::array::c_style | py::array::forcecast> tets,
                           py::array_t<uint8_t, py::array::c_style | py::array::forcecast> fixed,
                           py::array_t<int, py::array::c_style | py::array::forcecast> inject,
                           py::array_t<int64_t, py::array::c_style | py::array::forcecast> pPtr,
                           py::array_t<int, py::array::c_style | py::array::forcecast> pCol,
                           py::array_t<double, py::array::c_style | py::array::forcecast> pVal, py::object young,
                           py::object poisson)
    {
        if (nodes.size() % 3 != 0) throw std::runtime_error("add_multigrid_level: nodes must be n x 3");
        FemPhysicsMatrixFree::MultigridLevel l;
        l.mNodes.resize(nodes.size() / 3);
        const double* x = nodes.data();
        for (size_t i = 0; i < l.mNodes.size(); i++) l.mNodes[i] = Vector3R(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        l.mTets.assign(tets.data(), tets.data() + tets.size());
        l.mFixed.assign(fixed.data(), fixed.data() + fixed.size());
        l.mInject.assign(inject.data(), inject.data() + inject.size());
        l.mProlongPtr.assign(pPtr.data(), pPtr.data() + pPtr.size());
        l.mProlongCol.assign(pCol.data(), pCol.data() + pCol.size());
        l.mProlongVal.assign(pVal.data(), pVal.data() + pVal.size());
        if (young.is_none() != poisson.is_none())
            throw std::runtime_error("add_multigrid_level: give both coarse moduli arrays or neither");
        if (!young.is_none()) {
            auto e = py::cast<py::array_t<double, py::array::c_style | py::array::forcecast>>(young);
            auto n = py::cast<py::array_t<double, py::array::c_style | py::array::forcecast>>(poisson);
            l.mYoung.assign(e.data(), e.data() + e.size());
            l.mPoisson.assign(n.data(), n.data() + n.size());
        }
        mPhys->AddMultigridLevel(l);
    }

    void ClearMultigrid() { mPhys->ClearMultigrid(); }
    int NumMultigridLevels() const { return (int)mPhys->GetNumMultigridLevels(); }

    py::dict LastStats() const
    {
        const sfem_solve_stats& s = mPhys->GetLastSolveStats();
        py::dict d;
        d["newton_iters"] = s.newton_iters;
        d["cg_iters"] = s.cg_iters;
        d["converged"] = s.converged;
        d["fallbacks"] = s.fallbacks;
        d["reruns"] = s.reruns;
        d["rejects"] = s.rejects;
        d["res0"] = s.res0;
        d["res"] = s.res;
        d["ext_norm"] = s.ext_norm;
        d["seconds"] = s.seconds;
        d["hessvec_evals"] = s.hessvec_evals;
        d["force_evals"] = s.force_evals;
        return d;
    }

    real GetForceFraction() const { return mPhys->GetForceFraction(); }
    bool CheckForInversion() const { return mPhys->CheckForInversion(0); }

private:
    void InitFromXml(const std::string& path, py::dict* config)
    {
        Scene sc = LoadScene(path);
        FemConfig c = sc.config;
        if (config) {
            c = ParseConfig(*config);
        } else {
            c.mMaterial = MMT_NEO_HOOKEAN;
            mCfg.mSolver = NST_NEWTON_LS;
            c.mInnerIterations = 100000;
            c.mCustomConfig = &mCfg;
        }
        Prepare(sc.nodes, sc.tets);
        mBody.BuildBoundaryMesh(sc.tets, sc.nodes);
        mPhys.reset(new FemPhysicsMatrixFree(sc.tets, sc.nodes, c));
        if (!sc.tris.empty())
            mPhys->SetBoundaryConditionsSurface(std::vector<uint32>(sc.tris.begin(), sc.tris.end()), sc.pressure);
        // the scene's visual mesh (FemBody::LoadFromXml, FemBody.cpp:106-125) and
        // cables (PyNonlinearFEM.cpp:88-91)
        if (!sc.visualPath.empty() && !mBody.LoadVisualMesh(sc.visualPath.c_str(), (float)sc.scale))
            throw std::invalid_argument(path + ": cannot read the visual mesh " + sc.visualPath);
        for (const CableDescriptor& d : sc.cables) AddCreatedCable(CreateCable(d, sc.nodes, sc.tets));
    }

    void AddCreatedCable(const Cable& cb)
    {
        for (const SpringNode& n : cb.mCableNodes)
            if (n.elem < 0)
                throw std::invalid_argument("cable: a point lies outside the mesh (a free cable node, which the "
                                            "matrix-free path does not support)");
        mPhys->AddCable(cb);
    }

    FemConfig ParseConfig(py::dict config)
    {
        FemConfig c;
        c.mVerbose = false;
        if (config.contains("young")) c.mYoungsModulus = py::cast<real>(config["young"]);
        if (config.contains("poisson")) c.mPoissonRatio = py::cast<real>(config["poisson"]);
        if (config.contains("density")) c.mDensity = py::cast<real>(config["density"]);
        if (config.contains("simtype")) c.mSimType = (SimulationType)py::cast<int>(config["simtype"]);
        if (config.contains("substeps")) c.mNumSubsteps = py::cast<int>(config["substeps"]);
        if (config.contains("maxiter")) c.mOuterIterations = py::cast<int>(config["maxiter"]);
        if (config.contains("tol")) c.mAbsNewtonRsidualThreshold = py::cast<real>(config["tol"]);
        if (config.contains("mixed") && py::cast<bool>(config["mixed"]))
            throw std::invalid_argument("the mixed (pressure) formulation is not on the matrix-free path");
        if (config.contains("optimizer")) mCfg.mOptimizer = py::cast<bool>(config["optimizer"]);
        // the reference bindings hard-wire Neo-Hookean + NST_NEWTON_LS
        c.mMaterial = config.contains("material") ? parse_material(config["material"]) : MMT_NEO_HOOKEAN;
        mCfg.mSolver = config.contains("solver") ? parse_solver(config["solver"]) : NST_NEWTON_LS;
        if (config.contains("gravity")) c.mGravity = py::cast<real>(config["gravity"]);
        if (config.contains("zup")) c.mZUpAxis = py::cast<bool>(config["zup"]);
        if (config.contains("inner_iters")) c.mInnerIterations = py::cast<int>(config["inner_iters"]);
        else if (mCfg.mSolver != NST_NEWTON_CG) c.mInnerIterations = 100000;
        if (config.contains("force_step")) c.mForceApplicationStep = py::cast<real>(config["force_step"]);
        if (config.contains("rel_tol")) mCfg.mTolerance = py::cast<real>(config["rel_tol"]);
        if (config.contains("cg_tol")) mCfg.mCGTolerance = py::cast<real>(config["cg_tol"]);
        if (config.contains("verbose")) c.mVerbose = py::cast<bool>(config["verbose"]);
        if (config.contains("predictor")) mCfg.mPredictor = py::cast<int>(config["predictor"]);
        if (config.contains("precond")) {
            std::string p = py::cast<std::string>(config["precond"]);
            if (p == "none") mCfg.mPreconditioner = FemPhysicsMatrixFree::PC_NONE;
            else if (p == "jacobi") mCfg.mPreconditioner = FemPhysicsMatrixFree::PC_JACOBI;
            else if (p == "learned") mCfg.mPreconditioner = FemPhysicsMatrixFree::PC_LEARNED;
            else if (p == "multigrid") mCfg.mPreconditioner = FemPhysicsMatrixFree::PC_MULTIGRID;
            else if (p == "multigrid_learned") mCfg.mPreconditioner = FemPhysicsMatrixFree::PC_MULTIGRID_LEARNED;
            else throw std::runtime_error("precond: none, jacobi, learned, multigrid or multigrid_learned expected");
        }
        c.mCustomConfig = &mCfg;
        return c;
    }

    FemPhysicsMatrixFree::Config mCfg;
    std::unique_ptr<FemPhysicsMatrixFree> mPhys;
    FemBody mBody;
    std::vector<real> mGradMu, mGradLambda, mGradRho;
};

}  // namespace

PYBIND11_MODULE(pysolidfem, m)
{
    m.doc() = "SolidFEM-compatible bindings of the MI355X matrix-free Newton-CG solver";
    m.def("load_from_xml", &PyLoadFromXml);
    m.def("build_boundary_mesh", &PyBuildBoundaryMesh);
    py::class_<PyNonlinearFEM>(m, "NonlinearFEM")
        .def(py::init<py::array_t<int, py::array::c_style | py::array::forcecast>,
                      py::array_t<double, py::array::c_style | py::array::forcecast>,
                      py::array_t<int, py::array::c_style | py::array::forcecast>, py::dict>())
        .def(py::init<const std::string&>())
        .def(py::init<const std::string&, py::dict>())
        .def("save_to_vtk", &PyNonlinearFEM::SaveToVTK)
        .def("step", &PyNonlinearFEM::Step, py::arg("dt") = DT)
        .def("simulate_static", &PyNonlinearFEM::SimulateStatic)
        .def("get_nodes", &PyNonlinearFEM::GetNodes)
        .def("get_tets", &PyNonlinearFEM::GetTets)
        .def("set_lame_params", &PyNonlinearFEM::SetLameParams)
        .def("get_shear_modulus", &PyNonlinearFEM::GetShearModulus)
        .def("get_lame_lambda", &PyNonlinearFEM::GetLameLambda)
        .def("set_pressure", &PyNonlinearFEM::SetPressure)
        .def("set_external_forces", &PyNonlinearFEM::SetExternalForces)
        .def("set_preconditioner_blocks", &PyNonlinearFEM::SetPreconditionerBlocks)
        .def("add_multigrid_level", &PyNonlinearFEM::AddMultigridLevel, py::arg("nodes"), py::arg("tets"),
             py::arg("fixed"), py::arg("inject"), py::arg("p_ptr"), py::arg("p_col"), py::arg("p_val"),
             py::arg("young_e") = py::none(), py::arg("poisson_e") = py::none())
        .def("clear_multigrid", &PyNonlinearFEM::ClearMultigrid)
        .def("num_multigrid_levels", &PyNonlinearFEM::NumMultigridLevels)
        .def("last_stats", &PyNonlinearFEM::LastStats)
        .def("get_force_fraction", &PyNonlinearFEM::GetForceFraction)
        .def("check_for_inversion", &PyNonlinearFEM::CheckForInversion)
        .def("save_to_obj", &PyNonlinearFEM::SaveToOBJ)
        .def("get_boundary_mesh", &PyNonlinearFEM::GetBoundaryMesh)
        .def("get_visual_mesh", &PyNonlinearFEM::GetVisualMesh)
        .def("get_hessian", &PyNonlinearFEM::GetHessian)
        .def("compute_force_param_grads", &PyNonlinearFEM::ComputeForceParamGrads)
        .def("get_force_mu_grad", &PyNonlinearFEM::GetForceMuGrad)
        .def("get_force_lambda_grad", &PyNonlinearFEM::GetForceLambdaGrad)
        .def("get_force_rho_grad", &PyNonlinearFEM::GetForceRhoGrad)
        .def("set_cable_actuation", &PyNonlinearFEM::SetCableActuation)
        .def("add_cable", &PyNonlinearFEM::AddCable, py::arg("elem"), py::arg("bary"), py::arg("rest_length"),
             py::arg("stiffness") = 1000.0, py::arg("damping") = 0.0)
        .def("num_cables", &PyNonlinearFEM::NumCables)
        .def("get_cable_positions", &PyNonlinearFEM::GetCablePositions)
        .def("compute_energy", &PyNonlinearFEM::ComputeEnergy)
        .def("compute_elastic_energy", &PyNonlinearFEM::ComputeElasticEnergy);
}
