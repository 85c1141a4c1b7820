// FemBody.cpp -- boundary mesh, visual mesh, OBJ I/O and cable registration
// of the host API (the roles of Source/FemBody.cpp and Source/MeshTools.cpp).
#include "FemBody.h"

#include <algorithm>
#include <cfloat>
#include <fstream>
#include <sstream>

namespace FEM_SYSTEM {

namespace {

typedef Vector3T<float> Vec3f;

Vec3f to_float(const Vector3R& v) { return Vec3f((float)v[0], (float)v[1], (float)v[2]); }

// signed volume det[b - a, c - a, d - a] / 6 (ComputeTetVolume, FemBody.cpp:8-15)
real tet_volume(const Vector3R& a, const Vector3R& b, const Vector3R& c, const Vector3R& d)
{
    return Matrix3R(b - a, c - a, d - a).Determinant() / 6;
}

// the point of triangle abc closest to p: the Voronoi-region case analysis of
// C. Ericson, Real-Time Collision Detection (2005), section 5.1.5
Vec3f closest_point_on_triangle(const Vec3f& p, const Vec3f& a, const Vec3f& b, const Vec3f& c)
{
    const Vec3f ab = b - a, ac = c - a, ap = p - a;
    const float d1 = ab.Dot(ap), d2 = ac.Dot(ap);
    if (d1 <= 0 && d2 <= 0) return a;
    const Vec3f bp = p - b;
    const float d3 = ab.Dot(bp), d4 = ac.Dot(bp);
    if (d3 >= 0 && d4 <= d3) return b;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return a + (d1 / (d1 - d3)) * ab;
    const Vec3f cp = p - c;
    const float d5 = ab.Dot(cp), d6 = ac.Dot(cp);
    if (d6 >= 0 && d5 <= d6) return c;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return a + (d2 / (d2 - d6)) * ac;
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
    const float den = 1.0f / (va + vb + vc);
    const float v = vb * den, w = vc * den;
    return a + v * ab + w * ac;
}

// p = a + u (c - a) + v (b - a): weights w = 1 - u - v of a, v of b, u of c
// (the convention of BarycentricCoordinates, MeshTools.cpp:46-70)
void barycentric_in_plane(const Vec3f& p, const Vec3f& a, const Vec3f& b, const Vec3f& c, float& u, float& v, float& w)
{
    const Vec3f e0 = c - a, e1 = b - a, e2 = p - a;
    const float d00 = e0.Dot(e0), d01 = e0.Dot(e1), d02 = e0.Dot(e2), d11 = e1.Dot(e1), d12 = e1.Dot(e2);
    const float inv = 1.0f / (d00 * d11 - d01 * d01);
    u = (d11 * d02 - d01 * d12) * inv;
    v = (d00 * d12 - d01 * d02) * inv;
    w = 1.0f - u - v;
}

struct Face {
    uint32 i, j, k;   // sorted vertex ids
    uint32 elem;
    int face;         // the tet corner the face is opposite to
};

}  // namespace

void Mesh::ComputeNormals()
{
    normals.assign(vertices.size(), Vec3f());
    for (size_t t = 0; t + 2 < indices.size(); t += 3) {
        const Vec3f& a = vertices[indices[t]];
        const Vec3f n = (vertices[indices[t + 1]] - a).Cross(vertices[indices[t + 2]] - a);
        for (int k = 0; k < 3; k++) normals[indices[t + k]] += n;
    }
    for (Vec3f& n : normals) n.Normalize();
}

void FemBody::BuildBoundaryMesh(const std::vector<Tet>& tets, const std::vector<Node>& nodes)
{
    std::vector<Face> faces;
    faces.reserve(4 * tets.size());
    for (uint32 e = 0; e < (uint32)tets.size(); e++) {
        const uint32* t = tets[e].idx;
        for (int f = 0; f < 4; f++) {
            uint32 v[3], n = 0;
            for (int c = 0; c < 4; c++)
                if (c != f) v[n++] = t[c];
            std::sort(v, v + 3);
            faces.push_back(Face{v[0], v[1], v[2], e, f});
        }
    }
    std::stable_sort(faces.begin(), faces.end(), [](const Face& a, const Face& b) {
        if (a.i != b.i) return a.i < b.i;
        if (a.j != b.j) return a.j < b.j;
        return a.k < b.k;
    });
    std::vector<Face> boundary;
    std::vector<uint32> bverts;
    for (size_t f = 0; f < faces.size(); f++) {
        if (f + 1 < faces.size() && faces[f].i == faces[f + 1].i && faces[f].j == faces[f + 1].j &&
            faces[f].k == faces[f + 1].k) {
            f++;   // an interior face: shared with the next tetrahedron
            continue;
        }
        boundary.push_back(faces[f]);
        bverts.insert(bverts.end(), {faces[f].i, faces[f].j, faces[f].k});
    }
    std::sort(bverts.begin(), bverts.end());
    bverts.erase(std::unique(bverts.begin(), bverts.end()), bverts.end());
    mNodesToBoundaryMap.assign(nodes.size(), -1);
    mBoundaryToNodesMap.clear();
    mBoundaryMesh = Mesh();
    for (uint32 v : bverts) {
        mNodesToBoundaryMap[v] = (int)mBoundaryMesh.vertices.size();
        mBoundaryToNodesMap.push_back((int)v);
        mBoundaryMesh.vertices.push_back(to_float(nodes[v].pos));
    }
    for (const Face& f : boundary) {
        const uint32* t = tets[f.elem].idx;
        uint32 v[3], n = 0;
        for (int c = 0; c < 4; c++)
            if (c != f.face) v[n++] = t[c];
        const Vec3f a = to_float(nodes[v[0]].pos), b = to_float(nodes[v[1]].pos), c = to_float(nodes[v[2]].pos);
        const Vec3f inner = to_float(nodes[t[f.face]].pos);
        const bool flip = (b - a).Cross(c - a).Dot(inner - a) > 0;
        mBoundaryMesh.AddTriangle((uint32)mNodesToBoundaryMap[v[0]], (uint32)mNodesToBoundaryMap[v[1]],
                                  (uint32)mNodesToBoundaryMap[v[2]], flip);
    }
    mBoundaryMesh.ComputeNormals();
}

void FemBody::UpdateBoundaryMesh(const FemPhysicsMatrixFree& phys)
{
    const std::vector<Node>& nodes = phys.GetNodes();
    for (size_t i = 0; i < nodes.size() && i < mNodesToBoundaryMap.size(); i++)
        if (mNodesToBoundaryMap[i] >= 0) mBoundaryMesh.vertices[mNodesToBoundaryMap[i]] = to_float(nodes[i].pos);
    mBoundaryMesh.ComputeNormals();
}

bool FemBody::LoadVisualMesh(const char* path, float scale)
{
    if (!ImportMeshFromOBJ(path, mVisualMesh, scale)) return false;
    const Mesh& bm = mBoundaryMesh;
    mMapData.resize(mVisualMesh.vertices.size());
    for (size_t j = 0; j < mVisualMesh.vertices.size(); j++) {
        const Vec3f& p = mVisualMesh.vertices[j];
        float best = FLT_MAX;
        size_t tri = 0;
        Vec3f closest;
        for (size_t t = 0; t < bm.GetNumTriangles(); t++) {
            const Vec3f q = closest_point_on_triangle(p, bm.vertices[bm.indices[3 * t]], bm.vertices[bm.indices[3 * t + 1]],
                                                      bm.vertices[bm.indices[3 * t + 2]]);
            const float d = (p - q).Length();
            if (d < best) {
                best = d;
                tri = t;
                closest = q;
            }
        }
        MeshInterp m{};
        m.id1 = bm.indices[3 * tri];
        m.id2 = bm.indices[3 * tri + 1];
        m.id3 = bm.indices[3 * tri + 2];
        const Vec3f &a = bm.vertices[m.id1], &b = bm.vertices[m.id2], &c = bm.vertices[m.id3];
        Vec3f n = (b - a).Cross(c - a);
        n.Normalize();
        m.distance = n.Dot(p - closest);
        barycentric_in_plane(closest, a, b, c, m.u, m.v, m.w);
        mMapData[j] = m;
    }
    return true;
}

void FemBody::UpdateVisualMesh()
{
    for (size_t j = 0; j < mMapData.size(); j++) {
        const MeshInterp& m = mMapData[j];
        const Vec3f &a = mBoundaryMesh.vertices[m.id1], &b = mBoundaryMesh.vertices[m.id2], &c = mBoundaryMesh.vertices[m.id3];
        Vec3f n = (b - a).Cross(c - a);
        n.Normalize();
        mVisualMesh.vertices[j] = m.w * a + m.v * b + m.u * c + m.distance * n;
    }
    mVisualMesh.ComputeNormals();
}

void ExportMeshToOBJ(FILE* out, const Mesh& mesh)
{
    for (size_t j = 0; j < mesh.vertices.size(); j++) {
        const Vec3f n = j < mesh.normals.size() ? mesh.normals[j] : Vec3f();
        std::fprintf(out, "vn %f %f %f\n", n[0], n[1], n[2]);
        std::fprintf(out, "v %f %f %f\n", mesh.vertices[j][0], mesh.vertices[j][1], mesh.vertices[j][2]);
    }
    for (size_t j = 0; j + 2 < mesh.indices.size(); j += 3)
        std::fprintf(out, "f %u//%u %u//%u %u//%u\n", mesh.indices[j] + 1, mesh.indices[j] + 1, mesh.indices[j + 1] + 1,
                     mesh.indices[j + 1] + 1, mesh.indices[j + 2] + 1, mesh.indices[j + 2] + 1);
}

bool ImportMeshFromOBJ(const char* path, Mesh& mesh, float scale)
{
    std::ifstream in(path);
    if (!in) return false;
    mesh = Mesh();
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "v") {
            float x, y, z;
            if (ls >> x >> y >> z) mesh.vertices.push_back(Vec3f(scale * x, scale * y, scale * z));
        } else if (tag == "f") {
            std::string c;
            uint32 ids[3];
            int k = 0;
            while (k < 3 && ls >> c) ids[k++] = (uint32)std::stoul(c.substr(0, c.find('/'))) - 1;
            if (k == 3) mesh.indices.insert(mesh.indices.end(), {ids[0], ids[1], ids[2]});
        }
    }
    for (uint32 i : mesh.indices)
        if (i >= mesh.vertices.size()) return false;
    mesh.ComputeNormals();
    return true;
}

Cable CreateCable(const CableDescriptor& d, const std::vector<Node>& nodes, const std::vector<Tet>& tets, real scale)
{
    Cable cable;
    const int n = d.divs;
    const real div = d.length / n;
    cable.mCableNodes.resize(n + 1);
    cable.mCablePositions.resize(n + 1);
    for (int i = 0; i <= n; i++) {
        const Vector3R p = (i * div) * d.dir + d.offset;
        real best = 1e20, bestPos = 1e20;
        int elem = -1, elemPos = -1;
        for (uint32 e = 0; e < (uint32)tets.size(); e++) {
            const uint32* t = tets[e].idx;
            const Vector3R x0 = scale * nodes[t[0]].pos, x1 = scale * nodes[t[1]].pos, x2 = scale * nodes[t[2]].pos,
                           x3 = scale * nodes[t[3]].pos;
            const real vol = tet_volume(x0, x1, x2, x3);
            const real w[4] = {tet_volume(p, x1, x2, x3) / vol, tet_volume(x0, p, x2, x3) / vol,
                               tet_volume(x0, x1, p, x3) / vol, tet_volume(x0, x1, x2, p) / vol};
            const bool inside = w[0] > -0.001 && w[1] > -0.001 && w[2] > -0.001 && w[3] > -0.001;
            real score = 0;
            for (int k = 0; k < 4; k++) score += (100.0 * nodes[t[k]].pos - p).LengthSquared();
            if (score < best) { best = score; elem = (int)e; }
            if (inside && score < bestPos) { bestPos = score; elemPos = (int)e; }
        }
        if (elemPos >= 0) elem = elemPos;
        const uint32* t = tets[elem].idx;
        const Vector3R x0 = scale * nodes[t[0]].pos, x1 = scale * nodes[t[1]].pos, x2 = scale * nodes[t[2]].pos,
                       x3 = scale * nodes[t[3]].pos;
        const real vol = tet_volume(x0, x1, x2, x3);
        const real w[4] = {tet_volume(p, x1, x2, x3) / vol, tet_volume(x0, p, x2, x3) / vol, tet_volume(x0, x1, p, x3) / vol,
                           tet_volume(x0, x1, x2, p) / vol};
        const bool outside = w[0] < -0.1 || w[1] < -0.1 || w[2] < -0.1 || w[3] < -0.1;
        cable.mCableNodes[i].elem = d.useFreeCable && outside ? -elem : elem;
        cable.mCableNodes[i].bary = Vector3R(w[0], w[1], w[2]);
        cable.mCablePositions[i] = (1 / scale) * p;
    }
    cable.mActuation = 1;
    cable.mCableRestLength = div / scale;
    cable.mCableStiffness = d.stiffness;
    return cable;
}

}  // namespace FEM_SYSTEM
