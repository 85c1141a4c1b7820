// FemBody.h -- the host-side mesh tools that the reference's bindings build
// around the physics (Include/FemBody.h, Source/FemBody.cpp, Source/MeshTools.cpp):
// the boundary triangle mesh of a tetrahedral mesh, an OBJ visual mesh mapped
// onto it, OBJ export, and cable registration (CreateCable).
//
// Unlike FemBody::BuildBoundaryMesh (FemBody.cpp:263,310-313), every node
// index here is 32-bit: meshes above 65,535 nodes keep their boundary.
#pragma once

#include <cstdio>
#include <string>
#include <vector>

#include "FemPhysicsMatrixFree.h"

namespace FEM_SYSTEM {

// Mesh (Include/MeshTools.h:40-66): float vertices, normals, triangle indices
struct Mesh {
    std::vector<Vector3T<float>> vertices, normals;
    std::vector<uint32> indices;
    size_t GetNumTriangles() const { return indices.size() / 3; }
    // AddTriangle (MeshTools.h:51-62): b and c swapped when flip
    void AddTriangle(uint32 a, uint32 b, uint32 c, bool flip = false)
    {
        indices.push_back(a);
        indices.push_back(flip ? c : b);
        indices.push_back(flip ? b : c);
    }
    // ComputeNormals (MeshTools.cpp:279-318): per vertex the normalized sum of
    // the unnormalized normals of its triangles
    void ComputeNormals();
};

// CableDescriptor (FemDataStructures.h:196-204) as the <cable> element sets it
struct CableDescriptor {
    int divs = 10;
    real length = 1;
    Vector3R offset;
    Vector3R dir = Vector3R(1, 0, 0);
    bool useFreeCable = true;
    real stiffness = 1000;
};

class FemBody {
public:
    // BuildBoundaryMesh (FemBody.cpp:244-359): the faces that belong to one
    // tetrahedron only, sorted by their vertex triple, wound outward (away
    // from the tetrahedron's fourth node); the vertex buffer holds the
    // boundary nodes in increasing node order
    void BuildBoundaryMesh(const std::vector<Tet>& tets, const std::vector<Node>& nodes);
    // UpdateBoundaryMesh (FemBody.cpp:361-373): vertices at the physics' positions
    void UpdateBoundaryMesh(const FemPhysicsMatrixFree& phys);
    // LoadVisualMesh + MapVisualMesh (FemBody.cpp:157-216): an OBJ mesh, each
    // vertex tied to the closest boundary triangle (barycentric coordinates of
    // the closest point and the distance along the triangle's normal)
    bool LoadVisualMesh(const char* path, float scale);
    // UpdateVisualMesh (FemBody.cpp:218-239)
    void UpdateVisualMesh();
    bool HasVisualMesh() const { return !mVisualMesh.vertices.empty(); }
    const Mesh& GetBoundaryMesh() const { return mBoundaryMesh; }
    const Mesh& GetVisualMesh() const { return mVisualMesh; }
    const std::vector<int>& GetBoundaryNodes() const { return mBoundaryToNodesMap; }

private:
    struct MeshInterp {
        uint32 id1, id2, id3;
        float u, v, w, distance;
    };
    Mesh mBoundaryMesh, mVisualMesh;
    std::vector<int> mNodesToBoundaryMap, mBoundaryToNodesMap;
    std::vector<MeshInterp> mMapData;
};

// ExportMeshToOBJ (MeshTools.cpp:216-227): "vn" and "v" per vertex, "f a//a b//b c//c"
void ExportMeshToOBJ(FILE* out, const Mesh& mesh);
// ImportMeshFromOBJ (MeshTools.cpp:229-277): v and f records (the first index
// of each face corner), positions scaled
bool ImportMeshFromOBJ(const char* path, Mesh& mesh, float scale);

// CreateCable (FemBody.cpp:31-104): descriptor.divs + 1 points along the
// cable, each registered to the element of smallest score sum |100 X_j - p|^2
// among those whose barycentric weights all exceed -0.001 (else among all);
// a point further outside its element than 0.1 in a weight is a free cable
// node when descriptor.useFreeCable (elem = -element).
Cable CreateCable(const CableDescriptor& descriptor, const std::vector<Node>& nodes, const std::vector<Tet>& tets,
                  real scale = 1);

}  // namespace FEM_SYSTEM
