// FemIO.cpp -- .tet reading and legacy-VTK writing (see FemIO.h).
#include "FemIO.h"

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

namespace FEM_SYSTEM {
namespace IO {

bool LoadFromTet1File(const char* path, real scale, const Vector3R& offset, std::vector<Node>& nodes,
                      std::vector<Tet>& tets)
{
    std::ifstream in(path);
    if (!in) {
        std::printf("Couldn't open file %s\n", path);
        return false;
    }
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::istringstream ss(line.substr(1));
        if (line[0] == 'v') {
            real x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            Node n;
            n.pos = Vector3R(scale * x + offset[0], scale * z + offset[1], scale * y + offset[2]);   // y <-> z
            n.pos0 = n.pos;
            nodes.push_back(n);
        } else if (line[0] == 't') {
            Tet t;
            for (int k = 0; k < 4; k++) ss >> t.idx[k];
            tets.push_back(t);
        }
    }
    return true;
}

void ExportToVTKFile(std::ostream& out, const FemPhysicsMatrixFree& phys)
{
    const uint32 nn = phys.GetNumNodes(), ne = phys.GetNumElements();
    out << "# vtk DataFile Version 2.0\nUnstructured Grid\nASCII\nDATASET UNSTRUCTURED_GRID\n";
    out.precision(17);
    out << "POINTS " << nn << " double\n";
    for (uint32 i = 0; i < nn; i++) {
        const Vector3R v = phys.GetDeformedPosition(i);
        out << v[0] << " " << v[1] << " " << v[2] << "\n";
    }
    out << "\nCELLS " << ne << " " << 5 * ne << "\n";
    for (uint32 e = 0; e < ne; e++) {
        out << 4;
        for (uint32 k = 0; k < 4; k++) out << " " << phys.GetGlobalOriginalIndex(e, k);
        out << "\n";
    }
    out << "\nCELL_TYPES " << ne << "\n";
    for (uint32 e = 0; e < ne; e++) out << 10 << "\n";   // VTK_TETRA
    out << "\n";
}

void ExportToVTKHeatMap(std::ostream& out, const FemPhysicsMatrixFree& phys)
{
    ExportToVTKFile(out, phys);
    const uint32 nn = phys.GetNumNodes(), ne = phys.GetNumElements();
    out.precision(9);
    out << "POINT_DATA " << nn << "\nVECTORS displacement float\n";
    for (uint32 i = 0; i < nn; i++) {
        const Vector3R u = phys.GetDeformedPosition(i) - phys.GetInitialPosition(i);
        out << u[0] << " " << u[1] << " " << u[2] << "\n";
    }
    std::vector<Matrix3R> P;
    phys.GetElementStresses(P);
    out << "\nCELL_DATA " << ne << "\nTENSORS stress float\n";
    for (uint32 e = 0; e < ne; e++)
        for (int i = 0; i < 3; i++) out << P[e](i, 0) << " " << P[e](i, 1) << " " << P[e](i, 2) << "\n";
    out << "\n";
}

}  // namespace IO
}  // namespace FEM_SYSTEM
