// FemIO.h -- the file formats of Source/FemIO.cpp that the matrix-free path
// uses: .tet meshes in and legacy-VTK heat maps out.
#pragma once
#include <ostream>
#include <vector>

#include "FemPhysicsMatrixFree.h"

namespace FEM_SYSTEM {
namespace IO {

// IO::LoadFromTet1File (FemIO.cpp:506-581): 'v x y z' and 't a b c d' lines;
// y and z are swapped on read, positions become scale * pos + offset
bool LoadFromTet1File(const char* path, real scale, const Vector3R& offset, std::vector<Node>& nodes,
                      std::vector<Tet>& tets);

// IO::ExportToVTKFile (FemIO.cpp:351-399): deformed points, VTK_TETRA cells
void ExportToVTKFile(std::ostream& out, const FemPhysicsMatrixFree& phys);

// IO::ExportToVTKHeatMap (FemIO.cpp:317-349): the above plus POINT_DATA
// displacements and CELL_DATA first Piola-Kirchhoff stress tensors (the
// element stresses come from the device, sfem_element_stress)
void ExportToVTKHeatMap(std::ostream& out, const FemPhysicsMatrixFree& phys);

}  // namespace IO
}  // namespace FEM_SYSTEM
