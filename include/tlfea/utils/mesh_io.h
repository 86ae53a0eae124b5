// tlfea/utils/mesh_io.h -- mesh readers/generators of the drop-in API.
// FEAT10_remap_tetgen_indices / FEAT10_read_nodes / FEAT10_read_elements keep the names,
// arguments and adaptive 0/1 base of the reference's TetGen reader
// (src/utils/cpu_utils.cc:596-735); LoadVtuTet stands in for mophi::LoadVtu
// (examples/beam_linear_static.cc:69-99) for ASCII VTU files of TET10 (24) or TET4 (10).
#pragma once

#include <string>

#include "../types.h"

namespace tlfea {

void FEAT10_remap_tetgen_indices(const VectorXi& tetgen_elem, VectorXi& standard_elem);
int FEAT10_read_nodes(const std::string& filename, MatrixXR& nodes);
int FEAT10_read_elements(const std::string& filename, MatrixXi& elements);

// Returns the number of elements (0 on failure); nodes is n x 3, elements n_elem x nn.
int LoadVtuTet(const std::string& filename, MatrixXR& nodes, MatrixXi& elements);

// Structured Kuhn beam [0,lx]x[0,ly]x[0,lz] with nx*ny*nz hex cells split into 6
// positively oriented tets each (TET10 on the 2x lattice when tet10).  Node id
// (I*NY + J)*NZ + K, elements cell by cell with x outermost.
void GenerateKuhnBeam(int nx, int ny, int nz, Real lx, Real ly, Real lz, bool tet10, MatrixXR& nodes,
                      MatrixXi& elements);

// The reference declares its mesh utilities in tlfea::ANCFCPUUtils (cpu_utils.h:8-163,
// mesh_utils.h:242-275); its callers write ANCFCPUUtils::FEAT10_read_nodes(...).
namespace ANCFCPUUtils {
using tlfea::FEAT10_read_elements;
using tlfea::FEAT10_read_nodes;
using tlfea::FEAT10_remap_tetgen_indices;

// Legacy ASCII VTK (POINTS at x, y, z; CELLS; CELL_TYPES 24 or 10; POINT_DATA VECTORS
// displacement = (x, y, z) - nodes), the layout of mesh_utils.cc:962-1077.  Returns false
// when the file cannot be opened or the sizes disagree.
bool WriteFEAT10ToVTK(const std::string& filename, const MatrixXR& nodes, const MatrixXi& elements, const VectorXR& x,
                      const VectorXR& y, const VectorXR& z);
bool WriteFEAT4ToVTK(const std::string& filename, const MatrixXR& nodes, const MatrixXi& elements, const VectorXR& x,
                     const VectorXR& y, const VectorXR& z);
}  // namespace ANCFCPUUtils

}  // namespace tlfea
