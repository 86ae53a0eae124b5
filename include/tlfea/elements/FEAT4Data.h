// tlfea/elements/FEAT4Data.h -- 4-node tetrahedron (reference: src/elements/FEAT4Data.cuh).
#pragma once

#include "../utils/quadrature_utils.h"
#include "TetElementData.h"

namespace tlfea {

struct GPU_FEAT4_Data : public detail::TetElementData {
    static constexpr int N_NODES_PER_ELEM = Quadrature::N_NODE_T4_4;
    static constexpr int N_QP_PER_ELEM = Quadrature::N_QP_T4_1;
    GPU_FEAT4_Data(int num_elements, int num_nodes) : TetElementData(4, num_elements, num_nodes) {
        type = TYPE_T4;
        d_data = this;
    }
};

}  // namespace tlfea
