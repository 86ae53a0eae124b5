// tlfea/elements/FEAT10Data.h -- 10-node tetrahedron (reference: src/elements/FEAT10Data.cuh).
#pragma once

#include "../utils/quadrature_utils.h"
#include "TetElementData.h"

namespace tlfea {

struct GPU_FEAT10_Data : public detail::TetElementData {
    static constexpr int N_NODES_PER_ELEM = Quadrature::N_NODE_T10_10;
    static constexpr int N_QP_PER_ELEM = Quadrature::N_QP_T10_5;
    GPU_FEAT10_Data(int num_elements, int num_nodes) : TetElementData(10, num_elements, num_nodes) {
        type = TYPE_T10;
        d_data = this;
    }
};

}  // namespace tlfea
