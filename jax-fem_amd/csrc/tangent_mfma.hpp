// HEX27 elastic tangents on MFMA (tangent_mfma.hip).
#pragma once
#include "fem_common.hpp"

namespace fem {

// True for 27-node cells with 27 quadrature points and a LINEAR_ELASTIC or SIMP law.
bool hex27_tangent_mfma_applies(const fem_mesh_t& mesh, const fem_material_t& mat);
// precision FEM_F64: FP64 MFMA, Ke double; FEM_BF16X3: split-BF16 MFMA, Ke float.
int hex27_tangent_mfma(const fem_mesh_t& mesh, const fem_material_t& mat, int precision, void* Ke, hipStream_t s);

}  // namespace fem
