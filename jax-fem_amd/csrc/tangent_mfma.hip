// HEX27 tangents of the isotropic elastic laws on the matrix cores.
//
// Replaces the per-cell jacfwd of jax-fem (problem.py:262-266) for
// LINEAR_ELASTIC and SIMP materials on 27-node cells with 27 quadrature
// points.  With X[q][(a,i)] = dN_a/dx_i at quad q (27 x 81) and
// W_q = JxW_q E_q, all of the 81 x 81 cell tangent follows from one GEMM
//   Z = X^T diag(W) X                                  (81 x 81, inner dim 27)
// because lambda = c0 E and mu = c1 E share the modulus E:
//   K[(a,i),(b,k)] = c0 Z[(a,i),(b,k)] + c1 Z[(a,k),(b,i)]
//                  + delta_ik c1 (Z[(a,0),(b,0)] + Z[(a,1),(b,1)] + Z[(a,2),(b,2)]).
// Z is formed on MFMA in 16 x 16 tiles (81 padded to 96, 36 tiles, 9 per
// wave), staged in LDS and expanded into Ke[c][a][b][i][k] with coalesced
// stores.  Operand layouts: A[i][k] in lane i + 16 k (f64; for bf16 the four
// k of one lane are k = 4 (lane / 16) + e), B[k][j] in lane j + 16 k; the two
// instructions place the 16 x 16 result differently (see the stores).  Two arithmetic modes:
//   FP64 : v_mfma_f64_16x16x4f64, exact FP64 products and sums.
//   BF16x3: each operand split x = hi + lo (hi = bf16(x), lo = bf16(x - hi));
//          hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x16bf16_1k with FP32
//          accumulation (about 16 significant bits per product); Ke in FP32.
// One workgroup of 4 waves per cell, persistent over cells (LDS ~ 91 KB).
#include "element.hpp"
#include "tangent_mfma.hpp"

#include <algorithm>
#include <type_traits>

namespace fem {

namespace {

constexpr int kN = 27, kQ = 27, kM = 81, kTiles = 6;  // 81 padded to 6 tiles of 16
constexpr int kGrid = 1024;                            // persistent workgroups

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef short short4_t __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 (finite inputs)
__device__ inline unsigned short bf16_bits(float x) {
  unsigned int b = __float_as_uint(x);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return (unsigned short)(b >> 16);
}
__device__ inline float bf16_value(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

struct Hex27Lds {
  double dN[kQ][kN][3];
  double w[kQ];
  double W[kQ];
  CellLds<kN, kQ, double> cell;
  int32_t node[kN];
};

}  // namespace

template <bool kSimp, bool kBf16, typename TO>
__global__ __launch_bounds__(256) void k_tangent_hex27_mfma(fem_mesh_t mesh, fem_material_t mat,
                                                            TO* __restrict__ Ke) {
  using TZ = typename std::conditional<kBf16, float, double>::type;
  __shared__ Hex27Lds S;
  __shared__ TZ Z[kM][kM + 1];
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63;
  load_tables(mesh, S.dN, S.w);
  __syncthreads();
  const double c0 = mat.c[0], c1 = mat.c[1];
  for (int64_t c = blockIdx.x; c < mesh.num_cells; c += gridDim.x) {
    stage_geometry<kN, kQ, 256, double>(mesh, c, true, t, S.dN, S.w, S.cell, S.node);
    if (t < kQ) {
      const double E = kSimp ? MatElastic<true>::modulus(mat, mat.param[c * kQ + t]) : 1.0;
      S.W[t] = S.cell.jxw[t] * E;
    }
    __syncthreads();
    const double(*g)[kN][3] = S.cell.g;
    for (int tile = wave; tile < kTiles * kTiles; tile += 4) {
      const int tm = tile / kTiles, tn = tile % kTiles;
      const int m = tm * 16 + (lane & 15);      // A row (operand lane layout)
      const int n = tn * 16 + (lane & 15);      // B column
      if (!kBf16) {
        double4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 7; ++kk) {          // inner dim 27 padded to 28
          const int q = kk * 4 + (lane >> 4);
          const double a = (q < kQ && m < kM) ? S.W[q] * g[q][m / 3][m % 3] : 0.0;
          const double b = (q < kQ && n < kM) ? g[q][n / 3][n % 3] : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
        // f64 16x16x4 result: D[i][j] in lane j + 16 (i % 4), register i / 4
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = tm * 16 + (lane >> 4) + 4 * r, col = tn * 16 + (lane & 15);
          if (row < kM && col < kM) Z[row][col] = TZ(acc[r]);
        }
      } else {
        float4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {          // inner dim 27 padded to 32
          short4_t ahi, alo, bhi, blo;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int q = kk * 16 + 4 * (lane >> 4) + e;
            const float a = (q < kQ && m < kM) ? float(S.W[q] * g[q][m / 3][m % 3]) : 0.f;
            const float b = (q < kQ && n < kM) ? float(g[q][n / 3][n % 3]) : 0.f;
            const unsigned short ah = bf16_bits(a), bh = bf16_bits(b);
            ahi[e] = (short)ah;
            alo[e] = (short)bf16_bits(a - bf16_value(ah));
            bhi[e] = (short)bh;
            blo[e] = (short)bf16_bits(b - bf16_value(bh));
          }
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, bhi, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, blo, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, bhi, acc, 0, 0, 0);
        }
        // bf16 16x16x16 result: D[i][j] in lane j + 16 (i / 4), register i % 4
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = tm * 16 + 4 * (lane >> 4) + r, col = tn * 16 + (lane & 15);
          if (row < kM && col < kM) Z[row][col] = TZ(acc[r]);
        }
      }
    }
    __syncthreads();
    TO* out = Ke + (size_t)c * kM * kM;
    for (int e = t; e < kM * kM; e += 256) {
      const int k = e % 3, i = (e / 3) % 3, ab = e / 9;
      const int a = ab / kN, b = ab % kN;
      double v = c0 * double(Z[3 * a + i][3 * b + k]) + c1 * double(Z[3 * a + k][3 * b + i]);
      if (i == k)
        v += c1 * (double(Z[3 * a][3 * b]) + double(Z[3 * a + 1][3 * b + 1]) + double(Z[3 * a + 2][3 * b + 2]));
      out[e] = TO(v);
    }
    __syncthreads();
  }
}

bool hex27_tangent_mfma_applies(const fem_mesh_t& mesh, const fem_material_t& mat) {
  return mesh.nodes_per_cell == kN && mesh.num_quads == kQ && mesh.vec == 3 &&
         (mat.kind == FEM_MAT_LINEAR_ELASTIC || (mat.kind == FEM_MAT_SIMP && mat.has_param));
}

int hex27_tangent_mfma(const fem_mesh_t& mesh, const fem_material_t& mat, int precision, void* Ke, hipStream_t s) {
  const unsigned grid = (unsigned)std::min<int64_t>(mesh.num_cells, kGrid);
  const bool simp = mat.kind == FEM_MAT_SIMP;
  if (precision == FEM_F64) {
    if (simp)
      hipLaunchKernelGGL((k_tangent_hex27_mfma<true, false, double>), dim3(grid), dim3(256), 0, s, mesh, mat,
                         (double*)Ke);
    else
      hipLaunchKernelGGL((k_tangent_hex27_mfma<false, false, double>), dim3(grid), dim3(256), 0, s, mesh, mat,
                         (double*)Ke);
  } else {
    if (simp)
      hipLaunchKernelGGL((k_tangent_hex27_mfma<true, true, float>), dim3(grid), dim3(256), 0, s, mesh, mat,
                         (float*)Ke);
    else
      hipLaunchKernelGGL((k_tangent_hex27_mfma<false, true, float>), dim3(grid), dim3(256), 0, s, mesh, mat,
                         (float*)Ke);
  }
  FEM_LAUNCHED();
  return FEM_OK;
}

}  // namespace fem
