// Per-cell staging shared by the element kernels: configuration, LDS cell
// state, geometry staging (Jacobian, inverse, physical gradients, JxW) and
// the reference tables.
#pragma once
#include "materials.hpp"

namespace fem {

// Q is the number of quadrature points staged at a time: kernels walk the
// mesh's num_quads points in chunks of Q (one chunk when they are equal), so
// any quadrature order runs with the LDS footprint of Q points.
template <int N_, int Q_>
struct ElemCfg {
  static constexpr int N = N_;
  static constexpr int Q = Q_;
  static constexpr int TPC = (N_ * N_ <= 64) ? 64 : 256;  // threads per cell
  static constexpr int CPW = 256 / TPC;                   // cells per workgroup
};

// LDS state of one cell.
template <int N, int Q, typename T>
struct CellLds {
  double xd[N][3];
  T invJ[Q][9];
  T jxw[Q];
  T g[Q][N][3];
};

// Stage node coordinates, Jacobians and physical gradients of cell c for the
// first qn (default Q) staged quadrature points.
// Must be called by all threads of the workgroup (contains barriers).
template <int N, int Q, int TPC, typename T>
__device__ inline void stage_geometry(const fem_mesh_t& mesh, int64_t c, bool active, int t,
                                      const T (&dN)[Q][N][3], const double (&w)[Q], CellLds<N, Q, T>& L,
                                      int32_t (&node)[N], int qn = Q) {
  if (active && t < N) {
    const int32_t n = mesh.cells[c * N + t];
    node[t] = n;
    L.xd[t][0] = mesh.points[(int64_t)n * 3 + 0];
    L.xd[t][1] = mesh.points[(int64_t)n * 3 + 1];
    L.xd[t][2] = mesh.points[(int64_t)n * 3 + 2];
  }
  __syncthreads();
  if (active) {
    for (int q = t; q < qn; q += TPC) {
      T J[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) J[k] = T(0);
      for (int a = 0; a < N; ++a) {
        const T x0 = T(L.xd[a][0] - L.xd[0][0]);
        const T x1 = T(L.xd[a][1] - L.xd[0][1]);
        const T x2 = T(L.xd[a][2] - L.xd[0][2]);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const T d = dN[q][a][j];
          J[0 * 3 + j] += x0 * d;
          J[1 * 3 + j] += x1 * d;
          J[2 * 3 + j] += x2 * d;
        }
      }
      const T det = det3(J);
      inv3(J, det, L.invJ[q]);
      L.jxw[q] = det * T(w[q]);
    }
  }
  __syncthreads();
  if (active) {
    for (int it = t; it < qn * N; it += TPC) {
      const int q = it / N, a = it % N;
      const T* iJ = L.invJ[q];
      const T d0 = dN[q][a][0], d1 = dN[q][a][1], d2 = dN[q][a][2];
#pragma unroll
      for (int j = 0; j < 3; ++j) L.g[q][a][j] = d0 * iJ[0 * 3 + j] + d1 * iJ[1 * 3 + j] + d2 * iJ[2 * 3 + j];
    }
  }
  __syncthreads();
}

template <int N, int Q, typename T>
__device__ inline void load_tables(const fem_mesh_t& mesh, T (&dN)[Q][N][3], double (&w)[Q]) {
  T* flat = &dN[0][0][0];
  for (int i = threadIdx.x; i < Q * N * 3; i += blockDim.x) flat[i] = T(mesh.shape_grads_ref[i]);
  for (int i = threadIdx.x; i < Q; i += blockDim.x) w[i] = mesh.quad_weights[i];
}

// Reference gradients and weights of quadrature points q0 .. q0 + qn - 1 (qn <= Q).
template <int N, int Q, typename T>
__device__ inline void load_tables_chunk(const fem_mesh_t& mesh, int q0, int qn, T (&dN)[Q][N][3], double (&w)[Q]) {
  T* flat = &dN[0][0][0];
  const double* src = mesh.shape_grads_ref + (size_t)q0 * N * 3;
  for (int i = threadIdx.x; i < qn * N * 3; i += blockDim.x) flat[i] = T(src[i]);
  for (int i = threadIdx.x; i < qn; i += blockDim.x) w[i] = mesh.quad_weights[q0 + i];
}

// grad f at quad q from nodal values f[N][V] (LDS)
template <int N, int Q, int V, typename T>
__device__ inline void quad_grad(const CellLds<N, Q, T>& L, const T (&f)[N][V], int q, T (&G)[V][3]) {
#pragma unroll
  for (int i = 0; i < V; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) G[i][j] = T(0);
  for (int a = 0; a < N; ++a) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const T fa = f[a][i];
#pragma unroll
      for (int j = 0; j < 3; ++j) G[i][j] += fa * L.g[q][a][j];
    }
  }
}

}  // namespace fem
