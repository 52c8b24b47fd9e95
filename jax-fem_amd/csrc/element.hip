// Per-cell element kernels: geometry, tangent, residual, tangent-vector
// product, adjoint parameter gradient, body/surface loads.
//
// Replaces the vmapped per-cell work of jax-fem:
//   fe.py:112-141       get_shape_grads (Jacobian, inverse, physical grad N, JxW)
//   fe.py:143-183       get_face_shape_grads (Nanson scale)
//   fe.py:185-218       physical (surface) quadrature points
//   problem.py:189-259  laplace / mass / surface kernels
//   problem.py:261-356  value_and_jacfwd per cell (here: closed-form tangents)
//   solver.py:1386-1416 vjp of the residual w.r.t. parameters (adjoint)
//
// Layout: a cell is processed by TPC threads (one wave for HEX8, four for
// HEX27); the nodal gather (node ids, coordinates, solution) is staged in LDS,
// then per-quad Jacobians, physical shape gradients and material states are
// built in LDS and consumed by (a,b) block pairs or (a,i) rows.  Geometry is
// computed from coordinates relative to the cell's first node (differences in
// FP64, then cast), so FP32 instances keep a relative accuracy of ~eps32.
#include "element.hpp"
#include "tangent_mfma.hpp"

#include <algorithm>
#include <type_traits>

namespace fem {

// ------------------------------------------------------------------ tangent
// Quadrature points are staged Cfg::Q at a time; each thread keeps the blocks
// of its PPT (a, b) pairs in registers across the chunks.
template <class Cfg, class MAT, typename T>
__global__ __launch_bounds__(256) void k_cell_tangent(fem_mesh_t mesh, fem_material_t mat,
                                                      const double* __restrict__ u, T* __restrict__ Ke) {
  constexpr int N = Cfg::N, Q = Cfg::Q, TPC = Cfg::TPC, CPW = Cfg::CPW, V = MAT::V;
  constexpr int PPT = (N * N + TPC - 1) / TPC;  // (a, b) pairs per thread
  using St = typename MAT::template St<T>;
  __shared__ T s_dN[Q][N][3];
  __shared__ double s_w[Q];
  __shared__ CellLds<N, Q, T> s_cell[CPW];
  __shared__ T s_u[CPW][N][V];
  __shared__ St s_st[CPW][Q];
  __shared__ int32_t s_node[CPW][N];
  const int w = threadIdx.x / TPC, t = threadIdx.x % TPC;
  const int64_t c = (int64_t)blockIdx.x * CPW + w;
  const bool active = c < mesh.num_cells;
  const int Qt = mesh.num_quads;
  CellLds<N, Q, T>& L = s_cell[w];
  T K[PPT][V][V];
#pragma unroll
  for (int p = 0; p < PPT; ++p)
#pragma unroll
    for (int i = 0; i < V; ++i)
#pragma unroll
      for (int k = 0; k < V; ++k) K[p][i][k] = T(0);
  for (int q0 = 0; q0 < Qt; q0 += Q) {
    const int qn = min(Q, Qt - q0);
    __syncthreads();  // the previous chunk's tables and states are no longer read
    load_tables_chunk(mesh, q0, qn, s_dN, s_w);
    __syncthreads();
    stage_geometry<N, Q, TPC, T>(mesh, c, active, t, s_dN, s_w, L, s_node[w], qn);
    if (MAT::kNeedsU && active && t < N) {
#pragma unroll
      for (int i = 0; i < V; ++i) s_u[w][t][i] = T(u[(int64_t)s_node[w][t] * V + i]);
    }
    __syncthreads();
    if (active) {
      for (int q = t; q < qn; q += TPC) {
        T G[V][3];
        if (MAT::kNeedsU) {
          quad_grad<N, Q, V, T>(L, s_u[w], q, G);
        } else {
#pragma unroll
          for (int i = 0; i < V; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) G[i][j] = T(0);
        }
        s_st[w][q] = MAT::template prep<T>(mat, c * Qt + q0 + q, G);
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        const int pr = t + p * TPC;
        if (pr < N * N) {
          const int a = pr / N, b = pr % N;
          for (int q = 0; q < qn; ++q) MAT::tangent(s_st[w][q], L.g[q][a], L.g[q][b], L.jxw[q], K[p]);
        }
      }
    }
  }
  // HEX8: a cell's blocks are staged in LDS and stored lane-contiguously (each store
  // instruction writes 64 consecutive values of the cell's Ke instead of one value
  // of 64 blocks 72 bytes apart)
  constexpr bool kStage = N == 8;
  __shared__ T s_K[kStage ? CPW : 1][kStage ? N * N * V * V : 1];
  if (active) {
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      const int pr = t + p * TPC;
      if (pr < N * N) {
        T* out = kStage ? &s_K[w][pr * V * V] : Ke + ((size_t)c * N * N + pr) * (V * V);
#pragma unroll
        for (int i = 0; i < V; ++i)
#pragma unroll
          for (int k = 0; k < V; ++k) out[i * V + k] = K[p][i][k];
      }
    }
  }
  if (kStage) {
    __syncthreads();
    if (active) {
      T* out = Ke + (size_t)c * N * N * V * V;
      for (int e = t; e < N * N * V * V; e += TPC) out[e] = s_K[w][e];
    }
  }
}

// -------------------------------------------- residual / jvp (FP64 vectors)
// MODE 0: re = sum_q JxW P(grad u) grad N          (internal force)
// MODE 1: re = sum_q JxW dP/dF(grad u)[grad v] grad N   (tangent-vector product)
template <class Cfg, class MAT, typename T, int MODE>
__global__ __launch_bounds__(256) void k_cell_vector(fem_mesh_t mesh, fem_material_t mat,
                                                     const double* __restrict__ u, const double* __restrict__ v,
                                                     double* __restrict__ re) {
  constexpr int N = Cfg::N, Q = Cfg::Q, TPC = Cfg::TPC, CPW = Cfg::CPW, V = MAT::V;
  constexpr int RPT = (N * V + TPC - 1) / TPC;  // (a, i) rows per thread
  __shared__ T s_dN[Q][N][3];
  __shared__ double s_w[Q];
  __shared__ CellLds<N, Q, T> s_cell[CPW];
  __shared__ T s_u[CPW][N][V];
  __shared__ T s_v[CPW][N][V];
  __shared__ T s_P[CPW][Q][V][3];
  __shared__ int32_t s_node[CPW][N];
  const int w = threadIdx.x / TPC, t = threadIdx.x % TPC;
  const int64_t c = (int64_t)blockIdx.x * CPW + w;
  const bool active = c < mesh.num_cells;
  const int Qt = mesh.num_quads;
  CellLds<N, Q, T>& L = s_cell[w];
  T r[RPT];
#pragma unroll
  for (int p = 0; p < RPT; ++p) r[p] = T(0);
  for (int q0 = 0; q0 < Qt; q0 += Q) {
    const int qn = min(Q, Qt - q0);
    __syncthreads();
    load_tables_chunk(mesh, q0, qn, s_dN, s_w);
    __syncthreads();
    stage_geometry<N, Q, TPC, T>(mesh, c, active, t, s_dN, s_w, L, s_node[w], qn);
    if (active && t < N) {
      const int64_t n = s_node[w][t];
#pragma unroll
      for (int i = 0; i < V; ++i) {
        s_u[w][t][i] = T(u[n * V + i]);
        if (MODE == 1) s_v[w][t][i] = T(v[n * V + i]);
      }
    }
    __syncthreads();
    if (active) {
      for (int q = t; q < qn; q += TPC) {
        T G[V][3], P[V][3];
        quad_grad<N, Q, V, T>(L, s_u[w], q, G);
        const auto st = MAT::template prep<T>(mat, c * Qt + q0 + q, G);
        if (MODE == 0) {
          MAT::stress(st, G, P);
        } else {
          T dG[V][3];
          quad_grad<N, Q, V, T>(L, s_v[w], q, dG);
          MAT::jvp(st, dG, P);
        }
        const T jw = L.jxw[q];
#pragma unroll
        for (int i = 0; i < V; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) s_P[w][q][i][j] = P[i][j] * jw;
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int p = 0; p < RPT; ++p) {
        const int it = t + p * TPC;
        if (it < N * V) {
          const int a = it / V, i = it % V;
          for (int q = 0; q < qn; ++q)
            r[p] += s_P[w][q][i][0] * L.g[q][a][0] + s_P[w][q][i][1] * L.g[q][a][1] + s_P[w][q][i][2] * L.g[q][a][2];
        }
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int p = 0; p < RPT; ++p) {
    const int it = t + p * TPC;
    if (it < N * V) re[(size_t)c * N * V + it] = double(r[p]);
  }
}

// HEX8 cells with 8 quadrature points: one thread per (cell, quadrature point),
// 8 consecutive lanes per cell, everything in registers.  Each lane forms its
// point's Jacobian, physical gradients and P (or dP), and the per-node sums
// over the 8 points are lane shuffles (a fixed order).  Same results as
// k_cell_vector up to the summation order.
template <class MAT, int MODE>
__global__ __launch_bounds__(256) void k_cell_vector_q8(fem_mesh_t mesh, fem_material_t mat,
                                                        const double* __restrict__ u, const double* __restrict__ v,
                                                        double* __restrict__ re) {
  constexpr int N = 8, Q = 8, V = MAT::V;
  __shared__ double s_dN[Q][N][3];
  __shared__ double s_w[Q];
  load_tables(mesh, s_dN, s_w);
  __syncthreads();
  const int q = threadIdx.x & 7;
  const int64_t c = (int64_t)blockIdx.x * 32 + (threadIdx.x >> 3);
  const bool active = c < mesh.num_cells;
  const int64_t cc = active ? c : 0;  // idle lanes work on cell 0 and discard it (uniform shuffles)
  int32_t node[N];
#pragma unroll
  for (int a = 0; a < N; ++a) node[a] = mesh.cells[cc * N + a];
  const double* p0 = mesh.points + (int64_t)node[0] * 3;
  const double x00 = p0[0], x01 = p0[1], x02 = p0[2];
  double J[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) J[k] = 0.0;
#pragma unroll
  for (int a = 1; a < N; ++a) {
    const double* pa = mesh.points + (int64_t)node[a] * 3;
    const double d0 = pa[0] - x00, d1 = pa[1] - x01, d2 = pa[2] - x02;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double dn = s_dN[q][a][j];
      J[0 * 3 + j] += d0 * dn;
      J[1 * 3 + j] += d1 * dn;
      J[2 * 3 + j] += d2 * dn;
    }
  }
  const double det = det3(J);
  double iJ[9];
  inv3(J, det, iJ);
  const double jxw = det * s_w[q];
  double g[N][3];
#pragma unroll
  for (int a = 0; a < N; ++a) {
    const double e0 = s_dN[q][a][0], e1 = s_dN[q][a][1], e2 = s_dN[q][a][2];
#pragma unroll
    for (int j = 0; j < 3; ++j) g[a][j] = e0 * iJ[0 * 3 + j] + e1 * iJ[1 * 3 + j] + e2 * iJ[2 * 3 + j];
  }
  double G[V][3], P[V][3];
#pragma unroll
  for (int i = 0; i < V; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) G[i][j] = 0.0;
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const double ua = u[(int64_t)node[a] * V + i];
#pragma unroll
      for (int j = 0; j < 3; ++j) G[i][j] += ua * g[a][j];
    }
  const auto st = MAT::template prep<double>(mat, cc * Q + q, G);
  if (MODE == 0) {
    MAT::stress(st, G, P);
  } else {
    double dG[V][3];
#pragma unroll
    for (int i = 0; i < V; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) dG[i][j] = 0.0;
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int i = 0; i < V; ++i) {
        const double va = v[(int64_t)node[a] * V + i];
#pragma unroll
        for (int j = 0; j < 3; ++j) dG[i][j] += va * g[a][j];
      }
    MAT::jvp(st, dG, P);
  }
#pragma unroll
  for (int i = 0; i < V; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) P[i][j] *= jxw;
#pragma unroll
  for (int a = 0; a < N; ++a) {
    double r[V];
#pragma unroll
    for (int i = 0; i < V; ++i) r[i] = P[i][0] * g[a][0] + P[i][1] * g[a][1] + P[i][2] * g[a][2];
#pragma unroll
    for (int off = 1; off < 8; off <<= 1)
#pragma unroll
      for (int i = 0; i < V; ++i) r[i] += __shfl_xor(r[i], off, 64);
    if (active && q == a) {
#pragma unroll
      for (int i = 0; i < V; ++i) re[((size_t)c * N + a) * V + i] = r[i];
    }
  }
}

// ------------------------------------------------------- quadrature gradients
// G[c][q][i][j] = sum_a u[cells[c][a]][i] dN_a/dx_j (problem.py:200-203): one
// thread per (cell, quad), Jacobian from coordinates relative to the first node.
template <class Cfg, int V>
__global__ __launch_bounds__(256) void k_quad_grad(fem_mesh_t mesh, const double* __restrict__ u,
                                                   double* __restrict__ G, double* __restrict__ U) {
  constexpr int N = Cfg::N, Q = Cfg::Q;
  const int Qt = mesh.num_quads;
  // reference gradients from LDS when all points fit, else straight from the (cached) table
  __shared__ double s_dN[Q][N][3];
  __shared__ double s_w[Q];
  if (Qt <= Q) load_tables_chunk(mesh, 0, Qt, s_dN, s_w);
  __syncthreads();
  const double(*dN)[N][3] = Qt <= Q ? s_dN : reinterpret_cast<const double(*)[N][3]>(mesh.shape_grads_ref);
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= mesh.num_cells * Qt) return;
  const int64_t c = e / Qt;
  const int q = (int)(e - c * Qt);
  const int32_t* cn = mesh.cells + c * N;
  const double* p0 = mesh.points + (int64_t)cn[0] * 3;
  double J[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) J[k] = 0.0;
  for (int a = 1; a < N; ++a) {
    const double* pa = mesh.points + (int64_t)cn[a] * 3;
    const double d0 = pa[0] - p0[0], d1 = pa[1] - p0[1], d2 = pa[2] - p0[2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double dn = dN[q][a][j];
      J[0 * 3 + j] += d0 * dn;
      J[1 * 3 + j] += d1 * dn;
      J[2 * 3 + j] += d2 * dn;
    }
  }
  double iJ[9];
  inv3(J, det3(J), iJ);
  double Gl[V][3], Ul[V];
#pragma unroll
  for (int i = 0; i < V; ++i) {
    Ul[i] = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) Gl[i][j] = 0.0;
  }
  for (int a = 0; a < N; ++a) {
    const double e0 = dN[q][a][0], e1 = dN[q][a][1], e2 = dN[q][a][2];
    const double na = mesh.shape_vals[q * N + a];
    double g[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) g[j] = e0 * iJ[0 * 3 + j] + e1 * iJ[1 * 3 + j] + e2 * iJ[2 * 3 + j];
    const int64_t node = cn[a];
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const double ua = u[node * V + i];
      Ul[i] += ua * na;
#pragma unroll
      for (int j = 0; j < 3; ++j) Gl[i][j] += ua * g[j];
    }
  }
#pragma unroll
  for (int i = 0; i < V; ++i) {
    if (U) U[e * V + i] = Ul[i];
    if (G) {
#pragma unroll
      for (int j = 0; j < 3; ++j) G[(e * V + i) * 3 + j] = Gl[i][j];
    }
  }
}

// -------------------------------------------------- u-dependent mass term
// MODE 0 (tangent): Ke[c][a][b][i][k] += sum_q N_a N_b M[c][q][i][k] JxW   (Ke in TK)
// MODE 1 (jvp):     re[c][a][i] += sum_q N_a JxW sum_k M[c][q][i][k] v_q[k]
template <class Cfg, int V, int MODE, typename TK>
__global__ __launch_bounds__(256) void k_cell_mass(fem_mesh_t mesh, const double* __restrict__ M,
                                                   const double* __restrict__ v, TK* __restrict__ Ke,
                                                   double* __restrict__ re) {
  using T = double;
  constexpr int N = Cfg::N, Q = Cfg::Q, TPC = Cfg::TPC, CPW = Cfg::CPW;
  constexpr int PPT = (N * N + TPC - 1) / TPC, RPT = (N * V + TPC - 1) / TPC;
  __shared__ T s_dN[Q][N][3];
  __shared__ double s_w[Q];
  __shared__ CellLds<N, Q, T> s_cell[CPW];
  __shared__ int32_t s_node[CPW][N];
  __shared__ double s_mv[CPW][Q][V];   // MODE 1: JxW M v_q
  const int w = threadIdx.x / TPC, t = threadIdx.x % TPC;
  const int64_t c = (int64_t)blockIdx.x * CPW + w;
  const bool active = c < mesh.num_cells;
  const int Qt = mesh.num_quads;
  CellLds<N, Q, T>& L = s_cell[w];
  double K[MODE == 0 ? PPT : 1][V][V];
  double r[MODE == 1 ? RPT : 1];
#pragma unroll
  for (int p = 0; p < (MODE == 0 ? PPT : 1); ++p)
#pragma unroll
    for (int i = 0; i < V; ++i)
#pragma unroll
      for (int k = 0; k < V; ++k) K[p][i][k] = 0.0;
#pragma unroll
  for (int p = 0; p < (MODE == 1 ? RPT : 1); ++p) r[p] = 0.0;
  for (int q0 = 0; q0 < Qt; q0 += Q) {
    const int qn = min(Q, Qt - q0);
    __syncthreads();
    load_tables_chunk(mesh, q0, qn, s_dN, s_w);
    __syncthreads();
    stage_geometry<N, Q, TPC, T>(mesh, c, active, t, s_dN, s_w, L, s_node[w], qn);
    const double* sv = mesh.shape_vals + (size_t)q0 * N;
    if (MODE == 1) {
      if (active) {
        for (int q = t; q < qn; q += TPC) {
          double vq[V];
#pragma unroll
          for (int k = 0; k < V; ++k) vq[k] = 0.0;
          for (int b = 0; b < N; ++b) {
            const double nb = sv[q * N + b];
#pragma unroll
            for (int k = 0; k < V; ++k) vq[k] += nb * v[(int64_t)s_node[w][b] * V + k];
          }
          const double* Mq = M + ((size_t)c * Qt + q0 + q) * V * V;
#pragma unroll
          for (int i = 0; i < V; ++i) {
            double rr = 0.0;
#pragma unroll
            for (int k = 0; k < V; ++k) rr += Mq[i * V + k] * vq[k];
            s_mv[w][q][i] = rr * L.jxw[q];
          }
        }
      }
      __syncthreads();
      if (active) {
#pragma unroll
        for (int p = 0; p < RPT; ++p) {
          const int it = t + p * TPC;
          if (it < N * V) {
            const int a = it / V, i = it % V;
            for (int q = 0; q < qn; ++q) r[p] += sv[q * N + a] * s_mv[w][q][i];
          }
        }
      }
    } else if (active) {
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        const int pr = t + p * TPC;
        if (pr < N * N) {
          const int a = pr / N, b = pr % N;
          for (int q = 0; q < qn; ++q) {
            const double s = sv[q * N + a] * sv[q * N + b] * L.jxw[q];
            const double* Mq = M + ((size_t)c * Qt + q0 + q) * V * V;
#pragma unroll
            for (int i = 0; i < V; ++i)
#pragma unroll
              for (int k = 0; k < V; ++k) K[p][i][k] += s * Mq[i * V + k];
          }
        }
      }
    }
  }
  if (!active) return;
  if (MODE == 1) {
#pragma unroll
    for (int p = 0; p < RPT; ++p) {
      const int it = t + p * TPC;
      if (it < N * V) re[(size_t)c * N * V + it] += r[p];
    }
  } else {
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      const int pr = t + p * TPC;
      if (pr < N * N) {
        TK* out = Ke + ((size_t)c * N * N + pr) * (V * V);
#pragma unroll
        for (int i = 0; i < V; ++i)
#pragma unroll
          for (int k = 0; k < V; ++k) out[i * V + k] = TK(double(out[i * V + k]) + K[p][i][k]);
      }
    }
  }
}

// re[c][a][i] += sum_b sum_k K[c][a][b][i][k] v[node_b][k]: one thread per (c, a, i)
__global__ __launch_bounds__(256) void k_cell_matvec(fem_mesh_t mesh, const double* __restrict__ K,
                                                     const double* __restrict__ v, double* __restrict__ re) {
  const int N = mesh.nodes_per_cell, V = mesh.vec;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= mesh.num_cells * N * V) return;
  const int64_t c = e / (N * V);
  const int a = (int)((e / V) % N), i = (int)(e % V);
  const int32_t* cn = mesh.cells + c * N;
  const double* Ka = K + (((size_t)c * N + a) * N) * V * V + (size_t)i * V;
  double r = 0.0;
  for (int b = 0; b < N; ++b) {
    const int64_t nb = cn[b];
    for (int k = 0; k < V; ++k) r += Ka[(size_t)b * V * V + k] * v[nb * V + k];
  }
  re[e] += r;
}

// ------------------------------------------------------------------ adjoint
// g[c][q] = - JxW (dP/dtheta)(grad u) : grad lam   at quad q of cell c.
template <class Cfg, class MAT>
__global__ __launch_bounds__(256) void k_cell_adjoint(fem_mesh_t mesh, fem_material_t mat,
                                                      const double* __restrict__ u, const double* __restrict__ lam,
                                                      double* __restrict__ gout) {
  using T = double;
  constexpr int N = Cfg::N, Q = Cfg::Q, TPC = Cfg::TPC, CPW = Cfg::CPW, V = MAT::V;
  __shared__ T s_dN[Q][N][3];
  __shared__ double s_w[Q];
  __shared__ CellLds<N, Q, T> s_cell[CPW];
  __shared__ T s_u[CPW][N][V];
  __shared__ T s_l[CPW][N][V];
  __shared__ int32_t s_node[CPW][N];
  const int w = threadIdx.x / TPC, t = threadIdx.x % TPC;
  const int64_t c = (int64_t)blockIdx.x * CPW + w;
  const bool active = c < mesh.num_cells;
  const int Qt = mesh.num_quads;
  CellLds<N, Q, T>& L = s_cell[w];
  for (int q0 = 0; q0 < Qt; q0 += Q) {
    const int qn = min(Q, Qt - q0);
    __syncthreads();
    load_tables_chunk(mesh, q0, qn, s_dN, s_w);
    __syncthreads();
    stage_geometry<N, Q, TPC, T>(mesh, c, active, t, s_dN, s_w, L, s_node[w], qn);
    if (active && t < N) {
      const int64_t n = s_node[w][t];
#pragma unroll
      for (int i = 0; i < V; ++i) {
        s_u[w][t][i] = u[n * V + i];
        s_l[w][t][i] = lam[n * V + i];
      }
    }
    __syncthreads();
    if (active) {
      for (int q = t; q < qn; q += TPC) {
        T G[V][3], GL[V][3], dP[V][3];
        quad_grad<N, Q, V, T>(L, s_u[w], q, G);
        quad_grad<N, Q, V, T>(L, s_l[w], q, GL);
        const int64_t cq = c * Qt + q0 + q;
        const auto st = MAT::template prep<T>(mat, cq, G);
        MAT::dparam(mat, cq, st, G, dP);
        T s = 0;
#pragma unroll
        for (int i = 0; i < V; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) s += dP[i][j] * GL[i][j];
        gout[cq] = -s * L.jxw[q];
      }
    }
  }
}

// HEX8 with 8 quadrature points: one thread per (cell, quadrature point),
// registers only (the per-point gradient needs no reduction).
template <class MAT>
__global__ __launch_bounds__(256) void k_cell_adjoint_q8(fem_mesh_t mesh, fem_material_t mat,
                                                         const double* __restrict__ u, const double* __restrict__ lam,
                                                         double* __restrict__ gout) {
  constexpr int N = 8, Q = 8, V = MAT::V;
  __shared__ double s_dN[Q][N][3];
  __shared__ double s_w[Q];
  load_tables(mesh, s_dN, s_w);
  __syncthreads();
  const int q = threadIdx.x & 7;
  const int64_t c = (int64_t)blockIdx.x * 32 + (threadIdx.x >> 3);
  if (c >= mesh.num_cells) return;  // no cross-lane work below
  int32_t node[N];
#pragma unroll
  for (int a = 0; a < N; ++a) node[a] = mesh.cells[c * N + a];
  const double* p0 = mesh.points + (int64_t)node[0] * 3;
  double J[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) J[k] = 0.0;
#pragma unroll
  for (int a = 1; a < N; ++a) {
    const double* pa = mesh.points + (int64_t)node[a] * 3;
    const double d0 = pa[0] - p0[0], d1 = pa[1] - p0[1], d2 = pa[2] - p0[2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double dn = s_dN[q][a][j];
      J[0 * 3 + j] += d0 * dn;
      J[1 * 3 + j] += d1 * dn;
      J[2 * 3 + j] += d2 * dn;
    }
  }
  const double det = det3(J);
  double iJ[9];
  inv3(J, det, iJ);
  double G[V][3], GL[V][3], dP[V][3];
#pragma unroll
  for (int i = 0; i < V; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) G[i][j] = GL[i][j] = 0.0;
#pragma unroll
  for (int a = 0; a < N; ++a) {
    double g[3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
      g[j] = s_dN[q][a][0] * iJ[0 * 3 + j] + s_dN[q][a][1] * iJ[1 * 3 + j] + s_dN[q][a][2] * iJ[2 * 3 + j];
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const double ua = u[(int64_t)node[a] * V + i], la = lam[(int64_t)node[a] * V + i];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        G[i][j] += ua * g[j];
        GL[i][j] += la * g[j];
      }
    }
  }
  const auto st = MAT::template prep<double>(mat, c * Q + q, G);
  MAT::dparam(mat, c * Q + q, st, G, dP);
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < V; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) s += dP[i][j] * GL[i][j];
  gout[c * Q + q] = -s * det * s_w[q];
}

// ------------------------------------------------ geometry and body loads
template <class Cfg>
__global__ __launch_bounds__(256) void k_geometry(fem_mesh_t mesh, double* __restrict__ shape_grads,
                                                  double* __restrict__ JxW, double* __restrict__ qpts) {
  using T = double;
  constexpr int N = Cfg::N, Q = Cfg::Q, TPC = Cfg::TPC, CPW = Cfg::CPW;
  __shared__ T s_dN[Q][N][3];
  __shared__ double s_w[Q];
  __shared__ CellLds<N, Q, T> s_cell[CPW];
  __shared__ int32_t s_node[CPW][N];
  const int w = threadIdx.x / TPC, t = threadIdx.x % TPC;
  const int64_t c = (int64_t)blockIdx.x * CPW + w;
  const bool active = c < mesh.num_cells;
  const int Qt = mesh.num_quads;
  CellLds<N, Q, T>& L = s_cell[w];
  for (int q0 = 0; q0 < Qt; q0 += Q) {
    const int qn = min(Q, Qt - q0);
    __syncthreads();
    load_tables_chunk(mesh, q0, qn, s_dN, s_w);
    __syncthreads();
    stage_geometry<N, Q, TPC, T>(mesh, c, active, t, s_dN, s_w, L, s_node[w], qn);
    if (!active) continue;
    const size_t cq0 = (size_t)c * Qt + q0;
    if (shape_grads)
      for (int it = t; it < qn * N * 3; it += TPC) shape_grads[cq0 * N * 3 + it] = (&L.g[0][0][0])[it];
    if (JxW)
      for (int q = t; q < qn; q += TPC) JxW[cq0 + q] = L.jxw[q];
    if (qpts)
      for (int it = t; it < qn * 3; it += TPC) {
        const int q = it / 3, d = it % 3;
        double x = 0;
        for (int a = 0; a < N; ++a) x += mesh.shape_vals[(size_t)(q0 + q) * N + a] * L.xd[a][d];
        qpts[cq0 * 3 + it] = x;
      }
  }
}

// re[c][a][i] = sum_q m[c][q][i] N_a(q) JxW(c, q); the vec count is runtime (1..3).
template <class Cfg>
__global__ __launch_bounds__(256) void k_cell_load(fem_mesh_t mesh, int V, const double* __restrict__ m,
                                                   double* __restrict__ re) {
  using T = double;
  constexpr int N = Cfg::N, Q = Cfg::Q, TPC = Cfg::TPC, CPW = Cfg::CPW;
  constexpr int RPT = (N * 3 + TPC - 1) / TPC;  // (a, i) rows per thread for vec <= 3
  __shared__ T s_dN[Q][N][3];
  __shared__ double s_w[Q];
  __shared__ CellLds<N, Q, T> s_cell[CPW];
  __shared__ int32_t s_node[CPW][N];
  const int w = threadIdx.x / TPC, t = threadIdx.x % TPC;
  const int64_t c = (int64_t)blockIdx.x * CPW + w;
  const bool active = c < mesh.num_cells;
  const int Qt = mesh.num_quads;
  CellLds<N, Q, T>& L = s_cell[w];
  double r[RPT];
#pragma unroll
  for (int p = 0; p < RPT; ++p) r[p] = 0.0;
  for (int q0 = 0; q0 < Qt; q0 += Q) {
    const int qn = min(Q, Qt - q0);
    __syncthreads();
    load_tables_chunk(mesh, q0, qn, s_dN, s_w);
    __syncthreads();
    stage_geometry<N, Q, TPC, T>(mesh, c, active, t, s_dN, s_w, L, s_node[w], qn);
    if (!active) continue;
#pragma unroll
    for (int p = 0; p < RPT; ++p) {
      const int it = t + p * TPC;
      if (it < N * V) {
        const int a = it / V, i = it % V;
        for (int q = 0; q < qn; ++q)
          r[p] += m[((size_t)c * Qt + q0 + q) * V + i] * mesh.shape_vals[(size_t)(q0 + q) * N + a] * L.jxw[q];
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int p = 0; p < RPT; ++p) {
    const int it = t + p * TPC;
    if (it < N * V) re[(size_t)c * N * V + it] = r[p];
  }
}

// Transpose of k_cell_load: lq[c][q][i] = sum_a lam[node_a][i] N_a(q) JxW(c, q).
template <class Cfg>
__global__ __launch_bounds__(256) void k_cell_load_adjoint(fem_mesh_t mesh, int V, const double* __restrict__ lam,
                                                           double* __restrict__ lq) {
  using T = double;
  constexpr int N = Cfg::N, Q = Cfg::Q, TPC = Cfg::TPC, CPW = Cfg::CPW;
  __shared__ T s_dN[Q][N][3];
  __shared__ double s_w[Q];
  __shared__ CellLds<N, Q, T> s_cell[CPW];
  __shared__ int32_t s_node[CPW][N];
  const int w = threadIdx.x / TPC, t = threadIdx.x % TPC;
  const int64_t c = (int64_t)blockIdx.x * CPW + w;
  const bool active = c < mesh.num_cells;
  const int Qt = mesh.num_quads;
  CellLds<N, Q, T>& L = s_cell[w];
  for (int q0 = 0; q0 < Qt; q0 += Q) {
    const int qn = min(Q, Qt - q0);
    __syncthreads();
    load_tables_chunk(mesh, q0, qn, s_dN, s_w);
    __syncthreads();
    stage_geometry<N, Q, TPC, T>(mesh, c, active, t, s_dN, s_w, L, s_node[w], qn);
    if (!active) continue;
    for (int it = t; it < qn * V; it += TPC) {
      const int q = it / V, i = it % V;
      double r = 0;
      for (int a = 0; a < N; ++a) r += lam[(int64_t)s_node[w][a] * V + i] * mesh.shape_vals[(size_t)(q0 + q) * N + a];
      lq[((size_t)c * Qt + q0 + q) * V + i] = r * L.jxw[q];
    }
  }
}

// --------------------------------------------------------------- faces
// One thread per selected face (boundary sets are small).
template <int N>
__global__ void k_face(fem_mesh_t mesh, int64_t S, const int32_t* __restrict__ faces, double* __restrict__ fgrads,
                       double* __restrict__ nanson, double* __restrict__ fqpts, const double* __restrict__ tvals,
                       double* __restrict__ rf, const double* __restrict__ lam, double* __restrict__ lamq) {
  const int FQ = mesh.num_face_quads;
  const int V = mesh.vec;
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < S; s += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = faces[2 * s];
    const int f = faces[2 * s + 1];
    int32_t node[N];
    for (int a = 0; a < N; ++a) node[a] = mesh.cells[c * N + a];
    const double* x0 = mesh.points + (int64_t)node[0] * 3;
    if (rf)
      for (int k = 0; k < N * V; ++k) rf[s * N * V + k] = 0.0;
    for (int fq = 0; fq < FQ; ++fq) {
      const double* dN = mesh.face_shape_grads_ref + ((size_t)f * FQ + fq) * N * 3;
      const double* Nv = mesh.face_shape_vals + ((size_t)f * FQ + fq) * N;
      double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int a = 0; a < N; ++a) {
        const double* xa = mesh.points + (int64_t)node[a] * 3;
        for (int i = 0; i < 3; ++i) {
          const double xi = xa[i] - x0[i];
          for (int j = 0; j < 3; ++j) J[i * 3 + j] += xi * dN[a * 3 + j];
        }
      }
      const double det = det3(J);
      double iJ[9];
      inv3(J, det, iJ);
      const double* nrm = mesh.face_normals + f * 3;
      double nv[3];
      for (int j = 0; j < 3; ++j) nv[j] = nrm[0] * iJ[0 * 3 + j] + nrm[1] * iJ[1 * 3 + j] + nrm[2] * iJ[2 * 3 + j];
      const double ns = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]) * det * mesh.face_weights[f * FQ + fq];
      if (nanson) nanson[s * FQ + fq] = ns;
      if (fgrads)
        for (int a = 0; a < N; ++a)
          for (int j = 0; j < 3; ++j)
            fgrads[((s * FQ + fq) * N + a) * 3 + j] =
                dN[a * 3 + 0] * iJ[0 * 3 + j] + dN[a * 3 + 1] * iJ[1 * 3 + j] + dN[a * 3 + 2] * iJ[2 * 3 + j];
      if (fqpts)
        for (int d = 0; d < 3; ++d) {
          double x = 0;
          for (int a = 0; a < N; ++a) x += Nv[a] * mesh.points[(int64_t)node[a] * 3 + d];
          fqpts[(s * FQ + fq) * 3 + d] = x;
        }
      if (rf)
        for (int a = 0; a < N; ++a)
          for (int i = 0; i < V; ++i) rf[(s * N + a) * V + i] += tvals[(s * FQ + fq) * V + i] * Nv[a] * ns;
      if (lamq)
        for (int i = 0; i < V; ++i) {
          double acc = 0;
          for (int a = 0; a < N; ++a) acc += lam[(int64_t)node[a] * V + i] * Nv[a];
          lamq[(s * FQ + fq) * V + i] = acc * ns;
        }
    }
  }
}

// ----------------------------------------------------------- dispatching
template <class Cfg>
static int grid_cells(const fem_mesh_t& m) {
  return (int)ceil_div(m.num_cells, Cfg::CPW);
}

// Calls f(Cfg{}) for the supported (nodes per cell, quads) pairs.
template <class F>
static int with_cfg(const fem_mesh_t& m, F&& f) {
  // quadrature points are staged 8 (HEX8, 8 points) or 27 at a time, so any order runs
  const int N = m.nodes_per_cell, Q = m.num_quads;
  if (Q >= 1 && Q <= 4096) {
    if (N == 8 && Q == 8) return f(ElemCfg<8, 8>{});
    if (N == 8) return f(ElemCfg<8, 27>{});
    if (N == 27) return f(ElemCfg<27, 27>{});
  }
  set_error("unsupported element/quadrature combination: %d nodes, %d quadrature points "
            "(supported: HEX8 and HEX27 with 1 to 4096 points)", N, Q);
  return FEM_ERR_BAD_ARG;
}

template <class F>
static int with_mat(const fem_material_t& mat, int vec, F&& f) {
  switch (mat.kind) {
    case FEM_MAT_POISSON:
      if (vec != 1) break;
      return f(MatPoisson{});
    case FEM_MAT_LINEAR_ELASTIC:
      if (vec != 3) break;
      return f(MatElastic<false>{});
    case FEM_MAT_SIMP:
      if (vec != 3 || !mat.has_param) break;
      return f(MatElastic<true>{});
    case FEM_MAT_NEO_HOOKEAN:
      if (vec != 3) break;
      return f(MatNeoHookean{});
    case FEM_MAT_J2_PLASTIC:
      if (vec != 3 || !mat.state) break;
      return f(MatJ2{});
    case FEM_MAT_TABULATED:
      if (!mat.state) break;
      if (vec == 1) return f(MatTabulated<1>{});
      if (vec == 3) return f(MatTabulated<3>{});
      break;
    default:
      break;
  }
  set_error("unsupported material kind %d for vec=%d (param %d, state %s)", mat.kind, vec, mat.has_param,
            mat.state ? "set" : "NULL");
  return FEM_ERR_BAD_ARG;
}

static int check_mesh(const fem_mesh_t* m) {
  FEM_CHECK_ARG(m && m->points && m->cells && m->shape_grads_ref && m->quad_weights && m->shape_vals,
                "mesh descriptor has NULL fields");
  FEM_CHECK_ARG(m->num_cells >= 0 && m->num_nodes > 0, "mesh sizes invalid");
  FEM_CHECK_ARG(m->vec >= 1 && m->vec <= 4, "vec must be 1 to 4 (got %d)", m->vec);
  return FEM_OK;
}

__global__ void k_zero(int64_t n, double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = 0.0;
}

__global__ void k_zero_f(int64_t n, float* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = 0.f;
}

// FEM_MAT_NONE: the law contributes nothing; the output holds n zeros (FP64, or FP32 with f32).
static int zero_output(void* out, int64_t n, bool f32, void* stream) {
  if (n <= 0) return FEM_OK;
  const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(n, 256), 4096);
  if (f32)
    hipLaunchKernelGGL(k_zero_f, dim3(grid), dim3(256), 0, as_stream(stream), n, (float*)out);
  else
    hipLaunchKernelGGL(k_zero, dim3(grid), dim3(256), 0, as_stream(stream), n, (double*)out);
  FEM_LAUNCHED();
  return FEM_OK;
}

}  // namespace fem

using namespace fem;

extern "C" {

int fem_geometry(const fem_mesh_t* mesh, double* shape_grads, double* JxW, double* quad_points, void* stream) {
  if (int rc = check_mesh(mesh)) return rc;
  if (mesh->num_cells == 0) return FEM_OK;
  return with_cfg(*mesh, [&](auto cfg) -> int {
    using Cfg = decltype(cfg);
    hipLaunchKernelGGL(k_geometry<Cfg>, dim3(grid_cells<Cfg>(*mesh)), dim3(256), 0, as_stream(stream), *mesh,
                       shape_grads, JxW, quad_points);
    FEM_LAUNCHED();
    return FEM_OK;
  });
}

int fem_cell_tangent(const fem_mesh_t* mesh, const fem_material_t* mat, int32_t precision, const double* u,
                     void* Ke, void* stream) {
  if (int rc = check_mesh(mesh)) return rc;
  FEM_CHECK_ARG(mat && Ke && u, "fem_cell_tangent: NULL argument");
  FEM_CHECK_ARG(precision == FEM_F32 || precision == FEM_F64 || precision == FEM_BF16X3,
                "fem_cell_tangent: bad precision");
  FEM_CHECK_ARG(!mat->has_param || mat->param, "fem_cell_tangent: material param missing");
  FEM_CHECK_ARG(precision != FEM_BF16X3 || hex27_tangent_mfma_applies(*mesh, *mat),
                "fem_cell_tangent: FEM_BF16X3 needs HEX27 cells, 27 quadrature points and an elastic law");
  if (mesh->num_cells == 0) return FEM_OK;
  if (mat->kind == FEM_MAT_NONE)
    return zero_output(Ke, mesh->num_cells * mesh->nodes_per_cell * mesh->nodes_per_cell * mesh->vec * mesh->vec,
                       precision != FEM_F64, stream);
  if (precision != FEM_F32 && hex27_tangent_mfma_applies(*mesh, *mat))
    return hex27_tangent_mfma(*mesh, *mat, precision, Ke, as_stream(stream));
  return with_cfg(*mesh, [&](auto cfg) -> int {
    using Cfg = decltype(cfg);
    return with_mat(*mat, mesh->vec, [&](auto m) -> int {
      using MAT = decltype(m);
      if (precision == FEM_F32)
        hipLaunchKernelGGL((k_cell_tangent<Cfg, MAT, float>), dim3(grid_cells<Cfg>(*mesh)), dim3(256), 0,
                           as_stream(stream), *mesh, *mat, u, (float*)Ke);
      else
        hipLaunchKernelGGL((k_cell_tangent<Cfg, MAT, double>), dim3(grid_cells<Cfg>(*mesh)), dim3(256), 0,
                           as_stream(stream), *mesh, *mat, u, (double*)Ke);
      FEM_LAUNCHED();
      return FEM_OK;
    });
  });
}

static int cell_vector(const fem_mesh_t* mesh, const fem_material_t* mat, const double* u, const double* v,
                       double* re, int mode, void* stream) {
  if (int rc = check_mesh(mesh)) return rc;
  FEM_CHECK_ARG(mat && u && re && (mode == 0 || v), "cell vector kernel: NULL argument");
  FEM_CHECK_ARG(!mat->has_param || mat->param, "cell vector kernel: material param missing");
  if (mesh->num_cells == 0) return FEM_OK;
  if (mat->kind == FEM_MAT_NONE)
    return zero_output(re, mesh->num_cells * mesh->nodes_per_cell * mesh->vec, false, stream);
  if (mesh->nodes_per_cell == 8 && mesh->num_quads == 8)
    return with_mat(*mat, mesh->vec, [&](auto m) -> int {
      using MAT = decltype(m);
      const dim3 grid((unsigned)ceil_div(mesh->num_cells, 32));
      if (mode == 0)
        hipLaunchKernelGGL((k_cell_vector_q8<MAT, 0>), grid, dim3(256), 0, as_stream(stream), *mesh, *mat, u, v, re);
      else
        hipLaunchKernelGGL((k_cell_vector_q8<MAT, 1>), grid, dim3(256), 0, as_stream(stream), *mesh, *mat, u, v, re);
      FEM_LAUNCHED();
      return FEM_OK;
    });
  return with_cfg(*mesh, [&](auto cfg) -> int {
    using Cfg = decltype(cfg);
    return with_mat(*mat, mesh->vec, [&](auto m) -> int {
      using MAT = decltype(m);
      if (mode == 0)
        hipLaunchKernelGGL((k_cell_vector<Cfg, MAT, double, 0>), dim3(grid_cells<Cfg>(*mesh)), dim3(256), 0,
                           as_stream(stream), *mesh, *mat, u, v, re);
      else
        hipLaunchKernelGGL((k_cell_vector<Cfg, MAT, double, 1>), dim3(grid_cells<Cfg>(*mesh)), dim3(256), 0,
                           as_stream(stream), *mesh, *mat, u, v, re);
      FEM_LAUNCHED();
      return FEM_OK;
    });
  });
}

int fem_quad_grad(const fem_mesh_t* mesh, const double* u, double* G, double* U, void* stream) {
  if (int rc = check_mesh(mesh)) return rc;
  FEM_CHECK_ARG(u && (G || U), "fem_quad_grad: NULL argument");
  FEM_CHECK_ARG(mesh->vec == 1 || mesh->vec == 3, "fem_quad_grad: vec must be 1 or 3 (got %d)", mesh->vec);
  if (mesh->num_cells == 0) return FEM_OK;
  return with_cfg(*mesh, [&](auto cfg) -> int {
    using Cfg = decltype(cfg);
    const dim3 grid((unsigned)ceil_div(mesh->num_cells * mesh->num_quads, 256));
    if (mesh->vec == 1)
      hipLaunchKernelGGL((k_quad_grad<Cfg, 1>), grid, dim3(256), 0, as_stream(stream), *mesh, u, G, U);
    else
      hipLaunchKernelGGL((k_quad_grad<Cfg, 3>), grid, dim3(256), 0, as_stream(stream), *mesh, u, G, U);
    FEM_LAUNCHED();
    return FEM_OK;
  });
}

static int cell_mass(const fem_mesh_t* mesh, const double* M, int mode, int32_t precision, void* Ke,
                     const double* v, double* re, void* stream) {
  if (int rc = check_mesh(mesh)) return rc;
  FEM_CHECK_ARG(M && (mode == 0 ? Ke != nullptr : (v && re)), "mass term kernel: NULL argument");
  FEM_CHECK_ARG(precision == FEM_F32 || precision == FEM_F64, "fem_cell_mass_tangent: bad precision");
  FEM_CHECK_ARG(mesh->vec == 1 || mesh->vec == 3, "mass term kernel: vec must be 1 or 3 (got %d)", mesh->vec);
  if (mesh->num_cells == 0) return FEM_OK;
  return with_cfg(*mesh, [&](auto cfg) -> int {
    using Cfg = decltype(cfg);
    const dim3 grid(grid_cells<Cfg>(*mesh));
    hipStream_t s = as_stream(stream);
    auto go = [&](auto vc) {
      constexpr int V = decltype(vc)::value;
      if (mode == 1)
        hipLaunchKernelGGL((k_cell_mass<Cfg, V, 1, double>), grid, dim3(256), 0, s, *mesh, M, v, (double*)nullptr,
                           re);
      else if (precision == FEM_F32)
        hipLaunchKernelGGL((k_cell_mass<Cfg, V, 0, float>), grid, dim3(256), 0, s, *mesh, M, v, (float*)Ke, re);
      else
        hipLaunchKernelGGL((k_cell_mass<Cfg, V, 0, double>), grid, dim3(256), 0, s, *mesh, M, v, (double*)Ke, re);
    };
    if (mesh->vec == 1)
      go(std::integral_constant<int, 1>{});
    else
      go(std::integral_constant<int, 3>{});
    FEM_LAUNCHED();
    return FEM_OK;
  });
}

int fem_cell_matvec(const fem_mesh_t* mesh, const double* K, const double* v, double* re, void* stream) {
  if (int rc = check_mesh(mesh)) return rc;
  FEM_CHECK_ARG(K && v && re, "fem_cell_matvec: NULL argument");
  const int N = mesh->nodes_per_cell, V = mesh->vec;
  FEM_CHECK_ARG(N == 8 || N == 27, "fem_cell_matvec: %d nodes per cell", N);
  const int64_t rows = mesh->num_cells * N * V;
  if (rows == 0) return FEM_OK;
  hipLaunchKernelGGL(k_cell_matvec, dim3((unsigned)ceil_div(rows, 256)), dim3(256), 0, as_stream(stream), *mesh, K,
                     v, re);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_cell_mass_tangent(const fem_mesh_t* mesh, const double* M, int32_t precision, void* Ke, void* stream) {
  return cell_mass(mesh, M, 0, precision, Ke, nullptr, nullptr, stream);
}

int fem_cell_mass_jvp(const fem_mesh_t* mesh, const double* M, const double* v, double* re, void* stream) {
  return cell_mass(mesh, M, 1, FEM_F64, nullptr, v, re, stream);
}

int fem_cell_residual(const fem_mesh_t* mesh, const fem_material_t* mat, const double* u, double* re,
                      void* stream) {
  return cell_vector(mesh, mat, u, nullptr, re, 0, stream);
}

int fem_cell_jvp(const fem_mesh_t* mesh, const fem_material_t* mat, const double* u, const double* v, double* re,
                 void* stream) {
  return cell_vector(mesh, mat, u, v, re, 1, stream);
}

int fem_cell_adjoint(const fem_mesh_t* mesh, const fem_material_t* mat, const double* u, const double* lam,
                     double* g, void* stream) {
  if (int rc = check_mesh(mesh)) return rc;
  FEM_CHECK_ARG(mat && u && lam && g, "fem_cell_adjoint: NULL argument");
  FEM_CHECK_ARG(mat->has_param && mat->param, "fem_cell_adjoint: the material has no per-quad parameter");
  if (mesh->num_cells == 0) return FEM_OK;
  if (mesh->nodes_per_cell == 8 && mesh->num_quads == 8)
    return with_mat(*mat, mesh->vec, [&](auto m) -> int {
      using MAT = decltype(m);
      hipLaunchKernelGGL((k_cell_adjoint_q8<MAT>), dim3((unsigned)ceil_div(mesh->num_cells, 32)), dim3(256), 0,
                         as_stream(stream), *mesh, *mat, u, lam, g);
      FEM_LAUNCHED();
      return FEM_OK;
    });
  return with_cfg(*mesh, [&](auto cfg) -> int {
    using Cfg = decltype(cfg);
    return with_mat(*mat, mesh->vec, [&](auto m) -> int {
      using MAT = decltype(m);
      hipLaunchKernelGGL((k_cell_adjoint<Cfg, MAT>), dim3(grid_cells<Cfg>(*mesh)), dim3(256), 0, as_stream(stream),
                         *mesh, *mat, u, lam, g);
      FEM_LAUNCHED();
      return FEM_OK;
    });
  });
}

int fem_cell_load(const fem_mesh_t* mesh, const double* m, double* re, void* stream) {
  if (int rc = check_mesh(mesh)) return rc;
  FEM_CHECK_ARG(m && re, "fem_cell_load: NULL argument");
  if (mesh->num_cells == 0) return FEM_OK;
  return with_cfg(*mesh, [&](auto cfg) -> int {
    using Cfg = decltype(cfg);
    hipLaunchKernelGGL(k_cell_load<Cfg>, dim3(grid_cells<Cfg>(*mesh)), dim3(256), 0, as_stream(stream), *mesh,
                       mesh->vec, m, re);
    FEM_LAUNCHED();
    return FEM_OK;
  });
}

static int face_common(const fem_mesh_t* mesh, int64_t S, const int32_t* faces, double* fgrads, double* nanson,
                       double* fqpts, const double* t, double* rf, const double* lam, double* lamq, void* stream) {
  if (int rc = check_mesh(mesh)) return rc;
  FEM_CHECK_ARG(mesh->face_shape_vals && mesh->face_shape_grads_ref && mesh->face_weights && mesh->face_normals &&
                    mesh->num_face_quads > 0,
                "face tables missing from mesh descriptor");
  FEM_CHECK_ARG(S >= 0 && (S == 0 || faces), "face kernel: bad face list");
  if (S == 0) return FEM_OK;
  const int grid = (int)std::min<int64_t>(ceil_div(S, 128), 4096);
  if (mesh->nodes_per_cell == 8)
    hipLaunchKernelGGL(k_face<8>, dim3(grid), dim3(128), 0, as_stream(stream), *mesh, S, faces, fgrads, nanson, fqpts,
                       t, rf, lam, lamq);
  else if (mesh->nodes_per_cell == 27)
    hipLaunchKernelGGL(k_face<27>, dim3(grid), dim3(128), 0, as_stream(stream), *mesh, S, faces, fgrads, nanson,
                       fqpts, t, rf, lam, lamq);
  else
    FEM_CHECK_ARG(false, "face kernel: unsupported nodes per cell %d", mesh->nodes_per_cell);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_face_geometry(const fem_mesh_t* mesh, int64_t S, const int32_t* faces, double* face_shape_grads,
                      double* nanson, double* quad_points, void* stream) {
  return face_common(mesh, S, faces, face_shape_grads, nanson, quad_points, nullptr, nullptr, nullptr, nullptr, stream);
}

int fem_face_load(const fem_mesh_t* mesh, int64_t S, const int32_t* faces, const double* t, double* rf,
                  void* stream) {
  FEM_CHECK_ARG(t && rf, "fem_face_load: NULL argument");
  return face_common(mesh, S, faces, nullptr, nullptr, nullptr, t, rf, nullptr, nullptr, stream);
}

int fem_face_load_adjoint(const fem_mesh_t* mesh, int64_t S, const int32_t* faces, const double* lam, double* lq,
                          void* stream) {
  FEM_CHECK_ARG(lam && lq, "fem_face_load_adjoint: NULL argument");
  return face_common(mesh, S, faces, nullptr, nullptr, nullptr, nullptr, nullptr, lam, lq, stream);
}

int fem_cell_load_adjoint(const fem_mesh_t* mesh, const double* lam, double* lq, void* stream) {
  if (int rc = check_mesh(mesh)) return rc;
  FEM_CHECK_ARG(lam && lq, "fem_cell_load_adjoint: NULL argument");
  if (mesh->num_cells == 0) return FEM_OK;
  return with_cfg(*mesh, [&](auto cfg) -> int {
    using Cfg = decltype(cfg);
    hipLaunchKernelGGL(k_cell_load_adjoint<Cfg>, dim3(grid_cells<Cfg>(*mesh)), dim3(256), 0, as_stream(stream),
                       *mesh, mesh->vec, lam, lq);
    FEM_LAUNCHED();
    return FEM_OK;
  });
}

}  // extern "C"
