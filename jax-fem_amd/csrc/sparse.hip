// Global matrix and vector kernels over the SELL-64 block layout.
//
// Replaces jax-fem's host-side global assembly and BC handling:
//   problem.py:426-437  scatter-add of cell residuals into nodal residuals
//   solver.py:522-529   PETSc setValuesCOO (sum duplicates) + zeroRows(diag=1)
//   solver.py:290-304   apply_bc_vec (residual rows of Dirichlet dofs)
//   solver.py:65-68     CSR export / Jacobi diagonal of jax_solve
// All sums run in a fixed order (cells ascending), so results are
// bit-reproducible run to run.
#include "fem_common.hpp"

#include <algorithm>

namespace fem {

// ------------------------------------------------------- nodal gather
// V is a template parameter so the per-node sums stay in registers; the
// incident entries are visited two at a time to keep two gathers in flight.
template <int V>
__global__ __launch_bounds__(256) void k_gather_nodal(int64_t num_nodes, const int32_t* __restrict__ inc_ptr,
                                                      const int32_t* __restrict__ inc_idx,
                                                      const double* __restrict__ ent, const double* base,
                                                      double* out) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < num_nodes;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int b = inc_ptr[n], e = inc_ptr[n + 1];
    double acc[V], acc2[V];
#pragma unroll
    for (int i = 0; i < V; ++i) acc[i] = acc2[i] = 0.0;
    int t = b;
    for (; t + 1 < e; t += 2) {
      const double* p0 = ent + (int64_t)inc_idx[t] * V;
      const double* p1 = ent + (int64_t)inc_idx[t + 1] * V;
#pragma unroll
      for (int i = 0; i < V; ++i) {
        acc[i] += p0[i];
        acc2[i] += p1[i];
      }
    }
    if (t < e) {
      const double* p0 = ent + (int64_t)inc_idx[t] * V;
#pragma unroll
      for (int i = 0; i < V; ++i) acc[i] += p0[i];
    }
#pragma unroll
    for (int i = 0; i < V; ++i) out[n * V + i] = (base ? base[n * V + i] : 0.0) + (acc[i] + acc2[i]);
  }
}

// ------------------------------------------------------- matrix values
// One thread per slice lane: for every SELL slot k, sum the cell blocks that land in
// block (row, col_k) and store them (zeros in padding slots).  DIFF stores
// the row-sum block R_row = sum_k A(row, k) (accumulated in TS) in the
// diagonal slot instead of the diagonal block (difference form of the SpMV).
// dinv (may be NULL) receives 1 / A(row*V+c, row*V+c).  The slot bounds and
// the cell-block ids are read through windows (a lane's lists are contiguous).
template <int V, bool DIFF, typename TS, typename T>
__global__ __launch_bounds__(256) void k_gather_matrix(fem_pattern_t pat, const TS* __restrict__ Ke,
                                                       T* __restrict__ vals, T* __restrict__ dinv) {
  constexpr int VV = V * V;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= pat.num_slices * kSlice) return;
  const int64_t s = t >> 6;
  const int lane = (int)(t & 63);
  const int base = pat.slice_ptr[s];
  const int L = (pat.slice_ptr[s + 1] - base) >> 6;
  const int row = pat.sell_row[t];
  const bool real = row >= 0;
  const int rb = real ? pat.row_ptr[row] : 0;
  const int len = real ? pat.row_ptr[row + 1] - rb : 0;
  const int ds = real ? pat.diag_slot[row] : -1;
  T* vp = vals + (size_t)base * VV + lane;
  TS rsum[VV];
  TS diag[V];
#pragma unroll
  for (int v = 0; v < VV; ++v) rsum[v] = TS(0);
#pragma unroll
  for (int c = 0; c < V; ++c) diag[c] = TS(0);
  ListWindow<4> cptr(pat.contrib_ptr, pat.nnzb + 1);
  ListWindow<8> cidx(pat.contrib_idx, (int64_t)pat.num_cells * pat.nodes_per_cell * pat.nodes_per_cell);
  int c1 = len > 0 ? cptr[rb] : 0;
  for (int k = 0; k < L; ++k) {
    TS acc[VV];
#pragma unroll
    for (int v = 0; v < VV; ++v) acc[v] = TS(0);
    if (k < len) {
      const int c0 = c1;
      c1 = cptr[rb + k + 1];
      for (int ci = c0; ci < c1; ++ci) {
        const TS* src = Ke + (size_t)cidx[ci] * VV;
#pragma unroll
        for (int v = 0; v < VV; ++v) acc[v] += src[v];
      }
    }
#pragma unroll
    for (int v = 0; v < VV; ++v) rsum[v] += acc[v];
    if (k == ds) {
#pragma unroll
      for (int c = 0; c < V; ++c) diag[c] = acc[c * V + c];
    }
    if (!(DIFF && k == ds)) {
#pragma unroll
      for (int v = 0; v < VV; ++v) vp[((size_t)k * VV + v) * kSlice] = T(acc[v]);
    }
  }
  if (DIFF && ds >= 0) {
#pragma unroll
    for (int v = 0; v < VV; ++v) vp[((size_t)ds * VV + v) * kSlice] = T(rsum[v]);
  }
  if (dinv && real) {
#pragma unroll
    for (int c = 0; c < V; ++c) dinv[(int64_t)row * V + c] = diag[c] != TS(0) ? T(TS(1) / diag[c]) : T(0);
  }
}

// Zero row d of the matrix and put 1 on its diagonal (PETSc zeroRows, diag = 1).
// The row of an identity row sums to e_d, so the same values serve both
// forms; dinv[d] = 1.
template <int V, typename T>
__global__ void k_bc_rows(fem_pattern_t pat, T* __restrict__ vals, int64_t n_bc, const int32_t* __restrict__ dofs,
                          T* __restrict__ dinv) {
  constexpr int VV = V * V;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_bc; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = dofs[i];
    const int64_t row = d / V;
    const int ci = (int)(d % V);
    const int64_t pos = pat.sell_pos[row];
    const int64_t s = pos >> 6;
    const int lane = (int)(pos & 63);
    const int base = pat.slice_ptr[s];
    const int L = (pat.slice_ptr[s + 1] - base) >> 6;
    T* vp = vals + (size_t)base * VV + lane;
    for (int k = 0; k < L; ++k)
#pragma unroll
      for (int cj = 0; cj < V; ++cj) vp[((size_t)k * VV + ci * V + cj) * kSlice] = T(0);
    const int ds = pat.diag_slot[row];
    if (ds >= 0) vp[((size_t)ds * VV + ci * V + ci) * kSlice] = T(1);
    if (dinv) dinv[d] = T(1);
  }
}

template <int V, typename T>
__global__ void k_jacobi(fem_pattern_t pat, const T* __restrict__ vals, T* __restrict__ dinv) {
  constexpr int VV = V * V;
  for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < pat.num_rows;
       row += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pos = pat.sell_pos[row];
    const int64_t s = pos >> 6;
    const int lane = (int)(pos & 63);
    const int base = pat.slice_ptr[s];
    const T* vp = vals + (size_t)base * VV + lane;
    const int ds = pat.diag_slot[row];
#pragma unroll
    for (int c = 0; c < V; ++c) {
      const T d = ds >= 0 ? vp[((size_t)ds * VV + c * V + c) * kSlice] : T(0);
      dinv[row * V + c] = d != T(0) ? T(1) / d : T(0);
    }
  }
}

__global__ void k_residual_bc(int64_t n_bc, const int32_t* __restrict__ dofs, const double* __restrict__ g,
                              const double* __restrict__ u, double* __restrict__ res, double scale) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_bc; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = dofs[i];
    res[d] = u[d] - g[i] * scale;
  }
}

// ------------------------------------------------------------ SpMV
// One block per slice; the W waves split the slots and wave 0 combines them.
template <int V, bool DIFF, typename T>
__global__ __launch_bounds__(SliceCfg<T>::NT) void k_spmv(fem_pattern_t pat, const T* __restrict__ vals,
                                                          const T* __restrict__ x, T* __restrict__ y) {
  constexpr int W = SliceCfg<T>::W, U = SliceCfg<T>::U;
  __shared__ double red[W][kSlice][V];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t s = blockIdx.x;
  const int64_t row = pat.sell_row[s * kSlice + lane];
  const bool real = row >= 0;
  auto xat = [&](int64_t c, T (&xc)[V]) {
#pragma unroll
    for (int j = 0; j < V; ++j) xc[j] = x[c * V + j];
  };
  T xt[V];
  xat(real ? row : 0, xt);
  double xr[V], acc[V];
#pragma unroll
  for (int j = 0; j < V; ++j) xr[j] = double(xt[j]);
  sell_slice_partial<V, DIFF, W, U, T>(pat, vals, s, w, lane, row, xat, xr, acc);
  if (w > 0) {
#pragma unroll
    for (int i = 0; i < V; ++i) red[w][lane][i] = acc[i];
  }
  __syncthreads();
  if (w == 0 && real) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      double t = acc[i];
#pragma unroll
      for (int o = 1; o < W; ++o) t += red[o][lane][i];
      y[row * V + i] = T(t);
    }
  }
}

// Scalar CSR of the block matrix: row r = V*node + ci holds the blocks of node
// row `node` in column order, each contributing V consecutive columns.  In
// difference form the diagonal block is recovered as R - sum of the others
// (in FP64).
template <int V, typename T>
__global__ void k_export_csr(fem_pattern_t pat, const T* __restrict__ vals, int form, int32_t* __restrict__ indptr,
                             int32_t* __restrict__ indices, double* __restrict__ data) {
  constexpr int VV = V * V;
  for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < pat.num_rows;
       row += (int64_t)gridDim.x * blockDim.x) {
    const int rb = pat.row_ptr[row];
    const int len = pat.row_ptr[row + 1] - rb;
    const int ds = pat.diag_slot[row];
    const int64_t pos = pat.sell_pos[row];
    const int64_t s = pos >> 6;
    const int lane = (int)(pos & 63);
    const T* vp = vals + (size_t)pat.slice_ptr[s] * VV + lane;
    for (int ci = 0; ci < V; ++ci) {
      const int64_t start = (int64_t)VV * rb + (int64_t)ci * V * len;
      if (indptr) indptr[row * V + ci] = (int32_t)start;
      for (int k = 0; k < len; ++k)
        for (int cj = 0; cj < V; ++cj) {
          const int64_t pos = start + (int64_t)k * V + cj;
          if (indices) indices[pos] = pat.col_idx[rb + k] * V + cj;
          if (data && vals) {
            double v = double(vp[((size_t)k * VV + ci * V + cj) * kSlice]);
            if (form == FEM_FORM_DIFFERENCE && k == ds)
              for (int kk = 0; kk < len; ++kk)
                if (kk != ds) v -= double(vp[((size_t)kk * VV + ci * V + cj) * kSlice]);
            data[pos] = v;
          }
        }
    }
    if (indptr && row == pat.num_rows - 1) indptr[pat.num_rows * V] = (int32_t)((int64_t)VV * pat.row_ptr[pat.num_rows]);
  }
}

// ------------------------------------------------------- small vector ops
__global__ void k_copy_bc(int64_t n_bc, const int32_t* __restrict__ dofs, const double* __restrict__ x,
                          double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_bc; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = dofs[i];
    y[d] = x[d];
  }
}

__global__ __launch_bounds__(256) void k_dot(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                             double* __restrict__ partials) {
  __shared__ double smem[8];
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += x[i] * y[i];
  s = block_sum_256(s, smem);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_dot_finish(const double* __restrict__ partials, int n, double* __restrict__ out) {
  __shared__ double smem[8];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += 256) s += partials[i];
  s = block_sum_256(s, smem);
  if (threadIdx.x == 0) *out = s;
}

static int grid_for(int64_t n, int threads, int cap = 65536) {
  int64_t g = ceil_div(n, threads);
  if (g < 1) g = 1;
  return (int)(g > cap ? cap : g);
}

template <class F>
static int with_vt(int vec, int precision, F&& f) {
  if (vec == 1 && precision == FEM_F32) return f(std::integral_constant<int, 1>{}, float{});
  if (vec == 1 && precision == FEM_F64) return f(std::integral_constant<int, 1>{}, double{});
  if (vec == 3 && precision == FEM_F32) return f(std::integral_constant<int, 3>{}, float{});
  if (vec == 3 && precision == FEM_F64) return f(std::integral_constant<int, 3>{}, double{});
  // block sizes of several variables on one mesh (sum of their vec)
  if (vec == 2 && precision == FEM_F32) return f(std::integral_constant<int, 2>{}, float{});
  if (vec == 2 && precision == FEM_F64) return f(std::integral_constant<int, 2>{}, double{});
  if (vec == 4 && precision == FEM_F32) return f(std::integral_constant<int, 4>{}, float{});
  if (vec == 4 && precision == FEM_F64) return f(std::integral_constant<int, 4>{}, double{});
  set_error("unsupported vec=%d / precision=%d", vec, precision);
  return FEM_ERR_BAD_ARG;
}

static int check_pat(const fem_pattern_t* p) {
  FEM_CHECK_ARG(p && p->row_ptr && p->slice_ptr && p->sell_row && p->sell_pos && p->num_rows > 0,
                "pattern descriptor invalid");
  FEM_CHECK_ARG(p->num_slices == ceil_div(p->num_rows, kSlice), "pattern num_slices mismatch");
  return FEM_OK;
}

}  // namespace fem

using namespace fem;

// Implemented here, used by the PCG driver.
int fem_operator_apply_impl(const fem_operator_t* op, const double* x, double* y, hipStream_t s);

int fem_operator_apply_impl(const fem_operator_t* op, const double* x, double* y, hipStream_t s) {
  FEM_CHECK_ARG(op && op->mesh && op->mat && op->u && op->inc_ptr && op->inc_idx && op->cell_scratch && x && y,
                "fem_operator_apply: NULL argument");
  int rc = fem_cell_jvp(op->mesh, op->mat, op->u, x, op->cell_scratch, s);
  if (rc) return rc;
  if (op->mass_M && (rc = fem_cell_mass_jvp(op->mesh, op->mass_M, x, op->cell_scratch, s))) return rc;
  if (op->cell_K && (rc = fem_cell_matvec(op->mesh, op->cell_K, x, op->cell_scratch, s))) return rc;
  rc = fem_gather_nodal(op->mesh->num_nodes, op->mesh->vec, op->inc_ptr, op->inc_idx, op->cell_scratch, nullptr, y, s);
  if (rc) return rc;
  if (op->n_bc > 0) {
    hipLaunchKernelGGL(k_copy_bc, dim3(grid_for(op->n_bc, 256)), dim3(256), 0, s, op->n_bc, op->bc_dofs, x, y);
    FEM_LAUNCHED();
  }
  return FEM_OK;
}

extern "C" {

int fem_gather_nodal(int64_t num_nodes, int32_t vec, const int32_t* inc_ptr, const int32_t* inc_idx,
                     const double* ent, const double* base, double* out, void* stream) {
  FEM_CHECK_ARG(num_nodes > 0 && vec >= 1 && vec <= 4 && inc_ptr && inc_idx && ent && out,
                "fem_gather_nodal: bad arguments");
  const dim3 grid(grid_for(num_nodes, 256));
  hipStream_t s = as_stream(stream);
  if (vec == 3)
    hipLaunchKernelGGL(k_gather_nodal<3>, grid, dim3(256), 0, s, num_nodes, inc_ptr, inc_idx, ent, base, out);
  else if (vec == 1)
    hipLaunchKernelGGL(k_gather_nodal<1>, grid, dim3(256), 0, s, num_nodes, inc_ptr, inc_idx, ent, base, out);
  else if (vec == 2)
    hipLaunchKernelGGL(k_gather_nodal<2>, grid, dim3(256), 0, s, num_nodes, inc_ptr, inc_idx, ent, base, out);
  else
    hipLaunchKernelGGL(k_gather_nodal<4>, grid, dim3(256), 0, s, num_nodes, inc_ptr, inc_idx, ent, base, out);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_gather_matrix(const fem_pattern_t* pat, int32_t ke_precision, const void* Ke, int32_t precision, int32_t form,
                      void* vals, void* dinv, void* stream) {
  if (int rc = check_pat(pat)) return rc;
  FEM_CHECK_ARG(Ke && vals && pat->contrib_ptr && pat->contrib_idx && pat->diag_slot,
                "fem_gather_matrix: NULL argument");
  FEM_CHECK_ARG(pat->nnzb > 0 && pat->num_cells > 0 && pat->nodes_per_cell > 0,
                "fem_gather_matrix: pattern nnzb, num_cells and nodes_per_cell must be set");
  FEM_CHECK_ARG(ke_precision == precision || (ke_precision == FEM_F64 && precision == FEM_F32),
                "fem_gather_matrix: Ke precision must equal the matrix precision or be FP64");
  FEM_CHECK_ARG(form == FEM_FORM_STANDARD || form == FEM_FORM_DIFFERENCE, "fem_gather_matrix: bad form");
  const int64_t threads = pat->num_slices * kSlice;
  const dim3 grid((unsigned)ceil_div(threads, 256));
  return with_vt(pat->vec, precision, [&](auto vc, auto tv) -> int {
    constexpr int V = decltype(vc)::value;
    using T = decltype(tv);
    hipStream_t s = as_stream(stream);
    if (ke_precision == FEM_F64) {
      if (form == FEM_FORM_DIFFERENCE)
        hipLaunchKernelGGL((k_gather_matrix<V, true, double, T>), grid, dim3(256), 0, s, *pat, (const double*)Ke,
                           (T*)vals, (T*)dinv);
      else
        hipLaunchKernelGGL((k_gather_matrix<V, false, double, T>), grid, dim3(256), 0, s, *pat, (const double*)Ke,
                           (T*)vals, (T*)dinv);
    } else {
      if (form == FEM_FORM_DIFFERENCE)
        hipLaunchKernelGGL((k_gather_matrix<V, true, T, T>), grid, dim3(256), 0, s, *pat, (const T*)Ke, (T*)vals,
                           (T*)dinv);
      else
        hipLaunchKernelGGL((k_gather_matrix<V, false, T, T>), grid, dim3(256), 0, s, *pat, (const T*)Ke, (T*)vals,
                           (T*)dinv);
    }
    FEM_LAUNCHED();
    return FEM_OK;
  });
}

int fem_matrix_bc_rows(const fem_pattern_t* pat, int32_t precision, void* vals, int64_t n_bc, const int32_t* bc_dofs,
                       void* dinv, void* stream) {
  if (int rc = check_pat(pat)) return rc;
  FEM_CHECK_ARG(vals && (n_bc == 0 || bc_dofs) && pat->diag_slot, "fem_matrix_bc_rows: NULL argument");
  if (n_bc == 0) return FEM_OK;
  return with_vt(pat->vec, precision, [&](auto vc, auto tv) -> int {
    constexpr int V = decltype(vc)::value;
    using T = decltype(tv);
    hipLaunchKernelGGL((k_bc_rows<V, T>), dim3(grid_for(n_bc, 256)), dim3(256), 0, as_stream(stream), *pat, (T*)vals,
                       n_bc, bc_dofs, (T*)dinv);
    FEM_LAUNCHED();
    return FEM_OK;
  });
}

int fem_jacobi(const fem_pattern_t* pat, int32_t precision, const void* vals, void* dinv, void* stream) {
  if (int rc = check_pat(pat)) return rc;
  FEM_CHECK_ARG(vals && dinv && pat->diag_slot, "fem_jacobi: NULL argument");
  return with_vt(pat->vec, precision, [&](auto vc, auto tv) -> int {
    constexpr int V = decltype(vc)::value;
    using T = decltype(tv);
    hipLaunchKernelGGL((k_jacobi<V, T>), dim3(grid_for(pat->num_rows, 256)), dim3(256), 0, as_stream(stream), *pat,
                       (const T*)vals, (T*)dinv);
    FEM_LAUNCHED();
    return FEM_OK;
  });
}

int fem_residual_bc(int64_t n_bc, const int32_t* bc_dofs, const double* bc_vals, const double* u, double* res,
                    double scale, void* stream) {
  FEM_CHECK_ARG(n_bc >= 0 && (n_bc == 0 || (bc_dofs && bc_vals && u && res)), "fem_residual_bc: bad arguments");
  if (n_bc == 0) return FEM_OK;
  hipLaunchKernelGGL(k_residual_bc, dim3(grid_for(n_bc, 256)), dim3(256), 0, as_stream(stream), n_bc, bc_dofs,
                     bc_vals, u, res, scale);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_spmv(const fem_pattern_t* pat, int32_t precision, int32_t form, const void* vals, const void* x, void* y,
             void* stream) {
  if (int rc = check_pat(pat)) return rc;
  FEM_CHECK_ARG(vals && x && y && pat->sell_col, "fem_spmv: NULL argument");
  FEM_CHECK_ARG(form == FEM_FORM_STANDARD || form == FEM_FORM_DIFFERENCE, "fem_spmv: bad form");
  return with_vt(pat->vec, precision, [&](auto vc, auto tv) -> int {
    constexpr int V = decltype(vc)::value;
    using T = decltype(tv);
    const dim3 grid((unsigned)pat->num_slices), block(SliceCfg<T>::NT);
    if (form == FEM_FORM_DIFFERENCE)
      hipLaunchKernelGGL((k_spmv<V, true, T>), grid, block, 0, as_stream(stream), *pat, (const T*)vals, (const T*)x,
                         (T*)y);
    else
      hipLaunchKernelGGL((k_spmv<V, false, T>), grid, block, 0, as_stream(stream), *pat, (const T*)vals, (const T*)x,
                         (T*)y);
    FEM_LAUNCHED();
    return FEM_OK;
  });
}

int fem_export_csr(const fem_pattern_t* pat, int32_t precision, int32_t form, const void* vals, int32_t* indptr,
                   int32_t* indices, double* data, void* stream) {
  if (int rc = check_pat(pat)) return rc;
  FEM_CHECK_ARG(pat->col_idx && pat->diag_slot, "fem_export_csr: pattern without col_idx");
  FEM_CHECK_ARG((int64_t)pat->vec * pat->vec * pat->nnzb < (int64_t)INT32_MAX, "fem_export_csr: nnz exceeds int32");
  return with_vt(pat->vec, precision, [&](auto vc, auto tv) -> int {
    constexpr int V = decltype(vc)::value;
    using T = decltype(tv);
    hipLaunchKernelGGL((k_export_csr<V, T>), dim3(grid_for(pat->num_rows, 256)), dim3(256), 0, as_stream(stream),
                       *pat, (const T*)vals, form, indptr, indices, data);
    FEM_LAUNCHED();
    return FEM_OK;
  });
}

int fem_operator_apply(const fem_operator_t* op, const double* x, double* y, void* stream) {
  return fem_operator_apply_impl(op, x, y, as_stream(stream));
}

int fem_dot(int64_t n, const double* x, const double* y, double* out, void* ws, void* stream) {
  FEM_CHECK_ARG(n >= 0 && x && y && out && ws, "fem_dot: bad arguments");
  const int grid = grid_for(n, 256, 512);
  double* partials = reinterpret_cast<double*>(ws);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_dot, dim3(grid), dim3(256), 0, s, n, x, y, partials);
  hipLaunchKernelGGL(k_dot_finish, dim3(1), dim3(256), 0, s, (const double*)partials, grid, out);
  FEM_LAUNCHED();
  return FEM_OK;
}

}  // extern "C"
