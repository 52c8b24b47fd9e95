// Jacobi-preconditioned conjugate gradients on the SELL-64 block matrix.
//
// Replaces the default linear solve of jax-fem (solver.py:63-92, jax_solve:
// host CSR -> scipy -> BCOO conversion on every call, then BiCGSTAB with
// M = x / diag(A), tol = atol = 1e-10, maxiter 10000).  The Dirichlet rows of
// A are identity rows (solver.py:527-528) and the initial guess carries the
// exact boundary increment (solver.py:402-409), so the residual and search
// direction vanish on those rows and CG runs on the SPD free block.
//
// Two kernels per iteration, both ending in a deterministic grid reduction
// whose last block computes the CG scalars on the device:
//   A: p = z + beta p_old (ping-pong buffers), q = A p, <p, q> -> alpha
//   B: x += alpha p, r -= alpha q, z = D^-1 r, <r, z>, <r, r> -> beta, done
// Precision FEM_F64: everything FP64.  FEM_F32: matrix, r, z, p, q in FP32,
// x accumulated in FP64, and the recursive residual is replaced by the FP64
// true residual b - A x (A applied matrix-free in FP64) whenever it has
// dropped by `reliable_delta` and at convergence ("reliable updates").
#include "fem_common.hpp"

#include <algorithm>
#include <cmath>

int fem_operator_apply_impl(const fem_operator_t* op, const double* x, double* y, hipStream_t s);

namespace fem {

struct PcgScalars {
  double alpha, beta, rz, pq, rr, rr_true;
  int32_t done, iter, nan, pad;
};

template <int V, typename T>
__global__ __launch_bounds__(256) void k_pcg_spmv(fem_pattern_t pat, const T* __restrict__ vals,
                                                  const T* __restrict__ z, const T* __restrict__ p_old,
                                                  T* __restrict__ p_new, T* __restrict__ q, PcgScalars* sc,
                                                  double* __restrict__ partials, unsigned* __restrict__ counter,
                                                  int remap) {
  constexpr int VV = V * V;
  __shared__ double smem[8];
  if (sc->done) return;
  const T beta = T(sc->beta);
  const int nb = gridDim.x;
  const int blk = remap ? xcd_logical_block(blockIdx.x, nb) : blockIdx.x;
  const int64_t row = (int64_t)blk * 256 + threadIdx.x;
  T dot = T(0);
  if (row < pat.num_rows) {
    const int64_t s = row >> 6;
    const int lane = (int)(row & 63);
    const int base = pat.slice_ptr[s];
    const int L = (pat.slice_ptr[s + 1] - base) >> 6;
    const int32_t* cp = pat.sell_col + base + lane;
    const T* vp = vals + (size_t)base * VV + lane;
    T acc[V];
#pragma unroll
    for (int i = 0; i < V; ++i) acc[i] = T(0);
#pragma unroll 2
    for (int k = 0; k < L; ++k) {
      const int64_t c = cp[(size_t)k * kSlice];
      T xc[V];
#pragma unroll
      for (int j = 0; j < V; ++j) xc[j] = z[c * V + j] + beta * p_old[c * V + j];
#pragma unroll
      for (int i = 0; i < V; ++i)
#pragma unroll
        for (int j = 0; j < V; ++j) acc[i] += vp[((size_t)k * VV + i * V + j) * kSlice] * xc[j];
    }
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const T pn = z[row * V + i] + beta * p_old[row * V + i];
      p_new[row * V + i] = pn;
      q[row * V + i] = acc[i];
      dot += pn * acc[i];
    }
  }
  double d = block_sum_256(double(dot), smem);
  double mine[1] = {d}, tot[1];
  if (last_block_sums<1>(mine, partials, counter, tot, smem) && threadIdx.x == 0) {
    sc->pq = tot[0];
    sc->alpha = sc->rz / tot[0];
  }
}

template <typename T, typename TX>
__global__ __launch_bounds__(256) void k_pcg_update(int64_t n, TX* __restrict__ x, T* __restrict__ r,
                                                    T* __restrict__ z, const T* __restrict__ p,
                                                    const T* __restrict__ q, const T* __restrict__ dinv,
                                                    PcgScalars* sc, double* __restrict__ partials,
                                                    unsigned* __restrict__ counter, double tol2, int maxiter) {
  __shared__ double smem[8];
  if (sc->done) return;
  const double alpha_d = sc->alpha;
  const T alpha = T(alpha_d);
  T rz = T(0), rr = T(0);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += TX(alpha_d) * TX(p[i]);
    const T ri = r[i] - alpha * q[i];
    const T zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  double mine[2] = {block_sum_256(double(rz), smem), block_sum_256(double(rr), smem + 4)};
  double tot[2];
  if (last_block_sums<2>(mine, partials, counter, tot, smem) && threadIdx.x == 0) {
    const double rz_new = tot[0], rr_new = tot[1];
    sc->beta = rz_new / sc->rz;
    sc->rz = rz_new;
    sc->rr = rr_new;
    sc->iter += 1;
    const bool bad = !isfinite(rr_new) || !isfinite(rz_new);
    if (bad) sc->nan = 1;
    if (bad || rr_new <= tol2 || sc->iter >= maxiter) sc->done = 1;
  }
}

// r = (T)(b - y) and <b - y, b - y> in FP64 -> rr_true.
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_resid(int64_t n, const double* __restrict__ b, const double* __restrict__ y,
                                                   T* __restrict__ r, PcgScalars* sc, double* __restrict__ partials,
                                                   unsigned* __restrict__ counter) {
  __shared__ double smem[8];
  double rr = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double ri = b[i] - y[i];
    r[i] = T(ri);
    rr += ri * ri;
  }
  double mine[1] = {block_sum_256(rr, smem)}, tot[1];
  if (last_block_sums<1>(mine, partials, counter, tot, smem) && threadIdx.x == 0) sc->rr_true = tot[0];
}

// z = D^-1 r, rz, rr; init (beta = 0) or residual replacement (beta = rz_new / rz_old).
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_restart(int64_t n, const T* __restrict__ r, T* __restrict__ z,
                                                     const T* __restrict__ dinv, PcgScalars* sc,
                                                     double* __restrict__ partials, unsigned* __restrict__ counter,
                                                     int init) {
  __shared__ double smem[8];
  T rz = T(0), rr = T(0);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T zi = dinv[i] * r[i];
    z[i] = zi;
    rz += r[i] * zi;
    rr += r[i] * r[i];
  }
  double mine[2] = {block_sum_256(double(rz), smem), block_sum_256(double(rr), smem + 4)};
  double tot[2];
  if (last_block_sums<2>(mine, partials, counter, tot, smem) && threadIdx.x == 0) {
    sc->beta = init ? 0.0 : tot[0] / sc->rz;
    sc->rz = tot[0];
    sc->rr = tot[1];
    sc->done = 0;
    if (init) {
      sc->iter = 0;
      sc->nan = 0;
    }
  }
}

template <typename T>
__global__ void k_fill(int64_t n, T* __restrict__ v, T val) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = val;
}

template <typename TS, typename TD>
__global__ void k_cast(int64_t n, const TS* __restrict__ a, TD* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = TD(a[i]);
}

template <int V, typename T>
__global__ __launch_bounds__(256) void k_spmv_x(fem_pattern_t pat, const T* __restrict__ vals,
                                                const double* __restrict__ x, double* __restrict__ y) {
  constexpr int VV = V * V;
  const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (row >= pat.num_rows) return;
  const int64_t s = row >> 6;
  const int lane = (int)(row & 63);
  const int base = pat.slice_ptr[s];
  const int L = (pat.slice_ptr[s + 1] - base) >> 6;
  const int32_t* cp = pat.sell_col + base + lane;
  const T* vp = vals + (size_t)base * VV + lane;
  double acc[V];
#pragma unroll
  for (int i = 0; i < V; ++i) acc[i] = 0.0;
  for (int k = 0; k < L; ++k) {
    const int64_t c = cp[(size_t)k * kSlice];
#pragma unroll
    for (int i = 0; i < V; ++i)
#pragma unroll
      for (int j = 0; j < V; ++j) acc[i] += double(vp[((size_t)k * VV + i * V + j) * kSlice]) * x[c * V + j];
  }
#pragma unroll
  for (int i = 0; i < V; ++i) y[row * V + i] = acc[i];
}

constexpr int kUpdGridCap = 2048;

struct Ws {
  void* r;
  void* z;
  void* p0;
  void* p1;
  void* q;
  void* ones;
  double* y;
  double* partials;
  unsigned* counters;
  PcgScalars* sc;
};

static int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

static int64_t grid_a(const fem_pattern_t* pat) { return ceil_div(pat->num_rows, 256); }

static int64_t ws_layout(const fem_pattern_t* pat, int precision, void* base, Ws* w) {
  const int64_t n = pat->num_rows * pat->vec;
  const int64_t es = precision == FEM_F32 ? 4 : 8;
  const int64_t nparts = grid_a(pat) + 2 * kUpdGridCap + 64;  // kernel A, then kernel B's two sums
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* p = base ? reinterpret_cast<char*>(base) + off : nullptr;
    off += align256(bytes);
    return p;
  };
  Ws tmp;
  Ws* o = w ? w : &tmp;
  o->r = take(n * es);
  o->z = take(n * es);
  o->p0 = take(n * es);
  o->p1 = take(n * es);
  o->q = take(n * es);
  o->ones = take(n * es);
  o->y = reinterpret_cast<double*>(take(n * 8));
  o->partials = reinterpret_cast<double*>(take(nparts * 8));
  o->counters = reinterpret_cast<unsigned*>(take(4 * sizeof(unsigned)));
  o->sc = reinterpret_cast<PcgScalars*>(take(sizeof(PcgScalars)));
  return off;
}

template <int V, typename T>
static int pcg_run(const fem_pattern_t* pat, const T* vals, const T* dinv_in, const fem_operator_t* op,
                   const double* b, double* x, const fem_pcg_options_t* opt, void* ws_base,
                   PcgScalars* host_sc, fem_pcg_result_t* res, hipStream_t s) {
  const int64_t n = pat->num_rows * V;
  Ws w;
  ws_layout(pat, std::is_same<T, float>::value ? FEM_F32 : FEM_F64, ws_base, &w);
  T* r = (T*)w.r;
  T* z = (T*)w.z;
  T* pbuf[2] = {(T*)w.p0, (T*)w.p1};
  T* q = (T*)w.q;
  const T* dinv = dinv_in;
  if (!opt->precond || !dinv) {
    hipLaunchKernelGGL(k_fill<T>, dim3(256), dim3(256), 0, s, n, (T*)w.ones, T(1));
    FEM_LAUNCHED();
    dinv = (T*)w.ones;
  }
  const bool mixed = std::is_same<T, float>::value;
  const int gA = (int)grid_a(pat);
  const int gB = (int)std::min<int64_t>(ceil_div(n, 256), kUpdGridCap);
  unsigned* cnt = w.counters;
  FEM_HIP(hipMemsetAsync(cnt, 0, 4 * sizeof(unsigned), s));
  FEM_HIP(hipMemsetAsync(w.sc, 0, sizeof(PcgScalars), s));
  hipLaunchKernelGGL(k_fill<T>, dim3(256), dim3(256), 0, s, n, pbuf[0], T(0));
  FEM_LAUNCHED();

  // |b|^2 (reuse the residual kernel with y = 0)
  hipLaunchKernelGGL(k_fill<double>, dim3(256), dim3(256), 0, s, n, w.y, 0.0);
  FEM_LAUNCHED();
  hipLaunchKernelGGL(k_pcg_resid<T>, dim3(gB), dim3(256), 0, s, n, b, (const double*)w.y, r, w.sc, w.partials, cnt + 2);
  FEM_LAUNCHED();
  FEM_HIP(hipMemcpyAsync(host_sc, w.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
  FEM_HIP(hipStreamSynchronize(s));
  const double bb = host_sc->rr_true;
  const double tol2 = std::max(opt->rtol * opt->rtol * bb, opt->atol * opt->atol);
  res->b_norm = std::sqrt(bb);

  auto true_residual = [&]() -> int {
    if (mixed) {
      int rc = fem_operator_apply_impl(op, x, w.y, s);
      if (rc) return rc;
    } else {
      hipLaunchKernelGGL((k_spmv_x<V, T>), dim3(gA), dim3(256), 0, s, *pat, vals, (const double*)x, w.y);
      FEM_LAUNCHED();
    }
    hipLaunchKernelGGL(k_pcg_resid<T>, dim3(gB), dim3(256), 0, s, n, b, (const double*)w.y, r, w.sc, w.partials,
                       cnt + 2);
    FEM_LAUNCHED();
    return FEM_OK;
  };

  if (int rc = true_residual()) return rc;
  hipLaunchKernelGGL(k_pcg_restart<T>, dim3(gB), dim3(256), 0, s, n, (const T*)r, z, dinv, w.sc, w.partials,
                     cnt + 3, 1);
  FEM_LAUNCHED();
  FEM_HIP(hipMemcpyAsync(host_sc, w.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
  FEM_HIP(hipStreamSynchronize(s));
  double rr_true = host_sc->rr_true;
  double rr_ref = rr_true;
  res->iterations = 0;
  res->replacements = 0;
  res->status = FEM_OK;
  if (!std::isfinite(rr_true)) {
    res->status = FEM_ERR_NAN;
    res->res_norm = rr_true;
    return FEM_OK;
  }
  if (rr_true <= tol2) {
    res->res_norm = std::sqrt(rr_true);
    return FEM_OK;
  }
  const int check = std::max(1, opt->check_every);
  const double delta2 = mixed ? opt->reliable_delta * opt->reliable_delta : 0.0;
  int iter = 0;
  for (;;) {
    for (int m = 0; m < check; ++m) {
      const int cur = (iter + m) & 1;
      hipLaunchKernelGGL((k_pcg_spmv<V, T>), dim3(gA), dim3(256), 0, s, *pat, vals, (const T*)z,
                         (const T*)pbuf[cur], pbuf[cur ^ 1], q, w.sc, w.partials, cnt + 0, opt->xcd_remap);
      hipLaunchKernelGGL((k_pcg_update<T, double>), dim3(gB), dim3(256), 0, s, n, x, r, z, (const T*)pbuf[cur ^ 1],
                         (const T*)q, dinv, w.sc, w.partials + gA, cnt + 1, tol2, opt->maxiter);
    }
    FEM_LAUNCHED();
    FEM_HIP(hipMemcpyAsync(host_sc, w.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    iter = host_sc->iter;
    if (host_sc->nan) {
      res->status = FEM_ERR_NAN;
      break;
    }
    const bool conv_rec = host_sc->done && host_sc->rr <= tol2;
    const bool trigger = mixed && host_sc->rr < delta2 * rr_ref;
    if (conv_rec || trigger) {
      if (int rc = true_residual()) return rc;
      FEM_HIP(hipMemcpyAsync(host_sc, w.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
      FEM_HIP(hipStreamSynchronize(s));
      rr_true = host_sc->rr_true;
      if (rr_true <= tol2) break;
      if (!std::isfinite(rr_true)) {
        res->status = FEM_ERR_NAN;
        break;
      }
      // residual replacement: keep the search direction, restart the recurrences from b - A x
      hipLaunchKernelGGL(k_pcg_restart<T>, dim3(gB), dim3(256), 0, s, n, (const T*)r, z, dinv, w.sc, w.partials,
                         cnt + 3, 0);
      FEM_LAUNCHED();
      rr_ref = rr_true;
      res->replacements += 1;
    }
    if (iter >= opt->maxiter) {
      res->status = FEM_ERR_NOT_CONVERGED;
      break;
    }
  }
  // final true residual
  if (res->status == FEM_OK && !(rr_true <= tol2)) {
    if (int rc = true_residual()) return rc;
    FEM_HIP(hipMemcpyAsync(host_sc, w.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    rr_true = host_sc->rr_true;
  }
  if (res->status != FEM_OK && res->status != FEM_ERR_NAN) {
    if (int rc = true_residual()) return rc;
    FEM_HIP(hipMemcpyAsync(host_sc, w.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    rr_true = host_sc->rr_true;
  }
  res->iterations = iter;
  res->res_norm = std::sqrt(rr_true);
  return FEM_OK;
}

}  // namespace fem

using namespace fem;

extern "C" {

int64_t fem_pcg_ws_bytes(const fem_pattern_t* pat, int32_t precision) {
  if (!pat || pat->num_rows <= 0) return -1;
  return ws_layout(pat, precision, nullptr, nullptr) + 256;
}

int fem_pcg_solve(const fem_pattern_t* pat, const void* vals, const void* dinv, const fem_operator_t* op,
                  const double* b, double* x, const fem_pcg_options_t* opt, void* ws, void* host_status,
                  fem_pcg_result_t* result, void* stream) {
  FEM_CHECK_ARG(pat && pat->slice_ptr && pat->sell_col && pat->num_rows > 0, "fem_pcg_solve: bad pattern");
  FEM_CHECK_ARG(vals && b && x && opt && ws && host_status && result, "fem_pcg_solve: NULL argument");
  FEM_CHECK_ARG(opt->precision == FEM_F32 || opt->precision == FEM_F64, "fem_pcg_solve: bad precision");
  FEM_CHECK_ARG(opt->precision == FEM_F64 || op, "fem_pcg_solve: FP32 mode needs the FP64 operator");
  FEM_CHECK_ARG(opt->maxiter > 0 && opt->rtol >= 0 && opt->atol >= 0, "fem_pcg_solve: bad tolerances");
  FEM_CHECK_ARG(!opt->precond || dinv, "fem_pcg_solve: Jacobi requested without dinv");
  hipStream_t s = as_stream(stream);
  PcgScalars* hsc = reinterpret_cast<PcgScalars*>(host_status);
  const int V = pat->vec;
  if (V == 1 && opt->precision == FEM_F32)
    return pcg_run<1, float>(pat, (const float*)vals, (const float*)dinv, op, b, x, opt, ws, hsc, result, s);
  if (V == 1 && opt->precision == FEM_F64)
    return pcg_run<1, double>(pat, (const double*)vals, (const double*)dinv, op, b, x, opt, ws, hsc, result, s);
  if (V == 3 && opt->precision == FEM_F32)
    return pcg_run<3, float>(pat, (const float*)vals, (const float*)dinv, op, b, x, opt, ws, hsc, result, s);
  if (V == 3 && opt->precision == FEM_F64)
    return pcg_run<3, double>(pat, (const double*)vals, (const double*)dinv, op, b, x, opt, ws, hsc, result, s);
  set_error("fem_pcg_solve: unsupported vec %d", V);
  return FEM_ERR_BAD_ARG;
}

}  // extern "C"
