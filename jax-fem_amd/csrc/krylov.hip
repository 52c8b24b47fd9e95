// Jacobi-preconditioned BiCGSTAB on the FP64 standard-form matrix: the
// jax.scipy.sparse.linalg.bicgstab + Jacobi of jax_solve (solver.py:63-92),
// used where the tangent is not symmetric (CG does not apply).
//
// Per iteration: two SpMVs (fem_spmv), three fused dot passes with fixed
// per-block partials (deterministic sums), three vector updates, and
// single-thread scalar kernels; the scalars stay on the device and the host
// reads the status every check_every iterations.
#include "fem_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fem {
namespace {

constexpr int kBiBlocks = 512;

struct BiScalars {
  double rho, alpha, omega, beta;
  double rr, stop2, bb;
  double d0, d1;        // results of the last dot pass
  int32_t done, iter, nan, breakdown, maxiter, pad;
};
static_assert(sizeof(BiScalars) <= 128, "host status buffer");

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 4096)); }

// part[2*blk] = sum a_i b_i, part[2*blk+1] = sum c_i d_i over this block's strided range
__global__ __launch_bounds__(256) void k_bi_dot(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                const double* __restrict__ c, const double* __restrict__ d,
                                                double* __restrict__ part, const BiScalars* sc) {
  __shared__ double smem[4];
  if (sc && sc->done) return;
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    s0 += a[i] * b[i];
    if (c) s1 += c[i] * d[i];
  }
  s0 = block_sum_256(s0, smem);
  s1 = block_sum_256(s1, smem);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0;
    part[2 * blockIdx.x + 1] = s1;
  }
}

// sc->d0, sc->d1 = sums of the partials (one block, fixed order), then the stage's scalar update
// STAGE 0: start   bb = d0, rr = d1 (= <r, r>), rho = d1 (rhat = r)
// STAGE 1: alpha = rho / <rhat, v>
// STAGE 2: omega = <t, s> / <t, t>
// STAGE 3: rr = <r, r>, beta = (<rhat, r> / rho) (alpha / omega), rho = <rhat, r>; stop test
template <int STAGE>
__global__ __launch_bounds__(256) void k_bi_scalar(const double* __restrict__ part, int nb, BiScalars* sc,
                                                   double rtol, double atol, int maxiter) {
  __shared__ double smem[4];
  if (STAGE != 0 && sc->done) return;
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) {
    s0 += part[2 * i];
    s1 += part[2 * i + 1];
  }
  s0 = block_sum_256(s0, smem);
  s1 = block_sum_256(s1, smem);
  if (threadIdx.x != 0) return;
  sc->d0 = s0;
  sc->d1 = s1;
  if (STAGE == 0) {
    sc->bb = s0;
    sc->rr = s1;
    sc->rho = s1;
    sc->alpha = sc->omega = 1.0;
    sc->beta = 0.0;
    const double tol = fmax(rtol * sqrt(s0), atol);
    sc->stop2 = tol * tol;
    sc->iter = 0;
    sc->nan = !isfinite(s1);
    sc->breakdown = 0;
    sc->maxiter = maxiter;
    sc->done = sc->nan || s1 <= sc->stop2 || maxiter <= 0;
  } else if (STAGE == 1) {
    if (s0 == 0.0) {
      sc->breakdown = 1;
      sc->done = 1;
      return;
    }
    sc->alpha = sc->rho / s0;
  } else if (STAGE == 2) {
    sc->omega = s1 != 0.0 ? s0 / s1 : 0.0;
  } else {
    sc->rr = s0;
    sc->iter += 1;
    const bool bad = !isfinite(s0) || !isfinite(s1);
    sc->nan = bad;
    if (bad || s0 <= sc->stop2 || sc->iter >= sc->maxiter) {
      sc->done = 1;
      return;
    }
    if (sc->omega == 0.0 || sc->rho == 0.0) {
      sc->breakdown = 1;
      sc->done = 1;
      return;
    }
    sc->beta = (s1 / sc->rho) * (sc->alpha / sc->omega);
    sc->rho = s1;
  }
}

// p = r + beta (p - omega v); phat = D^-1 p
__global__ __launch_bounds__(256) void k_bi_p(int64_t n, const double* __restrict__ r, const double* __restrict__ v,
                                              const double* __restrict__ dinv, double* __restrict__ p,
                                              double* __restrict__ ph, const BiScalars* sc) {
  if (sc->done) return;
  const double beta = sc->beta, omega = sc->omega;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double pi = r[i] + beta * (p[i] - omega * v[i]);
    p[i] = pi;
    ph[i] = dinv ? dinv[i] * pi : pi;
  }
}

// s = r - alpha v; shat = D^-1 s
__global__ __launch_bounds__(256) void k_bi_s(int64_t n, const double* __restrict__ r, const double* __restrict__ v,
                                              const double* __restrict__ dinv, double* __restrict__ s,
                                              double* __restrict__ sh, const BiScalars* sc) {
  if (sc->done) return;
  const double alpha = sc->alpha;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double si = r[i] - alpha * v[i];
    s[i] = si;
    sh[i] = dinv ? dinv[i] * si : si;
  }
}

// x += alpha phat + omega shat; r = s - omega t
__global__ __launch_bounds__(256) void k_bi_xr(int64_t n, const double* __restrict__ ph,
                                               const double* __restrict__ sh, const double* __restrict__ s,
                                               const double* __restrict__ t, double* __restrict__ x,
                                               double* __restrict__ r, const BiScalars* sc) {
  if (sc->done) return;
  const double alpha = sc->alpha, omega = sc->omega;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * ph[i] + omega * sh[i];
    r[i] = s[i] - omega * t[i];
  }
}

__global__ __launch_bounds__(256) void k_bi_zero2(int64_t n, double* __restrict__ a, double* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = b[i] = 0.0;
}

// r = b - y
__global__ __launch_bounds__(256) void k_bi_resid(int64_t n, const double* __restrict__ b,
                                                  const double* __restrict__ y, double* __restrict__ r,
                                                  double* __restrict__ rhat) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double ri = b[i] - y[i];
    r[i] = ri;
    if (rhat) rhat[i] = ri;
  }
}

// y = A x, scalar CSR, one thread per row
__global__ __launch_bounds__(256) void k_csr_spmv(int64_t rows, const int32_t* __restrict__ ptr,
                                                  const int32_t* __restrict__ idx, const double* __restrict__ val,
                                                  const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows; i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) acc += val[k] * x[idx[k]];
    y[i] = acc;
  }
}

// reduced Jacobi: dr[j] = 1 / sum_k PT_jk^2 A_kk with A_kk = 1 / dinv[k]
__global__ __launch_bounds__(256) void k_reduced_dinv(int64_t m, const int32_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ idx,
                                                      const double* __restrict__ val,
                                                      const double* __restrict__ dinv, double* __restrict__ dr) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    double d = 0.0;
    for (int32_t k = ptr[j]; k < ptr[j + 1]; ++k) d += val[k] * val[k] / dinv[idx[k]];
    dr[j] = d != 0.0 ? 1.0 / d : 1.0;
  }
}

int read_scalars(const BiScalars* dev, void* host_status, hipStream_t s, BiScalars* out) {
  FEM_HIP(hipMemcpyAsync(host_status, dev, sizeof(BiScalars), hipMemcpyDeviceToHost, s));
  FEM_HIP(hipStreamSynchronize(s));
  memcpy(out, host_status, sizeof(BiScalars));
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" {

int fem_csr_spmv(int64_t rows, const int32_t* ptr, const int32_t* idx, const double* val, const double* x, double* y,
                 void* stream) {
  FEM_CHECK_ARG(rows >= 0 && (rows == 0 || (ptr && x && y)), "fem_csr_spmv: bad arguments");
  if (rows == 0) return FEM_OK;
  hipLaunchKernelGGL(k_csr_spmv, dim3(grid_for(rows)), dim3(256), 0, as_stream(stream), rows, ptr, idx, val, x, y);
  FEM_LAUNCHED();
  return FEM_OK;
}

int64_t fem_bicgstab_ws_bytes(const fem_pattern_t* pat, const fem_constraint_t* P) {
  if (!pat) return 0;
  const int64_t nf = std::max<int64_t>(pat->num_rows * pat->vec, 1);
  const int64_t n = P ? std::max<int64_t>(P->cols, 1) : nf;
  size_t b = 8 * align256(n * 8) + align256(2 * kBiBlocks * 8) + align256(sizeof(BiScalars));
  if (P) b += 2 * align256(nf * 8) + align256(n * 8);
  return (int64_t)b;
}

int fem_bicgstab_solve(const fem_pattern_t* pat, const double* vals, const double* dinv, const fem_constraint_t* P,
                       const double* b, double* x, const fem_pcg_options_t* opt, void* ws, void* host_status,
                       fem_pcg_result_t* result, void* stream) {
  FEM_CHECK_ARG(pat && vals && b && x && opt && ws && host_status && result, "fem_bicgstab_solve: NULL argument");
  FEM_CHECK_ARG(opt->precision == FEM_F64, "fem_bicgstab_solve: the matrix must be FP64 (standard form)");
  FEM_CHECK_ARG(!opt->precond || dinv, "fem_bicgstab_solve: Jacobi preconditioning needs dinv");
  FEM_CHECK_ARG(opt->maxiter >= 0 && opt->check_every >= 1, "fem_bicgstab_solve: bad options");
  const int64_t nfull = pat->num_rows * pat->vec;
  FEM_CHECK_ARG(!P || (P->rows == nfull && P->cols >= 0 && P->p_ptr && P->p_idx && P->p_val && P->pt_ptr &&
                       P->pt_idx && P->pt_val),
                "fem_bicgstab_solve: constraint does not match the pattern (%lld rows, %lld dofs)",
                P ? (long long)P->rows : 0LL, (long long)nfull);
  hipStream_t s = as_stream(stream);
  const int64_t n = P ? P->cols : nfull;
  char* w = static_cast<char*>(ws);
  auto take = [&](size_t bytes) {
    char* p = w;
    w += align256(bytes);
    return p;
  };
  double* r = (double*)take(n * 8);
  double* rhat = (double*)take(n * 8);
  double* p = (double*)take(n * 8);
  double* ph = (double*)take(n * 8);
  double* v = (double*)take(n * 8);
  double* sv = (double*)take(n * 8);
  double* sh = (double*)take(n * 8);
  double* t = (double*)take(n * 8);
  double* part = (double*)take(2 * kBiBlocks * 8);
  BiScalars* sc = (BiScalars*)take(sizeof(BiScalars));
  double* full0 = P ? (double*)take(nfull * 8) : nullptr;
  double* full1 = P ? (double*)take(nfull * 8) : nullptr;
  double* dred = P ? (double*)take(n * 8) : nullptr;
  const double* Dinv = opt->precond ? dinv : nullptr;
  const unsigned g = grid_for(n);
  *result = fem_pcg_result_t{};
  if (P && Dinv) {
    hipLaunchKernelGGL(k_reduced_dinv, dim3(g), dim3(256), 0, s, n, P->pt_ptr, P->pt_idx, P->pt_val, Dinv, dred);
    FEM_LAUNCHED();
    Dinv = dred;
  }
  // y = A x, or P^T A P x for a constrained system
  auto apply = [&](const double* in, double* out) -> int {
    if (!P) return fem_spmv(pat, FEM_F64, FEM_FORM_STANDARD, vals, in, out, stream);
    if (int rc = fem_csr_spmv(nfull, P->p_ptr, P->p_idx, P->p_val, in, full0, stream)) return rc;
    if (int rc = fem_spmv(pat, FEM_F64, FEM_FORM_STANDARD, vals, full0, full1, stream)) return rc;
    return fem_csr_spmv(n, P->pt_ptr, P->pt_idx, P->pt_val, full1, out, stream);
  };
  // r = rhat = b - A x; p = v = 0
  if (int rc = apply(x, t)) return rc;
  hipLaunchKernelGGL(k_bi_resid, dim3(g), dim3(256), 0, s, n, b, t, r, rhat);
  hipLaunchKernelGGL(k_bi_zero2, dim3(g), dim3(256), 0, s, n, p, v);
  hipLaunchKernelGGL(k_bi_dot, dim3(kBiBlocks), dim3(256), 0, s, n, b, b, r, r, part, (const BiScalars*)nullptr);
  hipLaunchKernelGGL(k_bi_scalar<0>, dim3(1), dim3(256), 0, s, part, kBiBlocks, sc, opt->rtol, opt->atol,
                     opt->maxiter);
  FEM_LAUNCHED();
  BiScalars h{};
  if (int rc = read_scalars(sc, host_status, s, &h)) return rc;
  int it = 0;
  while (!h.done) {
    hipLaunchKernelGGL(k_bi_p, dim3(g), dim3(256), 0, s, n, r, v, Dinv, p, ph, (const BiScalars*)sc);
    FEM_LAUNCHED();
    if (int rc = apply(ph, v)) return rc;
    hipLaunchKernelGGL(k_bi_dot, dim3(kBiBlocks), dim3(256), 0, s, n, rhat, v, (const double*)nullptr,
                       (const double*)nullptr, part, (const BiScalars*)sc);
    hipLaunchKernelGGL(k_bi_scalar<1>, dim3(1), dim3(256), 0, s, part, kBiBlocks, sc, 0.0, 0.0, 0);
    hipLaunchKernelGGL(k_bi_s, dim3(g), dim3(256), 0, s, n, r, v, Dinv, sv, sh, (const BiScalars*)sc);
    FEM_LAUNCHED();
    if (int rc = apply(sh, t)) return rc;
    hipLaunchKernelGGL(k_bi_dot, dim3(kBiBlocks), dim3(256), 0, s, n, t, sv, t, t, part, (const BiScalars*)sc);
    hipLaunchKernelGGL(k_bi_scalar<2>, dim3(1), dim3(256), 0, s, part, kBiBlocks, sc, 0.0, 0.0, 0);
    hipLaunchKernelGGL(k_bi_xr, dim3(g), dim3(256), 0, s, n, ph, sh, sv, t, x, r, (const BiScalars*)sc);
    hipLaunchKernelGGL(k_bi_dot, dim3(kBiBlocks), dim3(256), 0, s, n, r, r, rhat, r, part, (const BiScalars*)sc);
    hipLaunchKernelGGL(k_bi_scalar<3>, dim3(1), dim3(256), 0, s, part, kBiBlocks, sc, 0.0, 0.0, 0);
    FEM_LAUNCHED();
    ++it;
    if (it % opt->check_every == 0 || it >= opt->maxiter) {
      if (int rc = read_scalars(sc, host_status, s, &h)) return rc;
      if (it >= opt->maxiter) break;
    }
  }
  // true residual |b - A x|
  if (int rc = apply(x, t)) return rc;
  hipLaunchKernelGGL(k_bi_resid, dim3(g), dim3(256), 0, s, n, b, t, r, (double*)nullptr);
  hipLaunchKernelGGL(k_bi_dot, dim3(kBiBlocks), dim3(256), 0, s, n, r, r, (const double*)nullptr,
                     (const double*)nullptr, part, (const BiScalars*)nullptr);
  FEM_LAUNCHED();
  double hp[2 * kBiBlocks];
  FEM_HIP(hipMemcpyAsync(hp, part, sizeof(hp), hipMemcpyDeviceToHost, s));
  FEM_HIP(hipStreamSynchronize(s));
  if (int rc = read_scalars(sc, host_status, s, &h)) return rc;
  double tr = 0.0;
  for (int i = 0; i < kBiBlocks; ++i) tr += hp[2 * i];
  result->iterations = h.iter;
  result->b_norm = sqrt(h.bb);
  result->res_norm = sqrt(tr);
  const double tol = std::max(opt->rtol * result->b_norm, opt->atol);
  result->status = h.nan || !std::isfinite(tr) ? FEM_ERR_NAN : (result->res_norm > tol ? FEM_ERR_NOT_CONVERGED : FEM_OK);
  return result->status;
}

}  // extern "C"
