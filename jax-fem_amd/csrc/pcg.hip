// Jacobi-preconditioned conjugate gradients on the SELL-64 block matrix.
//
// Replaces the default linear solve of jax-fem (solver.py:63-92, jax_solve:
// host CSR -> scipy -> BCOO conversion on every call, then BiCGSTAB with
// M = x / diag(A), tol = atol = 1e-10, maxiter 10000).  The Dirichlet rows of
// A are identity rows (solver.py:527-528) and the initial guess carries the
// exact boundary increment (solver.py:402-409), so the residual and search
// direction vanish on those rows and CG runs on the SPD free block.
//
// Five launches per iteration k:
//   P_k : p = z + beta p (row-local, in place)
//   A_k : q = A p, per-block <p, q> (one 64-row slice per block, W waves
//         splitting its slots; the product gathers p only)
//   RA_k: one block: pq = sum of A's partials, alpha = rz / pq
//   B_k : x += alpha p, r -= alpha q, z = D^-1 r, per-block <r,z>, <r,r>
//   RB_k: one block: rz, rr; convergence / NaN / maxiter test; beta = rz / rz_old
// Forming p inside the product (gathering z and p_old per column) cost ~20% of
// the product's bandwidth; the separate P pass streams 3 vectors once.
// The reductions sum per-block partials in a fixed order (bit-reproducible);
// no grid-wide fence or atomic is used.  Scalars live in device memory and
// are written by one kernel and read by the next.
//
// Precision FEM_F64: everything FP64, standard matrix form.
// FEM_F32: FP32 matrix in difference form, FP32 r, z, p, q (FP32 storage,
// FP64 arithmetic inside the product); x accumulated in FP64; the recursive
// residual is replaced by the FP64 true residual b - A x
// (A applied matrix-free in FP64) each time it has dropped by reliable_delta
// and at convergence; beta keeps the value of the recursion across a
// replacement.
#include "fem_common.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

int fem_operator_apply_impl(const fem_operator_t* op, const double* x, double* y, hipStream_t s);

namespace fem {

struct PcgScalars {
  double rz;             // <r, z> of the current iteration
  double alpha, beta;
  double rr;             // <r, r> of the current iteration
  double rr_true;        // |b - A x|^2 of the last true-residual evaluation
  double pad0[3];
  int32_t done, iter, nan, pad1;
  int32_t pad2[4];
};
static_assert(sizeof(PcgScalars) <= 128, "host status buffer");

// Sum of `n` partials starting at p, identical in every block (blockDim == 256).
__device__ inline double block_reduce_partials(const double* __restrict__ p, int n, double* smem) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += p[i];
  return block_sum_256(acc, smem);
}



// Kernel P: p = z + beta p in place (beta = 0 at a start).
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_p(int64_t n, const T* __restrict__ z, T* __restrict__ p,
                                               const PcgScalars* sc) {
  if (sc->done) return;
  const T beta = T(sc->beta);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

// Kernel A: one 64-row slice per block; the W waves split the slots
// (sell_slice_partial) and wave 0 combines them through LDS, stores
// q = A p and the block's <p, q>.
template <int V, bool DIFF, typename T>
__global__ __launch_bounds__(SliceCfg<T>::NT) void k_pcg_a(fem_pattern_t pat, const T* __restrict__ vals,
                                                           const T* __restrict__ p, T* __restrict__ q,
                                                           const PcgScalars* sc, double* __restrict__ partA,
                                                           int remap) {
  constexpr int W = SliceCfg<T>::W, U = SliceCfg<T>::U, NT = SliceCfg<T>::NT;
  __shared__ double smem[NT / 64];
  __shared__ double red[W][kSlice][V];
  if (sc->done) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t s = remap ? xcd_logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const int64_t row = pat.sell_row[s * kSlice + lane];
  const bool real = row >= 0;
  auto xat = [&](int64_t c, T (&xc)[V]) {
#pragma unroll
    for (int j = 0; j < V; ++j) xc[j] = p[c * V + j];
  };
  T pr[V];
  xat(real ? row : 0, pr);
  double xr[V], acc[V];
#pragma unroll
  for (int j = 0; j < V; ++j) xr[j] = double(pr[j]);
  sell_slice_partial<V, DIFF, W, U, T>(pat, vals, s, w, lane, row, xat, xr, acc);
  if (w > 0) {
#pragma unroll
    for (int i = 0; i < V; ++i) red[w][lane][i] = acc[i];
  }
  __syncthreads();
  double dot = 0.0;
  if (w == 0 && real) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      double t = acc[i];
#pragma unroll
      for (int o = 1; o < W; ++o) t += red[o][lane][i];
      const T qi = T(t);
      q[row * V + i] = qi;
      dot += xr[i] * double(qi);
    }
  }
  const double d = block_sum<NT>(dot, smem);
  if (threadIdx.x == 0) partA[blockIdx.x] = d;
}

// RA: alpha = rz / sum(partA).  Eight independent loads per thread are in
// flight at once (the sum is latency-bound); the order of the sum is fixed.
__global__ __launch_bounds__(1024) void k_pcg_ra(const double* __restrict__ partA, int nA, PcgScalars* sc) {
  __shared__ double smem[16];
  double a[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) a[u] = 0.0;
  for (int base = 0; base < nA; base += 8 * 1024) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = base + u * 1024 + (int)threadIdx.x;
      if (i < nA) a[u] += partA[i];
    }
  }
  const double t = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  const double pq = block_sum<1024>(t, smem);
  if (sc->done) return;
  if (threadIdx.x == 0) sc->alpha = sc->rz / pq;
}

// RB: rz, rr of iteration k+1; stop test; beta.  nB <= 1024.
__global__ __launch_bounds__(1024) void k_pcg_rb(const double* __restrict__ partB, int nB, PcgScalars* sc, int k,
                                                 double tol2, int maxiter) {
  __shared__ double smem[16];
  const int i = threadIdx.x;
  const double a0 = i < nB ? partB[i] : 0.0;
  const double a1 = i < nB ? partB[nB + i] : 0.0;
  const double rz = block_sum<1024>(a0, smem);
  const double rr = block_sum<1024>(a1, smem);
  if (sc->done) return;
  if (threadIdx.x == 0) {
    const bool bad = !isfinite(rr) || !isfinite(rz);
    sc->beta = rz / sc->rz;
    sc->rz = rz;
    sc->rr = rr;
    sc->iter = k + 1;
    if (bad) sc->nan = 1;
    if (bad || rr <= tol2 || k + 1 >= maxiter) sc->done = 1;
  }
}

template <typename T, typename TX>
__global__ __launch_bounds__(256) void k_pcg_b(int64_t n, TX* __restrict__ x, T* __restrict__ r,
                                               T* __restrict__ z, const T* __restrict__ p,
                                               const T* __restrict__ q, const T* __restrict__ dinv,
                                               const PcgScalars* sc, double* __restrict__ partB) {
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
  const double s1 = block_sum_256(double(rz), smem);
  const double s2 = block_sum_256(double(rr), smem + 4);
  if (threadIdx.x == 0) {
    partB[blockIdx.x] = s1;
    partB[gridDim.x + blockIdx.x] = s2;
  }
}

// r = (T)(b - y) (y = NULL: r = b); per-block partials of |b - y|^2.
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_resid(int64_t n, const double* __restrict__ b, const double* __restrict__ y,
                                                   T* __restrict__ r, double* __restrict__ part) {
  __shared__ double smem[8];
  double rr = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double ri = b[i] - (y ? y[i] : 0.0);
    r[i] = T(ri);
    rr += ri * ri;
  }
  const double s = block_sum_256(rr, smem);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_pcg_finish_rr(const double* __restrict__ part, int n, PcgScalars* sc) {
  __shared__ double smem[8];
  const double s = block_reduce_partials(part, n, smem);
  if (threadIdx.x == 0) sc->rr_true = s;
}

// z = D^-1 r and the partials of <r,z>, <r,r> (start or residual replacement).
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_restart(int64_t n, const T* __restrict__ r, T* __restrict__ z,
                                                     const T* __restrict__ dinv, double* __restrict__ partR) {
  __shared__ double smem[8];
  T rz = T(0), rr = T(0);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T zi = dinv[i] * r[i];
    z[i] = zi;
    rz += r[i] * zi;
    rr += r[i] * r[i];
  }
  const double s1 = block_sum_256(double(rz), smem);
  const double s2 = block_sum_256(double(rr), smem + 4);
  if (threadIdx.x == 0) {
    partR[blockIdx.x] = s1;
    partR[gridDim.x + blockIdx.x] = s2;
  }
}

// Start (init = 1: beta = 0, counters reset) or residual replacement at
// iteration k (beta keeps the recursion's value; only rz and rr change).
__global__ __launch_bounds__(1024) void k_pcg_rr(const double* __restrict__ partR, int n, PcgScalars* sc, int init,
                                                 int k) {
  __shared__ double smem[16];
  double a0 = 0.0, a1 = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) {
    a0 += partR[i];
    a1 += partR[n + i];
  }
  const double rz = block_sum<1024>(a0, smem);
  const double rr = block_sum<1024>(a1, smem);
  if (threadIdx.x == 0) {
    sc->rz = rz;
    sc->rr = rr;
    sc->done = 0;
    sc->iter = k;
    if (init) {
      sc->beta = 0.0;
      sc->nan = 0;
    }
  }
}

template <typename T>
__global__ void k_fill(int64_t n, T* __restrict__ v, T val) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = val;
}

// y = A x for FP64 x, y (standard form values of precision T).
template <int V, typename T>
__global__ __launch_bounds__(256) void k_spmv_x(fem_pattern_t pat, const T* __restrict__ vals,
                                                const double* __restrict__ x, double* __restrict__ y) {
  constexpr int VV = V * V;
  const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (row >= pat.num_rows) return;
  const int64_t pos = pat.sell_pos[row];
  const int64_t s = pos >> 6;
  const int lane = (int)(pos & 63);
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

constexpr int kVecGrid = 1024;  // blocks of the vector kernels (partials per reduction)

struct EventPool {
  std::vector<hipEvent_t> e;
  bool ok = true;
  explicit EventPool(int n) : e(n, nullptr) {
    for (auto& x : e)
      if (hipEventCreate(&x) != hipSuccess) ok = false;
  }
  ~EventPool() {
    for (auto& x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

struct Ws {
  void* r;
  void* z;
  void* p;
  void* q;
  void* ones;
  double* y;
  double* partA;
  double* partB;
  double* partR;
  double* partT;
  PcgScalars* sc;
};

static int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

static int64_t grid_a(const fem_pattern_t* pat) { return pat->num_slices; }

static int64_t ws_layout(const fem_pattern_t* pat, int precision, void* base, Ws* w) {
  const int64_t n = pat->num_rows * pat->vec;
  const int64_t es = precision == FEM_F32 ? 4 : 8;
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
  o->p = take(n * es);
  o->q = take(n * es);
  o->ones = take(n * es);
  o->y = reinterpret_cast<double*>(take(n * 8));
  o->partA = reinterpret_cast<double*>(take(grid_a(pat) * 8));
  o->partB = reinterpret_cast<double*>(take(2 * kVecGrid * 8));
  o->partR = reinterpret_cast<double*>(take(2 * kVecGrid * 8));
  o->partT = reinterpret_cast<double*>(take(kVecGrid * 8));
  o->sc = reinterpret_cast<PcgScalars*>(take(sizeof(PcgScalars)));
  return off;
}

template <int V, typename T>
static int pcg_run(const fem_pattern_t* pat, const T* vals, const T* dinv_in, const fem_operator_t* op,
                   const double* b, double* x, const fem_pcg_options_t* opt, void* ws_base,
                   PcgScalars* host_sc, fem_pcg_result_t* res, hipStream_t s) {
  constexpr bool kMixed = std::is_same<T, float>::value;
  constexpr bool kDiff = kMixed;  // the FP32 matrix is stored in difference form
  const int64_t n = pat->num_rows * V;
  Ws w;
  ws_layout(pat, kMixed ? FEM_F32 : FEM_F64, ws_base, &w);
  T* r = (T*)w.r;
  T* z = (T*)w.z;
  T* p = (T*)w.p;
  T* q = (T*)w.q;
  const T* dinv = dinv_in;
  if (!opt->precond || !dinv) {
    hipLaunchKernelGGL(k_fill<T>, dim3(256), dim3(256), 0, s, n, (T*)w.ones, T(1));
    FEM_LAUNCHED();
    dinv = (T*)w.ones;
  }
  const int gA = (int)grid_a(pat);
  const int gV = (int)std::min<int64_t>(ceil_div(n, 256), kVecGrid);
  FEM_HIP(hipMemsetAsync(w.sc, 0, sizeof(PcgScalars), s));
  hipLaunchKernelGGL(k_fill<T>, dim3(256), dim3(256), 0, s, n, p, T(0));
  FEM_LAUNCHED();

  auto read_status = [&]() -> int {
    FEM_HIP(hipMemcpyAsync(host_sc, w.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    FEM_HIP(hipStreamSynchronize(s));
    return FEM_OK;
  };
  // r = b - y (y = NULL: r = b), |r|^2 -> rr_true
  auto residual_from = [&](const double* y) -> int {
    hipLaunchKernelGGL(k_pcg_resid<T>, dim3(gV), dim3(256), 0, s, n, b, y, r, w.partT);
    hipLaunchKernelGGL(k_pcg_finish_rr, dim3(1), dim3(256), 0, s, (const double*)w.partT, gV, w.sc);
    FEM_LAUNCHED();
    return read_status();
  };
  auto true_residual = [&]() -> int {
    if (kMixed) {
      if (int rc = fem_operator_apply_impl(op, x, w.y, s)) return rc;
    } else {
      hipLaunchKernelGGL((k_spmv_x<V, T>), dim3((unsigned)ceil_div(pat->num_rows, 256)), dim3(256), 0, s, *pat, vals,
                         (const double*)x, w.y);
      FEM_LAUNCHED();
    }
    return residual_from(w.y);
  };

  if (int rc = residual_from(nullptr)) return rc;  // |b|^2
  const double bb = host_sc->rr_true;
  const double tol2 = std::max(opt->rtol * opt->rtol * bb, opt->atol * opt->atol);
  res->b_norm = std::sqrt(bb);
  res->iterations = 0;
  res->replacements = 0;
  res->status = FEM_OK;
  res->timed_launches = 0;
  res->spmv_ms = res->update_ms = 0.0;

  if (int rc = true_residual()) return rc;
  double rr_true = host_sc->rr_true;
  if (!std::isfinite(rr_true)) {
    res->status = FEM_ERR_NAN;
    res->res_norm = rr_true;
    return FEM_OK;
  }
  if (rr_true <= tol2) {
    res->res_norm = std::sqrt(rr_true);
    return FEM_OK;
  }
  hipLaunchKernelGGL(k_pcg_restart<T>, dim3(gV), dim3(256), 0, s, n, (const T*)r, z, dinv, w.partR);
  hipLaunchKernelGGL(k_pcg_rr, dim3(1), dim3(1024), 0, s, (const double*)w.partR, gV, w.sc, 1, 0);
  FEM_LAUNCHED();
  double rr_ref = rr_true;
  const int check = std::max(1, opt->check_every);
  const double delta2 = kMixed ? opt->reliable_delta * opt->reliable_delta : 0.0;
  EventPool ev(opt->timing ? 4 * check : 0);
  if (!ev.ok) {
    set_error("fem_pcg_solve: hipEventCreate failed");
    return FEM_ERR_HIP;
  }
  double spmv_ms = 0.0, upd_ms = 0.0;
  int timed = 0;
  int iter = 0;
  for (;;) {
    const int iter0 = iter;
    for (int m = 0; m < check; ++m) {
      const int k = iter0 + m;
      if (opt->timing) FEM_HIP(hipEventRecord(ev.e[4 * m], s));
      hipLaunchKernelGGL(k_pcg_p<T>, dim3(gV), dim3(256), 0, s, n, (const T*)z, p, (const PcgScalars*)w.sc);
      if (opt->timing) FEM_HIP(hipEventRecord(ev.e[4 * m + 1], s));
      hipLaunchKernelGGL((k_pcg_a<V, kDiff, T>), dim3(gA), dim3(SliceCfg<T>::NT), 0, s, *pat, vals, (const T*)p, q,
                         (const PcgScalars*)w.sc, w.partA, opt->xcd_remap);
      if (opt->timing) FEM_HIP(hipEventRecord(ev.e[4 * m + 2], s));
      hipLaunchKernelGGL(k_pcg_ra, dim3(1), dim3(1024), 0, s, (const double*)w.partA, gA, w.sc);
      hipLaunchKernelGGL((k_pcg_b<T, double>), dim3(gV), dim3(256), 0, s, n, x, r, z, (const T*)p, (const T*)q, dinv,
                         (const PcgScalars*)w.sc, w.partB);
      hipLaunchKernelGGL(k_pcg_rb, dim3(1), dim3(1024), 0, s, (const double*)w.partB, gV, w.sc, k, tol2,
                         opt->maxiter);
      if (opt->timing) FEM_HIP(hipEventRecord(ev.e[4 * m + 3], s));
    }
    FEM_LAUNCHED();
    if (int rc = read_status()) return rc;
    iter = host_sc->iter;
    if (opt->timing) {
      // only the launches that ran an iteration (the rest returned on `done`)
      for (int m = 0; m < std::min(check, iter - iter0); ++m) {
        float t0 = 0.f, t1 = 0.f, t2 = 0.f;
        FEM_HIP(hipEventElapsedTime(&t0, ev.e[4 * m], ev.e[4 * m + 1]));
        FEM_HIP(hipEventElapsedTime(&t1, ev.e[4 * m + 1], ev.e[4 * m + 2]));
        FEM_HIP(hipEventElapsedTime(&t2, ev.e[4 * m + 2], ev.e[4 * m + 3]));
        spmv_ms += t1;
        upd_ms += t0 + t2;
        ++timed;
      }
    }
    if (host_sc->nan) {
      res->status = FEM_ERR_NAN;
      break;
    }
    const bool conv_rec = host_sc->done && host_sc->rr <= tol2;
    const bool trigger = kMixed && !host_sc->done && host_sc->rr < delta2 * rr_ref;
    if (conv_rec || trigger) {
      if (int rc = true_residual()) return rc;
      rr_true = host_sc->rr_true;
      if (rr_true <= tol2) break;
      if (!std::isfinite(rr_true)) {
        res->status = FEM_ERR_NAN;
        break;
      }
      if (iter >= opt->maxiter) {
        res->status = FEM_ERR_NOT_CONVERGED;
        break;
      }
      // residual replacement at iteration `iter`: keep the search direction
      hipLaunchKernelGGL(k_pcg_restart<T>, dim3(gV), dim3(256), 0, s, n, (const T*)r, z, dinv, w.partR);
      hipLaunchKernelGGL(k_pcg_rr, dim3(1), dim3(1024), 0, s, (const double*)w.partR, gV, w.sc, 0, iter);
      FEM_LAUNCHED();
      rr_ref = rr_true;
      res->replacements += 1;
      continue;
    }
    if (host_sc->done || iter >= opt->maxiter) {  // maxiter reached
      res->status = FEM_ERR_NOT_CONVERGED;
      break;
    }
  }
  if (!(rr_true <= tol2) || res->status != FEM_OK) {
    if (int rc = true_residual()) return rc;
    rr_true = host_sc->rr_true;
  }
  res->iterations = iter;
  res->res_norm = std::sqrt(rr_true);
  res->timed_launches = timed;
  res->spmv_ms = timed ? spmv_ms / timed : 0.0;
  res->update_ms = timed ? upd_ms / timed : 0.0;
  return FEM_OK;
}

// ------------------------------------------------- no-allocation variant
// The workspace of fem_pcg_solve_noalloc: a library-held device array, one
// copy per GPU (each device that loads this code object holds it until the
// process ends).  Host code never passes `scratch` itself to a kernel: its
// device address comes from hipGetSymbolAddress, looked up once per device.
static __device__ __align__(256) unsigned char scratch[FEM_PCG_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array.
static int scratch_address(void** out) {
  int dev = -1;
  FEM_HIP(hipGetDevice(&dev));
  FEM_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "fem_pcg_solve_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    FEM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    FEM_CHECK_ARG(p != nullptr, "fem_pcg_solve_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
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
  FEM_CHECK_ARG(pat && pat->slice_ptr && pat->sell_col && pat->sell_row && pat->sell_pos && pat->num_rows > 0,
                "fem_pcg_solve: bad pattern");
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
  if (V == 2 && opt->precision == FEM_F32)
    return pcg_run<2, float>(pat, (const float*)vals, (const float*)dinv, op, b, x, opt, ws, hsc, result, s);
  if (V == 2 && opt->precision == FEM_F64)
    return pcg_run<2, double>(pat, (const double*)vals, (const double*)dinv, op, b, x, opt, ws, hsc, result, s);
  if (V == 4 && opt->precision == FEM_F32)
    return pcg_run<4, float>(pat, (const float*)vals, (const float*)dinv, op, b, x, opt, ws, hsc, result, s);
  if (V == 4 && opt->precision == FEM_F64)
    return pcg_run<4, double>(pat, (const double*)vals, (const double*)dinv, op, b, x, opt, ws, hsc, result, s);
  set_error("fem_pcg_solve: unsupported vec %d", V);
  return FEM_ERR_BAD_ARG;
}

int fem_pcg_solve_noalloc(const fem_pattern_t* pat, const void* vals, const void* dinv, const fem_operator_t* op,
                          const double* b, double* x, const fem_pcg_options_t* opt, void* host_status,
                          fem_pcg_result_t* result, void* stream) {
  FEM_CHECK_ARG(pat && opt, "fem_pcg_solve_noalloc: NULL argument");
  FEM_CHECK_ARG(opt->timing == 0, "fem_pcg_solve_noalloc: timing=1 creates HIP events; use fem_pcg_solve");
  const int64_t need = fem_pcg_ws_bytes(pat, opt->precision);
  FEM_CHECK_ARG(need > 0, "fem_pcg_solve_noalloc: bad pattern");
  if (need > (int64_t)FEM_PCG_NOALLOC_SCRATCH_BYTES) {
    set_error("fem_pcg_solve_noalloc: the system needs %lld workspace bytes, more than the %lld-byte scratch "
              "(%lld rows of %d); use fem_pcg_solve with a caller workspace",
              (long long)need, (long long)FEM_PCG_NOALLOC_SCRATCH_BYTES, (long long)pat->num_rows, pat->vec);
    return FEM_ERR_CAPACITY;
  }
  void* ws = nullptr;
  if (int rc = scratch_address(&ws)) return rc;
  // every workspace entry the solve reads is written by it first (scalars and p are
  // cleared at its start), so the previous call's contents do not matter
  return fem_pcg_solve(pat, vals, dinv, op, b, x, opt, ws, host_status, result, stream);
}

}  // extern "C"
