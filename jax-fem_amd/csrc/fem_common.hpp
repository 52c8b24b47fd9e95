// Shared device/host helpers of libfem_mi355x (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "fem_mi355x.h"

namespace fem {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);

#define FEM_CHECK_ARG(cond, ...)                 \
  do {                                           \
    if (!(cond)) {                               \
      ::fem::set_error(__VA_ARGS__);             \
      return FEM_ERR_BAD_ARG;                    \
    }                                            \
  } while (0)

#define FEM_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::fem::set_error("HIP error %s at %s:%d", hipGetErrorString(e_),       \
                       __FILE__, __LINE__);                                  \
      return FEM_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

#define FEM_LAUNCHED() FEM_HIP(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kWave = 64;
constexpr int kSlice = 64;  // rows per SELL slice

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------- reductions
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum for blockDim.x == 256 (4 waves); result valid in all threads.
template <typename T>
__device__ inline T block_sum_256(T v, T* smem /* >= 4 */) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) smem[w] = v;
  __syncthreads();
  T t = (smem[0] + smem[1]) + (smem[2] + smem[3]);
  __syncthreads();
  return t;
}

// Block-wide sum for blockDim.x == NT (a multiple of 64, <= 1024); result
// valid in all threads.  smem holds >= NT/64 values.
template <int NT, typename T>
__device__ inline T block_sum(T v, T* smem) {
  constexpr int NW = NT / 64;
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) smem[w] = v;
  __syncthreads();
  T t = T(0);
#pragma unroll
  for (int i = 0; i < NW; ++i) t += smem[i];
  __syncthreads();
  return t;
}

// Slice-parallel SELL-64 row product.  A slice (64 rows) is processed by W
// waves; wave w accumulates the slots k = w, w + W, ... of its 64 rows, so
// each slice has W independent streams of coalesced value loads in flight.
// `xat(c, xc)` loads the V entries of the input vector at block column c.
// DIFF: difference form (diagonal slot holds the row-sum block; off-diagonal
// blocks multiply x_c - x_row).  `row` is the block row of this lane
// (sell_row; -1 on padding lanes, whose slots hold zeros).  The arithmetic is FP64 whatever the storage
// type T: the difference of two FP32 inputs is exact in FP64 and the products
// are accumulated without FP32 rounding (the kernel is HBM-bound, so the FP64
// FMAs cost no time).  The matrix stream (values, column ids) is read once per
// product with nontemporal loads, which keeps the gathered vector entries in
// the caches.  Returns this wave's partial sums.
// TX is the type `xat` loads (the storage type T unless stated).
template <int V, bool DIFF, int W, int U, typename T, typename XAt, typename TX = T>
__device__ inline void sell_slice_partial(const fem_pattern_t& pat, const T* __restrict__ vals, int64_t s, int w,
                                          int lane, int64_t row, const XAt& xat, const double (&xr)[V],
                                          double (&acc)[V]) {
  constexpr int VV = V * V;
  const int base = pat.slice_ptr[s];
  const int L = (pat.slice_ptr[s + 1] - base) >> 6;
  const int32_t* cp = pat.sell_col + base + lane;
  const T* vp = vals + (size_t)base * VV + lane;
#pragma unroll
  for (int i = 0; i < V; ++i) acc[i] = 0.0;
#pragma unroll U
  for (int k = w; k < L; k += W) {
    const int64_t c = __builtin_nontemporal_load(cp + (size_t)k * kSlice);
    TX xt[V];
    xat(c, xt);
    double xc[V];
#pragma unroll
    for (int j = 0; j < V; ++j) xc[j] = (DIFF && c != row) ? double(xt[j]) - xr[j] : double(xt[j]);
#pragma unroll
    for (int i = 0; i < V; ++i)
#pragma unroll
      for (int j = 0; j < V; ++j)
        acc[i] = fma(double(__builtin_nontemporal_load(vp + ((size_t)k * VV + i * V + j) * kSlice)), xc[j], acc[i]);
  }
}

// Slice split chosen by a sweep of (W, U) on the 1.6M-dof HEX8 elasticity
// matrix: FP32 4 waves x unroll 2, FP64 2 waves x unroll 2 (FP16 storage
// takes the FP32 split).
template <typename T>
struct SliceCfg {
  static constexpr int W = sizeof(T) <= 4 ? 4 : 2;
  static constexpr int U = 2;
  static constexpr int NT = 64 * W;
};

// XCD-contiguous block order: hardware dispatches block b to XCD b % 8; this
// maps the blocks of one XCD onto a contiguous range of logical block ids so
// neighbouring rows (which share gathered vector entries) stay in one L2.
__device__ inline int xcd_logical_block(int bid, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int xcd = bid & 7, local = bid >> 3;
  return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// ------------------------------------------------------------ 3x3 algebra
template <typename T>
__device__ inline T det3(const T* A) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
         A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// inverse of row-major 3x3 with given determinant
template <typename T>
__device__ inline void inv3(const T* A, T det, T* B) {
  const T id = T(1) / det;
  B[0] = (A[4] * A[8] - A[5] * A[7]) * id;
  B[1] = (A[2] * A[7] - A[1] * A[8]) * id;
  B[2] = (A[1] * A[5] - A[2] * A[4]) * id;
  B[3] = (A[5] * A[6] - A[3] * A[8]) * id;
  B[4] = (A[0] * A[8] - A[2] * A[6]) * id;
  B[5] = (A[2] * A[3] - A[0] * A[5]) * id;
  B[6] = (A[3] * A[7] - A[4] * A[6]) * id;
  B[7] = (A[1] * A[6] - A[0] * A[7]) * id;
  B[8] = (A[0] * A[4] - A[1] * A[3]) * id;
}

// ------------------------------------------------------- list windows
// A lane walking its own int32 list in order: entries come from W-entry windows
// fetched by 16-byte loads aligned to the array's address, one wide load per W
// entries.  Lanes of a wave walk different lists, so between two narrow reads of
// one lane its cache line is usually gone; a window keeps the next W entries in
// registers.  Windows that would run past [0, len) are read entry by entry.
template <int W>
struct ListWindow {
  static_assert(W % 4 == 0, "windows are whole 16-byte loads");
  const int32_t* __restrict__ a;
  int64_t len;
  int off;         // entries between the preceding 16-byte boundary and a[0]
  int64_t base;    // window start, in entries from that boundary
  int32_t v[W];

  __device__ ListWindow(const int32_t* p, int64_t n)
      : a(p), len(n), off((int)(((uintptr_t)p & 15) >> 2)), base(-1) {}

  __device__ int32_t operator[](int64_t i) {
    const int64_t s = i + off;
    const int64_t b = s & ~(int64_t)(W - 1);
    if (b != base) {
      base = b;
      const int64_t first = b - off;
      if (first >= 0 && first + W <= len) {
        const int4* q = reinterpret_cast<const int4*>(a + first);
#pragma unroll
        for (int j = 0; j < W / 4; ++j) {
          const int4 x = q[j];
          v[4 * j] = x.x;
          v[4 * j + 1] = x.y;
          v[4 * j + 2] = x.z;
          v[4 * j + 3] = x.w;
        }
      } else {
#pragma unroll
        for (int j = 0; j < W; ++j) v[j] = (first + j >= 0 && first + j < len) ? a[first + j] : 0;
      }
    }
    const int o = (int)(s - b);
    int32_t r = v[0];
#pragma unroll
    for (int j = 1; j < W; ++j) r = o == j ? v[j] : r;
    return r;
  }
};

}  // namespace fem
