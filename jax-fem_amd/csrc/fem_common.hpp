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

// "Last block" completion: every block publishes its partial, the last block
// to arrive sums all partials in index order (deterministic for a fixed grid)
// and runs `fin(total...)`.  counter must start at 0; the last block resets it.
template <int K, typename T>
__device__ inline bool last_block_sums(const T (&mine)[K], T* partials, unsigned* counter,
                                       T (&total)[K], T* smem /* >= 4*K */) {
  __shared__ bool s_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partials[(size_t)k * gridDim.x + blockIdx.x] = mine[k];
    __threadfence();
    unsigned t = atomicAdd(counter, 1u);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    T acc = 0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) {
      acc += __hip_atomic_load(&partials[(size_t)k * gridDim.x + i], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    total[k] = block_sum_256(acc, smem + 4 * k);
  }
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

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

}  // namespace fem
