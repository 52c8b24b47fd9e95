// Smoothed-aggregation algebraic multigrid on the device.
//
// The reference's `amgx_solver` option (solver.py:141-259, chosen in
// linear_solver, solver.py:261-284) hands the assembled PETSc matrix to AMGX.
// Here the hierarchy is built on the GPU from the assembled SELL-64 matrix and
// used as the preconditioner of FP64 CG:
//   setup  MIS-2 aggregation by max-propagation rounds; tentative prolongator
//          by modified Gram-Schmidt per aggregate (one wave per aggregate);
//          P = (I - omega D^-1 A) P0; Galerkin R (A P) by a block SpGEMM that
//          sorts each output row's block products in LDS (bitonic) and sums
//          every output block in sorted order; transpose by counting sort;
//          Gauss-Jordan inverse of the coarsest operator
//   cycle  V-cycle with Chebyshev smoothing (SELL-64 kernels on level 0, block
//          CSR kernels below), dense GEMV on the coarsest level
//   solve  CG whose scalars stay on the device; the host reads a status block
//          every `check_every` iterations
// No floating-point atomics: every result is bit-reproducible run to run.
#include "fem_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace fem {
namespace {

constexpr int kXCap = 1024;     // max blocks in one row of the left SpGEMM operand (sorted-pair kernel)
constexpr int kDotBlocks = 1024;
constexpr double kRankTol = 1e-10;

// ------------------------------------------------------------- priorities
__host__ __device__ inline uint32_t amg_hash(uint32_t x) {
  x *= 0x9E3779B1u;
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}
// unique 61-bit priority of node i (i < 2^31)
__device__ inline uint64_t amg_key(int64_t i) {
  return ((uint64_t)(amg_hash((uint32_t)i) >> 2) << 31) | (uint64_t)i;
}
// (rank, key): rank 2 = root, 1 = undecided, 0 = decided not root
__device__ inline uint64_t mis_tuple(int8_t st, int64_t j) {
  const uint64_t rank = st == 1 ? 2u : (st == 0 ? 1u : 0u);
  return (rank << 61) | amg_key(j);
}

inline unsigned grid_for(int64_t n, int nt) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, nt), 1 << 20));
}

// --------------------------------------------------------------- MIS-2
__global__ void k_mis_init(int64_t n, int8_t* __restrict__ state) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    state[i] = 0;
}

// t1[i] = max over the neighbourhood of i of the (rank, key) tuples.
__global__ void k_mis_prop1(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                            const int8_t* __restrict__ state, uint64_t* __restrict__ t1, int32_t* __restrict__ flag) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t m = 0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      const int64_t j = ci[e];
      m = max(m, mis_tuple(state[j], j));
    }
    t1[i] = m;
  }
}

// t2 = max of t1 over the neighbourhood (= the distance-2 maximum): an
// undecided node that is its own maximum becomes a root, one that sees a root
// within distance 2 is decided; *flag = 1 while undecided nodes remain.
__global__ void k_mis_prop2(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                            const uint64_t* __restrict__ t1, int8_t* __restrict__ state, int32_t* __restrict__ flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (state[i] != 0) continue;
    uint64_t m = 0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) m = max(m, t1[ci[e]]);
    if (m == mis_tuple(0, i))
      state[i] = 1;
    else if ((m >> 61) == 2)
      state[i] = 2;
    else
      atomicOr(flag, 1);
  }
}

__global__ void k_root_flags(int64_t n, const int8_t* __restrict__ state, int32_t* __restrict__ f) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    f[i] = state[i] == 1 ? 1 : 0;
}

// roots and their neighbours (roots are >= 3 apart: at most one root neighbour)
__global__ void k_agg_pass1(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                            const int8_t* __restrict__ state, const int32_t* __restrict__ root_id,
                            int32_t* __restrict__ agg1) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t a = -1;
    if (state[i] == 1) {
      a = root_id[i];
    } else {
      for (int e = rp[i]; e < rp[i + 1]; ++e) {
        const int j = ci[e];
        if (state[j] == 1) {
          a = root_id[j];
          break;
        }
      }
    }
    agg1[i] = a;
  }
}

// distance-2 nodes join the assigned neighbour of largest priority
__global__ void k_agg_pass2(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                            const int32_t* __restrict__ agg1, int32_t* __restrict__ agg, int32_t* __restrict__ err) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t a = agg1[i];
    if (a < 0) {
      uint64_t best = 0;
      for (int e = rp[i]; e < rp[i + 1]; ++e) {
        const int j = ci[e];
        if (agg1[j] >= 0) {
          const uint64_t k = amg_key(j) + 1;
          if (k > best) {
            best = k;
            a = agg1[j];
          }
        }
      }
      if (a < 0) atomicOr(err, 1);
    }
    agg[i] = a;
  }
}

// -------------------------------------------------------------- members
__global__ void k_count_keys(int64_t n, const int32_t* __restrict__ key, int32_t* __restrict__ cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[key[i]], 1);
}

__global__ void k_fill_members(int64_t n, const int32_t* __restrict__ agg, const int32_t* __restrict__ ptr,
                               int32_t* __restrict__ cursor, int32_t* __restrict__ memb) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int a = agg[i];
    memb[ptr[a] + atomicAdd(&cursor[a], 1)] = (int32_t)i;
  }
}

// insertion sort of each segment of v (segments are short); `aux` (may be
// NULL) is permuted alongside
__global__ void k_sort_segments(int64_t m, const int32_t* __restrict__ ptr, int32_t* __restrict__ v,
                                int32_t* __restrict__ aux) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < m; s += (int64_t)gridDim.x * blockDim.x) {
    const int b = ptr[s], e = ptr[s + 1];
    for (int i = b + 1; i < e; ++i) {
      const int32_t key = v[i];
      const int32_t ax = aux ? aux[i] : 0;
      int j = i - 1;
      while (j >= b && v[j] > key) {
        v[j + 1] = v[j];
        if (aux) aux[j + 1] = aux[j];
        --j;
      }
      v[j + 1] = key;
      if (aux) aux[j + 1] = ax;
    }
  }
}

// ------------------------------------------------------------ tentative
// One wave per aggregate.  The aggregate's rows of the near null space are
// written into P0 and orthonormalised there (each lane owns rows r = lane,
// lane + 64, ...; the column dot products are wave sums).
template <int BS, int NS>
__global__ __launch_bounds__(256) void k_tentative(int64_t n_agg, const int32_t* __restrict__ agg_ptr,
                                                   const int32_t* __restrict__ memb, const double* __restrict__ points,
                                                   const uint8_t* __restrict__ bc_mask, const double* __restrict__ B,
                                                   double* __restrict__ P0, double* __restrict__ Bc) {
  const int lane = threadIdx.x & 63;
  const int64_t a = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (a >= n_agg) return;  // wave-uniform
  const int b = agg_ptr[a];
  const int rows = (agg_ptr[a + 1] - b) * BS;
  const int first = memb[b];
  for (int r = lane; r < rows; r += 64) {
    const int node = memb[b + r / BS], comp = r % BS;
    double v[NS];
    if (B) {
#pragma unroll
      for (int j = 0; j < NS; ++j) v[j] = B[((int64_t)node * BS + comp) * NS + j];
    } else if (NS == 1) {
      v[0] = 1.0;
    } else {
      const double x = points[(int64_t)node * 3 + 0] - points[(int64_t)first * 3 + 0];
      const double y = points[(int64_t)node * 3 + 1] - points[(int64_t)first * 3 + 1];
      const double z = points[(int64_t)node * 3 + 2] - points[(int64_t)first * 3 + 2];
#pragma unroll
      for (int j = 0; j < NS; ++j) v[j] = 0.0;
      v[comp] = 1.0;
      // rotations: (-y, x, 0), (0, -z, y), (z, 0, -x)
      if (comp == 0) {
        v[3] = -y;
        v[5] = z;
      } else if (comp == 1) {
        v[3] = x;
        v[4] = -z;
      } else {
        v[4] = y;
        v[5] = -x;
      }
    }
    if (!B && bc_mask && bc_mask[(int64_t)node * BS + comp])
#pragma unroll
      for (int j = 0; j < NS; ++j) v[j] = 0.0;
    double* dst = P0 + ((int64_t)node * BS + comp) * NS;
#pragma unroll
    for (int j = 0; j < NS; ++j) dst[j] = v[j];
  }
  double R[NS][NS];
#pragma unroll
  for (int k = 0; k < NS; ++k)
#pragma unroll
    for (int j = 0; j < NS; ++j) R[k][j] = 0.0;
  auto at = [&](int r, int j) -> double& {
    return P0[((int64_t)memb[b + r / BS] * BS + r % BS) * NS + j];
  };
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    double s = 0.0;
    for (int r = lane; r < rows; r += 64) s += at(r, j) * at(r, j);
    const double n0 = sqrt(wave_sum(s));
#pragma unroll
    for (int k = 0; k < j; ++k) {
      double d = 0.0;
      for (int r = lane; r < rows; r += 64) d += at(r, k) * at(r, j);
      d = wave_sum(d);
      R[k][j] = d;
      for (int r = lane; r < rows; r += 64) at(r, j) -= d * at(r, k);
    }
    s = 0.0;
    for (int r = lane; r < rows; r += 64) s += at(r, j) * at(r, j);
    const double nv = sqrt(wave_sum(s));
    const bool keep = nv > kRankTol * n0 && nv > 0.0;
    const double scale = keep ? 1.0 / nv : 0.0;
    R[j][j] = keep ? nv : 0.0;
    for (int r = lane; r < rows; r += 64) at(r, j) *= scale;
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k)
#pragma unroll
      for (int j = 0; j < NS; ++j) Bc[(a * NS + k) * NS + j] = R[k][j];
  }
}

__global__ void k_bc_mask(int64_t n, uint8_t* __restrict__ mask) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    mask[i] = 0;
}
__global__ void k_bc_mask_set(int64_t n_bc, const int32_t* __restrict__ dofs, uint8_t* __restrict__ mask) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_bc; i += (int64_t)gridDim.x * blockDim.x)
    mask[dofs[i]] = 1;
}

// ------------------------------------------------------ matrix helpers
// One block per slice: KC slots of the slice's 64 rows are read coalesced into
// LDS, then each row's KC*V*V entries (contiguous in the block CSR) are stored
// as one contiguous run, so the stores fill whole cache lines.  DIFF: the same
// pass also writes the FP32 difference-form SELL copy (dout): off-diagonal
// entries rounded as they are read, and each row's FP64 row sums (summed over
// the slots in slot order by the thread that owns the (row, entry) pair)
// rounded once into the diagonal slot; with dinv, also 1 / the rounded diagonal
// entries (0 for a zero entry), as fem_amg_dinv forms them from the copy.
template <int V, typename TO, bool DIFF = false>
__global__ __launch_bounds__(256) void k_sell_to_bcsr(fem_pattern_t pat, const double* __restrict__ sv,
                                                      TO* __restrict__ out, float* __restrict__ dout,
                                                      double* __restrict__ dinv) {
  constexpr int VV = V * V, KC = 4;
  constexpr int NOWN = (kSlice * VV + 255) / 256;  // (row, entry) sums owned per thread
  __shared__ double tile[kSlice][KC * VV + 1];
  __shared__ int s_dst[kSlice], s_len[kSlice], s_ds[kSlice], s_row[kSlice];
  const int64_t s = blockIdx.x;
  const int base = pat.slice_ptr[s];
  const int L = (pat.slice_ptr[s + 1] - base) >> 6;
  if (threadIdx.x < kSlice) {
    const int row = pat.sell_row[s * kSlice + threadIdx.x];
    s_row[threadIdx.x] = row;
    s_dst[threadIdx.x] = row >= 0 ? pat.row_ptr[row] : 0;
    s_len[threadIdx.x] = row >= 0 ? pat.row_ptr[row + 1] - pat.row_ptr[row] : 0;
    s_ds[threadIdx.x] = row >= 0 ? pat.diag_slot[row] : -1;
  }
  const double* src = sv + (size_t)base * VV;
  float* dst = DIFF ? dout + (size_t)base * VV : nullptr;
  double rs[NOWN], dg[NOWN];  // row sums; diagonal entries (of the owned pairs on the block diagonal)
#pragma unroll
  for (int o = 0; o < NOWN; ++o) rs[o] = dg[o] = 0.0;
  for (int k0 = 0; k0 < L; k0 += KC) {
    const int nc = min(KC, L - k0) * VV;  // entries per row in this chunk
    __syncthreads();                      // row data ready, previous tile consumed
    for (int f = threadIdx.x; f < nc * kSlice; f += 256) {
      const int lane = f & (kSlice - 1), c = f >> 6;  // c = slot-in-chunk * VV + entry
      const size_t at = ((size_t)k0 * VV + c) * kSlice + lane;
      const double x = src[at];
      tile[lane][c] = x;
      if (DIFF && k0 + c / VV != s_ds[lane]) dst[at] = float(x);
    }
    __syncthreads();
    for (int f = threadIdx.x; f < nc * kSlice; f += 256) {
      const int lane = f / nc, c = f - lane * nc;
      if (k0 + c / VV < s_len[lane]) out[((size_t)s_dst[lane] + k0) * VV + c] = TO(tile[lane][c]);
    }
    if (DIFF) {
#pragma unroll
      for (int o = 0; o < NOWN; ++o) {
        const int p = threadIdx.x + 256 * o;
        if (p < kSlice * VV) {
          const int lane = p & (kSlice - 1), v = p >> 6;
          for (int kk = 0; kk < nc / VV; ++kk) rs[o] += tile[lane][kk * VV + v];
          const int kd = s_ds[lane] - k0;
          if (v % (V + 1) == 0 && kd >= 0 && kd < nc / VV) dg[o] = tile[lane][kd * VV + v];
        }
      }
    }
  }
  if (DIFF) {
#pragma unroll
    for (int o = 0; o < NOWN; ++o) {
      const int p = threadIdx.x + 256 * o;
      if (p < kSlice * VV) {
        const int lane = p & (kSlice - 1), v = p >> 6;
        if (s_ds[lane] >= 0) dst[((size_t)s_ds[lane] * VV + v) * kSlice + lane] = float(rs[o]);
        if (dinv && v % (V + 1) == 0 && s_row[lane] >= 0) {
          const double d = double(TO(dg[o]));
          dinv[(int64_t)s_row[lane] * V + v / (V + 1)] = d != 0.0 ? 1.0 / d : 0.0;
        }
      }
    }
  }
}

// A row without a diagonal block (not produced by the Galerkin products)
// gets dinv = 0, i.e. no smoothing on it.
// FP32 difference-form copy of a standard-form FP64 SELL matrix: off-diagonal
// blocks rounded, the diagonal slot holds the FP64 row sum rounded once.
// FP16 values of D^-1 A are stored times 2^8: entries down to 2.4e-7 stay normal
// FP16 numbers (smaller couplings would lose bits as subnormals), entries up to
// 255 stay finite; the kernels multiply the products back by 2^-8 (exact).
constexpr double kF16Scale = 256.0;

// Difference-form copy of a standard-form SELL matrix in TO: off-diagonal blocks
// rounded from FP64, the diagonal slot the FP64 row sum rounded once; with
// dinv, each block row a is scaled by dinv[row*V + a] first.
template <int V, typename TO>
__device__ inline void k_sell_to_diff_body(const fem_pattern_t& pat, const double* __restrict__ sv,
                                           const double* __restrict__ dinv, TO* __restrict__ out) {
  constexpr int VV = V * V;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= pat.num_slices * kSlice) return;
  const int64_t s = t >> 6;
  const int lane = (int)(t & 63);
  const int base = pat.slice_ptr[s];
  const int L = (pat.slice_ptr[s + 1] - base) >> 6;
  const int64_t row = pat.sell_row[t];
  const int ds = row >= 0 ? pat.diag_slot[row] : -1;
  const double* vp = sv + (size_t)base * VV + lane;
  TO* op = out + (size_t)base * VV + lane;
  double sc[V];
#pragma unroll
  for (int a = 0; a < V; ++a) sc[a] = (dinv && row >= 0) ? kF16Scale * dinv[row * V + a] : 1.0;
  double rs[VV];
#pragma unroll
  for (int v = 0; v < VV; ++v) rs[v] = 0.0;
  for (int k = 0; k < L; ++k)
#pragma unroll
    for (int v = 0; v < VV; ++v) {
      const double x = vp[((size_t)k * VV + v) * kSlice] * sc[v / V];
      rs[v] += x;
      if (k != ds) op[((size_t)k * VV + v) * kSlice] = (TO)x;
    }
  if (ds >= 0)
#pragma unroll
    for (int v = 0; v < VV; ++v) op[((size_t)ds * VV + v) * kSlice] = (TO)rs[v];
}

template <int V>
__global__ void k_sell_to_diff32(fem_pattern_t pat, const double* __restrict__ sv, float* __restrict__ out) {
  k_sell_to_diff_body<V, float>(pat, sv, nullptr, out);
}

// FP16 copy of D^-1 A in difference form (rows scaled by dinv before the rounding)
template <int V>
__global__ void k_sell_to_f16_scaled(fem_pattern_t pat, const double* __restrict__ sv, const double* __restrict__ dinv,
                                     _Float16* __restrict__ out) {
  k_sell_to_diff_body<V, _Float16>(pat, sv, dinv, out);
}

template <int BS, typename T>
__global__ void k_bcsr_dinv(int64_t rows, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                            T* __restrict__ v, int fix, double* __restrict__ dinv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows; i += (int64_t)gridDim.x * blockDim.x) {
    int k = -1;
    for (int e = rp[i]; e < rp[i + 1]; ++e)
      if (ci[e] == i) {
        k = e;
        break;
      }
    if (k < 0) {
#pragma unroll
      for (int c = 0; c < BS; ++c) dinv[i * BS + c] = 0.0;
      continue;
    }
#pragma unroll
    for (int c = 0; c < BS; ++c) {
      T& d = v[((size_t)k * BS + c) * BS + c];
      if (fix && d == T(0)) d = T(1);
      dinv[i * BS + c] = d != T(0) ? 1.0 / double(d) : 0.0;
    }
  }
}

// ------------------------------------------------------------ transpose
__global__ void k_count_cols(int64_t nnzb, const int32_t* __restrict__ ci, int32_t* __restrict__ cnt) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnzb; e += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[ci[e]], 1);
}

__global__ void k_t_fill(int64_t rows, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                         const int32_t* __restrict__ trp, int32_t* __restrict__ cursor, int32_t* __restrict__ tci,
                         int32_t* __restrict__ src) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows; i += (int64_t)gridDim.x * blockDim.x)
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
      const int c = ci[e];
      const int q = trp[c] + atomicAdd(&cursor[c], 1);
      tci[q] = (int32_t)i;
      src[q] = e;
    }
}

template <int BR, int BC, typename T>
__global__ void k_t_values(int64_t nnzb, const int32_t* __restrict__ src, const T* __restrict__ xv,
                           T* __restrict__ tv) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnzb; q += (int64_t)gridDim.x * blockDim.x) {
    const T* s = xv + (size_t)src[q] * BR * BC;
    T* d = tv + (size_t)q * BR * BC;
#pragma unroll
    for (int i = 0; i < BR; ++i)
#pragma unroll
      for (int j = 0; j < BC; ++j) d[j * BR + i] = s[i * BC + j];
  }
}

// --------------------------------------------------------------- SpGEMM
struct BView {
  int64_t rows;
  const int32_t* rp;
  const int32_t* ci;
};

__global__ void k_spgemm_pairs(int64_t rows, const int32_t* __restrict__ xrp, const int32_t* __restrict__ xci,
                               const int32_t* __restrict__ yrp, int32_t* __restrict__ mx,
                               int32_t* __restrict__ row_pairs) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
    int p = 0;
    for (int e = xrp[r]; e < xrp[r + 1]; ++e) p += yrp[xci[e] + 1] - yrp[xci[e]];
    if (row_pairs) row_pairs[r] = p;
    atomicMax(&mx[0], p);
    atomicMax(&mx[1], xrp[r + 1] - xrp[r]);
  }
}

// In-place inclusive scan of a[0..m) in LDS by NT threads (chunked);
// `part` holds NT ints.  Contains barriers: all threads must call it.
template <int NT>
__device__ inline void lds_inclusive_scan(int32_t* a, int m, int32_t* part) {
  const int t = threadIdx.x;
  const int chunk = (m + NT - 1) / NT;
  const int b = min(m, t * chunk), e = min(m, b + chunk);
  int32_t s = 0;
  for (int i = b; i < e; ++i) {
    s += a[i];
    a[i] = s;
  }
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int32_t acc = 0;
    for (int i = 0; i < NT; ++i) {
      const int32_t v = part[i];
      part[i] = acc;
      acc += v;
    }
  }
  __syncthreads();
  const int32_t off = part[t];
  for (int i = b; i < e; ++i) a[i] += off;
  __syncthreads();
}

// largest e in [0, xl) with pref[e] <= p (pref nondecreasing, pref[xl] > p)
__device__ inline int pair_row(const int32_t* pref, int xl, int p) {
  int lo = 0, hi = xl;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pref[mid] <= p)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// One workgroup per output row r of C = X Y.  The row's block products are
// enumerated as pairs p (X entry e, Y entry f), keyed (column of f, p) and
// bitonic-sorted in LDS; equal columns form one output block, summed in p
// order.  FILL = false: count the output blocks; FILL = true: write them.
// Plan of row r (FILL): output block u of the row is column cci[cb + u] and
// sums the products plan[2 q], plan[2 q + 1] (X entry, Y entry) for q in
// [plan_ptr[cb + u], plan_ptr[cb + u + 1]), in sorted order.
struct PlanOut {
  const int32_t* crp;       // C row pointers (FILL)
  int32_t* cci;             // C column ids
  const int32_t* pair_ptr;  // (rows+1) scan of the products per row
  int32_t* plan_ptr;        // (C nnzb + 1)
  int32_t* plan;            // (2 * total products)
};

template <int NT, int CAP, bool FILL>
__global__ __launch_bounds__(NT) void k_spgemm_row(BView X, BView Y, int32_t* __restrict__ counts, PlanOut o,
                                                   int32_t* __restrict__ err) {
  __shared__ uint64_t keys[CAP];
  __shared__ int32_t pos[CAP + 1];
  __shared__ int32_t pref[kXCap + 1];
  __shared__ int32_t part[NT];
  const int t = threadIdx.x;
  const int64_t r = blockIdx.x;
  const int xb = X.rp[r], xl = X.rp[r + 1] - xb;
  if (xl > kXCap) {  // uniform over the workgroup
    if (t == 0) atomicOr(err, 1);
    return;
  }
  for (int e = t; e < xl; e += NT) {
    const int c = X.ci[xb + e];
    pref[e + 1] = Y.rp[c + 1] - Y.rp[c];
  }
  if (t == 0) pref[0] = 0;
  __syncthreads();
  lds_inclusive_scan<NT>(pref + 1, xl, part);
  const int np = pref[xl];
  if (np > CAP) {
    if (t == 0) atomicOr(err, 2);
    return;
  }
  int P2 = 1;
  while (P2 < np) P2 <<= 1;
  for (int p = t; p < P2; p += NT) {
    uint64_t k = ~0ull;
    if (p < np) {
      const int e = pair_row(pref, xl, p);
      const int f = Y.rp[X.ci[xb + e]] + (p - pref[e]);
      k = ((uint64_t)(uint32_t)Y.ci[f] << 32) | (uint32_t)p;
    }
    keys[p] = k;
  }
  __syncthreads();
  for (int k = 2; k <= P2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < P2; i += NT) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint64_t a = keys[i], b = keys[ixj];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            keys[i] = b;
            keys[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  for (int p = t; p < np; p += NT) pos[p] = (p == 0 || (keys[p] >> 32) != (keys[p - 1] >> 32)) ? 1 : 0;
  __syncthreads();
  lds_inclusive_scan<NT>(pos, np, part);
  const int U = np > 0 ? pos[np - 1] : 0;
  if (!FILL) {
    if (t == 0) counts[r] = U;
    return;
  }
  const int64_t cb = o.crp[r];
  if (o.crp[r + 1] - cb != U) {
    if (t == 0) atomicOr(err, 4);
    return;
  }
  __shared__ int32_t seg[CAP + 1];  // output block u = sorted products [seg[u], seg[u+1])
  for (int p = t; p < np; p += NT)
    if (p == 0 || (keys[p] >> 32) != (keys[p - 1] >> 32)) {
      const int u = pos[p] - 1;
      seg[u] = p;
      o.cci[cb + u] = (int32_t)(keys[p] >> 32);
    }
  __syncthreads();
  const int64_t pb = o.pair_ptr[r];
  for (int u = t; u < U; u += NT) o.plan_ptr[cb + u] = (int32_t)(pb + seg[u]);
  if (t == 0) o.plan_ptr[cb + U] = o.pair_ptr[r + 1];
  for (int q = t; q < np; q += NT) {
    const int pp = (int)(keys[q] & 0xFFFFFFFFu);
    const int e = pair_row(pref, xl, pp);
    o.plan[2 * (pb + q)] = xb + e;
    o.plan[2 * (pb + q) + 1] = Y.rp[X.ci[xb + e]] + (pp - pref[e]);
  }
}

// Rows with many products: the output columns are first collected in an LDS
// hash table and ranked (sorted), so a product's sort key fits 32 bits
// (rank << 18 | p).  The products are then placed chunk by chunk (kBigCap at
// a time, each chunk bitonic-sorted): an output block's products are in
// (chunk, p) order, a fixed order.  Limits: 2^18 products and kBigHash / 2
// output blocks per row, kBigXCap blocks per row of X.
constexpr int kBigCap = 16384;   // products sorted at a time
constexpr int kBigHash = 4096;   // hash slots
constexpr int kBigXCap = 2048;   // blocks in one row of X
constexpr int kBigPairs = 1 << 18;

__device__ inline int hash_slot(int col) { return (int)(((uint32_t)col * 2654435761u) >> 20); }  // 12 bits

template <int NT, bool FILL>
__global__ __launch_bounds__(NT) void k_spgemm_row_big(BView X, BView Y, int32_t* __restrict__ counts, PlanOut o,
                                                       int32_t* __restrict__ err) {
  static_assert(kBigHash == 4096, "hash_slot yields 12 bits");
  __shared__ uint32_t keys[kBigCap];
  __shared__ int32_t tab[kBigHash];
  __shared__ int32_t rank[kBigHash];
  __shared__ int32_t ucol[kBigHash / 2];  // distinct columns; reused as running offsets
  __shared__ int32_t seg[kBigHash / 2 + 1];
  __shared__ int32_t cs[kBigHash / 2];    // segment start of a rank within the current chunk
  __shared__ int32_t pref[kBigXCap + 1];
  __shared__ int32_t part[NT];
  __shared__ int32_t flag;
  const int t = threadIdx.x;
  const int64_t r = blockIdx.x;
  const int xb = X.rp[r], xl = X.rp[r + 1] - xb;
  if (xl > kBigXCap) {
    if (t == 0) atomicOr(err, 1);
    return;
  }
  for (int e = t; e < xl; e += NT) {
    const int c = X.ci[xb + e];
    pref[e + 1] = Y.rp[c + 1] - Y.rp[c];
  }
  for (int i = t; i < kBigHash; i += NT) tab[i] = -1;
  if (t == 0) {
    pref[0] = 0;
    flag = 0;
  }
  __syncthreads();
  lds_inclusive_scan<NT>(pref + 1, xl, part);
  const int np = pref[xl];
  if (np >= kBigPairs) {
    if (t == 0) atomicOr(err, 2);
    return;
  }
  auto pair_col = [&](int p) -> int {
    const int e = pair_row(pref, xl, p);
    return Y.ci[Y.rp[X.ci[xb + e]] + (p - pref[e])];
  };
  auto find = [&](int c) -> int {  // slot of a column present in the table
    int h = hash_slot(c);
    while (tab[h] != c) h = (h + 1) & (kBigHash - 1);
    return h;
  };
  // 1. distinct output columns into the hash table (bounded probing)
  for (int p = t; p < np; p += NT) {
    const int c = pair_col(p);
    int h = hash_slot(c);
    int probes = 0;
    while (true) {
      const int prev = atomicCAS(&tab[h], -1, c);
      if (prev == -1 || prev == c) break;
      h = (h + 1) & (kBigHash - 1);
      if (++probes >= kBigHash) {
        flag = 1;
        break;
      }
    }
  }
  __syncthreads();
  // 2. compact and sort the distinct columns
  for (int i = t; i < kBigHash; i += NT) rank[i] = tab[i] >= 0 ? 1 : 0;
  __syncthreads();
  lds_inclusive_scan<NT>(rank, kBigHash, part);
  const int U = rank[kBigHash - 1];
  if (U > kBigHash / 2 || flag) {
    if (t == 0) atomicOr(err, 8);
    return;
  }
  int U2 = 1;
  while (U2 < U) U2 <<= 1;
  for (int i = t; i < kBigHash; i += NT)
    if (tab[i] >= 0) ucol[rank[i] - 1] = tab[i];
  __syncthreads();
  for (int i = U + t; i < U2; i += NT) ucol[i] = INT32_MAX;
  __syncthreads();
  for (int k = 2; k <= U2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < U2; i += NT) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const int a = ucol[i], b = ucol[ixj];
          if ((a > b) == ((i & k) == 0)) {
            ucol[i] = b;
            ucol[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  if (!FILL) {
    if (t == 0) counts[r] = U;
    return;
  }
  const int64_t cb = o.crp[r];
  if (o.crp[r + 1] - cb != U) {
    if (t == 0) atomicOr(err, 4);
    return;
  }
  // 3. rank of every occupied slot; products per rank
  for (int u = t; u < U; u += NT) {
    const int c = ucol[u];
    rank[find(c)] = u;
    o.cci[cb + u] = c;
    seg[u] = 0;
  }
  __syncthreads();
  for (int p = t; p < np; p += NT) atomicAdd(&seg[rank[find(pair_col(p))]], 1);
  __syncthreads();
  lds_inclusive_scan<NT>(seg, U, part);  // seg[u] = products of ranks <= u
  const int64_t pb = o.pair_ptr[r];
  for (int u = t; u < U; u += NT) {
    const int base = u == 0 ? 0 : seg[u - 1];
    o.plan_ptr[cb + u] = (int32_t)(pb + base);
    cs[u] = base;  // first free position of rank u
  }
  if (t == 0) o.plan_ptr[cb + U] = o.pair_ptr[r + 1];
  __syncthreads();
  for (int u = t; u < U; u += NT) seg[u] = cs[u];  // seg[u]: next position of rank u
  __syncthreads();
  // 4. place the products chunk by chunk
  for (int c0 = 0; c0 < np; c0 += kBigCap) {
    const int m = min(kBigCap, np - c0);
    int P2 = 1;
    while (P2 < m) P2 <<= 1;
    for (int q = t; q < P2; q += NT) {
      uint32_t k = 0xFFFFFFFFu;
      if (q < m) {
        const int p = c0 + q;
        k = ((uint32_t)rank[find(pair_col(p))] << 18) | (uint32_t)p;
      }
      keys[q] = k;
    }
    __syncthreads();
    for (int k = 2; k <= P2; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = t; i < P2; i += NT) {
          const int ixj = i ^ j;
          if (ixj > i) {
            const uint32_t a = keys[i], b = keys[ixj];
            if ((a > b) == ((i & k) == 0)) {
              keys[i] = b;
              keys[ixj] = a;
            }
          }
        }
        __syncthreads();
      }
    for (int q = t; q < m; q += NT)
      if (q == 0 || (keys[q] >> 18) != (keys[q - 1] >> 18)) cs[keys[q] >> 18] = q;
    __syncthreads();
    for (int q = t; q < m; q += NT) {
      const int u = (int)(keys[q] >> 18);
      const int pp = (int)(keys[q] & 0x3FFFFu);
      const int e = pair_row(pref, xl, pp);
      const int64_t pos = pb + seg[u] + (q - cs[u]);
      o.plan[2 * pos] = xb + e;
      o.plan[2 * pos + 1] = Y.rp[X.ci[xb + e]] + (pp - pref[e]);
    }
    __syncthreads();
    for (int q = t; q < m; q += NT)
      if (q == m - 1 || (keys[q] >> 18) != (keys[q + 1] >> 18)) {
        const int u = (int)(keys[q] >> 18);
        seg[u] += q - cs[u] + 1;
      }
    __syncthreads();
  }
}

// n consecutive entries in their storage type, with 16-byte loads where n and
// the entry size allow (p is then 16-byte aligned: blocks of 36 FP32 entries)
template <int n, typename T>
__device__ inline void load_raw(const T* __restrict__ p, T (&a)[n]) {
  if constexpr (sizeof(T) == 4 && n % 4 == 0) {
    const float4* p4 = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int k = 0; k < n / 4; ++k) {
      const float4 v = p4[k];
      a[4 * k] = v.x;
      a[4 * k + 1] = v.y;
      a[4 * k + 2] = v.z;
      a[4 * k + 3] = v.w;
    }
  } else if constexpr (sizeof(T) == 8 && n % 2 == 0) {
    const double2* p2 = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int k = 0; k < n / 2; ++k) {
      const double2 v = p2[k];
      a[2 * k] = v.x;
      a[2 * k + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < n; ++k) a[k] = p[k];
  }
}

// Load an R x C block (FP64 or FP32 storage) into FP64 registers.  Blocks whose
// size is a multiple of 16 bytes (given 256-byte aligned bases: FP64 blocks but
// 3x3 and 1x1, FP32 6x6 blocks) use 16-byte loads; FP32 blocks of an even
// count (3x6, 6x3) 8-byte loads.
template <int R, int C, typename T = double>
__device__ inline void load_block(const T* __restrict__ p, double (&a)[R][C]) {
  constexpr int n = R * C;
  if constexpr (sizeof(T) == 8 && n % 2 == 0 && n > 9) {
    const double2* p2 = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int k = 0; k < n / 2; ++k) {
      const double2 v = p2[k];
      a[(2 * k) / C][(2 * k) % C] = v.x;
      a[(2 * k + 1) / C][(2 * k + 1) % C] = v.y;
    }
  } else if constexpr (sizeof(T) == 4 && n % 4 == 0 && (n * 4) % 16 == 0 && n > 9) {
    const float4* p4 = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int k = 0; k < n / 4; ++k) {
      const float4 v = p4[k];
      a[(4 * k) / C][(4 * k) % C] = v.x;
      a[(4 * k + 1) / C][(4 * k + 1) % C] = v.y;
      a[(4 * k + 2) / C][(4 * k + 2) % C] = v.z;
      a[(4 * k + 3) / C][(4 * k + 3) % C] = v.w;
    }
  } else if constexpr (sizeof(T) == 4 && n % 2 == 0 && n > 9) {
    const float2* p2 = reinterpret_cast<const float2*>(p);
#pragma unroll
    for (int k = 0; k < n / 2; ++k) {
      const float2 v = p2[k];
      a[(2 * k) / C][(2 * k) % C] = v.x;
      a[(2 * k + 1) / C][(2 * k + 1) % C] = v.y;
    }
  } else {
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int j = 0; j < C; ++j) a[i][j] = double(p[i * C + j]);
  }
}

// C = X Y by its plan: one thread per output block, products in plan order.
template <int XR, int XC, int YC, typename T>
__global__ __launch_bounds__(256) void k_spgemm_numeric(int64_t nb, const int32_t* __restrict__ plan_ptr,
                                                        const int32_t* __restrict__ plan, const T* __restrict__ xv,
                                                        const T* __restrict__ yv, T* __restrict__ cv) {
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nb; b += (int64_t)gridDim.x * blockDim.x) {
    double acc[XR][YC];
#pragma unroll
    for (int i = 0; i < XR; ++i)
#pragma unroll
      for (int j = 0; j < YC; ++j) acc[i][j] = 0.0;
    const int q0 = plan_ptr[b], q1 = plan_ptr[b + 1];
    const int2* pl = reinterpret_cast<const int2*>(plan);
    if constexpr (sizeof(T) == 4) {
      // FP32 blocks, software-pipelined: NP products' blocks are issued together
      // and the next NP plan entries are fetched while they are multiplied (the
      // compiler's own unrolling keeps about one product in flight); products
      // are added in plan order.  NP = 1 where the accumulator alone takes 72
      // registers (6-row outputs)
      constexpr int NP = XR == 3 ? 2 : 1;
      int2 e[NP];
#pragma unroll
      for (int u = 0; u < NP; ++u) e[u] = q0 + u < q1 ? pl[q0 + u] : make_int2(0, 0);
      for (int q = q0; q < q1; q += NP) {
        T xa[NP][XR * XC], ya[NP][XC * YC];
#pragma unroll
        for (int u = 0; u < NP; ++u)
          if (u == 0 || q + u < q1) {
            load_raw<XR * XC, T>(xv + (size_t)e[u].x * XR * XC, xa[u]);
            load_raw<XC * YC, T>(yv + (size_t)e[u].y * XC * YC, ya[u]);
          }
#pragma unroll
        for (int u = 0; u < NP; ++u) e[u] = q + NP + u < q1 ? pl[q + NP + u] : make_int2(0, 0);
#pragma unroll
        for (int u = 0; u < NP; ++u)
          if (u == 0 || q + u < q1)
#pragma unroll
            for (int i = 0; i < XR; ++i)
#pragma unroll
              for (int j = 0; j < YC; ++j)
#pragma unroll
                for (int m = 0; m < XC; ++m)
                  acc[i][j] = fma(double(xa[u][i * XC + m]), double(ya[u][m * YC + j]), acc[i][j]);
      }
    } else {
      // several products' loads in flight per thread (the 6x3 x 3x6 case is
      // register-bound and runs best without)
      constexpr int kUnroll = (XR == 6 && XC == 3) ? 1 : 4;
#pragma unroll kUnroll
      for (int q = q0; q < q1; ++q) {
        const int2 ef = pl[q];
        double xa[XR][XC], ya[XC][YC];
        load_block<XR, XC, T>(xv + (size_t)ef.x * XR * XC, xa);
        load_block<XC, YC, T>(yv + (size_t)ef.y * XC * YC, ya);
#pragma unroll
        for (int i = 0; i < XR; ++i)
#pragma unroll
          for (int j = 0; j < YC; ++j)
#pragma unroll
            for (int m = 0; m < XC; ++m) acc[i][j] = fma(xa[i][m], ya[m][j], acc[i][j]);
      }
    }
    T* c = cv + (size_t)b * XR * YC;
#pragma unroll
    for (int i = 0; i < XR; ++i)
#pragma unroll
      for (int j = 0; j < YC; ++j) c[i * YC + j] = T(acc[i][j]);
  }
}

// Products with fewer output blocks than this split each block's plan over a wave.
constexpr int64_t kSplitBelow = 200000;

// C = X Y by its plan for products with few output blocks (the coarse levels,
// where one thread per block leaves the GPU nearly idle on long serial sums):
// YC lanes form one column group (lane c forms column c of the block: the group
// reads the X block together, lane c reads column c of the Y block), and with
// S > 1 the S groups of a wave split one block's plan (group g takes products
// g, g + S, ...); group 0 sums their columns in group order.  The order is
// fixed: bit-reproducible.
template <int XR, int XC, int YC, int S, typename T>
__global__ __launch_bounds__(256) void k_spgemm_numeric_g(int64_t nb, const int32_t* __restrict__ plan_ptr,
                                                          const int32_t* __restrict__ plan, const T* __restrict__ xv,
                                                          const T* __restrict__ yv, T* __restrict__ cv) {
  constexpr int G = YC, GPW = 64 / G;  // lanes per group, groups per wave (the last 64 - GPW*G lanes idle)
  static_assert(S == 1 || S == GPW, "a block is split over one group or over the whole wave");
  constexpr int BPW = GPW / S;  // output blocks per wave
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int grp = lane / G, c = lane - grp * G;
  const int64_t b = wave * BPW + grp / S;
  const int sg = grp % S;
  double acc[XR];
#pragma unroll
  for (int i = 0; i < XR; ++i) acc[i] = 0.0;
  if (grp < GPW && b < nb) {
    const int q1 = plan_ptr[b + 1];
#pragma unroll 2
    for (int q = plan_ptr[b] + sg; q < q1; q += S) {
      const int2 ef = reinterpret_cast<const int2*>(plan)[q];
      double xa[XR][XC], yc[XC];
      load_block<XR, XC, T>(xv + (size_t)ef.x * XR * XC, xa);
      const T* y = yv + (size_t)ef.y * XC * YC + c;
#pragma unroll
      for (int m = 0; m < XC; ++m) yc[m] = double(y[m * YC]);
#pragma unroll
      for (int i = 0; i < XR; ++i)
#pragma unroll
        for (int m = 0; m < XC; ++m) acc[i] = fma(xa[i][m], yc[m], acc[i]);
    }
  }
  if (S > 1) {
    // lanes c of groups 1 .. S-1 hand their columns to lane c of group 0, in group order
#pragma unroll
    for (int i = 0; i < XR; ++i) {
      double t = acc[i];
      for (int g = 1; g < S; ++g) t += __shfl(acc[i], c + g * G, 64);
      acc[i] = t;
    }
  }
  if (grp < GPW && sg == 0 && b < nb) {
    T* out = cv + (size_t)b * XR * YC + c;
#pragma unroll
    for (int i = 0; i < XR; ++i) out[i * YC] = T(acc[i]);
  }
}

// P = P0 - omega D^-1 (A P0) along the plan of A P0 (Y = P0 as block CSR, one
// block per row at column agg): one thread per block of P.
template <int BS, int NS, typename T>
__global__ __launch_bounds__(256) void k_smooth_numeric(int64_t nb, int64_t rows, const int32_t* __restrict__ crp,
                                                        const int32_t* __restrict__ cci,
                                                        const int32_t* __restrict__ plan_ptr,
                                                        const int32_t* __restrict__ plan, const T* __restrict__ av,
                                                        const double* __restrict__ p0, const int32_t* __restrict__ agg,
                                                        const double* __restrict__ dinv, double omega,
                                                        const double* __restrict__ lmax_dev, T* __restrict__ pv) {
  // lmax_dev: the level's lambda_max estimate on the device; omega = 1.5 / lambda_max
  // (against the textbook 4 / (3 lambda_max): CG iterations on the HEX8 cantilever 48 -> 47,
  // the SIMP cantilever 68 -> 68, the HEX27 cube 113 -> 107; 1.7 / lambda_max took 49 on the
  // cantilever)
  if (lmax_dev) omega = 1.5 / lmax_dev[0];
  const int32_t plan_len = plan_ptr[nb];
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nb; b += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = rows;  // row i: crp[i] <= b < crp[i+1]
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (crp[mid] <= b)
        lo = mid;
      else
        hi = mid;
    }
    const int64_t i = lo;
    double acc[BS][NS];
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
      for (int j = 0; j < NS; ++j) acc[r][j] = 0.0;
    const int q0 = plan_ptr[b], q1 = plan_ptr[b + 1];
    if constexpr (sizeof(T) == 4 && BS == 3) {
      // software-pipelined like k_spgemm_numeric: the next plan entry is fetched
      // during the multiply, sums in plan order (two products in flight measured
      // no faster: 7.41 against 7.44 ms of C2 AMG setup)
      const int2* pl = reinterpret_cast<const int2*>(plan);
      constexpr int NP = 1;
      int2 e[NP];
#pragma unroll
      for (int u = 0; u < NP; ++u) e[u] = q0 + u < q1 ? pl[q0 + u] : make_int2(0, 0);
      for (int q = q0; q < q1; q += NP) {
        T xa[NP][BS * BS];
        double ya[NP][BS * NS];
#pragma unroll
        for (int u = 0; u < NP; ++u)
          if (u == 0 || q + u < q1) {
            load_raw<BS * BS, T>(av + (size_t)e[u].x * BS * BS, xa[u]);
            load_raw<BS * NS, double>(p0 + (size_t)e[u].y * BS * NS, ya[u]);
          }
#pragma unroll
        for (int u = 0; u < NP; ++u) e[u] = q + NP + u < q1 ? pl[q + NP + u] : make_int2(0, 0);
#pragma unroll
        for (int u = 0; u < NP; ++u)
          if (u == 0 || q + u < q1)
#pragma unroll
            for (int r = 0; r < BS; ++r)
#pragma unroll
              for (int m = 0; m < BS; ++m)
#pragma unroll
                for (int j = 0; j < NS; ++j)
                  acc[r][j] = fma(double(xa[u][r * BS + m]), ya[u][m * NS + j], acc[r][j]);
      }
    } else {
      // the thread's (A block, P0 block) pairs are contiguous: read through a window
      ListWindow<4> pw(plan, 2 * (int64_t)plan_len);
      for (int q = q0; q < q1; ++q) {
        double xa[BS][BS], ya[BS][NS];
        load_block<BS, BS, T>(av + (size_t)pw[2 * (int64_t)q] * BS * BS, xa);
        load_block<BS, NS>(p0 + (size_t)pw[2 * (int64_t)q + 1] * BS * NS, ya);
#pragma unroll
        for (int r = 0; r < BS; ++r)
#pragma unroll
          for (int m = 0; m < BS; ++m)
#pragma unroll
            for (int j = 0; j < NS; ++j) acc[r][j] = fma(xa[r][m], ya[m][j], acc[r][j]);
      }
    }
    const bool own = cci[b] == agg[i];
    T* out = pv + (size_t)b * BS * NS;
#pragma unroll
    for (int r = 0; r < BS; ++r) {
      const double sc = -omega * dinv[i * BS + r];
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        double v = sc * acc[r][j];
        if (own) v += p0[(i * BS + r) * NS + j];
        out[r * NS + j] = T(v);
      }
    }
  }
}

// Sort each row's (column, source) pairs by column: one workgroup per row,
// bitonic in LDS (rows of at most kSortCap entries; longer rows, which the
// hierarchy does not produce, are sorted by one thread).
constexpr int kSortCap = 4096;
__global__ __launch_bounds__(256) void k_sort_rows(int64_t rows, const int32_t* __restrict__ ptr,
                                                   int32_t* __restrict__ col, int32_t* __restrict__ src) {
  __shared__ uint64_t k[kSortCap];
  const int64_t r = blockIdx.x;
  const int t = threadIdx.x;
  const int b = ptr[r], n = ptr[r + 1] - b;
  if (n <= 1) return;
  if (n > kSortCap) {
    if (t == 0)
      for (int i = b + 1; i < b + n; ++i) {
        const int32_t c = col[i], sv = src[i];
        int j = i - 1;
        while (j >= b && col[j] > c) {
          col[j + 1] = col[j];
          src[j + 1] = src[j];
          --j;
        }
        col[j + 1] = c;
        src[j + 1] = sv;
      }
    return;
  }
  int P2 = 1;
  while (P2 < n) P2 <<= 1;
  for (int i = t; i < P2; i += 256)
    k[i] = i < n ? (((uint64_t)(uint32_t)col[b + i] << 32) | (uint32_t)src[b + i]) : ~0ull;
  __syncthreads();
  for (int kk = 2; kk <= P2; kk <<= 1)
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = t; i < P2; i += 256) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint64_t a = k[i], c = k[ixj];
          if ((a > c) == ((i & kk) == 0)) {
            k[i] = c;
            k[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  for (int i = t; i < n; i += 256) {
    col[b + i] = (int32_t)(k[i] >> 32);
    src[b + i] = (int32_t)(k[i] & 0xFFFFFFFFu);
  }
}

// ------------------------------------------------------------ dense
__global__ void k_dense_zero(int64_t nn, double* __restrict__ M) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x)
    M[i] = 0.0;
}

// one workgroup per block row (rows <= 4096): its threads stride over the row's block entries
template <typename T>
__global__ __launch_bounds__(256) void k_dense_scatter(int64_t rows, int br, int bc, int64_t n,
                                                       const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                       const T* __restrict__ v, double* __restrict__ M) {
  const int64_t i = blockIdx.x;
  if (i >= rows) return;
  const int kb = rp[i], per = br * bc, nf = (rp[i + 1] - kb) * per;
  for (int f = threadIdx.x; f < nf; f += 256) {
    const int k = kb + f / per, e = f % per, a = e / bc, b = e % bc;
    M[(i * br + a) * n + (int64_t)ci[k] * bc + b] = double(v[(size_t)k * per + e]);
  }
}

// one Gauss-Jordan pivot step src -> dst (in-place inversion formula)
__global__ void k_gj_step(int64_t n, int64_t k, const double* __restrict__ src, double* __restrict__ dst) {
  const double d = 1.0 / src[k * n + k];
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, j = e % n;
    double v;
    if (i == k && j == k)
      v = d;
    else if (i == k)
      v = src[k * n + j] * d;
    else if (j == k)
      v = -src[i * n + k] * d;
    else
      v = src[e] - src[i * n + k] * src[k * n + j] * d;
    dst[e] = v;
  }
}

// the value of v held by lane l of the wave (l uniform across the wave)
__device__ inline double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The whole Gauss-Jordan elimination of a small dense block (n <= kGjSmall: the n x n
// block at (k0, k0) of a matrix of leading dimension ld, inverted in place) in one
// workgroup, each step the arithmetic of k_gj_step (so the inverse is the same bit for
// bit) without a launch per pivot.  The block stays in registers: wave w of 16 holds rows
// w, w + 16, w + 32, w + 48, lane j column j (16 waves: 50 us for a 64 x 64 block, 4 waves
// of 16 rows: 78 us).  Per pivot k only the pivot row and its reciprocal pivot go through
// LDS (the row's owner publishes both at the end of step k - 1, double-buffered, one
// barrier per pivot); entry (i, k) of a wave's own row comes from lane k of that wave.
// The loop is issue-bound (four waves per SIMD): one division per pivot instead of one per
// wave, and selects instead of branches.
constexpr int kGjSmall = 64;
constexpr int kGjWaves = 16;
__global__ __launch_bounds__(64 * kGjWaves) void k_gj_small(int n, int64_t ld, int64_t k0, double* __restrict__ a) {
  constexpr int kW = kGjWaves, kPer = kGjSmall / kW;
  __shared__ double rowbuf[2][kGjSmall + 1];  // the pivot row, then 1 / its pivot
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool col = lane < n;
  double m[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int i = w + kW * u;
    m[u] = (i < n && col) ? a[(k0 + i) * ld + k0 + lane] : 0.0;
  }
  if (w == 0) {
    rowbuf[0][lane] = m[0];
    const double p0 = readlane_f64(m[0], 0);
    if (lane == 0) rowbuf[0][kGjSmall] = 1.0 / p0;
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const double rkj = rowbuf[k & 1][lane];
    const double d = rowbuf[k & 1][kGjSmall];
    const double rd = rkj * d;  // the pivot row's new entries
    const bool pc = lane == k;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = w + kW * u;
      const double aik = readlane_f64(m[u], k);
      double v = fma(-(aik * rkj), d, m[u]);
      v = pc ? -aik * d : v;
      v = i == k ? (pc ? d : rd) : v;
      if (i < n && col) m[u] = v;
    }
    // the owner of row k + 1 publishes it and its reciprocal pivot for the next step
    if (k + 1 < n) {
#pragma unroll
      for (int u = 0; u < kPer; ++u)
        if (w + kW * u == k + 1) {
          rowbuf[(k + 1) & 1][lane] = m[u];
          const double piv = readlane_f64(m[u], k + 1);
          if (lane == 0) rowbuf[(k + 1) & 1][kGjSmall] = 1.0 / piv;
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int i = w + kW * u;
    if (i < n && col) a[(k0 + i) * ld + k0 + lane] = m[u];
  }
}

// Blocked Gauss-Jordan (n >= 2 kGjSmall): per pivot block K = [k0, k0 + bk), bk <= kGjSmall,
// with P = A_KK^-1 (k_gj_small, in place),
//   A_Kj <- P A_Kj,   A_iK <- -A_iK P,   A_ij <- A_ij - A_iK (P A_Kj)     (i, j outside K):
// the scalar steps' Schur complements, the pivots taken in the same order, a product of
// 64-wide panels per block instead of a launch per pivot.
// k_bgj_panel: colp (n x kGjSmall) <- the old column panel A_iK (i outside K),
//              rowp (bk x n) <- P A_Kj (j outside K);
// one workgroup per 64-column tile and group of 16 rows of P: those rows of P and the tile
// of A_Kj in LDS, 4 entries per thread (the first row group also copies 64 rows of the
// column panel).
__global__ __launch_bounds__(256) void k_bgj_panel(int64_t n, int64_t k0, int bk, const double* __restrict__ a,
                                                   double* __restrict__ colp, double* __restrict__ rowp) {
  constexpr int T = kGjSmall, RG = 16;
  __shared__ double sp[RG][T + 1];  // rows r0 .. r0 + 15 of P
  __shared__ double sa[T][T + 1];   // the tile of A_Kj
  const int64_t tj = blockIdx.x, j0 = tj * T, kb = k0 / T;
  const int r0 = blockIdx.y * RG;
  if (blockIdx.y == 0)
    for (int e = threadIdx.x; e < T * T; e += 256) {
      const int64_t i = j0 + e / T;
      const int c = e % T;
      if (i < n && c < bk && (i < k0 || i >= k0 + bk)) colp[i * T + c] = a[i * n + k0 + c];
    }
  if (tj == kb || r0 >= bk) return;
  for (int e = threadIdx.x; e < RG * T; e += 256) {
    const int r = e / T, q = e % T;
    sp[r][q] = (r0 + r < bk && q < bk) ? a[(k0 + r0 + r) * n + k0 + q] : 0.0;
  }
  for (int e = threadIdx.x; e < T * T; e += 256) {
    const int r = e / T, q = e % T;
    const int64_t j = j0 + q;
    sa[r][q] = (r < bk && j < n) ? a[(k0 + r) * n + j] : 0.0;
  }
  __syncthreads();
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int c = 0; c < bk; ++c) {
    const double x = sp[tr][c];
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] = fma(x, sa[c][tc + 16 * v], acc[v]);
  }
  const int r = r0 + tr;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int64_t j = j0 + tc + 16 * v;
    if (r < bk && j < n) rowp[r * n + j] = acc[v];
  }
}

// k_bgj_update: one workgroup per 64 x 64 tile of A (tiles aligned with the pivot blocks);
// the tile's column panel and row panel (or P) in LDS, 4 x 4 entries per thread.
__global__ __launch_bounds__(256) void k_bgj_update(int64_t n, int64_t k0, int bk, double* __restrict__ a,
                                                    const double* __restrict__ colp,
                                                    const double* __restrict__ rowp) {
  constexpr int T = kGjSmall;
  __shared__ double sc[T][T + 1];  // sc[r][c] = A_iK (old), i = i0 + r
  __shared__ double sr[T][T + 1];  // sr[c][q] = (P A_Kj)[c][j0 + q], or P[c][q] on the pivot column tile
  const int64_t kb = k0 / T, ti = blockIdx.y, tj = blockIdx.x, i0 = ti * T, j0 = tj * T;
  if (ti == kb) {
    // the pivot rows take P A_Kj; the pivot block keeps P
    if (tj == kb) return;
    for (int e = threadIdx.x; e < T * T; e += 256) {
      const int r = e / T;
      const int64_t j = j0 + e % T;
      if (r < bk && j < n) a[(k0 + r) * n + j] = rowp[r * n + j];
    }
    return;
  }
  const bool pivcol = tj == kb;
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  // the tile's own entries, loaded ahead of the product (the pivot column tile has none)
  double old[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t i = i0 + tr + 16 * u, j = j0 + tc + 16 * v;
      old[u][v] = (!pivcol && i < n && j < n) ? a[i * n + j] : 0.0;
    }
  for (int e = threadIdx.x; e < T * T; e += 256) {
    const int r = e / T, c = e % T;
    const int64_t i = i0 + r;
    sc[r][c] = (i < n && c < bk) ? colp[i * T + c] : 0.0;
    const int cc = e / T, q = e % T;
    const int64_t j = j0 + q;
    sr[cc][q] = (cc < bk && j < n) ? (pivcol ? a[(k0 + cc) * n + j] : rowp[cc * n + j]) : 0.0;
  }
  __syncthreads();
  double acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
  for (int c = 0; c < bk; ++c) {
    double x[4], y[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) x[u] = sc[tr + 16 * u][c];
#pragma unroll
    for (int v = 0; v < 4; ++v) y[v] = sr[c][tc + 16 * v];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] = fma(x[u], y[v], acc[u][v]);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t i = i0 + tr + 16 * u, j = j0 + tc + 16 * v;
      if (i < n && j < n) a[i * n + j] = pivcol ? -acc[u][v] : old[u][v] - acc[u][v];
    }
}

__global__ void k_copy(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// y = M x, one wave per row; four independent sums per lane keep four loads in flight
// (a fixed order: the result does not depend on timing)
__global__ __launch_bounds__(256) void k_dense_gemv(int64_t n, const double* __restrict__ M,
                                                    const double* __restrict__ x, double* __restrict__ y,
                                                    const int32_t* __restrict__ done) {
  if (done && *done) return;
  const int lane = threadIdx.x & 63;
  const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (i >= n) return;
  const double* row = M + i * n;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int64_t j = lane;
  for (; j + 192 < n; j += 256) {
    s0 = fma(row[j], x[j], s0);
    s1 = fma(row[j + 64], x[j + 64], s1);
    s2 = fma(row[j + 128], x[j + 128], s2);
    s3 = fma(row[j + 192], x[j + 192], s3);
  }
  for (; j < n; j += 64) s0 = fma(row[j], x[j], s0);
  const double s = wave_sum((s0 + s1) + (s2 + s3));
  if (lane == 0) y[i] = s;
}

// ---------------------------------------------------------- operator apply
enum { AP_STORE = 0, AP_ADD = 1, AP_RESID = 2, AP_CHEB = 3, AP_DSTORE = 4, AP_PQ = 5 };

// TX: storage of the input vector and of the Chebyshev direction, TY: of the
// output.  Arithmetic is FP64 whatever the storage; the V-cycle's vectors are
// FP64, the level-0 power iteration on the FP32 copy keeps FP32 iterates.
template <typename TX = double, typename TY = double>
struct ApplyArgsT {
  const TX* x;         // gathered input
  TY* y;               // output (CHEB: the new iterate)
  const double* rhs;   // RESID, CHEB
  const double* dinv;  // CHEB, DSTORE
  TX* d;               // CHEB: search direction (in/out)
  double c1, c2;       // CHEB: d = c1 d + c2 D^-1 (rhs - A x), y = x + d
  double* part;        // PQ: per-block partials of <x, A x>
  const int32_t* done; // NULL, or a flag: the kernel returns at once when it is set
                       // (V-cycles queued after the outer CG converged)
};
using ApplyArgs = ApplyArgsT<double, double>;

// The epilogue's own-entry operands (loaded ahead of the product where a kernel can
// overlap their latency with it) and the epilogue itself.
struct EpiOperands {
  double y, rhs, dinv, d, x;
};

template <int MODE, typename TX, typename TY>
__device__ inline EpiOperands epilogue_load(const ApplyArgsT<TX, TY>& a, int64_t idx) {
  EpiOperands e{0.0, 0.0, 0.0, 0.0, 0.0};
  if (MODE == AP_ADD) e.y = double(a.y[idx]);
  if (MODE == AP_RESID || MODE == AP_CHEB) e.rhs = a.rhs[idx];
  if (MODE == AP_CHEB || MODE == AP_DSTORE) e.dinv = a.dinv[idx];
  if (MODE == AP_CHEB && a.c1 != 0.0) e.d = double(a.d[idx]);
  if (MODE == AP_CHEB) e.x = double(a.x[idx]);
  return e;
}

template <int MODE, typename TX, typename TY>
__device__ inline void epilogue_store(const ApplyArgsT<TX, TY>& a, int64_t idx, double ax, const EpiOperands& e) {
  if (MODE == AP_STORE || MODE == AP_PQ) {
    a.y[idx] = TY(ax);
  } else if (MODE == AP_ADD) {
    a.y[idx] = TY(e.y + ax);
  } else if (MODE == AP_RESID) {
    a.y[idx] = TY(e.rhs - ax);
  } else if (MODE == AP_DSTORE) {
    a.y[idx] = TY(e.dinv * ax);
  } else {
    const double r = e.dinv * (e.rhs - ax);
    const double dn = a.c1 != 0.0 ? fma(a.c1, e.d, a.c2 * r) : a.c2 * r;
    a.d[idx] = TX(dn);
    a.y[idx] = TY(e.x + dn);
  }
}

template <int MODE, typename TX, typename TY>
__device__ inline void apply_epilogue(const ApplyArgsT<TX, TY>& a, int64_t idx, double ax) {
  epilogue_store<MODE>(a, idx, ax, epilogue_load<MODE>(a, idx));
}

// Block CSR operator, G lanes per block row: the lanes stride over the row's
// (block, column) pairs and accumulate every block row entry.
template <int BR, int BC, int G, int MODE, typename T, typename TX = double, typename TY = double>
__global__ __launch_bounds__(256) void k_bcsr_apply(int64_t rows, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ci, const T* __restrict__ v,
                                                    ApplyArgsT<TX, TY> a) {
  if (a.done && *a.done) return;
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t row = gid / G;
  const int sub = (int)(gid % G);
  double acc[BR];
#pragma unroll
  for (int i = 0; i < BR; ++i) acc[i] = 0.0;
  if (row < rows) {
    const int b = rp[row];
    const int nf = (rp[row + 1] - b) * BC;
    for (int f = sub; f < nf; f += G) {
      const int k = b + f / BC, j = f % BC;
      const double xv = double(a.x[(int64_t)ci[k] * BC + j]);
      const T* vb = v + (size_t)k * BR * BC + j;
#pragma unroll
      for (int i = 0; i < BR; ++i) acc[i] = fma(double(vb[i * BC]), xv, acc[i]);
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
    for (int i = 0; i < BR; ++i) acc[i] += __shfl_xor(acc[i], off, 64);
  if (row < rows && sub == 0) {
#pragma unroll
    for (int i = 0; i < BR; ++i) apply_epilogue<MODE>(a, row * BR + i, acc[i]);
  }
}

// Block CSR operator, a lane per block: G lanes per block row (G = 64: one
// wave per row) stride over the row's blocks, each lane reading whole blocks
// (16- or 8-byte loads from the row's contiguous block stream) times its
// BC-entry slice of x.  The group's partial rows meet in a shuffle reduction
// (every lane of the group then holds the bitwise identical sums) and lane i
// of the group finishes entry i of the row.  The lanes that finish an entry load
// its epilogue operands (the prolongation's y, the residual's rhs, the Chebyshev
// step's rhs, D^-1, d and x) before the block stream, so that their latency
// overlaps it (the level-1 Chebyshev step: 22.3 -> 18.6 us on the 1.6M-dof
// cantilever, although its four operands spill two registers at 6 waves per SIMD).
template <int BR, int BC, int G, int MODE, typename T, typename TX = double, typename TY = double>
__global__ __launch_bounds__(256, 6) void k_bcsr_apply_blk(int64_t rows, const int32_t* __restrict__ rp,
                                                           const int32_t* __restrict__ ci, const T* __restrict__ v,
                                                           ApplyArgsT<TX, TY> a) {
  static_assert(G >= BR && G <= 64 && (G & (G - 1)) == 0, "G: a power of two, at least the block rows");
  if (a.done && *a.done) return;
  const int sub = threadIdx.x & (G - 1);
  const int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) / G;
  const bool live = row < rows;  // a group is live or dead as a whole: its shuffles stay inside it
  const bool fin = live && sub < BR;
  const int64_t idx = row * BR + sub;
  EpiOperands e{0.0, 0.0, 0.0, 0.0, 0.0};
  if (fin) e = epilogue_load<MODE>(a, idx);
  double acc[BR];
#pragma unroll
  for (int i = 0; i < BR; ++i) acc[i] = 0.0;
  if (live) {
    const int k1 = rp[row + 1];
#pragma unroll 1
    for (int k = rp[row] + sub; k < k1; k += G) {
      // the block stays in its storage type until each product (FP32 blocks: half the
      // registers of an FP64 copy, so more waves stay resident)
      T blk[BR * BC];
      load_raw<BR * BC, T>(v + (size_t)k * BR * BC, blk);
      const TX* xp = a.x + (int64_t)ci[k] * BC;
      double xc[BC];
#pragma unroll
      for (int j = 0; j < BC; ++j) xc[j] = double(xp[j]);
#pragma unroll
      for (int i = 0; i < BR; ++i)
#pragma unroll
        for (int j = 0; j < BC; ++j) acc[i] = fma(double(blk[i * BC + j]), xc[j], acc[i]);
    }
  }
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < BR; ++i) {
    double si = acc[i];
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) si += __shfl_xor(si, off, 64);
    t = sub == i ? si : t;
  }
  if (fin) epilogue_store<MODE>(a, idx, t, e);
}

// Long block rows (R and the coarse operators, hundreds of (block, column) pairs):
// one 256-thread workgroup per block row, partial sums met in LDS.
template <int BR, int BC, int MODE, typename T, typename TX = double, typename TY = double>
__global__ __launch_bounds__(256) void k_bcsr_apply_row(int64_t rows, const int32_t* __restrict__ rp,
                                                        const int32_t* __restrict__ ci, const T* __restrict__ v,
                                                        ApplyArgsT<TX, TY> a) {
  __shared__ double smem[4][BR];
  if (a.done && *a.done) return;
  const int64_t row = blockIdx.x;
  const int w = threadIdx.x >> 6;
  EpiOperands e{0.0, 0.0, 0.0, 0.0, 0.0};
  if (threadIdx.x < BR) e = epilogue_load<MODE>(a, row * BR + threadIdx.x);
  double acc[BR];
#pragma unroll
  for (int i = 0; i < BR; ++i) acc[i] = 0.0;
  const int b = rp[row];
  const int nf = (rp[row + 1] - b) * BC;
  // unrolled so that several pairs' loads are in flight (the sums keep their order)
#pragma unroll 4
  for (int f = threadIdx.x; f < nf; f += 256) {
    const int k = b + f / BC, j = f % BC;
    const double xv = double(a.x[(int64_t)ci[k] * BC + j]);
    const T* vb = v + (size_t)k * BR * BC + j;
#pragma unroll
    for (int i = 0; i < BR; ++i) acc[i] = fma(double(vb[i * BC]), xv, acc[i]);
  }
  // the four waves' sums of all block rows meet in LDS behind one barrier
#pragma unroll
  for (int i = 0; i < BR; ++i) {
    const double si = wave_sum(acc[i]);
    if ((threadIdx.x & 63) == 0) smem[w][i] = si;
  }
  __syncthreads();
  if (threadIdx.x < BR) {
    const int i = threadIdx.x;
    epilogue_store<MODE>(a, row * BR + i, (smem[0][i] + smem[1][i]) + (smem[2][i] + smem[3][i]), e);
  }
}

// Slice split of the V-cycle's SELL kernels: FP32 storage with FP64 vectors 3 waves
// per slice (one thread per (row, component) in the epilogue; on the 1.6M-dof HEX8
// cantilever 3, 4, 6, 8 waves: Chebyshev step 122.5, 123.2, 130.3, 125.6 us, residual
// 111.9, 113.6, 125.9, 121.7 us), otherwise the split of the Jacobi path's kernels.
template <typename T>
struct AmgSliceCfg {
  static constexpr int W = sizeof(T) == 4 ? 3 : SliceCfg<T>::W;
  static constexpr int NT = 64 * W;
};

// Standard-form SELL-64 operator (level 0), FP64 vectors, one block per slice.
// DIFF: the matrix is in difference form (the diagonal slot holds the row
// sum); the own-row entries of the input then enter every product.
template <int V, int MODE, typename T, bool DIFF = false, typename TX = double, typename TY = double>
__global__ __launch_bounds__(AmgSliceCfg<T>::NT) void k_sell_apply(fem_pattern_t pat, const T* __restrict__ vals,
                                                                ApplyArgsT<TX, TY> a,
                                                                const int32_t* __restrict__ done) {
  // FP32 products unroll the slots by 4 (the residual: 124 -> 112 us on the 1.6M-dof
  // cantilever), the Chebyshev step, with its larger epilogue, by 2 (119.5 against 124.6 us);
  // FP16 storage keeps 2 (its residual: 74 us unrolled by 2, 78 us by 4)
  constexpr int W = AmgSliceCfg<T>::W;
  constexpr int U = sizeof(T) == 4 ? (MODE == AP_CHEB ? 2 : 4) : SliceCfg<T>::U;
  __shared__ double red[W][kSlice][V];
  if ((done && *done) || (a.done && *a.done)) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // XCD-contiguous slice order: neighbouring slices (which gather the same
  // vector entries) run on one XCD and share its L2
  const int64_t s = xcd_logical_block(blockIdx.x, gridDim.x);
  const int64_t row = pat.sell_row[s * kSlice + lane];
  const bool real = row >= 0;
  auto xat = [&](int64_t c, TX (&xc)[V]) {
#pragma unroll
    for (int j = 0; j < V; ++j) xc[j] = a.x[c * V + j];
  };
  double xr[V], acc[V];
#pragma unroll
  for (int j = 0; j < V; ++j) xr[j] = (DIFF && real) ? double(a.x[row * V + j]) : 0.0;
  constexpr bool kRhs = MODE == AP_RESID || MODE == AP_CHEB;
  // FP16 values hold D^-1 A (rows scaled by dinv): the epilogues undo or skip the scaling
  constexpr bool kScaled = sizeof(T) == 2;
  // The epilogue's own-row operands are loaded before the matrix stream, so their
  // latency overlaps it.  With a thread per (row, component) pair (kDist: 4 waves),
  // thread p finishes component p % V of lane p / V after the partial sums meet in
  // LDS, holding 4 operands; otherwise wave 0 finishes its lanes' rows (the outer
  // product's <p, q> partial is then one wave sum).
  constexpr bool kDist = MODE != AP_PQ && AmgSliceCfg<T>::NT >= kSlice * V;
  const int pl = threadIdx.x / V, pi = threadIdx.x % V;
  int64_t prow = -1;
  double p_rhs = 0.0, p_dinv = 0.0, p_d = 0.0, p_x = 0.0;
  if (kDist && threadIdx.x < kSlice * V) {
    prow = pat.sell_row[s * kSlice + pl];
    if (prow >= 0) {
      const int64_t idx = prow * V + pi;
      p_rhs = kRhs ? a.rhs[idx] : 0.0;
      p_dinv = (MODE == AP_CHEB || MODE == AP_DSTORE || (kScaled && MODE == AP_RESID)) ? a.dinv[idx] : 0.0;
      p_d = (MODE == AP_CHEB && a.c1 != 0.0) ? double(a.d[idx]) : 0.0;
      p_x = MODE == AP_CHEB ? double(a.x[idx]) : 0.0;
    }
  }
  double e_rhs[V], e_dinv[V], e_d[V], e_x[V];
  if (!kDist && w == 0 && real) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      const int64_t idx = row * V + i;
      e_rhs[i] = kRhs ? a.rhs[idx] : 0.0;
      e_dinv[i] = (MODE == AP_CHEB || MODE == AP_DSTORE || (kScaled && MODE == AP_RESID)) ? a.dinv[idx] : 0.0;
      e_d[i] = (MODE == AP_CHEB && a.c1 != 0.0) ? double(a.d[idx]) : 0.0;
      e_x[i] = (MODE == AP_CHEB || MODE == AP_PQ) ? double(a.x[idx]) : 0.0;
    }
  }
  sell_slice_partial<V, DIFF, W, U, T, decltype(xat), TX>(pat, vals, s, w, lane, row, xat, xr, acc);
  if (kDist || w > 0) {
#pragma unroll
    for (int i = 0; i < V; ++i) red[w][lane][i] = acc[i];
  }
  __syncthreads();
  if constexpr (kDist) {
    if (prow >= 0) {
      double t = red[0][pl][pi];
#pragma unroll
      for (int o = 1; o < W; ++o) t += red[o][pl][pi];
      if (kScaled) t *= 1.0 / kF16Scale;
      const int64_t idx = prow * V + pi;
      if (MODE == AP_CHEB) {
        const double r = kScaled ? fma(p_dinv, p_rhs, -t) : p_dinv * (p_rhs - t);
        const double dn = a.c1 != 0.0 ? fma(a.c1, p_d, a.c2 * r) : a.c2 * r;
        a.d[idx] = TX(dn);
        a.y[idx] = TY(p_x + dn);
      } else if (MODE == AP_RESID) {
        a.y[idx] = TY(kScaled ? p_rhs - t / p_dinv : p_rhs - t);
      } else if (MODE == AP_DSTORE) {
        a.y[idx] = TY(kScaled ? t : p_dinv * t);
      } else {
        apply_epilogue<MODE>(a, idx, t);
      }
    }
    return;
  }
  if (w != 0) return;
  double pq = 0.0;
  if (real) {
#pragma unroll
    for (int i = 0; i < V; ++i) {
      double t = acc[i];
#pragma unroll
      for (int o = 1; o < W; ++o) t += red[o][lane][i];
      if (kScaled) t *= 1.0 / kF16Scale;
      const int64_t idx = row * V + i;
      if (MODE == AP_CHEB) {
        const double r = kScaled ? fma(e_dinv[i], e_rhs[i], -t) : e_dinv[i] * (e_rhs[i] - t);
        const double dn = a.c1 != 0.0 ? fma(a.c1, e_d[i], a.c2 * r) : a.c2 * r;
        a.d[idx] = TX(dn);
        a.y[idx] = TY(e_x[i] + dn);
      } else if (MODE == AP_RESID) {
        a.y[idx] = TY(kScaled ? e_rhs[i] - t / e_dinv[i] : e_rhs[i] - t);
      } else if (MODE == AP_DSTORE) {
        a.y[idx] = TY(kScaled ? t : e_dinv[i] * t);
      } else {
        apply_epilogue<MODE>(a, idx, t);
      }
      if (MODE == AP_PQ) pq = fma(e_x[i], t, pq);
    }
  }
  if (MODE == AP_PQ) {
    pq = wave_sum(pq);
    if (lane == 0) a.part[s] = pq;
  }
}

// first Chebyshev step from a zero iterate: d = c2 D^-1 rhs, x = d
template <typename TX>
__global__ void k_cheb_first(int64_t n, const double* __restrict__ rhs, const double* __restrict__ dinv, double c2,
                             TX* __restrict__ d, TX* __restrict__ x, const int32_t* __restrict__ done) {
  if (done && *done) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const TX v = TX(c2 * dinv[i] * rhs[i]);
    d[i] = v;
    x[i] = v;
  }
}

// power-iteration start vector: hash-derived values in [0.5, 1.5)
template <typename T>
__global__ void k_start_vec(int64_t n, T* __restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = T(0.5 + (double)(amg_hash((uint32_t)i) >> 8) / 16777216.0);
}

// x = y / sqrt(*nrm2)
template <typename T>
__global__ void k_scale_by_norm(int64_t n, const T* __restrict__ y, const double* __restrict__ nrm2,
                                T* __restrict__ x) {
  const double inv = 1.0 / sqrt(*nrm2);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = T(double(y[i]) * inv);
}

// *out = sqrt(nrm[0] / nrm[1])
__global__ void k_lmax_ratio(const double* __restrict__ nrm, double* __restrict__ out) {
  if (threadIdx.x == 0) *out = sqrt(nrm[0] / nrm[1]);
}

// per-block partials of <x, x> for FP32 x (FP64 sums)
__global__ __launch_bounds__(256) void k_norm2_part_f32(int64_t n, const float* __restrict__ x,
                                                        double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[i];
    s = fma(v, v, s);
  }
  s = block_sum_256(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---------------------------------------------------------------- CG
struct AmgScalars {
  double rz, alpha, beta, rr, stop2, bb;
  int32_t done, iter, nan, first, maxiter, pad;
};

__global__ __launch_bounds__(256) void k_dot_part(int64_t n, const double* __restrict__ x,
                                                  const double* __restrict__ y, double* __restrict__ part,
                                                  const int32_t* __restrict__ done) {
  __shared__ double red[4];
  if (done && *done) return;
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s = fma(x[i], y[i], s);
  s = block_sum_256(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__device__ inline double sum_parts(const double* part, int np, double* red) {
  // four independent sums keep four loads in flight per thread (a fixed order)
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int i = threadIdx.x;
  for (; i + 768 < np; i += 1024) {
    s0 += part[i];
    s1 += part[i + 256];
    s2 += part[i + 512];
    s3 += part[i + 768];
  }
  for (; i < np; i += 256) s0 += part[i];
  return block_sum_256((s0 + s1) + (s2 + s3), red);
}

__global__ __launch_bounds__(256) void k_finish(const double* __restrict__ part, int np, double* __restrict__ out) {
  __shared__ double red[4];
  const double s = sum_parts(part, np, red);
  if (threadIdx.x == 0) *out = s;
}

// start: sc->rr holds <r0, r0>, sc->bb holds <b, b>
__global__ void k_cg_start(AmgScalars* sc, double rtol, double atol, int maxiter) {
  const double stop = fmax(rtol * sqrt(sc->bb), atol);
  sc->stop2 = stop * stop;
  sc->iter = 0;
  sc->first = 1;
  sc->maxiter = maxiter;
  sc->nan = !isfinite(sc->rr);
  sc->done = sc->nan || sc->rr <= sc->stop2 || maxiter <= 0;
}

__global__ __launch_bounds__(256) void k_cg_beta(const double* __restrict__ part, int np, AmgScalars* sc) {
  __shared__ double red[4];
  if (sc->done) return;
  const double rz = sum_parts(part, np, red);
  if (threadIdx.x == 0) {
    if (!isfinite(rz)) {
      sc->nan = 1;
      sc->done = 1;
      return;
    }
    sc->beta = sc->first ? 0.0 : rz / sc->rz;
    sc->rz = rz;
    sc->first = 0;
  }
}

__global__ void k_cg_p(int64_t n, const double* __restrict__ z, double* __restrict__ p, const AmgScalars* sc) {
  if (sc->done) return;
  const double beta = sc->beta;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = beta == 0.0 ? z[i] : fma(beta, p[i], z[i]);
}

__global__ __launch_bounds__(256) void k_cg_alpha(const double* __restrict__ part, int np, AmgScalars* sc) {
  __shared__ double red[4];
  if (sc->done) return;
  const double pq = sum_parts(part, np, red);
  if (threadIdx.x == 0) {
    if (!isfinite(pq) || pq == 0.0) {
      sc->nan = !isfinite(pq);
      sc->done = 1;
      sc->alpha = 0.0;
      return;
    }
    sc->alpha = sc->rz / pq;
  }
}

__global__ __launch_bounds__(256) void k_cg_xr(int64_t n, const double* __restrict__ p, const double* __restrict__ q,
                                               double* __restrict__ x, double* __restrict__ r,
                                               double* __restrict__ part, const AmgScalars* sc) {
  __shared__ double red[4];
  if (sc->done) return;
  const double alpha = sc->alpha;
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] = fma(alpha, p[i], x[i]);
    const double ri = fma(-alpha, q[i], r[i]);
    r[i] = ri;
    s = fma(ri, ri, s);
  }
  s = block_sum_256(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// after an FP64 residual replacement (sc->rr = |b - A x|^2): the stop test
__global__ void k_cg_replaced(AmgScalars* sc) {
  sc->nan = sc->nan || !isfinite(sc->rr);
  sc->done = sc->nan || sc->rr <= sc->stop2 || sc->iter >= sc->maxiter;
}

__global__ __launch_bounds__(256) void k_cg_rr(const double* __restrict__ part, int np, AmgScalars* sc) {
  __shared__ double red[4];
  if (sc->done) return;
  const double rr = sum_parts(part, np, red);
  if (threadIdx.x == 0) {
    sc->rr = rr;
    sc->iter += 1;
    if (!isfinite(rr)) {
      sc->nan = 1;
      sc->done = 1;
    } else if (rr <= sc->stop2 || sc->iter >= sc->maxiter) {
      sc->done = 1;
    }
  }
}

// ------------------------------------------------------------ dispatch
// Launch shapes of the block CSR operators, from the mean blocks per row:
//   - scalar (1x1) hierarchies: G lanes strided over a row's entries (k_bcsr_apply);
//   - long rows of few block rows (the restriction to and the operator of a small
//     coarse level): one 256-thread workgroup per row (k_bcsr_apply_row), since one
//     wave per row would leave most of the chip idle;
//   - rows of more than kWaveBlocks blocks (the level-1 operator: 45, the level-1
//     restriction: 195 on the 1.6M-dof cantilever): one wave per row, a lane per block;
//   - short rows (the prolongations: 4 to 5 blocks): kSmallG lanes per row (at least
//     the block's row count), a lane per block.
constexpr double kWaveBlocks = 12.0;
constexpr double kRowBlocks = 64.0;
constexpr int64_t kRowKernelRows = 4096;
constexpr int kSmallG = 4;

template <int BR, int BC, int MODE, typename T, typename TX, typename TY>
void bcsr_apply_shape(const fem_bcsr_t& M, const T* v, const ApplyArgsT<TX, TY>& a, hipStream_t s) {
  const double bpr = double(M.nnzb) / double(M.rows);
  constexpr int GS = kSmallG >= BR ? kSmallG : 8;
  if (bpr > kRowBlocks && M.rows < kRowKernelRows)
    hipLaunchKernelGGL((k_bcsr_apply_row<BR, BC, MODE, T, TX, TY>), dim3((unsigned)M.rows), dim3(256), 0, s, M.rows,
                       M.row_ptr, M.col_idx, v, a);
  else if (bpr > kWaveBlocks)
    hipLaunchKernelGGL((k_bcsr_apply_blk<BR, BC, 64, MODE, T, TX, TY>), dim3((unsigned)ceil_div(M.rows * 64, 256)),
                       dim3(256), 0, s, M.rows, M.row_ptr, M.col_idx, v, a);
  else
    hipLaunchKernelGGL((k_bcsr_apply_blk<BR, BC, GS, MODE, T, TX, TY>), dim3((unsigned)ceil_div(M.rows * GS, 256)),
                       dim3(256), 0, s, M.rows, M.row_ptr, M.col_idx, v, a);
}

template <int MODE, typename T, typename TX = double, typename TY = double>
int bcsr_apply(const fem_bcsr_t& M, const ApplyArgsT<TX, TY>& a, hipStream_t s) {
  if (M.rows == 0) return FEM_OK;
  const T* v = static_cast<const T*>(M.vals);
  if (M.br == 1 && M.bc == 1) {
    // lanes per row from the mean entries per row (a whole workgroup above 400)
    const double per_row = double(M.nnzb) / double(M.rows);
    if (per_row > 400.0)
      hipLaunchKernelGGL((k_bcsr_apply_row<1, 1, MODE, T, TX, TY>), dim3((unsigned)M.rows), dim3(256), 0, s, M.rows,
                         M.row_ptr, M.col_idx, v, a);
    else if (per_row > 48.0)
      hipLaunchKernelGGL((k_bcsr_apply<1, 1, 64, MODE, T, TX, TY>), dim3((unsigned)ceil_div(M.rows * 64, 256)),
                         dim3(256), 0, s, M.rows, M.row_ptr, M.col_idx, v, a);
    else
      hipLaunchKernelGGL((k_bcsr_apply<1, 1, 8, MODE, T, TX, TY>), dim3((unsigned)ceil_div(M.rows * 8, 256)),
                         dim3(256), 0, s, M.rows, M.row_ptr, M.col_idx, v, a);
  } else if (M.br == 3 && M.bc == 3) {
    bcsr_apply_shape<3, 3, MODE, T>(M, v, a, s);
  } else if (M.br == 6 && M.bc == 6) {
    bcsr_apply_shape<6, 6, MODE, T>(M, v, a, s);
  } else if (M.br == 3 && M.bc == 6) {
    bcsr_apply_shape<3, 6, MODE, T>(M, v, a, s);
  } else if (M.br == 6 && M.bc == 3) {
    bcsr_apply_shape<6, 3, MODE, T>(M, v, a, s);
  } else {
    set_error("block CSR operator: unsupported block shape %d x %d", M.br, M.bc);
    return FEM_ERR_BAD_ARG;
  }
  FEM_LAUNCHED();
  return FEM_OK;
}

template <int MODE, typename T, bool DIFF = false, typename TX = double, typename TY = double>
int sell_apply(const fem_pattern_t& pat, const void* vals, const ApplyArgsT<TX, TY>& a, const int32_t* done,
               hipStream_t s) {
  if (pat.num_slices == 0) return FEM_OK;
  const unsigned grid = (unsigned)pat.num_slices;
  const T* v = static_cast<const T*>(vals);
  if (pat.vec == 3)
    hipLaunchKernelGGL((k_sell_apply<3, MODE, T, DIFF, TX, TY>), dim3(grid), dim3(AmgSliceCfg<T>::NT), 0, s, pat, v, a,
                       done);
  else if (pat.vec == 1)
    hipLaunchKernelGGL((k_sell_apply<1, MODE, T, DIFF, TX, TY>), dim3(grid), dim3(AmgSliceCfg<T>::NT), 0, s, pat, v, a,
                       done);
  else
    FEM_CHECK_ARG(false, "AMG: vec must be 1 or 3");
  FEM_LAUNCHED();
  return FEM_OK;
}

int64_t level_n(const fem_amg_t& amg, int l) {
  if (l == 0) return amg.levels[0].pat->num_rows * amg.levels[0].pat->vec;
  return amg.levels[l].A.rows * amg.levels[l].A.br;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Chebyshev degree of level l: amg.degree on level 0, 2 amg.degree + 1 (at most 9) on
// the coarser levels, whose products cost a fifth of level 0's or less.  Against the
// level-0 degree (2) everywhere: on the 1.6M-dof HEX8 cantilever 47 -> 36 CG iterations
// (coarse degree 3, 4, 5, 6: 39, 37, 36, 35 iterations at 51.1, 50.7, 50.6, 50.9 ms per
// step, with its 882-unknown dense coarse level), the SIMP cantilever 68 -> 47, the HEX27
// cube 107 -> 77 (degree 4: 51 and 83).
inline int level_degree(const fem_amg_t& amg, int l) { return l == 0 ? amg.degree : std::min(2 * amg.degree + 1, 9); }

// Chebyshev on [lo, hi] = [lmax / 15, 1.1 lmax].  A sweep of the lower bound
// (lmax / 12 ... / 50, upper 1.05 ... 1.3 lmax) over the HEX8 cantilever
// (1.6M dofs), the SIMP cantilever, the HEX27 cube, the Neo-Hookean cube and the
// Poisson cube: lmax / 15 cut the CG iterations on every one against lmax / 30
// (53 -> 48, 76 -> 68, 125 -> 113, 261 -> 254, 13 -> 10); lmax / 20 took 47 on the
// cantilever but 130 on the HEX27 cube.
struct Cheb {
  double theta, delta, sigma;
  explicit Cheb(double lmax) {
    const double hi = 1.1 * lmax, lo = lmax / 15.0;
    theta = 0.5 * (hi + lo);
    delta = 0.5 * (hi - lo);
    sigma = theta / delta;
  }
};

template <typename T, typename TX = double, typename TY = double>
int level_apply(const fem_amg_t& amg, int l, int mode, const ApplyArgsT<TX, TY>& a, hipStream_t s) {
  const fem_amg_level_t& L = amg.levels[l];
  if (l == 0) {
    if (amg.level0_form == FEM_FORM_SCALED_DIFF16) {
      ApplyArgsT<TX, TY> b = a;
      b.dinv = L.dinv;   // the residual undoes the row scaling
      if (mode == AP_CHEB) return sell_apply<AP_CHEB, _Float16, true>(*L.pat, L.sell_vals, b, a.done, s);
      return sell_apply<AP_RESID, _Float16, true>(*L.pat, L.sell_vals, b, a.done, s);
    }
    if (amg.level0_form == FEM_FORM_DIFFERENCE) {
      if (mode == AP_CHEB) return sell_apply<AP_CHEB, T, true>(*L.pat, L.sell_vals, a, a.done, s);
      return sell_apply<AP_RESID, T, true>(*L.pat, L.sell_vals, a, a.done, s);
    }
    if (mode == AP_CHEB) return sell_apply<AP_CHEB, T>(*L.pat, L.sell_vals, a, a.done, s);
    return sell_apply<AP_RESID, T>(*L.pat, L.sell_vals, a, a.done, s);
  }
  if (mode == AP_CHEB) return bcsr_apply<AP_CHEB, T>(L.A, a, s);
  return bcsr_apply<AP_RESID, T>(L.A, a, s);
}

// Pre-smoothing (Chebyshev from a zero iterate), residual and restriction on
// level l, vectors stored as TX.  On return *xo is the smoothed iterate and
// *oo the other iterate buffer.  ev: see vcycle_impl.
template <typename T, typename TX>
int level_down(const fem_amg_t& amg, int l, int64_t n, const double* rhs, double* rhs_next, TX* x, TX* other,
               TX* dd, TX* res, TX** xo, TX** oo, hipStream_t s, hipEvent_t* ev, const int32_t* done) {
  const fem_amg_level_t& L = amg.levels[l];
  const Cheb cb(L.lmax);
  hipLaunchKernelGGL(k_cheb_first<TX>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, rhs, L.dinv, 1.0 / cb.theta, dd,
                     x, done);
  FEM_LAUNCHED();
  double rho = 1.0 / cb.sigma;
  const int deg = level_degree(amg, l);
  for (int k = 1; k < deg; ++k) {
    const double rn = 1.0 / (2.0 * cb.sigma - rho);
    ApplyArgsT<TX, TX> a{x, other, rhs, L.dinv, dd, rn * rho, 2.0 * rn / cb.delta, nullptr, done};
    const bool timed = ev && l == 0 && k == 1;
    if (timed) FEM_HIP(hipEventRecord(ev[4], s));
    if (int rc = level_apply<T>(amg, l, AP_CHEB, a, s)) return rc;
    if (timed) FEM_HIP(hipEventRecord(ev[5], s));
    rho = rn;
    std::swap(x, other);
  }
  ApplyArgsT<TX, TX> ar{x, res, rhs, nullptr, nullptr, 0.0, 0.0, nullptr, done};
  if (int rc = level_apply<T>(amg, l, AP_RESID, ar, s)) return rc;
  ApplyArgsT<TX, double> arr{res, rhs_next, nullptr, nullptr, nullptr, 0.0, 0.0, nullptr, done};
  if (int rc = bcsr_apply<AP_STORE, T>(L.R, arr, s)) return rc;
  *xo = x;
  *oo = other;
  return FEM_OK;
}

// Prolongation of the coarse correction xc into x and post-smoothing on level
// l.  z (level 0 only, FP64): the output of the last Chebyshev step.  ev: see
// vcycle_impl.
template <typename T, typename TX>
int level_up(const fem_amg_t& amg, int l, const double* rhs, const double* xc, TX* x, TX* other, TX* dd, double* z,
             TX** xo, hipStream_t s, hipEvent_t* ev, const int32_t* done) {
  const fem_amg_level_t& L = amg.levels[l];
  const Cheb cb(L.lmax);
  ApplyArgsT<double, TX> ap{xc, x, nullptr, nullptr, nullptr, 0.0, 0.0, nullptr, done};
  if (int rc = bcsr_apply<AP_ADD, T>(L.P, ap, s)) return rc;
  double rho = 1.0 / cb.sigma;
  const int deg = level_degree(amg, l);
  for (int k = 0; k < deg; ++k) {
    double c1 = 0.0, c2 = 1.0 / cb.theta;
    if (k > 0) {
      const double rn = 1.0 / (2.0 * cb.sigma - rho);
      c1 = rn * rho;
      c2 = 2.0 * rn / cb.delta;
      rho = rn;
    }
    const bool timed = ev && l == 0 && k == 0;
    const bool timed_last = ev && l == 0 && k > 0 && k == deg - 1;
    if (timed) FEM_HIP(hipEventRecord(ev[0], s));
    if (timed_last) FEM_HIP(hipEventRecord(ev[6], s));
    if (z && k == deg - 1) {
      ApplyArgsT<TX, double> a{x, z, rhs, L.dinv, dd, c1, c2, nullptr, done};
      if (int rc = level_apply<T>(amg, l, AP_CHEB, a, s)) return rc;
    } else {
      ApplyArgsT<TX, TX> a{x, other, rhs, L.dinv, dd, c1, c2, nullptr, done};
      if (int rc = level_apply<T>(amg, l, AP_CHEB, a, s)) return rc;
      std::swap(x, other);
    }
    if (timed) FEM_HIP(hipEventRecord(ev[1], s));
    if (timed_last) FEM_HIP(hipEventRecord(ev[7], s));
  }
  *xo = x;
  return FEM_OK;
}

// ev (may be NULL): events recorded around the whole cycle (ev[2], ev[3]) and
// around its level-0 Chebyshev SELL kernels: the first post-smoothing step
// (ev[0], ev[1]) and, with a level-0 degree of at least 2, the pre-smoothing
// step after k_cheb_first (ev[4], ev[5]) and the last post-smoothing step
// (ev[6], ev[7]).  With no levels ev[0], ev[1] bracket the dense solve.
// Vectors are FP64 on every level, whatever the storage T of the matrices:
// FP32 level-0 iterates (which move 20 MB less per level-0 product on the
// 1.6M-dof cantilever) round the differences x_j - x_i the difference form
// multiplies, and on the high-contrast SIMP cantilever (160x80x40, E ratio
// 10^3) the adjoint solve then took 84 instead of 76 iterations (106 with the
// power iteration on FP32 iterates as well).
template <typename T>
int vcycle_impl(const fem_amg_t& amg, const double* r, double* z, char* ws, hipStream_t s,
                hipEvent_t* ev = nullptr, const int32_t* done = nullptr) {
  if (ev) FEM_HIP(hipEventRecord(ev[2], s));
  const int NL = amg.num_levels;
  std::vector<double*> rhs(NL + 1), xa(NL), xb(NL), dd(NL), res(NL), cur(NL + 1);
  size_t off = 0;
  auto take = [&](int64_t n) {
    double* p = reinterpret_cast<double*>(ws + off);
    off += align256((size_t)std::max<int64_t>(n, 1) * 8);
    return p;
  };
  for (int l = 0; l < NL; ++l) {
    const int64_t n = level_n(amg, l);
    rhs[l] = l == 0 ? const_cast<double*>(r) : take(n);
    xa[l] = take(n);
    xb[l] = take(n);
    dd[l] = take(n);
    res[l] = take(n);
  }
  // with no levels the cycle is the dense solve of the whole matrix
  rhs[NL] = NL == 0 ? const_cast<double*>(r) : take(amg.coarse_n);
  cur[NL] = NL == 0 ? z : take(amg.coarse_n);
  // down
  for (int l = 0; l < NL; ++l) {
    const int64_t n = level_n(amg, l);
    double *x = nullptr, *o = nullptr;
    if (int rc = level_down<T, double>(amg, l, n, rhs[l], rhs[l + 1], xa[l], xb[l], dd[l], res[l], &x, &o, s,
                                       ev, done))
      return rc;
    cur[l] = x;
    xb[l] = o;
  }
  // coarsest
  if (ev && NL == 0) FEM_HIP(hipEventRecord(ev[0], s));
  hipLaunchKernelGGL(k_dense_gemv, dim3((unsigned)ceil_div(amg.coarse_n * 64, 256)), dim3(256), 0, s, amg.coarse_n,
                     amg.coarse_inv, rhs[NL], cur[NL], done);
  FEM_LAUNCHED();
  if (ev && NL == 0) FEM_HIP(hipEventRecord(ev[1], s));
  // up (level 0's last Chebyshev step writes the cycle's output z)
  for (int l = NL - 1; l >= 0; --l) {
    double* x = nullptr;
    if (int rc = level_up<T, double>(amg, l, rhs[l], cur[l + 1], cur[l], xb[l], dd[l], l == 0 ? z : nullptr, &x, s,
                                     ev, done))
      return rc;
    cur[l] = x;
  }
  if (ev) FEM_HIP(hipEventRecord(ev[3], s));
  return FEM_OK;
}

int check_amg(const fem_amg_t* amg) {
  FEM_CHECK_ARG(amg && (amg->levels || amg->num_levels == 0) && amg->num_levels >= 0 && amg->coarse_inv &&
                    amg->coarse_n > 0,
                "AMG hierarchy descriptor incomplete");
  FEM_CHECK_ARG(amg->precision == FEM_F32 || amg->precision == FEM_F64, "AMG precision must be FEM_F32 or FEM_F64");
  FEM_CHECK_ARG(amg->degree >= 1 && amg->degree <= 8, "AMG Chebyshev degree must be in [1, 8]");
  FEM_CHECK_ARG(amg->level0_form == FEM_FORM_STANDARD ||
                    ((amg->level0_form == FEM_FORM_DIFFERENCE || amg->level0_form == FEM_FORM_SCALED_DIFF16) &&
                     amg->precision == FEM_F32),
                "AMG level 0: the difference form is an FP32 storage form");
  if (amg->num_levels == 0) return FEM_OK;   // the dense coarse inverse is the whole preconditioner
  const fem_amg_level_t& L0 = amg->levels[0];
  FEM_CHECK_ARG(L0.pat && L0.sell_vals && L0.dinv, "AMG level 0 needs the SELL pattern, values and dinv");
  for (int l = 0; l < amg->num_levels; ++l) {
    const fem_amg_level_t& L = amg->levels[l];
    FEM_CHECK_ARG(L.dinv && L.P.vals && L.R.vals && L.lmax > 0.0, "AMG level %d incomplete", l);
    if (l > 0) FEM_CHECK_ARG(L.A.vals && L.A.br == L.A.bc, "AMG level %d operator missing", l);
    const int64_t next = l + 1 < amg->num_levels ? amg->levels[l + 1].A.rows * amg->levels[l + 1].A.br
                                                 : amg->coarse_n;
    FEM_CHECK_ARG(L.P.rows * L.P.br == level_n(*amg, l) && L.P.cols * L.P.bc == next &&
                      L.R.rows * L.R.br == next && L.R.cols * L.R.bc == level_n(*amg, l),
                  "AMG level %d: prolongator shapes do not match the levels", l);
  }
  return FEM_OK;
}

int64_t vcycle_ws(const fem_amg_t& amg) {
  size_t b = 0;
  for (int l = 0; l < amg.num_levels; ++l) b += 5 * align256((size_t)std::max<int64_t>(level_n(amg, l), 1) * 8);
  b += 2 * align256((size_t)std::max<int64_t>(amg.coarse_n, 1) * 8);
  return (int64_t)b;
}

BView view(const fem_bcsr_t& M) { return BView{M.rows, M.row_ptr, M.col_idx}; }

// numeric SpGEMM and prolongator smoothing on FP64 or FP32 block values
template <typename T>
int spgemm_numeric_t(const fem_bcsr_t& X, const fem_bcsr_t& Y, const fem_bcsr_t& C, const int32_t* plan_ptr,
                     const int32_t* plan, hipStream_t s);
template <typename T>
int smooth_t(const fem_bcsr_t& A, const fem_bcsr_t& P0, const fem_bcsr_t& P, const int32_t* plan_ptr,
             const int32_t* plan, const int32_t* agg, const double* dinv, double omega, const double* lmax_dev,
             hipStream_t s);

template <bool FILL>
int spgemm_dispatch(int64_t max_pairs, const fem_bcsr_t& X, const fem_bcsr_t& Y, int32_t* counts, const PlanOut& o,
                    int32_t* err, hipStream_t s) {
  const BView xv = view(X), yv = view(Y);
  const dim3 grid((unsigned)X.rows);
  if (max_pairs <= 256)
    hipLaunchKernelGGL((k_spgemm_row<64, 256, FILL>), grid, dim3(64), 0, s, xv, yv, counts, o, err);
  else if (max_pairs <= 1024)
    hipLaunchKernelGGL((k_spgemm_row<256, 1024, FILL>), grid, dim3(256), 0, s, xv, yv, counts, o, err);
  else if (max_pairs < kBigPairs)
    hipLaunchKernelGGL((k_spgemm_row_big<256, FILL>), grid, dim3(256), 0, s, xv, yv, counts, o, err);
  else {
    set_error("block SpGEMM: a row has %lld block products (capacity %d)", (long long)max_pairs, kBigPairs - 1);
    return FEM_ERR_CAPACITY;
  }
  FEM_LAUNCHED();
  return FEM_OK;
}

// read `bytes` (<= 128) from the device through the pinned status buffer
int read_dev(const void* dev, size_t bytes, void* host_status, hipStream_t s, void* out) {
  FEM_HIP(hipMemcpyAsync(host_status, dev, bytes, hipMemcpyDeviceToHost, s));
  FEM_HIP(hipStreamSynchronize(s));
  memcpy(out, host_status, bytes);
  return FEM_OK;
}
int read_int(const int32_t* dev, void* host_status, hipStream_t s, int32_t* out) {
  return read_dev(dev, sizeof(int32_t), host_status, s, out);
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" {

int64_t fem_amg_aggregate_ws_bytes(int64_t n) {
  const int64_t m = std::max<int64_t>(n, 1);
  return (int64_t)(align256(m) + align256(m * 8) + 2 * align256(m * 4) + align256((m + 1) * 4) +
                   align256((size_t)fem_scan_ws_bytes(m)) + 256);
}

int fem_amg_aggregate(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, int32_t* agg, int64_t* n_agg,
                      void* ws, void* host_status, void* stream) {
  FEM_CHECK_ARG(n >= 0 && agg && n_agg && ws && host_status && (n == 0 || (row_ptr && col_idx)),
                "fem_amg_aggregate: bad arguments");
  FEM_CHECK_ARG(n < (int64_t(1) << 31), "fem_amg_aggregate: more than 2^31 nodes");
  hipStream_t s = as_stream(stream);
  *n_agg = 0;
  if (n == 0) return FEM_OK;
  char* w = static_cast<char*>(ws);
  int8_t* state = reinterpret_cast<int8_t*>(w);
  w += align256(n);
  uint64_t* t1 = reinterpret_cast<uint64_t*>(w);
  w += align256(n * 8);
  int32_t* flags = reinterpret_cast<int32_t*>(w);
  w += align256(n * 4);
  int32_t* root_id = reinterpret_cast<int32_t*>(w);
  w += align256((n + 1) * 4);
  int32_t* flag = reinterpret_cast<int32_t*>(w);
  w += 256;
  int32_t* tmp = reinterpret_cast<int32_t*>(w);
  w += align256(n * 4);
  void* scan_ws = w;
  const unsigned g = grid_for(n, 256);
  hipLaunchKernelGGL(k_mis_init, dim3(g), dim3(256), 0, s, n, state);
  FEM_LAUNCHED();
  for (int round = 0;; ++round) {
    FEM_CHECK_ARG(round < 10000, "fem_amg_aggregate: MIS-2 rounds did not terminate");
    hipLaunchKernelGGL(k_mis_prop1, dim3(g), dim3(256), 0, s, n, row_ptr, col_idx, state, t1, flag);
    FEM_LAUNCHED();
    hipLaunchKernelGGL(k_mis_prop2, dim3(g), dim3(256), 0, s, n, row_ptr, col_idx, t1, state, flag);
    FEM_LAUNCHED();
    int32_t more = 0;
    if (int rc = read_int(flag, host_status, s, &more)) return rc;
    if (!more) break;
  }
  hipLaunchKernelGGL(k_root_flags, dim3(g), dim3(256), 0, s, n, state, flags);
  FEM_LAUNCHED();
  if (int rc = fem_exclusive_scan(n, flags, root_id, scan_ws, stream)) return rc;
  hipLaunchKernelGGL(k_agg_pass1, dim3(g), dim3(256), 0, s, n, row_ptr, col_idx, state, root_id, tmp);
  FEM_LAUNCHED();
  FEM_HIP(hipMemsetAsync(flag, 0, sizeof(int32_t), s));
  hipLaunchKernelGGL(k_agg_pass2, dim3(g), dim3(256), 0, s, n, row_ptr, col_idx, tmp, agg, flag);
  FEM_LAUNCHED();
  int32_t bad = 0, total = 0;
  if (int rc = read_int(flag, host_status, s, &bad)) return rc;
  if (int rc = read_int(root_id + n, host_status, s, &total)) return rc;
  if (bad) {
    set_error("fem_amg_aggregate: a node was left without an aggregate");
    return FEM_ERR_BAD_ARG;
  }
  *n_agg = total;
  return FEM_OK;
}

int fem_amg_members(int64_t n, const int32_t* agg, int64_t n_agg, int32_t* agg_ptr, int32_t* memb, void* ws,
                    void* stream) {
  FEM_CHECK_ARG(n >= 0 && n_agg >= 0 && agg_ptr && ws && (n == 0 || (agg && memb)), "fem_amg_members: bad arguments");
  hipStream_t s = as_stream(stream);
  char* w = static_cast<char*>(ws);
  int32_t* cnt = reinterpret_cast<int32_t*>(w);
  w += align256(std::max<int64_t>(n_agg, 1) * 4);
  void* scan_ws = w;
  FEM_HIP(hipMemsetAsync(cnt, 0, std::max<int64_t>(n_agg, 1) * 4, s));
  if (n > 0) {
    hipLaunchKernelGGL(k_count_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, n, agg, cnt);
    FEM_LAUNCHED();
  }
  if (int rc = fem_exclusive_scan(n_agg, cnt, agg_ptr, scan_ws, stream)) return rc;
  if (n == 0) return FEM_OK;
  FEM_HIP(hipMemsetAsync(cnt, 0, std::max<int64_t>(n_agg, 1) * 4, s));
  hipLaunchKernelGGL(k_fill_members, dim3(grid_for(n, 256)), dim3(256), 0, s, n, agg, agg_ptr, cnt, memb);
  FEM_LAUNCHED();
  hipLaunchKernelGGL(k_sort_segments, dim3(grid_for(n_agg, 256)), dim3(256), 0, s, n_agg, agg_ptr, memb,
                     (int32_t*)nullptr);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_amg_tentative(int64_t n, int32_t bs, int32_t ns, int64_t n_agg, const int32_t* agg_ptr, const int32_t* memb,
                      const double* points, const uint8_t* bc_mask, const double* B, double* P0, double* Bc,
                      void* stream) {
  FEM_CHECK_ARG(n >= 0 && n_agg >= 0 && agg_ptr && P0 && Bc && (n == 0 || memb), "fem_amg_tentative: bad arguments");
  FEM_CHECK_ARG(B || points || ns == 1, "fem_amg_tentative: level 0 with 6 modes needs the node coordinates");
  hipStream_t s = as_stream(stream);
  if (n_agg == 0) return FEM_OK;
  const unsigned grid = (unsigned)ceil_div(n_agg * 64, 256);
#define FEM_AMG_TENT(BS, NS)                                                                                  \
  if (bs == BS && ns == NS) {                                                                                 \
    hipLaunchKernelGGL((k_tentative<BS, NS>), dim3(grid), dim3(256), 0, s, n_agg, agg_ptr, memb, points, bc_mask, \
                       B, P0, Bc);                                                                            \
    FEM_LAUNCHED();                                                                                           \
    return FEM_OK;                                                                                            \
  }
  if (!B) {
    FEM_CHECK_ARG((bs == 3 && ns == 6) || (bs == 1 && ns == 1), "fem_amg_tentative: level 0 needs vec 3 (6 modes) "
                  "or vec 1 (1 mode)");
  }
  FEM_AMG_TENT(3, 6)
  FEM_AMG_TENT(6, 6)
  FEM_AMG_TENT(1, 1)
#undef FEM_AMG_TENT
  set_error("fem_amg_tentative: unsupported block size %d with %d modes", bs, ns);
  return FEM_ERR_BAD_ARG;
}

int fem_amg_bc_mask(int64_t n, int64_t n_bc, const int32_t* bc_dofs, uint8_t* mask, void* stream) {
  FEM_CHECK_ARG(n >= 0 && n_bc >= 0 && mask && (n_bc == 0 || bc_dofs), "fem_amg_bc_mask: bad arguments");
  hipStream_t s = as_stream(stream);
  if (n == 0) return FEM_OK;
  hipLaunchKernelGGL(k_bc_mask, dim3(grid_for(n, 256)), dim3(256), 0, s, n, mask);
  FEM_LAUNCHED();
  if (n_bc > 0) {
    hipLaunchKernelGGL(k_bc_mask_set, dim3(grid_for(n_bc, 256)), dim3(256), 0, s, n_bc, bc_dofs, mask);
    FEM_LAUNCHED();
  }
  return FEM_OK;
}

int fem_amg_sell_to_bcsr(const fem_pattern_t* pat, const double* sell_vals, void* bcsr_vals, int32_t precision,
                         void* stream) {
  FEM_CHECK_ARG(pat && sell_vals && bcsr_vals, "fem_amg_sell_to_bcsr: NULL argument");
  FEM_CHECK_ARG(precision == FEM_F32 || precision == FEM_F64, "fem_amg_sell_to_bcsr: bad precision");
  hipStream_t s = as_stream(stream);
  if (pat->num_rows == 0) return FEM_OK;
  FEM_CHECK_ARG(pat->num_slices == ceil_div(pat->num_rows, kSlice), "fem_amg_sell_to_bcsr: num_slices mismatch");
  FEM_CHECK_ARG(pat->vec == 1 || pat->vec == 3, "fem_amg_sell_to_bcsr: vec must be 1 or 3");
  const dim3 g((unsigned)pat->num_slices);
  auto go = [&](auto t) {
    using TO = decltype(t);
    TO* o = static_cast<TO*>(bcsr_vals);
    if (pat->vec == 3)
      hipLaunchKernelGGL((k_sell_to_bcsr<3, TO>), g, dim3(256), 0, s, *pat, sell_vals, o, (float*)nullptr,
                         (double*)nullptr);
    else
      hipLaunchKernelGGL((k_sell_to_bcsr<1, TO>), g, dim3(256), 0, s, *pat, sell_vals, o, (float*)nullptr,
                         (double*)nullptr);
  };
  if (precision == FEM_F32)
    go(float{});
  else
    go(double{});
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_amg_sell_to_f32_copies(const fem_pattern_t* pat, const double* sell_vals, float* bcsr_vals, float* diff_vals,
                               double* dinv, void* stream) {
  FEM_CHECK_ARG(pat && sell_vals && bcsr_vals && diff_vals, "fem_amg_sell_to_f32_copies: NULL argument");
  hipStream_t s = as_stream(stream);
  if (pat->num_rows == 0) return FEM_OK;
  FEM_CHECK_ARG(pat->num_slices == ceil_div(pat->num_rows, kSlice), "fem_amg_sell_to_f32_copies: num_slices mismatch");
  FEM_CHECK_ARG(pat->vec == 1 || pat->vec == 3, "fem_amg_sell_to_f32_copies: vec must be 1 or 3");
  const dim3 g((unsigned)pat->num_slices);
  if (pat->vec == 3)
    hipLaunchKernelGGL((k_sell_to_bcsr<3, float, true>), g, dim3(256), 0, s, *pat, sell_vals, bcsr_vals, diff_vals,
                       dinv);
  else
    hipLaunchKernelGGL((k_sell_to_bcsr<1, float, true>), g, dim3(256), 0, s, *pat, sell_vals, bcsr_vals, diff_vals,
                       dinv);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_amg_sell_to_f32_diff(const fem_pattern_t* pat, const double* sell_vals, float* out, void* stream) {
  FEM_CHECK_ARG(pat && sell_vals && out, "fem_amg_sell_to_f32_diff: NULL argument");
  hipStream_t s = as_stream(stream);
  if (pat->num_slices == 0) return FEM_OK;
  const unsigned g = (unsigned)ceil_div(pat->num_slices * kSlice, 256);
  if (pat->vec == 3)
    hipLaunchKernelGGL(k_sell_to_diff32<3>, dim3(g), dim3(256), 0, s, *pat, sell_vals, out);
  else if (pat->vec == 1)
    hipLaunchKernelGGL(k_sell_to_diff32<1>, dim3(g), dim3(256), 0, s, *pat, sell_vals, out);
  else
    FEM_CHECK_ARG(false, "fem_amg_sell_to_f32_diff: vec must be 1 or 3");
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_amg_sell_to_f16_scaled(const fem_pattern_t* pat, const double* sell_vals, const double* dinv, void* out,
                               void* stream) {
  FEM_CHECK_ARG(pat && sell_vals && dinv && out, "fem_amg_sell_to_f16_scaled: NULL argument");
  hipStream_t s = as_stream(stream);
  if (pat->num_slices == 0) return FEM_OK;
  const unsigned g = (unsigned)ceil_div(pat->num_slices * kSlice, 256);
  _Float16* o = static_cast<_Float16*>(out);
  if (pat->vec == 3)
    hipLaunchKernelGGL(k_sell_to_f16_scaled<3>, dim3(g), dim3(256), 0, s, *pat, sell_vals, dinv, o);
  else if (pat->vec == 1)
    hipLaunchKernelGGL(k_sell_to_f16_scaled<1>, dim3(g), dim3(256), 0, s, *pat, sell_vals, dinv, o);
  else
    FEM_CHECK_ARG(false, "fem_amg_sell_to_f16_scaled: vec must be 1 or 3");
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_amg_dinv(const fem_bcsr_t* A, int32_t fix_zero, double* dinv, void* stream) {
  FEM_CHECK_ARG(A && dinv && A->br == A->bc && (A->rows == 0 || (A->row_ptr && A->col_idx && A->vals)),
                "fem_amg_dinv: bad arguments");
  hipStream_t s = as_stream(stream);
  if (A->rows == 0) return FEM_OK;
  const unsigned g = grid_for(A->rows, 256);
  FEM_CHECK_ARG(A->br == 1 || A->br == 3 || A->br == 6, "fem_amg_dinv: unsupported block size %d", A->br);
  auto go = [&](auto t) {
    using T = decltype(t);
    T* v = const_cast<T*>(static_cast<const T*>(A->vals));
    if (A->br == 1)
      hipLaunchKernelGGL((k_bcsr_dinv<1, T>), dim3(g), dim3(256), 0, s, A->rows, A->row_ptr, A->col_idx, v,
                         fix_zero, dinv);
    else if (A->br == 3)
      hipLaunchKernelGGL((k_bcsr_dinv<3, T>), dim3(g), dim3(256), 0, s, A->rows, A->row_ptr, A->col_idx, v,
                         fix_zero, dinv);
    else
      hipLaunchKernelGGL((k_bcsr_dinv<6, T>), dim3(g), dim3(256), 0, s, A->rows, A->row_ptr, A->col_idx, v,
                         fix_zero, dinv);
  };
  if (A->f32)
    go(float{});
  else
    go(double{});
  FEM_LAUNCHED();
  return FEM_OK;
}

}  // extern "C"

extern "C" {

int64_t fem_amg_lmax_ws_bytes(int64_t n) {
  n = std::max<int64_t>(n, 1);
  return (int64_t)(2 * align256(n * 8) + align256(kDotBlocks * 8) + 256);  // iterates, partials, 3 scalars
}

int fem_amg_lmax(const fem_bcsr_t* A, const fem_pattern_t* pat, const void* sell_vals, int32_t sell_f32,
                 const double* dinv, int32_t steps, double* lmax, void* ws, void* stream) {
  FEM_CHECK_ARG(dinv && lmax && ws && steps >= 1, "fem_amg_lmax: bad arguments");
  const int64_t n = pat ? (pat->num_rows > 0 ? pat->num_rows * pat->vec : 0) : (A ? A->rows * A->br : 0);
  // the estimate is formed on the device in the workspace's last slot and read back
  double* dev = reinterpret_cast<double*>(static_cast<char*>(ws) + 2 * align256(std::max<int64_t>(n, 1) * 8) +
                                          align256(kDotBlocks * 8)) + 2;
  if (int rc = fem_amg_lmax_async(A, pat, sell_vals, sell_f32, dinv, steps, dev, ws, stream)) return rc;
  hipStream_t s = as_stream(stream);
  FEM_HIP(hipMemcpyAsync(lmax, dev, sizeof(double), hipMemcpyDeviceToHost, s));
  FEM_HIP(hipStreamSynchronize(s));
  return FEM_OK;
}

int fem_amg_lmax_async(const fem_bcsr_t* A, const fem_pattern_t* pat, const void* sell_vals, int32_t sell_f32,
                       const double* dinv, int32_t steps, double* lmax_dev, void* ws, void* stream) {
  FEM_CHECK_ARG(dinv && lmax_dev && ws && steps >= 1, "fem_amg_lmax: bad arguments");
  FEM_CHECK_ARG(pat ? (sell_vals != nullptr && pat->num_rows > 0) : (A && A->br == A->bc && A->rows > 0),
                "fem_amg_lmax: give a SELL matrix (pat, sell_vals) or a square block CSR matrix");
  hipStream_t s = as_stream(stream);
  const int64_t n = pat ? pat->num_rows * pat->vec : A->rows * A->br;
  char* w = static_cast<char*>(ws);
  char* xb = w;
  w += align256(n * 8);
  char* yb = w;
  w += align256(n * 8);
  double* part = reinterpret_cast<double*>(w);
  w += align256(kDotBlocks * 8);
  double* nrm = reinterpret_cast<double*>(w);
  const unsigned g = grid_for(n, 256);
  // the iterates of the FP32 difference-form SELL matrix are FP32 (FP64 arithmetic and
  // norms): the estimate needs a few digits, and each product moves 13 MB less on the
  // 1.6M-dof cantilever
  auto run = [&](auto tag) -> int {
    using TV = decltype(tag);
    TV* x = reinterpret_cast<TV*>(xb);
    TV* y = reinterpret_cast<TV*>(yb);
    auto norm2 = [&](const TV* v, double* out) {
      if constexpr (sizeof(TV) == 4)
        hipLaunchKernelGGL(k_norm2_part_f32, dim3(kDotBlocks), dim3(256), 0, s, n, v, part);
      else
        hipLaunchKernelGGL(k_dot_part, dim3(kDotBlocks), dim3(256), 0, s, n, v, v, part, (const int32_t*)nullptr);
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, s, part, kDotBlocks, out);
    };
    hipLaunchKernelGGL(k_start_vec<TV>, dim3(g), dim3(256), 0, s, n, x);
    FEM_LAUNCHED();
    // y_k = (D^-1 A)^k x_0, rescaled to unit norm every kRescale steps (the iterate
    // grows like lmax^k); lmax = |y_s| / |y_{s-1}|, the norm of the last step's
    // product of a unit vector
    constexpr int kRescale = 8;
    for (int k = 0; k < steps; ++k) {
      const bool last = k == steps - 1;
      if (last) norm2(x, nrm + 1);
      ApplyArgsT<TV, TV> a{x, y, nullptr, dinv, nullptr, 0.0, 0.0, nullptr};
      int rc;
      if constexpr (sizeof(TV) == 4) {
        rc = sell_apply<AP_DSTORE, float, true>(*pat, static_cast<const float*>(sell_vals), a, nullptr, s);
      } else {
        rc = !pat ? (A->f32 ? bcsr_apply<AP_DSTORE, float>(*A, a, s) : bcsr_apply<AP_DSTORE, double>(*A, a, s))
             : sell_f32 == 2 ? sell_apply<AP_DSTORE, _Float16, true>(*pat, sell_vals, a, nullptr, s)
                             : sell_apply<AP_DSTORE, double>(*pat, static_cast<const double*>(sell_vals), a, nullptr, s);
      }
      if (rc) return rc;
      if (last) {
        norm2(y, nrm);
      } else if (k % kRescale == kRescale - 1) {
        norm2(y, nrm);
        hipLaunchKernelGGL(k_scale_by_norm<TV>, dim3(g), dim3(256), 0, s, n, y, nrm, y);
      }
      FEM_LAUNCHED();
      std::swap(x, y);
    }
    return FEM_OK;
  };
  if (int rc = (pat && sell_f32 == 1) ? run(float{}) : run(double{})) return rc;
  // lmax = sqrt(|y_s|^2 / |y_{s-1}|^2)
  hipLaunchKernelGGL(k_lmax_ratio, dim3(1), dim3(64), 0, s, nrm, lmax_dev);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_bcsr_spgemm_symbolic(const fem_bcsr_t* X, const fem_bcsr_t* Y, int32_t* row_counts, int32_t* row_pairs,
                             int64_t* max_pairs, void* ws, void* host_status, void* stream) {
  FEM_CHECK_ARG(X && Y && row_counts && max_pairs && ws && host_status, "fem_bcsr_spgemm_symbolic: NULL argument");
  FEM_CHECK_ARG(X->bc == Y->br && X->cols == Y->rows, "fem_bcsr_spgemm_symbolic: inner dimensions differ");
  hipStream_t s = as_stream(stream);
  *max_pairs = 0;
  if (X->rows == 0) return FEM_OK;
  int32_t* mx = static_cast<int32_t*>(ws);  // [0] max products, [1] max X row, [2] error flags
  FEM_HIP(hipMemsetAsync(mx, 0, 3 * sizeof(int32_t), s));
  hipLaunchKernelGGL(k_spgemm_pairs, dim3(grid_for(X->rows, 256)), dim3(256), 0, s, X->rows, X->row_ptr,
                     X->col_idx, Y->row_ptr, mx, row_pairs);
  FEM_LAUNCHED();
  int32_t h[3];
  if (int rc = read_dev(mx, sizeof(h), host_status, s, h)) return rc;
  const int xcap = h[0] <= 1024 ? kXCap : kBigXCap;
  if (h[1] > xcap) {
    set_error("block SpGEMM: a row of the left operand has %d blocks (capacity %d)", h[1], xcap);
    return FEM_ERR_CAPACITY;
  }
  if (int rc = spgemm_dispatch<false>(h[0], *X, *Y, row_counts, PlanOut{}, mx + 2, s)) return rc;
  if (int rc = read_dev(mx, sizeof(h), host_status, s, h)) return rc;
  if (h[2]) {
    set_error("block SpGEMM: capacity exceeded in the symbolic pass (flags %d)", h[2]);
    return FEM_ERR_CAPACITY;
  }
  *max_pairs = h[0];
  return FEM_OK;
}

int fem_bcsr_spgemm_plan(const fem_bcsr_t* X, const fem_bcsr_t* Y, int64_t max_pairs, fem_bcsr_t* C,
                         const int32_t* pair_ptr, int32_t* plan_ptr, int32_t* plan, void* ws, void* host_status,
                         void* stream) {
  FEM_CHECK_ARG(X && Y && C && pair_ptr && plan_ptr && ws && host_status && C->row_ptr,
                "fem_bcsr_spgemm_plan: NULL argument");
  FEM_CHECK_ARG(X->bc == Y->br && X->cols == Y->rows && C->rows == X->rows && C->cols == Y->cols &&
                    C->br == X->br && C->bc == Y->bc,
                "fem_bcsr_spgemm_plan: shapes do not match");
  hipStream_t s = as_stream(stream);
  if (X->rows == 0) return FEM_OK;
  FEM_CHECK_ARG(C->nnzb == 0 || (C->col_idx && plan), "fem_bcsr_spgemm_plan: output arrays missing");
  int32_t* err = static_cast<int32_t*>(ws);
  FEM_HIP(hipMemsetAsync(err, 0, sizeof(int32_t), s));
  const PlanOut o{C->row_ptr, const_cast<int32_t*>(C->col_idx), pair_ptr, plan_ptr, plan};
  if (int rc = spgemm_dispatch<true>(max_pairs, *X, *Y, nullptr, o, err, s)) return rc;
  int32_t e = 0;
  if (int rc = read_int(err, host_status, s, &e)) return rc;
  if (e) {
    set_error("block SpGEMM: plan pass disagrees with the symbolic pass (flags %d)", e);
    return FEM_ERR_CAPACITY;
  }
  return FEM_OK;
}

int fem_bcsr_spgemm_numeric(const fem_bcsr_t* X, const fem_bcsr_t* Y, fem_bcsr_t* C, const int32_t* plan_ptr,
                            const int32_t* plan, void* stream) {
  FEM_CHECK_ARG(X && Y && C && plan_ptr, "fem_bcsr_spgemm_numeric: NULL argument");
  FEM_CHECK_ARG(X->bc == Y->br && C->br == X->br && C->bc == Y->bc, "fem_bcsr_spgemm_numeric: shapes do not match");
  hipStream_t s = as_stream(stream);
  if (C->nnzb == 0) return FEM_OK;
  FEM_CHECK_ARG(plan && X->vals && Y->vals && C->vals, "fem_bcsr_spgemm_numeric: values missing");
  FEM_CHECK_ARG(X->f32 == Y->f32 && X->f32 == C->f32, "fem_bcsr_spgemm_numeric: operands of different precisions");
  if (X->f32) return spgemm_numeric_t<float>(*X, *Y, *C, plan_ptr, plan, s);
  return spgemm_numeric_t<double>(*X, *Y, *C, plan_ptr, plan, s);
}

int fem_amg_smooth_prolongator(const fem_bcsr_t* A, const fem_bcsr_t* P0, fem_bcsr_t* P, const int32_t* plan_ptr,
                               const int32_t* plan, const int32_t* agg, const double* dinv, double omega,
                               void* stream) {
  FEM_CHECK_ARG(A && P0 && P && plan_ptr && agg && dinv, "fem_amg_smooth_prolongator: NULL argument");
  FEM_CHECK_ARG(A->br == A->bc && P0->br == A->bc && P->rows == A->rows && P->br == P0->br && P->bc == P0->bc &&
                    P0->nnzb == P0->rows,
                "fem_amg_smooth_prolongator: shapes do not match");
  FEM_CHECK_ARG(!P0->f32 && A->f32 == P->f32, "fem_amg_smooth_prolongator: P0 is FP64, A and P of one precision");
  hipStream_t s = as_stream(stream);
  if (P->nnzb == 0) return FEM_OK;
  FEM_CHECK_ARG(plan && A->vals && P0->vals && P->vals && P->col_idx, "fem_amg_smooth_prolongator: arrays missing");
  if (A->f32) return smooth_t<float>(*A, *P0, *P, plan_ptr, plan, agg, dinv, omega, nullptr, s);
  return smooth_t<double>(*A, *P0, *P, plan_ptr, plan, agg, dinv, omega, nullptr, s);
}

int fem_amg_smooth_prolongator_dev(const fem_bcsr_t* A, const fem_bcsr_t* P0, fem_bcsr_t* P,
                                   const int32_t* plan_ptr, const int32_t* plan, const int32_t* agg,
                                   const double* dinv, const double* lmax_dev, void* stream) {
  FEM_CHECK_ARG(A && P0 && P && plan_ptr && agg && dinv && lmax_dev,
                "fem_amg_smooth_prolongator_dev: NULL argument");
  FEM_CHECK_ARG(A->br == A->bc && P0->br == A->bc && P->rows == A->rows && P->br == P0->br && P->bc == P0->bc &&
                    P0->nnzb == P0->rows,
                "fem_amg_smooth_prolongator_dev: shapes do not match");
  FEM_CHECK_ARG(!P0->f32 && A->f32 == P->f32, "fem_amg_smooth_prolongator_dev: P0 is FP64, A and P of one precision");
  hipStream_t s = as_stream(stream);
  if (P->nnzb == 0) return FEM_OK;
  FEM_CHECK_ARG(plan && A->vals && P0->vals && P->vals && P->col_idx, "fem_amg_smooth_prolongator_dev: arrays missing");
  if (A->f32) return smooth_t<float>(*A, *P0, *P, plan_ptr, plan, agg, dinv, 0.0, lmax_dev, s);
  return smooth_t<double>(*A, *P0, *P, plan_ptr, plan, agg, dinv, 0.0, lmax_dev, s);
}

}  // extern "C"

namespace fem {
namespace {
template <typename T>
int spgemm_numeric_t(const fem_bcsr_t& Xd, const fem_bcsr_t& Yd, const fem_bcsr_t& Cd, const int32_t* plan_ptr,
                     const int32_t* plan, hipStream_t s) {
  const fem_bcsr_t *X = &Xd, *Y = &Yd, *C = &Cd;
  const T* xv = static_cast<const T*>(X->vals);
  const T* yv = static_cast<const T*>(Y->vals);
  T* cv = static_cast<T*>(const_cast<void*>(C->vals));
  // one thread per output block; products with few output blocks (the coarse levels) split
  // each block's plan over a whole wave (YC lanes per column group).  On level 0 of the
  // 1.6M-dof cantilever the grouped form with one block per group is slower (A P: 3.50 ->
  // 4.00 ms; it issues about four times the load instructions per product)
  const bool split = C->nnzb < kSplitBelow;
#define FEM_AMG_NUM(XR, XC, YC)                                                                               \
  if (X->br == XR && X->bc == XC && Y->bc == YC) {                                                            \
    constexpr int GPW = 64 / (YC);                                                                            \
    const int64_t waves = split ? C->nnzb : ceil_div(C->nnzb, GPW);                                           \
    const dim3 grid((unsigned)ceil_div(waves, 4));                                                            \
    if (split)                                                                                                \
      hipLaunchKernelGGL((k_spgemm_numeric_g<XR, XC, YC, GPW, T>), grid, dim3(256), 0, s, C->nnzb, plan_ptr,   \
                         plan, xv, yv, cv);                                                                   \
    else                                                                                                      \
      hipLaunchKernelGGL((k_spgemm_numeric<XR, XC, YC, T>), dim3(grid_for(C->nnzb, 256)), dim3(256), 0, s,    \
                         C->nnzb, plan_ptr, plan, xv, yv, cv);                                                \
    FEM_LAUNCHED();                                                                                           \
    return FEM_OK;                                                                                            \
  }
  FEM_AMG_NUM(1, 1, 1)
  FEM_AMG_NUM(3, 3, 6)
  FEM_AMG_NUM(6, 3, 6)
  FEM_AMG_NUM(6, 6, 6)
#undef FEM_AMG_NUM
  set_error("block SpGEMM: unsupported block shapes %d x %d times %d x %d", X->br, X->bc, Y->br, Y->bc);
  return FEM_ERR_BAD_ARG;
}

template <typename T>
int smooth_t(const fem_bcsr_t& Ad, const fem_bcsr_t& P0d, const fem_bcsr_t& Pd, const int32_t* plan_ptr,
             const int32_t* plan, const int32_t* agg, const double* dinv, double omega, const double* lmax_dev,
             hipStream_t s) {
  const fem_bcsr_t *A = &Ad, *P0 = &P0d, *P = &Pd;
  const T* av = static_cast<const T*>(A->vals);
  const double* p0 = static_cast<const double*>(P0->vals);
  T* pv = static_cast<T*>(const_cast<void*>(P->vals));
  const unsigned g = grid_for(P->nnzb, 256);
#define FEM_AMG_SMOOTH(BS, NS)                                                                                  \
  if (P->br == BS && P->bc == NS) {                                                                             \
    hipLaunchKernelGGL((k_smooth_numeric<BS, NS, T>), dim3(g), dim3(256), 0, s, P->nnzb, P->rows, P->row_ptr,      \
                       P->col_idx, plan_ptr, plan, av, p0, agg, dinv, omega, lmax_dev, pv);                     \
    FEM_LAUNCHED();                                                                                             \
    return FEM_OK;                                                                                              \
  }
  FEM_AMG_SMOOTH(3, 6)
  FEM_AMG_SMOOTH(6, 6)
  FEM_AMG_SMOOTH(1, 1)
#undef FEM_AMG_SMOOTH
  set_error("fem_amg_smooth_prolongator: unsupported block shape %d x %d", P->br, P->bc);
  return FEM_ERR_BAD_ARG;
}
}  // namespace
}  // namespace fem

extern "C" {

int fem_bcsr_transpose_structure(const fem_bcsr_t* X, fem_bcsr_t* T, int32_t* src, void* ws, void* stream) {
  FEM_CHECK_ARG(X && T && ws && T->row_ptr, "fem_bcsr_transpose_structure: NULL argument");
  FEM_CHECK_ARG(T->rows == X->cols && T->cols == X->rows && T->nnzb == X->nnzb && T->br == X->bc && T->bc == X->br,
                "fem_bcsr_transpose_structure: output descriptor does not match X^T");
  hipStream_t s = as_stream(stream);
  char* w = static_cast<char*>(ws);
  int32_t* cnt = reinterpret_cast<int32_t*>(w);
  w += align256(std::max<int64_t>(X->cols, 1) * 4);
  void* scan_ws = w;
  int32_t* trp = const_cast<int32_t*>(T->row_ptr);
  int32_t* tci = const_cast<int32_t*>(T->col_idx);
  FEM_HIP(hipMemsetAsync(cnt, 0, std::max<int64_t>(X->cols, 1) * 4, s));
  if (X->nnzb > 0) {
    hipLaunchKernelGGL(k_count_cols, dim3(grid_for(X->nnzb, 256)), dim3(256), 0, s, X->nnzb, X->col_idx, cnt);
    FEM_LAUNCHED();
  }
  if (int rc = fem_exclusive_scan(X->cols, cnt, trp, scan_ws, stream)) return rc;
  if (X->nnzb == 0) return FEM_OK;
  FEM_CHECK_ARG(tci && src, "fem_bcsr_transpose_structure: output arrays missing");
  FEM_HIP(hipMemsetAsync(cnt, 0, std::max<int64_t>(X->cols, 1) * 4, s));
  hipLaunchKernelGGL(k_t_fill, dim3(grid_for(X->rows, 256)), dim3(256), 0, s, X->rows, X->row_ptr, X->col_idx, trp,
                     cnt, tci, src);
  FEM_LAUNCHED();
  if (T->rows > 0) {
    hipLaunchKernelGGL(k_sort_rows, dim3((unsigned)T->rows), dim3(256), 0, s, T->rows, trp, tci, src);
    FEM_LAUNCHED();
  }
  return FEM_OK;
}

int fem_bcsr_transpose_values(const fem_bcsr_t* X, fem_bcsr_t* T, const int32_t* src, void* stream) {
  FEM_CHECK_ARG(X && T && (X->nnzb == 0 || (src && X->vals && T->vals)), "fem_bcsr_transpose_values: NULL argument");
  FEM_CHECK_ARG(T->nnzb == X->nnzb && T->br == X->bc && T->bc == X->br, "fem_bcsr_transpose_values: shapes differ");
  hipStream_t s = as_stream(stream);
  if (X->nnzb == 0) return FEM_OK;
  FEM_CHECK_ARG(X->f32 == T->f32, "fem_bcsr_transpose_values: X and X^T of different precisions");
  const unsigned g = grid_for(X->nnzb, 256);
#define FEM_AMG_T(BR, BC)                                                                                   \
  if (X->br == BR && X->bc == BC) {                                                                         \
    if (X->f32)                                                                                             \
      hipLaunchKernelGGL((k_t_values<BR, BC, float>), dim3(g), dim3(256), 0, s, X->nnzb, src,                \
                         static_cast<const float*>(X->vals), static_cast<float*>(const_cast<void*>(T->vals)));    \
    else                                                                                                    \
      hipLaunchKernelGGL((k_t_values<BR, BC, double>), dim3(g), dim3(256), 0, s, X->nnzb, src,               \
                         static_cast<const double*>(X->vals), static_cast<double*>(const_cast<void*>(T->vals))); \
    FEM_LAUNCHED();                                                                                         \
    return FEM_OK;                                                                                          \
  }
  FEM_AMG_T(3, 6)
  FEM_AMG_T(6, 6)
  FEM_AMG_T(1, 1)
  FEM_AMG_T(6, 3)
  FEM_AMG_T(3, 3)
#undef FEM_AMG_T
  set_error("fem_bcsr_transpose_values: unsupported block shape %d x %d", X->br, X->bc);
  return FEM_ERR_BAD_ARG;
}

int fem_amg_coarse_inverse(const fem_bcsr_t* A, double* inv, void* ws, void* stream) {
  FEM_CHECK_ARG(A && inv && ws && A->rows > 0, "fem_amg_coarse_inverse: bad arguments");
  const int64_t n = A->rows * A->br;
  FEM_CHECK_ARG(n <= 4096 && A->cols * A->bc == n, "fem_amg_coarse_inverse: %lld unknowns (at most 4096, square)",
                (long long)n);
  hipStream_t s = as_stream(stream);
  double* a = inv;
  double* b = static_cast<double*>(ws);
  const unsigned g = grid_for(n * n, 256);
  hipLaunchKernelGGL(k_dense_zero, dim3(g), dim3(256), 0, s, n * n, a);
  if (A->f32)
    hipLaunchKernelGGL(k_dense_scatter<float>, dim3((unsigned)A->rows), dim3(256), 0, s, A->rows, A->br, A->bc,
                       n, A->row_ptr, A->col_idx, static_cast<const float*>(A->vals), a);
  else
    hipLaunchKernelGGL(k_dense_scatter<double>, dim3((unsigned)A->rows), dim3(256), 0, s, A->rows, A->br,
                       A->bc, n, A->row_ptr, A->col_idx, static_cast<const double*>(A->vals), a);
  FEM_LAUNCHED();
  if (n <= kGjSmall) {
    hipLaunchKernelGGL(k_gj_small, dim3(1), dim3(64 * kGjWaves), 0, s, (int)n, n, (int64_t)0, a);
    FEM_LAUNCHED();
    return FEM_OK;
  }
  if (n >= 2 * kGjSmall) {
    // blocked: the panels (n * kGjSmall + kGjSmall * n doubles) fit in ws (n * n doubles)
    double* colp = b;
    double* rowp = b + n * kGjSmall;
    const unsigned tiles = (unsigned)ceil_div(n, kGjSmall);
    for (int64_t k0 = 0; k0 < n; k0 += kGjSmall) {
      const int bk = (int)std::min<int64_t>(kGjSmall, n - k0);
      hipLaunchKernelGGL(k_gj_small, dim3(1), dim3(64 * kGjWaves), 0, s, bk, n, k0, a);
      hipLaunchKernelGGL(k_bgj_panel, dim3(tiles, kGjSmall / 16), dim3(256), 0, s, n, k0, bk, a, colp, rowp);
      hipLaunchKernelGGL(k_bgj_update, dim3(tiles, tiles), dim3(256), 0, s, n, k0, bk, a, colp, rowp);
    }
    FEM_LAUNCHED();
    return FEM_OK;
  }
  for (int64_t k = 0; k < n; ++k) {
    hipLaunchKernelGGL(k_gj_step, dim3(g), dim3(256), 0, s, n, k, a, b);
    std::swap(a, b);
  }
  FEM_LAUNCHED();
  if (a != inv) {
    hipLaunchKernelGGL(k_copy, dim3(g), dim3(256), 0, s, n * n, a, inv);
    FEM_LAUNCHED();
  }
  return FEM_OK;
}

int fem_amg_to_f32(int64_t n, const double* in, float* out, void* stream) {
  FEM_CHECK_ARG(n >= 0 && (n == 0 || (in && out)), "fem_amg_to_f32: bad arguments");
  if (n == 0) return FEM_OK;
  hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), n, in, out);
  FEM_LAUNCHED();
  return FEM_OK;
}

int64_t fem_amg_ws_bytes(const fem_amg_t* amg) {
  if (!amg || amg->num_levels < 0 || (amg->num_levels > 0 && (!amg->levels || !amg->levels[0].pat))) return 0;
  return vcycle_ws(*amg);
}

int fem_amg_vcycle(const fem_amg_t* amg, const double* r, double* z, void* ws, void* stream) {
  if (int rc = check_amg(amg)) return rc;
  FEM_CHECK_ARG(r && z && ws && r != z, "fem_amg_vcycle: bad arguments");
  hipStream_t s = as_stream(stream);
  char* w = static_cast<char*>(ws);
  return amg->precision == FEM_F32 ? vcycle_impl<float>(*amg, r, z, w, s) : vcycle_impl<double>(*amg, r, z, w, s);
}

int64_t fem_amg_solve_ws_bytes(const fem_pattern_t* pat, const fem_amg_t* amg) {
  if (!pat) return 0;
  const int64_t n = std::max<int64_t>(pat->num_rows * pat->vec, 1);
  const int64_t np = std::max<int64_t>(pat->num_slices, kDotBlocks);
  return fem_amg_ws_bytes(amg) + (int64_t)(4 * align256(n * 8) + align256(np * 8) + align256(sizeof(AmgScalars)));
}

int fem_amg_pcg_solve(const fem_pattern_t* pat, const double* vals, const float* vals32, const fem_amg_t* amg,
                      const double* b, double* x, const fem_amg_options_t* opt, void* ws, void* host_status,
                      fem_pcg_result_t* result, void* stream) {
  if (int rc = check_amg(amg)) return rc;
  FEM_CHECK_ARG(pat && vals && b && x && opt && ws && host_status && result, "fem_amg_pcg_solve: NULL argument");
  const bool outer32 = vals32 != nullptr;
  FEM_CHECK_ARG(!outer32 || (opt->reliable_delta > 0.0 && opt->reliable_delta < 1.0),
                "fem_amg_pcg_solve: reliable_delta must be in (0, 1) with an FP32 outer operator");
  FEM_CHECK_ARG(amg->num_levels == 0 ? amg->coarse_n == pat->num_rows * pat->vec
                                     : amg->levels[0].pat->num_rows == pat->num_rows && amg->levels[0].pat->vec == pat->vec,
                "fem_amg_pcg_solve: the hierarchy was built for another pattern");
  FEM_CHECK_ARG(opt->maxiter >= 0 && opt->check_every >= 1, "fem_amg_pcg_solve: bad options");
  hipStream_t s = as_stream(stream);
  const int64_t n = pat->num_rows * pat->vec;
  char* w = static_cast<char*>(ws);
  char* cyc = w;
  w += vcycle_ws(*amg);
  double* r = reinterpret_cast<double*>(w);
  w += align256(std::max<int64_t>(n, 1) * 8);
  double* z = reinterpret_cast<double*>(w);
  w += align256(std::max<int64_t>(n, 1) * 8);
  double* p = reinterpret_cast<double*>(w);
  w += align256(std::max<int64_t>(n, 1) * 8);
  double* q = reinterpret_cast<double*>(w);
  w += align256(std::max<int64_t>(n, 1) * 8);
  double* part = reinterpret_cast<double*>(w);
  w += align256(std::max<int64_t>(pat->num_slices, kDotBlocks) * 8);
  AmgScalars* sc = reinterpret_cast<AmgScalars*>(w);
  const int32_t* done = &sc->done;
  const unsigned g = grid_for(n, 256);
  *result = fem_pcg_result_t{};
  // r = b - A x with the FP64 matrix, sc->rr = |r|^2
  auto true_residual = [&]() -> int {
    ApplyArgs ar{x, r, b, nullptr, nullptr, 0.0, 0.0, nullptr};
    if (int rc = sell_apply<AP_RESID, double>(*pat, vals, ar, nullptr, s)) return rc;
    hipLaunchKernelGGL(k_dot_part, dim3(kDotBlocks), dim3(256), 0, s, n, r, r, part, (const int32_t*)nullptr);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, s, part, kDotBlocks, &sc->rr);
    FEM_LAUNCHED();
    return FEM_OK;
  };
  hipLaunchKernelGGL(k_dot_part, dim3(kDotBlocks), dim3(256), 0, s, n, b, b, part, (const int32_t*)nullptr);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, s, part, kDotBlocks, &sc->bb);
  FEM_LAUNCHED();
  if (int rc = true_residual()) return rc;
  hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(1), 0, s, sc, opt->rtol, opt->atol, opt->maxiter);
  FEM_LAUNCHED();
  AmgScalars h{};
  if (int rc = read_dev(sc, sizeof(AmgScalars), host_status, s, &h)) return rc;
  constexpr int kEv = 8;
  hipEvent_t ev[kEv] = {};
  // the three level-0 Chebyshev SELL launches of a cycle are timed where level 0 has them
  const bool three = amg->num_levels > 0 && amg->degree >= 2;
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      for (int i = 0; i < kEv; ++i)
        if (e[i]) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  if (opt->timing)
    for (int i = 0; i < kEv; ++i) FEM_HIP(hipEventCreate(&ev[i]));
  double cheb_ms = 0.0, cycle_ms = 0.0;
  int samples = 0;
  int it = 0;
  // FP32 outer operator: the recursive residual is replaced by the FP64 true
  // residual when it has fallen by reliable_delta since the last replacement,
  // and when it meets the tolerance (convergence is decided on the true one)
  double rr_repl = h.rr;
  const double delta2 = opt->reliable_delta * opt->reliable_delta;
  int replacements = 0;
  while (!h.done) {
    hipEvent_t* e = opt->timing ? ev : nullptr;
    if (int rc = (amg->precision == FEM_F32 ? vcycle_impl<float>(*amg, r, z, cyc, s, e, done)
                                            : vcycle_impl<double>(*amg, r, z, cyc, s, e, done)))
      return rc;
    hipLaunchKernelGGL(k_dot_part, dim3(kDotBlocks), dim3(256), 0, s, n, r, z, part, done);
    hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(256), 0, s, part, kDotBlocks, sc);
    hipLaunchKernelGGL(k_cg_p, dim3(g), dim3(256), 0, s, n, z, p, sc);
    FEM_LAUNCHED();
    ApplyArgs apq{p, q, nullptr, nullptr, nullptr, 0.0, 0.0, part};
    if (int rc = outer32 ? sell_apply<AP_PQ, float, true>(*pat, vals32, apq, done, s)
                         : sell_apply<AP_PQ, double>(*pat, vals, apq, done, s))
      return rc;
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(256), 0, s, part, (int)pat->num_slices, sc);
    hipLaunchKernelGGL(k_cg_xr, dim3(kDotBlocks), dim3(256), 0, s, n, p, q, x, r, part, sc);
    hipLaunchKernelGGL(k_cg_rr, dim3(1), dim3(256), 0, s, part, kDotBlocks, sc);
    FEM_LAUNCHED();
    ++it;
    if (it % opt->check_every == 0 || it >= opt->maxiter) {
      if (int rc = read_dev(sc, sizeof(AmgScalars), host_status, s, &h)) return rc;
      // (a cycle queued after convergence exits at once: h.iter < it, no sample)
      if (opt->timing && h.iter >= it) {
        float a = 0.f, b = 0.f;
        FEM_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
        FEM_HIP(hipEventElapsedTime(&b, ev[2], ev[3]));
        if (three) {
          float c = 0.f, d = 0.f;
          FEM_HIP(hipEventElapsedTime(&c, ev[4], ev[5]));
          FEM_HIP(hipEventElapsedTime(&d, ev[6], ev[7]));
          a = (a + c + d) / 3.f;
        }
        cheb_ms += a;
        cycle_ms += b;
        ++samples;
      }
      if (outer32 && !h.nan && (h.done ? h.iter < opt->maxiter : h.rr < delta2 * rr_repl)) {
        if (int rc = true_residual()) return rc;
        hipLaunchKernelGGL(k_cg_replaced, dim3(1), dim3(1), 0, s, sc);
        FEM_LAUNCHED();
        if (int rc = read_dev(sc, sizeof(AmgScalars), host_status, s, &h)) return rc;
        rr_repl = h.rr;
        ++replacements;
      }
      if (it >= opt->maxiter && !h.done) break;
    }
  }
  if (samples) {
    result->spmv_ms = cheb_ms / samples;
    result->update_ms = cycle_ms / samples;
    result->timed_launches = samples;
  }
  // true residual
  ApplyArgs af{x, q, b, nullptr, nullptr, 0.0, 0.0, nullptr};
  if (int rc = sell_apply<AP_RESID, double>(*pat, vals, af, nullptr, s)) return rc;
  hipLaunchKernelGGL(k_dot_part, dim3(kDotBlocks), dim3(256), 0, s, n, q, q, part, (const int32_t*)nullptr);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, s, part, kDotBlocks, &sc->rz);
  FEM_LAUNCHED();
  AmgScalars fin{};
  if (int rc = read_dev(sc, sizeof(AmgScalars), host_status, s, &fin)) return rc;
  result->iterations = fin.iter;
  result->replacements = replacements;
  result->b_norm = sqrt(fin.bb);
  result->res_norm = sqrt(fin.rz);
  result->status = fin.nan ? FEM_ERR_NAN : (fin.rr > fin.stop2 ? FEM_ERR_NOT_CONVERGED : FEM_OK);
  return FEM_OK;
}

}  // extern "C"
