// Sparsity pattern of the global tangent, built on the device once per mesh.
//
// Replaces the host-side COO pattern of jax_fem/problem.py:86-107 (I, J from
// onp.repeat over the cell dof indices, C*(N*vec)^2 int64 entries each) and the
// PETSc preallocation of solver.py:472-478.  Instead of materialising COO, the
// node graph is derived from node -> cell incidence: the columns of node row i
// are the nodes of the cells incident to i.  Per row the (column, cell-block)
// candidates are sorted; the sorted cell-block ids double as the "sum the
// duplicates" gather list that turns per-cell tangents into matrix values in
// the same order PETSc's setValuesCOO visits them (cells in order).
#include "fem_common.hpp"

#include <cstdarg>
#include <mutex>

namespace fem {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

// ------------------------------------------------------------ incidence
__global__ void k_count(int64_t n_items, const int32_t* __restrict__ conn, int32_t* __restrict__ cnt) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n_items;
       t += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[conn[t]], 1);
}

__global__ void k_fill_incidence(int64_t n_items, const int32_t* __restrict__ conn,
                                 const int32_t* __restrict__ ptr, int32_t* __restrict__ cursor,
                                 int32_t* __restrict__ idx) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n_items;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t node = conn[t];
    const int32_t pos = atomicAdd(&cursor[node], 1);
    idx[ptr[node] + pos] = (int32_t)t;
  }
}

// Insertion sort of each node's (short) incidence list -> ascending c*N + a.
__global__ void k_sort_lists(int64_t num_nodes, const int32_t* __restrict__ ptr, int32_t* __restrict__ idx) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < num_nodes;
       n += (int64_t)gridDim.x * blockDim.x) {
    const int b = ptr[n], e = ptr[n + 1];
    for (int i = b + 1; i < e; ++i) {
      const int32_t v = idx[i];
      int j = i - 1;
      while (j >= b && idx[j] > v) {
        idx[j + 1] = idx[j];
        --j;
      }
      idx[j + 1] = v;
    }
  }
}

// ------------------------------------------------------------------ scan
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline int32_t block_exclusive_scan_256(int32_t v, int32_t* smem /*256+*/, int32_t& total) {
  // Hillis-Steele on 256 values in shared memory.
  smem[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    int32_t a = threadIdx.x >= off ? smem[threadIdx.x - off] : 0;
    __syncthreads();
    smem[threadIdx.x] += a;
    __syncthreads();
  }
  total = smem[kScanThreads - 1];
  const int32_t incl = smem[threadIdx.x];
  __syncthreads();
  return incl - v;
}

__global__ void k_scan_tiles(int64_t n, const int32_t* __restrict__ in, int32_t* __restrict__ out,
                             int32_t* __restrict__ tile_sums) {
  __shared__ int32_t smem[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int32_t v[kScanItems];
  int32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0;
    s += v[i];
  }
  int32_t total;
  int32_t run = block_exclusive_scan_256(s, smem, total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Single block: exclusive scan of tile sums in place; writes the grand total.
__global__ void k_scan_sums(int64_t ntiles, int32_t* __restrict__ sums, int32_t* __restrict__ grand) {
  __shared__ int32_t smem[kScanThreads];
  int32_t carry = 0;
  for (int64_t b = 0; b < ntiles; b += kScanThreads) {
    const int64_t i = b + threadIdx.x;
    int32_t v = i < ntiles ? sums[i] : 0;
    int32_t total;
    int32_t ex = block_exclusive_scan_256(v, smem, total);
    if (i < ntiles) sums[i] = ex + carry;
    carry += total;
  }
  if (threadIdx.x == 0) *grand = carry;
}

__global__ void k_scan_add(int64_t n, int32_t* __restrict__ out, const int32_t* __restrict__ sums) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  const int32_t add = sums[blockIdx.x];
  for (int i = threadIdx.x; i < kScanTile; i += kScanThreads)
    if (base + i < n) out[base + i] += add;
}

// ------------------------------------------------------------ row patterns
constexpr int kMaxCand = 2048;  // (column, cell-block) candidates per row

__global__ __launch_bounds__(64) void k_pattern_rows(int64_t num_nodes, int32_t N,
                                                     const int32_t* __restrict__ cells,
                                                     const int32_t* __restrict__ inc_ptr,
                                                     const int32_t* __restrict__ inc_idx,
                                                     int32_t* __restrict__ contrib_idx,
                                                     int32_t* __restrict__ row_len,
                                                     int32_t* __restrict__ err) {
  __shared__ unsigned long long keys[kMaxCand];
  for (int64_t row = blockIdx.x; row < num_nodes; row += gridDim.x) {
    const int beg = inc_ptr[row];
    const int deg = inc_ptr[row + 1] - beg;
    const int M = deg * N;
    if (M > kMaxCand) {
      if (threadIdx.x == 0) {
        atomicMax(err, M);
        row_len[row] = 0;
      }
      continue;
    }
    int P = 1;
    while (P < M) P <<= 1;
    for (int t = threadIdx.x; t < P; t += 64) {
      unsigned long long key = ~0ull;
      if (t < M) {
        const int32_t inc = inc_idx[beg + t / N];  // c*N + a
        const int b = t % N;
        const int32_t c = inc / N;
        const uint32_t col = (uint32_t)cells[(int64_t)c * N + b];
        const uint32_t cid = (uint32_t)inc * (uint32_t)N + (uint32_t)b;  // c*N*N + a*N + b
        key = ((unsigned long long)col << 32) | cid;
      }
      keys[t] = key;
    }
    __syncthreads();
    // bitonic sort ascending
    for (int k = 2; k <= P; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int t = threadIdx.x; t < P; t += 64) {
          const int l = t ^ j;
          if (l > t) {
            const unsigned long long a = keys[t], b = keys[l];
            const bool up = (t & k) == 0;
            if ((a > b) == up) {
              keys[t] = b;
              keys[l] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    int cnt = 0;
    for (int t = threadIdx.x; t < M; t += 64) {
      const unsigned long long kk = keys[t];
      contrib_idx[(int64_t)beg * N + t] = (int32_t)(kk & 0xffffffffull);
      cnt += (t == 0 || (keys[t - 1] >> 32) != (kk >> 32)) ? 1 : 0;
    }
    cnt = wave_sum(cnt);
    if (threadIdx.x == 0) row_len[row] = cnt;
    __syncthreads();
  }
}

__global__ void k_pattern_fill(int64_t num_nodes, int32_t N, const int32_t* __restrict__ cells,
                               const int32_t* __restrict__ inc_ptr, const int32_t* __restrict__ contrib_idx,
                               const int32_t* __restrict__ row_ptr, int32_t* __restrict__ col_idx,
                               int32_t* __restrict__ contrib_ptr, int32_t* __restrict__ diag_slot) {
  const int64_t NN = (int64_t)N * N;
  for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < num_nodes;
       row += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b0 = (int64_t)inc_ptr[row] * N;
    const int64_t b1 = (int64_t)inc_ptr[row + 1] * N;
    int slot = row_ptr[row] - 1;
    int32_t prev = -1;
    int dslot = -1;
    for (int64_t t = b0; t < b1; ++t) {
      const int32_t cid = contrib_idx[t];
      const int64_t c = cid / NN;
      const int b = (int)(cid % N);
      const int32_t col = cells[c * N + b];
      if (col != prev) {
        ++slot;
        col_idx[slot] = col;
        contrib_ptr[slot] = (int32_t)t;
        if (col == (int32_t)row) dslot = slot - row_ptr[row];
        prev = col;
      }
    }
    diag_slot[row] = dslot;
    if (row == num_nodes - 1) contrib_ptr[row_ptr[num_nodes]] = (int32_t)b1;
  }
}

// One block per window of kSortWindow rows: stable rank of each row by
// decreasing length (padding rows, length -1, go last).  Each thread ranks
// kSortWindow / kSortThreads rows against the whole window held in LDS.
constexpr int kSortWindow = 2048;
constexpr int kSortThreads = 1024;
__global__ __launch_bounds__(kSortThreads) void k_sell_order(int64_t num_rows, int64_t num_lanes, int sort,
                                                             const int32_t* __restrict__ row_ptr,
                                                             int32_t* __restrict__ sell_row,
                                                             int32_t* __restrict__ sell_pos) {
  __shared__ int len[kSortWindow];
  const int64_t base = blockIdx.x * (int64_t)kSortWindow;
  for (int i = threadIdx.x; i < kSortWindow; i += kSortThreads) {
    const int64_t r = base + i;
    len[i] = r < num_rows ? row_ptr[r + 1] - row_ptr[r] : -1;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kSortWindow; i += kSortThreads) {
    const int my = len[i];
    int rank = i;
    if (sort) {
      rank = 0;
      for (int j = 0; j < kSortWindow; ++j) {
        const int l = len[j];
        rank += (l > my) || (l == my && j < i);
      }
    }
    const int64_t r = base + i;
    const int64_t pos = base + rank;
    // real rows rank first in their window, so their positions are < num_rows <= num_lanes
    if (pos < num_lanes) sell_row[pos] = my >= 0 ? (int32_t)r : -1;
    if (my >= 0) sell_pos[r] = (int32_t)pos;
  }
}

__global__ void k_slice_sizes(const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ sell_row,
                              int32_t* __restrict__ slice_blocks, int64_t num_slices) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < num_slices;
       s += (int64_t)gridDim.x * blockDim.x) {
    int L = 0;
    for (int l = 0; l < kSlice; ++l) {
      const int r = sell_row[s * kSlice + l];
      if (r >= 0) L = max(L, row_ptr[r + 1] - row_ptr[r]);
    }
    slice_blocks[s] = L * kSlice;
  }
}

// Padding slots hold col = row (col 0 on padding lanes) and zero values.
__global__ void k_sell_fill(int64_t num_slices, const int32_t* __restrict__ row_ptr,
                            const int32_t* __restrict__ col_idx, const int32_t* __restrict__ slice_ptr,
                            const int32_t* __restrict__ sell_row, int32_t* __restrict__ sell_col) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < num_slices * kSlice;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t / kSlice;
    const int lane = (int)(t % kSlice);
    const int base = slice_ptr[s];
    const int L = (slice_ptr[s + 1] - base) / kSlice;
    const int row = sell_row[t];
    const bool real = row >= 0;
    const int rb = real ? row_ptr[row] : 0;
    const int len = real ? row_ptr[row + 1] - rb : 0;
    for (int k = 0; k < L; ++k)
      sell_col[base + (int64_t)k * kSlice + lane] = k < len ? col_idx[rb + k] : (real ? row : 0);
  }
}

static int grid_for(int64_t n, int threads, int cap = 65536) {
  int64_t g = ceil_div(n, threads);
  if (g < 1) g = 1;
  return (int)(g > cap ? cap : g);
}

}  // namespace fem

using namespace fem;

extern "C" {

const char* fem_last_error(void) { return g_err.c_str(); }
int fem_version(void) { return 1; }

int64_t fem_scan_ws_bytes(int64_t n) { return (ceil_div(n, kScanTile) + 2) * 4 + 256; }

int fem_exclusive_scan(int64_t n, const int32_t* in, int32_t* out, void* ws, void* stream) {
  FEM_CHECK_ARG(n >= 0 && out && ws && (n == 0 || in), "fem_exclusive_scan: bad arguments");
  hipStream_t s = as_stream(stream);
  const int64_t ntiles = ceil_div(n, kScanTile);
  int32_t* sums = reinterpret_cast<int32_t*>(ws);
  if (ntiles > 0) {
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, n, in, out, sums);
    FEM_LAUNCHED();
  }
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, ntiles, sums, out + n);
  FEM_LAUNCHED();
  if (ntiles > 0) {
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, n, out, sums);
    FEM_LAUNCHED();
  }
  return FEM_OK;
}

int fem_incidence(int64_t num_entities, int32_t N, const int32_t* conn, int64_t num_nodes,
                  int32_t* inc_ptr, int32_t* inc_idx, void* ws, void* stream) {
  FEM_CHECK_ARG(num_entities >= 0 && N > 0 && num_nodes > 0 && (conn || num_entities == 0) && inc_ptr && inc_idx &&
                    ws,
                "fem_incidence: bad arguments");
  FEM_CHECK_ARG(num_entities * N < (int64_t)INT32_MAX, "fem_incidence: too many entries for int32 indices");
  hipStream_t s = as_stream(stream);
  int32_t* cnt = reinterpret_cast<int32_t*>(ws);
  void* scan_ws = reinterpret_cast<char*>(ws) + ((num_nodes * 4 + 255) / 256) * 256;
  const int64_t items = num_entities * N;
  FEM_HIP(hipMemsetAsync(cnt, 0, num_nodes * sizeof(int32_t), s));
  if (items > 0) {
    hipLaunchKernelGGL(k_count, dim3(grid_for(items, 256)), dim3(256), 0, s, items, conn, cnt);
    FEM_LAUNCHED();
  }
  int rc = fem_exclusive_scan(num_nodes, cnt, inc_ptr, scan_ws, stream);
  if (rc) return rc;
  FEM_HIP(hipMemsetAsync(cnt, 0, num_nodes * sizeof(int32_t), s));
  if (items > 0) {
    hipLaunchKernelGGL(k_fill_incidence, dim3(grid_for(items, 256)), dim3(256), 0, s, items, conn, inc_ptr,
                       cnt, inc_idx);
    FEM_LAUNCHED();
  }
  hipLaunchKernelGGL(k_sort_lists, dim3(grid_for(num_nodes, 256)), dim3(256), 0, s, num_nodes, inc_ptr,
                     inc_idx);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_pattern_rows(const fem_mesh_t* mesh, const int32_t* inc_ptr, const int32_t* inc_idx,
                     int32_t* contrib_idx, int32_t* row_len, int32_t* err_flag, void* stream) {
  FEM_CHECK_ARG(mesh && inc_ptr && inc_idx && contrib_idx && row_len && err_flag,
                "fem_pattern_rows: bad arguments");
  const int N = mesh->nodes_per_cell;
  FEM_CHECK_ARG((int64_t)mesh->num_cells * N * N < (int64_t)INT32_MAX,
                "fem_pattern_rows: cell-block ids exceed int32");
  hipStream_t s = as_stream(stream);
  FEM_HIP(hipMemsetAsync(err_flag, 0, sizeof(int32_t), s));
  hipLaunchKernelGGL(k_pattern_rows, dim3(grid_for(mesh->num_nodes, 1, 1 << 20)), dim3(64), 0, s,
                     mesh->num_nodes, (int32_t)N, mesh->cells, inc_ptr, inc_idx, contrib_idx, row_len, err_flag);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_pattern_fill(const fem_mesh_t* mesh, const int32_t* inc_ptr, const int32_t* contrib_idx,
                     const int32_t* row_ptr, int32_t* col_idx, int32_t* contrib_ptr, int32_t* diag_slot,
                     void* stream) {
  FEM_CHECK_ARG(mesh && inc_ptr && contrib_idx && row_ptr && col_idx && contrib_ptr && diag_slot,
                "fem_pattern_fill: bad arguments");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_pattern_fill, dim3(grid_for(mesh->num_nodes, 256)), dim3(256), 0, s, mesh->num_nodes,
                     mesh->nodes_per_cell, mesh->cells, inc_ptr, contrib_idx, row_ptr, col_idx, contrib_ptr,
                     diag_slot);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_sell_order(int64_t num_rows, int32_t sort, const int32_t* row_ptr, int32_t* sell_row, int32_t* sell_pos,
                   void* stream) {
  FEM_CHECK_ARG(num_rows > 0 && row_ptr && sell_row && sell_pos, "fem_sell_order: bad arguments");
  const int64_t lanes = ceil_div(num_rows, kSlice) * kSlice;
  hipLaunchKernelGGL(k_sell_order, dim3((unsigned)ceil_div(num_rows, kSortWindow)), dim3(kSortThreads), 0,
                     as_stream(stream), num_rows, lanes, sort ? 1 : 0, row_ptr, sell_row, sell_pos);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_sell_slice_sizes(int64_t num_rows, const int32_t* row_ptr, const int32_t* sell_row, int32_t* slice_blocks,
                         void* stream) {
  FEM_CHECK_ARG(num_rows > 0 && row_ptr && sell_row && slice_blocks, "fem_sell_slice_sizes: bad arguments");
  const int64_t ns = ceil_div(num_rows, kSlice);
  hipLaunchKernelGGL(k_slice_sizes, dim3(grid_for(ns, 256)), dim3(256), 0, as_stream(stream), row_ptr, sell_row,
                     slice_blocks, ns);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_sell_fill(const fem_pattern_t* pat, int32_t* sell_col, void* stream) {
  FEM_CHECK_ARG(pat && pat->row_ptr && pat->col_idx && pat->slice_ptr && pat->sell_row && sell_col,
                "fem_sell_fill: bad arguments");
  FEM_CHECK_ARG(pat->num_slices == ceil_div(pat->num_rows, kSlice), "fem_sell_fill: num_slices mismatch");
  hipLaunchKernelGGL(k_sell_fill, dim3(grid_for(pat->num_slices * kSlice, 256)), dim3(256), 0,
                     as_stream(stream), pat->num_slices, pat->row_ptr, pat->col_idx, pat->slice_ptr, pat->sell_row,
                     sell_col);
  FEM_LAUNCHED();
  return FEM_OK;
}

}  // extern "C"
