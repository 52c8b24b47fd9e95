// Single-pass HEX8 assembly: per-cell tangents (and optionally the residual)
// computed and summed straight into the SELL-64 matrix, with no per-cell
// tangent buffer in between.
//
// Replaces, in one kernel, what jax-fem does in two host-visible stages:
//   problem.py:447-460  compute_newton_vars: vmapped jacfwd per cell -> problem.V
//   solver.py:522-526   PETSc setValuesCOO: sum the COO duplicates into AIJ
//   problem.py:426-437  scatter of the cell residuals (optional, same pass)
//
// Work split.  A workgroup owns one SELL slice (64 block rows, lane l = row
// sell_row[64 s + l]), so every matrix block it writes is its own and no
// atomics are needed.  The plan lists the slice's (cell, local node a, row
// lane) pairs in a fixed order (cells in chunk order, see k_slice_reorder) and
// re-expresses each slot's contributions as (pair, column node b) entries.
// The pairs are walked in chunks of at most 16 cells / 32 pairs; per chunk the
// cells' geometry is staged in LDS, one thread per (pair, b) forms the 3x3
// block K_ab (and the pair's residual contribution), and the thread owning a
// slot (wave w owns the slots w, w + 8, ... of its lane's row) adds its
// blocks in pair order.  The sums do not depend on scheduling
// (bit-reproducible).  After the last chunk every lane stores its slots: one
// coalesced 64-lane store per matrix value.  The difference form's row sum and
// the Jacobi inverse diagonal are formed from the accumulators in a fixed order.
//
// A cell's Jacobian is recomputed once per slice that touches it (about 4.5
// times for a structured HEX8 box with 64-row slices); each tangent block is
// computed exactly once.
#include "element.hpp"

#include <algorithm>
#include <type_traits>

namespace fem {

constexpr int kAsmThreads = 512;   // 8 waves per slice
constexpr int kAsmWaves = kAsmThreads / 64;
constexpr int kSliceIncCap = 4096;  // incident (cell, node) entries of one slice's rows

// ------------------------------------------------------------------ plan
// Unique incident cells of each slice's rows, ascending.  FILL = false writes
// the count per slice; FILL = true writes the cells at out_ptr[s].
template <bool FILL>
__global__ __launch_bounds__(256) void k_slice_cells(int64_t num_slices, int32_t N,
                                                     const int32_t* __restrict__ sell_row,
                                                     const int32_t* __restrict__ inc_ptr,
                                                     const int32_t* __restrict__ inc_idx,
                                                     const int32_t* __restrict__ out_ptr, int32_t* __restrict__ out,
                                                     int32_t* __restrict__ pair_out, int32_t* __restrict__ err) {
  __shared__ int32_t buf[kSliceIncCap];
  __shared__ int32_t s_off[65];
  __shared__ int32_t s_scan[256];
  const int64_t s = blockIdx.x;
  const int t = threadIdx.x;
  if (s >= num_slices) return;
  int row = -1, b = 0, e = 0;
  if (t < 64) {
    row = sell_row[s * 64 + t];
    if (row >= 0) {
      b = inc_ptr[row];
      e = inc_ptr[row + 1];
    }
    s_off[t + 1] = e - b;
  }
  if (t == 0) s_off[0] = 0;
  __syncthreads();
  if (t == 0)
    for (int i = 1; i <= 64; ++i) s_off[i] += s_off[i - 1];
  __syncthreads();
  const int total = s_off[64];
  if (!FILL && t == 0) pair_out[s] = total <= kSliceIncCap ? total : 0;  // the slice's (cell, row node) pairs
  if (total > kSliceIncCap) {
    if (t == 0) {
      atomicMax(err, total);
      if (!FILL) out[s] = 0;
    }
    return;
  }
  if (t < 64)
    for (int j = b; j < e; ++j) buf[s_off[t] + (j - b)] = inc_idx[j] / N;
  int P = 1;
  while (P < total) P <<= 1;
  for (int i = total + t; i < P; i += 256) buf[i] = INT32_MAX;
  __syncthreads();
  // bitonic sort of buf[0, P)
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < P; i += 256) {
        const int l = i ^ j;
        if (l > i) {
          const int32_t x = buf[i], y = buf[l];
          const bool up = (i & k) == 0;
          if ((x > y) == up) {
            buf[i] = y;
            buf[l] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  // unique entries: each thread takes a contiguous run of the sorted list
  const int per = (total + 255) / 256;
  const int i0 = min(total, t * per), i1 = min(total, i0 + per);
  int cnt = 0;
  for (int i = i0; i < i1; ++i) cnt += (i == 0 || buf[i] != buf[i - 1]) ? 1 : 0;
  s_scan[t] = cnt;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int a = t >= off ? s_scan[t - off] : 0;
    __syncthreads();
    s_scan[t] += a;
    __syncthreads();
  }
  if (!FILL) {
    if (t == 255) out[s] = s_scan[255];
    return;
  }
  int pos = out_ptr[s] + s_scan[t] - cnt;
  for (int i = i0; i < i1; ++i)
    if (i == 0 || buf[i] != buf[i - 1]) out[pos++] = buf[i];
}

// contrib_loc[t] = (position of the cell in its slice's cell list) * N*N + a*N + b
// for every contribution t of every row (one thread per slice lane).
__global__ __launch_bounds__(256) void k_contrib_local(fem_pattern_t pat, const int32_t* __restrict__ sc_ptr,
                                                       const int32_t* __restrict__ sc, int32_t* __restrict__ loc,
                                                       int32_t* __restrict__ err) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= pat.num_slices * kSlice) return;
  const int row = pat.sell_row[g];
  if (row < 0) return;
  const int64_t s = g >> 6;
  const int NN = pat.nodes_per_cell * pat.nodes_per_cell;
  const int32_t* list = sc + sc_ptr[s];
  const int n = sc_ptr[s + 1] - sc_ptr[s];
  const int t0 = pat.contrib_ptr[pat.row_ptr[row]], t1 = pat.contrib_ptr[pat.row_ptr[row + 1]];
  for (int t = t0; t < t1; ++t) {
    const int32_t e = pat.contrib_idx[t];
    const int32_t c = e / NN;
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (list[mid] < c) lo = mid + 1; else hi = mid;
    }
    if (lo >= n || list[lo] != c) {
      atomicMax(err, 1);
      return;
    }
    loc[t] = lo * NN + (e - c * NN);
  }
}

// Chunk order of each slice's cell list.  A cell is ranked by the parity of the
// lowest slice lane among its nodes, then by id: on a structured HEX8 box in
// natural row order, the cells of one parity touch every row of a z-line slice
// once, so the assembly's fixed-size cell chunks keep every row lane busy
// (in id order a chunk covers half the rows).  Rewrites the slice's cell list
// in that order and re-expresses contrib_loc in it, each slot's contributions
// sorted by position (the order the assembly sums them in).  One 64-thread
// workgroup per slice; integer LDS minima only, so the result is deterministic.
__global__ __launch_bounds__(64) void k_slice_reorder(fem_pattern_t pat, const int32_t* __restrict__ inc_ptr,
                                                      const int32_t* __restrict__ inc_idx,
                                                      const int32_t* __restrict__ sc_ptr, int32_t* __restrict__ sc,
                                                      const int32_t* __restrict__ sp_ptr, int32_t* __restrict__ sp,
                                                      int32_t* __restrict__ loc, const int32_t* __restrict__ err) {
  __shared__ int32_t s_cell[kSliceIncCap];
  __shared__ int32_t s_rank[kSliceIncCap];
  __shared__ int32_t s_key[kSliceIncCap];
  __shared__ int32_t s_cnt[64];
  if (*err) return;  // the plan is refused (a slice over capacity or a cell missing from its list)
  const int64_t s = blockIdx.x;
  const int lane = threadIdx.x;
  const int c0 = sc_ptr[s], n = sc_ptr[s + 1] - c0;
  const int N = pat.nodes_per_cell, NN = N * N;
  for (int i = lane; i < n; i += 64) {
    s_cell[i] = sc[c0 + i];
    s_rank[i] = 64;  // lowest lane among the cell's nodes, 64 = none yet
  }
  __syncthreads();
  const int row = pat.sell_row[s * kSlice + lane];
  auto find = [&](int32_t c) {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_cell[mid] < c) lo = mid + 1; else hi = mid;
    }
    return lo;
  };
  if (row >= 0)
    for (int j = inc_ptr[row]; j < inc_ptr[row + 1]; ++j) {
      const int p = find(inc_idx[j] / N);
      if (p < n) atomicMin(&s_rank[p], lane);
    }
  __syncthreads();
  // parity-0 cells first, each group in id order: rank = group offset + count of the
  // group's cells before it (per-lane runs, then a scan over the 64 lanes)
  const int per = (n + 63) / 64;
  const int i0 = min(n, lane * per), i1 = min(n, i0 + per);
  int odd = 0;
  for (int i = i0; i < i1; ++i) odd += s_rank[i] & 1;
  s_cnt[lane] = odd;
  __syncthreads();
  if (lane == 0)
    for (int l = 1; l < 64; ++l) s_cnt[l] += s_cnt[l - 1];
  __syncthreads();
  const int n_odd = s_cnt[63], n_even = n - n_odd;
  int o = lane ? s_cnt[lane - 1] : 0;  // odd cells before i0
  int e = i0 - o;                      // even cells before i0
  for (int i = i0; i < i1; ++i) {
    const bool is_odd = s_rank[i] & 1;
    s_rank[i] = is_odd ? n_even + o++ : e++;
  }
  __syncthreads();
  for (int i = lane; i < n; i += 64) sc[c0 + s_rank[i]] = s_cell[i];
  // the slice's (cell, local node a, row lane) pairs, keyed (rank * 8 + a) * 64 + lane and
  // sorted: the order the assembly forms and sums the blocks in
  const int p0 = sp_ptr[s], np = sp_ptr[s + 1] - p0;
  const int ib = row >= 0 ? inc_ptr[row] : 0, ie = row >= 0 ? inc_ptr[row + 1] : 0;
  __syncthreads();
  s_cnt[lane] = ie - ib;
  __syncthreads();
  if (lane == 0)
    for (int l = 1; l < 64; ++l) s_cnt[l] += s_cnt[l - 1];
  __syncthreads();
  {
    int pos = lane ? s_cnt[lane - 1] : 0;
    for (int j = ib; j < ie; ++j) {
      const int32_t e = inc_idx[j];
      const int q = find(e / N);
      s_key[pos++] = ((q < n ? s_rank[q] : 0) * 8 + e % N) * 64 + lane;
    }
  }
  int P2 = 1;
  while (P2 < np) P2 <<= 1;
  for (int i = np + lane; i < P2; i += 64) s_key[i] = INT32_MAX;
  __syncthreads();
  for (int k = 2; k <= P2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < P2; i += 64) {
        const int l = i ^ j;
        if (l > i) {
          const int32_t x = s_key[i], y = s_key[l];
          if ((x > y) == ((i & k) == 0)) {
            s_key[i] = y;
            s_key[l] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = lane; i < np; i += 64) sp[p0 + i] = s_key[i];
  if (row < 0) return;
  auto pair_of = [&](int32_t key) {
    int lo = 0, hi = np;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_key[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
  };
  const int rb = pat.row_ptr[row], len = pat.row_ptr[row + 1] - rb;
  for (int k = 0; k < len; ++k) {
    const int t0 = pat.contrib_ptr[rb + k], t1 = pat.contrib_ptr[rb + k + 1];
    for (int t = t0; t < t1; ++t) {
      const int v = loc[t];  // in [0, n * NN): k_contrib_local found every contribution's cell
      const int a = (v % NN) / N, b = v % N;
      loc[t] = pair_of((s_rank[v / NN] * 8 + a) * 64 + lane) * N + b;
    }
    for (int t = t0 + 1; t < t1; ++t) {  // insertion sort of the slot's few entries
      const int v = loc[t];
      int u = t - 1;
      while (u >= t0 && loc[u] > v) {
        loc[u + 1] = loc[u];
        --u;
      }
      loc[u + 1] = v;
    }
  }
}

// ------------------------------------------------------------- assembly
template <class MAT>
struct alignas(8) StRow {
  typename MAT::template St<double> s[8];
  double pad;  // odd stride between cells
};

// Isotropic linear laws (elasticity, SIMP): lambda = c0 E and mu = c1 E at every
// point, so K_ab = c0 M + c1 M^T + c1 tr(M) I with the single weighted sum
// M = sum_q JxW E g_a (x) g_b (the closed form of MatElastic::tangent).
template <class MAT>
struct IsoLinear : std::false_type {
  __device__ static double E(const fem_material_t&, int64_t) { return 1.0; }
};
template <bool S>
struct IsoLinear<MatElastic<S>> : std::true_type {
  // the modulus factor E of lambda = c0 E, mu = c1 E at point cq
  __device__ static double E(const fem_material_t& m, int64_t cq) {
    return S ? MatElastic<S>::modulus(m, m.param[cq]) : 1.0;
  }
};

constexpr int kAsmMaxK = 4;               // slots per thread: rows of at most 8 * 4 = 32 blocks
constexpr int kResWave0 = kAsmWaves - 3;  // waves summing the residual's components
constexpr int kAsmGS = 8 * 8 * 3 + 1;     // odd stride (in doubles) between cells' gradients

// Chunk sizes: at most CH cells and PM (cell, row node) pairs per chunk (on a
// z-line slice every cell holds two of the slice's nodes, so PM = 2 CH).
template <class MAT>
struct AsmCfg {
  using St = typename MAT::template St<double>;
  static constexpr bool kBig = sizeof(St) > 4 * sizeof(double);  // Neo-Hookean, J2: large point states
  static constexpr int CH = kBig ? 16 : 32;
  static constexpr int PM = 2 * CH;
};
constexpr int kSlicePairCap = 1024;  // pairs (and cells) of one slice, staged in LDS

// The slice's (cell, local node a, row lane) pairs are walked in chunks; per
// chunk:
//   A1  one thread per (cell, node): node coordinates (and u) to LDS;
//   A2  one thread per (cell, quadrature point): Jacobian, inverse, JxW, the 8
//       physical gradients, the material state (and the weighted stress);
//   B1  one thread per (pair, column node b): the 3x3 block K_ab of the pair's
//       cell (the pair's 8 threads share g_a and read the 8 g_b side by side)
//       and, for b < V, component b of the pair's residual contribution, all
//       staged in LDS;
//   B2  every owner thread (row lane, slot) adds the staged blocks of its
//       slots, in pair order (which is the plan's fixed cell order).
template <class MAT, typename TO, bool DIFF, bool RES>
__global__ __launch_bounds__(kAsmThreads) void k_assemble_hex8(
    fem_mesh_t mesh, fem_material_t mat, fem_pattern_t pat, const int32_t* __restrict__ sc_ptr,
    const int32_t* __restrict__ sc, const int32_t* __restrict__ sp_ptr, const int32_t* __restrict__ sp,
    const int32_t* __restrict__ cloc, const double* __restrict__ u, TO* __restrict__ vals, TO* __restrict__ dinv,
    const double* __restrict__ base, double* __restrict__ res) {
  constexpr int N = 8, Q = 8, V = MAT::V, VV = V * V, MAXK = kAsmMaxK, GS = kAsmGS;
  constexpr int CH = AsmCfg<MAT>::CH, PM = AsmCfg<MAT>::PM;
  constexpr bool kIso = IsoLinear<MAT>::value;
  __shared__ double s_dN[Q][N][3];
  __shared__ double s_w[Q];
  __shared__ double s_g[CH][GS];
  __shared__ double s_wq[CH][Q + 1];                        // JxW (elastic laws: JxW E)
  __shared__ StRow<MAT> s_st[kIso ? 1 : CH];
  __shared__ double s_K[PM][N][VV];                         // staged blocks of the chunk's pairs
  __shared__ double s_R[RES ? PM : 1][V];                   // staged residual contributions
  __shared__ int64_t s_c[CH];
  __shared__ int32_t s_key[kSlicePairCap];   // the slice's pair keys
  __shared__ int32_t s_cid[kSlicePairCap];   // the slice's cells, in rank order
  // A1 stages coordinates and u in s_buf; with the residual, A2's weighted stresses reuse it
  constexpr int kXU = CH * N * (3 + V);
  constexpr int kP = RES ? CH * (Q * V * 3 + 1) : 0;
  __shared__ double s_buf[kXU > kP ? kXU : kP];
  double(*s_x)[N][3] = reinterpret_cast<double(*)[N][3]>(s_buf);
  double(*s_u)[N][V] = reinterpret_cast<double(*)[N][V]>(s_buf + CH * N * 3);
  double(*s_P)[Q * V * 3 + 1] = reinterpret_cast<double(*)[Q * V * 3 + 1]>(s_buf);
  // the difference form's row sums are staged in s_K after the last chunk, half the waves at a time
  constexpr int kHalf = kAsmWaves / 2;
  static_assert(kHalf * kSlice * VV <= PM * N * VV, "row-sum staging does not fit in s_K");
  double(*s_red)[kSlice][VV] = reinterpret_cast<double(*)[kSlice][VV]>(&s_K[0][0][0]);

  const int64_t s = blockIdx.x;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const double iso_c1 = kIso ? mat.c[1] : 0.0;  // elastic laws: the slots sum M; K is formed at the end
  load_tables(mesh, s_dN, s_w);
  const int sbase = pat.slice_ptr[s];
  const int L = (pat.slice_ptr[s + 1] - sbase) >> 6;
  const int row = pat.sell_row[s * kSlice + lane];
  const bool real = row >= 0;
  const int rb = real ? pat.row_ptr[row] : 0;
  const int len = real ? pat.row_ptr[row + 1] - rb : 0;
  const int ds = real ? pat.diag_slot[row] : -1;
  const int c0 = sc_ptr[s], ncs = sc_ptr[s + 1] - c0;
  const int p_base = sp_ptr[s], np = sp_ptr[s + 1] - p_base;
  // np, ncs <= kSlicePairCap: fem_assemble refuses plans with larger slices
  for (int i = t; i < np; i += kAsmThreads) s_key[i] = sp[p_base + i];
  for (int i = t; i < ncs; i += kAsmThreads) s_cid[i] = sc[c0 + i];

  double acc[MAXK][VV];
  // cur: next contribution of each owned slot; nxt: its entry (pair * 8 + b), loaded one
  // contribution ahead
  int cur[MAXK], cend[MAXK], nxt[MAXK];
#pragma unroll
  for (int m = 0; m < MAXK; ++m) {
    const int k = w + kAsmWaves * m;
    cur[m] = cend[m] = 0;
    if (k < len) {
      cur[m] = pat.contrib_ptr[rb + k];
      cend[m] = pat.contrib_ptr[rb + k + 1];
    }
    nxt[m] = cur[m] < cend[m] ? cloc[cur[m]] : INT32_MAX;
#pragma unroll
    for (int v = 0; v < VV; ++v) acc[m][v] = 0.0;
  }
  // residual: component ri of the row is summed by wave kResWave0 + ri from the diagonal
  // slot's list, whose entries are exactly the row's pairs
  const int ri = w - kResWave0;
  const bool resw = RES && real && ri >= 0 && ri < V;
  int rcur = 0, rend = 0, rnxt = INT32_MAX;
  if (resw) {
    rcur = pat.contrib_ptr[rb + ds];
    rend = pat.contrib_ptr[rb + ds + 1];
    rnxt = rcur < rend ? cloc[rcur] : INT32_MAX;
  }
  double r = 0.0;
  __syncthreads();

  for (int p0 = 0; p0 < np;) {
    // ---- the chunk: pairs [p0, p1), cells of ranks [r0, r0 + ncell)
    const int r0 = s_key[p0] >> 9;
    int lo = p0 + 1, hi = min(np, p0 + PM);  // first pair of rank >= r0 + CH, at most p0 + PM
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((s_key[mid] >> 9) < r0 + CH) lo = mid + 1; else hi = mid;
    }
    const int p1 = lo;
    const int ncell = (s_key[p1 - 1] >> 9) - r0 + 1;
    // ---- A1: node coordinates (and u)
    if (t < ncell * N) {
      const int cc = t / N, a = t % N;
      const int64_t c = s_cid[r0 + cc];
      if (a == 0) s_c[cc] = c;
      const int64_t na = mesh.cells[c * N + a];
#pragma unroll
      for (int j = 0; j < 3; ++j) s_x[cc][a][j] = mesh.points[na * 3 + j];
      if (MAT::kNeedsU || RES)
#pragma unroll
        for (int i = 0; i < V; ++i) s_u[cc][a][i] = u[na * V + i];
    }
    __syncthreads();
    // ---- A2: per (cell, quadrature point) geometry and material state
    double Pw[RES ? V : 1][3];
    const int acc_cc = t / Q, acc_q = t % Q;
    const bool a2 = t < ncell * Q;
    if (a2) {
      const int cc = acc_cc, q = acc_q;
      const int64_t c = s_c[cc];
      const double x00 = s_x[cc][0][0], x01 = s_x[cc][0][1], x02 = s_x[cc][0][2];
      double J[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) J[k] = 0.0;
#pragma unroll
      for (int a = 1; a < N; ++a) {
        const double d0 = s_x[cc][a][0] - x00, d1 = s_x[cc][a][1] - x01, d2 = s_x[cc][a][2] - x02;
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
      double G[V][3];
#pragma unroll
      for (int i = 0; i < V; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) G[i][j] = 0.0;
#pragma unroll
      for (int a = 0; a < N; ++a) {
        const double e0 = s_dN[q][a][0], e1 = s_dN[q][a][1], e2 = s_dN[q][a][2];
        double g[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          g[j] = e0 * iJ[0 * 3 + j] + e1 * iJ[1 * 3 + j] + e2 * iJ[2 * 3 + j];
          s_g[cc][(q * N + a) * 3 + j] = g[j];
        }
        if (MAT::kNeedsU || RES) {
#pragma unroll
          for (int i = 0; i < V; ++i) {
            const double ua = s_u[cc][a][i];
#pragma unroll
            for (int j = 0; j < 3; ++j) G[i][j] += ua * g[j];
          }
        }
      }
      const auto st = MAT::template prep<double>(mat, c * Q + q, G);
      if constexpr (!kIso) s_st[cc].s[q] = st;
      s_wq[cc][q] = kIso ? jxw * IsoLinear<MAT>::E(mat, c * Q + q) : jxw;  // elastic laws: JxW E
      if (RES) {
        double P[V][3];
        MAT::stress(st, G, P);
#pragma unroll
        for (int i = 0; i < V; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) Pw[i][j] = P[i][j] * jxw;
      }
    }
    if (RES) {
      // the weighted stresses overwrite the coordinates' storage once every thread has read it
      __syncthreads();
      if (a2)
#pragma unroll
        for (int i = 0; i < V; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) s_P[acc_cc][(acc_q * V + i) * 3 + j] = Pw[i][j];
    }
    __syncthreads();
    // ---- B1: one block per (pair, column node)
    if (t < (p1 - p0) * N) {
      const int pp = t / N, b = t % N;
      const int key = s_key[p0 + pp];
      const int cc = (key >> 9) - r0, a = (key >> 6) & 7;
      double K[V][V];
#pragma unroll
      for (int i = 0; i < V; ++i)
#pragma unroll
        for (int k = 0; k < V; ++k) K[i][k] = 0.0;
      if constexpr (kIso) {
        // M = sum_q JxW E g_a (x) g_b
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const double* ga = &s_g[cc][(q * N + a) * 3];
          const double* gb = &s_g[cc][(q * N + b) * 3];
          const double wq = s_wq[cc][q];
          const double gb0 = gb[0], gb1 = gb[1], gb2 = gb[2];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const double wa = wq * ga[i];
            K[i][0] += wa * gb0;
            K[i][1] += wa * gb1;
            K[i][2] += wa * gb2;
          }
        }
      } else {
#pragma unroll 2
        for (int q = 0; q < Q; ++q)
          MAT::tangent(s_st[cc].s[q], &s_g[cc][(q * N + a) * 3], &s_g[cc][(q * N + b) * 3], s_wq[cc][q], K);
      }
#pragma unroll
      for (int i = 0; i < V; ++i)
#pragma unroll
        for (int k = 0; k < V; ++k) s_K[pp][b][i * V + k] = K[i][k];
      if (RES && b < V) {
        double rr = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
          for (int j = 0; j < 3; ++j) rr += s_P[cc][(q * V + b) * 3 + j] * s_g[cc][(q * N + a) * 3 + j];
        s_R[RES ? pp : 0][b] = rr;
      }
    }
    __syncthreads();
    // ---- B2: owners add their slots' blocks of this chunk, in pair order
    const int lim = p1 * N;
#pragma unroll
    for (int m = 0; m < MAXK; ++m) {
      int e = nxt[m];
      while (e < lim) {
        ++cur[m];
        const int en = cur[m] < cend[m] ? cloc[cur[m]] : INT32_MAX;
        const double* kb = &s_K[e / N - p0][e % N][0];
#pragma unroll
        for (int v = 0; v < VV; ++v) acc[m][v] += kb[v];
        e = en;
      }
      nxt[m] = e;
    }
    if (RES && resw) {
      while (rnxt < lim) {
        ++rcur;
        const int en = rcur < rend ? cloc[rcur] : INT32_MAX;
        r += s_R[RES ? rnxt / N - p0 : 0][ri];
        rnxt = en;
      }
    }
    __syncthreads();
    p0 = p1;
  }
  if constexpr (kIso) {
    // the slots summed M: K = c0 M + c1 M^T + c1 tr(M) I
    const double c0m = mat.c[0];
#pragma unroll
    for (int m = 0; m < MAXK; ++m) {
      double M[9];
#pragma unroll
      for (int v = 0; v < 9; ++v) M[v] = acc[m][v];
      const double tr = iso_c1 * (M[0] + M[4] + M[8]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[m][i * 3 + k] = c0m * M[i * 3 + k] + iso_c1 * M[k * 3 + i] + (i == k ? tr : 0.0);
    }
  }

  // ---- difference form: row sums of all slots, combined in a fixed order
  double rsum[DIFF ? VV : 1];
  if (DIFF) {
#pragma unroll
    for (int v = 0; v < VV; ++v) {
      double p = 0.0;
#pragma unroll
      for (int m = 0; m < MAXK; ++m) p += acc[m][v];
      rsum[v] = p;
    }
    // round 1: waves kHalf.. stage, waves 0..kHalf-1 add (pairs w, w + kHalf)
    if (w >= kHalf)
#pragma unroll
      for (int v = 0; v < VV; ++v) s_red[w - kHalf][lane][v] = rsum[v];
    __syncthreads();
    if (w < kHalf)
#pragma unroll
      for (int v = 0; v < VV; ++v) rsum[v] += s_red[w][lane][v];
    __syncthreads();
    // round 2: the pair sums, read by every thread
    if (w < kHalf)
#pragma unroll
      for (int v = 0; v < VV; ++v) s_red[w][lane][v] = rsum[v];
    __syncthreads();
#pragma unroll
    for (int v = 0; v < VV; ++v) {
      double sum = s_red[0][lane][v];
#pragma unroll
      for (int o = 1; o < kHalf; ++o) sum += s_red[o][lane][v];
      rsum[v] = sum;
    }
  }

  // ---- stores: one coalesced 64-lane store per value of each owned slot
  TO* vp = vals + (size_t)sbase * VV + lane;
#pragma unroll
  for (int m = 0; m < MAXK; ++m) {
    const int k = w + kAsmWaves * m;
    if (k >= L) continue;
    if (k == ds && dinv) {
#pragma unroll
      for (int c = 0; c < V; ++c) {
        const double d = acc[m][c * V + c];
        dinv[(int64_t)row * V + c] = d != 0.0 ? TO(1.0 / d) : TO(0);
      }
    }
#pragma unroll
    for (int v = 0; v < VV; ++v) vp[((size_t)k * VV + v) * kSlice] = TO((DIFF && k == ds) ? rsum[v] : acc[m][v]);
  }
  if (RES && resw) {
    const int64_t idx = (int64_t)row * V + ri;
    res[idx] = (base ? base[idx] : 0.0) + r;
  }
}

// ------------------------------------------------------------- dispatch
template <class F>
static int with_mat8(const fem_material_t& mat, int vec, F&& f) {
  switch (mat.kind) {
    case FEM_MAT_POISSON:
      if (vec == 1) return f(MatPoisson{});
      break;
    case FEM_MAT_LINEAR_ELASTIC:
      if (vec == 3) return f(MatElastic<false>{});
      break;
    case FEM_MAT_SIMP:
      if (vec == 3 && mat.has_param) return f(MatElastic<true>{});
      break;
    case FEM_MAT_NEO_HOOKEAN:
      if (vec == 3) return f(MatNeoHookean{});
      break;
    case FEM_MAT_J2_PLASTIC:
      if (vec == 3 && mat.state) return f(MatJ2{});
      break;
    default:
      break;
  }
  set_error("fem_assemble: material kind %d (vec=%d) is assembled by the two-pass path "
            "(fem_cell_tangent + fem_gather_matrix)", mat.kind, vec);
  return FEM_ERR_BAD_ARG;
}

}  // namespace fem

using namespace fem;

extern "C" {

int fem_assembly_plan_counts(const fem_pattern_t* pat, const int32_t* inc_ptr, const int32_t* inc_idx,
                             int32_t* slice_cell_counts, int32_t* slice_pair_counts, int32_t* err_flag,
                             void* stream) {
  FEM_CHECK_ARG(pat && inc_ptr && inc_idx && slice_cell_counts && slice_pair_counts && err_flag,
                "fem_assembly_plan_counts: NULL argument");
  if (pat->num_slices == 0) return FEM_OK;
  hipStream_t st = as_stream(stream);
  FEM_HIP(hipMemsetAsync(err_flag, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_slice_cells<false>, dim3((unsigned)pat->num_slices), dim3(256), 0, st, pat->num_slices,
                     pat->nodes_per_cell, pat->sell_row, inc_ptr, inc_idx, (const int32_t*)nullptr,
                     slice_cell_counts, slice_pair_counts, err_flag);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_assembly_plan_fill(const fem_pattern_t* pat, const int32_t* inc_ptr, const int32_t* inc_idx,
                           const int32_t* slice_cell_ptr, const int32_t* slice_pair_ptr, int32_t* slice_cells,
                           int32_t* slice_pairs, int32_t* contrib_loc, int32_t* err_flag, void* stream) {
  FEM_CHECK_ARG(pat && inc_ptr && inc_idx && slice_cell_ptr && slice_pair_ptr && slice_cells && slice_pairs &&
                    contrib_loc && err_flag,
                "fem_assembly_plan_fill: NULL argument");
  if (pat->num_slices == 0) return FEM_OK;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_slice_cells<true>, dim3((unsigned)pat->num_slices), dim3(256), 0, st, pat->num_slices,
                     pat->nodes_per_cell, pat->sell_row, inc_ptr, inc_idx, slice_cell_ptr, slice_cells,
                     (int32_t*)nullptr, err_flag);
  FEM_LAUNCHED();
  const int64_t lanes = pat->num_slices * kSlice;
  hipLaunchKernelGGL(k_contrib_local, dim3((unsigned)ceil_div(lanes, 256)), dim3(256), 0, st, *pat, slice_cell_ptr,
                     slice_cells, contrib_loc, err_flag);
  FEM_LAUNCHED();
  hipLaunchKernelGGL(k_slice_reorder, dim3((unsigned)pat->num_slices), dim3(64), 0, st, *pat, inc_ptr, inc_idx,
                     slice_cell_ptr, slice_cells, slice_pair_ptr, slice_pairs, contrib_loc, err_flag);
  FEM_LAUNCHED();
  return FEM_OK;
}

int fem_assemble(const fem_mesh_t* mesh, const fem_material_t* mat, const fem_pattern_t* pat,
                 const fem_assembly_plan_t* plan, const double* u, int32_t precision, int32_t form, void* vals,
                 void* dinv, const double* base, double* res, void* stream) {
  FEM_CHECK_ARG(mesh && mat && pat && plan && u && vals, "fem_assemble: NULL argument");
  FEM_CHECK_ARG(mesh->nodes_per_cell == 8 && mesh->num_quads == 8,
                "fem_assemble: single-pass assembly covers HEX8 cells with 8 quadrature points");
  FEM_CHECK_ARG(pat->nodes_per_cell == 8 && pat->vec == mesh->vec && pat->num_cells == mesh->num_cells,
                "fem_assemble: pattern and mesh disagree");
  FEM_CHECK_ARG(plan->slice_cell_ptr && plan->slice_cells && plan->slice_pair_ptr && plan->slice_pairs &&
                    plan->contrib_loc,
                "fem_assemble: plan not built");
  FEM_CHECK_ARG(plan->max_slots >= 0 && plan->max_slots <= kAsmWaves * kAsmMaxK,
                "fem_assemble: a slice has %d blocks per row (at most %d)", plan->max_slots, kAsmWaves * kAsmMaxK);
  FEM_CHECK_ARG(plan->max_pairs >= 0 && plan->max_pairs <= kSlicePairCap,
                "fem_assemble: a slice has %d (cell, row node) pairs (at most %d)", plan->max_pairs, kSlicePairCap);
  FEM_CHECK_ARG((precision == FEM_F64 && form == FEM_FORM_STANDARD) ||
                    (precision == FEM_F32 && form == FEM_FORM_DIFFERENCE),
                "fem_assemble: FP64 standard-form or FP32 difference-form storage");
  FEM_CHECK_ARG(!mat->has_param || mat->param, "fem_assemble: material param missing");
  if (pat->num_slices == 0) return FEM_OK;
  hipStream_t st = as_stream(stream);
  const dim3 grid((unsigned)pat->num_slices), block(kAsmThreads);
  return with_mat8(*mat, mesh->vec, [&](auto m) -> int {
    using MAT = decltype(m);
    auto go = [&](auto res_c) {
      constexpr bool R = decltype(res_c)::value;
      if (precision == FEM_F64)
        hipLaunchKernelGGL((k_assemble_hex8<MAT, double, false, R>), grid, block, 0, st, *mesh, *mat, *pat,
                           plan->slice_cell_ptr, plan->slice_cells, plan->slice_pair_ptr, plan->slice_pairs,
                           plan->contrib_loc, u, (double*)vals, (double*)dinv, base, res);
      else
        hipLaunchKernelGGL((k_assemble_hex8<MAT, float, true, R>), grid, block, 0, st, *mesh, *mat, *pat,
                           plan->slice_cell_ptr, plan->slice_cells, plan->slice_pair_ptr, plan->slice_pairs,
                           plan->contrib_loc, u, (float*)vals, (float*)dinv, base, res);
    };
    if (res)
      go(std::true_type{});
    else
      go(std::false_type{});
    FEM_LAUNCHED();
    return FEM_OK;
  });
}

}  // extern "C"
