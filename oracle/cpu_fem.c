/*
 * cpu_fem.c - multi-threaded CPU restatement of the reference's HEX8
 * linear-elasticity Newton step, used as the CPU baseline of the benchmark.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): bench.py's cpu_baseline
 * leg and the tests call it; the product never does.  It restates, in FP64:
 *   fe.py:112-141        shape gradients and JxW of trilinear hexahedra
 *                        (2x2x2 Gauss points on [0,1]^3, basis.py:56)
 *   problem.py:189-214,  the per-cell tangent of the laplace kernel for
 *   262-266              isotropic elasticity (the jacfwd of sigma : grad v,
 *                        written in closed form), with a per-cell modulus
 *                        scale E_c (lambda E_c, mu E_c)
 *   solver.py:469-528    COO -> CSR with duplicates summed, zeroRows(diag 1)
 *   solver.py:63-92      jax_solve: BiCGSTAB with M = x / diag(A), stopping
 *                        when |r|^2 <= max(tol^2 |b|^2, atol^2)
 * plus Jacobi-preconditioned CG with the same stopping rule.  Matrices are
 * scalar CSR (int64 row pointers, int32 columns) with sorted columns, the
 * layout PETSc AIJ and scipy produce.  Cell loops and vector kernels are
 * OpenMP-parallel; the tangent's duplicates are summed with atomic adds, so the
 * last bits of a matrix value depend on the thread schedule.
 *
 * Build: make -C oracle  (gcc -O3 -fopenmp, no GPU code).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int cfem_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------- pattern */
static int cmp_i32(const void* a, const void* b) {
  const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

enum { kMaxInc = 256 };  /* incident cells per node */

/* Node graph rows: nodes sharing a cell with node i, sorted (node_ptr n+1, node_col).
 * Pass node_col = NULL to get the row lengths in node_ptr[1..n] (not yet scanned). */
static int node_rows(int64_t n_nodes, int64_t n_cells, const int32_t* cells, const int64_t* inc_ptr,
                     const int32_t* inc_cell, int64_t* node_ptr, int32_t* node_col) {
  int bad = 0;
#pragma omp parallel
  {
    int32_t buf[kMaxInc * 8];
#pragma omp for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n_nodes; ++i) {
      int m = 0;
      for (int64_t t = inc_ptr[i]; t < inc_ptr[i + 1] && t - inc_ptr[i] < kMaxInc; ++t) {
        const int32_t* cn = cells + (int64_t)inc_cell[t] * 8;
        for (int a = 0; a < 8; ++a) buf[m++] = cn[a];
      }
      if (inc_ptr[i + 1] - inc_ptr[i] > kMaxInc) bad = 1;
      qsort(buf, m, sizeof(int32_t), cmp_i32);
      int u = 0;
      for (int k = 0; k < m; ++k)
        if (k == 0 || buf[k] != buf[k - 1]) buf[u++] = buf[k];
      if (node_col)
        memcpy(node_col + node_ptr[i], buf, (size_t)u * sizeof(int32_t));
      else
        node_ptr[i + 1] = u;
    }
  }
  (void)n_cells;
  return bad;
}

/* Scalar CSR pattern of vec components per node (dof = vec * node + comp): returns nnz
 * (-1: out of memory, -2: a node with more than 256 incident cells); with rowptr/col NULL
 * it only counts. */
int64_t cfem_csr_pattern(int64_t n_nodes, int64_t n_cells, const int32_t* cells, int32_t vec, int64_t* rowptr,
                         int32_t* col) {
  int64_t* inc_ptr = calloc((size_t)n_nodes + 1, sizeof(int64_t));
  int32_t* inc_cell = malloc((size_t)n_cells * 8 * sizeof(int32_t));
  int64_t* node_ptr = calloc((size_t)n_nodes + 1, sizeof(int64_t));
  if (!inc_ptr || !inc_cell || !node_ptr) return -1;
  for (int64_t c = 0; c < n_cells; ++c)
    for (int a = 0; a < 8; ++a) inc_ptr[cells[c * 8 + a] + 1]++;
  for (int64_t i = 0; i < n_nodes; ++i) inc_ptr[i + 1] += inc_ptr[i];
  int64_t* fill = malloc((size_t)n_nodes * sizeof(int64_t));
  memcpy(fill, inc_ptr, (size_t)n_nodes * sizeof(int64_t));
  for (int64_t c = 0; c < n_cells; ++c)
    for (int a = 0; a < 8; ++a) inc_cell[fill[cells[c * 8 + a]]++] = (int32_t)c;
  free(fill);
  if (node_rows(n_nodes, n_cells, cells, inc_ptr, inc_cell, node_ptr, NULL)) {
    free(inc_ptr);
    free(inc_cell);
    free(node_ptr);
    return -2;
  }
  for (int64_t i = 0; i < n_nodes; ++i) node_ptr[i + 1] += node_ptr[i];
  const int64_t nnz = node_ptr[n_nodes] * vec * vec;
  if (rowptr && col) {
    int32_t* node_col = malloc((size_t)node_ptr[n_nodes] * sizeof(int32_t));
    node_rows(n_nodes, n_cells, cells, inc_ptr, inc_cell, node_ptr, node_col);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_nodes; ++i) {
      const int64_t len = node_ptr[i + 1] - node_ptr[i];
      for (int a = 0; a < vec; ++a) {
        const int64_t r = i * vec + a;
        const int64_t base = node_ptr[i] * vec * vec + a * len * vec;
        rowptr[r] = base;
        for (int64_t k = 0; k < len; ++k)
          for (int b = 0; b < vec; ++b) col[base + k * vec + b] = node_col[node_ptr[i] + k] * vec + b;
      }
    }
    rowptr[n_nodes * vec] = nnz;
    free(node_col);
  }
  free(inc_ptr);
  free(inc_cell);
  free(node_ptr);
  return nnz;
}

/* ------------------------------------------------------------ assembly */
static const double kNodes[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

static double det3(const double* A) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
}

static int64_t find_col(const int32_t* col, int64_t lo, int64_t hi, int32_t c) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (col[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

/* vals (nnz) = sum over cells of the elastic cell tangents (lambda E_c, mu E_c; E may be
 * NULL for E_c = 1), summed into the CSR slots of the pattern. */
void cfem_assemble_elastic(int64_t n_cells, const double* points, const int32_t* cells, double lam, double mu,
                           const double* E, const int64_t* rowptr, const int32_t* col, double* vals) {
  const double g = 0.5 - sqrt(3.0) / 6.0;
  const double gp[2] = {g, 1.0 - g};
  double dN[8][8][3];  /* [q][a][d] reference gradients at the 8 Gauss points (weights 1/8) */
  for (int q = 0; q < 8; ++q) {
    const double xi[3] = {gp[q >> 2], gp[(q >> 1) & 1], gp[q & 1]};
    for (int a = 0; a < 8; ++a)
      for (int d = 0; d < 3; ++d) {
        double v = 1.0;
        for (int e = 0; e < 3; ++e) {
          const double s = kNodes[a][e];
          v *= (e == d) ? (2.0 * s - 1.0) : (s * xi[e] + (1.0 - s) * (1.0 - xi[e]));
        }
        dN[q][a][d] = v;
      }
  }
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n_cells; ++c) {
    const int32_t* cn = cells + c * 8;
    double K[24][24];
    memset(K, 0, sizeof(K));
    const double s = E ? E[c] : 1.0;
    for (int q = 0; q < 8; ++q) {
      double J[9] = {0};
      for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) J[i * 3 + j] += points[(int64_t)cn[a] * 3 + i] * dN[q][a][j];
      const double det = det3(J), id = 1.0 / det;
      double iJ[9];
      iJ[0] = (J[4] * J[8] - J[5] * J[7]) * id;
      iJ[1] = (J[2] * J[7] - J[1] * J[8]) * id;
      iJ[2] = (J[1] * J[5] - J[2] * J[4]) * id;
      iJ[3] = (J[5] * J[6] - J[3] * J[8]) * id;
      iJ[4] = (J[0] * J[8] - J[2] * J[6]) * id;
      iJ[5] = (J[2] * J[3] - J[0] * J[5]) * id;
      iJ[6] = (J[3] * J[7] - J[4] * J[6]) * id;
      iJ[7] = (J[1] * J[6] - J[0] * J[7]) * id;
      iJ[8] = (J[0] * J[4] - J[1] * J[3]) * id;
      double G[8][3];
      for (int a = 0; a < 8; ++a)
        for (int j = 0; j < 3; ++j) G[a][j] = dN[q][a][0] * iJ[j] + dN[q][a][1] * iJ[3 + j] + dN[q][a][2] * iJ[6 + j];
      const double w = det * 0.125, l = w * lam * s, m = w * mu * s;
      for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
          const double gg = m * (G[a][0] * G[b][0] + G[a][1] * G[b][1] + G[a][2] * G[b][2]);
          for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k)
              K[a * 3 + i][b * 3 + k] += l * G[a][i] * G[b][k] + m * G[a][k] * G[b][i] + (i == k ? gg : 0.0);
        }
    }
    for (int a = 0; a < 8; ++a)
      for (int i = 0; i < 3; ++i) {
        const int64_t r = (int64_t)cn[a] * 3 + i;
        for (int b = 0; b < 8; ++b)
          for (int k = 0; k < 3; ++k) {
            const int64_t p = find_col(col, rowptr[r], rowptr[r + 1], cn[b] * 3 + k);
#pragma omp atomic
            vals[p] += K[a * 3 + i][b * 3 + k];
          }
      }
  }
}

/* PETSc zeroRows(rows, diag = 1): the rows keep their pattern. */
void cfem_zero_rows(const int64_t* rowptr, const int32_t* col, double* vals, int64_t n_rows, const int64_t* rows) {
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < n_rows; ++t) {
    const int64_t r = rows[t];
    for (int64_t p = rowptr[r]; p < rowptr[r + 1]; ++p) vals[p] = (col[p] == r) ? 1.0 : 0.0;
  }
}

/* ------------------------------------------------------------- Krylov */
static void spmv(int64_t n, const int64_t* rp, const int32_t* col, const double* v, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) s += v[p] * x[col[p]];
    y[i] = s;
  }
}

static double dot(int64_t n, const double* x, const double* y) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

void cfem_spmv(int64_t n, const int64_t* rp, const int32_t* col, const double* v, const double* x, double* y) {
  spmv(n, rp, col, v, x, y);
}

/* BiCGSTAB of jax.scipy.sparse.linalg.bicgstab with M = x / diag(A) (solver.py:63-92).
 * Stops when |r|^2 <= max(tol^2 |b|^2, atol^2) or after maxiter; with fixed > 0 it runs
 * exactly `fixed` iterations (a timing block).  x holds x0.  Returns the iterations. */
int64_t cfem_bicgstab(int64_t n, const int64_t* rp, const int32_t* col, const double* v, const double* b, double* x,
                      double tol, double atol, int64_t maxiter, int64_t fixed) {
  double* w = malloc((size_t)n * 8 * 8);
  double *r = w, *rh = w + n, *p = w + 2 * n, *q = w + 3 * n, *ph = w + 4 * n, *s = w + 5 * n, *sh = w + 6 * n,
         *t = w + 7 * n, *dinv = NULL;
  dinv = malloc((size_t)n * 8);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double d = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
      if (col[k] == i) d = v[k];
    dinv[i] = 1.0 / d;
  }
  spmv(n, rp, col, v, x, q);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    r[i] = b[i] - q[i];
    rh[i] = r[i];
    p[i] = r[i];
    q[i] = r[i];
  }
  const double bb = dot(n, b, b);
  const double a2 = fmax(tol * tol * bb, atol * atol);
  double alpha = 1.0, rho = 1.0, omega = 1.0;
  int64_t k = 0;
  const int64_t lim = fixed > 0 ? fixed : maxiter;
  while (k < lim && (fixed > 0 || dot(n, r, r) > a2)) {
    const double rho_ = dot(n, rh, r);
    const double beta = rho_ / rho * alpha / omega;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      p[i] = r[i] + beta * (p[i] - omega * q[i]);
      ph[i] = p[i] * dinv[i];
    }
    spmv(n, rp, col, v, ph, q);
    alpha = rho_ / dot(n, rh, q);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) s[i] = r[i] - alpha * q[i];
    if (fixed == 0 && dot(n, s, s) < a2) {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) {
        x[i] += alpha * ph[i];
        r[i] = s[i];
      }
      ++k;
      break;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) sh[i] = s[i] * dinv[i];
    spmv(n, rp, col, v, sh, t);
    omega = dot(n, t, s) / dot(n, t, t);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * ph[i] + omega * sh[i];
      r[i] = s[i] - omega * t[i];
    }
    rho = rho_;
    ++k;
  }
  free(w);
  free(dinv);
  return k;
}

/* Jacobi-preconditioned CG, same stopping rule and timing mode as cfem_bicgstab. */
int64_t cfem_pcg(int64_t n, const int64_t* rp, const int32_t* col, const double* v, const double* b, double* x,
                 double tol, double atol, int64_t maxiter, int64_t fixed) {
  double* w = malloc((size_t)n * 8 * 5);
  double *r = w, *z = w + n, *p = w + 2 * n, *q = w + 3 * n, *dinv = w + 4 * n;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double d = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
      if (col[k] == i) d = v[k];
    dinv[i] = 1.0 / d;
  }
  spmv(n, rp, col, v, x, q);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    r[i] = b[i] - q[i];
    z[i] = r[i] * dinv[i];
    p[i] = z[i];
  }
  const double a2 = fmax(tol * tol * dot(n, b, b), atol * atol);
  double rz = dot(n, r, z);
  int64_t k = 0;
  const int64_t lim = fixed > 0 ? fixed : maxiter;
  while (k < lim && (fixed > 0 || dot(n, r, r) > a2)) {
    spmv(n, rp, col, v, p, q);
    const double alpha = rz / dot(n, p, q);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = r[i] * dinv[i];
    }
    const double rz2 = dot(n, r, z);
    const double beta = rz2 / rz;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz = rz2;
    ++k;
  }
  free(w);
  return k;
}
