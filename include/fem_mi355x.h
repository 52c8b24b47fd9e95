/*
 * fem_mi355x.h - C ABI of the MI355X (gfx950) finite-element library.
 *
 * The library replaces the device work of jax-fem's assembly -> linear solve ->
 * adjoint path (jax_fem/fe.py, jax_fem/problem.py, jax_fem/solver.py).  It is a
 * set of stateless launchers over caller-owned device buffers: every pointer
 * named here is device memory unless its comment says "host"; nothing is
 * allocated inside except where a comment says so.  Every call is enqueued on
 * the HIP stream passed as `stream` (a hipStream_t, NULL = default stream) and
 * returns an int status (FEM_OK = 0, negative on error).  On error,
 * fem_last_error() returns a message (thread-local, valid until the next call
 * on the same thread).  Calls that read results back to the host (sizes,
 * solver statistics) synchronise the stream; they say so below.
 *
 * Conventions (same as the reference):
 *   dof index      = vec * node + component              (problem.py:89)
 *   cell dof index = vec * a + component, a = local node   (problem.py:197, 211)
 *   per-cell tangent K[c][a*vec+i][b*vec+j] = dR_{a,i} / du_{b,j}  (problem.py:262-266)
 *   residual R(u) = sum_q P(grad u) : grad v JxW + mass + surface terms
 *                                                        (problem.py:189-259, 296)
 *
 * Matrix storage.  The global tangent has a vec x vec block structure over the
 * node graph.  The node-graph pattern is kept as block CSR (row_ptr/col_idx,
 * columns sorted), which expands exactly to the scalar CSR pattern the
 * reference builds with PETSc from its COO I/J (problem.py:86-107,
 * solver.py:472-478).  Values live in a sliced-ELL layout over 64-row slices
 * (SELL-64): row `row` occupies lane l = sell_pos[row] % 64 of slice
 * s = sell_pos[row] / 64 (sell_row is the inverse map; rows are in natural
 * order unless length-sorted within 2048-row windows, see fem_sell_order).
 * Block k of the row is stored at block position slice_ptr[s] + 64*k + l, and
 * value v (0..vec*vec-1, row-major in the block) of it at
 * vals[slice_ptr[s]*vec*vec + (k*vec*vec + v)*64 + l].  Slots past a row's
 * length hold zero values and the row itself as column.
 */
#ifndef FEM_MI355X_H
#define FEM_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  FEM_OK = 0,
  FEM_ERR_BAD_ARG = -1,       /* invalid argument / unsupported combination   */
  FEM_ERR_NOT_CONVERGED = -2, /* Krylov solve hit maxiter                     */
  FEM_ERR_NAN = -3,           /* non-finite value met in a solve              */
  FEM_ERR_HIP = -4,           /* HIP runtime error                            */
  FEM_ERR_CAPACITY = -5       /* a fixed per-row capacity was exceeded        */
};

/* ---- enums ----------------------------------------------------------- */
enum { FEM_F32 = 0, FEM_F64 = 1,            /* arithmetic precision          */
       FEM_BF16X3 = 2 };                     /* fem_cell_tangent only: HEX27 elastic
                                                tangents from split-BF16 MFMA
                                                products (hi*hi + hi*lo + lo*hi,
                                                FP32 accumulation), Ke in FP32 */
enum { FEM_HEX8 = 8, FEM_HEX27 = 27 };       /* = nodes per cell              */
enum {                                       /* storage form of the matrix    */
  FEM_FORM_STANDARD = 0,      /* block (i, j) holds A_ij                      */
  FEM_FORM_DIFFERENCE = 1,    /* diagonal slot holds R_i = sum_j A_ij and the
                                 product is y_i = sum_{j!=i} A_ij (x_j - x_i)
                                 + R_i x_i (exact on translations)            */
  FEM_FORM_SCALED_DIFF16 = 2  /* AMG level 0 only: FP16 values of D^-1 A (rows
                                 scaled by the point-Jacobi inverse diagonal)
                                 in difference form                          */
};
enum {                                       /* constitutive laws (tensor_map)*/
  FEM_MAT_NONE = 0,           /* no laplace term (a problem whose weak form is all
                                 universal kernels): the cell kernels write zeros
                                                                  (vec 1 to 4) */
  FEM_MAT_POISSON = 1,        /* P = c0 * grad u                      (vec 1) */
  FEM_MAT_LINEAR_ELASTIC = 2, /* sigma = c0 tr(eps) I + 2 c1 eps      (vec 3) */
  FEM_MAT_SIMP = 3,           /* as LINEAR_ELASTIC with lambda = c0 E(t),
                                 mu = c1 E(t), E(t) = c2 + (c3 - c2) t^c4, t =
                                 per-quad parameter                  (vec 3) */
  FEM_MAT_NEO_HOOKEAN = 4,    /* P = d psi / dF, psi = mu/2 (J^-2/3 I1 - 3) +
                                 kappa/2 (J - 1)^2, mu = c0 s, kappa = c1 s,
                                 s = per-quad parameter or 1         (vec 3) */
  FEM_MAT_J2_PLASTIC = 5,     /* small-strain J2 perfect plasticity, radial
                                 return from the per-quad state (sigma_old,
                                 eps_old): lambda = c0, mu = c1, yield stress
                                 c2 (tests/benchmarks/plasticity/
                                 test_plasticity.py:32-64)          (vec 3) */
  FEM_MAT_TABULATED = 6       /* any tensor map: P(grad u) and its derivative
                                 read per quadrature point from `state`,
                                 evaluated beforehand at the current u (the
                                 value_and_jacfwd of problem.py:262-266):
                                 state + cq * FEM_TAB_STRIDE(vec) holds P[V][3],
                                 D[V][3][V][3] = dP_ij / dG_kl, then
                                 dP/dtheta[V][3] (read when has_param), cq =
                                 cell * Q + quad                  (vec 1, 3) */
};
#define FEM_TAB_STRIDE(vec) (6 * (vec) + 9 * (vec) * (vec))

/* ---- plain descriptor structs (device pointers) ------------------------ */
typedef struct {
  int32_t nodes_per_cell;        /* N: 8 (HEX8) or 27 (HEX27)                 */
  int32_t num_quads;             /* Q                                         */
  int32_t vec;                   /* 1 or 3; 1 to 4 (the sum over the variables of
                                    a several-variable problem) with FEM_MAT_NONE */
  int32_t num_face_quads;        /* FQ                                        */
  int64_t num_cells;
  int64_t num_nodes;
  const double*  points;         /* (num_nodes, 3)                            */
  const int32_t* cells;          /* (num_cells, N), meshio/VTK node order     */
  const double*  shape_vals;     /* (Q, N)     basis.py:141-175               */
  const double*  shape_grads_ref;/* (Q, N, 3)                                 */
  const double*  quad_weights;   /* (Q)                                       */
  const double*  face_shape_vals;/* (6, FQ, N) basis.py:178-250               */
  const double*  face_shape_grads_ref; /* (6, FQ, N, 3)                      */
  const double*  face_weights;   /* (6, FQ)                                   */
  const double*  face_normals;   /* (6, 3)                                    */
} fem_mesh_t;

typedef struct {
  int32_t kind;                  /* FEM_MAT_*                                 */
  int32_t has_param;             /* 1: `param` holds one value per (cell, quad) */
  double  c[6];                  /* constants, meaning per kind above         */
  const double* param;           /* (num_cells, Q) or NULL                    */
  const double* state;           /* J2_PLASTIC: (num_cells, Q, 18) = sigma_old
                                    (3x3 row-major) then eps_old; TABULATED:
                                    (num_cells, Q, FEM_TAB_STRIDE(vec));
                                    else NULL                                */
} fem_material_t;

typedef struct {
  int64_t num_rows;              /* block rows (= mesh nodes)                 */
  int64_t nnzb;                  /* stored blocks                             */
  int64_t num_slices;            /* ceil(num_rows / 64)                       */
  int64_t sell_blocks;           /* slice_ptr[num_slices]                     */
  int32_t vec;                   /* block size                                */
  int32_t num_cells;             /* cells feeding the contributions          */
  int32_t nodes_per_cell;
  int32_t pad;
  const int32_t* row_ptr;        /* (num_rows + 1)                            */
  const int32_t* col_idx;        /* (nnzb), sorted within a row               */
  const int32_t* diag_slot;      /* (num_rows) slot of the diagonal block, -1 */
  const int32_t* contrib_ptr;    /* (nnzb + 1) into contrib_idx               */
  const int32_t* contrib_idx;    /* (num_cells*N*N) cell-block id c*N*N+a*N+b */
  const int32_t* slice_ptr;      /* (num_slices + 1) in blocks                */
  const int32_t* sell_col;       /* (sell_blocks) column node per SELL block  */
  const int32_t* sell_row;       /* (num_slices*64) block row held by each slice
                                    lane, -1 for padding lanes               */
  const int32_t* sell_pos;       /* (num_rows) lane s*64+l holding each row   */
} fem_pattern_t;

/* FP64 matrix-free tangent operator y = dR/du(u) x on free dofs, y = x on
 * Dirichlet dofs; used for the FP64 residual of iterative refinement. */
typedef struct {
  const fem_mesh_t*     mesh;
  const fem_material_t* mat;
  const double*  u;              /* (num_nodes*vec) linearisation point       */
  const int32_t* inc_ptr;        /* node -> (cell, local node) incidence      */
  const int32_t* inc_idx;
  int64_t        n_bc;
  const int32_t* bc_dofs;        /* (n_bc) unique Dirichlet dofs              */
  double*        cell_scratch;   /* (num_cells * N * vec) doubles             */
  const double*  mass_M;         /* NULL, or dm/du of a u-dependent mass term
                                    (C, Q, vec, vec), added as fem_cell_mass_jvp */
  const double*  cell_K;         /* NULL, or per-cell matrices (C, N, N, vec,
                                    vec) added as fem_cell_matvec            */
} fem_operator_t;

typedef struct {
  double  rtol;                  /* stop when |r| <= max(rtol |b|, atol)      */
  double  atol;
  int32_t maxiter;
  int32_t check_every;           /* iterations between host status reads      */
  int32_t precision;             /* FEM_F64: plain FP64 PCG. FEM_F32: FP32 PCG,
                                    x accumulated in FP64, FP64 true-residual
                                    replacement through `op` (required)       */
  int32_t precond;               /* 1: Jacobi, 0: none                        */
  double  reliable_delta;        /* FP32 mode: replace the residual when the
                                    recursive |r| fell by this factor         */
  int32_t xcd_remap;             /* 1: XCD-contiguous block order in SpMV     */
  int32_t timing;                /* 1: time the iteration phases with HIP
                                    events (adds a record per launch)         */
} fem_pcg_options_t;

typedef struct {
  int32_t iterations;
  int32_t replacements;          /* FP64 residual replacements (FP32 mode)    */
  int32_t status;                /* FEM_OK / FEM_ERR_NOT_CONVERGED / NAN      */
  int32_t pad;
  double  b_norm;                /* |b|                                       */
  double  res_norm;              /* final true residual norm |b - A x|        */
  double  spmv_ms;               /* timing=1: mean duration of the SpMV kernel
                                    (A p and the partials of <p, A p>)       */
  double  update_ms;             /* timing=1: mean duration of the rest of an
                                    iteration (p update, reductions, x/r/z)  */
  int32_t timed_launches;        /* launches averaged into spmv_ms            */
  int32_t pad2;
} fem_pcg_result_t;

/* ---- library --------------------------------------------------------- */
const char* fem_last_error(void);
int fem_version(void);

/* ---- pattern construction (problem.py:86-107, solver.py:472-478) ------- */
/* Node -> (cell, local node) incidence lists, sorted ascending by c*N+a.
 * inc_ptr (num_nodes+1), inc_idx (num_entities*N); ws >= fem_scan_ws_bytes(num_nodes)
 * + 4*num_nodes bytes.  Also used for boundary faces (conn = face cells). */
int fem_incidence(int64_t num_entities, int32_t N, const int32_t* conn,
                  int64_t num_nodes, int32_t* inc_ptr, int32_t* inc_idx,
                  void* ws, void* stream);
/* Workspace bytes for fem_exclusive_scan over n items. */
int64_t fem_scan_ws_bytes(int64_t n);
/* out[0..n] = exclusive prefix sum of in[0..n-1]; out[n] = total. */
int fem_exclusive_scan(int64_t n, const int32_t* in, int32_t* out, void* ws, void* stream);
/* Per node row: sort the (column, cell-block) candidates of its incident cells,
 * write the sorted cell-block ids to contrib_idx[inc_ptr[row]*N ...] and the
 * number of distinct columns to row_len.  FEM_ERR_CAPACITY past 2048
 * candidates per row. */
int fem_pattern_rows(const fem_mesh_t* mesh, const int32_t* inc_ptr, const int32_t* inc_idx,
                     int32_t* contrib_idx, int32_t* row_len, int32_t* err_flag, void* stream);
/* Fill col_idx, contrib_ptr, diag_slot from row_ptr (scan of row_len). */
int fem_pattern_fill(const fem_mesh_t* mesh, const int32_t* inc_ptr, const int32_t* contrib_idx,
                     const int32_t* row_ptr, int32_t* col_idx, int32_t* contrib_ptr,
                     int32_t* diag_slot, void* stream);
/* Row order of the SELL slices.  sort = 0: natural order.  sort = 1: within
 * each window of 2048 rows, rows sorted stably by decreasing length
 * (SELL-C-sigma), so a slice holds rows of equal length and carries little
 * padding.  sell_row (num_slices*64), sell_pos (num_rows). */
int fem_sell_order(int64_t num_rows, int32_t sort, const int32_t* row_ptr, int32_t* sell_row, int32_t* sell_pos,
                   void* stream);
/* 64*max(row_len) per slice (rows taken in sell_row order), to be scanned into slice_ptr. */
int fem_sell_slice_sizes(int64_t num_rows, const int32_t* row_ptr, const int32_t* sell_row, int32_t* slice_blocks,
                         void* stream);
/* Fill sell_col from the block CSR pattern (needs row_ptr, col_idx, slice_ptr, sell_row). */
int fem_sell_fill(const fem_pattern_t* pat, int32_t* sell_col, void* stream);

/* ---- element kernels (fe.py:112-218, problem.py:189-356) --------------- */
/* shape_grads (C,Q,N,3), JxW (C,Q), quad points (C,Q,3); any output may be NULL. */
int fem_geometry(const fem_mesh_t* mesh, double* shape_grads, double* JxW,
                 double* quad_points, void* stream);
/* Per-cell tangent blocks Ke[c][a][b][vec][vec] in `precision` (float/double;
 * FEM_BF16X3 writes float).  HEX27 cells with 27 quadrature points and a
 * LINEAR_ELASTIC or SIMP law run on MFMA (FP64 MFMA for FEM_F64). */
int fem_cell_tangent(const fem_mesh_t* mesh, const fem_material_t* mat, int32_t precision,
                     const double* u, void* Ke, void* stream);
/* u and its gradient at the quadrature points (problem.py:200-203, 228):
 * G[c][q][i][j] = sum_a u[cells[c][a]][i] * dN_a/dx_j (FP64, (C, Q, vec, 3)),
 * U[c][q][i] = sum_a u[cells[c][a]][i] * N_a(q) (FP64, (C, Q, vec)).  Either
 * output may be NULL. */
int fem_quad_grad(const fem_mesh_t* mesh, const double* u, double* G, double* U, void* stream);
/* Tangent of a mass term whose map depends on u (the jacfwd of the mass
 * kernel, problem.py:216-236, 262-266): adds to the per-cell tangents
 * Ke[c][a][b][i][k] += sum_q N_a(q) N_b(q) M[c][q][i][k] JxW(c, q), where
 * M = dm/du (FP64, (C, Q, vec, vec)) was evaluated at u beforehand; Ke is in
 * `precision` (FEM_F32 or FEM_F64), laid out as fem_cell_tangent writes it. */
int fem_cell_mass_tangent(const fem_mesh_t* mesh, const double* M, int32_t precision, void* Ke, void* stream);
/* Its product with a vector: re[c][a][i] += sum_q N_a(q) JxW(c, q)
 * sum_k M[c][q][i][k] v_q[k], v_q = sum_b N_b(q) v[cells[c][b]] (FP64). */
int fem_cell_mass_jvp(const fem_mesh_t* mesh, const double* M, const double* v, double* re, void* stream);
/* Product of per-cell matrices with a nodal vector (the matrix-free form of
 * a per-cell weak form's jacfwd, problem.py:289-301):
 * re[c][a][i] += sum_b sum_k K[c][a][b][i][k] v[cells[c][b]][k] (FP64, K laid
 * out as fem_cell_tangent writes Ke). */
int fem_cell_matvec(const fem_mesh_t* mesh, const double* K, const double* v, double* re, void* stream);
/* Per-cell internal-force residual re[c][a][vec] (FP64). */
int fem_cell_residual(const fem_mesh_t* mesh, const fem_material_t* mat, const double* u,
                      double* re, void* stream);
/* Per-cell tangent-vector product re[c][a][vec] = K_c(u) v_c (FP64). */
int fem_cell_jvp(const fem_mesh_t* mesh, const fem_material_t* mat, const double* u,
                 const double* v, double* re, void* stream);
/* Per-cell load re[c][a][i] = sum_q m[c][q][i] N_a(q) JxW(c,q) (mass_map term, FP64). */
int fem_cell_load(const fem_mesh_t* mesh, const double* m, double* re, void* stream);
/* Transpose of fem_cell_load: lq[c][q][i] = sum_a lam[cells[c][a]][i] N_a(q) JxW(c,q)
 * (parameter gradient of a mass_map term). */
int fem_cell_load_adjoint(const fem_mesh_t* mesh, const double* lam, double* lq, void* stream);
/* Adjoint kernel (solver.py:1386-1416): g[c][q] = -JxW lam_q . dP/dtheta(grad u_q) grad N,
 * i.e. -lambda^T dR/dtheta_{c,q} for the per-quad material parameter. */
int fem_cell_adjoint(const fem_mesh_t* mesh, const fem_material_t* mat, const double* u,
                     const double* lam, double* g, void* stream);
/* Boundary faces (fe.py:143-183, 199-218): faces (S,2) = (cell, local face).
 * Outputs (any may be NULL): face shape grads (S,FQ,N,3), nanson scale (S,FQ),
 * physical quad points (S,FQ,3). */
int fem_face_geometry(const fem_mesh_t* mesh, int64_t S, const int32_t* faces,
                      double* face_shape_grads, double* nanson, double* quad_points, void* stream);
/* Face load rf[s][a][i] = sum_fq t[s][fq][i] Nf_a(fq) nanson(s,fq) (surface_map term). */
int fem_face_load(const fem_mesh_t* mesh, int64_t S, const int32_t* faces, const double* t,
                  double* rf, void* stream);
/* Transpose of fem_face_load: lq[s][fq][i] = sum_a lam[node_a][i] Nf_a(fq) nanson(s,fq). */
int fem_face_load_adjoint(const fem_mesh_t* mesh, int64_t S, const int32_t* faces, const double* lam,
                          double* lq, void* stream);

/* ---- scatter / global matrix (problem.py:426-437, solver.py:522-529) --- */
/* out[n*vec+i] = base[n*vec+i] (0 if base is NULL; base may alias out)
 *              + sum over incidences t of node n of ent[inc_idx[t]*vec+i],
 * summed in incidence order (problem.py:426-437 scatter-add, deterministic). */
int fem_gather_nodal(int64_t num_nodes, int32_t vec, const int32_t* inc_ptr, const int32_t* inc_idx,
                     const double* ent, const double* base, double* out, void* stream);
/* SELL values (in `precision`) = sum of the cell blocks of Ke (in `ke_precision`, equal
 * to `precision` or FP64), duplicates summed in cell order then rounded once
 * (PETSc setValuesCOO, solver.py:522-526), stored in `form`; dinv (may be
 * NULL) receives 1/A[d,d] in `precision`. */
int fem_gather_matrix(const fem_pattern_t* pat, int32_t ke_precision, const void* Ke, int32_t precision,
                      int32_t form, void* vals, void* dinv, void* stream);
/* Dirichlet row elimination (PETSc zeroRows, diag = 1, solver.py:527-528): row d zeroed,
 * A[d,d] = 1, dinv[d] = 1 (dinv may be NULL).  Valid for both forms. */
int fem_matrix_bc_rows(const fem_pattern_t* pat, int32_t precision, void* vals, int64_t n_bc,
                       const int32_t* bc_dofs, void* dinv, void* stream);
/* Jacobi M = x / diag(A) (solver.py:75-77): dinv[d] = 1/A[d,d] (0 where the
 * diagonal is absent), in `precision`.  Standard-form matrices only (the
 * diagonal slot of a difference-form matrix holds row sums; fem_gather_matrix
 * returns the true D^-1 for those). */
int fem_jacobi(const fem_pattern_t* pat, int32_t precision, const void* vals, void* dinv, void* stream);
/* res[d] = u[d] - scale*g[d] on Dirichlet dofs (solver.py:290-304). */
int fem_residual_bc(int64_t n_bc, const int32_t* bc_dofs, const double* bc_vals, const double* u,
                    double* res, double scale, void* stream);
/* y = A x in `precision` (x, y of that precision), matrix stored in `form`
 * (the BCOO product inside jax_solve, solver.py:67-84).  FP64 arithmetic
 * for both storage precisions. */
int fem_spmv(const fem_pattern_t* pat, int32_t precision, int32_t form, const void* vals, const void* x,
             void* y, void* stream);
/* Scalar CSR export, the PETSc AIJ / getValuesCSR layout of get_A
 * (solver.py:540-553): indptr (vec*rows+1), indices (nnz), data (nnz, FP64). */
int fem_export_csr(const fem_pattern_t* pat, int32_t precision, int32_t form, const void* vals,
                   int32_t* indptr, int32_t* indices, double* data, void* stream);
/* y = dR/du(u) x (FP64, matrix-free, BC rows = identity). */
int fem_operator_apply(const fem_operator_t* op, const double* x, double* y, void* stream);

/* ---- single-pass HEX8 assembly (problem.py:447-460 and solver.py:522-526 in
 *      one kernel, optionally with the residual scatter of problem.py:426-437)
 * The cell tangents are formed and summed straight into the SELL values: a
 * workgroup owns one 64-row slice and walks the cells incident to its rows in
 * chunks; no per-cell tangent buffer exists.  The plan (built once per pattern)
 * lists each slice's incident cells and re-expresses every contribution as a
 * position in that list. */
typedef struct {
  const int32_t* slice_cell_ptr; /* (num_slices + 1)                           */
  const int32_t* slice_cells;    /* unique incident cells of each slice, in chunk
                                    order: cells whose lowest slice lane is even
                                    first, each group by ascending id         */
  const int32_t* slice_pair_ptr; /* (num_slices + 1)                           */
  const int32_t* slice_pairs;    /* the slice's (cell, local node a, row lane)
                                    pairs, sorted keys (rank * 8 + a) * 64 +
                                    lane, rank = position in slice_cells      */
  const int32_t* contrib_loc;    /* (num_cells*N*N), aligned with contrib_idx:
                                    pair * N + b (pair = index in the slice's
                                    pair list), ascending within each slot    */
  int32_t        max_slots;      /* largest slot count of a slice (<= 32)     */
  int32_t        max_pairs;      /* largest pair count of a slice (<= 1024)   */
} fem_assembly_plan_t;

/* Plan, pass 1: unique incident cells and (cell, row node) pairs per slice
 * (counts, to be scanned into slice_cell_ptr and slice_pair_ptr).  err_flag > 0
 * after the call: a slice's rows have more than 4096 incidences (use the
 * two-pass path). */
int fem_assembly_plan_counts(const fem_pattern_t* pat, const int32_t* inc_ptr, const int32_t* inc_idx,
                             int32_t* slice_cell_counts, int32_t* slice_pair_counts, int32_t* err_flag,
                             void* stream);
/* Plan, pass 2: slice_cells, slice_pairs and contrib_loc (err_flag as above, or
 * 1 when a contribution's cell is missing from its slice list). */
int fem_assembly_plan_fill(const fem_pattern_t* pat, const int32_t* inc_ptr, const int32_t* inc_idx,
                           const int32_t* slice_cell_ptr, const int32_t* slice_pair_ptr, int32_t* slice_cells,
                           int32_t* slice_pairs, int32_t* contrib_loc, int32_t* err_flag, void* stream);
/* Assemble the tangent of HEX8 cells with 8 quadrature points at u into the
 * SELL values: FP64 standard form (precision FEM_F64, form FEM_FORM_STANDARD)
 * or FP32 difference form (FEM_F32, FEM_FORM_DIFFERENCE); the cell tangents
 * are FP64 and each matrix value is their FP64 sum in the plan's chunk
 * order, rounded once.  dinv (may be NULL): Jacobi inverse diagonal in `precision`.  res (may
 * be NULL): R(u) without Dirichlet rows, = base (may be NULL) + the scattered
 * internal force, FP64.  Dirichlet rows are applied afterwards by
 * fem_matrix_bc_rows. */
int fem_assemble(const fem_mesh_t* mesh, const fem_material_t* mat, const fem_pattern_t* pat,
                 const fem_assembly_plan_t* plan, const double* u, int32_t precision, int32_t form, void* vals,
                 void* dinv, const double* base, double* res, void* stream);

/* ---- Krylov solve (solver.py:63-92, 261-284) -------------------------- */
/* Workspace bytes for fem_pcg_solve. */
int64_t fem_pcg_ws_bytes(const fem_pattern_t* pat, int32_t precision);
/* Jacobi-PCG on the row-eliminated tangent.  b, x are FP64 (x holds x0 on
 * entry).  vals/dinv are in opt->precision; an FP64 matrix is in standard
 * form, an FP32 matrix in difference form.  `op` is required for FEM_F32
 * and ignored for FEM_F64.  `host_status` is pinned host memory of at least
 * 128 bytes.  Synchronises the stream.  Returns FEM_OK or the solve status. */
int fem_pcg_solve(const fem_pattern_t* pat, const void* vals, const void* dinv,
                  const fem_operator_t* op, const double* b, double* x,
                  const fem_pcg_options_t* opt, void* ws, void* host_status,
                  fem_pcg_result_t* result /* host */, void* stream);

/* Bytes of the library-held device scratch of fem_pcg_solve_noalloc (one
 * array per GPU the library runs on, held until the process ends). */
#define FEM_PCG_NOALLOC_SCRATCH_BYTES (128LL << 20)
/* No counterpart in the reference: fem_pcg_solve for callers that cannot
 * allocate device memory (serving loops and harnesses that forbid hipMalloc).
 * Same arguments and results as fem_pcg_solve without `ws`: the workspace is
 * the library's scratch array, so the call allocates nothing.  Systems whose
 * fem_pcg_ws_bytes exceeds FEM_PCG_NOALLOC_SCRATCH_BYTES (about 2.79 M FP64
 * dofs, 4.78 M in FEM_F32) return FEM_ERR_CAPACITY; opt->timing must be 0.
 * Every call on one GPU uses that GPU's single array: calls that overlap on
 * one GPU (two host threads, two streams) race. */
int fem_pcg_solve_noalloc(const fem_pattern_t* pat, const void* vals, const void* dinv,
                          const fem_operator_t* op, const double* b, double* x,
                          const fem_pcg_options_t* opt, void* host_status,
                          fem_pcg_result_t* result /* host */, void* stream);

/* Linear multipoint constraints u = P u_r (the P_mat of solver.py:404-406,
 * 545-551, 1290-1339): P (rows x cols) and its transpose as scalar CSR. */
typedef struct {
  int64_t rows;                  /* N: dofs of the mesh                       */
  int64_t cols;                  /* M: reduced dofs                           */
  const int32_t* p_ptr;          /* P: (N+1), (nnz), (nnz)                    */
  const int32_t* p_idx;
  const double*  p_val;
  const int32_t* pt_ptr;         /* P^T: (M+1), (nnz), (nnz)                  */
  const int32_t* pt_idx;
  const double*  pt_val;
} fem_constraint_t;

/* y = A x for a scalar CSR matrix (FP64): the products with P and P^T. */
int fem_csr_spmv(int64_t rows, const int32_t* ptr, const int32_t* idx, const double* val, const double* x,
                 double* y, void* stream);

/* Jacobi-preconditioned BiCGSTAB, the method of jax_solve (solver.py:63-92),
 * for tangents that are not symmetric and for constrained systems.  FP64
 * standard-form matrix only (opt->precision = FEM_F64); dinv from
 * fem_gather_matrix or fem_jacobi (required when opt->precond).  With P ==
 * NULL, b and x (x holds x0) are FP64 vectors of all dofs.  With a
 * constraint P, the system is the reduced P^T A P x_r = b_r of get_A
 * (solver.py:545-551), b and x have P->cols entries, and the Jacobi
 * preconditioner is sum_k P_kj^2 A_kk.  Uses rtol, atol, maxiter,
 * check_every and precond of `opt`.  `host_status` is pinned host memory of
 * at least 128 bytes.  Synchronises.  Returns FEM_OK or the solve status
 * (result->status). */
int64_t fem_bicgstab_ws_bytes(const fem_pattern_t* pat, const fem_constraint_t* P);
int fem_bicgstab_solve(const fem_pattern_t* pat, const double* vals, const double* dinv,
                       const fem_constraint_t* P, const double* b, double* x, const fem_pcg_options_t* opt,
                       void* ws, void* host_status, fem_pcg_result_t* result /* host */, void* stream);

/* ---- smoothed-aggregation AMG (the amgx_solver option, solver.py:141-259,
 *      261-284) ------------------------------------------------------------
 * The reference hands the assembled PETSc matrix to AMGX.  Here the hierarchy
 * is built on the GPU from the assembled matrix: MIS-2 aggregation of the
 * block graph, tentative prolongators from the near null space (rigid-body
 * modes for vec 3, constants for vec 1), one damped-Jacobi smoothing step of
 * the prolongator, Galerkin products by a deterministic block SpGEMM, and a
 * dense inverse of the coarsest operator.  The preconditioner is a V-cycle
 * with Chebyshev smoothing inside FP64 CG.  Setup calls run in FP64; the
 * host reads sizes between them (the library allocates nothing), and every
 * result is bit-reproducible run to run (no floating-point atomics). */
typedef struct {
  int64_t        rows;           /* block rows                                */
  int64_t        cols;           /* block columns                             */
  int64_t        nnzb;           /* stored blocks                             */
  int32_t        br, bc;         /* block shape (1, 3 or 6)                   */
  int32_t        f32;            /* 0: vals are FP64, 1: FP32 (the setup of an
                                    FP32 hierarchy; in a fem_amg_t the levels'
                                    storage follows fem_amg_t.precision)     */
  int32_t        pad;
  const int32_t* row_ptr;        /* (rows+1)                                  */
  const int32_t* col_idx;        /* (nnzb), ascending within a row            */
  const void*    vals;           /* (nnzb*br*bc), row-major blocks            */
} fem_bcsr_t;

typedef struct {
  const fem_pattern_t* pat;      /* level 0: SELL pattern of the matrix       */
  const void*    sell_vals;      /* level 0: standard-form SELL values        */
  fem_bcsr_t     A;              /* levels >= 1: Galerkin operator            */
  fem_bcsr_t     P;              /* prolongator from the next level           */
  fem_bcsr_t     R;              /* restriction P^T                           */
  const double*  dinv;           /* (rows*bs) inverse diagonal, FP64          */
  double         lmax;           /* largest eigenvalue estimate of D^-1 A     */
} fem_amg_level_t;

typedef struct {
  int32_t        num_levels;     /* smoothed levels (>= 0; with 0 the matrix
                                    itself is the coarsest level and the cycle
                                    is its dense inverse, levels may be NULL) */
  int32_t        precision;      /* value storage of A, P, R and the level-0
                                    SELL copy: FEM_F32 or FEM_F64; vectors and
                                    arithmetic are FP64                       */
  int32_t        degree;         /* Chebyshev degree of pre/post smoothing on
                                    level 0; coarser levels: 2 * degree + 1 (<= 9) */
  int32_t        level0_form;    /* form of the level-0 SELL values: FEM_FORM_
                                    STANDARD, FEM_FORM_DIFFERENCE (FP32), or
                                    FEM_FORM_SCALED_DIFF16 (FP16 D^-1 A, with
                                    the level's dinv; FP32 hierarchy)        */
  const fem_amg_level_t* levels; /* host array [num_levels]                   */
  int64_t        coarse_n;       /* unknowns of the coarsest level            */
  const double*  coarse_inv;     /* (coarse_n x coarse_n) dense inverse       */
} fem_amg_t;

typedef struct {
  double  rtol;                  /* stop when |r| <= max(rtol |b|, atol)      */
  double  atol;
  int32_t maxiter;
  int32_t check_every;           /* iterations between host status reads      */
  int32_t timing;                /* 1: at every status read, time the last
                                    V-cycle and its level-0 Chebyshev SELL
                                    kernels (the pre-smoothing step and both
                                    post-smoothing steps at degree >= 2) with
                                    HIP events: result update_ms = the mean
                                    cycle, spmv_ms = the mean Chebyshev step */
  int32_t pad;
  double  reliable_delta;        /* FP32 outer operator: replace the residual
                                    when |r| fell by this factor             */
} fem_amg_options_t;

/* Workspace bytes of fem_amg_aggregate on an n-node graph. */
int64_t fem_amg_aggregate_ws_bytes(int64_t n);
/* MIS-2 aggregation of the graph (row_ptr, col_idx; self loops included):
 * agg[i] = aggregate of node i, *n_agg (host) = number of aggregates.
 * host_status: pinned host memory >= 128 bytes.  Synchronises the stream. */
int fem_amg_aggregate(int64_t n, const int32_t* row_ptr, const int32_t* col_idx, int32_t* agg, int64_t* n_agg,
                      void* ws, void* host_status, void* stream);
/* Members of each aggregate: agg_ptr (n_agg+1), memb (n) ascending within an
 * aggregate.  ws >= 4*n_agg rounded up to 256, + fem_scan_ws_bytes(n_agg). */
int fem_amg_members(int64_t n, const int32_t* agg, int64_t n_agg, int32_t* agg_ptr, int32_t* memb, void* ws,
                    void* stream);
/* Tentative prolongator by modified Gram-Schmidt per aggregate: P0 (n x bs x
 * ns, the one block of node i's row, column agg[i]) and the coarse near null
 * space Bc (n_agg*ns x ns).  Level 0 (B == NULL): rigid-body modes (vec 3,
 * ns 6, rotations about the aggregate's first node) or constants (vec 1, ns
 * 1) from `points` (n x 3); dofs with bc_mask[d] != 0 get zero rows.  Coarser
 * levels: the rows of B (n*bs x ns). */
int fem_amg_tentative(int64_t n, int32_t bs, int32_t ns, int64_t n_agg, const int32_t* agg_ptr,
                      const int32_t* memb, const double* points, const uint8_t* bc_mask, const double* B,
                      double* P0, double* Bc, void* stream);
/* Dirichlet mask: mask[d] = 1 for d in bc_dofs, else 0 (n dofs). */
int fem_amg_bc_mask(int64_t n, int64_t n_bc, const int32_t* bc_dofs, uint8_t* mask, void* stream);
/* FP32 difference-form copy (off-diagonal blocks rounded, diagonal slot = the
 * FP64 row sum rounded once) of a standard-form FP64 SELL matrix. */
int fem_amg_sell_to_f32_diff(const fem_pattern_t* pat, const double* sell_vals, float* out, void* stream);
/* Block CSR copy (FEM_F64 or FEM_F32 values) of a standard-form FP64 SELL
 * matrix; the pattern's row_ptr/col_idx are the CSR structure. */
int fem_amg_sell_to_bcsr(const fem_pattern_t* pat, const double* sell_vals, void* bcsr_vals,
                         int32_t precision, void* stream);
/* Both FP32 copies above (block CSR and difference-form SELL) in one pass over
 * the FP64 SELL values: the level-0 inputs of an FP32 hierarchy.  dinv (n, may
 * be NULL): the inverse diagonal of the block-CSR copy, as fem_amg_dinv forms
 * it (1 / the rounded diagonal entries, 0 for a zero entry). */
int fem_amg_sell_to_f32_copies(const fem_pattern_t* pat, const double* sell_vals, float* bcsr_vals,
                               float* diff_vals, double* dinv, void* stream);
/* dinv = 1 / diag(A) (FP64 A); on levels >= 1 (fix_zero = 1) a zero
 * diagonal entry (an empty coarse row) is first set to 1 in A. */
int fem_amg_dinv(const fem_bcsr_t* A, int32_t fix_zero, double* dinv, void* stream);
/* Largest-eigenvalue estimate of D^-1 A by `steps` power steps from a fixed
 * start vector.  A is the SELL matrix (pat, sell_vals) if pat != NULL --
 * sell_f32 = 0: FP64 standard form, 1: FP32 difference form, 2: FP16 scaled
 * difference form (FEM_FORM_SCALED_DIFF16) -- else the FP64 block CSR matrix
 * A.  ws >= fem_amg_lmax_ws_bytes(unknowns).  Synchronises the stream. */
int64_t fem_amg_lmax_ws_bytes(int64_t n);
int fem_amg_lmax(const fem_bcsr_t* A, const fem_pattern_t* pat, const void* sell_vals, int32_t sell_f32,
                 const double* dinv,
                 int32_t steps, double* lmax, void* ws, void* stream);
/* The same estimate written to device memory (lmax_dev[0]) without
 * synchronising: a hierarchy's setup reads all levels' estimates back once. */
int fem_amg_lmax_async(const fem_bcsr_t* A, const fem_pattern_t* pat, const void* sell_vals, int32_t sell_f32,
                       const double* dinv, int32_t steps, double* lmax_dev, void* ws, void* stream);
/* C = X Y for block CSR X (br x k) and Y (k x bc), FP64.  The structure is
 * computed once and kept as a plan, so a new matrix costs one numeric pass:
 *   symbolic  row_counts[r] = blocks of row r of C; row_pairs[r] (may be
 *             NULL) = block products of row r; *max_pairs (host) = their
 *             maximum (an input of the plan pass)
 *   plan      given C->row_ptr (scan of row_counts) and pair_ptr (scan of
 *             row_pairs): writes C->col_idx (ascending), plan_ptr (C->nnzb+1)
 *             and plan (2 int32 per product: X entry, Y entry), each output
 *             block's products in a fixed (sorted) order
 *   numeric   C->vals from X->vals and Y->vals along the plan (one thread per
 *             output block, products summed in plan order: bit-reproducible)
 * Rows of 2^18 products or more, of more than 2048 output blocks, or
 * left-operand rows of more than 2048 blocks return FEM_ERR_CAPACITY.  ws >= 256 bytes of device memory.
 * Symbolic and plan synchronise. */
int fem_bcsr_spgemm_symbolic(const fem_bcsr_t* X, const fem_bcsr_t* Y, int32_t* row_counts, int32_t* row_pairs,
                             int64_t* max_pairs, void* ws, void* host_status, void* stream);
int fem_bcsr_spgemm_plan(const fem_bcsr_t* X, const fem_bcsr_t* Y, int64_t max_pairs, fem_bcsr_t* C,
                         const int32_t* pair_ptr, int32_t* plan_ptr, int32_t* plan, void* ws, void* host_status,
                         void* stream);
int fem_bcsr_spgemm_numeric(const fem_bcsr_t* X, const fem_bcsr_t* Y, fem_bcsr_t* C, const int32_t* plan_ptr,
                            const int32_t* plan, void* stream);
/* Smoothed prolongator P = P0 - omega D^-1 A P0 (FP64): P0 is the tentative
 * prolongator as block CSR (one block per row, columns agg), P has the
 * structure and (plan_ptr, plan) the plan of the product A P0. */
int fem_amg_smooth_prolongator(const fem_bcsr_t* A, const fem_bcsr_t* P0, fem_bcsr_t* P, const int32_t* plan_ptr,
                               const int32_t* plan, const int32_t* agg, const double* dinv, double omega,
                               void* stream);
/* The same with omega = 1.5 / lmax_dev[0] formed on the device from the
 * level's fem_amg_lmax_async estimate. */
int fem_amg_smooth_prolongator_dev(const fem_bcsr_t* A, const fem_bcsr_t* P0, fem_bcsr_t* P,
                                   const int32_t* plan_ptr, const int32_t* plan, const int32_t* agg,
                                   const double* dinv, const double* lmax_dev, void* stream);
/* T = X^T (block CSR; T->rows = X->cols, blocks bc x br).  Structure:
 * T->row_ptr, T->col_idx (ascending) and src (nnzb: the X entry of each T
 * entry); ws >= 4*X->cols rounded up to 256, + fem_scan_ws_bytes(X->cols).
 * Values (FP64): T->vals from X->vals along src. */
int fem_bcsr_transpose_structure(const fem_bcsr_t* X, fem_bcsr_t* T, int32_t* src, void* ws, void* stream);
int fem_bcsr_transpose_values(const fem_bcsr_t* X, fem_bcsr_t* T, const int32_t* src, void* stream);
/* Dense inverse of a FP64 block CSR matrix (n = rows*br <= 4096) by
 * Gauss-Jordan without pivoting (SPD); ws >= 8*n*n bytes. */
int fem_amg_coarse_inverse(const fem_bcsr_t* A, double* inv, void* ws, void* stream);
/* FP16 scaled difference-form copy of a standard-form FP64 SELL matrix:
 * value (a, b) of block (i, j) becomes dinv[i*vec + a] * A_ij[a][b], the
 * diagonal slot the scaled row-sum block (FEM_FORM_SCALED_DIFF16). */
int fem_amg_sell_to_f16_scaled(const fem_pattern_t* pat, const double* sell_vals, const double* dinv, void* out,
                               void* stream);
/* out[i] = (float)in[i]. */
int fem_amg_to_f32(int64_t n, const double* in, float* out, void* stream);
/* z = M r, one V-cycle; ws >= fem_amg_ws_bytes(amg). */
int64_t fem_amg_ws_bytes(const fem_amg_t* amg);
int fem_amg_vcycle(const fem_amg_t* amg, const double* r, double* z, void* ws, void* stream);
/* AMG-preconditioned CG on the row-eliminated tangent (b, x FP64, x holds
 * x0).  The outer operator is the FP64 standard-form SELL matrix `vals`, or,
 * if vals32 != NULL, the FP32 difference-form copy vals32 with FP64 residual
 * replacement through `vals` (opt->reliable_delta; convergence is decided on
 * the FP64 true residual; result->replacements counts them).  Uses
 * ws >= fem_amg_solve_ws_bytes(pat, amg).  Synchronises.  Returns FEM_OK
 * when the solve ran; its outcome (FEM_OK / FEM_ERR_NOT_CONVERGED /
 * FEM_ERR_NAN) is result->status. */
int64_t fem_amg_solve_ws_bytes(const fem_pattern_t* pat, const fem_amg_t* amg);
int fem_amg_pcg_solve(const fem_pattern_t* pat, const double* vals, const float* vals32, const fem_amg_t* amg,
                      const double* b, double* x, const fem_amg_options_t* opt, void* ws, void* host_status,
                      fem_pcg_result_t* result /* host */, void* stream);

/* ---- vector helpers -------------------------------------------------- */
/* out = sum_i x_i y_i (FP64), deterministic; ws >= 8 * 512 bytes (the
 * residual norms of the Newton loop, solver.py:1319-1333). */
int fem_dot(int64_t n, const double* x, const double* y, double* out, void* ws, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FEM_MI355X_H */
