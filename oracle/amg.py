"""CPU restatement of the device smoothed-aggregation AMG preconditioner.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): tests compare the device
hierarchy and V-cycle with these functions on the same inputs.

The reference reaches algebraic multigrid through AMGX (`amgx_solver`,
solver.py:141-259; AMGX's own classical AMG lives in the third-party pyamgx /
NVIDIA AMGX packages, absent here).  The device path offers the same option key
with a smoothed-aggregation hierarchy instead, so parity for this row is at the
solution level (the solve reaches the same tolerance as jax_solve,
solver.py:63-92) and every stage below is pinned by what the mathematics fixes:
  aggregation   every node in exactly one aggregate; roots pairwise at graph
                distance >= 3; every node within distance 2 of its root
  tentative P   orthonormal columns per aggregate (P0^T P0 = I on non-empty
                columns) and P0 Bc = B (the near null space is reproduced)
  Galerkin      Ac = P^T A P against a dense product on small cases
  V-cycle       symmetric and positive definite (<M x, y> = <x, M y>,
                <M x, x> > 0) so it is a valid CG preconditioner
The algorithm (not a restatement of any reference file):
  * MIS-2 roots by two max-propagations of (state, hash priority, index)
    tuples per round; aggregates = root + distance-1 nodes, distance-2 nodes
    join the neighbour aggregate of largest priority.
  * Near null space: rigid-body modes (vec 3, coordinates relative to the
    aggregate's first node) or constants (vec 1); rows of Dirichlet dofs zero.
    Tentative P by modified Gram-Schmidt per aggregate.
  * P = (I - omega D^-1 A) P0, omega = 1.5 / lambda, lambda from 15 power
    steps on D^-1 A; Ac = P^T A P; zero diagonal entries of Ac set to 1.
  * Coarsening stops at <= coarse_limit(nnz0) unknowns (dense inverse: 512,
    up to 1024 while n^3 <= 8 nnz0, nnz0 the level-0 operator's entries) or
    10 levels.
  * V(2,2) cycle (degree 5 on the coarser levels), Chebyshev smoothing on
    [lambda/15, 1.1 lambda].
"""
import numpy as np
import scipy.sparse as sp

COARSE_MAX = 1024
COARSE_SMALL = 512
MAX_LEVELS = 10


def coarse_limit(nnz0):
    """Largest dense coarsest level: 512 unknowns, and up to 1024 while n^3 <= 8 nnz0."""
    n = COARSE_SMALL
    while n < COARSE_MAX and (n + 1) ** 3 <= 8 * nnz0:
        n += 1
    return n
POWER_STEPS = 15
CHEB_LOW = 1.0 / 15.0
CHEB_HIGH = 1.1
RANK_TOL = 1e-10


def hash32(i):
    """32-bit integer mix (murmur3 finaliser) of node ids."""
    x = np.asarray(i, dtype=np.uint64) & np.uint64(0xFFFFFFFF)
    x = (x * np.uint64(0x9E3779B1)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x85EBCA6B)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(13)
    x = (x * np.uint64(0xC2B2AE35)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    return x


def priority_key(n):
    """61-bit unique keys: (hash >> 2) << 31 | index (index < 2^31)."""
    i = np.arange(n, dtype=np.uint64)
    return ((hash32(i) >> np.uint64(2)) << np.uint64(31)) | i


def _segmax(vals, row_ptr):
    n = len(row_ptr) - 1
    out = np.zeros(n, dtype=vals.dtype)
    nz = np.flatnonzero(np.diff(row_ptr) > 0)
    out[nz] = np.maximum.reduceat(vals, row_ptr[:-1][nz])
    return out


def mis2_aggregate(row_ptr, col_idx):
    """(agg, n_agg): aggregate id of every node of the graph (row_ptr, col_idx incl. self)."""
    row_ptr = np.asarray(row_ptr, np.int64)
    col_idx = np.asarray(col_idx, np.int64)
    n = len(row_ptr) - 1
    key = priority_key(n)
    state = np.zeros(n, np.int8)          # 0 undecided, 1 root, 2 not root
    rounds = 0
    while (state == 0).any():
        rank = np.where(state == 1, 2, np.where(state == 0, 1, 0)).astype(np.uint64)
        t = (rank << np.uint64(61)) | key
        t1 = _segmax(t[col_idx], row_ptr)
        t2 = _segmax(t1[col_idx], row_ptr)
        und = state == 0
        new_root = und & (t2 == t)
        near_root = und & ~new_root & ((t2 >> np.uint64(61)) == 2)
        state[new_root] = 1
        state[near_root] = 2
        rounds += 1
        assert rounds < 1000
    root = state == 1
    root_id = np.cumsum(root) - 1
    agg1 = np.where(root, root_id, -1)
    # distance 1: the (unique) root neighbour
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    nb_root = root[col_idx]
    agg1_full = agg1.copy()
    sel = nb_root & ~root[rows]
    agg1_full[rows[sel]] = root_id[col_idx[sel]]
    # distance 2: the assigned neighbour of largest priority key
    agg = agg1_full.copy()
    todo = agg1_full < 0
    if todo.any():
        cand = np.where(agg1_full[col_idx] >= 0, key[col_idx] + np.uint64(1), np.uint64(0))
        best = _segmax(cand, row_ptr)
        for r in np.flatnonzero(todo):
            seg = slice(row_ptr[r], row_ptr[r + 1])
            j = col_idx[seg][np.flatnonzero(cand[seg] == best[r])[0]]
            agg[r] = agg1_full[j]
    assert (agg >= 0).all()
    return agg.astype(np.int32), int(root.sum())


def members(agg, n_agg):
    """CSR of aggregates: agg_ptr (n_agg+1), member nodes ascending."""
    order = np.argsort(agg, kind='stable')
    counts = np.bincount(agg, minlength=n_agg)
    ptr = np.concatenate([[0], np.cumsum(counts)])
    return ptr.astype(np.int32), order.astype(np.int32)


def near_null_space(points, vec, bc_mask):
    """B (n*vec, NS): constants (vec 1) or rigid-body modes (vec 3), raw coordinates;
    Dirichlet rows zero.  The tentative step shifts rotations per aggregate."""
    n = points.shape[0]
    if vec == 1:
        B = np.ones((n, 1))
    else:
        x, y, z = points[:, 0], points[:, 1], points[:, 2]
        B = np.zeros((n * 3, 6))
        B[0::3, 0] = 1
        B[1::3, 1] = 1
        B[2::3, 2] = 1
        B[0::3, 3], B[1::3, 3] = -y, x
        B[1::3, 4], B[2::3, 4] = -z, y
        B[0::3, 5], B[2::3, 5] = z, -x
    B[np.asarray(bc_mask, bool)] = 0
    return B


def level0_null_space(points, vec, bc_mask, agg_ptr, memb):
    """Rigid-body modes with rotations about each aggregate's first node (Dirichlet rows zero)."""
    n = points.shape[0]
    if vec == 1:
        B = np.ones((n, 1))
    else:
        first = memb[agg_ptr[:-1]]
        owner = np.empty(n, np.int64)
        owner[memb] = np.repeat(np.arange(len(agg_ptr) - 1), np.diff(agg_ptr))
        rel = points - points[first[owner]]
        B = near_null_space(rel, 3, np.zeros(n * 3, bool))
    B[np.asarray(bc_mask, bool)] = 0
    return B


def tentative(agg_ptr, memb, B, bs):
    """Modified Gram-Schmidt per aggregate: P0 blocks (n, bs, NS) and Bc (n_agg*NS, NS)."""
    ns = B.shape[1]
    n = len(memb)
    n_agg = len(agg_ptr) - 1
    P0 = np.zeros((n, bs, ns))
    Bc = np.zeros((n_agg * ns, ns))
    for a in range(n_agg):
        nodes = memb[agg_ptr[a]:agg_ptr[a + 1]]
        dofs = (nodes[:, None] * bs + np.arange(bs)).ravel()
        V = B[dofs].copy()
        Q = np.zeros_like(V)
        R = np.zeros((ns, ns))
        for j in range(ns):
            v = V[:, j].copy()
            n0 = np.sqrt(v @ v)
            for k in range(j):
                R[k, j] = Q[:, k] @ v
                v -= R[k, j] * Q[:, k]
            nv = np.sqrt(v @ v)
            if nv > RANK_TOL * n0 and nv > 0:
                Q[:, j] = v / nv
                R[j, j] = nv
        P0[nodes] = Q.reshape(len(nodes), bs, ns)
        Bc[a * ns:(a + 1) * ns] = R
    return P0, Bc


def bcsr_to_scipy(row_ptr, col_idx, vals):
    """Block CSR (blocks br x bc) to a scipy CSR matrix."""
    vals = np.asarray(vals)
    nb, br, bc = vals.shape
    rows = len(row_ptr) - 1
    ncols = (int(np.max(col_idx)) + 1) if nb else 0
    return sp.bsr_matrix((vals, np.asarray(col_idx), np.asarray(row_ptr)),
                         shape=(rows * br, max(ncols, 1) * bc)).tocsr()


def p0_matrix(P0, agg, n_agg):
    n, bs, ns = P0.shape
    return sp.bsr_matrix((P0, agg.astype(np.int64), np.arange(n + 1)), shape=(n * bs, n_agg * ns)).tocsr()


def start_vector(n):
    """Deterministic power-iteration start: hash-derived values in [0.5, 1.5)."""
    return 0.5 + (hash32(np.arange(n)) >> np.uint64(8)).astype(np.float64) / 2.0 ** 24


def lmax(A, dinv, steps=POWER_STEPS):
    """Power iteration on D^-1 A: returns |D^-1 A x| for the normalised x of the last step."""
    x = start_vector(A.shape[0])
    x /= np.sqrt(x @ x)
    lam = 0.0
    for _ in range(steps):
        y = dinv * (A @ x)
        lam = np.sqrt(y @ y)
        x = y / lam
    return lam


def safe_dinv(A):
    d = A.diagonal()
    return np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 0.0)


def node_graph(ptr, idx, n_cols=None):
    """0/1 scipy CSR of a node graph (row_ptr, col_idx)."""
    n = len(ptr) - 1
    return sp.csr_matrix((np.ones(len(idx), np.int64), np.asarray(idx), np.asarray(ptr)),
                         shape=(n, n if n_cols is None else n_cols))


class Hierarchy:
    """Smoothed-aggregation hierarchy of the row-eliminated matrix A (scipy CSR,
    identity rows on Dirichlet dofs).  (graph_ptr, graph_idx) is the block
    pattern of A (the FE node graph, Dirichlet nodes included); the coarse
    graphs are the structural patterns of the Galerkin products, as the device
    builds them (blocks that cancel to zero stay in the pattern)."""

    def __init__(self, A, graph_ptr, graph_idx, points, vec, bc_mask, lams=None):
        self.levels = []
        bs = vec
        A = A.tocsr()
        ptr, idx = np.asarray(graph_ptr), np.asarray(graph_idx)
        n_nodes = len(ptr) - 1
        B = None
        lvl = 0
        # level 0's operator has the FE block pattern: vec^2 entries per node-graph entry
        n_dense = coarse_limit(len(idx) * vec * vec)
        while A.shape[0] > n_dense and len(self.levels) < MAX_LEVELS - 1:
            agg, n_agg = mis2_aggregate(ptr, idx)
            if n_agg >= n_nodes:
                break
            aptr, memb = members(agg, n_agg)
            if B is None:
                B = level0_null_space(points, vec, bc_mask, aptr, memb)
            P0, Bc = tentative(aptr, memb, B, bs)
            dinv = safe_dinv(A)
            lam = lams[lvl] if lams is not None else lmax(A, dinv)
            omega = 1.5 / lam
            P0m = p0_matrix(P0, agg, n_agg)
            P = (P0m - omega * (sp.diags(dinv) @ (A @ P0m))).tocsr()
            Ac = (P.T @ A @ P).tocsr()
            d = Ac.diagonal()
            zero = d == 0
            if zero.any():
                Ac = (Ac + sp.diags(zero.astype(float))).tocsr()
            GA = node_graph(ptr, idx)
            GP0 = sp.csr_matrix((np.ones(n_nodes, np.int64), (np.arange(n_nodes), agg)), shape=(n_nodes, n_agg))
            GP = (GA @ GP0).tocsr()
            GAc = (GP.T @ (GA @ GP)).tocsr()
            GAc.sort_indices()
            self.levels.append(dict(A=A, dinv=dinv, lam=lam, P=P, agg=agg, n_agg=n_agg, P0=P0, Bc=Bc, B=B,
                                    bs=bs, graph=(ptr, idx), p_graph=GP))
            ns = B.shape[1]
            ptr, idx = GAc.indptr, GAc.indices
            n_nodes = n_agg
            A, B, bs = Ac, Bc, ns
            lvl += 1
        self.coarse = A
        self.coarse_inv = np.linalg.inv(A.toarray())

    def smooth(self, L, x, rhs, degree, zero_start):
        A, d, lam = L['A'], L['dinv'], L['lam']
        hi, lo = CHEB_HIGH * lam, CHEB_LOW * lam
        theta, delta = (hi + lo) / 2, (hi - lo) / 2
        sigma = theta / delta
        rho = 1.0 / sigma
        r = d * rhs if zero_start else d * (rhs - A @ x)
        dd = r / theta
        for k in range(degree):
            x = x + dd
            if k == degree - 1:
                break
            r = d * (rhs - A @ x)
            rn = 1.0 / (2 * sigma - rho)
            dd = rn * rho * dd + (2 * rn / delta) * r
            rho = rn
        return x

    def vcycle(self, r, l=0, pre=2, post=2):
        """V-cycle; pre/post are level 0's Chebyshev degrees, the coarser levels smooth
        with 2 degree + 1 (at most 9), as the device cycle does."""
        if l == len(self.levels):
            return self.coarse_inv @ r
        L = self.levels[l]
        dpre, dpost = (pre, post) if l == 0 else (min(2 * pre + 1, 9), min(2 * post + 1, 9))
        x = self.smooth(L, np.zeros_like(r), r, dpre, True)
        rc = L['P'].T @ (r - L['A'] @ x)
        x = x + L['P'] @ self.vcycle(rc, l + 1, pre, post)
        return self.smooth(L, x, r, dpost, False)


def pcg(A, b, M, x0=None, rtol=1e-10, atol=0.0, maxiter=1000):
    """Preconditioned CG (returns x, iterations)."""
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - A @ x
    z = M(r)
    p = z.copy()
    rz = r @ z
    stop = max(rtol * np.linalg.norm(b), atol)
    k = 0
    while np.linalg.norm(r) > stop and k < maxiter:
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        z = M(r)
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
        k += 1
    return x, k
