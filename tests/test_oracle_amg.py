"""Pins of the CPU smoothed-aggregation restatement (oracle/amg.py) by what the
mathematics fixes: the MIS-2 structure of the aggregates, orthonormal
tentative prolongators that reproduce the near null space, a symmetric
positive definite V-cycle, and CG with it reaching the direct solution."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import pytest

from oracle import amg as OA
from oracle import fem as OF


def _cantilever(nx, vec=3):
    pts, cells = OF.box_mesh(nx, nx // 2, nx // 2, 2., 1., 1.)
    left = lambda p: p[0] < 1e-9
    if vec == 3:
        mat = OF.Material('elastic', **dict(zip(('lmbda', 'mu'), OF.lame(70e3, 0.3))))
        o = OF.Problem(pts, cells, 3, mat=mat, bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.] * 3])
        bi = OF.boundary_inds(o.points, o.cells, o.face_inds, lambda p: p[0] > 2 - 1e-9)
        o.add_surface(bi, np.broadcast_to(np.array([0., 0., 1.]), (len(bi), o.fw.shape[1], 3)))
    else:
        o = OF.Problem(pts, cells, 1, mat=OF.Material('poisson', k=1.0), bc_info=[[left], [0], [lambda p: 0.]])
        o.add_mass(np.ones((len(cells), o.sv.shape[0], 1)))
    A = o.tangent(np.zeros(o.n)).tocsr()
    b = -o.residual(np.zeros(o.n))
    mask = np.zeros(o.n, bool)
    mask[o.bc_dofs] = True
    # FE node graph (all nodes, Dirichlet nodes included)
    ones = np.ones((len(cells), 8, 8))
    G = sp.csr_matrix((ones.ravel(), (np.repeat(cells, 8, axis=1).ravel(), np.tile(cells, (1, 8)).ravel())),
                      shape=(len(pts), len(pts)))
    G.sort_indices()
    return o, A, b, mask, G


@pytest.fixture(scope='module')
def elastic():
    return _cantilever(12)


def _distances_from(G, src, limit):
    dist = {int(src): 0}
    frontier = [int(src)]
    for d in range(1, limit + 1):
        nxt = []
        for i in frontier:
            for j in G.indices[G.indptr[i]:G.indptr[i + 1]]:
                if int(j) not in dist:
                    dist[int(j)] = d
                    nxt.append(int(j))
        frontier = nxt
    return dist


def test_mis2_aggregates_structure(elastic):
    _, _, _, _, G = elastic
    agg, n_agg = OA.mis2_aggregate(G.indptr, G.indices)
    assert agg.min() == 0 and agg.max() == n_agg - 1 and len(np.unique(agg)) == n_agg
    key = OA.priority_key(G.shape[0])
    for a in range(n_agg):
        nodes = np.flatnonzero(agg == a)
        # the root is the first node whose whole neighbourhood is in the aggregate and is of largest key
        d = None
        for r in nodes:
            dist = _distances_from(G, r, 2)
            if all(int(i) in dist for i in nodes):
                d = dist
                break
        assert d is not None, f"aggregate {a} is not within distance 2 of any member"
    # roots (aggregate ids follow root order) are pairwise >= 3 apart
    ptr, idx = G.indptr, G.indices
    GG = (G @ G).tocsr()
    roots = []
    for a in range(n_agg):
        nodes = np.flatnonzero(agg == a)
        # a root has every neighbour in its own aggregate
        cand = [r for r in nodes if np.all(agg[idx[ptr[r]:ptr[r + 1]]] == a)]
        assert cand
        roots.append(max(cand, key=lambda r: key[r]))
    roots = np.array(roots)
    sub = GG[roots][:, roots].toarray()
    np.fill_diagonal(sub, 0)
    assert not sub.any(), "two roots within distance 2"


def test_tentative_orthonormal_and_reproduces_null_space(elastic):
    o, A, _, mask, G = elastic
    agg, n_agg = OA.mis2_aggregate(G.indptr, G.indices)
    aptr, memb = OA.members(agg, n_agg)
    B = OA.level0_null_space(o.points, 3, mask, aptr, memb)
    P0, Bc = OA.tentative(aptr, memb, B, 3)
    P0m = OA.p0_matrix(P0, agg, n_agg)
    # P0 Bc = B: the near null space is in the range of P0
    np.testing.assert_allclose(P0m @ Bc, B, atol=1e-12)
    # orthonormal columns (aggregates touching the clamp may lose modes: zero columns)
    Gm = (P0m.T @ P0m).toarray()
    d = np.diag(Gm)
    assert set(np.round(d, 12)) <= {0.0, 1.0}
    np.testing.assert_allclose(Gm - np.diag(d), 0, atol=1e-12)
    # rigid-body modes of the free structure are in the null space of the unconstrained operator
    K = o.tangent(np.zeros(o.n), bc=False)
    Bfull = OA.near_null_space(o.points, 3, np.zeros(o.n, bool))
    assert np.abs(K @ Bfull).max() < 1e-8 * np.abs(K).max()


def test_vcycle_symmetric_positive_definite(elastic):
    o, A, b, mask, G = elastic
    H = OA.Hierarchy(A, G.indptr, G.indices, o.points, 3, mask)
    assert len(H.levels) >= 1 and H.coarse.shape[0] <= OA.COARSE_MAX
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal((2, o.n))
    x[mask] = 0
    y[mask] = 0
    Mx, My = H.vcycle(x), H.vcycle(y)
    assert abs(Mx @ y - x @ My) <= 1e-10 * np.sqrt((Mx @ Mx) * (y @ y))
    assert Mx @ x > 0 and My @ y > 0
    assert np.abs(Mx[mask]).max() == 0.0   # Dirichlet components stay zero


@pytest.mark.parametrize('vec', [3, 1])
def test_amg_cg_matches_direct_solve(vec):
    o, A, b, mask, G = _cantilever(12, vec)
    H = OA.Hierarchy(A, G.indptr, G.indices, o.points, vec, mask)
    x, k = OA.pcg(A, b, H.vcycle, rtol=1e-12)
    ref = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - ref) <= 1e-9 * np.linalg.norm(ref)
    d = A.diagonal()
    _, kj = OA.pcg(A, b, lambda r: r / d, rtol=1e-12)
    # elasticity: far fewer iterations than Jacobi; the scalar problem on this small
    # grid is cheap for Jacobi too, there the bound is the size-independent AMG rate
    assert (k < kj / 2) if vec == 3 else (k <= 30), (k, kj)


def test_galerkin_coarse_operator_is_p_transpose_a_p(elastic):
    o, A, _, mask, G = elastic
    H = OA.Hierarchy(A, G.indptr, G.indices, o.points, 3, mask)
    L0 = H.levels[0]
    P = L0['P'].toarray()
    Ac = P.T @ A.toarray() @ P
    Ac[np.diag_indices_from(Ac)] += (np.diag(Ac) == 0)
    nxt = H.levels[1]['A'] if len(H.levels) > 1 else H.coarse
    np.testing.assert_allclose(nxt.toarray(), Ac, atol=1e-9 * np.abs(Ac).max())
    # the structural coarse graph covers every nonzero block of Ac
    gptr, gidx = (H.levels[1]['graph'] if len(H.levels) > 1 else (None, None))
    if gptr is not None:
        S = OA.node_graph(gptr, gidx).toarray() > 0
        nz = np.abs(Ac).reshape(S.shape[0], 6, S.shape[0], 6).max(axis=(1, 3)) > 0
        assert not (nz & ~S).any()
