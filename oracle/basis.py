"""Reference-element tables, restated for the oracle.

Follows jax_fem/basis.py (get_elements 19-114, get_shape_vals_and_grads
141-175, get_face_shape_vals_and_grads 178-250) with basix replaced by a
generic construction: shape functions are obtained by inverting the
Vandermonde matrix of the tensor monomial basis at the nodes (no closed-form
Lagrange factors), quadrature is numpy's Gauss-Legendre mapped to [0, 1].
Conventions (node order = meshio/VTK, quadrature order = basix tensor order
with the first coordinate slowest, basix facet order) are those of the
reference; they are data, written out below.
"""
import numpy as np

# meshio node positions on [0,1]^3
HEX8_NODES = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                       [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float)
_h = 0.5
HEX27_NODES = np.array([
    [0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1],
    [_h, 0, 0], [1, _h, 0], [_h, 1, 0], [0, _h, 0], [_h, 0, 1], [1, _h, 1], [_h, 1, 1], [0, _h, 1],
    [0, 0, _h], [1, 0, _h], [1, 1, _h], [0, 1, _h],
    [0, _h, _h], [1, _h, _h], [_h, 0, _h], [_h, 1, _h], [_h, _h, 0], [_h, _h, 1], [_h, _h, _h]])

# basix hexahedron: vertices, facets (basix vertex ids), outward normals
BASIX_VERTS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0],
                        [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]], dtype=float)
BASIX_FACETS = [[0, 1, 2, 3], [0, 1, 4, 5], [0, 2, 4, 6], [1, 3, 5, 7], [2, 3, 6, 7], [4, 5, 6, 7]]
FACET_NORMALS = np.array([[0, 0, -1], [0, -1, 0], [-1, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
# re_order[:8] of basis.py:53: meshio vertex j = basix vertex RE_ORDER8[j]
RE_ORDER8 = [0, 1, 3, 2, 4, 5, 7, 6]

DEFAULT_ORDER = {'HEX8': 2, 'HEX27': 10}
DEGREE = {'HEX8': 1, 'HEX27': 2}


def nodes(ele_type):
    return {'HEX8': HEX8_NODES, 'HEX27': HEX27_NODES}[ele_type]


def _monomials(p, x):
    """Tensor monomials x^i y^j z^k (i,j,k <= p) and their gradients at points x (P,3)."""
    exps = [(i, j, k) for i in range(p + 1) for j in range(p + 1) for k in range(p + 1)]
    P = x.shape[0]
    val = np.empty((P, len(exps)))
    grad = np.empty((P, len(exps), 3))
    for m, (i, j, k) in enumerate(exps):
        xi, yj, zk = x[:, 0] ** i, x[:, 1] ** j, x[:, 2] ** k
        val[:, m] = xi * yj * zk
        grad[:, m, 0] = (i * x[:, 0] ** max(i - 1, 0) if i else 0.) * yj * zk
        grad[:, m, 1] = xi * (j * x[:, 1] ** max(j - 1, 0) if j else 0.) * zk
        grad[:, m, 2] = xi * yj * (k * x[:, 2] ** max(k - 1, 0) if k else 0.)
    return val, grad


def tabulate(ele_type, pts):
    """values (P,N), reference gradients (P,N,3) of the Lagrange basis in meshio order."""
    p = DEGREE[ele_type]
    X = nodes(ele_type)
    Vn, _ = _monomials(p, X)            # (N, M), square
    coef = np.linalg.inv(Vn)            # columns: coefficients of each nodal function
    v, g = _monomials(p, np.asarray(pts, float).reshape(-1, 3))
    return v @ coef, np.einsum('pmd,mn->pnd', g, coef)


def gauss01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return (x + 1) / 2, w / 2


def quad_rule(order, dim):
    n = (order + 2) // 2
    t, w = gauss01(n)
    grids = np.meshgrid(*([t] * dim), indexing='ij')
    wg = np.meshgrid(*([w] * dim), indexing='ij')
    pts = np.stack([g.reshape(-1) for g in grids], axis=1)
    return pts, np.prod(np.stack([g.reshape(-1) for g in wg], axis=1), axis=1)


def shape_vals_and_grads(ele_type, quadrature_order=None):
    """(Q,N) values, (Q,N,3) reference gradients, (Q,) weights (basis.py:141-175)."""
    order = DEFAULT_ORDER[ele_type] if quadrature_order is None else quadrature_order
    pts, w = quad_rule(order, 3)
    v, g = tabulate(ele_type, pts)
    return v, g, w


def face_shape_vals_and_grads(ele_type, quadrature_order=None):
    """Face tables (basis.py:178-250): vals (6,FQ,N), grads (6,FQ,N,3), weights (6,FQ),
    normals (6,3), face vertex ids in meshio numbering (6,4)."""
    order = DEFAULT_ORDER[ele_type] if quadrature_order is None else quadrature_order
    pts2, w2 = quad_rule(order, 2)
    s, t = pts2[:, 0], pts2[:, 1]
    phi = np.stack([(1 - s) * (1 - t), s * (1 - t), (1 - s) * t, s * t], axis=1)
    fpts, fw, finds = [], [], []
    for facet in BASIX_FACETS:
        V = BASIX_VERTS[facet]
        fpts.append(phi @ V)
        fw.append(w2 * np.linalg.norm(np.cross(V[1] - V[0], V[2] - V[0])))
        finds.append([RE_ORDER8.index(b) for b in facet])
    fpts = np.stack(fpts)
    v, g = tabulate(ele_type, fpts.reshape(-1, 3))
    nf, fq = fpts.shape[:2]
    return (v.reshape(nf, fq, -1), g.reshape(nf, fq, -1, 3), np.stack(fw), FACET_NORMALS.copy(),
            np.array(finds))
