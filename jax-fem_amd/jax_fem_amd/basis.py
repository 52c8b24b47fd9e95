"""Reference-element tables for the hexahedral elements on the device path.

Mirrors ``jax_fem/basis.py`` (get_elements 19-114, get_shape_vals_and_grads
141-175, get_face_shape_vals_and_grads 178-250) without basix: the tables are
built from closed-form tensor-product Lagrange polynomials and Gauss-Legendre
rules on the reference cube [0, 1]^3, using basix's conventions:

* quadrature: Gauss-Jacobi(0) = Gauss-Legendre with ``np = (order + 2) // 2``
  points per direction, tensor product with the first coordinate varying
  slowest, 1-D points ascending;
* nodes are in meshio/VTK order (the order ``re_order`` maps onto, basis.py:53
  and basis.py:60-61);
* faces are basix's hexahedron facets (z=0, y=0, x=0, x=1, y=1, z=1) with
  outward normals, quadrature points mapped through the bilinear facet map.

Only HEX8 and HEX27 are supported; the other element types of the reference
raise ``NotImplementedError`` (as basis.py:112 does for unknown types).
"""
import numpy as onp

# Default quadrature orders of basis.py:56 and basis.py:64.
_DEFAULT_ORDER = {'HEX8': 2, 'HEX27': 10}
_DEGREE = {'HEX8': 1, 'HEX27': 2}

# Reference coordinates of the cell nodes in meshio/VTK order, as multiples of
# 1/degree along each axis.
_HEX8_NODES = onp.array([
    [0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=onp.int64)

# HEX27 (meshio ``hexahedron27``): 8 vertices, 12 edge midpoints in VTK edge
# order, 6 face centres ordered x=0, x=1, y=0, y=1, z=0, z=1, then the centre.
# This is the geometric reading of the reference's re_order (basis.py:60-61)
# against basix's sub-entity numbering of the degree-2 hexahedron.
_HEX27_NODES = onp.array([
    [0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0],
    [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2],
    [1, 0, 0], [2, 1, 0], [1, 2, 0], [0, 1, 0],
    [1, 0, 2], [2, 1, 2], [1, 2, 2], [0, 1, 2],
    [0, 0, 1], [2, 0, 1], [2, 2, 1], [0, 2, 1],
    [0, 1, 1], [2, 1, 1], [1, 0, 1], [1, 2, 1],
    [1, 1, 0], [1, 1, 2], [1, 1, 1]], dtype=onp.int64)

# basix hexahedron facets in basix vertex numbering, then the outward normals.
_BASIX_HEX_VERTS = onp.array([
    [0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0],
    [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]], dtype=onp.float64)
_BASIX_HEX_FACETS = onp.array([
    [0, 1, 2, 3], [0, 1, 4, 5], [0, 2, 4, 6],
    [1, 3, 5, 7], [2, 3, 6, 7], [4, 5, 6, 7]], dtype=onp.int64)
_HEX_FACET_NORMALS = onp.array([
    [0., 0., -1.], [0., -1., 0.], [-1., 0., 0.],
    [1., 0., 0.], [0., 1., 0.], [0., 0., 1.]])
# basix vertex b sits at meshio vertex _BASIX_TO_MESHIO[b] (inverse of re_order[:8]).
_BASIX_TO_MESHIO = onp.array([0, 1, 3, 2, 4, 5, 7, 6], dtype=onp.int64)


def get_elements(ele_type):
    """(degree, default quadrature order, node lattice) for ``ele_type``.

    Counterpart of basis.py:19-114 restricted to the hexahedra.
    """
    if ele_type == 'HEX8':
        return 1, _DEFAULT_ORDER['HEX8'], _HEX8_NODES
    if ele_type == 'HEX27':
        return 2, _DEFAULT_ORDER['HEX27'], _HEX27_NODES
    raise NotImplementedError(
        f"ele_type {ele_type!r} is not on the device path (HEX8 and HEX27 only)")


def gauss_legendre_01(npts):
    """Gauss-Legendre points (ascending) and weights on [0, 1]."""
    x, w = onp.polynomial.legendre.leggauss(npts)
    return 0.5 * (x + 1.), 0.5 * w


def points_per_axis(quadrature_order):
    """basix Gauss-Jacobi: (m + 2) // 2 points per direction for degree m."""
    return (int(quadrature_order) + 2) // 2


def _lagrange_1d(degree, t):
    """1-D Lagrange values and derivatives on equispaced nodes of [0, 1].

    Returns arrays of shape (len(t), degree + 1) for node index 0..degree.
    """
    t = onp.asarray(t, dtype=onp.float64)
    if degree == 1:
        val = onp.stack([1. - t, t], axis=-1)
        der = onp.stack([-onp.ones_like(t), onp.ones_like(t)], axis=-1)
    elif degree == 2:
        # nodes 0, 1/2, 1
        val = onp.stack([(2. * t - 1.) * (t - 1.), 4. * t * (1. - t), t * (2. * t - 1.)], axis=-1)
        der = onp.stack([4. * t - 3., 4. - 8. * t, 4. * t - 1.], axis=-1)
    else:
        raise NotImplementedError(degree)
    return val, der


def tabulate(ele_type, pts):
    """Shape values (P, N) and reference gradients (P, N, 3) at points (P, 3)."""
    degree, _, lattice = get_elements(ele_type)
    pts = onp.asarray(pts, dtype=onp.float64).reshape(-1, 3)
    v = [None] * 3
    d = [None] * 3
    for ax in range(3):
        v[ax], d[ax] = _lagrange_1d(degree, pts[:, ax])
    ix, iy, iz = lattice[:, 0], lattice[:, 1], lattice[:, 2]
    vx, vy, vz = v[0][:, ix], v[1][:, iy], v[2][:, iz]
    dx, dy, dz = d[0][:, ix], d[1][:, iy], d[2][:, iz]
    vals = vx * vy * vz
    grads = onp.stack([dx * vy * vz, vx * dy * vz, vx * vy * dz], axis=-1)
    return vals, grads


def volume_quadrature(quadrature_order):
    """Tensor Gauss rule on [0,1]^3, first coordinate slowest (basix order)."""
    t, w = gauss_legendre_01(points_per_axis(quadrature_order))
    X, Y, Z = onp.meshgrid(t, t, t, indexing='ij')
    WX, WY, WZ = onp.meshgrid(w, w, w, indexing='ij')
    pts = onp.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=-1)
    return pts, (WX * WY * WZ).ravel()


def face_quadrature(quadrature_order):
    """Tensor Gauss rule on the unit square, first coordinate slowest."""
    t, w = gauss_legendre_01(points_per_axis(quadrature_order))
    S, T = onp.meshgrid(t, t, indexing='ij')
    WS, WT = onp.meshgrid(w, w, indexing='ij')
    return onp.stack([S.ravel(), T.ravel()], axis=-1), (WS * WT).ravel()


def get_shape_vals_and_grads(ele_type, quadrature_rule=None, quadrature_order=None):
    """Shape values (Q, N), reference gradients (Q, N, 3), weights (Q,).

    Counterpart of basis.py:141-175.
    """
    _check_rule(quadrature_rule)
    _, default_order, _ = get_elements(ele_type)
    order = default_order if quadrature_order is None else quadrature_order
    pts, w = volume_quadrature(order)
    vals, grads = tabulate(ele_type, pts)
    return vals, grads, w


def get_face_shape_vals_and_grads(ele_type, quadrature_rule=None, quadrature_order=None):
    """Face tables, counterpart of basis.py:178-250.

    Returns face_shape_vals (6, FQ, N), face_shape_grads_ref (6, FQ, N, 3),
    face_weights (6, FQ), face_normals (6, 3), face_inds (6, 4) (meshio vertex
    numbers of each facet).
    """
    _check_rule(quadrature_rule)
    _, default_order, _ = get_elements(ele_type)
    order = default_order if quadrature_order is None else quadrature_order
    st, w2 = face_quadrature(order)
    s, t = st[:, 0], st[:, 1]
    # bilinear facet map with basix quadrilateral vertex order
    phi = onp.stack([(1. - s) * (1. - t), s * (1. - t), (1. - s) * t, s * t], axis=-1)
    face_pts, face_w = [], []
    for facet in _BASIX_HEX_FACETS:
        verts = _BASIX_HEX_VERTS[facet]
        face_pts.append(phi @ verts)
        e1, e2 = verts[1] - verts[0], verts[2] - verts[0]
        face_w.append(w2 * onp.linalg.norm(onp.cross(e1, e2)))
    face_pts = onp.stack(face_pts)
    nf, fq, _ = face_pts.shape
    vals, grads = tabulate(ele_type, face_pts.reshape(-1, 3))
    face_vals = vals.reshape(nf, fq, -1)
    face_grads = grads.reshape(nf, fq, -1, 3)
    face_inds = _BASIX_TO_MESHIO[_BASIX_HEX_FACETS]
    return face_vals, face_grads, onp.stack(face_w), _HEX_FACET_NORMALS.copy(), face_inds


def _check_rule(rule):
    if rule not in (None, 'default', 'gauss_jacobi', 'GJ'):
        raise NotImplementedError(
            f"quadrature_rule {rule!r}: only the default Gauss-Jacobi rule is supported")
