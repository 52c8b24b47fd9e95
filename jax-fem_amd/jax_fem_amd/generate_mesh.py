"""Structured hexahedral meshes and the ``Mesh`` container.

Counterpart of ``jax_fem/generate_mesh.py``: ``Mesh`` (13-55), ``box_mesh``
(151-189, node/cell numbering reproduced exactly), ``get_meshio_cell_type``
(83-117). meshio is not used: generators return a small object with the two
attributes callers read from a ``meshio.Mesh`` (``points`` and ``cells_dict``).
``box_mesh_hex27`` is new: the reference only reaches HEX27 through gmsh.
"""
import numpy as onp

from .basis import get_elements, get_face_shape_vals_and_grads, _HEX27_NODES


class MeshioLike:
    """The two fields of ``meshio.Mesh`` that jax-fem callers use."""

    def __init__(self, points, cells_dict):
        self.points = points
        self.cells_dict = cells_dict


class Mesh:
    """Mesh manager: ``points`` (num_total_nodes, dim), ``cells`` (num_cells, num_nodes).

    Counterpart of generate_mesh.py:13-27. Arrays are kept as host numpy
    arrays (float64 points, int64 cells); the device copies live in the
    problem's discretisation.
    """

    def __init__(self, points, cells, ele_type=None):
        self.points = onp.asarray(points, dtype=onp.float64)
        self.cells = onp.asarray(cells, dtype=onp.int64)
        self.ele_type = ele_type

    def count_selected_faces(self, location_fn):
        """Number of faces whose vertices all satisfy ``location_fn`` (generate_mesh.py:29-55)."""
        from .fe import evaluate_location_fn
        ele_type = self.ele_type or ('HEX8' if self.cells.shape[1] == 8 else 'HEX27')
        _, _, _, _, face_inds = get_face_shape_vals_and_grads(ele_type)
        face_nodes = self.cells[:, face_inds]  # (C, 6, 4)
        flags = evaluate_location_fn(location_fn, self.points, onp.arange(len(self.points)))
        return int(onp.count_nonzero(onp.all(flags[face_nodes], axis=-1)))


def get_meshio_cell_type(ele_type):
    """meshio cell-type string for ``ele_type`` (generate_mesh.py:83-117)."""
    table = {'TET4': 'tetra', 'TET10': 'tetra10', 'HEX8': 'hexahedron',
             'HEX27': 'hexahedron27', 'HEX20': 'hexahedron20', 'TRI3': 'triangle',
             'TRI6': 'triangle6', 'QUAD4': 'quad', 'QUAD8': 'quad8'}
    if ele_type not in table:
        raise NotImplementedError
    return table[ele_type]


def box_mesh(Nx, Ny, Nz, domain_x, domain_y, domain_z):
    """HEX8 box with Nx*Ny*Nz cells (generate_mesh.py:151-189).

    Node id = (i*(Ny+1) + j)*(Nz+1) + k for lattice point (i, j, k); cells in
    VTK vertex order, cell id = (i*Ny + j)*Nz + k.
    """
    x = onp.linspace(0, domain_x, Nx + 1)
    y = onp.linspace(0, domain_y, Ny + 1)
    z = onp.linspace(0, domain_z, Nz + 1)
    xv, yv, zv = onp.meshgrid(x, y, z, indexing='ij')
    points = onp.stack((xv, yv, zv), axis=3).reshape(-1, 3)
    ids = onp.arange(len(points)).reshape(Nx + 1, Ny + 1, Nz + 1)
    corners = [ids[:-1, :-1, :-1], ids[1:, :-1, :-1], ids[1:, 1:, :-1], ids[:-1, 1:, :-1],
               ids[:-1, :-1, 1:], ids[1:, :-1, 1:], ids[1:, 1:, 1:], ids[:-1, 1:, 1:]]
    cells = onp.stack(corners, axis=3).reshape(-1, 8)
    return MeshioLike(points, {'hexahedron': cells})


def box_mesh_hex27(Nx, Ny, Nz, domain_x, domain_y, domain_z):
    """HEX27 box with Nx*Ny*Nz cells on a (2Nx+1)(2Ny+1)(2Nz+1) node lattice.

    Lattice numbering and cell order follow ``box_mesh``; the 27 nodes of each
    cell are in meshio ``hexahedron27`` order.
    """
    mx, my, mz = 2 * Nx + 1, 2 * Ny + 1, 2 * Nz + 1
    x = onp.linspace(0, domain_x, mx)
    y = onp.linspace(0, domain_y, my)
    z = onp.linspace(0, domain_z, mz)
    xv, yv, zv = onp.meshgrid(x, y, z, indexing='ij')
    points = onp.stack((xv, yv, zv), axis=3).reshape(-1, 3)
    I, J, K = onp.meshgrid(onp.arange(Nx), onp.arange(Ny), onp.arange(Nz), indexing='ij')
    I, J, K = I.reshape(-1, 1), J.reshape(-1, 1), K.reshape(-1, 1)
    lx, ly, lz = _HEX27_NODES[:, 0], _HEX27_NODES[:, 1], _HEX27_NODES[:, 2]
    cells = ((2 * I + lx) * my + (2 * J + ly)) * mz + (2 * K + lz)
    return MeshioLike(points, {'hexahedron27': cells.astype(onp.int64)})
