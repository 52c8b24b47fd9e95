"""Problem: the user-facing finite-element problem (jax_fem/problem.py).

Same dataclass fields and hooks as the reference (problem.py:15-54): a
subclass provides ``get_tensor_map`` and optionally ``get_mass_map``,
``get_surface_maps``, ``custom_init`` and ``set_params``; hooks are detected
with ``hasattr`` exactly as in problem.py:277-296, 311-323 and 340-346.

State lives on the device (``DeviceProblem``): the matrix pattern is built
once, per-cell kernels run on the whole mesh in one launch (no 20-way cell
batching, problem.py:359-396) and nothing of size C*(N*vec)^2 is copied to the
host.  The reference's COO arrays (``I``, ``J``, ``V``) are available as
on-demand exports for small meshes.
"""
from dataclasses import dataclass
from typing import Any

import numpy as onp

from .fe import FiniteElement
from .generate_mesh import Mesh


@dataclass
class Problem:
    mesh: Mesh
    vec: int
    dim: int
    ele_type: str = 'HEX8'
    quadrature_rule: Any = None
    quadrature_order: int = None
    dirichlet_bc_info: list = None
    location_fns: list = None
    additional_info: tuple = ()

    def __post_init__(self):
        if isinstance(self.mesh, list):
            if len(self.mesh) != 1:
                raise NotImplementedError("multi-variable problems are outside the device path")
            unwrap = lambda x: x[0] if isinstance(x, list) else x
            self.vec, self.ele_type = unwrap(self.vec), unwrap(self.ele_type)
            self.quadrature_rule, self.quadrature_order = unwrap(self.quadrature_rule), unwrap(self.quadrature_order)
            self.dirichlet_bc_info = unwrap(self.dirichlet_bc_info) if (
                isinstance(self.dirichlet_bc_info, list) and len(self.dirichlet_bc_info) == 1) else self.dirichlet_bc_info
            self.mesh = self.mesh[0]
        self.mesh = [self.mesh]
        self.vec = [self.vec]
        self.ele_type = [self.ele_type]
        self.quadrature_rule = [self.quadrature_rule]
        self.quadrature_order = [self.quadrature_order]
        self.dirichlet_bc_info = [self.dirichlet_bc_info]
        self.num_vars = 1
        self.fes = [FiniteElement(mesh=self.mesh[0], vec=self.vec[0], dim=self.dim, ele_type=self.ele_type[0],
                                  quadrature_rule=self.quadrature_rule[0],
                                  quadrature_order=self.quadrature_order[0],
                                  dirichlet_bc_info=self.dirichlet_bc_info[0])]
        fe = self.fes[0]
        self.cells_list = [fe.cells]
        self.num_cells = fe.num_cells
        self.boundary_inds_list = fe.get_boundary_conditions_inds(self.location_fns)
        self.offset = [0]
        self.num_total_dofs_all_vars = fe.num_total_dofs
        self.num_nodes_cumsum = onp.cumsum([0, fe.num_nodes])
        self.cells_list_face_list = [[fe.cells[bi[:, 0]]] for bi in self.boundary_inds_list]
        self.internal_vars = ()
        self.internal_vars_surfaces = [() for _ in range(len(self.boundary_inds_list))]
        self._dev = None
        self.custom_init(*self.additional_info)
        # surface-map count check of problem.py:340-346
        num_surfaces = len(self.boundary_inds_list)
        if hasattr(self, 'get_surface_maps'):
            assert num_surfaces == len(self.get_surface_maps())
        elif hasattr(self, 'get_universal_kernels_surface'):
            assert num_surfaces == len(self.get_universal_kernels_surface())
        else:
            assert num_surfaces == 0, "Missing definitions for surface integral"

    # ------------------------------------------------------------- hooks
    def custom_init(self, *args):
        """Child classes may override (problem.py:184-187)."""
        pass

    def set_params(self, params):
        """Child classes override to map ``params`` to ``internal_vars`` (problem.py:493-556)."""
        raise NotImplementedError("Child class must implement this function!")

    # ------------------------------------------------------------ device
    def device(self):
        """The problem's device state (pattern and buffers are built on first use)."""
        if self._dev is None:
            from ._device import DeviceProblem
            self._dev = DeviceProblem(self)
        return self._dev

    def unflatten_fn_sol_list(self, dofs):
        return [dofs.reshape(self.fes[0].num_total_nodes, self.fes[0].vec)]

    def unflatten_fn_dof(self, cell_sol_flat):
        """(num_nodes*vec,) cell dofs -> [(num_nodes, vec)] (problem.py:114)."""
        return [cell_sol_flat.reshape(self.fes[0].num_nodes, self.fes[0].vec)]

    # --------------------------------------------------- geometry fields
    @property
    def JxW(self):
        """(C, 1, Q) (problem.py:143)."""
        return self.fes[0].JxW[:, None, :]

    @property
    def shape_grads(self):
        return self.fes[0].shape_grads

    @property
    def v_grads_JxW(self):
        fe = self.fes[0]
        return fe.shape_grads[:, :, :, None, :] * fe.JxW[:, :, None, None, None]

    @property
    def physical_quad_points(self):
        return self.fes[0].get_physical_quad_points()

    @property
    def selected_face_shape_grads(self):
        return [self.fes[0].get_face_shape_grads(bi)[0] for bi in self.boundary_inds_list]

    @property
    def nanson_scale(self):
        return [self.fes[0].get_face_shape_grads(bi)[1][:, None, :] for bi in self.boundary_inds_list]

    @property
    def selected_face_shape_vals(self):
        fe = self.fes[0]
        return [fe.face_shape_vals[bi[:, 1]] for bi in self.boundary_inds_list]

    @property
    def physical_surface_quad_points(self):
        return [self.fes[0].get_physical_surface_quad_points(bi) for bi in self.boundary_inds_list]

    # ------------------------------------------------ residual / Newton
    def _flat(self, sol_list):
        import torch
        dev = self.device()
        s = sol_list[0] if isinstance(sol_list, (list, tuple)) else sol_list
        return torch.as_tensor(s).detach().to(device=dev.dm.device, dtype=torch.float64).reshape(-1).contiguous()

    def compute_residual(self, sol_list):
        """Residual list [(num_nodes, vec)] without Dirichlet rows (problem.py:462-475)."""
        dev = self.device()
        dev.prepare()
        res = dev.residual_unconstrained(self._flat(sol_list))
        return self.unflatten_fn_sol_list(res)

    def newton_update(self, sol_list):
        """Residual list; the FP64 tangent is assembled on the device (problem.py:477-491)."""
        dev = self.device()
        dev.prepare()
        u = self._flat(sol_list)
        from . import _lib
        # residual and FP64 tangent at u, in one pass where the single-pass assembly applies
        dev.tangent(u, _lib.FEM_F64, residual=True)
        self._last_u = u
        return self.unflatten_fn_sol_list(dev.last_residual)

    # ------------------------------------------- reference COO exports
    @property
    def I(self):
        """COO row indices of the reference, cells then faces (problem.py:95-107); small meshes only."""
        return onp.hstack([onp.repeat(x[:, :, None], x.shape[1], axis=2).reshape(-1) for x in self._coo_inds()])

    @property
    def J(self):
        return onp.hstack([onp.repeat(x[:, None, :], x.shape[1], axis=1).reshape(-1) for x in self._coo_inds()])

    @property
    def V(self):
        """Tangent values in COO order after newton_update (problem.py:453-458): the cell
        blocks (laplace, mass and universal cell kernels), then the face blocks of each
        surface (zero for a surface map that does not depend on u)."""
        u = getattr(self, '_last_u', None)
        if u is None:
            raise AttributeError("V is available after newton_update")
        dev = self.device()
        cells = dev.cell_tangents(u, faces=False).reshape(-1).cpu().numpy()
        faces = [f.reshape(-1).cpu().numpy() for f in dev.face_tangents(u)]
        return onp.hstack([cells] + faces)

    def _coo_inds(self):
        inds = self._cell_dof_inds()
        return [inds] + [inds[bi[:, 0]] for bi in self.boundary_inds_list]

    def _cell_dof_inds(self):
        fe = self.fes[0]
        return (fe.vec * fe.cells[:, :, None] + onp.arange(fe.vec)[None, None, :]).reshape(fe.num_cells, -1)
