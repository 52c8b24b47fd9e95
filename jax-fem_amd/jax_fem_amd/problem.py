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
        if isinstance(self.mesh, list) and len(self.mesh) > 1:
            self._init_variables()
        else:
            if isinstance(self.mesh, list):
                unwrap = lambda x: x[0] if isinstance(x, list) else x
                self.vec, self.ele_type = unwrap(self.vec), unwrap(self.ele_type)
                self.quadrature_rule, self.quadrature_order = unwrap(self.quadrature_rule), unwrap(self.quadrature_order)
                self.dirichlet_bc_info = unwrap(self.dirichlet_bc_info) if (
                    isinstance(self.dirichlet_bc_info, list) and len(self.dirichlet_bc_info) == 1) \
                    else self.dirichlet_bc_info
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
            self._fe_device = self.fes[0]
            self.offset = [0]
            self.num_total_dofs_all_vars = self.fes[0].num_total_dofs
        self._finish_init()

    def _init_variables(self):
        """Several variables (problem.py:58-119): lists of meshes, vecs, element types, BC infos.

        The device path takes variables that share one mesh and one element: their dofs
        form one node block of size sum(vec) inside the device (node-major, the variables'
        components side by side), while every array the user sees keeps the reference's
        layout: variable i at offset[i] + vec_i * node + comp (problem.py:82-90).
        """
        n = len(self.mesh)
        per_var = lambda x: list(x) if isinstance(x, list) else [x] * n
        self.vec, self.ele_type = per_var(self.vec), per_var(self.ele_type)
        self.quadrature_rule, self.quadrature_order = per_var(self.quadrature_rule), per_var(self.quadrature_order)
        self.dirichlet_bc_info = per_var(self.dirichlet_bc_info)
        self.num_vars = n
        self.fes = [FiniteElement(mesh=self.mesh[i], vec=self.vec[i], dim=self.dim, ele_type=self.ele_type[i],
                                  quadrature_rule=self.quadrature_rule[i], quadrature_order=self.quadrature_order[i],
                                  dirichlet_bc_info=self.dirichlet_bc_info[i]) for i in range(n)]
        f0 = self.fes[0]
        for fe in self.fes[1:]:
            if (fe.ele_type != f0.ele_type or fe.num_quads != f0.num_quads or
                    not onp.array_equal(fe.cells, f0.cells) or not onp.array_equal(fe.points, f0.points)):
                raise NotImplementedError("variables on different meshes or elements (e.g. Taylor-Hood pairs) are "
                                          "outside the device path: the variables must share one mesh and element")
        if sum(self.vec) > 4:
            raise NotImplementedError(f"the variables' components add up to {sum(self.vec)}; the device path "
                                      "takes at most 4 per node")
        for hook in ('get_tensor_map', 'get_mass_map', 'get_surface_maps'):
            if hasattr(self, hook):
                raise NotImplementedError(f"several variables on the device path take their weak form from "
                                          f"get_universal_kernel(s_surface), not {hook}")
        self._fe_device = FiniteElement(mesh=self.mesh[0], vec=sum(self.vec), dim=self.dim,
                                        ele_type=self.ele_type[0], quadrature_rule=self.quadrature_rule[0],
                                        quadrature_order=self.quadrature_order[0], dirichlet_bc_info=None)
        Nn = f0.num_total_nodes
        self.offset = list(onp.cumsum([0] + [Nn * v for v in self.vec[:-1]]))
        self.num_total_dofs_all_vars = Nn * sum(self.vec)

    def _finish_init(self):
        fe = self.fes[0]
        self.cells_list = [f.cells for f in self.fes]
        self.num_cells = fe.num_cells
        self.boundary_inds_list = fe.get_boundary_conditions_inds(self.location_fns)
        self.num_nodes_cumsum = onp.cumsum([0] + [f.num_nodes for f in self.fes])
        self.cells_list_face_list = [[f.cells[bi[:, 0]] for f in self.fes] for bi in self.boundary_inds_list]
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
        """Device dof vector -> [(num_nodes, vec_i)] per variable (problem.py:116-117)."""
        if self.num_vars == 1:
            return [dofs.reshape(self.fes[0].num_total_nodes, self.fes[0].vec)]
        blk = dofs.reshape(self.fes[0].num_total_nodes, sum(self.vec))
        offs = onp.cumsum([0] + self.vec)
        return [blk[:, offs[i]:offs[i + 1]].contiguous() for i in range(self.num_vars)]

    def unflatten_fn_dof(self, cell_sol_flat):
        """(num_nodes*vec + ...,) cell dofs -> [(num_nodes, vec_i)] (problem.py:114)."""
        out, k = [], 0
        for fe in self.fes:
            out.append(cell_sol_flat[k:k + fe.num_nodes * fe.vec].reshape(fe.num_nodes, fe.vec))
            k += fe.num_nodes * fe.vec
        return out

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
        """[(num_nodes, vec_i)] -> the device dof vector (node blocks of sum(vec) values)."""
        import torch
        dev = self.device()
        as_dev = lambda s: torch.as_tensor(s).detach().to(device=dev.dm.device, dtype=torch.float64)
        if self.num_vars == 1:
            s = sol_list[0] if isinstance(sol_list, (list, tuple)) else sol_list
            return as_dev(s).reshape(-1).contiguous()
        Nn = self.fes[0].num_total_nodes
        return torch.cat([as_dev(s).reshape(Nn, v) for s, v in zip(sol_list, self.vec)], dim=1).reshape(-1)

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
        """(C, N*vec + ...) reference dof numbers of each cell, variables concatenated (problem.py:86-92)."""
        return onp.hstack([(fe.vec * fe.cells[:, :, None] + onp.arange(fe.vec)[None, None, :] + o).reshape(
            fe.num_cells, -1) for fe, o in zip(self.fes, self.offset)])
