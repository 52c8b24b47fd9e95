"""FiniteElement: reference-element tables, Dirichlet data, boundary faces.

Counterpart of ``jax_fem/fe.py``.  Host-side work (tables, Dirichlet and
boundary-face discovery from user location functions) stays in numpy as in the
reference's setup; the per-cell geometry (fe.py:112-218) runs in the device
geometry kernel and is computed on first access.
"""
import inspect
from dataclasses import dataclass
from typing import Any, Optional

import numpy as onp

from .basis import get_face_shape_vals_and_grads, get_shape_vals_and_grads
from .generate_mesh import Mesh


def _num_args(fn):
    code = getattr(fn, '__code__', None)
    if code is not None:
        n = code.co_argcount
        if inspect.ismethod(fn):
            n -= 1
        return n
    return len(inspect.signature(fn).parameters)


def _pointwise(fn, points, inds, nargs):
    """fn evaluated point by point (always correct, slow)."""
    out = []
    for i in range(len(points)):
        args = (points[i],) if nargs == 1 else (points[i], inds[i])
        out.append(onp.asarray(_to_numpy(fn(*args))).reshape(()))
    return onp.array(out)


def _to_numpy(x):
    try:
        import torch
        if torch.is_tensor(x):
            return x.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return onp.asarray(x)


def _vectorised(fn, points, inds, nargs):
    """Evaluate fn on all points at once, passing point[d] as a coordinate array.

    Works for the usual location/value functions built from element-wise
    operations (isclose, comparisons, logical ops, arithmetic).  The result is
    checked against point-wise evaluation on a sample of points before use.
    """
    import torch
    n = len(points)
    candidates = [onp.ascontiguousarray(points.T), torch.as_tensor(onp.ascontiguousarray(points.T))]
    ind_c = [inds, torch.as_tensor(inds)]
    rng = onp.random.default_rng(0)
    sample = rng.choice(n, size=min(n, 24), replace=False)
    ref = _pointwise(fn, points[sample], inds[sample], nargs)
    for P, I in zip(candidates, ind_c):
        try:
            out = fn(P) if nargs == 1 else fn(P, I)
            out = _to_numpy(out)
            out = onp.broadcast_to(out, (n,)) if onp.ndim(out) == 0 else onp.asarray(out).reshape(-1)
            if out.shape != (n,):
                continue
            if ref.dtype == bool or out.dtype == bool:
                ok = onp.array_equal(out[sample].astype(bool), ref.astype(bool))
            else:
                ok = onp.allclose(out[sample], ref.astype(onp.float64), rtol=1e-12, atol=0)
            if ok:
                return out
        except Exception:
            continue
    return None


def evaluate_point_fn(fn, points, inds=None):
    """fn(point[, ind]) over all points -> array (num_points,)."""
    points = onp.asarray(points, dtype=onp.float64)
    inds = onp.arange(len(points)) if inds is None else onp.asarray(inds)
    nargs = _num_args(fn)
    if nargs not in (1, 2):
        raise ValueError(f"Wrong number of arguments for location_fn: must be 1 or 2, get {nargs}")
    if len(points) == 0:
        return onp.zeros((0,))
    out = _vectorised(fn, points, inds, nargs)
    if out is None:
        out = _pointwise(fn, points, inds, nargs)
    return out


def evaluate_location_fn(fn, points, inds=None):
    return evaluate_point_fn(fn, points, inds).astype(bool)


@dataclass
class FiniteElement:
    """One finite-element variable (fe.py:20-110)."""
    mesh: Mesh
    vec: int
    dim: int
    ele_type: str
    quadrature_rule: Any = None
    quadrature_order: Optional[int] = None
    dirichlet_bc_info: Optional[list] = None

    def __post_init__(self):
        if self.dim != 3:
            raise NotImplementedError("the device path is three-dimensional (dim=3)")
        self.points = onp.asarray(self.mesh.points, dtype=onp.float64)
        self.cells = onp.asarray(self.mesh.cells, dtype=onp.int64)
        self.num_cells = len(self.cells)
        self.num_total_nodes = len(self.points)
        self.num_total_dofs = self.num_total_nodes * self.vec
        self.shape_vals, self.shape_grads_ref, self.quad_weights = get_shape_vals_and_grads(
            self.ele_type, self.quadrature_rule, self.quadrature_order)
        (self.face_shape_vals, self.face_shape_grads_ref, self.face_quad_weights, self.face_normals,
         self.face_inds) = get_face_shape_vals_and_grads(self.ele_type, self.quadrature_rule, self.quadrature_order)
        self.num_quads = self.shape_vals.shape[0]
        self.num_nodes = self.shape_vals.shape[1]
        self.num_faces = self.face_shape_vals.shape[0]
        self.num_face_quads = self.face_quad_weights.shape[1]
        if self.cells.shape[1] != self.num_nodes:
            raise ValueError(f"cells have {self.cells.shape[1]} nodes, {self.ele_type} needs {self.num_nodes}")
        self.node_inds_list, self.vec_inds_list, self.vals_list = \
            self.Dirichlet_boundary_conditions(self.dirichlet_bc_info)
        self._dev_mesh = None
        self._geom = None

    # ------------------------------------------------------------- device
    def device_mesh(self):
        if self._dev_mesh is None:
            from ._device import DeviceMesh
            self._dev_mesh = DeviceMesh(self)
        return self._dev_mesh

    def get_shape_grads(self, points=None):
        """(shape_grads (C,Q,N,3), JxW (C,Q)) from the device geometry kernel (fe.py:112-141)."""
        if points is not None and points is not self.points:
            raise NotImplementedError("geometry with moved points is outside the device path")
        g, jxw, _ = self.device_mesh().geometry(want_grads=True, want_jxw=True)
        return g, jxw

    @property
    def shape_grads(self):
        if self._geom is None:
            self._geom = self.get_shape_grads()
        return self._geom[0]

    @property
    def JxW(self):
        if self._geom is None:
            self._geom = self.get_shape_grads()
        return self._geom[1]

    def get_face_shape_grads(self, boundary_inds, points=None):
        """(face shape grads (S,FQ,N,3), nanson scale (S,FQ)) (fe.py:143-183)."""
        g, ns, _ = self.device_mesh().face_geometry(boundary_inds, want_grads=True, want_nanson=True)
        return g, ns

    def get_physical_quad_points(self, points=None):
        """(C,Q,3) physical quadrature points (fe.py:185-197)."""
        _, _, x = self.device_mesh().geometry(want_qpts=True)
        return x

    def get_physical_surface_quad_points(self, boundary_inds, points=None):
        """(S,FQ,3) physical face quadrature points (fe.py:199-218)."""
        _, _, x = self.device_mesh().face_geometry(boundary_inds, want_qpts=True)
        return x

    # ------------------------------------------------------------ BC data
    def Dirichlet_boundary_conditions(self, dirichlet_bc_info):
        """node_inds_list, vec_inds_list, vals_list (fe.py:220-258)."""
        node_inds_list, vec_inds_list, vals_list = [], [], []
        if dirichlet_bc_info is not None:
            location_fns, vecs, value_fns = dirichlet_bc_info
            assert len(location_fns) == len(value_fns) and len(value_fns) == len(vecs)
            for i in range(len(location_fns)):
                flags = evaluate_location_fn(location_fns[i], self.points)
                node_inds = onp.argwhere(flags).reshape(-1)
                vec_inds = onp.ones_like(node_inds, dtype=onp.int32) * vecs[i]
                values = evaluate_point_fn(value_fns[i], self.points[node_inds]).astype(onp.float64).reshape(-1) \
                    if len(node_inds) else onp.zeros((0,))
                node_inds_list.append(node_inds)
                vec_inds_list.append(vec_inds)
                vals_list.append(values)
        return node_inds_list, vec_inds_list, vals_list

    def update_Dirichlet_boundary_conditions(self, dirichlet_bc_info):
        """Reset Dirichlet data (fe.py:260-269)."""
        self.node_inds_list, self.vec_inds_list, self.vals_list = \
            self.Dirichlet_boundary_conditions(dirichlet_bc_info)

    def get_boundary_conditions_inds(self, location_fns):
        """[(S,2) (cell, local face)] per location function, row-major order (fe.py:271-322)."""
        boundary_inds_list = []
        if location_fns is not None:
            for fn in location_fns:
                flags = evaluate_location_fn(fn, self.points)
                face_nodes = self.cells[:, self.face_inds]  # (C, 6, 4)
                on = onp.all(flags[face_nodes], axis=-1)
                boundary_inds_list.append(onp.argwhere(on))
        return boundary_inds_list

    # --------------------------------------------------- post-processing
    def _quad_kernel(self, sol, want_grad):
        """u or grad u at the quadrature points by fem_quad_grad, for a solution that
        does not carry autograd history; None when torch must do it (differentiable)."""
        import torch
        from ._lib import check, ptr, stream_ptr
        dm = self.device_mesh()
        if (torch.is_tensor(sol) and sol.requires_grad) or self.vec not in (1, 3):
            return None
        u = torch.as_tensor(sol, dtype=torch.float64, device=dm.device).reshape(-1).contiguous()
        out = dm.empty(dm.C * dm.Q, self.vec, 3) if want_grad else dm.empty(dm.C * dm.Q, self.vec)
        check(dm.L.fem_quad_grad(dm.ref, ptr(u), ptr(out) if want_grad else None, None if want_grad else ptr(out),
                                 stream_ptr()), 'fem_quad_grad')
        return out.reshape((dm.C, dm.Q, self.vec, 3) if want_grad else (dm.C, dm.Q, self.vec))

    def convert_from_dof_to_quad(self, sol):
        """(C,Q,vec) solution at quadrature points (fe.py:324-341)."""
        import torch
        out = self._quad_kernel(sol, False)
        if out is not None:
            return out
        dm = self.device_mesh()
        sol = torch.as_tensor(sol, dtype=torch.float64, device=dm.device)
        sv = dm.shape_vals
        return torch.einsum('cav,qa->cqv', sol[dm.cells.long()], sv)

    def convert_from_dof_to_face_quad(self, sol, boundary_inds):
        """(S,FQ,vec) solution at face quadrature points (fe.py:343-364)."""
        import torch
        dm = self.device_mesh()
        sol = torch.as_tensor(sol, dtype=torch.float64, device=dm.device)
        bi = torch.as_tensor(onp.asarray(boundary_inds), device=dm.device).long()
        fsv = dm.face_shape_vals[bi[:, 1]]  # (S,FQ,N)
        return torch.einsum('sav,sqa->sqv', sol[dm.cells.long()[bi[:, 0]]], fsv)

    def sol_to_grad(self, sol):
        """(C,Q,vec,dim) solution gradient at quadrature points (fe.py:366-382)."""
        import torch
        out = self._quad_kernel(sol, True)
        if out is not None:
            return out
        dm = self.device_mesh()
        sol = torch.as_tensor(sol, dtype=torch.float64, device=dm.device)
        return torch.einsum('cav,cqad->cqvd', sol[dm.cells.long()], self.shape_grads)
