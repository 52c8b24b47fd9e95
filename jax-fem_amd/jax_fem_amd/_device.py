"""Device state of a problem: mesh descriptors, matrix pattern, buffers, calls.

Everything numerical here is a call into libfem_mi355x (csrc/); torch only
allocates device memory, supplies the stream and evaluates the user's own
mass/surface callbacks at quadrature points.
"""
import ctypes as C
import time
import logging

import numpy as onp
import torch
from torch.func import vmap

from . import _lib
from ._lib import check, ptr, stream_ptr
from .materials import identify

logger = logging.getLogger('jax_fem_amd')


def _i32(x, device):
    return torch.as_tensor(onp.ascontiguousarray(onp.asarray(x, dtype=onp.int32)), device=device)


def _f64(x, device):
    return torch.as_tensor(onp.ascontiguousarray(onp.asarray(x, dtype=onp.float64)), device=device)


class DeviceMesh:
    """Device copies of one FiniteElement's mesh and reference tables."""

    def __init__(self, fe):
        L = _lib.lib()
        self.L = L
        self.device = torch.device('cuda', torch.cuda.current_device())
        d = self.device
        self.points = _f64(fe.points, d)
        self.cells = _i32(fe.cells, d)
        self.shape_vals = _f64(fe.shape_vals, d)
        self.shape_grads_ref = _f64(fe.shape_grads_ref, d)
        self.quad_weights = _f64(fe.quad_weights, d)
        self.face_shape_vals = _f64(fe.face_shape_vals, d)
        self.face_shape_grads_ref = _f64(fe.face_shape_grads_ref, d)
        self.face_weights = _f64(fe.face_quad_weights, d)
        self.face_normals = _f64(fe.face_normals, d)
        self.C, self.N = fe.cells.shape
        self.Q = fe.num_quads
        self.FQ = fe.num_face_quads
        self.V = fe.vec
        self.Nn = fe.num_total_nodes
        m = _lib.FemMesh()
        m.nodes_per_cell, m.num_quads, m.vec, m.num_face_quads = self.N, self.Q, self.V, self.FQ
        m.num_cells, m.num_nodes = self.C, self.Nn
        for name in ('points', 'cells', 'shape_vals', 'shape_grads_ref', 'quad_weights', 'face_shape_vals',
                     'face_shape_grads_ref', 'face_weights', 'face_normals'):
            setattr(m, name, getattr(self, name).data_ptr())
        self.desc = m
        self.ref = C.byref(m)

    def empty(self, *shape, dtype=torch.float64):
        return torch.empty(shape, dtype=dtype, device=self.device)

    # -------------------------------------------------------- geometry
    def geometry(self, want_grads=False, want_jxw=False, want_qpts=False):
        g = self.empty(self.C, self.Q, self.N, 3) if want_grads else None
        j = self.empty(self.C, self.Q) if want_jxw else None
        x = self.empty(self.C, self.Q, 3) if want_qpts else None
        check(self.L.fem_geometry(self.ref, ptr(g), ptr(j), ptr(x), stream_ptr()), 'fem_geometry')
        return g, j, x

    def faces_tensor(self, boundary_inds):
        return _i32(onp.asarray(boundary_inds).reshape(-1, 2), self.device)

    def face_geometry(self, boundary_inds, want_grads=False, want_nanson=False, want_qpts=False):
        f = boundary_inds if torch.is_tensor(boundary_inds) else self.faces_tensor(boundary_inds)
        S = f.shape[0]
        g = self.empty(S, self.FQ, self.N, 3) if want_grads else None
        ns = self.empty(S, self.FQ) if want_nanson else None
        x = self.empty(S, self.FQ, 3) if want_qpts else None
        check(self.L.fem_face_geometry(self.ref, S, ptr(f), ptr(g), ptr(ns), ptr(x), stream_ptr()),
              'fem_face_geometry')
        return g, ns, x

    def incidence(self, conn):
        """Node -> (entity, local node) incidence of an (E, N) int32 connectivity."""
        E, N = conn.shape
        inc_ptr = self.empty(self.Nn + 1, dtype=torch.int32)
        inc_idx = self.empty(max(E * N, 1), dtype=torch.int32)
        wsb = 4 * self.Nn + 512 + int(self.L.fem_scan_ws_bytes(self.Nn))
        ws = self.empty(wsb, dtype=torch.uint8)
        check(self.L.fem_incidence(E, N, ptr(conn), self.Nn, ptr(inc_ptr), ptr(inc_idx), ptr(ws), stream_ptr()),
              'fem_incidence')
        return inc_ptr, inc_idx

    def scan(self, x):
        n = x.numel()
        out = self.empty(n + 1, dtype=torch.int32)
        ws = self.empty(int(self.L.fem_scan_ws_bytes(n)), dtype=torch.uint8)
        check(self.L.fem_exclusive_scan(n, ptr(x), ptr(out), ptr(ws), stream_ptr()), 'fem_exclusive_scan')
        return out


def _eval_quad_map(fn, V, u, x, ivs):
    """Evaluate a user point map fn(u (V,), x (3,), *iv) at P points -> (P, V) float64.

    Tries torch.func.vmap, then a transposed batch (point[d] -> row d), then a
    point loop; the batched results are checked against point evaluation.
    """
    P = x.shape[0]
    dev = x.device

    def fix(r):
        r = torch.as_tensor(r, dtype=torch.float64, device=dev)
        if r.shape == (V,):
            r = r.expand(P, V)
        return r.reshape(P, V)

    sample = onp.unique(onp.linspace(0, P - 1, num=min(P, 7)).astype(int))

    def point(i):
        return torch.as_tensor(fn(u[i], x[i], *[iv[i] for iv in ivs]), dtype=torch.float64,
                               device=dev).reshape(V)

    ref = torch.stack([point(int(i)) for i in sample]) if P else None

    def ok(r):
        return P == 0 or torch.allclose(r[sample], ref, rtol=1e-12, atol=1e-300)

    try:
        r = fix(vmap(fn)(u, x, *ivs))
        if ok(r):
            return r
    except Exception:
        pass
    try:
        tr = lambda t: t.transpose(0, -1) if t.dim() > 1 else t
        r = fn(tr(u), tr(x), *[tr(iv) for iv in ivs])
        r = torch.as_tensor(r, dtype=torch.float64, device=dev)
        r = r.expand(P, V) if r.shape == (V,) else r.transpose(0, -1).reshape(P, V)
        if ok(r):
            return r
    except Exception:
        pass
    return torch.stack([point(i) for i in range(P)]) if P else torch.zeros((0, V), dtype=torch.float64, device=dev)


def _surface_map_kernel(smap, V):
    """The surface kernel of problem.py:238-259 for a surface map that depends on u:
    sum_q m(u_q, x_q, *iv_q) N_a(q) nanson_q, as a per-face function."""
    def kernel(cell_sol_flat, x, face_shape_vals, face_shape_grads, nanson, *ivs):
        N = face_shape_vals.shape[1]
        u = face_shape_vals @ cell_sol_flat.reshape(N, V)                    # (FQ, V)
        m = torch.func.vmap(lambda a, b, *c: _as_vector(smap(a, b, *c)).reshape(V))(u, x, *ivs)
        return ((face_shape_vals * nanson[0][:, None])[:, :, None] * m[:, None, :]).sum(0).reshape(-1)
    return kernel


def dm_size(dev):
    """Entries of the per-cell tangent buffer: C * (N * V)^2."""
    return dev.dm.C * (dev.dm.N * dev.V) ** 2


def _as_vector(r):
    """A point map's result as one tensor (maps may return lists of scalars)."""
    if isinstance(r, (list, tuple)):
        return torch.stack([torch.as_tensor(e, dtype=torch.float64) for e in r])
    return r


def _flatten_iv(iv, lead, device):
    """(lead..., rest) internal variable -> (prod(lead), rest) float64 device tensor."""
    t = torch.as_tensor(iv)
    if t.device != device or t.dtype != torch.float64:
        t = t.to(device=device, dtype=torch.float64)
    return t.reshape(int(onp.prod(lead)), *t.shape[len(lead):])


class DeviceProblem:
    """Pattern, boundary data, loads and solver buffers of one Problem."""

    def __init__(self, problem):
        fe = problem._fe_device     # fes[0], or for several variables one element with vec = sum(vec)
        self.problem = problem
        self.fe = fe
        self.dm = fe.device_mesh()
        dm = self.dm
        self.L = dm.L
        self.V = fe.vec
        self.n = dm.Nn * self.V
        # several variables: variable i holds components voff[i] .. voff[i] + vec_i - 1 of each
        # node block; cperm[k] = position in the device's per-cell vector of entry k of the
        # reference's per-cell vector (variables concatenated, problem.py:86-92)
        self.vecs = [f.vec for f in problem.fes]
        self.voff = onp.cumsum([0] + self.vecs[:-1])
        self.cperm = onp.concatenate([(onp.arange(dm.N)[:, None] * self.V + o + onp.arange(v)[None, :]).reshape(-1)
                                      for v, o in zip(self.vecs, self.voff)])
        self.cperm_inv = onp.argsort(self.cperm)
        self.inc_ptr, self.inc_idx = dm.incidence(dm.cells)
        self._build_pattern()
        self.surfaces = []
        for bi in problem.boundary_inds_list:
            faces = dm.faces_tensor(bi)
            conn = dm.cells[faces[:, 0].long()].contiguous() if faces.shape[0] else \
                torch.zeros((0, dm.N), dtype=torch.int32, device=dm.device)
            fptr, fidx = dm.incidence(conn)
            _, _, xq = dm.face_geometry(faces, want_qpts=True)
            self.surfaces.append(dict(faces=faces, inc_ptr=fptr, inc_idx=fidx, xq=xq))
        self._bc_sources = None
        self._mat_cache = {}
        self.buffers = {}
        self.material = None
        self.param = None
        self.load = None
        self.mass_dyn = None
        self.cell_terms = []
        self.dynamic_surfaces = set()
        self._constraint = None
        self.host_status = torch.empty(128, dtype=torch.uint8).pin_memory()
        self._amg = None
        self._amg_uses = 0
        self._asm_plan = None
        self.last_residual = None

    # ------------------------------------------------------------ pattern
    def _build_pattern(self):
        dm, L = self.dm, self.L
        N, Nn = dm.N, dm.Nn
        contrib_idx = dm.empty(max(dm.C * N * N, 1), dtype=torch.int32)
        row_len = dm.empty(Nn, dtype=torch.int32)
        err = dm.empty(1, dtype=torch.int32)
        check(L.fem_pattern_rows(dm.ref, ptr(self.inc_ptr), ptr(self.inc_idx), ptr(contrib_idx), ptr(row_len),
                                 ptr(err), stream_ptr()), 'fem_pattern_rows')
        if int(err.item()) > 0:
            raise _lib.FemError(f"a node row has {int(err.item())} candidate blocks (capacity 2048)")
        if int(row_len.sum(dtype=torch.int64).item()) >= 2 ** 31:
            raise _lib.FemError("the block pattern has more than 2^31 blocks (int32 row pointers)")
        row_ptr = dm.scan(row_len)
        nnzb = int(row_ptr[-1].item())
        col_idx = dm.empty(nnzb, dtype=torch.int32)
        contrib_ptr = dm.empty(nnzb + 1, dtype=torch.int32)
        diag_slot = dm.empty(Nn, dtype=torch.int32)
        check(L.fem_pattern_fill(dm.ref, ptr(self.inc_ptr), ptr(contrib_idx), ptr(row_ptr), ptr(col_idx),
                                 ptr(contrib_ptr), ptr(diag_slot), stream_ptr()), 'fem_pattern_fill')
        ns = (Nn + 63) // 64
        sell_row = dm.empty(ns * 64, dtype=torch.int32)
        sell_pos = dm.empty(Nn, dtype=torch.int32)
        slice_blocks = dm.empty(ns, dtype=torch.int32)
        # natural row order unless its slices carry more than 5% padding (rows of mixed
        # lengths, e.g. HEX27 lattices); then rows are length-sorted within 2048-row windows
        for sort in (0, 1):
            check(L.fem_sell_order(Nn, sort, ptr(row_ptr), ptr(sell_row), ptr(sell_pos), stream_ptr()),
                  'fem_sell_order')
            check(L.fem_sell_slice_sizes(Nn, ptr(row_ptr), ptr(sell_row), ptr(slice_blocks), stream_ptr()),
                  'fem_sell_slice_sizes')
            if int(slice_blocks.sum(dtype=torch.int64).item()) >= 2 ** 31:
                raise _lib.FemError("the SELL-64 layout needs more than 2^31 blocks (int32 slice pointers)")
            slice_ptr = dm.scan(slice_blocks)
            sell_blocks = int(slice_ptr[-1].item())
            if sell_blocks <= 1.05 * nnzb:
                break
        self.sell_sorted = bool(sort)
        sell_col = dm.empty(sell_blocks, dtype=torch.int32)
        p = _lib.FemPattern()
        p.num_rows, p.nnzb, p.num_slices, p.sell_blocks = Nn, nnzb, ns, sell_blocks
        p.vec, p.num_cells, p.nodes_per_cell = self.V, dm.C, N
        self.pat_tensors = dict(row_ptr=row_ptr, col_idx=col_idx, diag_slot=diag_slot, contrib_ptr=contrib_ptr,
                                contrib_idx=contrib_idx, slice_ptr=slice_ptr, sell_col=sell_col, sell_row=sell_row,
                                sell_pos=sell_pos)
        for k, t in self.pat_tensors.items():
            setattr(p, k, t.data_ptr())
        self.pat = p
        self.pref = C.byref(p)
        check(L.fem_sell_fill(self.pref, ptr(sell_col), stream_ptr()), 'fem_sell_fill')
        self.nnzb = nnzb
        self.sell_blocks = sell_blocks
        logger.debug("pattern: %d block rows, %d blocks (%d padded), vec %d", Nn, nnzb, sell_blocks, self.V)

    # --------------------------------------------------------- preparation
    def refresh_bc(self):
        fes = self.problem.fes
        sources = tuple((id(a), id(b), id(c)) for fe in fes
                        for a, b, c in zip(fe.node_inds_list, fe.vec_inds_list, fe.vals_list))
        if sources == self._bc_sources:
            return
        dofs, vals = [], []
        for fe, off in zip(fes, self.voff):
            for node_inds, vec_inds, v in zip(fe.node_inds_list, fe.vec_inds_list, fe.vals_list):
                node_inds = onp.asarray(node_inds).reshape(-1)
                dofs.append(node_inds * self.V + off + onp.asarray(vec_inds).reshape(-1))
                vals.append(onp.broadcast_to(onp.asarray(v, dtype=onp.float64).reshape(-1), node_inds.shape))
        if dofs:
            d = onp.concatenate(dofs)
            v = onp.concatenate(vals)
            # the last Dirichlet set wins on a shared dof (solver.py:297-300 applies them in order)
            rd = d[::-1]
            u, first = onp.unique(rd, return_index=True)
            d, v = u, v[::-1][first]
        else:
            d, v = onp.zeros((0,), onp.int64), onp.zeros((0,))
        self.bc_dofs = _i32(d, self.dm.device)
        self.bc_vals = _f64(v, self.dm.device)
        self.bc_dofs_long = self.bc_dofs.long()
        self.n_bc = int(d.shape[0])
        self._bc_sources = sources

    def prepare(self):
        """Material, per-quad parameter, constant loads and BC data from the problem's current state."""
        pb, dm = self.problem, self.dm
        self.refresh_bc()
        ivs = list(pb.internal_vars) if pb.internal_vars is not None else []
        C_, Q = dm.C, dm.Q
        iv_shapes = tuple(tuple(torch.as_tensor(iv).shape[2:]) for iv in ivs)
        for iv in ivs:
            shp = tuple(torch.as_tensor(iv).shape)
            if shp[:2] != (C_, Q):
                raise ValueError(f"internal variable of shape {shp} must be (num_cells, num_quads, ...) = ({C_}, {Q}, ...)")
        if not hasattr(pb, 'get_tensor_map') and not hasattr(pb, 'get_universal_kernel'):
            raise NotImplementedError("problems without get_tensor_map or get_universal_kernel are outside the "
                                      "device path")
        key = iv_shapes
        if key not in self._mat_cache:
            if hasattr(pb, 'get_tensor_map'):
                self._mat_cache[key] = identify(pb.get_tensor_map(), self.V, iv_shapes)
            else:   # no laplace term: the cell kernels contribute zeros
                from .materials import NoLaw
                self._mat_cache[key] = NoLaw(self.V)
        mat = self._mat_cache[key]
        self.material = mat
        # a closed-form law recognised from a user callable is re-checked at the real
        # quadrature states of every residual and tangent (_check_law)
        self._law_check = None
        if hasattr(pb, 'get_tensor_map') and not mat.tabulated:
            from .materials import Material
            fn = pb.get_tensor_map()
            if not isinstance(fn, Material):
                self._law_check = dict(fn=fn, key=key, ivs=[_flatten_iv(iv, (C_, Q), dm.device).detach()
                                                            for iv in ivs])
        if mat.has_param:
            k = getattr(mat, 'param_index', 0)
            if k >= len(ivs):
                raise ValueError(f"{mat!r} reads internal variable {k}, the problem has {len(ivs)}")
            self.param_index = k
            self.param = _flatten_iv(ivs[k], (C_, Q), dm.device).reshape(C_, Q).contiguous()
        else:
            self.param_index = None
            self.param = None
        self.tab_ivs = None
        if mat.has_state:
            sig = _flatten_iv(ivs[mat.sig_index], (C_, Q, 3, 3), dm.device).reshape(C_, Q, 9)
            eps = _flatten_iv(ivs[mat.eps_index], (C_, Q, 3, 3), dm.device).reshape(C_, Q, 9)
            self.state = torch.cat([sig, eps], dim=2).contiguous()
        elif mat.tabulated:
            # per-quad table of P and dP/dG, filled at the current u before each launch
            if 'tab' not in self.buffers:
                self.buffers['tab'] = dm.empty(max(C_ * Q * _lib.FEM_TAB_STRIDE(self.V), 1))
            self.state = self.buffers['tab']
            self.tab_ivs = [_flatten_iv(iv, (C_, Q), dm.device).detach() for iv in ivs]
        else:
            self.state = None
        self.mat_desc = mat.to_struct(self.param.data_ptr() if self.param is not None else None,
                                      self.state.data_ptr() if self.state is not None else None)
        self.mref = C.byref(self.mat_desc)
        self.cell_terms = []     # per-cell weak-form terms evaluated by torch (universal kernels)
        if hasattr(pb, 'get_universal_kernel'):
            self.cell_terms.append(self._cell_term(pb.get_universal_kernel(), ivs))
        self.load = self._constant_loads(ivs)
        self.dynamic_surfaces = {i for i, _, _ in self._pending_faces}
        for i, fn, siv in self._pending_faces:
            self.cell_terms.append(self._face_term(i, fn, siv))
        if hasattr(pb, 'get_universal_kernels_surface'):
            for i, fn in enumerate(pb.get_universal_kernels_surface()):
                siv = pb.internal_vars_surfaces[i] if i < len(pb.internal_vars_surfaces) else ()
                self.cell_terms.append(self._face_term(i, fn, siv))

    def _constant_loads(self, ivs):
        pb, dm = self.problem, self.dm
        V, dev = self.V, dm.device
        total = None
        self.mass_dyn = None
        self._pending_faces = []
        if hasattr(pb, 'get_mass_map'):
            if 'xq' not in self.buffers:
                self.buffers['xq'] = dm.geometry(want_qpts=True)[2].reshape(-1, 3)
            xq = self.buffers['xq']
            civ = [_flatten_iv(iv, (dm.C, dm.Q), dev) for iv in ivs]
            fn = pb.get_mass_map()
            m = self._eval_u_independent(fn, xq, civ, 'mass_map', allow_u=True)
            if m is None:
                # m depends on u: evaluated at u_q with its derivative before each launch
                hit = self._mat_cache.get('mass_pm')
                if hit is None or hit[0] is not fn:
                    from .materials import PointMap
                    hit = (fn, PointMap(lambda u, x, *iv: _as_vector(fn(u, x, *iv)), (V,), (V,), 'mass map'))
                    self._mat_cache['mass_pm'] = hit
                self.mass_dyn = dict(pm=hit[1], rest=[xq] + [c.detach() for c in civ])
            else:
                cell = self.cell_buffer()
                check(self.L.fem_cell_load(dm.ref, ptr(m), ptr(cell), stream_ptr()), 'fem_cell_load')
                total = dm.empty(self.n)
                check(self.L.fem_gather_nodal(dm.Nn, V, ptr(self.inc_ptr), ptr(self.inc_idx), ptr(cell), None,
                                              ptr(total), stream_ptr()), 'fem_gather_nodal')
        if self.surfaces and not hasattr(pb, 'get_surface_maps'):
            if not hasattr(pb, 'get_universal_kernels_surface'):
                raise AssertionError("Missing definitions for surface integral")   # problem.py:338-344
        elif self.surfaces:
            smaps = pb.get_surface_maps()
            for i, srf in enumerate(self.surfaces):
                S = srf['faces'].shape[0]
                if S == 0:
                    continue
                siv = pb.internal_vars_surfaces[i] if i < len(pb.internal_vars_surfaces) else ()
                civ = [_flatten_iv(iv, (S, dm.FQ), dev) for iv in siv]
                t = self._eval_u_independent(smaps[i], srf['xq'].reshape(-1, 3), civ, 'surface_map', allow_u=True)
                if t is None:
                    # depends on u: the surface kernel of problem.py:238-259 as a per-face term
                    self._pending_faces.append((i, _surface_map_kernel(smaps[i], V), siv))
                    continue
                rf = dm.empty(S * dm.N * V)
                check(self.L.fem_face_load(dm.ref, S, ptr(srf['faces']), ptr(t), ptr(rf), stream_ptr()),
                      'fem_face_load')
                if total is None:
                    total = torch.zeros(self.n, dtype=torch.float64, device=dev)
                check(self.L.fem_gather_nodal(dm.Nn, V, ptr(srf['inc_ptr']), ptr(srf['inc_idx']), ptr(rf),
                                              ptr(total), ptr(total), stream_ptr()), 'fem_gather_nodal')
        return total

    def _eval_u_independent(self, fn, xq, civ, what, allow_u=False):
        """fn at u = 0 on every point; None (allow_u) or NotImplementedError when it depends on u.

        The map is evaluated at every point twice more, at two independent random u with
        distinct components per point (seeded), so a dependence on any combination of the
        components or one confined to part of the domain is seen: m(u) == m(0) at all
        points for both draws is required for the map to be a constant load."""
        P = xq.shape[0]
        u0 = torch.zeros((P, self.V), dtype=torch.float64, device=xq.device)
        m0 = _eval_quad_map(fn, self.V, u0, xq, civ).contiguous()
        if P:
            gen = torch.Generator(device=xq.device).manual_seed(2024)
            scale = max(float(m0.abs().max()), 1.0)
            for _ in range(2):
                ur = torch.randn((P, self.V), dtype=torch.float64, device=xq.device, generator=gen)
                mr = _eval_quad_map(fn, self.V, ur, xq, civ)
                if float((mr - m0).abs().max()) > 1e-13 * scale:
                    if allow_u:
                        return None
                    raise NotImplementedError(f"a {what} that depends on u is outside the device path")
        return m0

    # ------------------------------------------------------------ buffers
    def cell_buffer(self):
        if 'cell' not in self.buffers:
            self.buffers['cell'] = self.dm.empty(self.dm.C * self.dm.N * self.V)
        return self.buffers['cell']

    def matrix_buffers(self, precision):
        """(SELL values, inverse diagonal) of the matrix in `precision`."""
        key = ('mat', precision)
        if key not in self.buffers:
            dt = torch.float32 if precision == _lib.FEM_F32 else torch.float64
            vals = self.dm.empty(max(self.sell_blocks * self.V * self.V, 1), dtype=dt)
            dinv = self.dm.empty(self.n, dtype=dt)
            self.buffers[key] = (vals, dinv)
        return self.buffers[key]

    def ke_buffer(self, precision):
        """Per-cell tangent blocks (C, N, N, V, V) of the two-pass assembly."""
        key = ('ke', precision)
        if key not in self.buffers:
            dt = torch.float32 if precision == _lib.FEM_F32 else torch.float64
            dm = self.dm
            self.buffers[key] = dm.empty(max(dm.C * dm.N * dm.N * self.V * self.V, 1), dtype=dt)
        return self.buffers[key]

    _FUSED_KINDS = (_lib.FEM_MAT_POISSON, _lib.FEM_MAT_LINEAR_ELASTIC, _lib.FEM_MAT_SIMP, _lib.FEM_MAT_NEO_HOOKEAN,
                    _lib.FEM_MAT_J2_PLASTIC)

    def assembly_plan(self):
        """The single-pass assembly's plan (slice cell lists, slice-local contribution ids), built
        once per pattern; None where it does not apply (not HEX8 with 8 quadrature points, rows of
        more than 32 blocks, slices with more than 4096 incidences)."""
        if self._asm_plan is None:
            self._asm_plan = False
            dm, L, s = self.dm, self.L, stream_ptr()
            sp = self.pat_tensors['slice_ptr']
            ns = self.pat.num_slices
            if dm.N == 8 and dm.Q == 8 and dm.C > 0 and ns > 0:
                max_slots = int(((sp[1:] - sp[:-1]) // 64).max().item())
                if max_slots <= 32:
                    counts = dm.empty(ns, dtype=torch.int32)
                    pcounts = dm.empty(ns, dtype=torch.int32)
                    err = dm.empty(1, dtype=torch.int32)
                    check(L.fem_assembly_plan_counts(self.pref, ptr(self.inc_ptr), ptr(self.inc_idx), ptr(counts),
                                                     ptr(pcounts), ptr(err), s), 'fem_assembly_plan_counts')
                    max_pairs = int(pcounts.max().item()) if ns else 0
                    if int(err.item()) == 0 and max_pairs <= 1024:
                        sc_ptr = dm.scan(counts)
                        sp_ptr = dm.scan(pcounts)
                        sc = dm.empty(max(int(sc_ptr[-1].item()), 1), dtype=torch.int32)
                        sp = dm.empty(max(int(sp_ptr[-1].item()), 1), dtype=torch.int32)
                        loc = dm.empty(max(dm.C * dm.N * dm.N, 1), dtype=torch.int32)
                        check(L.fem_assembly_plan_fill(self.pref, ptr(self.inc_ptr), ptr(self.inc_idx), ptr(sc_ptr),
                                                       ptr(sp_ptr), ptr(sc), ptr(sp), ptr(loc), ptr(err), s),
                              'fem_assembly_plan_fill')
                        if int(err.item()) == 0:
                            plan = _lib.FemAssemblyPlan()
                            plan.slice_cell_ptr, plan.slice_cells, plan.contrib_loc = (
                                sc_ptr.data_ptr(), sc.data_ptr(), loc.data_ptr())
                            plan.slice_pair_ptr, plan.slice_pairs = sp_ptr.data_ptr(), sp.data_ptr()
                            plan.max_slots, plan.max_pairs = max_slots, max_pairs
                            self._asm_plan = (plan, (sc_ptr, sc, sp_ptr, sp, loc))
                        else:
                            logger.info("single-pass assembly plan refused (code %d): two-pass assembly",
                                        int(err.item()))
        return self._asm_plan[0] if self._asm_plan else None

    def fused_applies(self, ke_precision):
        """True when the single-pass kernel assembles this problem's tangent."""
        return (ke_precision == _lib.FEM_F64 and self.material.kind in self._FUSED_KINDS
                and self.mass_dyn is None and not self.cell_terms and self.assembly_plan() is not None)

    # -------------------------------------------------------------- compute
    def quad_grad(self, u):
        """grad u at every quadrature point, (C*Q, V, 3) FP64 (problem.py:200-203)."""
        dm = self.dm
        G = dm.empty(dm.C * dm.Q, self.V, 3)
        check(self.L.fem_quad_grad(dm.ref, ptr(u), ptr(G), None, stream_ptr()), "fem_quad_grad")
        return G

    def _cell_term(self, fn, ivs):
        """get_universal_kernel (problem.py:275-283): fn(cell_sol_flat, x, cell_shape_grads,
        cell_JxW, cell_v_grads_JxW, *cell_internal_vars) -> (N*V,) per cell.  With several
        variables the shape gradients repeat per variable, (Q, N*nvars, 3), and cell_JxW is
        (nvars, Q), as problem.py:130-182 concatenates them."""
        dm = self.dm
        if 'uk_geo' not in self.buffers:
            nv = len(self.vecs)
            g, j, x = dm.geometry(want_grads=True, want_jxw=True, want_qpts=True)
            g = torch.cat([g] * nv, dim=2) if nv > 1 else g
            vg = g[:, :, :, None, :] * j[:, :, None, None, None]
            self.buffers['uk_geo'] = [x, g, j[:, None, :].expand(-1, nv, -1).contiguous(), vg]
        rest = self.buffers['uk_geo'] + [_flatten_iv(iv, (dm.C,), dm.device).detach() for iv in ivs]
        return dict(pm=self._term_map(('cell',), fn), rest=rest, ids=None, fn=fn, ivs=ivs, lead=(dm.C,))

    def _face_term(self, i, fn, siv):
        """get_universal_kernels_surface (problem.py:305-327): fn(cell_sol_flat, x,
        face_shape_vals, face_shape_grads, face_nanson_scale, *ivs) -> (N*V,) per face."""
        dm = self.dm
        srf = self.surfaces[i]
        key = ('uk_face', i)
        if key not in self.buffers:
            nv = len(self.vecs)
            faces = srf['faces']
            g, ns, x = dm.face_geometry(faces, want_grads=True, want_nanson=True, want_qpts=True)
            fv = dm.face_shape_vals[faces[:, 1].long()].contiguous()
            if nv > 1:
                fv, g = torch.cat([fv] * nv, dim=2), torch.cat([g] * nv, dim=2)
            self.buffers[key] = [x, fv, g, ns[:, None, :].expand(-1, nv, -1).contiguous()]
        S = srf['faces'].shape[0]
        rest = self.buffers[key] + [_flatten_iv(iv, (S,), dm.device).detach() for iv in siv]
        return dict(pm=self._term_map(('face', i), fn), rest=rest, ids=srf['faces'][:, 0].long(), fn=fn,
                    ivs=siv, lead=(S,), surface=i)

    def _term_map(self, slot, fn):
        """The PointMap of one weak-form term (one per slot, kept while the user function is the same object)."""
        from .materials import PointMap
        cache = self._mat_cache.setdefault('terms', {})
        hit = cache.get(slot)
        if hit is None or hit[0] is not fn:
            nv = self.dm.N * self.V
            hit = (fn, PointMap(lambda cs, *r: _as_vector(fn(cs, *r)), (nv,), (nv,), 'universal kernel'))
            cache[slot] = hit
        return hit[1]

    def cell_dofs_reference(self, u):
        """(C, N*V) cell dof vectors of the device vector u in the reference's per-cell order."""
        dm = self.dm
        cs = u.reshape(-1, self.V)[dm.cells.long()].reshape(dm.C, dm.N * self.V)
        return cs if len(self.vecs) == 1 else cs[:, torch.as_tensor(self.cperm, device=cs.device)]

    def _cell_terms_eval(self, u, jac):
        """Sum of the per-cell terms: r (C, N, V) and, with `jac`, K (C, N, N, V, V) FP64.
        The user's kernels see the reference's per-cell order; their values are moved to
        the device's node-block order here."""
        dm = self.dm
        N, V = dm.N, self.V
        cs = self.cell_dofs_reference(u)
        inv = torch.as_tensor(self.cperm_inv, device=dm.device) if len(self.vecs) > 1 else None
        r = torch.zeros((dm.C, N, V), dtype=torch.float64, device=dm.device)
        K = torch.zeros((dm.C, N, N, V, V), dtype=torch.float64, device=dm.device) if jac else None
        with torch.no_grad():
            for t in self.cell_terms:
                x = cs if t['ids'] is None else cs[t['ids']]
                y, d = t['pm'](x, t['rest'], jac)
                if inv is not None:
                    y = y[:, inv]
                    d = d[:, inv][:, :, inv] if jac else d
                ids = torch.arange(dm.C, device=dm.device) if t['ids'] is None else t['ids']
                r.index_add_(0, ids, y.reshape(-1, N, V))
                if jac:
                    K.index_add_(0, ids, d.reshape(-1, N, V, N, V).permute(0, 1, 3, 2, 4))
        return r, K

    def quad_vals(self, u):
        """u at every quadrature point, (C*Q, V) FP64."""
        dm = self.dm
        U = dm.empty(dm.C * dm.Q, self.V)
        check(self.L.fem_quad_grad(dm.ref, ptr(u), None, ptr(U), stream_ptr()), 'fem_quad_grad')
        return U

    def _mass_eval(self, u, jac):
        """m(u_q, x_q, *iv) and, with `jac`, M = dm/du at every quadrature point."""
        md = self.mass_dyn
        with torch.no_grad():
            m, M = md['pm'](self.quad_vals(u), md['rest'], jac)
        if jac:
            self.buffers['mass_M'] = M.contiguous()
        return m.contiguous()

    def _base_load(self, u):
        """Constant loads plus, for a mass map that depends on u, its nodal vector at u."""
        if self.mass_dyn is None:
            return self.load
        dm = self.dm
        m = self._mass_eval(u, False)
        cm = self.buffers.get('mass_cell')
        if cm is None:
            cm = self.buffers['mass_cell'] = dm.empty(dm.C * dm.N * self.V)
        check(self.L.fem_cell_load(dm.ref, ptr(m), ptr(cm), stream_ptr()), 'fem_cell_load')
        out = dm.empty(self.n)
        check(self.L.fem_gather_nodal(dm.Nn, self.V, ptr(self.inc_ptr), ptr(self.inc_idx), ptr(cm), ptr(self.load),
                                      ptr(out), stream_ptr()), 'fem_gather_nodal')
        return out

    def _cell_terms_tangent(self, u, Ke, faces=True):
        if not self.cell_terms:
            return
        if not faces:   # the cell kernels only (the cell part of the reference's V)
            terms, self.cell_terms = self.cell_terms, [t for t in self.cell_terms if 'surface' not in t]
            try:
                if self.cell_terms:
                    Ke[:dm_size(self)] += self._cell_terms_eval(u, True)[1].reshape(-1).to(Ke.dtype)
            finally:
                self.cell_terms = terms
            return
        K = self._cell_terms_eval(u, True)[1].reshape(-1)
        self.buffers['cell_K'] = K
        Ke[:K.numel()] += K.to(Ke.dtype)

    def face_tangents(self, u):
        """Per-face tangents [(S, N*V, N*V)] of each surface in the reference's jac layout
        (compute_face with jac, problem.py:398-424, 457-458): the jacfwd of the surface kernel,
        zero for a surface map that does not depend on u."""
        dm = self.dm
        N, V = dm.N, self.V
        cs = self.cell_dofs_reference(u)
        out = [torch.zeros((srf['faces'].shape[0], N * V, N * V), dtype=torch.float64, device=dm.device)
               for srf in self.surfaces]
        with torch.no_grad():
            for t in self.cell_terms:
                if 'surface' in t and t['ids'].numel():
                    out[t['surface']] += t['pm'](cs[t['ids']], t['rest'], True)[1].reshape(-1, N * V, N * V)
        return out

    def _mass_tangent(self, u, Ke, ke_storage):
        if self.mass_dyn is None:
            return
        self._mass_eval(u, True)
        check(self.L.fem_cell_mass_tangent(self.dm.ref, ptr(self.buffers['mass_M']), ke_storage, ptr(Ke),
                                           stream_ptr()), 'fem_cell_mass_tangent')

    def _tabulate(self, u, jac):
        """Fill the tabulated law's table at u: P, and with `jac` also D = dP/dG.
        A residual leaves D as the last tangent filled it (the operator of a
        solve is the tangent at its linearisation point)."""
        if not self.material.tabulated:
            return
        from .materials import evaluate_points
        dm = self.dm
        n, V = dm.C * dm.Q, self.V
        G = self.quad_grad(u)
        with torch.no_grad():
            P, D = evaluate_points(self.material, G, self.tab_ivs, jac)
        tab = self.state[:n * _lib.FEM_TAB_STRIDE(V)].view(n, _lib.FEM_TAB_STRIDE(V))
        tab[:, :3 * V] = P.reshape(n, 3 * V)
        if jac:
            tab[:, 3 * V:3 * V + 9 * V * V] = D.reshape(n, 9 * V * V)

    def _check_law(self, u):
        """Compare the closed-form law recognised from the user's tensor map with the map itself
        at real quadrature states: the 48 points of largest |grad u| and 48 spread over the mesh.
        On a difference the problem switches to the tabulated law (the map itself) for good."""
        lc = self._law_check
        if lc is None or lc.get('u') is u:
            return
        lc['u'] = u
        G = self.quad_grad(u)
        n = G.shape[0]
        if n == 0:
            return
        k = min(n, 48)
        mag = G.reshape(n, -1).abs().amax(1)
        idx = torch.unique(torch.cat([torch.topk(mag, k).indices,
                                      torch.linspace(0, n - 1, k, device=G.device).long()]))
        # the map is evaluated where identify() probed it: on the CPU, in float64 (the samples
        # come over in one copy); states with non-finite entries are left to the solve, which
        # reports them
        parts = [G[idx].reshape(len(idx), -1)] + [t[idx].reshape(len(idx), -1).to(G.dtype) for t in lc['ivs']]
        flat = torch.cat(parts, 1).cpu()
        o = parts[0].shape[1]
        Gs, ivs = flat[:, :o].reshape((len(idx),) + tuple(G.shape[1:])), []
        for t in lc['ivs']:
            w = parts[len(ivs) + 1].shape[1]
            ivs.append(flat[:, o:o + w].reshape((len(idx),) + tuple(t.shape[1:])).to(t.dtype))
            o += w
        fin = torch.isfinite(Gs.reshape(len(idx), -1)).all(1)
        for t in ivs:
            fin &= torch.isfinite(t.reshape(len(idx), -1)).all(1)
        if not bool(fin.all()):
            Gs, ivs = Gs[fin], [t[fin] for t in ivs]
            if Gs.shape[0] == 0:
                return
        from .materials import PointMap, Tabulated
        pm = self._mat_cache.get(('law_pm', lc['key']))
        if pm is None:
            pm = self._mat_cache[('law_pm', lc['key'])] = PointMap(lc['fn'], (self.V, 3), (self.V, 3), 'tensor map')
        mat = self.material
        law = lambda g, *r: mat(g, *r)
        with torch.no_grad():
            Pu = pm(Gs, ivs)[0]
            Pd = vmap(law)(Gs, *ivs).reshape(Pu.shape)
            # stiffness scale of the law (|P| per unit |grad u| near the reference state), so
            # that states at or near grad u = 0 compare against rounding of that size; formed
            # once per preparation
            k = lc.get('stiffness')
            if k is None:
                gen = torch.Generator().manual_seed(7)
                G1 = 1e-6 * (torch.rand((1, self.V, 3), generator=gen, dtype=torch.float64) - 0.5)
                k = lc['stiffness'] = float(vmap(law)(G1, *[t[:1] for t in ivs]).abs().max()) / float(G1.abs().max())
        scale = max(float(Pu.abs().max()), float(Pd.abs().max()), k * max(float(Gs.abs().max()), 1e-3), 1e-300)
        diff = float((Pu - Pd).abs().max())
        if diff <= 1e-9 * scale or not (diff == diff and scale < float('inf')):
            return   # agreement, or non-finite states that the solve itself reports
        logger.warning("the tensor map differs from %r at the quadrature states of this solve (relative "
                       "difference %.3g): switching to the tabulated law", self.material, diff / scale)
        self._mat_cache[lc['key']] = Tabulated(lc['fn'], self.V)
        self.prepare()

    def residual_unconstrained(self, u):
        """R(u) without Dirichlet rows, FP64 (problem.py:439-445)."""
        dm = self.dm
        self._check_law(u)
        self._tabulate(u, False)
        cell = self.cell_buffer()
        check(self.L.fem_cell_residual(dm.ref, self.mref, ptr(u), ptr(cell), stream_ptr()), 'fem_cell_residual')
        if self.cell_terms:
            cell[:dm.C * dm.N * self.V] += self._cell_terms_eval(u, False)[0].reshape(-1)
        base = self._base_load(u)
        res = dm.empty(self.n)
        check(self.L.fem_gather_nodal(dm.Nn, self.V, ptr(self.inc_ptr), ptr(self.inc_idx), ptr(cell),
                                      ptr(base), ptr(res), stream_ptr()), 'fem_gather_nodal')
        return res

    def cell_internal_force(self, u):
        """Per-cell internal force K_e(u) u_e as (C, N, V), FP64 (the laplace-kernel value of problem.py:189-214)."""
        dm = self.dm
        self._tabulate(u, False)
        cell = self.cell_buffer()
        check(self.L.fem_cell_residual(dm.ref, self.mref, ptr(u), ptr(cell), stream_ptr()), 'fem_cell_residual')
        return cell[:dm.C * dm.N * self.V].reshape(dm.C, dm.N, self.V).clone()

    def residual(self, u, scale=1.0):
        """R(u) with Dirichlet rows u - g (solver.py:290-304)."""
        return self.apply_residual_bc(u, self.residual_unconstrained(u), scale)

    def apply_residual_bc(self, u, res, scale=1.0):
        """res[d] = u[d] - scale g[d] on the Dirichlet dofs, in place (apply_bc_vec, solver.py:290-304)."""
        check(self.L.fem_residual_bc(self.n_bc, ptr(self.bc_dofs), ptr(self.bc_vals), ptr(u), ptr(res),
                                     float(scale), stream_ptr()), 'fem_residual_bc')
        return res

    def tangent(self, u, precision, ke_precision=None, residual=False, fused=None):
        """Assemble the row-eliminated tangent in `precision` (problem.py:447-460, solver.py:522-529).

        ke_precision: arithmetic of the per-cell tangents (default: FP64 for an FP32 matrix, else
        `precision`); FP64 cell tangents are summed in FP64 and rounded once into an FP32
        matrix.  FEM_BF16X3 (HEX27 elastic laws) forms them from split-BF16 MFMA products and
        stores them in FP32.  HEX8 problems with a built-in law assemble in one pass
        (fem_assemble: no per-cell tangent buffer); the rest through per-cell tangents and a
        gather.  With `residual`, R(u) without Dirichlet rows is formed too (from the same pass
        when it is single) and kept in self.last_residual.
        """
        dm = self.dm
        ke_precision = _lib.FEM_F64 if ke_precision is None and precision == _lib.FEM_F32 else \
            (precision if ke_precision is None else ke_precision)
        vals, dinv = self.matrix_buffers(precision)
        s = stream_ptr()
        # FP32 matrices are stored in difference form (exact on translations), FP64 in standard form
        form = _lib.FEM_FORM_DIFFERENCE if precision == _lib.FEM_F32 else _lib.FEM_FORM_STANDARD
        self._check_law(u)
        self._tabulate(u, True)
        if fused is None:
            fused = self.fused_applies(ke_precision)
        Ke, res = None, None
        if fused:
            plan = self.assembly_plan()
            base, res = None, None
            if residual:
                base, res = self._base_load(u), dm.empty(self.n)
            check(self.L.fem_assemble(dm.ref, self.mref, self.pref, C.byref(plan), ptr(u), precision, form,
                                      ptr(vals), ptr(dinv), ptr(base), ptr(res), s), 'fem_assemble')
        else:
            ke_storage = _lib.FEM_F32 if ke_precision == _lib.FEM_BF16X3 else ke_precision
            Ke = self.ke_buffer(ke_storage)
            check(self.L.fem_cell_tangent(dm.ref, self.mref, ke_precision, ptr(u), ptr(Ke), s), 'fem_cell_tangent')
            self._mass_tangent(u, Ke, ke_storage)
            self._cell_terms_tangent(u, Ke)
            check(self.L.fem_gather_matrix(self.pref, ke_storage, ptr(Ke), precision, form, ptr(vals), ptr(dinv),
                                           s), 'fem_gather_matrix')
            if residual:
                res = self.residual_unconstrained(u)
        check(self.L.fem_matrix_bc_rows(self.pref, precision, ptr(vals), self.n_bc, ptr(self.bc_dofs), ptr(dinv), s),
              'fem_matrix_bc_rows')
        self.current = (precision, Ke, vals, dinv)
        self.form = form
        self.last_residual = res
        return self.current

    def operator(self, u):
        op = _lib.FemOperator()
        op.mesh = C.pointer(self.dm.desc)
        op.mat = C.pointer(self.mat_desc)
        op.u = u.data_ptr()
        op.inc_ptr, op.inc_idx = self.inc_ptr.data_ptr(), self.inc_idx.data_ptr()
        op.n_bc = self.n_bc
        op.bc_dofs = self.bc_dofs.data_ptr() if self.n_bc else None
        op.cell_scratch = self.cell_buffer().data_ptr()
        # the mass term's dm/du is the one of the last tangent (the operator's linearisation point)
        op.mass_M = self.buffers['mass_M'].data_ptr() if self.mass_dyn is not None else None
        op.cell_K = self.buffers['cell_K'].data_ptr() if self.cell_terms else None
        self._op_keep = (op, u)
        return op

    def apply_operator(self, u, x):
        self._tabulate(u, True)
        if self.mass_dyn is not None:
            self._mass_eval(u, True)
        if self.cell_terms:
            self.buffers['cell_K'] = self._cell_terms_eval(u, True)[1].reshape(-1)
        op = self.operator(u)
        y = self.dm.empty(self.n)
        check(self.L.fem_operator_apply(C.byref(op), ptr(x), ptr(y), stream_ptr()), 'fem_operator_apply')
        return y

    def solve(self, u, b, x0, cfg):
        """Jacobi-PCG (or AMG-CG for cfg['method'] == 'amg') on the current tangent; returns (x, result struct)."""
        if cfg.get('method') == 'amg':
            return self.amg_solve(b, x0, cfg)
        if cfg.get('method') == 'bicgstab':
            return self.bicgstab(b, x0, cfg, constrained=cfg.get('constrained', False))
        precision, Ke, vals, dinv = self.current
        o = _lib.FemPcgOptions()
        o.rtol, o.atol, o.maxiter = cfg['tol'], cfg['atol'], cfg['maxiter']
        o.check_every, o.precision, o.precond = cfg['check_every'], precision, 1 if cfg['precond'] else 0
        o.reliable_delta, o.xcd_remap = cfg['reliable_delta'], 1 if cfg['xcd_remap'] else 0
        o.timing = 1 if cfg.get('timing', False) else 0
        key = ('pcg_ws', precision)
        if key not in self.buffers:
            nb = int(self.L.fem_pcg_ws_bytes(self.pref, precision))
            self.buffers[key] = self.dm.empty(nb, dtype=torch.uint8)
        ws = self.buffers[key]
        op = self.operator(u) if precision == _lib.FEM_F32 else None
        x = x0.clone()
        res = _lib.FemPcgResult()
        check(self.L.fem_pcg_solve(self.pref, ptr(vals), ptr(dinv), C.byref(op) if op is not None else None,
                                   ptr(b), ptr(x), C.byref(o), ptr(ws), C.c_void_p(self.host_status.data_ptr()),
                                   C.byref(res), stream_ptr()), 'fem_pcg_solve')
        return x, res

    def constraint(self):
        """Device CSR of the problem's P_mat and its transpose (solver.py:545-551), or None."""
        P = getattr(self.problem, 'P_mat', None)
        if P is None:
            return None
        key = id(P)
        if self._constraint is None or self._constraint[0] != key:
            import scipy.sparse as sp
            Pc = sp.csr_matrix(P)
            if Pc.shape[0] != self.n:
                raise ValueError(f"P_mat has {Pc.shape[0]} rows, the problem {self.n} dofs")
            Pc.sort_indices()
            Pt = Pc.T.tocsr()
            Pt.sort_indices()
            d = self.dm.device
            arrs = [torch.as_tensor(a, device=d) for a in (Pc.indptr.astype(onp.int32), Pc.indices.astype(onp.int32),
                                                            Pc.data.astype(onp.float64), Pt.indptr.astype(onp.int32),
                                                            Pt.indices.astype(onp.int32), Pt.data.astype(onp.float64))]
            c = _lib.FemConstraint()
            c.rows, c.cols = Pc.shape
            c.p_ptr, c.p_idx, c.p_val, c.pt_ptr, c.pt_idx, c.pt_val = [a.data_ptr() for a in arrs]
            self._constraint = (key, c, arrs)
        return self._constraint[1]

    def p_apply(self, x_r, transpose=False):
        """P x_r (full dofs), or P^T x with `transpose` (reduced dofs)."""
        c = self.constraint()
        arrs = self._constraint[2]
        rows = c.cols if transpose else c.rows
        ptr_, idx, val = (arrs[3], arrs[4], arrs[5]) if transpose else (arrs[0], arrs[1], arrs[2])
        y = self.dm.empty(int(rows))
        check(self.L.fem_csr_spmv(int(rows), ptr(ptr_), ptr(idx), ptr(val), ptr(x_r.contiguous()), ptr(y),
                                  stream_ptr()), 'fem_csr_spmv')
        return y

    def bicgstab(self, b, x0, cfg, constrained=False):
        """FP64 Jacobi-BiCGSTAB on the current FP64 tangent (jax_solve's method, solver.py:63-92);
        with `constrained`, on the reduced P^T A P of get_A (solver.py:545-551)."""
        precision, _, vals, dinv = self.current
        o = _lib.FemPcgOptions()
        o.rtol, o.atol, o.maxiter = cfg['tol'], cfg['atol'], cfg['maxiter']
        o.check_every, o.precision, o.precond = cfg['check_every'], precision, 1 if cfg['precond'] else 0
        cons = self.constraint() if constrained else None
        cref = C.byref(cons) if cons is not None else None
        nb = int(self.L.fem_bicgstab_ws_bytes(self.pref, cref))
        if self.buffers.get('bicgstab_ws_n') != nb:
            self.buffers['bicgstab_ws'] = self.dm.empty(nb, dtype=torch.uint8)
            self.buffers['bicgstab_ws_n'] = nb
        x = x0.clone()
        res = _lib.FemPcgResult()
        rc = self.L.fem_bicgstab_solve(self.pref, ptr(vals), ptr(dinv), cref, ptr(b), ptr(x), C.byref(o),
                                       ptr(self.buffers['bicgstab_ws']), C.c_void_p(self.host_status.data_ptr()),
                                       C.byref(res), stream_ptr())
        if rc not in (_lib.FEM_OK, _lib.FEM_ERR_NOT_CONVERGED, _lib.FEM_ERR_NAN):
            check(rc, 'fem_bicgstab_solve')
        return x, res

    @property
    def nonsymmetric(self):
        """True when a user map or weak form evaluated so far has a non-symmetric derivative."""
        pms = [getattr(self.material, '_pm', None)]
        if self.mass_dyn is not None:
            pms.append(self.mass_dyn['pm'])
        pms += [t['pm'] for t in self.cell_terms]
        return any(pm is not None and pm.nonsymmetric for pm in pms)

    def drop_amg_hierarchy(self):
        """Make the next AMG solve rebuild its numeric hierarchy (setup_reuse restarts)."""
        self._amg_uses = 1 << 30

    def amg_setup(self, cfg):
        """Smoothed-aggregation hierarchy of the current FP64 matrix."""
        precision, _, vals, _ = self.current
        if precision != _lib.FEM_F64:
            raise _lib.FemError("the AMG hierarchy is built from the FP64 matrix")
        if len(self.vecs) > 1 or self.V not in (1, 3):
            raise NotImplementedError("the AMG hierarchy takes one scalar or one vector variable (vec 1 or 3); "
                                      "solve several variables with 'jax_solver'")
        from .amg import DeviceAMG
        if self._amg is None:
            self._amg = DeviceAMG(self)
        return self._amg.setup(vals, cfg['amg_precision'], cfg['degree'], cfg.get('level0', 'fp32'))

    def amg_solve(self, b, x0, cfg):
        """AMG-CG on the current FP64 matrix.  With cfg['setup_reuse'] = k > 0 the
        hierarchy of an earlier matrix serves as the preconditioner of up to k
        further solves (the outer CG always uses the current matrix)."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        amg = self._amg
        reuse = (cfg.get('setup_reuse', 0) > 0 and amg is not None and amg.amg is not None
                 and amg.precision == cfg['amg_precision'] and self._amg_uses < cfg['setup_reuse']
                 and amg.level0_f16 == (cfg.get('level0', 'fp32') == 'fp16')
                 and amg.structure_current())
        if reuse:
            self._amg_uses += 1
        else:
            amg = self.amg_setup(cfg)
            self._amg_uses = 0
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        out = amg.solve(self.current[2], b, x0, cfg)
        self.amg_timing = dict(setup_s=t1 - t0, solve_s=time.perf_counter() - t1)
        return out

    def spmv(self, x):
        """y = A x with the current matrix, in its precision and storage form."""
        prec, _, vals, _ = self.current
        y = torch.empty_like(x)
        check(self.L.fem_spmv(self.pref, prec, self.form, ptr(vals), ptr(x), ptr(y), stream_ptr()), 'fem_spmv')
        return y

    def dot(self, x, y):
        if 'dot_ws' not in self.buffers:
            self.buffers['dot_ws'] = self.dm.empty(8192, dtype=torch.uint8)
            self.buffers['dot_out'] = self.dm.empty(1)
        out = self.buffers['dot_out']
        check(self.L.fem_dot(x.numel(), ptr(x), ptr(y), ptr(out), ptr(self.buffers['dot_ws']), stream_ptr()),
              'fem_dot')
        return float(out.item())

    def norm(self, x):
        return self.dot(x, x) ** 0.5

    def param_grad(self, u, lam):
        """-lambda^T dR/dtheta per (cell, quad) for the material parameter (solver.py:1386-1416)."""
        dm = self.dm
        g = dm.empty(dm.C, dm.Q)
        check(self.L.fem_cell_adjoint(dm.ref, self.mref, ptr(u), ptr(lam), ptr(g), stream_ptr()),
              'fem_cell_adjoint')
        return g

    def load_adjoint(self, lam):
        dm = self.dm
        lq = dm.empty(dm.C, dm.Q, self.V)
        check(self.L.fem_cell_load_adjoint(dm.ref, ptr(lam), ptr(lq), stream_ptr()), 'fem_cell_load_adjoint')
        return lq

    def face_load_adjoint(self, i, lam):
        dm = self.dm
        srf = self.surfaces[i]
        S = srf['faces'].shape[0]
        lq = dm.empty(S, dm.FQ, self.V)
        if S == 0:
            return lq
        check(self.L.fem_face_load_adjoint(dm.ref, S, ptr(srf['faces']), ptr(lam), ptr(lq), stream_ptr()),
              'fem_face_load_adjoint')
        return lq

    # ------------------------------------------------------------- export
    def export_csr(self):
        """Scalar CSR (indptr, indices, data) of the current matrix on the host."""
        prec, _, vals, _ = self.current
        nnz = self.nnzb * self.V * self.V
        indptr = self.dm.empty(self.n + 1, dtype=torch.int32)
        indices = self.dm.empty(max(nnz, 1), dtype=torch.int32)
        data = self.dm.empty(max(nnz, 1))
        check(self.L.fem_export_csr(self.pref, prec, self.form, ptr(vals), ptr(indptr), ptr(indices), ptr(data),
                                    stream_ptr()),
              'fem_export_csr')
        return indptr.cpu().numpy(), indices[:nnz].cpu().numpy(), data[:nnz].cpu().numpy()

    def cell_tangents(self, u, precision=_lib.FEM_F64, faces=True):
        """Per-cell tangents (C, N*V, N*V) in the reference's jac layout (problem.py:262-266, 453).
        With faces=False the per-face terms of u-dependent surface maps and surface kernels are
        left out (the reference keeps them as separate face blocks, problem.py:457-458)."""
        dm = self.dm
        dt = torch.float32 if precision in (_lib.FEM_F32, _lib.FEM_BF16X3) else torch.float64
        Ke = dm.empty(dm.C * dm.N * dm.N * self.V * self.V, dtype=dt)
        self._tabulate(u, True)
        check(self.L.fem_cell_tangent(dm.ref, self.mref, precision, ptr(u), ptr(Ke), stream_ptr()),
              'fem_cell_tangent')
        self._mass_tangent(u, Ke, _lib.FEM_F32 if dt == torch.float32 else _lib.FEM_F64)
        self._cell_terms_tangent(u, Ke, faces)
        N, V = dm.N, self.V
        K = Ke.reshape(dm.C, N, N, V, V).permute(0, 1, 3, 2, 4).reshape(dm.C, N * V, N * V)
        if len(self.vecs) > 1:    # the reference's per-cell order (variables concatenated)
            p = torch.as_tensor(self.cperm, device=K.device)
            K = K[:, p][:, :, p]
        return K
