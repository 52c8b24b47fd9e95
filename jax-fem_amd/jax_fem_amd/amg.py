"""Device smoothed-aggregation AMG: hierarchy setup and AMG-preconditioned CG.

Device counterpart of jax-fem's `amgx_solver` option (solver.py:141-259,
selected in linear_solver, solver.py:261-284).  The reference hands the
assembled PETSc matrix to AMGX; here every stage runs in libfem_mi355x on the
GPU and this module only sequences the calls, sizing the buffers between them
(the library allocates nothing):

  level 0   the assembled FP64 SELL matrix (row-eliminated, standard form) and
            its block CSR copy; the FE block pattern is the aggregation graph
  per level dinv -> MIS-2 aggregates -> tentative P0 (rigid-body modes or
            constants, Dirichlet rows zero) -> lambda_max(D^-1 A) -> A P0 ->
            P = P0 - omega D^-1 A P0 -> R = P^T -> Ac = R (A P)
  coarsest  dense Gauss-Jordan inverse (<= 512 unknowns, <= 1024 on large problems)

The structure (aggregates, tentative prolongators, patterns, product plans)
is built once per pattern and Dirichlet set; each new matrix then costs the
numeric passes only.

`precision='mixed'` stores the cycle's matrices (level-0 SELL copy, A, P, R)
in FP32; the outer CG keeps the FP64 operator, so the solution accuracy is
that of the FP64 solve.
"""
import ctypes as C
import logging

import torch

from . import _lib
from ._lib import check, ptr, stream_ptr

logger = logging.getLogger('jax_fem_amd')

COARSE_MAX = 1024   # unknowns of the coarsest (dense) level: at most this many,
COARSE_SMALL = 512  # and at most this many where coarse_limit says so


def coarse_limit(nnz0):
    """Largest dense coarsest level for a level-0 operator of nnz0 entries: 512 unknowns,
    and up to 1024 while n^3 <= 8 nnz0.  A dense level of n unknowns costs a blocked
    inverse per setup (about 0.1 ms per 64 unknowns) and saves the smoothed level it
    replaces (a dozen small launches per V-cycle): that pays where the fine operator, and
    with it the iteration count, is large (the 1.6M-dof cantilever: 882 unknowns dense,
    CG 39 -> 37 iterations), not on small problems that converge in a few iterations."""
    n = COARSE_SMALL
    while n < COARSE_MAX and (n + 1) ** 3 <= 8 * nnz0:
        n += 1
    return n
MAX_LEVELS = 10
POWER_STEPS = 15
DEGREE = 2          # Chebyshev degree of pre- and post-smoothing


class Bcsr:
    """Block CSR matrix held in device tensors."""

    def __init__(self, rows, cols, br, bc, row_ptr, col_idx, vals):
        self.rows, self.cols, self.br, self.bc = int(rows), int(cols), int(br), int(bc)
        self.row_ptr, self.col_idx, self.vals = row_ptr, col_idx, vals
        self.nnzb = int(col_idx.numel()) if col_idx is not None else 0

    def struct(self, vals=None):
        v = self.vals if vals is None else vals
        s = _lib.FemBcsr()
        s.rows, s.cols, s.nnzb, s.br, s.bc = self.rows, self.cols, self.nnzb, self.br, self.bc
        s.f32 = 1 if v is not None and v.dtype == torch.float32 else 0
        s.row_ptr = self.row_ptr.data_ptr()
        s.col_idx = self.col_idx.data_ptr() if self.nnzb else None
        s.vals = v.data_ptr() if self.nnzb else None
        return s

    def host(self):
        """(row_ptr, col_idx, vals (nnzb, br, bc)) as numpy arrays (tests)."""
        return (self.row_ptr.cpu().numpy(), self.col_idx.cpu().numpy(),
                self.vals.reshape(-1, self.br, self.bc).cpu().numpy())


class DeviceAMG:
    """Hierarchy of one DeviceProblem's current FP64 matrix."""

    def __init__(self, dev):
        self.dev = dev
        self.L = dev.L
        self._mask = None
        self._f32 = {}
        self._vbuf = {}
        self.vals32 = None
        self._diff_fresh = False
        self.levels = []
        self.amg = None
        self.level0_f16 = False

    # -------------------------------------------------------------- helpers
    def _empty(self, n, dtype=torch.float64):
        return torch.empty(max(int(n), 1), dtype=dtype, device=self.dev.dm.device)

    def _scan(self, counts):
        return self.dev.dm.scan(counts)

    def _aggregate(self, n, row_ptr, col_idx):
        ws = self._empty(self.L.fem_amg_aggregate_ws_bytes(n), torch.uint8)
        agg = self._empty(n, torch.int32)
        n_agg = C.c_int64(0)
        check(self.L.fem_amg_aggregate(n, ptr(row_ptr), ptr(col_idx), ptr(agg), C.byref(n_agg), ptr(ws),
                                       C.c_void_p(self.dev.host_status.data_ptr()), stream_ptr()),
              'fem_amg_aggregate')
        return agg[:n], int(n_agg.value)

    def _members(self, n, agg, n_agg):
        agg_ptr = self._empty(n_agg + 1, torch.int32)
        memb = self._empty(n, torch.int32)
        nb = ((4 * max(n_agg, 1) + 255) // 256) * 256 + int(self.L.fem_scan_ws_bytes(n_agg))
        ws = self._empty(nb, torch.uint8)
        check(self.L.fem_amg_members(n, ptr(agg), n_agg, ptr(agg_ptr), ptr(memb), ptr(ws), stream_ptr()),
              'fem_amg_members')
        return agg_ptr, memb

    def _spgemm_plan(self, X, Y):
        """Structure of C = X Y and its product plan (once per pattern)."""
        xs, ys = X.struct(), Y.struct()
        counts = self._empty(X.rows, torch.int32)
        pairs = self._empty(X.rows, torch.int32)
        mp = C.c_int64(0)
        ws = self._empty(256, torch.uint8)
        hs = C.c_void_p(self.dev.host_status.data_ptr())
        check(self.L.fem_bcsr_spgemm_symbolic(C.byref(xs), C.byref(ys), ptr(counts), ptr(pairs), C.byref(mp),
                                              ptr(ws), hs, stream_ptr()), 'fem_bcsr_spgemm_symbolic')
        rp = self._scan(counts[:X.rows])
        pair_ptr = self._scan(pairs[:X.rows])
        nnzb = int(rp[-1].item())
        total = int(pair_ptr[-1].item())
        out = Bcsr(X.rows, Y.cols, X.br, Y.bc, rp, self._empty(nnzb, torch.int32)[:nnzb],
                   self._empty(nnzb * X.br * Y.bc)[:nnzb * X.br * Y.bc])
        plan_ptr = self._empty(nnzb + 1, torch.int32)
        plan = self._empty(2 * total, torch.int32)
        cs = out.struct()
        check(self.L.fem_bcsr_spgemm_plan(C.byref(xs), C.byref(ys), mp.value, C.byref(cs), ptr(pair_ptr),
                                          ptr(plan_ptr), ptr(plan), ptr(ws), hs, stream_ptr()),
              'fem_bcsr_spgemm_plan')
        return out, (plan_ptr, plan)

    def _spgemm_numeric(self, X, Y, out, plan):
        xs, ys, cs = X.struct(), Y.struct(), out.struct()
        check(self.L.fem_bcsr_spgemm_numeric(C.byref(xs), C.byref(ys), C.byref(cs), ptr(plan[0]), ptr(plan[1]),
                                             stream_ptr()), 'fem_bcsr_spgemm_numeric')

    def _transpose_structure(self, X):
        T = Bcsr(X.cols, X.rows, X.bc, X.br, self._empty(X.cols + 1, torch.int32),
                 self._empty(X.nnzb, torch.int32)[:X.nnzb], self._empty(X.nnzb * X.br * X.bc))
        src = self._empty(X.nnzb, torch.int32)
        nb = ((4 * max(X.cols, 1) + 255) // 256) * 256 + int(self.L.fem_scan_ws_bytes(X.cols))
        ws = self._empty(nb, torch.uint8)
        xs, ts = X.struct(), T.struct()
        check(self.L.fem_bcsr_transpose_structure(C.byref(xs), C.byref(ts), ptr(src), ptr(ws), stream_ptr()),
              'fem_bcsr_transpose_structure')
        return T, src

    def _transpose_values(self, X, T, src):
        xs, ts = X.struct(), T.struct()
        check(self.L.fem_bcsr_transpose_values(C.byref(xs), C.byref(ts), ptr(src), stream_ptr()),
              'fem_bcsr_transpose_values')

    def _dinv(self, A, fix):
        d = self._empty(A.rows * A.br)
        s = A.struct()
        check(self.L.fem_amg_dinv(C.byref(s), 1 if fix else 0, ptr(d), stream_ptr()), 'fem_amg_dinv')
        return d

    def _lmax(self, A, dinv, sell_vals=None):
        """Power estimate of lambda_max(D^-1 A); level 0 runs on the SELL matrix (the FP32
        difference-form copy for an FP32 hierarchy, which halves the bytes of the 15 products)."""
        s = A.struct()
        ws = self._empty(self.L.fem_amg_lmax_ws_bytes(A.rows * A.br), torch.uint8)
        lam = C.c_double(0.0)
        pat = self.dev.pref if sell_vals is not None else None
        f32 = 0 if sell_vals is None else {torch.float32: 1, torch.float16: 2}.get(sell_vals.dtype, 0)
        check(self.L.fem_amg_lmax(C.byref(s), pat, ptr(sell_vals), f32, ptr(dinv), POWER_STEPS, C.byref(lam),
                                  ptr(ws), stream_ptr()), 'fem_amg_lmax')
        return float(lam.value)

    def _lmax_async(self, A, dinv, out, sell_vals=None):
        """_lmax's estimate written to the device scalar `out` (a 1-element FP64 tensor) without
        a host synchronisation."""
        s = A.struct()
        ws = self._empty(self.L.fem_amg_lmax_ws_bytes(A.rows * A.br), torch.uint8)
        pat = self.dev.pref if sell_vals is not None else None
        f32 = 0 if sell_vals is None else {torch.float32: 1, torch.float16: 2}.get(sell_vals.dtype, 0)
        check(self.L.fem_amg_lmax_async(C.byref(s), pat, ptr(sell_vals), f32, ptr(dinv), POWER_STEPS, ptr(out),
                                        ptr(ws), stream_ptr()), 'fem_amg_lmax_async')

    def _sell_diff32_buffer(self, vals64):
        dev = self.dev
        n = dev.sell_blocks * dev.V * dev.V
        out = self._f32.get('sell')
        if out is None or out.numel() != n:
            out = torch.empty(n, dtype=torch.float32, device=vals64.device)
            self._f32['sell'] = out
        return out

    def _sell_diff32(self, vals64):
        """FP32 difference-form copy of the FP64 SELL matrix (V-cycle level 0 and outer CG)."""
        out = self._sell_diff32_buffer(vals64)
        check(self.L.fem_amg_sell_to_f32_diff(self.dev.pref, ptr(vals64), ptr(out), stream_ptr()),
              'fem_amg_sell_to_f32_diff')
        return out

    def _sell_f16(self, vals64, dinv):
        """FP16 copy of D^-1 A in difference form (the level-0 smoother matrix of level0='fp16')."""
        dev = self.dev
        n = dev.sell_blocks * dev.V * dev.V
        out = self._f32.get('sell16')
        if out is None or out.numel() != n:
            out = torch.empty(n, dtype=torch.float16, device=vals64.device)
            self._f32['sell16'] = out
        check(self.L.fem_amg_sell_to_f16_scaled(dev.pref, ptr(vals64), ptr(dinv), ptr(out), stream_ptr()),
              'fem_amg_sell_to_f16_scaled')
        return out

    def _values(self, M, dtype):
        """Point M.vals at M's value buffer of `dtype`: the structure allocates FP64 values; an FP32
        hierarchy (FP32 Galerkin products) keeps a second, FP32 buffer, allocated on first use."""
        bufs = M.__dict__.setdefault('_bufs', {})
        if M.vals is not None:
            bufs.setdefault(M.vals.dtype, M.vals)
        if dtype not in bufs:
            n = M.nnzb * M.br * M.bc
            bufs[dtype] = self._empty(n, dtype)[:n]
        M.vals = bufs[dtype]
        return M.vals

    def _to_f32(self, t, key):
        out = self._f32.get(key)
        if out is None or out.numel() != t.numel():
            out = torch.empty(t.numel(), dtype=torch.float32, device=t.device)
            self._f32[key] = out
        check(self.L.fem_amg_to_f32(t.numel(), ptr(t), ptr(out), stream_ptr()), 'fem_amg_to_f32')
        return out

    # --------------------------------------------------------------- setup
    def structure_current(self):
        """True when the cached structure matches the problem's pattern and Dirichlet set."""
        dev = self.dev
        src = getattr(self, '_struct_src', None)
        return src is not None and src[0] is dev.pat_tensors['row_ptr'] and (
            src[1] is dev.bc_dofs or (src[1].shape == dev.bc_dofs.shape and torch.equal(src[1], dev.bc_dofs)))

    def _structure(self):
        """Value-independent part of the hierarchy, kept while the pattern and
        the Dirichlet set are unchanged: aggregates, tentative prolongators,
        the patterns of P, R, A P and R (A P), their product plans and the
        transpose map.  Level 0's graph is the FE block pattern."""
        dev, L = self.dev, self.L
        if self.structure_current():
            return self._struct     # same pattern and Dirichlet set (values may differ)
        V = dev.V
        ns = 6 if V == 3 else 1
        self._mask = self._empty(dev.n, torch.uint8)
        check(L.fem_amg_bc_mask(dev.n, dev.n_bc, ptr(dev.bc_dofs) if dev.n_bc else None, ptr(self._mask),
                                stream_ptr()), 'fem_amg_bc_mask')
        pt = dev.pat_tensors
        A = Bcsr(dev.dm.Nn, dev.dm.Nn, V, V, pt['row_ptr'], pt['col_idx'], self._empty(dev.nnzb * V * V))
        B, bs = None, V
        levels = []
        n_dense = coarse_limit(dev.nnzb * V * V)
        while A.rows * A.br > n_dense and len(levels) < MAX_LEVELS - 1:
            agg, n_agg = self._aggregate(A.rows, A.row_ptr, A.col_idx)
            if n_agg >= A.rows:
                break
            agg_ptr, memb = self._members(A.rows, agg, n_agg)
            P0 = self._empty(A.rows * bs * ns)
            Bc = self._empty(n_agg * ns * ns)
            check(L.fem_amg_tentative(A.rows, bs, ns, n_agg, ptr(agg_ptr), ptr(memb),
                                      ptr(dev.dm.points) if B is None else None,
                                      ptr(self._mask) if B is None else None, ptr(B), ptr(P0), ptr(Bc),
                                      stream_ptr()), 'fem_amg_tentative')
            P0m = Bcsr(A.rows, n_agg, bs, ns, torch.arange(A.rows + 1, dtype=torch.int32, device=P0.device),
                       agg, P0[:A.rows * bs * ns])
            P, plan_w0 = self._spgemm_plan(A, P0m)          # A P0 on the pattern of P
            R, src = self._transpose_structure(P)
            W, plan_w = self._spgemm_plan(A, P)
            Ac, plan_c = self._spgemm_plan(R, W)
            levels.append(dict(A=A, agg=agg, n_agg=n_agg, P0=P0, P0m=P0m, Bc=Bc, P=P, R=R, R_src=src, W=W,
                               plans=(plan_w0, plan_w, plan_c)))
            A, B, bs = Ac, Bc, ns
        n_c = A.rows * A.br
        if n_c > 4096:
            raise _lib.FemError(f"AMG coarsening stalled at {n_c} unknowns (the dense coarse solve takes <= 4096)")
        self._struct = (levels, A)
        self._struct_src = (dev.pat_tensors['row_ptr'], dev.bc_dofs)
        return self._struct

    def setup(self, vals64, precision=_lib.FEM_F64, degree=DEGREE, level0='fp32'):
        """Hierarchy of the current FP64 matrix (standard-form SELL values `vals64`).
        level0 ('fp32' or 'fp16', FP32 hierarchies only): storage of the level-0
        smoother matrix, FP32 difference form or FP16 D^-1 A in difference form."""
        dev, L = self.dev, self.L
        levels, coarse = self._structure()
        # an FP32 hierarchy forms its Galerkin products from FP32 values (FP64 arithmetic):
        # level 0's block-CSR copy, P, R, A P and the coarse operators are FP32
        dt = torch.float32 if precision == _lib.FEM_F32 else torch.float64
        for lv in levels:
            for name in ('A', 'P', 'R', 'W'):
                self._values(lv[name], dt)
        # a matrix without levels is its own coarse matrix: its dense inverse is formed from the
        # FP64 values, so the cycle is the exact solve in either precision
        self._values(coarse, dt if levels else torch.float64)
        A0 = levels[0]['A'] if levels else coarse   # no levels: the matrix is the coarsest level
        prec0 = _lib.FEM_F32 if A0.vals.dtype == torch.float32 else _lib.FEM_F64
        f16 = level0 == 'fp16' and precision == _lib.FEM_F32 and bool(levels)
        sv32 = dinv0 = None
        if precision == _lib.FEM_F32 and levels and not f16:
            # level 0's FP32 block-CSR copy (the Galerkin products' input), its FP32
            # difference-form SELL copy (the V-cycle's level-0 matrix) and its inverse
            # diagonal in one pass
            sv32 = self._sell_diff32_buffer(vals64)
            dinv0 = self._empty(A0.rows * A0.br)
            check(L.fem_amg_sell_to_f32_copies(dev.pref, ptr(vals64), ptr(A0.vals), ptr(sv32), ptr(dinv0),
                                               stream_ptr()), 'fem_amg_sell_to_f32_copies')
        else:
            check(L.fem_amg_sell_to_bcsr(dev.pref, ptr(vals64), ptr(A0.vals), prec0, stream_ptr()),
                  'fem_amg_sell_to_bcsr')
        # each level's lambda_max estimate stays on the device (the prolongator smoothing forms
        # omega = 1.5 / lambda_max there); the host reads them all back once, after the loop
        lam_dev = self._empty(max(len(levels), 1))
        for l, lv in enumerate(levels):
            A = lv['A']
            lv['dinv'] = dinv0 if (l == 0 and dinv0 is not None) else self._dinv(A, fix=l > 0)
            if l == 0 and f16:
                sv32 = self._sell_f16(vals64, lv['dinv'])
            lam_l = lam_dev[l:l + 1]
            self._lmax_async(A, lv['dinv'], lam_l, (vals64 if sv32 is None else sv32) if l == 0 else None)
            plan_w0, plan_w, plan_c = lv['plans']
            P = lv['P']
            as_, p0s, ps = A.struct(), lv['P0m'].struct(), P.struct()
            check(L.fem_amg_smooth_prolongator_dev(C.byref(as_), C.byref(p0s), C.byref(ps), ptr(plan_w0[0]),
                                                   ptr(plan_w0[1]), ptr(lv['agg']), ptr(lv['dinv']), ptr(lam_l),
                                                   stream_ptr()), 'fem_amg_smooth_prolongator_dev')
            self._transpose_values(P, lv['R'], lv['R_src'])
            self._spgemm_numeric(A, P, lv['W'], plan_w)
            nxt = levels[l + 1]['A'] if l + 1 < len(levels) else coarse
            self._spgemm_numeric(lv['R'], lv['W'], nxt, plan_c)
        if levels:
            for lv, lam in zip(levels, lam_dev.cpu().tolist()):
                lv['lam'] = lam
        self._dinv(coarse, fix=True)
        n_c = coarse.rows * coarse.br
        if getattr(self, '_coarse_n', None) != n_c:
            self.coarse_inv = self._empty(n_c * n_c)
            self._gj_ws = self._empty(n_c * n_c)
            self._coarse_n = n_c
        cs = coarse.struct()
        check(L.fem_amg_coarse_inverse(C.byref(cs), ptr(self.coarse_inv), ptr(self._gj_ws), stream_ptr()),
              'fem_amg_coarse_inverse')
        self.levels, self.coarse = levels, coarse
        self._descriptor(vals64, precision, degree, sv32)
        return self

    def _descriptor(self, vals64, precision, degree, sv32=None):
        dev = self.dev
        f32 = precision == _lib.FEM_F32
        keep = []
        arr = (_lib.FemAmgLevel * len(self.levels))()
        self.vals32, self._diff_fresh = None, False
        for l, lv in enumerate(self.levels):
            e = arr[l]
            if l == 0:
                e.pat = C.pointer(dev.pat)
                sv = (sv32 if sv32 is not None else self._sell_diff32(vals64)) if f32 else vals64
                keep.append(sv)
                e.sell_vals = sv.data_ptr()
                self.vals32 = sv if f32 else None
                self._diff_fresh = f32
            for name in ('A', 'P', 'R'):
                M = lv[name]
                v = M.vals     # in the hierarchy's precision
                keep.append(v)
                setattr(e, name, M.struct(v))
            e.dinv = lv['dinv'].data_ptr()
            e.lmax = lv['lam']
        amg = _lib.FemAmg()
        amg.num_levels = len(self.levels)
        amg.precision = _lib.FEM_F32 if f32 else _lib.FEM_F64
        amg.degree = degree
        self.level0_f16 = sv32 is not None and sv32.dtype == torch.float16
        amg.level0_form = (_lib.FEM_FORM_SCALED_DIFF16 if self.level0_f16 else _lib.FEM_FORM_DIFFERENCE) \
            if (f32 and self.levels) else _lib.FEM_FORM_STANDARD
        amg.levels = arr if self.levels else None
        amg.coarse_n = self.coarse.rows * self.coarse.br
        amg.coarse_inv = self.coarse_inv.data_ptr()
        self._keep = (keep, arr)
        self.amg = amg
        self.precision = amg.precision

    def describe(self):
        return [dict(rows=lv['A'].rows * lv['A'].br, nnz=lv['A'].nnzb * lv['A'].br * lv['A'].bc,
                     aggregates=lv['n_agg'], lmax=lv['lam'], p_nnz=lv['P'].nnzb * lv['P'].br * lv['P'].bc)
                for lv in self.levels] + [dict(rows=self.coarse.rows * self.coarse.br)]

    # --------------------------------------------------------------- apply
    def vcycle(self, r):
        """z = M r (one V-cycle)."""
        z = torch.empty_like(r)
        ws = self._empty(self.L.fem_amg_ws_bytes(C.byref(self.amg)), torch.uint8)
        check(self.L.fem_amg_vcycle(C.byref(self.amg), ptr(r), ptr(z), ptr(ws), stream_ptr()), 'fem_amg_vcycle')
        return z

    def solve(self, vals64, b, x0, cfg):
        """AMG-preconditioned CG on the FP64 operator; returns (x, result struct)."""
        dev = self.dev
        o = _lib.FemAmgOptions()
        o.rtol, o.atol, o.maxiter, o.check_every = cfg['tol'], cfg['atol'], cfg['maxiter'], cfg['check_every']
        o.timing = 1 if cfg.get('timing', False) else 0
        o.reliable_delta = cfg.get('reliable_delta', 0.01)
        nb = int(self.L.fem_amg_solve_ws_bytes(dev.pref, C.byref(self.amg)))
        key = ('amg_ws', nb)
        if dev.buffers.get('amg_ws_key') != key:
            dev.buffers['amg_ws'] = self._empty(nb, torch.uint8)
            dev.buffers['amg_ws_key'] = key
        if self.vals32 is not None and not self._diff_fresh:
            # a reused hierarchy: level 0 and the outer operator follow the matrix
            if self.level0_f16:
                self._sell_f16(vals64, self.levels[0]['dinv'])
            else:
                self._sell_diff32(vals64)
        self._diff_fresh = False
        x = x0.clone()
        res = _lib.FemPcgResult()
        outer32 = self.vals32 if (cfg.get('outer', 'fp64') == 'fp32' and not self.level0_f16) else None
        check(self.L.fem_amg_pcg_solve(dev.pref, ptr(vals64), ptr(outer32), C.byref(self.amg), ptr(b), ptr(x),
                                       C.byref(o),
                                       ptr(dev.buffers['amg_ws']), C.c_void_p(dev.host_status.data_ptr()),
                                       C.byref(res), stream_ptr()), 'fem_amg_pcg_solve')
        return x, res
