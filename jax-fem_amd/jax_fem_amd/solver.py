"""Newton solver, linear-solver options and the adjoint wrapper (jax_fem/solver.py).

``solver(problem, solver_options)`` and ``ad_wrapper(problem, solver_options,
adjoint_solver_options)`` keep the reference's signatures, option layout
(solver.py:1052-1106) and Newton logic (solver.py:1270-1356).  The linear
solve is Jacobi-PCG on the device (csrc/pcg.hip) in place of JAX BiCGSTAB;
``ad_wrapper`` returns a function whose output carries a
``torch.autograd.Function`` backward: the transposed tangent system is solved
for lambda and -lambda^T dR/dtheta is formed per quadrature point by the
adjoint kernel (solver.py:1362-1455).
"""
import logging
import math
import time

import numpy as onp
import torch

from . import _lib

logger = logging.getLogger('jax_fem_amd')

_METHOD_KEYS = frozenset({'newton', 'arc_length', 'dynamic_relax'})
_LINEAR_OPTION_KEYS = frozenset({'jax_solver', 'amgx_solver', 'spsolve_solver', 'petsc_solver', 'custom_solver'})
_NEWTON_OPTION_KEYS = frozenset({'tol', 'rel_tol', 'line_search_flag', 'initial_guess'})

# Keys of {'jax_solver': {...}}.  `precond` is the reference's; the rest select
# the device PCG: tol/atol/maxiter default to jax_solve's (solver.py:78-84),
# precision 'fp64' (plain FP64 PCG) or 'mixed' (FP32 matrix and PCG, FP64
# refinement residual and iterate).
JAX_SOLVER_DEFAULTS = dict(precond=True, precision='fp64', tol=1e-10, atol=1e-10, maxiter=10000,
                           check_every=16, reliable_delta=0.01, xcd_remap=True, assembly=None, timing=False,
                           krylov='auto')
_ASSEMBLY = {None: None, 'fp32': _lib.FEM_F32, 'fp64': _lib.FEM_F64, 'bf16x3': _lib.FEM_BF16X3}
_PRECISIONS = {'fp64': _lib.FEM_F64, 'mixed': _lib.FEM_F32, 'fp32': _lib.FEM_F32}

# Keys of {'amgx_solver': {...}} (solver.py:251-259 reads only `cfg_path`, an
# AMGX JSON file).  The device path builds a smoothed-aggregation hierarchy and
# runs AMG-preconditioned FP64 CG; `precision` 'mixed' stores the hierarchy in
# FP32; `setup_reuse` = k lets one hierarchy precondition k further solves
# (Newton steps).  From an AMGX config file only solver.tolerance and
# solver.max_iters are taken (as atol and maxiter, AMGX's ABSOLUTE convergence).
AMG_SOLVER_DEFAULTS = dict(cfg_path=None, precision='fp64', tol=1e-10, atol=1e-10, maxiter=1000, check_every=4,
                           degree=2, timing=False, setup_reuse=0, outer='fp64', reliable_delta=0.01, level0='fp32')


def _resolve_solver_options(solver_options):
    """(method, cfg); legacy flat dicts become Newton (solver.py:1088-1106)."""
    opts = solver_options or {}
    methods = [m for m in _METHOD_KEYS if m in opts]
    if not methods:
        linear = {k: opts[k] for k in _LINEAR_OPTION_KEYS if k in opts}
        cfg = {k: opts[k] for k in _NEWTON_OPTION_KEYS if k in opts}
        if linear:
            cfg['linear'] = linear
        return 'newton', cfg
    if len(methods) > 1:
        raise ValueError(f"Pick one nonlinear method, got {methods}.")
    method = methods[0]
    if not isinstance(opts[method], dict):
        raise ValueError(f"solver_options['{method}'] must be a dict.")
    return method, opts[method]


def linear_config(linear_options):
    """Device PCG configuration from a linear-options dict (solver.py:261-284)."""
    linear_options = linear_options or {}
    present = linear_options.keys() & _LINEAR_OPTION_KEYS
    if not present:
        user = {}
    elif 'jax_solver' in linear_options:
        user = dict(linear_options['jax_solver'] or {})
    elif 'amgx_solver' in linear_options:
        return _amg_config(dict(linear_options['amgx_solver'] or {}))
    else:
        raise NotImplementedError(
            f"linear solver {sorted(present)} is not available on the device path; "
            "use 'jax_solver' or 'amgx_solver'")
    unknown = set(user) - set(JAX_SOLVER_DEFAULTS)
    if unknown:
        raise ValueError(f"unknown jax_solver options {sorted(unknown)}")
    cfg = dict(JAX_SOLVER_DEFAULTS)
    cfg.update(user)
    if cfg['precision'] not in _PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}, got {cfg['precision']!r}")
    cfg['precision_code'] = _PRECISIONS[cfg['precision']]
    if cfg['assembly'] not in _ASSEMBLY:
        raise ValueError(f"assembly must be 'fp32', 'fp64' or 'bf16x3', got {cfg['assembly']!r}")
    cfg['assembly_code'] = _ASSEMBLY[cfg['assembly']]
    if cfg['krylov'] not in ('auto', 'cg', 'bicgstab'):
        raise ValueError(f"krylov must be 'auto', 'cg' or 'bicgstab', got {cfg['krylov']!r}")
    cfg['method'] = 'bicgstab' if cfg['krylov'] == 'bicgstab' else 'pcg'
    return cfg


def _amg_config(user):
    unknown = set(user) - set(AMG_SOLVER_DEFAULTS)
    if unknown:
        raise ValueError(f"unknown amgx_solver options {sorted(unknown)}")
    cfg = dict(AMG_SOLVER_DEFAULTS)
    cfg.update(user)
    if cfg['cfg_path'] is not None:
        import json
        with open(cfg['cfg_path']) as f:
            amgx = json.load(f).get('solver', {})
        if 'tolerance' in amgx:
            cfg['atol'], cfg['tol'] = float(amgx['tolerance']), 0.0
        if 'max_iters' in amgx:
            cfg['maxiter'] = int(amgx['max_iters'])
    if cfg['precision'] not in ('fp64', 'mixed'):
        raise ValueError(f"amgx_solver precision must be 'fp64' or 'mixed', got {cfg['precision']!r}")
    if cfg['outer'] not in ('fp64', 'fp32') or (cfg['outer'] == 'fp32' and cfg['precision'] != 'mixed'):
        raise ValueError("amgx_solver outer must be 'fp64', or 'fp32' with precision 'mixed'")
    if cfg['level0'] not in ('fp32', 'fp16') or (cfg['level0'] == 'fp16' and
                                                (cfg['precision'] != 'mixed' or cfg['outer'] != 'fp64')):
        raise ValueError("amgx_solver level0 must be 'fp32', or 'fp16' with precision 'mixed' and outer 'fp64'")
    cfg['method'] = 'amg'
    cfg['amg_precision'] = _lib.FEM_F32 if cfg['precision'] == 'mixed' else _lib.FEM_F64
    cfg['precision_code'] = _lib.FEM_F64      # the outer CG operator is the FP64 matrix
    cfg['assembly_code'] = None
    cfg['precond'] = True
    return cfg


def linear_solve(dev, u, b, x0, cfg):
    """Solve the current tangent system; mirrors jax_solve's post-check (solver.py:86-90).

    A tangent that is not symmetric (a user map or weak form whose derivative is
    not) is solved by FP64 Jacobi-BiCGSTAB, jax_solve's own method; CG and
    AMG-CG need symmetry."""
    if cfg.get('method') != 'bicgstab' and dev.nonsymmetric:
        if cfg.get('krylov', 'auto') == 'cg':
            raise ValueError("the tangent is not symmetric; 'krylov': 'cg' does not apply")
        logger.info("the tangent is not symmetric: FP64 Jacobi-BiCGSTAB")
        cfg = dict(cfg, method='bicgstab', precision='fp64', precision_code=_lib.FEM_F64, precond=True,
                   check_every=cfg.get('check_every', 16), xcd_remap=cfg.get('xcd_remap', True),
                   reliable_delta=cfg.get('reliable_delta', 0.01))
    if cfg.get('method') == 'bicgstab' and dev.current[0] != _lib.FEM_F64:
        dev.tangent(u, _lib.FEM_F64)          # BiCGSTAB runs on the FP64 matrix
    x, info = dev.solve(u, b, x0, cfg)
    err = info.res_norm
    logger.debug("%s (%s): %d iterations, %d FP64 residual replacements, |b - A x| = %.3g",
                 {'amg': 'AMG-CG', 'bicgstab': 'BiCGSTAB'}.get(cfg.get('method'), 'PCG'), cfg['precision'],
                 info.iterations,
                 info.replacements, err)
    if info.status == _lib.FEM_ERR_NOT_CONVERGED:
        logger.warning("the linear solve reached maxiter=%d, |b - A x| = %.3g", cfg['maxiter'], err)
    assert info.status != _lib.FEM_ERR_NAN and math.isfinite(err) and err < 0.1, \
        f"JAX linear solver failed to converge with err = {err}"
    return x, info


def _bc_increment(dev, dofs):
    """x0 = g - u on Dirichlet dofs, 0 elsewhere (solver.py:402-409)."""
    x0 = torch.zeros_like(dofs)
    if dev.n_bc:
        x0[dev.bc_dofs_long] = dev.bc_vals - dofs[dev.bc_dofs_long]
    return x0


def _line_search(dev, dofs, inc, full=lambda v: v, reduce=lambda v: v):
    """Step halving on |R| (solver.py:424-463)."""
    def res_norm(alpha):
        return dev.norm(reduce(dev.residual(full(dofs + alpha * inc))))
    alpha = 1.
    rn = res_norm(alpha)
    for _ in range(3):
        alpha *= 0.5
        rh = res_norm(alpha)
        if rh > rn:
            alpha *= 2.
            break
        rn = rh
    return dofs + alpha * inc


def solver(problem, solver_options={}):
    """Solve the nonlinear problem with Newton's method (solver.py:1112-1356).

    Returns ``sol_list`` = [tensor (num_nodes, vec)] on the device.
    """
    method, cfg = _resolve_solver_options(solver_options)
    if method != 'newton':
        raise NotImplementedError(f"the '{method}' nonlinear method is outside the device path")
    lin = linear_config(cfg.get('linear', {}))
    logger.info("Solving the nonlinear problem...")
    wall0 = time.perf_counter()
    dev = problem.device()
    dev.prepare()
    timing = {'local_assembly': 0., 'global_matrix': 0., 'linear': 0.}
    if lin.get('method') == 'amg' and len(dev.vecs) > 1:
        raise NotImplementedError("the AMG hierarchy takes one scalar or one vector variable (vec 1 or 3); "
                                  "solve several variables with 'jax_solver'")
    constrained = dev.constraint() is not None
    if constrained:
        # linear multipoint constraints u = P u_r (solver.py:1290-1339): Newton on the reduced dofs,
        # the reduced system P^T A P solved by FP64 Jacobi-BiCGSTAB as get_A + jax_solve do
        if lin.get('method') == 'amg':
            raise NotImplementedError("amgx_solver with P_mat constraints is outside the device path")
        lin = dict(lin, method='bicgstab', constrained=True, precision='fp64', precision_code=_lib.FEM_F64)
        full = dev.p_apply
        reduce = lambda r: dev.p_apply(r, transpose=True)
    else:
        full = reduce = lambda v: v
    if 'initial_guess' in cfg:
        dofs = problem._flat(cfg['initial_guess']).clone()
        if constrained and dofs.numel() != dev.constraint().cols:
            raise ValueError("with P_mat, initial_guess holds the reduced dofs")
    else:
        dofs = torch.zeros(dev.constraint().cols if constrained else dev.n, dtype=torch.float64,
                           device=dev.dm.device)
    rel_tol = cfg.get('rel_tol', 1e-8)
    tol = cfg.get('tol', 1e-6)

    # the first residual and tangent at the same u: one pass where the single-pass assembly
    # applies (newton_update, problem.py:477-491)
    t0 = time.perf_counter()
    have_tangent = not constrained and dev.fused_applies(
        lin['assembly_code'] if lin['assembly_code'] is not None else _lib.FEM_F64)
    if have_tangent:
        u_full = full(dofs)
        dev.tangent(u_full, lin['precision_code'], lin['assembly_code'], residual=True)
        res = dev.apply_residual_bc(u_full, dev.last_residual)
        torch.cuda.synchronize()
        timing['global_matrix'] += time.perf_counter() - t0
    else:
        res = reduce(dev.residual(full(dofs)))
    res_val = dev.norm(res)
    res_val0 = res_val
    rel = res_val / res_val0 if res_val0 != 0 else float('nan')
    logger.info("  iter 0: nonlinear residual L2 norm = %.3g", res_val)
    n_iters = 0
    stats = []
    while rel > rel_tol and res_val > tol:
        n_iters += 1
        t0 = time.perf_counter()
        u_full = full(dofs)
        if not have_tangent:
            dev.tangent(u_full, lin['precision_code'], lin['assembly_code'])
            torch.cuda.synchronize()
            timing['global_matrix'] += time.perf_counter() - t0
        have_tangent = False
        b = -res
        x0 = reduce(_bc_increment(dev, u_full))
        t0 = time.perf_counter()
        inc, info = linear_solve(dev, u_full, b, x0, lin)
        timing['linear'] += time.perf_counter() - t0
        stats.append(dict(iterations=info.iterations, replacements=info.replacements, res_norm=info.res_norm,
                          spmv_ms=info.spmv_ms, update_ms=info.update_ms))
        if cfg.get('line_search_flag', False):
            dofs = _line_search(dev, dofs, inc, full, reduce)
        else:
            dofs = dofs + inc
        t0 = time.perf_counter()
        res = reduce(dev.residual(full(dofs)))
        res_val = dev.norm(res)
        timing['local_assembly'] += time.perf_counter() - t0
        rel = res_val / res_val0
        logger.info("  iter %d: nonlinear residual L2 norm = %.3g (relative %.3g), PCG %d its",
                    n_iters, res_val, rel, info.iterations)
    assert math.isfinite(res_val), "res_val contains NaN, stop the program!"
    assert math.isfinite(dev.dot(dofs, dofs)), "dofs contains NaN, stop the program!"
    dofs = full(dofs)
    problem._last_u = dofs
    problem._solve_stats = dict(newton_iterations=n_iters, linear=stats, timing=timing,
                                wall=time.perf_counter() - wall0, res_norm=res_val)
    logger.info("Solve took %.3f s (%d Newton iterations)", time.perf_counter() - wall0, n_iters)
    return problem.unflatten_fn_sol_list(dofs)


def get_A(problem):
    """The device tangent of the last assembly (solver.py:540-553), as the DeviceProblem."""
    return problem.device()


# ------------------------------------------------------------------ adjoint
def _leaves(params):
    if isinstance(params, (list, tuple)):
        return list(params), (lambda xs, t=type(params): t(xs))
    return [params], (lambda xs: xs[0])


def implicit_vjp(problem, sol_list, params, v_list, adjoint_solver_options):
    """Adjoint vector-Jacobian product (solver.py:1362-1418).

    Solves A^T lambda = v on the row-eliminated tangent at u* (A^T acts as K on
    the free dofs, whose block is symmetric for every device law; the Dirichlet
    rows of the constraint carry no parameter dependence), then returns
    -lambda^T dR/dparams through the adjoint kernels and ``set_params``.
    """
    leaves, rebuild = _leaves(params)
    req = []
    with torch.enable_grad():
        for x in leaves:
            if torch.is_tensor(x) and x.is_floating_point():
                req.append(x.detach().clone().requires_grad_(True))
            else:
                req.append(x)
        problem.set_params(rebuild(req))
        ivs = list(problem.internal_vars) if problem.internal_vars is not None else []
        sivs = [list(s) for s in problem.internal_vars_surfaces]
    dev = problem.device()
    dev.prepare()
    u = problem._flat(sol_list)
    lin = linear_config(adjoint_solver_options)
    dev.tangent(u, lin['precision_code'], lin['assembly_code'])
    v = torch.as_tensor(v_list[0] if isinstance(v_list, (list, tuple)) else v_list).to(
        device=u.device, dtype=torch.float64).reshape(-1).clone()
    if dev.n_bc:
        v[dev.bc_dofs_long] = 0.
    if dev.nonsymmetric:
        raise NotImplementedError("the adjoint of a tangent that is not symmetric (A^T lambda = v) is outside the "
                                  "device path")
    if dev.constraint() is not None:
        # solver.py:1402-1412: v_r = P^T v, A_r lambda_r = v_r, lambda = P lambda_r
        lin_c = dict(lin, method='bicgstab', constrained=True, precision='fp64', precision_code=_lib.FEM_F64)
        v_r = dev.p_apply(v, transpose=True)
        lam_r, info = linear_solve(dev, u, v_r, torch.zeros_like(v_r), lin_c)
        lam = dev.p_apply(lam_r)
    else:
        lam, info = linear_solve(dev, u, v, torch.zeros_like(v), lin)
    problem._adjoint_stats = dict(iterations=info.iterations, replacements=info.replacements,
                                  res_norm=info.res_norm, spmv_ms=info.spmv_ms, update_ms=info.update_ms)
    outs, cots = [], []
    if dev.material.has_param:
        k = dev.param_index
        g = dev.param_grad(u, lam)
        if torch.is_tensor(ivs[k]) and ivs[k].requires_grad:
            outs.append(ivs[k])
            cots.append(g.reshape(ivs[k].shape).to(ivs[k].dtype))
    from ._device import _eval_quad_map, _flatten_iv
    dm = dev.dm
    if dev.material.tabulated and any(torch.is_tensor(t) and t.requires_grad for t in ivs):
        # lambda^T R = sum_q JxW P(grad u_q, theta_q) : grad lambda_q, so -lambda^T dR/dtheta is
        # the vjp of the tensor map with cotangent -JxW grad lambda at each point
        gu, gl = dev.quad_grad(u), dev.quad_grad(lam)
        if 'jxw' not in dev.buffers:
            dev.buffers['jxw'] = dm.geometry(want_jxw=True)[1].reshape(-1)
        with torch.enable_grad():
            civ = [_flatten_iv(t, (dm.C, dm.Q), dm.device) for t in ivs]
            P = dev.material._pm.form(gu, *civ)
        if P.requires_grad:
            outs.append(P)
            cots.append(-dev.buffers['jxw'][:, None, None] * gl)
    if dev.cell_terms:
        # per-cell / per-face weak-form terms: vjp of each with cotangent -lambda on its cell's dofs
        V = dev.V
        cl = dm.cells.long()
        u_c = dev.cell_dofs_reference(u)
        lam_c = dev.cell_dofs_reference(lam)
        for t in dev.cell_terms:
            if not any(torch.is_tensor(x) and x.requires_grad for x in t['ivs']):
                continue
            with torch.enable_grad():
                rest = t['rest'][:4] + [_flatten_iv(x, t['lead'], dm.device) for x in t['ivs']]
                ids = t['ids']
                y = t['pm'].form(u_c if ids is None else u_c[ids], *rest)
            outs.append(y)
            cots.append(-(lam_c if ids is None else lam_c[ids]))
    if hasattr(problem, 'get_mass_map') and any(torch.is_tensor(t) and t.requires_grad for t in ivs):
        lq = dev.load_adjoint(lam).reshape(-1, dev.V)
        xq = dev.buffers['xq']
        with torch.enable_grad():
            civ = [_flatten_iv(t, (dm.C, dm.Q), dm.device) for t in ivs]
            uq = dev.quad_vals(u) if dev.mass_dyn is not None else torch.zeros_like(lq)
            m = _eval_quad_map(problem.get_mass_map(), dev.V, uq, xq, civ)
        if m.requires_grad:
            outs.append(m)
            cots.append(-lq)
    for i, siv in enumerate(sivs):
        if not any(torch.is_tensor(t) and t.requires_grad for t in siv):
            continue
        if not hasattr(problem, 'get_surface_maps') or i in dev.dynamic_surfaces:
            continue   # universal or u-dependent surface terms: their vjp is in the per-cell terms above
        srf = dev.surfaces[i]
        S = srf['faces'].shape[0]
        lq = dev.face_load_adjoint(i, lam).reshape(-1, dev.V)
        with torch.enable_grad():
            civ = [_flatten_iv(t, (S, dm.FQ), dm.device) for t in siv]
            t = _eval_quad_map(problem.get_surface_maps()[i], dev.V, torch.zeros_like(lq),
                               srf['xq'].reshape(-1, 3), civ)
        if t.requires_grad:
            outs.append(t)
            cots.append(-lq)
    targets = [x for x in req if torch.is_tensor(x) and x.requires_grad]
    grads = torch.autograd.grad(outs, targets, grad_outputs=cots, allow_unused=True) if outs and targets else \
        [None] * len(targets)
    it = iter(grads)
    out = []
    for x in req:
        if torch.is_tensor(x) and x.requires_grad:
            g = next(it)
            out.append(torch.zeros_like(x) if g is None else g)
        else:
            out.append(None)
    return out


class _ImplicitSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, problem, solver_options, adjoint_solver_options, rebuild, *leaves):
        problem.set_params(rebuild(list(leaves)))
        sol = solver(problem, solver_options)[0]
        ctx.problem = problem
        ctx.adjoint_solver_options = adjoint_solver_options
        ctx.rebuild = rebuild
        ctx.n_leaves = len(leaves)
        ctx.nontensor = [None if torch.is_tensor(x) else x for x in leaves]
        tensors = [x for x in leaves if torch.is_tensor(x)]
        ctx.save_for_backward(sol, *tensors)
        return sol

    @staticmethod
    def backward(ctx, v):
        sol, *tensors = ctx.saved_tensors
        it = iter(tensors)
        leaves = [next(it) if x is None else x for x in ctx.nontensor]
        logger.info("Running backward and solving the adjoint problem...")
        grads = implicit_vjp(ctx.problem, [sol], ctx.rebuild(leaves), [v], ctx.adjoint_solver_options)
        return (None, None, None, None, *grads)


def ad_wrapper(problem, solver_options={}, adjoint_solver_options={}):
    """Differentiable forward solve (solver.py:1421-1455): ``fwd_pred(params) -> sol_list``."""
    def fwd_pred(params):
        leaves, rebuild = _leaves(params)
        sol = _ImplicitSolve.apply(problem, solver_options, adjoint_solver_options, rebuild, *leaves)
        return [sol]
    return fwd_pred
