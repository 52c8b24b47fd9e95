"""Mass maps that depend on u (reaction terms, implicit time steps) on the device:
m(u_q, x_q, *iv) and dm/du are evaluated at the quadrature points before each
launch; the HIP kernels add the mass term to the residual, the cell tangents
and the matrix-free operator.  Checked against the oracle (numpy map,
complex-step dm/du) and, for parameter gradients, against central differences."""
import numpy as np
import pytest
import torch

import oracle_setup
from oracle import fem as OF
from problems import (HEAT_DT, LinearPoisson, UserMapProblem, box, heat_np, heat_torch, isclose, reaction_np,
                      reaction_torch, spring_np, spring_torch)
from test_gpu_parity import oracle_problem, rel

from jax_fem_amd import _lib
from jax_fem_amd import materials as M
from jax_fem_amd.solver import ad_wrapper, solver

pytestmark = pytest.mark.gpu


def _oracle(p, mat, fn, ivs=()):
    # Dirichlet data and surface loads from the problem's own definition; the mass map
    # depends on u and stays in the oracle's residual and tangent
    o = oracle_setup.oracle_problem(p, mat, mass=False)
    o.set_mass_map(fn, ivs)
    return o


class Reaction(LinearPoisson):
    def get_mass_map(self):
        return reaction_torch


def _reaction_problem():
    left = lambda x: isclose(x[0], 0.)
    return Reaction(box(8, 6, 5), vec=1, dim=3, dirichlet_bc_info=[[left], [0], [lambda x: 0.]])


@pytest.mark.parametrize('opts', [{'jax_solver': {'precision': 'fp64'}}, {'jax_solver': {'precision': 'mixed'}},
                                  {'amgx_solver': {'precision': 'mixed'}}], ids=['fp64', 'mixed', 'amg'])
def test_reaction_diffusion_vs_oracle(gpu, opts):
    p = _reaction_problem()
    sol = solver(p, opts)[0].reshape(-1).cpu().numpy()
    assert p._solve_stats['newton_iterations'] >= 3
    assert p.device().mass_dyn is not None
    ref = _oracle(p, OF.Material('poisson', k=1.), reaction_np).newton(linear='direct')
    assert np.linalg.norm(sol - ref) <= 1e-8 * np.linalg.norm(ref)


def test_mass_term_cell_tangent_residual_and_operator_vs_oracle(gpu):
    """Elasticity with a nonlinear elastic foundation (m = k |u|^2 u): residual, assembled
    matrix and the matrix-free operator against the oracle at a random state."""
    left = lambda x: isclose(x[0], 0.)

    class P(UserMapProblem):
        tensor_map = M.LinearElastic(E=70e3, nu=0.3)

        def get_mass_map(self):
            return spring_torch
    p = P(box(4, 3, 3), vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda x: 0.] * 3])
    dev = p.device()
    dev.prepare()
    u = 0.1 * np.random.default_rng(3).standard_normal(dev.n)
    ut = torch.as_tensor(u, device='cuda')
    lm, mu = OF.lame(70e3, 0.3)
    o = _oracle(p, OF.Material('elastic', lmbda=lm, mu=mu), spring_np)
    r = p.compute_residual([ut.reshape(-1, 3)])[0].reshape(-1).cpu().numpy()
    assert rel(r, o.residual(u, bc=False)) < 1e-12
    for precision, tol in ((_lib.FEM_F64, 1e-12), (_lib.FEM_F32, 2e-6)):
        dev.tangent(ut, precision, _lib.FEM_F64 if precision == _lib.FEM_F32 else None)
        data = dev.export_csr()[2]
        A = o.tangent(u)
        assert rel(data, A.data) < tol
    x = torch.as_tensor(np.random.default_rng(4).standard_normal(dev.n), device='cuda')
    assert rel(dev.apply_operator(ut, x).cpu().numpy(), A @ x.cpu().numpy()) < 1e-12


def test_spring_foundation_newton_mixed_vs_oracle(gpu):
    left, right = (lambda x: isclose(x[0], 0.)), (lambda x: isclose(x[0], 2.))

    class P(UserMapProblem):
        tensor_map = M.LinearElastic(E=70e3, nu=0.3)

        def get_mass_map(self):
            return spring_torch

        def get_surface_maps(self):
            return [lambda u, x: torch.tensor([0., 0., 400.], dtype=torch.float64)]
    p = P(box(8, 4, 4, 2., 1., 1.), vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda x: 0.] * 3],
          location_fns=[right])
    sol = solver(p, {'jax_solver': {'precision': 'mixed'}})[0].reshape(-1).cpu().numpy()
    assert p._solve_stats['newton_iterations'] >= 3
    lm, mu = OF.lame(70e3, 0.3)
    ref = _oracle(p, OF.Material('elastic', lmbda=lm, mu=mu), spring_np).newton(linear='direct')
    assert np.linalg.norm(sol - ref) <= 1e-8 * np.linalg.norm(ref)


def test_backward_euler_heat_steps_vs_oracle(gpu):
    """Three implicit heat steps, u_old carried as a per-quad internal variable
    (the reference's transient pattern), against the oracle stepping the same way."""
    left, right = (lambda x: isclose(x[0], 0.)), (lambda x: isclose(x[0], 1.))

    class Heat(LinearPoisson):
        # internal variables reach the tensor map too (problem.py:208), which ignores them here
        def get_tensor_map(self):
            return lambda u_grad, u_old: u_grad

        def get_mass_map(self):
            return heat_torch
    p = Heat(box(8, 4, 4), vec=1, dim=3, dirichlet_bc_info=[[left, right], [0, 0], [lambda x: 0., lambda x: 1.]])
    C, Q = p.num_cells, p.fes[0].num_quads
    uq = torch.zeros((C, Q, 1), dtype=torch.float64)
    o = _oracle(p, OF.Material('poisson', k=1.), heat_np)
    uq_o = np.zeros((C, Q, 1))
    uo = np.zeros(o.n)
    for step in range(3):
        p.internal_vars = [uq]
        sol = solver(p, {'jax_solver': {'precision': 'fp64', 'tol': 1e-12, 'atol': 0.}})[0]
        dev = p.device()
        uq = dev.quad_vals(sol.reshape(-1)).reshape(C, Q, 1).cpu()
        o.set_mass_map(heat_np, (uq_o,))
        uo = o.newton(dofs=uo, linear='direct')
        uq_o = np.einsum('qa,cav->cqv', o.sv, uo.reshape(-1, 1)[o.cells])
        assert np.linalg.norm(sol.reshape(-1).cpu().numpy() - uo) <= 1e-9 * np.linalg.norm(uo), step
    # the temperature front advances: interior values grow from step to step
    assert 0.0 < float(uq.mean()) < 1.0


def test_reaction_coefficient_gradient_fd(gpu):
    """m = theta u^3 - f with theta per quadrature point: the adjoint gradient of
    sum(u) against central differences."""
    left = lambda x: isclose(x[0], 0.)

    class P(LinearPoisson):
        def get_tensor_map(self):
            return lambda u_grad, th: u_grad

        def get_mass_map(self):
            return lambda u, x, th: th * u ** 3 - 10. * x[0:1]

        def set_params(self, th):
            self.internal_vars = [th]
    p = P(box(6, 4, 3), vec=1, dim=3, dirichlet_bc_info=[[left], [0], [lambda x: 0.]])
    C, Q = p.num_cells, p.fes[0].num_quads
    th = torch.as_tensor(np.random.default_rng(5).uniform(0.5, 2.0, (C, Q, 1)), device='cuda')
    opts = {'jax_solver': {'tol': 1e-13, 'atol': 0.}, 'rel_tol': 1e-12, 'tol': 1e-13}
    fwd = ad_wrapper(p, opts, {'jax_solver': {'tol': 1e-13, 'atol': 0.}})
    t = th.clone().requires_grad_(True)
    fwd(t)[0].sum().backward()
    g = t.grad
    assert p.device().mass_dyn is not None
    h = 1e-5
    for (c, q) in [(0, 0), (17, 3), (C - 1, Q - 1)]:
        tp, tm = th.clone(), th.clone()
        tp[c, q, 0] += h
        tm[c, q, 0] -= h
        fd = float((fwd(tp)[0].sum() - fwd(tm)[0].sum()) / (2 * h))
        assert abs(fd - float(g[c, q, 0])) <= 1e-5 * max(abs(fd), 1e-3 * float(g.abs().max())), (c, q)
