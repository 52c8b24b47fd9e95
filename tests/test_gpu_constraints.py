"""Linear multipoint constraints u = P u_r (P_mat, solver.py:404-406, 545-551,
1290-1339, 1402-1412): periodic Poisson on the device (reduced P^T A P solved by
FP64 Jacobi-BiCGSTAB) against the oracle's reduced Newton with direct solves,
and the constrained adjoint against central differences."""
import numpy as np
import pytest
import torch

import oracle_setup
from oracle import fem as OF
from problems import box, isclose, periodic_p_mat
from test_gpu_parity import oracle_problem

from jax_fem_amd.problem import Problem
from jax_fem_amd.solver import ad_wrapper, solver

pytestmark = pytest.mark.gpu


class PeriodicPoisson(Problem):
    # applications/periodic_bc/example.py:119-133 in 3D
    def get_tensor_map(self):
        return lambda x, theta: x * theta

    def get_mass_map(self):
        def mass_map(u, x, theta):
            dx, dy = x[0] - 0.5, x[1] - 0.5
            val = x[0] * torch.sin(5.0 * np.pi * x[1]) + 1.0 * torch.exp(-(dx * dx + dy * dy) / 0.02)
            return torch.stack([-val])
        return mass_map

    def set_params(self, theta):
        self.internal_vars = [theta]


def _problem():
    mesh = box(10, 8, 4)
    bottom, top = (lambda x: isclose(x[2], 0.)), (lambda x: isclose(x[2], 1.))
    p = PeriodicPoisson(mesh, vec=1, dim=3, dirichlet_bc_info=[[bottom, top], [0, 0], [lambda x: 0.] * 2])
    left, right = (lambda x: isclose(x[0], 0.)), (lambda x: isclose(x[0], 1.))
    p.P_mat = periodic_p_mat(mesh.points, 1, [(left, right, lambda x: x + np.array([1., 0., 0.]), 0)])
    return p


def test_periodic_poisson_vs_oracle(gpu):
    p = _problem()
    C, Q = p.num_cells, p.fes[0].num_quads
    theta = torch.full((C, Q), 1.0, dtype=torch.float64)
    p.set_params(theta)
    sol = solver(p, {'jax_solver': {'tol': 1e-12, 'atol': 0.}})[0].reshape(-1).cpu().numpy()
    dev = p.device()
    o = oracle_setup.oracle_problem(p, OF.Material('poisson', k=1.))
    ref = o.newton_constrained(p.P_mat)
    assert np.linalg.norm(sol - ref) <= 1e-9 * np.linalg.norm(ref)
    # the solution is periodic: u(0, y, z) = u(1, y, z)
    pts = p.fes[0].points
    left = np.flatnonzero(np.isclose(pts[:, 0], 0.))
    right = np.flatnonzero(np.isclose(pts[:, 0], 1.))
    order = lambda ids: ids[np.lexsort((pts[ids, 2], pts[ids, 1]))]
    assert np.array_equal(sol[order(left)], sol[order(right)])
    assert np.abs(sol).max() > 1e-3


def test_periodic_adjoint_fd(gpu):
    p = _problem()
    C, Q = p.num_cells, p.fes[0].num_quads
    theta = torch.as_tensor(np.random.default_rng(3).uniform(0.5, 1.5, (C, Q)), device='cuda')
    opts = {'jax_solver': {'tol': 1e-13, 'atol': 0.}}
    fwd = ad_wrapper(p, opts, opts)
    t = theta.clone().requires_grad_(True)
    J = lambda u: (u[:, 0] ** 2).sum()
    J(fwd(t)[0]).backward()
    g = t.grad
    h = 1e-6
    for (c, q) in [(0, 0), (37, 5), (C - 1, Q - 1)]:
        tp, tm = theta.clone(), theta.clone()
        tp[c, q] += h
        tm[c, q] -= h
        fd = float((J(fwd(tp)[0]) - J(fwd(tm)[0])) / (2 * h))
        assert abs(fd - float(g[c, q])) <= 1e-5 * max(abs(fd), 1e-3 * float(g.abs().max())), (c, q)
