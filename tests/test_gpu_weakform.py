"""Per-cell and per-face weak forms written by the user (get_universal_kernel,
get_universal_kernels_surface) and surface maps that depend on u, on the device:
the kernel and its jacfwd are evaluated per cell by torch on the GPU; the HIP
path assembles, solves and applies the matrix-free operator (per-cell matrices
through fem_cell_matvec).  Checked against the oracle's batched numpy kernels
with complex-step tangents, and gradients against central differences."""
import numpy as np
import pytest
import torch

from oracle import fem as OF
from problems import (box, energy_kernel_np, energy_kernel_torch, isclose, reaction_np, robin_kernel_np,
                      robin_kernel_torch, robin_map_torch)
from test_gpu_parity import oracle_problem, rel

from jax_fem_amd import _lib
from jax_fem_amd.problem import Problem
from jax_fem_amd.solver import ad_wrapper, solver

pytestmark = pytest.mark.gpu

LEFT = lambda x: isclose(x[0], 0.)
RIGHT = lambda x: isclose(x[0], 1.)


class EnergyForm(Problem):
    """-lap u + u^3 = 10 x_0 written only as a universal kernel."""
    def get_universal_kernel(self):
        return energy_kernel_torch(torch.as_tensor(self.fes[0].shape_vals, device='cuda'))


class EnergyRobin(EnergyForm):
    def get_universal_kernels_surface(self):
        return [robin_kernel_torch]


class EnergyRobinMap(EnergyForm):
    def get_surface_maps(self):
        return [robin_map_torch]


def _oracle(p, face=False):
    dev = p.device()
    o = oracle_problem(p, OF.Material('poisson', k=0.))
    o.bc_dofs, o.bc_vals = dev.bc_dofs.cpu().numpy().astype(np.int64), dev.bc_vals.cpu().numpy()
    o.add_cell_kernel(energy_kernel_np(o.sv))
    if face:
        o.add_face_kernel(p.boundary_inds_list[0], robin_kernel_np)
    return o


@pytest.mark.parametrize('opts', [{'jax_solver': {'precision': 'fp64'}}, {'jax_solver': {'precision': 'mixed'}},
                                  {'amgx_solver': {'precision': 'mixed'}}], ids=['fp64', 'mixed', 'amg'])
def test_universal_kernel_newton_vs_oracle(gpu, opts):
    p = EnergyForm(box(8, 6, 5), vec=1, dim=3, dirichlet_bc_info=[[LEFT], [0], [lambda x: 0.]])
    sol = solver(p, opts)[0].reshape(-1).cpu().numpy()
    assert p._solve_stats['newton_iterations'] >= 3 and p.device().cell_terms
    ref = _oracle(p).newton(linear='direct')
    assert np.linalg.norm(sol - ref) <= 1e-8 * np.linalg.norm(ref)


def test_universal_kernel_matrix_and_operator_vs_oracle(gpu):
    p = EnergyRobin(box(4, 3, 3), vec=1, dim=3, dirichlet_bc_info=[[LEFT], [0], [lambda x: 0.]],
                    location_fns=[RIGHT])
    dev = p.device()
    dev.prepare()
    u = 0.5 * np.random.default_rng(2).standard_normal(dev.n)
    ut = torch.as_tensor(u, device='cuda')
    o = _oracle(p, face=True)
    r = p.compute_residual([ut.reshape(-1, 1)])[0].reshape(-1).cpu().numpy()
    assert rel(r, o.residual(u, bc=False)) < 1e-12
    dev.tangent(ut, _lib.FEM_F64)
    A = o.tangent(u)
    assert rel(dev.export_csr()[2], A.data) < 1e-12
    x = torch.as_tensor(np.random.default_rng(3).standard_normal(dev.n), device='cuda')
    assert rel(dev.apply_operator(ut, x).cpu().numpy(), A @ x.cpu().numpy()) < 1e-12


def test_robin_as_surface_kernel_and_as_u_dependent_map(gpu):
    """The same Robin condition as a universal surface kernel and as a surface map that
    depends on u: same solution, equal to the oracle's."""
    bc = [[LEFT], [0], [lambda x: 0.]]
    sols = []
    for cls in (EnergyRobin, EnergyRobinMap):
        p = cls(box(6, 4, 4), vec=1, dim=3, dirichlet_bc_info=bc, location_fns=[RIGHT])
        sols.append(solver(p, {'jax_solver': {'precision': 'mixed'}})[0].reshape(-1).cpu().numpy())
    assert p.device().dynamic_surfaces == {0}
    ref = _oracle(p, face=True).newton(linear='direct')
    for s in sols:
        assert np.linalg.norm(s - ref) <= 1e-8 * np.linalg.norm(ref)


def test_universal_kernel_parameter_gradient_fd(gpu):
    """theta per quadrature point inside a universal kernel (theta u^3): the adjoint
    gradient of sum(u) against central differences."""
    class P(Problem):
        def get_universal_kernel(self):
            return energy_kernel_torch(torch.as_tensor(self.fes[0].shape_vals, device='cuda'))

        def set_params(self, th):
            self.internal_vars = [th]
    p = P(box(6, 4, 3), vec=1, dim=3, dirichlet_bc_info=[[LEFT], [0], [lambda x: 0.]])
    C, Q = p.num_cells, p.fes[0].num_quads
    th = torch.as_tensor(np.random.default_rng(5).uniform(0.5, 2.0, (C, Q, 1)), device='cuda')
    opts = {'jax_solver': {'tol': 1e-13, 'atol': 0.}, 'rel_tol': 1e-12, 'tol': 1e-13}
    fwd = ad_wrapper(p, opts, {'jax_solver': {'tol': 1e-13, 'atol': 0.}})
    t = th.clone().requires_grad_(True)
    fwd(t)[0].sum().backward()
    g = t.grad
    h = 1e-5
    for (c, q) in [(0, 0), (17, 3), (C - 1, Q - 1)]:
        tp, tm = th.clone(), th.clone()
        tp[c, q, 0] += h
        tm[c, q, 0] -= h
        fd = float((fwd(tp)[0].sum() - fwd(tm)[0].sum()) / (2 * h))
        assert abs(fd - float(g[c, q, 0])) <= 1e-5 * max(abs(fd), 1e-3 * float(g.abs().max())), (c, q)
