"""Tensor maps outside the closed-form laws on the device (the tabulated law):
the map's P and dP/dG are evaluated at every quadrature point and read by the
HIP kernels.  Checked against the closed-form kernels on laws both paths cover,
and against the oracle (numpy map, complex-step derivatives) on laws only the
tabulated path covers."""
import numpy as np
import pytest
import torch

import oracle_setup
from oracle import fem as OF
from problems import (UserMapProblem, box, diffusion_np, diffusion_torch, fiber_torch, isclose, scaled_svk_np,
                      scaled_svk_torch, svk_np, svk_torch)
from test_gpu_parity import _fixture_hyper, distorted_box, make_material_problem, oracle_problem, rel

from jax_fem_amd import _lib
from jax_fem_amd import materials as M
from jax_fem_amd.generate_mesh import Mesh, box_mesh_hex27
from jax_fem_amd.solver import ad_wrapper, solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('name,ele', [('elastic', 'HEX8'), ('elastic', 'HEX27'), ('neo_hookean', 'HEX8'),
                                      ('poisson', 'HEX8'), ('simp', 'HEX8')])
def test_tabulated_matches_closed_form_kernels(gpu, name, ele):
    """The same law through the closed-form kernels and through the table."""
    p, o, u = make_material_problem(name, ele)
    dev = p.device()
    dev.prepare()
    assert not dev.material.tabulated
    ut = torch.as_tensor(u, device='cuda')
    K = dev.cell_tangents(ut, _lib.FEM_F64).cpu().numpy()
    r = p.compute_residual([ut.reshape(-1, p.fes[0].vec)])[0].cpu().numpy()
    x = torch.as_tensor(np.random.default_rng(2).standard_normal(dev.n), device='cuda')
    y = dev.apply_operator(ut, x).cpu().numpy()
    closed = dev.material
    type(p).get_tensor_map = lambda self: M.Tabulated(closed)
    dev._mat_cache.clear()
    dev.prepare()
    assert dev.material.tabulated
    Kt = dev.cell_tangents(ut, _lib.FEM_F64).cpu().numpy()
    rt = p.compute_residual([ut.reshape(-1, p.fes[0].vec)])[0].cpu().numpy()
    yt = dev.apply_operator(ut, x).cpu().numpy()
    assert rel(Kt, K) < 1e-12 and rel(rt, r) < 1e-12 and rel(yt, y) < 1e-12
    K32 = dev.cell_tangents(ut, _lib.FEM_F32).cpu().numpy()
    assert rel(K32, K) < 2e-6


def _svk_problem(n=(6, 3, 3), load=0.4):
    left, right = (lambda x: isclose(x[0], 0.)), (lambda x: isclose(x[0], 2.))

    class P(UserMapProblem):
        tensor_map = staticmethod(svk_torch)

        def get_surface_maps(self):
            return [lambda u, x: torch.tensor([0., 0., load], dtype=torch.float64)]
    return P(box(*n, 2., 1., 1.), vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda x: 0.] * 3],
             location_fns=[right])


@pytest.mark.parametrize('ele', ['HEX8', 'HEX27'])
def test_svk_cell_kernels_vs_oracle(gpu, ele):
    if ele == 'HEX8':
        mesh, order = distorted_box((3, 3, 4), seed=7), None
    else:
        m = box_mesh_hex27(2, 2, 2, 1., 1., 1.)
        mesh = Mesh(m.points + 0.02 * np.random.default_rng(7).uniform(-1, 1, m.points.shape),
                    m.cells_dict['hexahedron27'])
        order = 4

    class P(UserMapProblem):
        tensor_map = staticmethod(svk_torch)
    p = P(mesh, vec=3, dim=3, ele_type=ele, quadrature_order=order,
          dirichlet_bc_info=[[lambda x: isclose(x[0], 0.)] * 3, [0, 1, 2], [lambda x: 0.] * 3])
    dev = p.device()
    dev.prepare()
    assert dev.material.tabulated
    u = 0.05 * np.random.default_rng(8).standard_normal(p.fes[0].num_total_dofs)
    ut = torch.as_tensor(u, device='cuda')
    o = oracle_problem(p, OF.Material('user', fn=svk_np))
    U = u.reshape(-1, 3)
    assert rel(dev.cell_tangents(ut).cpu().numpy(), OF.cell_tangent(o.mat, U, o.cells, o.G, o.JxW)) < 1e-12
    r = p.compute_residual([ut.reshape(-1, 3)])[0].cpu().numpy()
    assert rel(r, OF.scatter(OF.cell_residual(o.mat, U, o.cells, o.G, o.JxW), o.cells, len(o.points))) < 1e-12
    G = dev.quad_grad(ut).cpu().numpy().reshape(o.G.shape[0], o.G.shape[1], 3, 3)
    assert rel(G, np.einsum('cav,cqaj->cqvj', U[o.cells], o.G)) < 1e-13


@pytest.mark.parametrize('opts', [{'jax_solver': {'precision': 'fp64'}}, {'jax_solver': {'precision': 'mixed'}},
                                  {'amgx_solver': {'precision': 'mixed'}}], ids=['fp64', 'mixed', 'amg'])
def test_svk_newton_vs_oracle(gpu, opts):
    p = _svk_problem()
    sol = solver(p, opts)[0].reshape(-1).cpu().numpy()
    assert p._solve_stats['newton_iterations'] >= 3   # a nonlinear solve
    o = oracle_setup.oracle_problem(p, OF.Material('user', fn=svk_np))
    ref = o.newton(linear='direct')
    assert np.linalg.norm(sol - ref) <= 1e-8 * np.linalg.norm(ref)


def test_nonlinear_diffusion_vs_oracle(gpu):
    left, right = (lambda x: isclose(x[0], 0.)), (lambda x: isclose(x[0], 1.))

    class P(UserMapProblem):
        tensor_map = staticmethod(diffusion_torch)

        def get_mass_map(self):
            # a source, so that the gradient varies and the flux law's nonlinearity acts
            return lambda u, x: torch.stack([-3. - 2. * x[1]])
    p = P(box(6, 5, 4), vec=1, dim=3,
          dirichlet_bc_info=[[left, right], [0, 0], [lambda x: 0., lambda x: 0.6]])
    sol = solver(p, {'jax_solver': {'precision': 'mixed'}})[0].reshape(-1).cpu().numpy()
    assert p._solve_stats['newton_iterations'] >= 3
    o = oracle_problem(p, OF.Material('user', fn=diffusion_np))
    dev = p.device()
    o.bc_dofs, o.bc_vals = dev.bc_dofs.cpu().numpy().astype(np.int64), dev.bc_vals.cpu().numpy()
    o.load = dev.load.cpu().numpy().reshape(-1, 1)
    ref = o.newton(linear='direct')
    assert np.linalg.norm(sol - ref) <= 1e-8 * np.linalg.norm(ref)


def test_tabulated_adjoint_scalar_parameter_vs_oracle(gpu):
    """Gradient of f.u through a map with one value per quadrature point: the vjp of
    the map against the oracle's -lambda^T dR/dtheta with complex-step dP/dtheta."""
    left, right = (lambda x: isclose(x[0], 0.)), (lambda x: isclose(x[0], 2.))

    class P(UserMapProblem):
        tensor_map = staticmethod(scaled_svk_torch)

        def get_surface_maps(self):
            return [lambda u, x: torch.tensor([0., 0., 0.3], dtype=torch.float64)]
    mesh = box(6, 3, 3, 2., 1., 1.)
    p = P(mesh, vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda x: 0.] * 3], location_fns=[right])
    C, Q = p.num_cells, p.fes[0].num_quads
    theta = torch.as_tensor(np.random.default_rng(10).uniform(0.4, 1.1, (C, Q)), device='cuda')
    opts = {'jax_solver': {'tol': 1e-12, 'atol': 0.}}
    fwd = ad_wrapper(p, opts, opts)
    th = theta.clone().requires_grad_(True)
    u = fwd(th)[0]
    dev = p.device()
    assert dev.material.tabulated
    J = -torch.dot(dev.load, u.reshape(-1))
    J.backward()
    g = th.grad.cpu().numpy()
    o = oracle_problem(p, OF.Material('user', fn=scaled_svk_np), theta=theta.cpu().numpy())
    o.bc_dofs, o.bc_vals = dev.bc_dofs.cpu().numpy().astype(np.int64), dev.bc_vals.cpu().numpy()
    o.load = dev.load.cpu().numpy().reshape(-1, 3)
    uo = o.newton(linear='direct')
    go, _ = o.adjoint_theta(uo, -o.load.reshape(-1))
    assert rel(g, go) < 1e-7


def test_tabulated_adjoint_vector_internal_variable_fd(gpu):
    """A fibre direction per quadrature point (3 values): gradient of f.u with respect
    to it against central differences on a few seeded points."""
    left, right = (lambda x: isclose(x[0], 0.)), (lambda x: isclose(x[0], 2.))

    class P(UserMapProblem):
        tensor_map = staticmethod(fiber_torch)

        def get_surface_maps(self):
            return [lambda u, x: torch.tensor([0., 0., 0.3], dtype=torch.float64)]
    p = P(box(4, 2, 2, 2., 1., 1.), vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda x: 0.] * 3],
          location_fns=[right])
    C, Q = p.num_cells, p.fes[0].num_quads
    a = np.random.default_rng(11).standard_normal((C, Q, 3))
    a /= np.linalg.norm(a, axis=2, keepdims=True)
    a = torch.as_tensor(a, device='cuda')
    # Newton to 1e-11 (relative): the differences below resolve J to about 1e-12
    opts = {'jax_solver': {'tol': 1e-13, 'atol': 0.}, 'rel_tol': 1e-11, 'tol': 1e-13}
    fwd = ad_wrapper(p, opts, {'jax_solver': {'tol': 1e-13, 'atol': 0.}})

    def compliance(av):
        u = fwd(av)[0].reshape(-1)
        return -torch.dot(p.device().load, u)

    def J(av):
        return float(compliance(av))
    at = a.clone().requires_grad_(True)
    Jt = compliance(at)
    Jt.backward()
    g = at.grad
    h = 1e-5
    for (c, q, k) in [(0, 0, 2), (3, 5, 0), (C - 1, Q - 1, 1), (7, 2, 2)]:
        ap, am = a.clone(), a.clone()
        ap[c, q, k] += h
        am[c, q, k] -= h
        fd = (J(ap) - J(am)) / (2 * h)
        assert abs(fd - float(g[c, q, k])) <= 1e-4 * max(abs(fd), 1e-3 * float(g.abs().max())), (c, q, k)


def test_fenicsx_hyperelasticity_fixture_tabulated(gpu):
    """The FEniCSx hyperelasticity fixture with the Neo-Hookean law (P written out:
    torch's second derivatives of an energy through linalg.det are NaN at F = I)
    forced through the table."""
    from problems import HyperElasticity
    orig = HyperElasticity.get_tensor_map
    HyperElasticity.get_tensor_map = lambda self: M.Tabulated(M.NeoHookean(E=1e3, nu=0.3))
    try:
        sol, ref, decimal = _fixture_hyper('fp64')
    finally:
        HyperElasticity.get_tensor_map = orig
    got = sol.cpu().numpy().reshape(ref.shape).astype(np.float32)
    np.testing.assert_array_almost_equal(ref, got, decimal=decimal)


@pytest.mark.parametrize('ele', ['HEX8', 'HEX27'])
def test_quad_point_utilities_kernel_vs_torch(gpu, ele):
    """sol_to_grad / convert_from_dof_to_quad (fe.py:324-382): the fem_quad_grad kernel (plain
    tensors) against the differentiable torch contraction (tensors with autograd history)."""
    if ele == 'HEX8':
        mesh, order = distorted_box((3, 4, 3), seed=11), None
    else:
        m = box_mesh_hex27(2, 2, 1, 1., 1., 1.)
        mesh, order = Mesh(m.points, m.cells_dict['hexahedron27']), 4

    class P(UserMapProblem):
        tensor_map = staticmethod(svk_torch)
    p = P(mesh, vec=3, dim=3, ele_type=ele, quadrature_order=order)
    fe = p.fes[0]
    u = torch.as_tensor(np.random.default_rng(12).standard_normal((fe.num_total_nodes, 3)), device='cuda')
    g_k, q_k = fe.sol_to_grad(u), fe.convert_from_dof_to_quad(u)
    ur = u.clone().requires_grad_(True)
    g_t, q_t = fe.sol_to_grad(ur), fe.convert_from_dof_to_quad(ur)
    assert g_t.requires_grad and not g_k.requires_grad
    assert rel(g_k.cpu().numpy(), g_t.detach().cpu().numpy()) < 1e-13
    assert rel(q_k.cpu().numpy(), q_t.detach().cpu().numpy()) < 1e-14
