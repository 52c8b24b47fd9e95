"""Edge cases of the device path: zero load, fully constrained problems, a
single cell (HEX8 and HEX27), a surface set that selects no face, a
parameter field with a NaN, and sizes that leave a partial last slice."""
import numpy as np
import pytest
import torch

from oracle import fem as OF
from oracle_setup import oracle_problem
from problems import LoadedElastic, box, isclose

from jax_fem_amd import materials as M
from jax_fem_amd.generate_mesh import Mesh, box_mesh_hex27
from jax_fem_amd.problem import Problem
from jax_fem_amd.solver import solver

pytestmark = pytest.mark.gpu


class Elastic(Problem):
    def get_tensor_map(self):
        return M.LinearElastic(E=70e3, nu=0.3)


def _clamp(x0=0.):
    left = lambda x: isclose(x[0], x0)
    return [[left] * 3, [0, 1, 2], [lambda x: 0.] * 3]


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_zero_load_gives_zero_solution(gpu, precision):
    p = Elastic(box(4, 3, 3), vec=3, dim=3, dirichlet_bc_info=_clamp())
    sol = solver(p, {'jax_solver': {'precision': precision}})[0]
    assert p._solve_stats['newton_iterations'] == 0
    assert torch.count_nonzero(sol) == 0


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_fully_constrained_problem(gpu, precision):
    # every dof is Dirichlet: the solution is the prescribed field
    allp = lambda x: True
    vals = [lambda x: 0.1 * x[0], lambda x: -0.2 * x[1], lambda x: 0.05]
    p = Elastic(box(3, 2, 2), vec=3, dim=3, dirichlet_bc_info=[[allp] * 3, [0, 1, 2], vals])
    sol = solver(p, {'jax_solver': {'precision': precision}})[0].cpu().numpy()
    pts = p.fes[0].points
    ref = np.stack([0.1 * pts[:, 0], -0.2 * pts[:, 1], 0.05 + 0 * pts[:, 0]], 1)
    np.testing.assert_allclose(sol, ref, rtol=0, atol=1e-14)


@pytest.mark.parametrize('ele', ['HEX8', 'HEX27'])
def test_single_cell(gpu, ele):
    if ele == 'HEX8':
        mesh, order = box(1, 1, 1), None
    else:
        m = box_mesh_hex27(1, 1, 1, 1., 1., 1.)
        mesh, order = Mesh(m.points, m.cells_dict['hexahedron27']), 4
    p = LoadedElastic(mesh, vec=3, dim=3, ele_type=ele, quadrature_order=order, dirichlet_bc_info=_clamp(),
                      location_fns=[lambda x: isclose(x[0], 1.)])
    dev = p.device()
    assert dev.dm.C == 1 and dev.pat.num_slices == 1
    sol = solver(p, {'jax_solver': {'precision': 'fp64', 'tol': 1e-13, 'atol': 0.}})[0].cpu().numpy()
    fe = p.fes[0]
    o = oracle_problem(p, OF.Material('elastic', **dict(zip(('lmbda', 'mu'), OF.lame(70e3, 0.3)))))
    assert np.array_equal(o.bc_dofs, dev.bc_dofs.cpu().numpy())
    ref = o.newton(linear='direct').reshape(-1, 3)
    assert np.linalg.norm(sol - ref) / np.linalg.norm(ref) < 1e-10


def test_surface_set_without_faces(gpu):
    # a location function that matches no face: zero surface load, nothing else changes
    p = LoadedElastic(box(4, 2, 2, 2., 1., 1.), vec=3, dim=3, dirichlet_bc_info=_clamp(),
                      location_fns=[lambda x: isclose(x[0], 5.)])
    assert len(p.boundary_inds_list[0]) == 0
    sol = solver(p, {'jax_solver': {'precision': 'mixed'}})[0]
    assert torch.count_nonzero(sol) == 0


def test_nan_parameter_is_reported(gpu):
    from problems import SIMPElasticity
    p = SIMPElasticity(box(4, 2, 2, 2., 1., 1.), vec=3, dim=3, dirichlet_bc_info=_clamp(),
                       location_fns=[lambda x: isclose(x[0], 2.)])
    theta = torch.full((p.num_cells,), 0.5, dtype=torch.float64)
    theta[3] = float('nan')
    p.set_params(theta)
    with pytest.raises(AssertionError):
        solver(p, {'jax_solver': {'precision': 'fp64'}})


@pytest.mark.parametrize('n', [(1, 1, 61), (5, 3, 2)])
def test_partial_last_slice(gpu, n):
    # 248 and 72 nodes: the last 64-row slice is partly padding lanes
    p = LoadedElastic(box(*n, 1., 1., float(n[2])), vec=3, dim=3, dirichlet_bc_info=_clamp(),
                      location_fns=[lambda x: isclose(x[0], 1.)])
    dev = p.device()
    assert dev.dm.Nn % 64 != 0
    for precision in ('fp64', 'mixed'):
        sol = solver(p, {'jax_solver': {'precision': precision, 'tol': 1e-12, 'atol': 0.}})[0].cpu().numpy()
        fe = p.fes[0]
        o = oracle_problem(p, OF.Material('elastic', **dict(zip(('lmbda', 'mu'), OF.lame(70e3, 0.3)))))
        ref = o.newton(linear='direct').reshape(-1, 3)
        assert np.linalg.norm(sol - ref) / np.linalg.norm(ref) < 1e-8, precision
