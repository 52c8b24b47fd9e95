"""Several variables on one mesh (problem.py:58-119): coupled thermoelasticity
(u, T) through get_universal_kernel, against the oracle's MultiProblem and the
closed forms that pin it (tests/test_oracle_multivar.py)."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import fem as OF
from problems import TM_LM, TM_MU, TM_ALPHA, box, thermo_np, thermo_torch

from jax_fem_amd.problem import Problem
from jax_fem_amd.solver import solver

pytestmark = pytest.mark.gpu

KAPPA = TM_ALPHA * (2. * TM_MU + 3. * TM_LM)


class ThermoElastic(Problem):
    # applications/thermal_mechanical_full/example.py:24-102 in 3D (T_old = 0, eps_old = 0)
    def custom_init(self):
        self.fe_u, self.fe_T = self.fes
        self.sv = torch.as_tensor(self.fe_T.shape_vals, device='cuda')

    def get_universal_kernel(self):
        return thermo_torch(self.sv, KAPPA)


def _plane(d, x=0.):
    return lambda p: np.isclose(p[d], x)


def _problem(n, bc_u, bc_T):
    mesh = box(*n)
    return ThermoElastic([mesh, mesh], vec=[3, 1], dim=3, ele_type=['HEX8', 'HEX8'],
                         dirichlet_bc_info=[bc_u, bc_T])


def _oracle(p, bc_u, bc_T):
    o = OF.MultiProblem(p.fes[0].points, p.fes[0].cells, [3, 1], bc_infos=[bc_u, bc_T])
    o.add_cell_kernel(thermo_np(o.sv, KAPPA))
    return o


def test_coupled_thermoelastic_newton_vs_oracle(gpu):
    left = _plane(0)
    bc_u = [[left] * 3, [0, 1, 2], [lambda p: 0.] * 3]
    bc_T = [[_plane(0, 1.), _plane(1)], [0, 0], [lambda p: 10., lambda p: -5. * p[0]]]
    p = _problem((4, 3, 2), bc_u, bc_T)
    assert p.num_vars == 2 and list(p.offset) == [0, 3 * p.fes[0].num_total_nodes]
    u, T = solver(p, {'jax_solver': {'tol': 1e-12, 'atol': 0.}})
    assert u.shape == (p.fes[0].num_total_nodes, 3) and T.shape == (p.fes[0].num_total_nodes, 1)
    assert p.device().nonsymmetric          # the coupling blocks are not transposes: BiCGSTAB
    o = _oracle(p, bc_u, bc_T)
    uo, To = o.sol_list(o.newton())
    assert np.linalg.norm(u.cpu().numpy() - uo) <= 1e-8 * np.linalg.norm(uo)
    assert np.linalg.norm(T.cpu().numpy() - To) <= 1e-8 * np.linalg.norm(To)
    assert np.abs(uo).max() > 0 and np.abs(To).max() > 0
    # the reference's COO export (problem.py:95-97, 447-458) in its variable-major dof numbering
    dofs = p._flat([u, T])
    p.newton_update([u, T])
    A = sp.coo_matrix((p.V, (p.I, p.J)), shape=(o.n, o.n)).tocsr()
    A.sum_duplicates()
    Ao = o.tangent(o.newton(), bc=False)
    assert np.abs((A - Ao).toarray()).max() <= 1e-12 * np.abs(Ao.toarray()).max()
    assert torch.equal(p._flat(p.unflatten_fn_sol_list(dofs)), dofs)


def test_uniform_heating_on_device(gpu):
    theta = 40.
    bc_u = [[_plane(0), _plane(1), _plane(2)], [0, 1, 2], [lambda p: 0.] * 3]
    bc_T = [[lambda p: True], [0], [lambda p: theta]]
    p = _problem((3, 2, 2), bc_u, bc_T)
    u, T = solver(p, {'jax_solver': {'tol': 1e-13, 'atol': 0.}})
    e = KAPPA * theta / (3. * TM_LM + 2. * TM_MU)
    np.testing.assert_allclose(u.cpu().numpy(), e * p.fes[0].points, rtol=0, atol=1e-10 * e)
    np.testing.assert_allclose(T.cpu().numpy()[:, 0], theta, rtol=1e-14)


def test_several_variables_refuse_the_amg_path(gpu):
    bc_u = [[_plane(0)] * 3, [0, 1, 2], [lambda p: 0.] * 3]
    p = _problem((2, 2, 2), bc_u, [[_plane(0)], [0], [lambda p: 1.]])
    with pytest.raises(NotImplementedError):
        solver(p, {'amgx_solver': {}})
