"""GPU tests of the public entry points around the Newton solve: newton_update,
the reference's COO value export Problem.V (cell blocks, then face blocks),
get_A, the line-search branch, and the re-check of a closed-form law
recognised from a user tensor map (problem.py:172-182, 447-491;
solver.py:161-176, 233-234, 424-463, 540-553)."""
import logging

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle_setup
from oracle import fem as OF
from problems import box, isclose, threshold_np, threshold_torch

from jax_fem_amd import materials as M
from jax_fem_amd.problem import Problem
from jax_fem_amd.solver import get_A, solver

pytestmark = pytest.mark.gpu

H_ROBIN = 3.0


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


class NHRobin(Problem):
    """Neo-Hookean with a body force and a Robin face term s(u) = h u on x = 1 (a surface map
    that depends on u, so the reference's face blocks of V are nonzero)."""
    def get_tensor_map(self):
        return M.NeoHookean(E=10., nu=0.3)

    def get_mass_map(self):
        return lambda u, x: torch.tensor([0.1, -0.2, 0.3], dtype=torch.float64)

    def get_surface_maps(self):
        return [lambda u, x: H_ROBIN * u]


def _nh_robin(n=(3, 2, 2)):
    left = lambda x: isclose(x[0], 0.)
    return NHRobin(box(*n, 1.5, 1., 1.), vec=3, dim=3,
                   dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda x: 0.] * 3],
                   location_fns=[lambda x: isclose(x[0], 1.5)])


def _oracle_nh(p):
    mu, kappa = 10. / 2.6, 10. / (3. * 0.4)
    o = oracle_setup.oracle_problem(p, OF.Material('neo_hookean', mu=mu, kappa=kappa))
    bi = OF.boundary_inds(o.points, o.cells, o.face_inds, p.location_fns[0])
    # the Robin term as a face kernel: per face sum_q N_a h u_q nanson_q (problem.py:238-259)
    fv = o.fv[bi[:, 1]]
    _, ns = OF.face_shape_grads(o.points, o.cells, bi, o.fg, o.fw, o.fn)

    def robin(cs, x, fvals, fgrads, nanson):
        u = np.einsum('sqa,sav->sqv', fvals, cs.reshape(len(cs), -1, 3))
        return np.einsum('sqa,sqv,sq->sav', fvals, H_ROBIN * u, nanson[:, 0]).reshape(len(cs), -1)
    o.add_face_kernel(bi, robin)
    return o, bi, fv, ns


def test_newton_update_V_and_get_A_vs_oracle(gpu):
    p = _nh_robin()
    o, bi, fv, ns = _oracle_nh(p)
    u = 0.05 * np.random.default_rng(4).standard_normal(o.n)
    ut = torch.as_tensor(u, device='cuda').reshape(-1, 3)
    res = p.newton_update([ut])[0].cpu().numpy().reshape(-1)
    assert rel(res, o.residual(u, bc=False)) < 1e-12
    # get_A: the row-eliminated tangent of that newton_update (PETSc zeroRows, diag 1)
    indptr, indices, data = get_A(p).export_csr()
    A = o.tangent(u)
    assert np.array_equal(indptr, A.indptr) and np.array_equal(indices, A.indices)
    assert rel(data, A.data) < 1e-12
    # V in the reference's COO order: sum over (I, J) is the tangent before zeroRows
    V = p.V
    I, J = p.I, p.J
    assert V.shape == I.shape == J.shape
    Acoo = sp.coo_matrix((V, (I, J)), shape=(o.n, o.n)).tocsr()
    Acoo.sum_duplicates()
    Acoo.sort_indices()
    A0 = o.tangent(u, bc=False)
    assert rel(Acoo.toarray(), A0.toarray()) < 1e-12
    # the cell blocks: laplace and mass kernels, nothing from the faces
    C_, nv = o.cells.shape[0], 24
    cells = V[:C_ * nv * nv].reshape(C_, nv, nv)
    assert rel(cells, OF.cell_tangent(o.mat, u.reshape(-1, 3), o.cells, o.G, o.JxW)) < 1e-12
    # the face blocks: the Robin term's jacfwd, h sum_q N_a N_b nanson delta_ik in closed form
    faces = V[C_ * nv * nv:].reshape(len(bi), 8, 3, 8, 3)
    Mf = H_ROBIN * np.einsum('sqa,sqb,sq->sab', fv, fv, ns)
    assert rel(faces, np.einsum('sab,ik->saibk', Mf, np.eye(3))) < 1e-12


def test_line_search_branch_vs_oracle(gpu):
    p = _nh_robin((4, 2, 2))
    sol = solver(p, {'jax_solver': {'precision': 'fp64', 'tol': 1e-12, 'atol': 0.}, 'line_search_flag': True})[0]
    assert p._solve_stats['newton_iterations'] >= 2
    o, *_ = _oracle_nh(p)
    ref = o.newton(linear='direct')
    assert np.linalg.norm(sol.cpu().numpy().reshape(-1) - ref) <= 1e-9 * np.linalg.norm(ref)


class Threshold(Problem):
    def get_tensor_map(self):
        return threshold_torch


class _Records(logging.Handler):
    def __init__(self):
        super().__init__(logging.WARNING)
        self.messages = []

    def emit(self, record):
        self.messages.append(record.getMessage())


def test_law_identification_guard_switches_to_the_user_map(gpu):
    # the law equals linear elasticity at every state identify() probes and stiffens above
    # eps:eps = 100, which the solve crosses: the device law is re-checked at the solve's
    # quadrature states and replaced by the map itself (tabulated), matching the oracle
    bottom, top = (lambda x: isclose(x[2], 0.)), (lambda x: isclose(x[2], 1.))
    bc = [[bottom] * 3 + [top], [0, 1, 2, 2], [lambda x: 0.] * 3 + [lambda x: 20.]]
    p = Threshold(box(3, 3, 3), vec=3, dim=3, dirichlet_bc_info=bc)
    dev = p.device()
    dev.prepare()
    assert isinstance(dev.material, M.LinearElastic)
    log, rec = logging.getLogger('jax_fem_amd'), _Records()
    log.addHandler(rec)
    try:
        sol = solver(p, {'jax_solver': {'precision': 'fp64', 'tol': 1e-12, 'atol': 0.}})[0]
    finally:
        log.removeHandler(rec)
    assert dev.material.tabulated
    assert any('switching to the tabulated law' in m for m in rec.messages)
    o = oracle_setup.oracle_problem(p, OF.Material('user', fn=threshold_np))
    ref = o.newton(linear='direct')
    assert np.linalg.norm(sol.cpu().numpy().reshape(-1) - ref) <= 1e-8 * np.linalg.norm(ref)
