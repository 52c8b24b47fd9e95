"""The oracle's several-variable problem (problem.py:58-119 layout, universal
kernels) pinned by closed forms of coupled thermoelasticity and by the
single-variable oracle when the coupling is switched off."""
import numpy as np

from oracle import fem as OF
from problems import (TM_ALPHA, TM_CAP, TM_COND, TM_DT, TM_LM, TM_MU, TM_T0, thermo_np)

KAPPA = TM_ALPHA * (2. * TM_MU + 3. * TM_LM)


def _plane(d):
    return lambda p: np.isclose(p[d], 0.)


def _linear(e, d):
    return lambda p: e * p[d]


def _problem(n, bc_u, bc_T, kappa=KAPPA):
    pts, cells = OF.box_mesh(*n, 1., 1., 1.)
    o = OF.MultiProblem(pts, cells, [3, 1], bc_infos=[bc_u, bc_T])
    o.add_cell_kernel(thermo_np(o.sv, kappa))
    return o


def test_layout_offsets_and_pattern():
    o = _problem((2, 2, 1), None, None)
    Nn = len(o.points)
    assert list(o.offset) == [0, 3 * Nn] and o.n == 4 * Nn
    inds = o.cell_inds()
    c0 = o.cells[0]
    # variable 0 then variable 1 (problem.py:86-92): 3 c0[a] + i, then 3 Nn + c0[a]
    np.testing.assert_array_equal(inds[0, :24], (3 * c0[:, None] + np.arange(3)).reshape(-1))
    np.testing.assert_array_equal(inds[0, 24:], 3 * Nn + c0)
    I, J = o.coo_pattern()
    assert I.shape == (len(o.cells) * 32 * 32,) and I[32] == inds[0, 1] and J[1] == inds[0, 1]


def test_uniform_heating_expands_freely():
    # T = theta everywhere, u pinned on three symmetry planes: the exact solution is the
    # free thermal expansion u = e x with 3 lmbda e + 2 mu e = kappa theta (sigma = 0)
    theta = 40.
    planes = [_plane(d) for d in range(3)]
    bc_u = [planes, [0, 1, 2], [lambda p: 0.] * 3]
    bc_T = [[lambda p: True], [0], [lambda p: theta]]
    o = _problem((3, 2, 2), bc_u, bc_T)
    u, T = o.sol_list(o.newton())
    e = KAPPA * theta / (3. * TM_LM + 2. * TM_MU)
    np.testing.assert_allclose(u, e * o.points, rtol=0, atol=1e-12 * e)
    np.testing.assert_allclose(T[:, 0], theta, rtol=1e-14)


def test_prescribed_expansion_sets_a_uniform_temperature():
    # u = e x prescribed at every node: the heat equation C/dt T + kappa T0 tr(eps)/dt - k lap T = 0
    # with natural boundaries has the uniform solution T = -3 kappa T0 e / C
    e = 1e-3
    bc_u = [[lambda p: True] * 3, [0, 1, 2], [_linear(e, d) for d in range(3)]]
    o = _problem((2, 3, 2), bc_u, None)
    u, T = o.sol_list(o.newton())
    np.testing.assert_allclose(T[:, 0], -3. * KAPPA * TM_T0 * e / TM_CAP, rtol=1e-10)


def test_without_coupling_equals_the_single_variable_oracle():
    left, right = (lambda p: np.isclose(p[0], 0.)), (lambda p: np.isclose(p[0], 1.))
    bc_u = [[left] * 3 + [right], [0, 1, 2, 0], [lambda p: 0.] * 3 + [lambda p: 0.01]]
    bc_T = [[left], [0], [lambda p: 10. * (1. + p[1])]]
    o = _problem((3, 3, 2), bc_u, bc_T, kappa=0.)
    u, T = o.sol_list(o.newton())
    el = OF.Problem(o.points, o.cells, 3, bc_info=bc_u, mat=OF.Material('elastic', lmbda=TM_LM, mu=TM_MU))
    ue = el.newton(linear='direct').reshape(-1, 3)
    assert np.linalg.norm(u - ue) <= 1e-10 * np.linalg.norm(ue)
    heat = OF.Problem(o.points, o.cells, 1, bc_info=bc_T, mat=OF.Material('poisson', k=TM_COND))
    heat.set_mass_map(lambda uq, xq: TM_CAP / TM_DT * uq)
    Th = heat.newton(linear='direct').reshape(-1, 1)
    assert np.linalg.norm(T - Th) <= 1e-10 * np.linalg.norm(Th)
