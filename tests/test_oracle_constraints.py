"""Pins of the oracle's constrained Newton (u = P u_r, reduced system P^T A P):
identity P reproduces the unconstrained solve, the periodic solution is
periodic and annihilates the reduced residual, and the constraint matrix ties
exactly the mapped node pairs."""
import numpy as np
import scipy.sparse as sp

from oracle import fem as OF
from problems import periodic_p_mat


def _poisson(nx=6, ny=4, nz=3):
    pts, cells = OF.box_mesh(nx, ny, nz, 1., 1., 1.)
    bc = [[lambda p: abs(p[2]) < 1e-9, lambda p: abs(p[2] - 1.) < 1e-9], [0, 0], [lambda p: 0., lambda p: 0.]]
    o = OF.Problem(pts, cells, 1, mat=OF.Material('poisson', k=1.), bc_info=bc)
    xq = o.quad_points()
    o.add_mass(-(xq[..., 0:1] * np.sin(5 * np.pi * xq[..., 1:2]) + 1.))
    return o, pts


def test_identity_constraint_is_the_plain_solve():
    o, _ = _poisson()
    ref = o.newton(linear='direct')
    got = o.newton_constrained(sp.identity(o.n, format='csr'))
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_periodic_solution():
    o, pts = _poisson()
    left, right = (lambda x: abs(x[0]) < 1e-9), (lambda x: abs(x[0] - 1.) < 1e-9)
    P = periodic_p_mat(pts, 1, [(left, right, lambda x: x + np.array([1., 0., 0.]), 0)])
    il, ir = np.flatnonzero(np.isclose(pts[:, 0], 0.)), np.flatnonzero(np.isclose(pts[:, 0], 1.))
    assert P.shape == (len(pts), len(pts) - len(ir)) and P.nnz == len(pts)
    u = o.newton_constrained(P)
    key = lambda ids: ids[np.lexsort((pts[ids, 2], pts[ids, 1]))]
    assert np.array_equal(u[key(il)], u[key(ir)])
    assert np.abs(P.T @ o.residual(u)).max() < 1e-9 * np.abs(o.load).max()
    # differs from the solution without periodicity (natural conditions on x = 0, 1)
    assert np.abs(u - o.newton(linear='direct')).max() > 1e-4
