"""The oracle's adjoint against the one ad_wrapper output the reference holds.

docs/source/learn/compute_gradients/example.ipynb (cells 11-31) solves a
Neo-Hookean unit cube on a 5x5x5 HEX8 mesh with mu, kappa scaled by a
per-quadrature-point density rho = 0.5 (E = 1e6, nu = 0.3), u = (0, 0, 1) on
z = 0, surface map (0, 0, 1e3) on z = 1, J = sum(u^2), and prints the
ad_wrapper gradient drho[0, 0] = 0.002266954599447443 (cell 31).  The gmsh
box's cell 0 is the corner cell at the origin; its quadrature point 0 is the
one nearest the origin, (0.0423, 0.0423, 0.0423).  box_mesh numbers that cell
and point 0 as well, and the four bottom-corner points are equivalent by the
problem's symmetry.
"""
import numpy as np

from oracle import fem as OF

DRHO_00 = 0.002266954599447443      # example.ipynb cell 31 (AD)
DRHO_00_FD = 0.0022666187078357325  # same cell, forward difference with h = 1e-3


def notebook_problem(rho=0.5, n=5):
    pts, cells = OF.box_mesh(n, n, n, 1., 1., 1.)
    E, nu = 1e6, 0.3
    mu, kappa = E / (2 * (1 + nu)), E / (3 * (1 - 2 * nu))
    bottom = lambda p: np.isclose(p[2], 0., atol=1e-5)
    top = lambda p: np.isclose(p[2], 1., atol=1e-5)
    bc = [[bottom] * 3, [0, 1, 2], [lambda p: 0., lambda p: 0., lambda p: 1.]]
    o = OF.Problem(pts, cells, 3, bc_info=bc, mat=OF.Material('neo_hookean', mu=mu, kappa=kappa))
    o.theta = np.full(o.JxW.shape, rho)
    bi = OF.boundary_inds(o.points, o.cells, o.face_inds, top)
    o.add_surface(bi, np.broadcast_to(np.array([0., 0., 1e3]), (len(bi), o.fw.shape[1], 3)))
    return o


def test_oracle_adjoint_matches_the_reference_ad_wrapper_value():
    o = notebook_problem()
    u = o.newton(linear='direct')
    g, _ = o.adjoint_theta(u, 2. * u)          # v = dJ/du = 2u
    assert abs(g[0, 0] - DRHO_00) <= 1e-9 * DRHO_00
    # the symmetry-equivalent points of the other bottom corners
    xq = o.quad_points()
    bottom_corners = np.all(np.minimum(xq, 1. - xq) < 0.05, axis=-1) & (xq[..., 2] < 0.05)
    assert bottom_corners.sum() == 4
    assert np.all(np.abs(g[bottom_corners] - DRHO_00) <= 1e-9 * DRHO_00)
    # the notebook's forward difference (h = 1e-3 relative on rho[0, 0]) agrees to its truncation error
    h = 1e-3
    o2 = notebook_problem()
    o2.theta[0, 0] *= 1. + h
    J0, J1 = float(u @ u), float(np.sum(o2.newton(linear='direct') ** 2))
    fd = (J1 - J0) / (h * 0.5)
    assert abs(fd - DRHO_00_FD) <= 1e-4 * DRHO_00_FD
