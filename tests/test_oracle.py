"""CPU pins of the oracle (oracle/): golden vectors, closed forms, invariants.

Each test is chosen so that a plausible mistake in the oracle (a dropped term,
a transposed operand, a wrong sign or index) fails it; none retypes the
oracle's own formula.  The fixtures in tests/golden/ are the reference's
FEniCSx solutions (tests/benchmarks/*/fenicsx/ of jax-fem).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from conftest import golden
from oracle import basis as OB
from oracle import fem as OF

LAME = OF.lame(70e3, 0.3)


def _close(fn_coord, val):
    return lambda p: np.isclose(p[fn_coord], val, atol=1e-5)


# ---------------------------------------------------------------- basis
@pytest.mark.parametrize('ele', ['HEX8', 'HEX27'])
def test_basis_kronecker_and_partition_of_unity(ele):
    X = OB.nodes(ele)
    v, g = OB.tabulate(ele, X)
    np.testing.assert_allclose(v, np.eye(len(X)), atol=1e-12)
    pts = np.random.default_rng(0).random((50, 3))
    v, g = OB.tabulate(ele, pts)
    np.testing.assert_allclose(v.sum(1), 1., atol=1e-12)
    np.testing.assert_allclose(g.sum(1), 0., atol=1e-11)
    # gradients against central differences of the values
    h = 1e-6
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        fd = (OB.tabulate(ele, pts + e)[0] - OB.tabulate(ele, pts - e)[0]) / (2 * h)
        np.testing.assert_allclose(g[..., d], fd, atol=1e-7)


def test_basis_reproduces_linear_and_quadratic_fields():
    pts = np.random.default_rng(1).random((20, 3))
    for ele, f in (('HEX8', lambda x: 1 + 2 * x[:, 0] - x[:, 1] + 3 * x[:, 2] + x[:, 0] * x[:, 1] * x[:, 2]),
                   ('HEX27', lambda x: (x[:, 0] ** 2) * x[:, 1] * (x[:, 2] ** 2) - x[:, 1] ** 2 + 0.5)):
        v, _ = OB.tabulate(ele, pts)
        np.testing.assert_allclose(v @ f(OB.nodes(ele)), f(pts), atol=1e-12)


@pytest.mark.parametrize('order', [2, 3, 4, 5])
def test_gauss_exactness(order):
    pts, w = OB.quad_rule(order, 3)
    n = (order + 2) // 2
    assert len(w) == n ** 3
    np.testing.assert_allclose(w.sum(), 1., atol=1e-14)
    # exact for x^a y^b z^c, a, b, c <= 2n - 1
    for a, b, c in ((2 * n - 1, 0, 1), (1, 2 * n - 1, 2 * n - 2), (2 * n - 1, 2 * n - 1, 2 * n - 1)):
        np.testing.assert_allclose((w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c).sum(),
                                   1. / ((a + 1) * (b + 1) * (c + 1)), rtol=1e-13)


def test_hex8_default_quadrature_points():
    _, _, w = OB.shape_vals_and_grads('HEX8')
    pts, _ = OB.quad_rule(2, 3)
    assert len(w) == 8
    np.testing.assert_allclose(np.sort(np.unique(np.round(pts, 12))), [0.5 - np.sqrt(3) / 6, 0.5 + np.sqrt(3) / 6])
    # basix tensor order: first coordinate slowest
    assert pts[0, 0] == pts[1, 0] and pts[0, 2] != pts[1, 2]


def test_face_tables_area_and_normals():
    fv, fg, fw, fn, finds = OB.face_shape_vals_and_grads('HEX8')
    np.testing.assert_allclose(fw.sum(1), 1.)
    X = OB.HEX8_NODES
    for f in range(6):
        # the four face vertices lie on the plane with outward normal fn[f]
        c = X[finds[f]] @ fn[f]
        assert np.ptp(c) == 0 and c[0] == (1. if fn[f].sum() > 0 else 0.)
        # only the face's own vertices are non-zero on the face
        others = np.setdiff1d(np.arange(8), finds[f])
        np.testing.assert_allclose(fv[f][:, others], 0., atol=1e-14)


# ------------------------------------------------------------- geometry
def test_affine_geometry_closed_form():
    pts, cells = OF.box_mesh(3, 2, 2, 1.5, 1., 0.5)
    A = np.array([[1.2, 0.3, 0.], [0.1, 0.9, 0.2], [0., -0.4, 1.1]])
    x = pts @ A.T + np.array([0.3, -1., 2.])
    _, g, w = OB.shape_vals_and_grads('HEX8')
    G0, JxW0 = OF.shape_grads(pts, cells, g, w)
    G, JxW = OF.shape_grads(x, cells, g, w)
    h = (0.5, 0.5, 0.25)
    np.testing.assert_allclose(JxW0, np.prod(h) / 8, rtol=1e-14)
    np.testing.assert_allclose(JxW, JxW0 * np.linalg.det(A), rtol=1e-13)
    # grad_x N = grad_X N A^-1 under x = A X + b
    np.testing.assert_allclose(G, G0 @ np.linalg.inv(A), atol=1e-12)
    np.testing.assert_allclose(G0[..., 0][:, 0, 1], 1 / h[0] * (1 - (0.5 - np.sqrt(3) / 6)) ** 2, rtol=1e-12)


def test_distorted_volume_and_cylinder_area():
    d = golden('linear_elasticity_cylinder')
    p = OF.Problem(d['points'], d['cells'], 3)
    bi = OF.boundary_inds(p.points, p.cells, p.face_inds, _close(2, 10.))
    _, ns = OF.face_shape_grads(p.points, p.cells, bi, p.fg, p.fw, p.fn)
    # the fixture mesh is an extrusion of its top face over height 10 (cells of
    # straight vertical edges): volume = area * height; the faceted disc of
    # radius 5 under-fills the circle
    np.testing.assert_allclose(p.JxW.sum(), 10. * ns.sum(), rtol=1e-12)
    assert 0.97 * 25 * np.pi < ns.sum() < 25 * np.pi
    np.testing.assert_allclose(ns.sum(), float(d['surface_area']), rtol=1e-10)
    np.testing.assert_allclose(ns.sum(), 76.53668647, rtol=1e-9)


# ------------------------------------------------------------- kernels
def _distorted_mesh(nx=2, ny=2, nz=1, seed=0):
    pts, cells = OF.box_mesh(nx, ny, nz, 1., 1., 1.)
    pts = pts + 0.08 * np.random.default_rng(seed).uniform(-1, 1, pts.shape)
    return pts, cells


MATS = {
    'poisson': (1, OF.Material('poisson', k=1.7), None),
    'elastic': (3, OF.Material('elastic', lmbda=LAME[0], mu=LAME[1]), None),
    'simp': (3, OF.Material('simp', lmbda1=0.577, mu1=0.385, Emin=1e-3, E0=1., p=3.), 'theta'),
    'neo_hookean': (3, OF.Material('neo_hookean', mu=3.846, kappa=8.333), None),
}


@pytest.mark.parametrize('name', list(MATS))
def test_tangent_is_derivative_of_residual(name):
    V, mat, th = MATS[name]
    pts, cells = _distorted_mesh()
    _, g, w = OB.shape_vals_and_grads('HEX8')
    G, JxW = OF.shape_grads(pts, cells, g, w)
    rng = np.random.default_rng(3)
    theta = rng.uniform(0.4, 0.9, JxW.shape) if th else None
    u = 0.05 * rng.standard_normal((len(pts), V))
    K = OF.cell_tangent(mat, u, cells, G, JxW, theta)
    h = 1e-6
    for c in range(len(cells)):
        for b in range(8 * V):
            up, um = u.copy(), u.copy()
            up[cells[c, b // V], b % V] += h
            um[cells[c, b // V], b % V] -= h
            fd = (OF.cell_residual(mat, up, cells, G, JxW, theta)[c] -
                  OF.cell_residual(mat, um, cells, G, JxW, theta)[c]).reshape(-1) / (2 * h)
            np.testing.assert_allclose(K[c][:, b], fd, rtol=1e-6, atol=1e-7 * np.abs(K).max())


def test_elastic_tangent_rigid_body_null_space_and_symmetry():
    pts, cells = _distorted_mesh()
    _, g, w = OB.shape_vals_and_grads('HEX8')
    G, JxW = OF.shape_grads(pts, cells, g, w)
    K = OF.cell_tangent(MATS['elastic'][1], np.zeros((len(pts), 3)), cells, G, JxW)
    np.testing.assert_allclose(K, np.swapaxes(K, 1, 2), atol=1e-9 * np.abs(K).max())
    for c in range(len(cells)):
        X = pts[cells[c]]
        modes = [np.tile(e, (8, 1)) for e in np.eye(3)]
        modes += [np.cross(a, X) for a in np.eye(3)]          # infinitesimal rotations
        for m in modes:
            np.testing.assert_allclose(K[c] @ m.reshape(-1), 0., atol=1e-9 * np.abs(K).max())


def test_patch_test_linear_field():
    pts, cells = _distorted_mesh(3, 3, 3, seed=5)
    _, g, w = OB.shape_vals_and_grads('HEX8')
    G, JxW = OF.shape_grads(pts, cells, g, w)
    B = np.array([[0.01, 0.002, -0.003], [0.004, -0.02, 0.001], [0., 0.003, 0.015]])
    u = pts @ B.T
    r = OF.scatter(OF.cell_residual(MATS['elastic'][1], u, cells, G, JxW), cells, len(pts))
    base, _ = OF.box_mesh(3, 3, 3, 1., 1., 1.)
    interior = np.all((base > 1e-9) & (base < 1 - 1e-9), axis=1)
    assert interior.sum() == 8
    np.testing.assert_allclose(r[interior], 0., atol=1e-9 * np.abs(r).max())


def test_neo_hookean_stress_is_gradient_of_energy():
    # psi of jax-fem tests/benchmarks/hyperelasticity/test_hyper_elasticity.py:17-26
    mu, kappa = 3.846, 8.333

    def psi(F):
        J = torch.det(F)
        Jinv = J ** (-2. / 3.)
        I1 = torch.sum(F * F)
        return mu / 2. * (Jinv * I1 - 3.) + kappa / 2. * (J - 1.) ** 2

    rng = np.random.default_rng(7)
    G = 0.2 * rng.standard_normal((5, 3, 3))
    P = MATS['neo_hookean'][1].stress(G)
    for i in range(5):
        F = torch.tensor(G[i] + np.eye(3), dtype=torch.float64, requires_grad=True)
        psi(F).backward()
        np.testing.assert_allclose(P[i], F.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_simp_dstress_dtheta_central_difference():
    mat = MATS['simp'][1]
    rng = np.random.default_rng(9)
    G = 0.01 * rng.standard_normal((4, 3, 3))
    th = rng.uniform(0.3, 0.9, 4)
    h = 1e-6
    fd = (mat.stress(G, th + h) - mat.stress(G, th - h)) / (2 * h)
    np.testing.assert_allclose(mat.dstress_dtheta(G, th), fd, rtol=1e-7, atol=1e-14)


# ------------------------------------------------------ global matrix
@pytest.mark.parametrize('dims,V', [((3, 2, 2), 1), ((2, 3, 2), 3)])
def test_csr_pattern_closed_form_and_dense_assembly(dims, V):
    pts, cells = OF.box_mesh(*dims, 1., 1., 1.)
    n = len(pts) * V
    rng = np.random.default_rng(2)
    Ke = rng.standard_normal((len(cells), 8 * V, 8 * V))
    A = OF.csr_from_cells(Ke, cells, V, n)
    assert A.nnz == np.prod([3 * d + 1 for d in dims]) * V * V
    assert np.all(np.diff(A.indptr) > 0) and A.has_sorted_indices
    D = np.zeros((n, n))
    for c in range(len(cells)):                       # brute-force loop assembly
        dofs = [V * cells[c, a] + i for a in range(8) for i in range(V)]
        for r in range(8 * V):
            for s in range(8 * V):
                D[dofs[r], dofs[s]] += Ke[c, r, s]
    np.testing.assert_allclose(A.toarray(), D, rtol=1e-14, atol=1e-14)
    Z = OF.zero_rows(A, np.array([0, 5]))
    assert Z.nnz == A.nnz
    np.testing.assert_array_equal(Z.toarray()[0], np.eye(n)[0])
    np.testing.assert_array_equal(Z.toarray()[5], np.eye(n)[5])


def test_krylov_solvers_match_direct_solve():
    pts, cells = OF.box_mesh(4, 3, 3, 1., 1., 1.)
    p = OF.Problem(pts, cells, 3, mat=MATS['elastic'][1],
                   bc_info=[[_close(0, 0.)] * 3, [0, 1, 2], [lambda x: 0.] * 3])
    A = p.tangent(np.zeros(p.n))
    b = np.random.default_rng(4).standard_normal(p.n)
    b[p.bc_dofs] = 0.
    x = spla.spsolve(A.tocsc(), b)
    for solve in (OF.pcg_jacobi, OF.bicgstab_jax):
        y, k = solve(A, b, np.zeros(p.n), A.diagonal(), tol=1e-12, atol=0.)
        assert 0 < k < 2000
        np.testing.assert_allclose(y, x, rtol=0, atol=1e-9 * np.abs(x).max())


# ------------------------------------------------------ FEniCSx fixtures
def test_fixture_linear_poisson():
    d = golden('linear_poisson')
    p = OF.Problem(d['points'], d['cells'], 1, mat=OF.Material('poisson', k=1.),
                   bc_info=[[_close(0, 0.), _close(0, 1.)], [0, 0], [lambda x: 0., lambda x: 1.]])
    u = p.newton(linear='direct')
    np.testing.assert_allclose(u, d['sol'][:, 0], atol=1e-5)
    np.testing.assert_allclose(u, d['points'][:, 0], atol=1e-9)     # closed form u = x


def test_fixture_linear_elasticity_cube():
    d = golden('linear_elasticity_cube')
    p = OF.Problem(d['points'], d['cells'], 3, mat=MATS['elastic'][1],
                   bc_info=[[_close(0, 0.)] * 3, [0, 1, 2], [lambda x: 1.] * 3])
    p.add_mass(np.broadcast_to(-np.array([0., 10., 10.]), p.JxW.shape + (3,)))
    bi = OF.boundary_inds(p.points, p.cells, p.face_inds, _close(0, 1.))
    p.add_surface(bi, np.broadcast_to(-np.array([10., 0., 0.]), (len(bi), p.fw.shape[1], 3)))
    u = p.newton(linear='direct')
    np.testing.assert_allclose(u.reshape(-1, 3), d['sol'], atol=1e-5)


def test_fixture_linear_elasticity_cylinder():
    d = golden('linear_elasticity_cylinder')
    p = OF.Problem(d['points'], d['cells'], 3, mat=MATS['elastic'][1],
                   bc_info=[[_close(2, 0.)] * 3, [0, 1, 2], [lambda x: 0.] * 3])
    xq = p.quad_points()
    p.add_mass(-np.stack([1e3 * xq[..., 0], 2e3 * xq[..., 1], 3e3 * xq[..., 2]], -1))
    bi = OF.boundary_inds(p.points, p.cells, p.face_inds, _close(2, 10.))
    xf = OF.face_quad_points(p.points, p.cells, bi, p.fv)
    p.add_surface(bi, -np.stack([1e3 * xf[..., 0] ** 2 + 1e3 * xf[..., 1] ** 2, 0 * xf[..., 0], 0 * xf[..., 0]], -1))
    u = p.newton(linear='direct')
    np.testing.assert_allclose(u.reshape(-1, 3), d['sol'], atol=1e-3)


def test_fixture_hyperelasticity_solution_and_traction():
    d = golden('hyperelasticity')
    E, nu = 1e3, 0.3
    mat = OF.Material('neo_hookean', mu=E / (2 * (1 + nu)), kappa=E / (3 * (1 - 2 * nu)))
    bottom, top = _close(2, 0.), _close(2, 10.)
    p = OF.Problem(d['points'], d['cells'], 3, mat=mat,
                   bc_info=[[bottom] * 3 + [top] * 3, [0, 1, 2, 0, 1, 2], [lambda x: 0.] * 5 + [lambda x: 1.]])
    u = p.newton(linear='direct')
    np.testing.assert_allclose(u.reshape(-1, 3), d['sol'], atol=1e-5)
    bi = OF.boundary_inds(p.points, p.cells, p.face_inds, top)
    Gf, ns = OF.face_shape_grads(p.points, p.cells, bi, p.fg, p.fw, p.fn)
    Gu = np.einsum('sav,sqaj->sqvj', u.reshape(-1, 3)[p.cells[bi[:, 0]]], Gf)
    traction = ((mat.stress(Gu) @ np.array([0., 0., 1.])) * ns[..., None]).sum((0, 1))
    np.testing.assert_allclose(traction[2], float(d['traction']), rtol=1e-6)


# --------------------------------------------------------------- adjoint
def test_adjoint_matches_central_differences_and_self_adjointness():
    pts, cells = OF.box_mesh(4, 2, 2, 2., 1., 1.)
    mat = MATS['simp'][1]
    rng = np.random.default_rng(1)
    theta_c = 0.5 + rng.uniform(-0.05, 0.05, len(cells))
    left = _close(0, 0.)

    def build(th):
        p = OF.Problem(pts, cells, 3, mat=mat, theta=np.repeat(th[:, None], 8, 1),
                       bc_info=[[left] * 3, [0, 1, 2], [lambda x: 0.] * 3])
        bi = OF.boundary_inds(p.points, p.cells, p.face_inds, _close(0, 2.))
        p.add_surface(bi, np.broadcast_to(np.array([0., 0., 1.]), (len(bi), p.fw.shape[1], 3)))
        return p

    def compliance(th):
        p = build(th)
        u = p.newton(linear='direct')
        return -p.load.reshape(-1) @ u, u, p

    J0, u, p = compliance(theta_c)
    f = -p.load.reshape(-1)
    g_q, lam = p.adjoint_theta(u, f)
    free = np.setdiff1d(np.arange(p.n), p.bc_dofs)
    np.testing.assert_allclose(lam[free], u[free], rtol=1e-10, atol=1e-14)      # lambda = u for compliance
    g = g_q.sum(1)
    # closed form: dC/dtheta_e = -u_e^T dK_e/dtheta_e u_e
    h = 1e-6
    for e in rng.choice(len(cells), 6, replace=False):
        tp, tm = theta_c.copy(), theta_c.copy()
        tp[e] += h
        tm[e] -= h
        fd = (compliance(tp)[0] - compliance(tm)[0]) / (2 * h)
        np.testing.assert_allclose(g[e], fd, rtol=1e-5)
        assert g[e] < 0


def test_j2_tangent_is_derivative_of_return_map():
    E, nu = 70e3, 0.3
    lm, mu = OF.lame(E, nu)
    m = OF.Material('j2', lmbda=lm, mu=mu, sig0=250.)
    rng = np.random.default_rng(0)
    sym = lambda a: 0.5 * (a + np.swapaxes(a, -1, -2))
    G = 0.01 * rng.standard_normal((6, 3, 3))
    st = np.concatenate([sym(100 * rng.standard_normal((6, 3, 3))).reshape(6, 9),
                         sym(1e-3 * rng.standard_normal((6, 3, 3))).reshape(6, 9)], 1)
    st[0] = 0.
    G[0] *= 1e-4                                   # an elastic point
    beta = m._j2(G, st)[2]
    assert beta[0] == 0 and np.all(beta[1:] > 0)   # and plastic ones
    A = m.tangent(G, st)
    h = 1e-7
    for k in range(3):
        for l in range(3):
            E_ = np.zeros((3, 3))
            E_[k, l] = h
            fd = (m.stress(G + E_, st) - m.stress(G - E_, st)) / (2 * h)
            np.testing.assert_allclose(A[..., k, l], fd, rtol=1e-5, atol=1e-6 * np.abs(A).max())
    # returned stresses lie on the yield surface
    s = m.stress(G, st)
    sd = s - np.trace(s, axis1=-2, axis2=-1)[..., None, None] / 3 * np.eye(3)
    np.testing.assert_allclose(np.sqrt(1.5 * (sd * sd).sum((-2, -1)))[1:], 250., rtol=1e-12)


def test_fixture_plasticity_solution_and_average_stress():
    d = golden('plasticity')
    lm, mu = OF.lame(70e3, 0.3)
    pts, cells = d['points'], d['cells']
    top, bottom = _close(2, 10.), _close(2, 0.)
    probe = OF.Problem(pts, cells, 3)
    C, Q = probe.JxW.shape
    state = np.zeros((C, Q, 18))
    u = None
    avg = []
    # load history of test_plasticity.py:105-145: top face pulled to each disp in turn
    for disp in d['disps']:
        bc = [[bottom] * 3 + [top] * 3, [0, 1, 2, 0, 1, 2], [lambda x: 0.] * 5 + [lambda x: float(disp)]]
        p = OF.Problem(pts, cells, 3, mat=OF.Material('j2', lmbda=lm, mu=mu, sig0=250.), theta=state, bc_info=bc)
        u = p.newton(linear='direct')
        # update_stress_strain (test_plasticity.py:72-77)
        Gu = np.einsum('cav,cqaj->cqvj', u.reshape(-1, 3)[cells], p.G)
        sig = p.mat.stress(Gu, state)
        state = np.concatenate([sig.reshape(C, Q, 9), (0.5 * (Gu + np.swapaxes(Gu, -1, -2))).reshape(C, Q, 9)], 2)
        avg.append((sig[..., 2, 2] * p.JxW).sum() / p.JxW.sum())
    np.testing.assert_allclose(u.reshape(-1, 3), d['sol'], atol=1e-5)
    np.testing.assert_allclose(avg, d['avg_stresses'], atol=1e-3)


def test_explicit_waveguide_matches_the_wave_solution():
    # applications/explicit_dynamics/example.py: the residual-only path under central differences
    import explicit_waveguide as W
    pts, cells = OF.box_mesh(W.NX, W.NY, W.NZ, W.LX, W.LY, W.LZ)
    lm, mu = OF.lame(W.E, W.NU)
    o = OF.Problem(pts, cells, 3, mat=OF.Material('elastic', lmbda=lm, mu=mu),
                   bc_info=[W.locations(np.isclose), [0, 1, 2], [lambda p: 0.] * 3])
    c, k, steps, dt = W.time_grid(lm, mu)
    em = W.RHO * np.einsum('qi,eq->ei', o.sv, o.JxW)
    nm = np.zeros(len(pts))
    np.add.at(nm, cells.reshape(-1), em.reshape(-1))
    mass = np.repeat(nm[:, None], 3, 1)
    fixed = np.zeros((len(pts), 3), bool)
    fixed.reshape(-1)[o.bc_dofs] = True
    u0 = np.zeros((len(pts), 3))
    u0[:, 0] = W.AMPLITUDE * np.sin(k * pts[:, 0])
    acc = lambda u: np.where(fixed, 0.0, -o.residual(u.reshape(-1), bc=False).reshape(-1, 3) / mass)
    hist = np.array(W.run(acc, u0, fixed, dt, steps, np))
    speed_err, l2 = W.wave_checks(hist[:, :, 0], nm, pts[:, 0], c, k, dt)
    assert steps == 640 and speed_err < 1e-4 and l2 < 1e-3, (steps, speed_err, l2)
