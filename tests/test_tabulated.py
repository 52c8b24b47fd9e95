"""Tensor maps outside the closed-form laws (the tabulated law), on the CPU:
the oracle's complex-step derivatives against the closed-form tangents they
replace, the torch point evaluation against the oracle, and the host rules
(fallback of `identify`, symmetry requirement)."""
import numpy as np
import pytest
import torch

from oracle import fem as OF
from problems import (diffusion_np, diffusion_torch, scaled_svk_np, scaled_svk_torch, svk_np, svk_torch,
                      fiber_torch)

from jax_fem_amd import materials as M


def _grads(n, V, scale, seed=0):
    return np.random.default_rng(seed).uniform(-scale, scale, (n, V, 3))


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


@pytest.mark.parametrize('kind', ['elastic', 'neo_hookean', 'poisson', 'simp'])
def test_complex_step_tangent_matches_closed_forms(kind):
    """The oracle's generic ('user') tangent is the complex-step derivative of the stress;
    on the closed-form laws it must reproduce their hand-derived tangents."""
    lm, mu = OF.lame(70e3, 0.3)
    law = {'elastic': OF.Material('elastic', lmbda=lm, mu=mu),
           'neo_hookean': OF.Material('neo_hookean', mu=10 / 2.6, kappa=10 / 1.2),
           'poisson': OF.Material('poisson', k=2.5),
           'simp': OF.Material('simp', lmbda1=0.57, mu1=0.38, Emin=1e-3, E0=1., p=3.)}[kind]
    V = 1 if kind == 'poisson' else 3
    G = _grads(40, V, 0.2)
    theta = np.random.default_rng(1).uniform(0.2, 1.0, 40) if kind == 'simp' else None
    user = OF.Material('user', fn=lambda g, th: law.stress(g, th))
    assert _rel(user.stress(G, theta), law.stress(G, theta)) == 0.0
    assert _rel(user.tangent(G, theta), law.tangent(G, theta)) < 1e-13
    if kind == 'simp':
        assert _rel(user.dstress_dtheta(G, theta), law.dstress_dtheta(G, theta)) < 1e-13


def test_svk_restatement_invariants():
    """SVK: zero stress under rigid rotations, the linear-elastic tangent at G = 0,
    a major-symmetric tangent (P derives from a potential)."""
    rng = np.random.default_rng(2)
    q = rng.standard_normal((10, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1).reshape(-1, 3, 3)
    assert np.abs(svk_np(R - np.eye(3))).max() < 1e-12
    user = OF.Material('user', fn=svk_np)
    lin = OF.Material('elastic', lmbda=40., mu=25.)
    G0 = np.zeros((1, 3, 3))
    assert _rel(user.tangent(G0), lin.tangent(G0)) < 1e-14
    A = user.tangent(_grads(30, 3, 0.3)).reshape(30, 9, 9)
    assert _rel(A, np.swapaxes(A, 1, 2)) < 1e-13


@pytest.mark.parametrize('case', ['svk', 'diffusion', 'scaled_svk'])
def test_point_evaluation_matches_oracle(case):
    """P and D = dP/dG from the torch point evaluation (vmap + forward-mode jvps)
    against the oracle's numpy map and its complex-step derivative."""
    fn, fnp, V = {'svk': (svk_torch, svk_np, 3), 'diffusion': (diffusion_torch, diffusion_np, 1),
                  'scaled_svk': (scaled_svk_torch, scaled_svk_np, 3)}[case]
    G = _grads(300, V, 0.25, seed=3)
    theta = np.random.default_rng(4).uniform(0.1, 1.2, 300)
    ivs = [torch.as_tensor(theta)] if case == 'scaled_svk' else []
    mat = M.identify(fn, V, iv_shapes=((),) if ivs else ())
    assert isinstance(mat, M.Tabulated)
    P, D = M.evaluate_points(mat, torch.as_tensor(G), ivs, jac=True)
    o = OF.Material('user', fn=fnp)
    th = theta if ivs else None
    assert _rel(P.numpy(), o.stress(G, th)) < 1e-14
    assert _rel(D.numpy(), o.tangent(G, th)) < 1e-13
    P2, D2 = M.evaluate_points(mat, torch.as_tensor(G), ivs, jac=False)
    assert D2 is None and torch.equal(P2, P)


def test_point_evaluation_broadcasting_map():
    """A map that only runs on batches cannot be checked per point and is refused;
    one written with broadcasting torch operations is evaluated."""
    def batched(u_grad):
        if u_grad.dim() != 3:
            raise RuntimeError("batched map")
        return diffusion_np_torch(u_grad)

    def diffusion_np_torch(g):
        return (1. + 4. * (g * g).sum((-2, -1)))[:, None, None] * g
    G = torch.as_tensor(_grads(50, 1, 0.3, seed=5))
    # per-point calls fail for this map, so the first-call check cannot run: evaluate_points
    # raises rather than using a form it could not check
    with pytest.raises(Exception):
        M.evaluate_points(M.Tabulated(batched, 1), G, [], jac=True)
    # a map that works both per point and on a batch
    both = lambda g: (1. + 4. * (g * g).sum((-2, -1), keepdim=True)) * g
    mat = M.Tabulated(both, 1)
    P, D = M.evaluate_points(mat, G, [], jac=True)
    o = OF.Material('user', fn=diffusion_np)
    assert _rel(P.numpy(), o.stress(G.numpy())) < 1e-14 and _rel(D.numpy(), o.tangent(G.numpy())) < 1e-13


def test_nonsymmetric_tangent_is_flagged():
    A = torch.as_tensor(np.random.default_rng(6).standard_normal((3, 3, 3, 3)))
    mat = M.Tabulated(lambda g: torch.einsum('ijkl,kl->ij', A, g), 3)
    with pytest.raises(NotImplementedError, match="not symmetric"):
        M.evaluate_points(mat, torch.as_tensor(_grads(8, 3, 0.1)), [], jac=True, symmetric='require')
    mat = M.Tabulated(lambda g: torch.einsum('ijkl,kl->ij', A, g), 3)
    M.evaluate_points(mat, torch.as_tensor(_grads(8, 3, 0.1)), [], jac=True)
    assert mat._pm.nonsymmetric
    sym = M.Tabulated(M.LinearElastic(E=1., nu=0.3), 3)
    M.evaluate_points(sym, torch.as_tensor(_grads(8, 3, 0.1)), [], jac=True)
    assert not sym._pm.nonsymmetric


def test_identify_falls_back_to_tabulated():
    assert isinstance(M.identify(svk_torch, 3), M.Tabulated)
    assert isinstance(M.identify(diffusion_torch, 1), M.Tabulated)
    assert isinstance(M.identify(fiber_torch, 3, iv_shapes=((3,),)), M.Tabulated)
    assert isinstance(M.identify(lambda G: G ** 3, 3), M.Tabulated)
    # closed-form laws are still recognised, and an explicit descriptor is taken as it is
    assert isinstance(M.identify(M.LinearElastic(E=1., nu=0.3), 3), M.LinearElastic)
    forced = M.identify(M.Tabulated(M.LinearElastic(E=1., nu=0.3)), 3)
    assert isinstance(forced, M.Tabulated) and forced.vec == 3
    with pytest.raises(ValueError):
        M.identify(M.Tabulated(svk_torch, 1), 3)


def test_non_finite_derivative_is_refused():
    """torch.func second derivatives through linalg.det are NaN at F = I: the table
    refuses them with a message instead of handing NaN to the solver."""
    from problems import HyperElasticity
    mat = M.Tabulated(HyperElasticity.get_tensor_map(None), 3)
    G = torch.as_tensor(_grads(6, 3, 0.1))
    P, D = M.evaluate_points(mat, G, [], jac=True)      # finite away from F = I
    assert torch.isfinite(D).all()
    with pytest.raises(NotImplementedError, match="not finite"):
        M.evaluate_points(mat, torch.zeros(4, 3, 3, dtype=torch.float64), [], jac=True)


def test_oracle_mass_tangent_is_the_residual_derivative():
    """The oracle's tangent with a u-dependent mass term (complex-step dm/du) against
    central differences of its residual."""
    from problems import spring_np, reaction_np
    for vec, fn, mat in ((3, spring_np, OF.Material('elastic', lmbda=1e3, mu=8e2)),
                         (1, reaction_np, OF.Material('poisson', k=1.))):
        pts, cells = OF.box_mesh(3, 2, 2, 1., 1., 1.)
        o = OF.Problem(pts, cells, vec, mat=mat)
        o.set_mass_map(fn)
        rng = np.random.default_rng(7)
        u, v = 0.3 * rng.standard_normal(o.n), rng.standard_normal(o.n)
        h = 1e-6
        fd = (o.residual(u + h * v, bc=False) - o.residual(u - h * v, bc=False)) / (2 * h)
        assert _rel(o.tangent(u, bc=False) @ v, fd) < 1e-7


def test_mass_point_map_matches_complex_step():
    from problems import spring_torch, spring_np
    rng = np.random.default_rng(8)
    U, X = rng.standard_normal((200, 3)), rng.uniform(0, 1, (200, 3))
    pm = M.PointMap(spring_torch, (3,), (3,), 'mass map')
    m, D = pm(torch.as_tensor(U), [torch.as_tensor(X)], jac=True)
    assert _rel(m.numpy(), spring_np(U, X)) < 1e-14
    assert _rel(D.numpy(), OF.mass_map_derivative(spring_np, U, X, ())) < 1e-13


def test_oracle_kernel_terms_match_closed_forms():
    """The oracle's universal-kernel terms (complex-step tangents) on a weak form that the
    closed-form oracle also writes: Poisson + cubic reaction + source, and a Robin face term."""
    from problems import energy_kernel_np, reaction_np, robin_kernel_np, ROBIN_H, ROBIN_UINF
    pts, cells = OF.box_mesh(3, 2, 2, 1., 1., 1.)
    a = OF.Problem(pts, cells, 1, mat=OF.Material('poisson', k=1.))
    a.set_mass_map(reaction_np)
    b = OF.Problem(pts, cells, 1, mat=OF.Material('poisson', k=0.))
    b.add_cell_kernel(energy_kernel_np(b.sv))
    u = 0.4 * np.random.default_rng(9).standard_normal(a.n)
    assert _rel(b.residual(u, bc=False), a.residual(u, bc=False)) < 1e-13
    assert _rel(b.tangent(u, bc=False).toarray(), a.tangent(u, bc=False).toarray()) < 1e-13
    # Robin face term = a u-independent surface load h u_inf plus h * (face mass matrix)
    bi = OF.boundary_inds(pts, cells, a.face_inds, lambda p: p[0] > 1 - 1e-9)
    c = OF.Problem(pts, cells, 1, mat=OF.Material('poisson', k=0.))
    c.add_face_kernel(bi, robin_kernel_np)
    d = OF.Problem(pts, cells, 1, mat=OF.Material('poisson', k=0.))
    d.add_surface(bi, np.full((len(bi), d.fw.shape[1], 1), -ROBIN_H * ROBIN_UINF))
    r0 = c.residual(np.zeros(c.n), bc=False)
    assert _rel(r0, d.residual(np.zeros(d.n), bc=False)) < 1e-14
    Kc = c.tangent(u, bc=False)
    assert _rel(c.residual(u, bc=False) - r0, Kc @ u) < 1e-13       # the face term is affine in u
    ones = np.ones(c.n)
    area = 1.0                                                      # the x = 1 face of the unit box
    assert abs(ones @ (Kc @ ones) - ROBIN_H * area) < 1e-12
