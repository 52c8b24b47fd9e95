"""GPU parity: every device stage against the FP64 oracle on the same inputs.

Tolerances: FP64 paths 1e-12 relative (geometry, tangents, residuals, matrix
values) and solver tolerance for solutions; FP32 tangents 2e-6 relative to the
largest entry; bit-exact for the sparsity pattern.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle_setup
from conftest import golden
from oracle import basis as OB
from oracle import fem as OF
from problems import (LinearPoisson, LinearElasticityCube, LinearElasticityCylinder, HyperElasticity,
                      Plasticity, SIMPElasticity, LoadedElastic, box, isclose)

from jax_fem_amd import _lib
from jax_fem_amd.generate_mesh import Mesh, box_mesh_hex27
from jax_fem_amd.problem import Problem
from jax_fem_amd.solver import solver, ad_wrapper
from jax_fem_amd import materials as M

pytestmark = pytest.mark.gpu


def _rng(seed=0):
    return np.random.default_rng(seed)


def distorted_box(n=(3, 4, 5), seed=0, amp=0.15):
    mesh = box(*n)
    pts = mesh.points.copy()
    h = 1.0 / max(n)
    pts += amp * h * _rng(seed).uniform(-1, 1, pts.shape)
    return Mesh(pts, mesh.cells)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


class ElasticLE(Problem):
    def get_tensor_map(self):
        return M.LinearElastic(E=70e3, nu=0.3)


class NH(Problem):
    def get_tensor_map(self):
        return M.NeoHookean(E=10., nu=0.3)


def oracle_problem(p, mat, theta=None):
    fe = p.fes[0]
    return OF.Problem(fe.points, fe.cells, fe.vec, ele_type=fe.ele_type, quadrature_order=fe.quadrature_order,
                      mat=mat, theta=theta)


# ------------------------------------------------------------------ geometry
@pytest.mark.parametrize('ele', ['HEX8', 'HEX27'])
def test_geometry_vs_oracle(gpu, ele):
    if ele == 'HEX8':
        mesh, order = distorted_box(), None
    else:
        m = box_mesh_hex27(2, 3, 2, 1., 1.5, 1.)
        pts = m.points + 0.03 * _rng(1).uniform(-1, 1, m.points.shape)
        mesh, order = Mesh(pts, m.cells_dict['hexahedron27']), 4
    p = ElasticLE(mesh, vec=3, dim=3, ele_type=ele, quadrature_order=order)
    g, jxw = p.fes[0].get_shape_grads()
    sv, sg, w = OB.shape_vals_and_grads(ele, order)
    G, J = OF.shape_grads(mesh.points, mesh.cells, sg, w)
    assert rel(g.cpu().numpy(), G) < 1e-12
    assert rel(jxw.cpu().numpy(), J) < 1e-12
    x = p.fes[0].get_physical_quad_points().cpu().numpy()
    assert rel(x, OF.quad_points(mesh.points, mesh.cells, sv)) < 1e-13


def test_face_geometry_cylinder_area(gpu):
    d = golden('linear_elasticity_cylinder')
    mesh = Mesh(d['points'], d['cells'])
    top = lambda p: isclose(p[2], 10.)
    p = LinearElasticityCylinder(mesh, vec=3, dim=3, dirichlet_bc_info=None, location_fns=[top])
    bi = p.boundary_inds_list[0]
    g, ns = p.fes[0].get_face_shape_grads(bi)
    area = float(ns.sum())
    np.testing.assert_almost_equal(area, float(d['surface_area']), decimal=4)  # test_linear_elasticity_cylinder.py:263
    fv, fg, fw, fn, fi = OB.face_shape_vals_and_grads('HEX8')
    G, NS = OF.face_shape_grads(mesh.points, mesh.cells, bi, fg, fw, fn)
    assert rel(ns.cpu().numpy(), NS) < 1e-12 and rel(g.cpu().numpy(), G) < 1e-12


# ------------------------------------------------------------------- pattern
@pytest.mark.parametrize('case', ['box_v1', 'box_v3', 'cylinder_v3', 'hex27_v3'])
def test_pattern_bit_exact(gpu, case):
    if case == 'hex27_v3':
        m = box_mesh_hex27(2, 2, 3, 1., 1., 1.)
        mesh, ele, vec, order = Mesh(m.points, m.cells_dict['hexahedron27']), 'HEX27', 3, 4
    elif case == 'cylinder_v3':
        d = golden('linear_elasticity_cylinder')
        mesh, ele, vec, order = Mesh(d['points'], d['cells']), 'HEX8', 3, None
    else:
        mesh, ele, vec, order = box(3, 4, 5), 'HEX8', (1 if case == 'box_v1' else 3), None
    cls = LinearPoisson if vec == 1 else ElasticLE
    p = cls(mesh, vec=vec, dim=3, ele_type=ele, quadrature_order=order)
    dev = p.device()
    dev.prepare()
    u = torch.zeros(dev.n, dtype=torch.float64, device='cuda')
    dev.tangent(u, _lib.FEM_F64)
    indptr, indices, data = dev.export_csr()
    n = dev.n
    A = sp.coo_matrix((np.ones(len(p.I)), (p.I, p.J)), shape=(n, n)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    assert np.array_equal(indptr, A.indptr) and np.array_equal(indices, A.indices)


def test_sell_order_is_a_permutation_without_padding(gpu):
    # SELL-C-sigma: sell_row / sell_pos are inverse permutations of the block rows and the
    # length-sorted slices of a HEX27 lattice (rows of 27, 45, 75, 125 blocks) carry little padding
    m = box_mesh_hex27(6, 5, 7, 1., 1., 1.)
    p = ElasticLE(Mesh(m.points, m.cells_dict['hexahedron27']), vec=3, dim=3, ele_type='HEX27', quadrature_order=4)
    dev = p.device()
    assert dev.sell_sorted
    t = dev.pat_tensors
    rows = t['sell_row'].cpu().numpy()
    pos = t['sell_pos'].cpu().numpy()
    Nn = dev.dm.Nn
    real = rows >= 0
    assert real.sum() == Nn and np.array_equal(np.sort(rows[real]), np.arange(Nn))
    assert np.array_equal(rows[pos], np.arange(Nn))
    # each row stays inside its window of 2048 rows; the order is a stable sort by decreasing length
    assert np.all(pos // 2048 == np.arange(Nn) // 2048)
    lens = np.diff(t['row_ptr'].cpu().numpy())
    order = np.concatenate([w + np.argsort(-lens[w:w + 2048], kind='stable') for w in range(0, Nn, 2048)])
    assert np.array_equal(rows[real], order)
    L = np.zeros(len(rows), np.int64)
    L[:Nn] = lens[order]
    expected = int((L.reshape(-1, 64).max(1) * 64).sum())
    unsorted = np.zeros(len(rows), np.int64)
    unsorted[:Nn] = lens
    assert dev.sell_blocks == expected < int((unsorted.reshape(-1, 64).max(1) * 64).sum())


@pytest.mark.parametrize('n', [(6, 5, 7), (32, 16, 16)])
def test_sell_order_rule(gpu, n):
    # natural row order unless its slices pad more than 5% (small boxes, where boundary
    # rows of other lengths are frequent, get sorted; the 1.6M-dof benchmark box does not)
    dev = ElasticLE(box(*n), vec=3, dim=3).device()
    lens = np.diff(dev.pat_tensors['row_ptr'].cpu().numpy())
    Nn = dev.dm.Nn
    L = np.zeros(((Nn + 63) // 64) * 64, np.int64)
    L[:Nn] = lens
    natural = int((L.reshape(-1, 64).max(1) * 64).sum())
    assert dev.sell_sorted == (natural > 1.05 * dev.nnzb)
    rows = dev.pat_tensors['sell_row'].cpu().numpy()
    if not dev.sell_sorted:
        assert dev.sell_blocks == natural
        assert np.array_equal(rows[:Nn], np.arange(Nn)) and np.all(rows[Nn:] == -1)


def test_pattern_nnz_closed_form(gpu):
    # HEX8: prod(3 n_d + 1) V^2 ; HEX27: prod(8 n_d + 1) V^2
    p = ElasticLE(box(4, 3, 5), vec=3, dim=3)
    assert p.device().nnzb * 9 == 13 * 10 * 16 * 9
    m = box_mesh_hex27(2, 3, 2, 1., 1., 1.)
    p = ElasticLE(Mesh(m.points, m.cells_dict['hexahedron27']), vec=3, dim=3, ele_type='HEX27', quadrature_order=4)
    assert p.device().nnzb * 9 == 17 * 25 * 17 * 9


# ------------------------------------------------------------- cell tangents
MATS = {
    'poisson': (1, lambda: M.Poisson(2.5), lambda: OF.Material('poisson', k=2.5)),
    'elastic': (3, lambda: M.LinearElastic(E=70e3, nu=0.3), lambda: OF.Material('elastic', **dict(zip(('lmbda', 'mu'), OF.lame(70e3, 0.3))))),
    'simp': (3, lambda: M.SIMPElastic(E0=1., Emin=1e-3, nu=0.3, penal=3.),
             lambda: OF.Material('simp', **dict(zip(('lmbda1', 'mu1'), OF.lame(1., 0.3))), Emin=1e-3, E0=1., p=3.)),
    'neo_hookean': (3, lambda: M.NeoHookean(E=10., nu=0.3),
                    lambda: OF.Material('neo_hookean', mu=10. / 2.6, kappa=10. / (3 * 0.4))),
}


def make_material_problem(name, ele='HEX8', seed=3):
    vec, dmat, omat = MATS[name]
    if ele == 'HEX8':
        mesh, order = distorted_box((3, 3, 4), seed=seed), None
    else:
        m = box_mesh_hex27(2, 2, 2, 1., 1., 1.)
        mesh, order = Mesh(m.points + 0.02 * _rng(seed).uniform(-1, 1, m.points.shape), m.cells_dict['hexahedron27']), 4

    class P(Problem):
        def get_tensor_map(self):
            return dmat()

        def set_params(self, theta):
            self.internal_vars = [theta]

    p = P(mesh, vec=vec, dim=3, ele_type=ele, quadrature_order=order,
          dirichlet_bc_info=[[lambda x: isclose(x[0], 0.)] * vec, list(range(vec)), [lambda x: 0.] * vec])
    theta = None
    if name == 'simp':
        theta = _rng(seed).uniform(0.3, 1.0, (p.num_cells, p.fes[0].num_quads))
        p.set_params(torch.as_tensor(theta))
    amp = 0.05 if ele == 'HEX8' else 0.01  # keeps det F > 0 on the finer HEX27 node spacing
    u = amp * _rng(seed + 1).standard_normal(p.fes[0].num_total_dofs)
    return p, oracle_problem(p, omat(), theta), u


@pytest.mark.parametrize('name', list(MATS))
@pytest.mark.parametrize('ele', ['HEX8', 'HEX27'])
def test_cell_tangent_and_residual_vs_oracle(gpu, name, ele):
    p, o, u = make_material_problem(name, ele)
    dev = p.device()
    dev.prepare()
    ut = torch.as_tensor(u, device='cuda')
    K64 = dev.cell_tangents(ut, _lib.FEM_F64).cpu().numpy()
    K32 = dev.cell_tangents(ut, _lib.FEM_F32).cpu().numpy()
    U = u.reshape(-1, p.fes[0].vec)
    Ko = OF.cell_tangent(o.mat, U, o.cells, o.G, o.JxW, o.theta)
    assert rel(K64, Ko) < 1e-12
    assert rel(K32, Ko) < 2e-6
    r = p.compute_residual([ut.reshape(-1, p.fes[0].vec)])[0].cpu().numpy()
    ro = OF.scatter(OF.cell_residual(o.mat, U, o.cells, o.G, o.JxW, o.theta), o.cells, len(o.points))
    assert rel(r, ro) < 1e-12
    # matrix values (with Dirichlet rows) against scipy COO->CSR + zeroRows of the oracle
    dev.tangent(ut, _lib.FEM_F64)
    indptr, indices, data = dev.export_csr()
    o.bc_dofs = dev.bc_dofs.cpu().numpy().astype(np.int64)
    A = o.tangent(u)
    assert np.array_equal(indptr, A.indptr) and np.array_equal(indices, A.indices)
    assert rel(data, A.data) < 1e-12
    # matrix-free FP64 operator == assembled FP64 matrix
    x = torch.as_tensor(_rng(9).standard_normal(dev.n), device='cuda')
    y = dev.apply_operator(ut, x).cpu().numpy()
    assert rel(y, A @ x.cpu().numpy()) < 1e-12


@pytest.mark.parametrize('name', ['elastic', 'simp'])
def test_hex27_bf16x3_mfma_tangent(gpu, name):
    """Split-BF16 MFMA tangents (hi*hi + hi*lo + lo*hi, FP32 accumulation): about 16
    significant bits per product, checked against the FP64 oracle; symmetric and with
    rigid translations in the null space to the same accuracy."""
    p, o, u = make_material_problem(name, 'HEX27')
    dev = p.device()
    dev.prepare()
    ut = torch.as_tensor(u, device='cuda')
    Kb = dev.cell_tangents(ut, _lib.FEM_BF16X3).double().cpu().numpy()
    Ko = OF.cell_tangent(o.mat, u.reshape(-1, 3), o.cells, o.G, o.JxW, o.theta)
    err = rel(Kb, Ko)
    assert 1e-9 < err < 1e-4, err          # not FP64-exact, but far better than one BF16 product (4e-3)
    assert rel(Kb, np.swapaxes(Kb, 1, 2)) < 1e-4
    t = np.tile([0.3, -1.1, 0.7], 27)
    assert np.abs(Kb @ t).max() < 1e-4 * np.abs(Kb).max()


def test_hex27_bf16x3_mixed_solve(gpu):
    """Mixed PCG on a BF16x3-assembled HEX27 matrix, refined against the FP64 residual."""
    from problems import LoadedElastic
    m = box_mesh_hex27(4, 3, 3, 2., 1., 1.)
    mesh = Mesh(m.points, m.cells_dict['hexahedron27'])
    left = lambda x: isclose(x[0], 0.)
    p = LoadedElastic(mesh, vec=3, dim=3, ele_type='HEX27', quadrature_order=4,
                      dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda x: 0.] * 3],
                      location_fns=[lambda x: isclose(x[0], 2.)])
    ref = solver(p, {'jax_solver': {'precision': 'fp64', 'tol': 1e-12, 'atol': 0.}})[0].cpu().numpy()
    sol = solver(p, {'jax_solver': {'precision': 'mixed', 'assembly': 'bf16x3'}})[0].cpu().numpy()
    assert np.linalg.norm(sol - ref) / np.linalg.norm(ref) < 1e-8


def test_difference_form_spmv(gpu):
    """FP32 matrices are stored in difference form: y_i = sum_j A_ij (x_j - x_i) + R_i x_i.
    Same product as the FP64 matrix up to FP32 rounding, and exact on translations."""
    mesh = distorted_box((4, 4, 5), seed=5)
    p = ElasticLE(mesh, vec=3, dim=3)
    dev = p.device()
    dev.prepare()
    u = torch.zeros(dev.n, dtype=torch.float64, device='cuda')
    x = torch.as_tensor(_rng(2).standard_normal(dev.n), device='cuda')
    dev.tangent(u, _lib.FEM_F64)
    y64 = dev.spmv(x).cpu().numpy()
    dev.tangent(u, _lib.FEM_F32, _lib.FEM_F64)
    assert dev.form == _lib.FEM_FORM_DIFFERENCE
    y32 = dev.spmv(x.float()).double().cpu().numpy()
    assert rel(y32, y64) < 1e-5
    t = torch.tensor([0.3, -1.2, 0.7], device='cuda').repeat(dev.dm.Nn)
    yt = dev.spmv(t).double().cpu().numpy()
    scale = np.abs(y64).max()
    assert np.abs(yt).max() <= 1e-9 * scale  # the standard FP32 form leaves ~1e-7 * scale here
    # the exported FP32 difference-form matrix reproduces the FP64 one to FP32 rounding
    data32 = dev.export_csr()[2]
    dev.tangent(u, _lib.FEM_F64)
    data64 = dev.export_csr()[2]
    assert rel(data32, data64) < 1e-6


def test_assembly_deterministic(gpu):
    p, o, u = make_material_problem('neo_hookean')
    dev = p.device()
    dev.prepare()
    ut = torch.as_tensor(u, device='cuda')
    outs = []
    for _ in range(2):
        dev.tangent(ut, _lib.FEM_F32)
        outs.append(dev.export_csr()[2].copy())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize('shift', [1, 2, 3])
def test_gather_with_unaligned_index_lists(gpu, shift):
    """fem_gather_matrix reads the slot bounds and the cell-block ids through
    16-byte windows aligned to each list's address.  Lists that start 4, 8 or 12
    bytes past a 16-byte boundary (a C caller's arrays need not be aligned) give
    the same matrix, bit for bit, in both storage forms."""
    import ctypes as C
    p, o, u = make_material_problem('neo_hookean')
    dev = p.device()
    dev.prepare()
    dev.tangent(torch.as_tensor(u, device='cuda'), _lib.FEM_F64, fused=False)   # the two-pass path keeps Ke
    Ke = dev.current[1]
    L = _lib.lib()
    pat = _lib.FemPattern.from_buffer_copy(dev.pat)
    keep = []
    for name in ('contrib_ptr', 'contrib_idx'):
        a = dev.pat_tensors[name]
        buf = torch.full((a.numel() + 8,), -7, dtype=torch.int32, device='cuda')
        view = buf[shift:shift + a.numel()]
        view.copy_(a)
        assert view.data_ptr() % 16 == 4 * shift
        setattr(pat, name, view.data_ptr())
        keep.append(buf)
    for precision, form, dt in ((_lib.FEM_F64, _lib.FEM_FORM_STANDARD, torch.float64),
                                (_lib.FEM_F32, _lib.FEM_FORM_DIFFERENCE, torch.float32)):
        outs = []
        for pref in (dev.pref, C.byref(pat)):
            vals = torch.full((dev.sell_blocks * p.fes[0].vec ** 2,), 3., dtype=dt, device='cuda')
            dinv = torch.zeros(dev.n, dtype=dt, device='cuda')
            _lib.check(L.fem_gather_matrix(pref, _lib.FEM_F64, _lib.ptr(Ke), precision, form, _lib.ptr(vals),
                                           _lib.ptr(dinv), _lib.stream_ptr()), 'fem_gather_matrix')
            torch.cuda.synchronize()
            outs.append((vals, dinv))
        assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1]), (precision, form)
        assert torch.isfinite(outs[0][0]).all() and bool((outs[0][1] != 0).any())


# --------------------------------------------------------- end-to-end fixtures
def _fixture_poisson(precision, method='jax_solver'):
    d = golden('linear_poisson')
    mesh = Mesh(d['points'], d['cells'])
    bc = [[lambda p: isclose(p[0], 0.), lambda p: isclose(p[0], 1.)], [0, 0], [lambda p: 0., lambda p: 1.]]
    p = LinearPoisson(mesh, vec=1, dim=3, dirichlet_bc_info=bc)
    sol = solver(p, {method: {'precision': precision}})[0]
    return sol, d['sol'], 5


def _fixture_cube(precision, method='jax_solver'):
    d = golden('linear_elasticity_cube')
    mesh = Mesh(d['points'], d['cells'])
    left = lambda p: isclose(p[0], 0.)
    bc = [[left] * 3, [0, 1, 2], [lambda p: 1.] * 3]
    p = LinearElasticityCube(mesh, vec=3, dim=3, dirichlet_bc_info=bc, location_fns=[lambda p: isclose(p[0], 1.)])
    sol = solver(p, {method: {'precision': precision}})[0]
    return sol, d['sol'], 5


def _fixture_cylinder(precision, method='jax_solver'):
    d = golden('linear_elasticity_cylinder')
    mesh = Mesh(d['points'], d['cells'])
    bottom = lambda p: isclose(p[2], 0.)
    bc = [[bottom] * 3, [0, 1, 2], [lambda p: 0.] * 3]
    p = LinearElasticityCylinder(mesh, vec=3, dim=3, dirichlet_bc_info=bc, location_fns=[lambda p: isclose(p[2], 10.)])
    sol = solver(p, {method: {'precision': precision}})[0]
    return sol, d['sol'], 3


def _fixture_hyper(precision, method='jax_solver'):
    d = golden('hyperelasticity')
    mesh = Mesh(d['points'], d['cells'])
    top, bottom = (lambda p: isclose(p[2], 10.)), (lambda p: isclose(p[2], 0.))
    bc = [[bottom] * 3 + [top] * 3, [0, 1, 2, 0, 1, 2], [lambda p: 0.] * 5 + [lambda p: 1.]]
    p = HyperElasticity(mesh, vec=3, dim=3, dirichlet_bc_info=bc)
    sol = solver(p, {method: {'precision': precision}})[0]
    # traction on the top face (test_hyper_elasticity.py:36-70, 121-127)
    bi = p.fes[0].get_boundary_conditions_inds([top])[0]
    g, ns = p.fes[0].get_face_shape_grads(bi)
    cells = torch.as_tensor(p.fes[0].cells, device='cuda')
    ug = torch.einsum('sav,sqaj->sqvj', sol[cells[torch.as_tensor(bi[:, 0], device='cuda')]], g)
    P = torch.func.vmap(p.get_tensor_map())(ug.reshape(-1, 3, 3)).reshape(ug.shape)
    trac = (P @ torch.tensor([0., 0., 1.], dtype=torch.float64, device='cuda'))
    tz = float((trac * ns[..., None]).sum((0, 1))[2])
    np.testing.assert_almost_equal(tz, float(d['traction']), decimal=5)
    return sol, d['sol'], 5


@pytest.mark.parametrize('method', ['jax_solver', 'amgx_solver'])
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_fenicsx_plasticity_fixture(gpu, precision, method):
    """J2 plasticity, 5 load steps with the state carried by internal variables
    (test_plasticity.py:94-161): solution at decimal 5, average sigma_zz at decimal 3."""
    d = golden('plasticity')
    mesh = Mesh(d['points'], d['cells'])
    top, bottom = (lambda p: isclose(p[2], 10.)), (lambda p: isclose(p[2], 0.))

    def bc(disp):
        return [[bottom] * 3 + [top] * 3, [0, 1, 2, 0, 1, 2], [lambda p: 0.] * 5 + [lambda p: disp]]

    problem = Plasticity(mesh, vec=3, dim=3, dirichlet_bc_info=bc(float(d['disps'][0])))
    avg = []
    for disp in d['disps']:
        problem.fe.update_Dirichlet_boundary_conditions(bc(float(disp)))
        sol = solver(problem, {method: {'precision': precision}})[0]
        problem.update_stress_strain(sol)
        avg.append(float(problem.compute_avg_stress()[2, 2]))
    assert isinstance(problem.device().material, M.J2Plastic)
    got = sol.cpu().numpy().astype(np.float32)
    np.testing.assert_array_almost_equal(d['sol'], got, decimal=5)
    np.testing.assert_array_almost_equal(d['avg_stresses'], np.array(avg), decimal=3)


def test_j2_tangent_and_residual_vs_oracle(gpu):
    """J2 tangents and residuals on a distorted mesh with a random stress/strain history that
    puts most quadrature points on the plastic branch."""
    mesh = distorted_box((3, 3, 4), seed=4)
    p = Plasticity(mesh, vec=3, dim=3, dirichlet_bc_info=[[lambda x: isclose(x[0], 0.)] * 3, [0, 1, 2],
                                                         [lambda x: 0.] * 3])
    C, Q = p.num_cells, p.fes[0].num_quads
    rng = _rng(5)
    sym = lambda a: 0.5 * (a + np.swapaxes(a, -1, -2))
    sig = sym(150. * rng.standard_normal((C, Q, 3, 3)))
    eps = sym(2e-3 * rng.standard_normal((C, Q, 3, 3)))
    p.internal_vars = [torch.as_tensor(sig), torch.as_tensor(eps)]
    u = 2e-3 * rng.standard_normal(p.fes[0].num_total_dofs)
    dev = p.device()
    dev.prepare()
    ut = torch.as_tensor(u, device='cuda')
    lm, mu = M.lame(70e3, 0.3)
    om = OF.Material('j2', lmbda=lm, mu=mu, sig0=250.)
    state = np.concatenate([sig.reshape(C, Q, 9), eps.reshape(C, Q, 9)], 2)
    o = oracle_problem(p, om, state)
    U = u.reshape(-1, 3)
    Gu = np.einsum('cav,cqaj->cqvj', U[o.cells], o.G)
    assert (om._j2(Gu, state)[2] > 0).mean() > 0.5
    K64 = dev.cell_tangents(ut, _lib.FEM_F64).cpu().numpy()
    Ko = OF.cell_tangent(om, U, o.cells, o.G, o.JxW, state)
    assert rel(K64, Ko) < 1e-12
    r = p.compute_residual([ut.reshape(-1, 3)])[0].cpu().numpy()
    ro = OF.scatter(OF.cell_residual(om, U, o.cells, o.G, o.JxW, state), o.cells, len(o.points))
    assert rel(r, ro) < 1e-12
    x = torch.as_tensor(_rng(9).standard_normal(dev.n), device='cuda')
    dev.tangent(ut, _lib.FEM_F64)
    o.bc_dofs = dev.bc_dofs.cpu().numpy().astype(np.int64)
    A = o.tangent(u)
    assert rel(dev.apply_operator(ut, x).cpu().numpy(), A @ x.cpu().numpy()) < 1e-12


@pytest.mark.parametrize('method', ['jax_solver', 'amgx_solver'])
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
@pytest.mark.parametrize('case', ['poisson', 'cube', 'cylinder', 'hyper'])
def test_fenicsx_fixtures(gpu, case, precision, method):
    fn = {'poisson': _fixture_poisson, 'cube': _fixture_cube, 'cylinder': _fixture_cylinder,
          'hyper': _fixture_hyper}[case]
    sol, ref, decimal = fn(precision, method)
    got = sol.cpu().numpy().reshape(ref.shape).astype(np.float32)  # save_sol writes float32 (utils.py:45)
    np.testing.assert_array_almost_equal(ref, got, decimal=decimal)


# -------------------------------------------------------------------- solvers
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_solution_vs_oracle_direct(gpu, precision):
    mesh = box(12, 6, 6, 2., 1., 1.)
    left = lambda p: isclose(p[0], 0.)
    p = LoadedElastic(mesh, vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.] * 3],
                      location_fns=[lambda p: isclose(p[0], 2.)])
    sol = solver(p, {'jax_solver': {'precision': precision, 'tol': 1e-12, 'atol': 1e-14}})[0].cpu().numpy()
    o = oracle_problem(p, OF.Material('elastic', **dict(zip(('lmbda', 'mu'), OF.lame(70e3, 0.3)))))
    o.bc_dofs, o.bc_vals = p.device().bc_dofs.cpu().numpy().astype(np.int64), p.device().bc_vals.cpu().numpy()
    bi = p.boundary_inds_list[0]
    o.add_surface(bi, np.broadcast_to(np.array([0., 0., 1.]), (len(bi), o.fw.shape[1], 3)))
    ref = o.newton(linear='direct').reshape(-1, 3)
    err = np.linalg.norm(sol - ref) / np.linalg.norm(ref)
    assert err < 1e-8, err


def test_pcg_bit_reproducible(gpu):
    outs = [_fixture_cube('mixed')[0].cpu().numpy() for _ in range(2)]
    assert np.array_equal(outs[0], outs[1])


def test_pcg_not_converged_raises(gpu):
    d = golden('linear_elasticity_cube')
    mesh = Mesh(d['points'], d['cells'])
    left = lambda p: isclose(p[0], 0.)
    p = LinearElasticityCube(mesh, vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 1.] * 3],
                             location_fns=[lambda p: isclose(p[0], 1.)])
    with pytest.raises(AssertionError, match="failed to converge"):
        solver(p, {'jax_solver': {'maxiter': 1}})


# -------------------------------------------------------------------- adjoint
def simp_setup(n=(8, 4, 4), seed=0):
    mesh = box(*n, 2., 1., 1.)
    left = lambda p: isclose(p[0], 0.)
    load = lambda p: isclose(p[0], 2.) & (p[2] <= 0.25 + 1e-5)
    p = SIMPElasticity(mesh, vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.] * 3],
                       location_fns=[load])
    theta = torch.as_tensor(0.5 + _rng(seed).uniform(-0.05, 0.05, p.num_cells), device='cuda')
    return p, theta


def compliance(p, sol):
    # J = F . u with the (constant) load vector: load is -R_surface, so F.u = -load.u
    dev = p.device()
    return -torch.dot(dev.load, sol.reshape(-1))


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_adjoint_simp_vs_oracle_and_fd(gpu, precision):
    p, theta = simp_setup()
    opts = {'jax_solver': {'precision': precision, 'tol': 1e-12, 'atol': 1e-14}}
    fwd = ad_wrapper(p, opts, opts)
    th = theta.clone().requires_grad_(True)
    J = compliance(p, fwd(th)[0])
    J.backward()
    g = th.grad.cpu().numpy()
    # oracle: same problem, adjoint by sparse direct solves
    o = oracle_setup.oracle_problem(p, OF.Material('simp', **dict(zip(('lmbda1', 'mu1'), OF.lame(1., 0.3))),
                                                   Emin=1e-3, E0=1., p=3.),
                                    theta=np.repeat(theta.cpu().numpy()[:, None], 8, axis=1))
    u = o.newton(linear='direct')
    gq, lam = o.adjoint_theta(u, -o.load.reshape(-1))
    go = gq.sum(1)
    assert rel(g, go) < 1e-7
    # compliance is self-adjoint: lambda = u on free dofs
    free = np.setdiff1d(np.arange(o.n), o.bc_dofs)
    assert rel(lam[free], u[free]) < 1e-9
    # closed form dJ/dtheta_e = -3 theta^2 (E0 - Emin) u_e^T K1_e u_e (< 0)
    assert np.all(g < 0)
    # central differences on seeded cells
    h = 1e-6
    cells = _rng(1).choice(p.num_cells, 4, replace=False)
    for c in cells:
        tp, tm = theta.clone(), theta.clone()
        tp[c] += h
        tm[c] -= h
        Jp = float(compliance(p, fwd(tp)[0]))
        Jm = float(compliance(p, fwd(tm)[0]))
        fd = (Jp - Jm) / (2 * h)
        assert abs(fd - g[c]) <= 1e-4 * abs(g).max(), (c, fd, g[c])
