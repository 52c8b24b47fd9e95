"""CPU tests of the product's host logic and of the C-ABI library surface."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from conftest import ROOT
from oracle import basis as OB
from oracle import fem as OF
from problems import LinearElasticityCube, LinearPoisson, HyperElasticity, Plasticity, SIMPElasticity, box, isclose

from jax_fem_amd import _lib, basis, materials as M
from jax_fem_amd.generate_mesh import box_mesh, box_mesh_hex27, Mesh
from jax_fem_amd.solver import _resolve_solver_options, linear_config


# ------------------------------------------------------------------ library
def test_library_exports_every_header_symbol():
    lib = _lib.load()
    header = open(os.path.join(ROOT, 'include', 'fem_mi355x.h')).read()
    names = set(re.findall(r'^\s*(?:const char\*|int64_t|int)\s+(fem_\w+)\s*\(', header, re.M))
    assert names, "no declarations parsed"
    assert names == set(_lib.EXPORTED), names ^ set(_lib.EXPORTED)
    for n in names:
        assert getattr(lib, n) is not None
    assert lib.fem_version() == 1


def test_struct_sizes_match_header_layout():
    # plain C layout: ints first, then 64-bit fields / pointers
    assert ctypes.sizeof(_lib.FemMesh) == 4 * 4 + 2 * 8 + 9 * 8
    assert ctypes.sizeof(_lib.FemMaterial) == 8 + 6 * 8 + 8 + 8
    assert ctypes.sizeof(_lib.FemPattern) == 4 * 8 + 4 * 4 + 9 * 8
    assert ctypes.sizeof(_lib.FemPcgOptions) == 8 + 8 + 4 * 4 + 8 + 8
    assert ctypes.sizeof(_lib.FemPcgResult) == 16 + 16 + 16 + 8
    assert ctypes.sizeof(_lib.FemBcsr) == 3 * 8 + 4 * 4 + 3 * 8
    assert ctypes.sizeof(_lib.FemAmgLevel) == 8 + 8 + 3 * ctypes.sizeof(_lib.FemBcsr) + 8 + 8
    assert ctypes.sizeof(_lib.FemAmg) == 4 * 4 + 8 + 8 + 8
    assert ctypes.sizeof(_lib.FemAmgOptions) == 8 + 8 + 4 * 4 + 8
    assert ctypes.sizeof(_lib.FemOperator) == 5 * 8 + 8 + 8 + 8 + 8 + 8


def test_amgx_solver_options(tmp_path):
    cfg = linear_config({'amgx_solver': {}})
    assert cfg['method'] == 'amg' and cfg['precision_code'] == _lib.FEM_F64
    assert cfg['amg_precision'] == _lib.FEM_F64
    assert linear_config({'amgx_solver': {'precision': 'mixed'}})['amg_precision'] == _lib.FEM_F32
    with pytest.raises(ValueError):
        linear_config({'amgx_solver': {'smoother': 'JACOBI'}})
    with pytest.raises(ValueError):
        linear_config({'amgx_solver': {'precision': 'fp16'}})
    with pytest.raises(ValueError):
        linear_config({'amgx_solver': {'outer': 'fp32'}})     # an FP32 outer operator needs precision 'mixed'
    assert linear_config({'amgx_solver': {'precision': 'mixed', 'level0': 'fp16'}})['level0'] == 'fp16'
    with pytest.raises(ValueError):
        linear_config({'amgx_solver': {'level0': 'fp16'}})    # FP16 level 0 needs the FP32 hierarchy
    with pytest.raises(ValueError):
        linear_config({'amgx_solver': {'precision': 'mixed', 'level0': 'fp16', 'outer': 'fp32'}})
    # an AMGX config file (solver.py:158-187 layout): tolerance and max_iters are taken
    path = tmp_path / 'amgx.json'
    path.write_text('{"config_version": 2, "solver": {"solver": "PCG", "tolerance": 1e-9, "max_iters": 77}}')
    cfg = linear_config({'amgx_solver': {'cfg_path': str(path)}})
    assert cfg['atol'] == 1e-9 and cfg['tol'] == 0.0 and cfg['maxiter'] == 77
    assert linear_config({})['method'] == 'pcg'


def test_compute_entry_fails_loudly_without_gpu():
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(_lib.FemError, match="no GPU"):
        _lib.lib()


def test_noalloc_pcg_refusals_before_any_launch():
    # the checks of fem_pcg_solve_noalloc run on the host, before any HIP call
    lib = _lib.load()
    pat = _lib.FemPattern()
    pat.vec, pat.num_rows = 3, 1 << 20
    pat.num_slices = pat.num_rows // 64
    opt = _lib.FemPcgOptions()
    opt.precision, opt.timing = _lib.FEM_F64, 0
    res = _lib.FemPcgResult()
    call = lambda: lib.fem_pcg_solve_noalloc(ctypes.byref(pat), None, None, None, None, None, ctypes.byref(opt),
                                             None, ctypes.byref(res), None)
    need = lib.fem_pcg_ws_bytes(ctypes.byref(pat), _lib.FEM_F64)
    assert need > _lib.FEM_PCG_NOALLOC_SCRATCH_BYTES      # 3.1 M FP64 dofs: above the scratch
    assert call() == _lib.FEM_ERR_CAPACITY and b'scratch' in lib.fem_last_error()
    opt.timing = 1
    assert call() == _lib.FEM_ERR_BAD_ARG and b'timing' in lib.fem_last_error()
    # FP32 workspaces are smaller: the same system fits
    assert lib.fem_pcg_ws_bytes(ctypes.byref(pat), _lib.FEM_F32) <= _lib.FEM_PCG_NOALLOC_SCRATCH_BYTES


# -------------------------------------------------------------------- basis
@pytest.mark.parametrize('ele,order', [('HEX8', None), ('HEX8', 4), ('HEX27', 4), ('HEX27', None)])
def test_product_basis_matches_oracle(ele, order):
    a = basis.get_shape_vals_and_grads(ele, quadrature_order=order)
    b = OB.shape_vals_and_grads(ele, order)
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=0, atol=1e-12)
    fa = basis.get_face_shape_vals_and_grads(ele, quadrature_order=order)
    fb = OB.face_shape_vals_and_grads(ele, order)
    for x, y in zip(fa, fb):
        np.testing.assert_allclose(x, y, rtol=0, atol=1e-12)


def test_hex27_face_inds_and_defaults():
    v, g, w = basis.get_shape_vals_and_grads('HEX27')
    assert v.shape == (216, 27)  # basis.py:64: order 10 -> 6x6x6
    v, g, w = basis.get_shape_vals_and_grads('HEX27', quadrature_order=4)
    assert v.shape == (27, 27)
    with pytest.raises(NotImplementedError):
        basis.get_shape_vals_and_grads('TET4')


def test_box_mesh_numbering():
    m = box_mesh(2, 3, 4, 1., 2., 3.)
    cells = m.cells_dict['hexahedron']
    assert m.points.shape == (3 * 4 * 5, 3) and cells.shape == (24, 8)
    # node id = (i*(Ny+1) + j)*(Nz+1) + k (generate_mesh.py:176-187)
    nid = lambda i, j, k: (i * 4 + j) * 5 + k
    assert list(cells[0]) == [nid(0, 0, 0), nid(1, 0, 0), nid(1, 1, 0), nid(0, 1, 0),
                              nid(0, 0, 1), nid(1, 0, 1), nid(1, 1, 1), nid(0, 1, 1)]
    p, c = OF.box_mesh(2, 3, 4, 1., 2., 3.)
    assert np.array_equal(p, m.points) and np.array_equal(c, cells)


def test_hex27_box_is_lagrange_consistent():
    m = box_mesh_hex27(2, 2, 1, 2., 2., 1.)
    pts, cells = m.points, m.cells_dict['hexahedron27']
    # node coordinates of each cell = its corner box mapped through the reference node positions
    for c in cells:
        X = pts[c]
        lo, hi = X[:8].min(0), X[:8].max(0)
        np.testing.assert_allclose(X, lo + OB.HEX27_NODES * (hi - lo), atol=1e-14)


# -------------------------------------------------------- BC and boundaries
def test_dirichlet_and_boundary_discovery_match_oracle():
    mesh = box(3, 4, 2)
    left, right = (lambda p: isclose(p[0], 0.)), (lambda p: isclose(p[0], 1.))
    bc = [[left, left, right], [0, 2, 1], [lambda p: 1., lambda p: p[1] + 2., lambda p: 0.5]]
    p = LinearElasticityCube(mesh, vec=3, dim=3, dirichlet_bc_info=bc, location_fns=[right])
    fe = p.fes[0]
    d, v = OF.dirichlet(mesh.points, 3, bc)
    got_d = np.concatenate([n * 3 + c for n, c in zip(fe.node_inds_list, fe.vec_inds_list)])
    got_v = np.concatenate(fe.vals_list)
    order = np.argsort(got_d)
    assert np.array_equal(got_d[order], d)
    np.testing.assert_allclose(got_v[order], v)
    bi = OF.boundary_inds(mesh.points, mesh.cells, OB.face_shape_vals_and_grads('HEX8')[4], right)
    assert np.array_equal(p.boundary_inds_list[0], bi)


def test_location_fn_with_index_argument():
    mesh = box(2, 2, 2)
    sel = lambda p, ind: np.isin(ind, np.array([0, 5, 7]))
    p = LinearPoisson(mesh, vec=1, dim=3, dirichlet_bc_info=[[sel], [0], [lambda p: 0.]])
    assert list(p.fes[0].node_inds_list[0]) == [0, 5, 7]


def test_surface_count_assertion():
    mesh = box(2, 2, 2)
    with pytest.raises(AssertionError):
        LinearPoisson(mesh, vec=1, dim=3, location_fns=[lambda p: isclose(p[0], 0.)])


def test_coo_exports_match_reference_formula():
    mesh = box(2, 1, 1)
    p = LinearElasticityCube(mesh, vec=3, dim=3, location_fns=[lambda p: isclose(p[0], 1.)])
    I, J = OF.coo_pattern(mesh.cells, 3)
    nf = len(p.boundary_inds_list[0]) * 24 * 24
    assert np.array_equal(p.I[:len(I)], I) and np.array_equal(p.J[:len(J)], J) and len(p.I) == len(I) + nf


# ------------------------------------------------------------------ options
def test_solver_options_resolution():
    assert _resolve_solver_options({}) == ('newton', {})
    assert _resolve_solver_options({'tol': 1e-5, 'jax_solver': {}}) == \
        ('newton', {'tol': 1e-5, 'linear': {'jax_solver': {}}})
    with pytest.raises(ValueError):
        _resolve_solver_options({'newton': {}, 'arc_length': {}})
    with pytest.raises(ValueError):
        _resolve_solver_options({'newton': 3})
    cfg = linear_config({})
    assert cfg['precond'] and cfg['tol'] == 1e-10 and cfg['maxiter'] == 10000 and cfg['precision'] == 'fp64'
    with pytest.raises(NotImplementedError):
        linear_config({'petsc_solver': {}})
    with pytest.raises(ValueError):
        linear_config({'jax_solver': {'bogus': 1}})


# ------------------------------------------------------- material probing
def test_identify_reference_tensor_maps():
    mesh = box(1, 1, 1)
    m = M.identify(LinearPoisson(mesh, vec=1, dim=3).get_tensor_map(), 1)
    assert isinstance(m, M.Poisson) and abs(m.k - 1.) < 1e-12
    from types import SimpleNamespace
    m = M.identify(LinearElasticityCube.get_tensor_map(SimpleNamespace(dim=3)), 3)
    lm, mu = OF.lame(70e3, 0.3)
    assert isinstance(m, M.LinearElastic) and abs(m.lmbda - lm) < 1e-6 * lm and abs(m.mu - mu) < 1e-6 * mu
    m = M.identify(HyperElasticity(mesh, vec=3, dim=3).get_tensor_map(), 3)
    assert isinstance(m, M.NeoHookean)
    assert abs(m.mu - 1e3 / 2.6) < 1e-6 and abs(m.kappa - 1e3 / 1.2) < 1e-6
    p = SimpleNamespace(E0=1., Emin=1e-3, penal=3.)
    m = M.identify(SIMPElasticity.get_tensor_map(p), 3, iv_shapes=((1,),))
    assert isinstance(m, M.SIMPElastic)
    assert abs(m.Emin - 1e-3) < 1e-9 and abs(m.penal - 3.) < 1e-9
    # E field: stress(u_grad, E) linear in E
    Efield = lambda G, E: M.LinearElastic(E=1., nu=0.3)(G) * E
    m = M.identify(Efield, 3, iv_shapes=((),))
    assert isinstance(m, M.SIMPElastic) and m.Emin == 0. and abs(m.penal - 1.) < 1e-9
    # a map outside the closed-form laws becomes the tabulated law
    assert isinstance(M.identify(lambda G: G ** 3, 3), M.Tabulated)


def test_material_descriptors_agree_with_oracle():
    G = torch.tensor(np.random.default_rng(0).uniform(-0.1, 0.1, (5, 3, 3)))
    lm, mu = OF.lame(70e3, 0.3)
    np.testing.assert_allclose(M.LinearElastic(E=70e3, nu=0.3)(G).numpy(),
                               OF.Material('elastic', lmbda=lm, mu=mu).stress(G.numpy()), rtol=1e-13)
    np.testing.assert_allclose(M.NeoHookean(E=10., nu=0.3)(G).numpy(),
                               OF.Material('neo_hookean', mu=10 / 2.6, kappa=10 / 1.2).stress(G.numpy()), rtol=1e-12)


def test_identify_reference_j2_return_map():
    # the reference's plasticity tensor map (sigma_old, eps_old as internal variables)
    p = Plasticity.__new__(Plasticity)
    p.dim = 3
    mat = M.identify(p.get_tensor_map(), 3, ((3, 3), (3, 3)))
    assert isinstance(mat, M.J2Plastic)
    lm, mu = M.lame(70e3, 0.3)
    assert abs(mat.lmbda - lm) < 1e-6 * lm and abs(mat.mu - mu) < 1e-6 * mu and abs(mat.sig0 - 250.) < 1e-9
    assert (mat.sig_index, mat.eps_index) == (0, 1)
    # swapped argument order is found too
    f = p.get_tensor_map()
    mat2 = M.identify(lambda G, e, s: f(G, s, e), 3, ((3, 3), (3, 3)))
    assert (mat2.sig_index, mat2.eps_index) == (1, 0)
    # the descriptor reproduces the map on plastic states (and the oracle's return map)
    rng = np.random.default_rng(3)
    sym = lambda a: 0.5 * (a + a.T)
    o = OF.Material('j2', lmbda=lm, mu=mu, sig0=250.)
    for _ in range(5):
        G = torch.as_tensor(0.02 * rng.standard_normal((3, 3)))
        so = torch.as_tensor(sym(200 * rng.standard_normal((3, 3))))
        eo = torch.as_tensor(sym(1e-3 * rng.standard_normal((3, 3))))
        np.testing.assert_allclose(mat(G, so, eo).numpy(), f(G, so, eo).numpy(), rtol=1e-12, atol=1e-9)
        st = np.concatenate([so.numpy().reshape(9), eo.numpy().reshape(9)])
        np.testing.assert_allclose(o.stress(G.numpy(), st), f(G, so, eo).numpy(), rtol=1e-12, atol=1e-9)


def test_mass_map_u_dependence_probe():
    # a map that depends on u only through component differences, one confined to part of
    # the domain, and a constant: only the last is a constant load
    from jax_fem_amd._device import DeviceProblem

    class _Fake:
        V = 3
    xq = torch.as_tensor(np.random.default_rng(0).uniform(0., 1., (400, 3)))
    diff = lambda u, x: 10. * torch.stack([u[0] - u[1], u[1] - u[2], u[2] - u[0]]) - 1.
    local = lambda u, x: torch.where(x[0] > 0.97, u, torch.zeros_like(u)) + 1.
    const = lambda u, x: torch.stack([x[0], 2. + 0. * x[1], x[2] ** 2])
    probe = lambda fn: DeviceProblem._eval_u_independent(_Fake(), fn, xq, [], 'mass_map', allow_u=True)
    assert probe(diff) is None
    assert int((xq[:, 0] > 0.97).sum()) > 0 and probe(local) is None
    m = probe(const)
    assert m is not None and torch.equal(m, torch.stack([xq[:, 0], 2. + 0. * xq[:, 1], xq[:, 2] ** 2], 1))
    with pytest.raises(NotImplementedError):
        DeviceProblem._eval_u_independent(_Fake(), diff, xq, [], 'surface_map', allow_u=False)


def test_dense_coarse_limit_matches_the_oracle_rule():
    # the dense coarsest level: 512 unknowns, up to 1024 while n^3 <= 8 nnz of level 0; the
    # device setup and the oracle hierarchy must stop coarsening at the same size
    from oracle import amg as OA
    from jax_fem_amd.amg import coarse_limit
    c1 = 33 ** 3 * 27                 # the 32^3 Poisson cube: about 27 entries per row
    c2 = 14_340_865 * 9               # the 1.6M-dof cantilever's level-0 entries
    for nnz0 in (0, 1, c1, 67_000_000, c2, 10 ** 9, 10 ** 12):
        n = coarse_limit(nnz0)
        assert n == OA.coarse_limit(nnz0)
        assert 512 <= n <= 1024
        assert n == 512 or n ** 3 <= 8 * nnz0
        assert n == 1024 or (n + 1) ** 3 > 8 * nnz0
    assert coarse_limit(c1) == 512        # the Poisson cube keeps its 858-unknown level smoothed
    assert coarse_limit(c2) >= 882        # the cantilever's third level (882 unknowns) is dense
