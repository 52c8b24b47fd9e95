"""Device smoothed-aggregation AMG (the `amgx_solver` option) against the CPU
restatement oracle/amg.py on the same assembled matrix, and AMG-CG solves
against a direct solve of the exported matrix."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from oracle import amg as OA
from problems import LinearPoisson, LoadedElastic, box, isclose

from jax_fem_amd import _lib
from jax_fem_amd.amg import DeviceAMG
from jax_fem_amd.solver import linear_config, solver

pytestmark = pytest.mark.gpu


def _elastic(nx):
    clamp = [[lambda x: isclose(x[0], 0.)] * 3, [0, 1, 2], [lambda x: 0.] * 3]
    return LoadedElastic(box(nx, nx // 2, nx // 2, 2., 1., 1.), vec=3, dim=3, dirichlet_bc_info=clamp,
                         location_fns=[lambda x: isclose(x[0], 2.)])


class _Source(LinearPoisson):
    def get_mass_map(self):
        return lambda u, x: torch.stack([-(1. + x[0] * x[1])])


def _poisson(nx):
    bc = [[lambda x: isclose(x[0], 0.)], [0], [lambda x: 0.]]
    return _Source(box(nx, nx, nx // 2), vec=1, dim=3, dirichlet_bc_info=bc)


def _assembled(p):
    dev = p.device()
    dev.prepare()
    u = torch.zeros(dev.n, dtype=torch.float64, device=dev.dm.device)
    dev.tangent(u, _lib.FEM_F64)
    indptr, indices, data = dev.export_csr()
    A = sp.csr_matrix((data, indices, indptr), shape=(dev.n, dev.n))
    gp = dev.pat_tensors['row_ptr'].cpu().numpy()
    gi = dev.pat_tensors['col_idx'].cpu().numpy()
    mask = np.zeros(dev.n, bool)
    mask[dev.bc_dofs.cpu().numpy()] = True
    b = -dev.residual(u)
    return dev, A, (gp, gi), mask, np.asarray(p.fes[0].points), b


def _cfg(precision='fp64', **kw):
    return linear_config({'amgx_solver': dict(precision=precision, **kw)})


def _bcsr(M):
    rp, ci, v = M.host()
    return sp.bsr_matrix((v, ci, rp), shape=(M.rows * M.br, M.cols * M.bc)).tocsr()


def _rel(a, b):
    a = a.toarray() if sp.issparse(a) else a
    b = b.toarray() if sp.issparse(b) else b
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize('make', [lambda: _elastic(12), lambda: _poisson(16)], ids=['elastic', 'poisson'])
def test_aggregation_bit_exact(gpu, make):
    p = make()
    dev, A, (gp, gi), mask, pts, b = _assembled(p)
    amg = DeviceAMG(dev)
    agg, n_agg = amg._aggregate(dev.dm.Nn, dev.pat_tensors['row_ptr'], dev.pat_tensors['col_idx'])
    ragg, rn = OA.mis2_aggregate(gp, gi)
    assert n_agg == rn
    np.testing.assert_array_equal(agg.cpu().numpy(), ragg)
    aptr, memb = amg._members(dev.dm.Nn, agg, n_agg)
    rptr, rmemb = OA.members(ragg, rn)
    np.testing.assert_array_equal(aptr[:rn + 1].cpu().numpy(), rptr)
    np.testing.assert_array_equal(memb[:dev.dm.Nn].cpu().numpy(), rmemb)


@pytest.mark.parametrize('make,vec', [(lambda: _elastic(32), 3), (lambda: _poisson(40), 1)],
                         ids=['elastic', 'poisson'])
def test_hierarchy_matches_oracle(gpu, make, vec):
    p = make()
    dev, A, (gp, gi), mask, pts, b = _assembled(p)
    amg = dev.amg_setup(_cfg())
    lams = [lv['lam'] for lv in amg.levels]
    H = OA.Hierarchy(A, gp, gi, pts, vec, mask, lams=lams)
    assert len(H.levels) == len(amg.levels) >= 1
    for l, (lv, ref) in enumerate(zip(amg.levels, H.levels)):
        np.testing.assert_array_equal(lv['agg'].cpu().numpy(), ref['agg'])
        lam_ref = OA.lmax(ref['A'], ref['dinv'])
        assert abs(lv['lam'] - lam_ref) <= 1e-9 * lam_ref, (l, lv['lam'], lam_ref)
        n = lv['A'].rows * lv['A'].br
        Bc = lv['Bc'][:lv['n_agg'] * ref['Bc'].shape[1] ** 2].cpu().numpy().reshape(ref['Bc'].shape)
        np.testing.assert_allclose(Bc, ref['Bc'], atol=1e-12 * max(1.0, np.abs(ref['Bc']).max()))
        assert _rel(_bcsr(lv['P']), ref['P']) < 1e-12, l
        assert _rel(_bcsr(lv['R']), ref['P'].T) < 1e-12, l
        nxt = amg.levels[l + 1]['A'] if l + 1 < len(amg.levels) else amg.coarse
        ref_next = H.levels[l + 1]['A'] if l + 1 < len(H.levels) else H.coarse
        assert _rel(_bcsr(nxt), ref_next) < 1e-11, l
        assert n == ref['A'].shape[0]
    inv = amg.coarse_inv[:H.coarse.shape[0] ** 2].cpu().numpy().reshape(H.coarse.shape)
    assert _rel(inv, H.coarse_inv) < 1e-8


# the FP32 hierarchy forms its Galerkin products and Jacobi diagonals from FP32 values: their
# rounding (1e-7), amplified through the coarse levels' solves, moves the cycle's output by a
# few 1e-5 against the FP64 oracle
@pytest.mark.parametrize('precision,tol', [('fp64', 1e-11), ('mixed', 1e-4)])
def test_vcycle_matches_oracle(gpu, precision, tol):
    p = _elastic(32)
    dev, A, (gp, gi), mask, pts, b = _assembled(p)
    amg = dev.amg_setup(_cfg(precision))
    H = OA.Hierarchy(A, gp, gi, pts, 3, mask, lams=[lv['lam'] for lv in amg.levels])
    rng = np.random.default_rng(11)
    r = rng.standard_normal(dev.n)
    r[mask] = 0.
    z = amg.vcycle(torch.as_tensor(r, device=dev.dm.device)).cpu().numpy()
    ref = H.vcycle(r)
    assert np.linalg.norm(z - ref) <= tol * np.linalg.norm(ref)
    assert np.abs(z[mask]).max() == 0.0


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
@pytest.mark.parametrize('make', [lambda: _elastic(16), lambda: _poisson(20)], ids=['elastic', 'poisson'])
def test_amg_solve_matches_direct(gpu, make, precision):
    p = make()
    dev, A, _, mask, _, b = _assembled(p)
    ref = spla.spsolve(A.tocsc(), b.cpu().numpy())
    sol = solver(p, {'amgx_solver': {'precision': precision, 'tol': 1e-12, 'atol': 0.}})[0]
    x = sol.reshape(-1).cpu().numpy()
    assert np.linalg.norm(x - ref) <= 1e-9 * np.linalg.norm(ref)
    its = p._solve_stats['linear'][0]['iterations']
    solver(p, {'jax_solver': {'precision': 'fp64', 'tol': 1e-12, 'atol': 0.}})
    its_j = p._solve_stats['linear'][0]['iterations']
    assert (its < its_j / 2) if dev.V == 3 else (its <= 30), (its, its_j)


def test_amg_solve_bit_reproducible(gpu):
    p = _elastic(16)
    opts = {'amgx_solver': {'precision': 'mixed'}}
    a = solver(p, opts)[0].clone()
    b = solver(p, opts)[0]
    assert torch.equal(a, b)


def test_amg_adjoint_matches_jacobi_pcg(gpu):
    from problems import SIMPElasticity
    from jax_fem_amd.solver import ad_wrapper
    clamp = [[lambda x: isclose(x[0], 0.)] * 3, [0, 1, 2], [lambda x: 0.] * 3]
    p = SIMPElasticity(box(16, 8, 8, 2., 1., 1.), vec=3, dim=3, dirichlet_bc_info=clamp,
                       location_fns=[lambda x: isclose(x[0], 2.)])
    theta = (torch.rand(p.num_cells, dtype=torch.float64, generator=torch.Generator().manual_seed(5)) * 0.8
             + 0.2).to('cuda')
    grads = []
    for opts in ({'jax_solver': {'tol': 1e-12, 'atol': 0.}}, {'amgx_solver': {'tol': 1e-12, 'atol': 0.}}):
        fwd = ad_wrapper(p, opts, opts)
        th = theta.clone().requires_grad_(True)
        u = fwd(th)[0]
        c = (u * u).sum()
        c.backward()
        grads.append(th.grad.clone())
    assert torch.linalg.norm(grads[1] - grads[0]) <= 1e-8 * torch.linalg.norm(grads[0])


def test_amg_hex27_matches_jacobi_pcg(gpu):
    # 16^3 HEX27 cells: rows of R (A P) have more than 16384 block products, so the
    # product plans are built in several sorted chunks
    from jax_fem_amd.generate_mesh import Mesh, box_mesh_hex27
    m = box_mesh_hex27(16, 16, 16, 1., 1., 1.)
    clamp = [[lambda x: isclose(x[2], 0.)] * 3, [0, 1, 2], [lambda x: 0.] * 3]
    p = LoadedElastic(Mesh(m.points, m.cells_dict['hexahedron27']), vec=3, dim=3, ele_type='HEX27',
                      quadrature_order=4, dirichlet_bc_info=clamp, location_fns=[lambda x: isclose(x[2], 1.)])
    ref = solver(p, {'jax_solver': {'precision': 'fp64', 'tol': 1e-12, 'atol': 0.}})[0].clone()
    its_j = p._solve_stats['linear'][0]['iterations']
    x = solver(p, {'amgx_solver': {'precision': 'mixed', 'tol': 1e-12, 'atol': 0.}})[0]
    its = p._solve_stats['linear'][0]['iterations']
    assert torch.linalg.norm(x - ref) <= 1e-9 * torch.linalg.norm(ref)
    assert its < its_j / 2, (its, its_j)
    y = solver(p, {'amgx_solver': {'precision': 'mixed', 'tol': 1e-12, 'atol': 0.}})[0]
    assert torch.equal(x, y)


def test_amg_mixed_outer_fp32_replacements(gpu):
    # precision 'mixed': the outer CG runs on the FP32 difference-form matrix and
    # replaces its residual by the FP64 one; the FP64 path needs no replacement
    p = _elastic(16)
    dev, A, _, mask, _, b = _assembled(p)
    ref = spla.spsolve(A.tocsc(), b.cpu().numpy())
    out = {}
    for precision in ('fp64', 'mixed'):
        x = solver(p, {'amgx_solver': {'precision': precision, 'tol': 1e-12, 'atol': 0.,
                                       'outer': 'fp32' if precision == 'mixed' else 'fp64'}})[0]
        st = p._solve_stats['linear'][0]
        out[precision] = (x.reshape(-1).cpu().numpy(), st['iterations'], st['replacements'], st['res_norm'])
    assert out['fp64'][2] == 0 and out['mixed'][2] >= 2, out
    for precision, (x, its, repl, res) in out.items():
        assert np.linalg.norm(x - ref) <= 1e-9 * np.linalg.norm(ref), precision
    assert out['mixed'][1] <= out['fp64'][1] + 5
    # setup reuse: the FP32 copy follows the current matrix (the tangent is re-assembled per solve)
    opts = {'amgx_solver': {'precision': 'mixed', 'outer': 'fp32', 'tol': 1e-12, 'atol': 0., 'setup_reuse': 3}}
    x1 = solver(p, opts)[0].reshape(-1).cpu().numpy()
    x2 = solver(p, opts)[0].reshape(-1).cpu().numpy()
    assert np.linalg.norm(x2 - ref) <= 1e-9 * np.linalg.norm(ref)
    assert np.linalg.norm(x1 - ref) <= 1e-9 * np.linalg.norm(ref)


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_amg_small_problem_dense_coarse_solve(gpu, precision):
    # 4x2x2 cells: 135 unknowns, below the coarse size, so the hierarchy has no
    # smoothed level and the preconditioner is the dense inverse of the matrix
    p = _elastic(4)
    dev, A, _, mask, _, b = _assembled(p)
    assert dev.n <= 512
    amg = dev.amg_setup(_cfg(precision))
    assert amg.levels == [] and amg.amg.num_levels == 0
    ref = spla.spsolve(A.tocsc(), b.cpu().numpy())
    r = b.cpu().numpy().copy()
    z = amg.vcycle(torch.as_tensor(r, device=dev.dm.device)).cpu().numpy()
    assert np.linalg.norm(z - ref) <= 1e-10 * np.linalg.norm(ref)
    sol = solver(p, {'amgx_solver': {'precision': precision, 'tol': 1e-12, 'atol': 0.}})[0]
    x = sol.reshape(-1).cpu().numpy()
    assert np.linalg.norm(x - ref) <= 1e-10 * np.linalg.norm(ref)
    assert p._solve_stats['linear'][0]['iterations'] <= 2


def test_amg_level0_fp16_scaled_smoother(gpu):
    # level0='fp16': the level-0 smoother reads FP16 D^-1 A in difference form; the cycle
    # differs from the FP32 one by the FP16 rounding, the solve reaches the same solution
    p = _elastic(16)
    dev, A, (gp, gi), mask, pts, b = _assembled(p)
    amg32 = dev.amg_setup(_cfg('mixed'))
    r = torch.as_tensor(np.random.default_rng(12).standard_normal(dev.n), device=dev.dm.device)
    r[torch.as_tensor(mask, device=r.device)] = 0.
    z32 = amg32.vcycle(r).clone()
    lam32 = amg32.levels[0]['lam']
    amg16 = dev.amg_setup(_cfg('mixed', level0='fp16'))
    assert amg16.level0_f16 and amg16.amg.level0_form == _lib.FEM_FORM_SCALED_DIFF16
    assert abs(amg16.levels[0]['lam'] - lam32) <= 2e-3 * lam32
    z16 = amg16.vcycle(r)
    assert torch.linalg.norm(z16 - z32) <= 5e-3 * torch.linalg.norm(z32)
    assert float(z16[torch.as_tensor(mask, device=r.device)].abs().max()) == 0.0
    ref = spla.spsolve(A.tocsc(), b.cpu().numpy())
    its = {}
    for level0 in ('fp32', 'fp16'):
        x = solver(p, {'amgx_solver': {'precision': 'mixed', 'level0': level0, 'tol': 1e-12, 'atol': 0.}})[0]
        its[level0] = p._solve_stats['linear'][0]['iterations']
        assert np.linalg.norm(x.reshape(-1).cpu().numpy() - ref) <= 1e-9 * np.linalg.norm(ref)
    assert its['fp16'] <= its['fp32'] + 3, its


def test_sell_to_bcsr_on_a_length_sorted_pattern(gpu):
    # HEX27 rows have 27 to 125 blocks, so the SELL rows are length-sorted within windows
    # (sell_row is not the identity): the block-CSR copy must still put each row's blocks
    # in pattern order, bit for bit
    from jax_fem_amd._lib import check, ptr, stream_ptr
    from jax_fem_amd.generate_mesh import Mesh, box_mesh_hex27
    m = box_mesh_hex27(4, 3, 3, 1., 1., 1.)
    clamp = [[lambda x: isclose(x[2], 0.)] * 3, [0, 1, 2], [lambda x: 0.] * 3]
    p = LoadedElastic(Mesh(m.points, m.cells_dict['hexahedron27']), vec=3, dim=3, ele_type='HEX27',
                      quadrature_order=4, dirichlet_bc_info=clamp, location_fns=[lambda x: isclose(x[2], 1.)])
    dev, A, (gp, gi), mask, pts, b = _assembled(p)
    assert dev.sell_sorted
    sell_row = dev.pat_tensors['sell_row'].cpu().numpy()
    assert not np.array_equal(sell_row[:dev.dm.Nn], np.arange(dev.dm.Nn))
    vals = dev.current[2]
    for precision, dtype in ((_lib.FEM_F64, torch.float64), (_lib.FEM_F32, torch.float32)):
        out = torch.empty(dev.nnzb * 9, dtype=dtype, device='cuda')
        check(dev.L.fem_amg_sell_to_bcsr(dev.pref, ptr(vals), ptr(out), precision, stream_ptr()),
              'fem_amg_sell_to_bcsr')
        B = sp.bsr_matrix((out.cpu().numpy().astype(np.float64).reshape(-1, 3, 3), gi, gp),
                          shape=(dev.n, dev.n)).tocsr()
        B.sort_indices()
        assert np.array_equal(B.indptr, A.indptr) and np.array_equal(B.indices, A.indices)
        # FP32: each value rounded once from the FP64 matrix
        ref = A.data if dtype == torch.float64 else A.data.astype(np.float32).astype(np.float64)
        assert np.array_equal(B.data, ref)


@pytest.mark.parametrize('mesh', ['hex8', 'hex27_sorted', 'poisson'])
def test_fused_f32_copies_match_the_separate_copies(gpu, mesh):
    # fem_amg_sell_to_f32_copies writes the FP32 block-CSR copy and the FP32 difference-form
    # SELL copy in one pass: both bit for bit the outputs of the separate copy kernels
    # (padding lanes, length-sorted rows and the vec-1 pattern included)
    from jax_fem_amd._lib import check, ptr, stream_ptr
    if mesh == 'hex27_sorted':
        from jax_fem_amd.generate_mesh import Mesh, box_mesh_hex27
        m = box_mesh_hex27(4, 3, 3, 1., 1., 1.)
        clamp = [[lambda x: isclose(x[2], 0.)] * 3, [0, 1, 2], [lambda x: 0.] * 3]
        p = LoadedElastic(Mesh(m.points, m.cells_dict['hexahedron27']), vec=3, dim=3, ele_type='HEX27',
                          quadrature_order=4, dirichlet_bc_info=clamp, location_fns=[lambda x: isclose(x[2], 1.)])
    else:
        p = _elastic(6) if mesh == 'hex8' else _poisson(6)
    dev = _assembled(p)[0]
    if mesh == 'hex27_sorted':
        assert dev.sell_sorted
    vals = dev.current[2]
    V = dev.V
    bcsr = torch.empty(dev.nnzb * V * V, dtype=torch.float32, device='cuda')
    diff = torch.empty(dev.sell_blocks * V * V, dtype=torch.float32, device='cuda')
    check(dev.L.fem_amg_sell_to_bcsr(dev.pref, ptr(vals), ptr(bcsr), _lib.FEM_F32, stream_ptr()),
          'fem_amg_sell_to_bcsr')
    check(dev.L.fem_amg_sell_to_f32_diff(dev.pref, ptr(vals), ptr(diff), stream_ptr()), 'fem_amg_sell_to_f32_diff')
    bcsr2 = torch.full_like(bcsr, float('nan'))
    diff2 = torch.full_like(diff, float('nan'))
    dinv2 = torch.full((dev.n,), float('nan'), dtype=torch.float64, device='cuda')
    check(dev.L.fem_amg_sell_to_f32_copies(dev.pref, ptr(vals), ptr(bcsr2), ptr(diff2), ptr(dinv2), stream_ptr()),
          'fem_amg_sell_to_f32_copies')
    # the inverse diagonal as fem_amg_dinv forms it from the block-CSR copy
    d = DeviceAMG(dev)
    from jax_fem_amd.amg import Bcsr
    rp, ci = dev.pat_tensors['row_ptr'], dev.pat_tensors['col_idx']
    M = Bcsr(dev.n // V, dev.n // V, V, V, rp, ci, bcsr)
    dinv = d._dinv(M, fix=False)
    torch.cuda.synchronize()
    assert torch.equal(bcsr2.view(torch.int32), bcsr.view(torch.int32))
    assert torch.equal(diff2.view(torch.int32), diff.view(torch.int32))
    assert torch.equal(dinv2.view(torch.int64), dinv.view(torch.int64))


@pytest.mark.parametrize('nb', [4, 10, 12, 22, 30, 147])
def test_coarse_inverse_against_numpy(gpu, nb):
    # the dense coarse solve's inverse: one workgroup in LDS up to 64 unknowns (nb <= 10 blocks
    # of 6), a launch per pivot from 65 to 127 (nb = 12), 64-wide pivot blocks from 128 (nb = 22:
    # 132 unknowns, a last block of 4; nb = 147: 882 unknowns, the 1.6M-dof cantilever's coarse
    # level); all against numpy's inverse of the same SPD matrix
    from jax_fem_amd._lib import check, ptr, stream_ptr
    from jax_fem_amd.amg import Bcsr
    rng = np.random.default_rng(nb)
    n = 6 * nb
    G = rng.standard_normal((n, n))
    Adense = G @ G.T + n * np.eye(n)
    blocks = Adense.reshape(nb, 6, nb, 6).transpose(0, 2, 1, 3).reshape(-1)
    dev = torch.device('cuda')
    rp = torch.arange(0, nb * nb + 1, nb, dtype=torch.int32, device=dev)
    ci = torch.arange(nb, dtype=torch.int32, device=dev).repeat(nb)
    M = Bcsr(nb, nb, 6, 6, rp, ci, torch.tensor(blocks, dtype=torch.float64, device=dev))
    s = M.struct()
    inv = torch.empty(n * n, dtype=torch.float64, device=dev)
    ws = torch.empty(n * n, dtype=torch.float64, device=dev)
    L = _lib.load()
    check(L.fem_amg_coarse_inverse(C.byref(s), ptr(inv), ptr(ws), stream_ptr()), 'fem_amg_coarse_inverse')
    got = inv.cpu().numpy().reshape(n, n)
    ref = np.linalg.inv(Adense)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_device_lmax_and_omega_match_the_host_path(gpu, precision):
    # the setup keeps each level's lambda_max estimate on the device and forms omega there
    # (fem_amg_lmax_async, fem_amg_smooth_prolongator_dev): the same estimate as the
    # synchronising fem_amg_lmax, and the same smoothed prolongator, bit for bit, as
    # fem_amg_smooth_prolongator with omega = 1.5 / lambda_max passed from the host
    from jax_fem_amd._lib import check, ptr, stream_ptr
    p = _elastic(16)
    dev = _assembled(p)[0]
    amg = dev.amg_setup(_cfg(precision))
    for l, lv in enumerate(amg.levels):
        sell = (amg.vals32 if amg.vals32 is not None else dev.current[2]) if l == 0 else None
        lam_host = amg._lmax(lv['A'], lv['dinv'], sell)
        assert lam_host == lv['lam']
        P = lv['P']
        ref = torch.empty_like(P.vals)
        as_, p0s, ps = lv['A'].struct(), lv['P0m'].struct(), P.struct(ref)
        plan_w0 = lv['plans'][0]
        check(amg.L.fem_amg_smooth_prolongator(C.byref(as_), C.byref(p0s), C.byref(ps), ptr(plan_w0[0]),
                                               ptr(plan_w0[1]), ptr(lv['agg']), ptr(lv['dinv']),
                                               1.5 / lam_host, stream_ptr()), 'fem_amg_smooth_prolongator')
        torch.cuda.synchronize()
        bits = torch.int32 if P.vals.dtype == torch.float32 else torch.int64
        assert torch.equal(ref.view(bits), P.vals.view(bits))
