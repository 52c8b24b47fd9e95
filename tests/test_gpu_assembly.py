"""Single-pass HEX8 assembly (fem_assemble) against the two-pass path (per-cell
tangents + gather) and the oracle: every built-in law, both storage forms, the
fused residual, an unstructured mesh, and run-to-run bit reproducibility
(problem.py:426-460, solver.py:522-529)."""
import numpy as np
import pytest
import torch

from conftest import golden
from oracle import fem as OF
from problems import Plasticity, isclose
from test_gpu_parity import MATS, distorted_box, make_material_problem, rel

from jax_fem_amd import _lib
from jax_fem_amd import materials as M
from jax_fem_amd.generate_mesh import Mesh
from jax_fem_amd.problem import Problem

pytestmark = pytest.mark.gpu


def _both(dev, ut, precision):
    """(single-pass, two-pass) SELL values and inverse diagonals."""
    out = []
    for fused in (True, False):
        dev.tangent(ut, precision, fused=fused, residual=True)
        _, _, vals, dinv = dev.current
        out.append((vals.clone(), dinv.clone(), dev.last_residual.clone()))
    return out


@pytest.mark.parametrize('name', list(MATS))
def test_single_pass_matches_two_pass_and_oracle(gpu, name):
    p, o, u = make_material_problem(name)
    dev = p.device()
    dev.prepare()
    assert dev.fused_applies(_lib.FEM_F64)
    ut = torch.as_tensor(u, device='cuda')
    (v1, d1, r1), (v2, d2, r2) = _both(dev, ut, _lib.FEM_F64)
    scale = float(v2.abs().max())
    assert float((v1 - v2).abs().max()) <= 1e-13 * scale
    assert float((d1 - d2).abs().max()) <= 1e-13 * float(d2.abs().max())
    assert float((r1 - r2).abs().max()) <= 1e-12 * float(r2.abs().max())
    # against the oracle's COO -> CSR + zeroRows, and its residual
    dev.tangent(ut, _lib.FEM_F64)
    indptr, indices, data = dev.export_csr()
    o.bc_dofs = dev.bc_dofs.cpu().numpy().astype(np.int64)
    A = o.tangent(u)
    assert np.array_equal(indptr, A.indptr) and np.array_equal(indices, A.indices)
    assert rel(data, A.data) < 1e-12
    assert rel(r1.cpu().numpy(), o.residual(u, bc=False)) < 1e-12
    # FP32 difference form: the same FP64 sums rounded once
    (f1, e1, _), (f2, e2, _) = _both(dev, ut, _lib.FEM_F32)
    assert float((f1.double() - f2.double()).abs().max()) <= 1e-6 * float(f2.abs().max())
    assert float((e1.double() - e2.double()).abs().max()) <= 1e-6 * float(e2.abs().max())


def test_single_pass_j2_with_state(gpu):
    # a random stress/strain history that puts most quadrature points on the plastic branch
    p = Plasticity(distorted_box((3, 3, 4), seed=4), vec=3, dim=3,
                   dirichlet_bc_info=[[lambda x: isclose(x[0], 0.)] * 3, [0, 1, 2], [lambda x: 0.] * 3])
    C, Q = p.num_cells, p.fes[0].num_quads
    rng = np.random.default_rng(5)
    sym = lambda a: 0.5 * (a + np.swapaxes(a, -1, -2))
    p.internal_vars = [torch.as_tensor(sym(150. * rng.standard_normal((C, Q, 3, 3)))),
                       torch.as_tensor(sym(2e-3 * rng.standard_normal((C, Q, 3, 3))))]
    u = 2e-3 * rng.standard_normal(p.fes[0].num_total_dofs)
    dev = p.device()
    dev.prepare()
    assert isinstance(dev.material, M.J2Plastic)
    assert dev.fused_applies(_lib.FEM_F64)
    ut = torch.as_tensor(u, device='cuda')
    (v1, d1, r1), (v2, d2, r2) = _both(dev, ut, _lib.FEM_F64)
    assert float((v1 - v2).abs().max()) <= 1e-13 * float(v2.abs().max())
    assert float((r1 - r2).abs().max()) <= 1e-12 * float(r2.abs().max())


def test_single_pass_bit_reproducible(gpu):
    p, o, u = make_material_problem('neo_hookean')
    dev = p.device()
    dev.prepare()
    ut = torch.as_tensor(u, device='cuda')
    outs = []
    for _ in range(3):
        dev.tangent(ut, _lib.FEM_F64, residual=True)
        outs.append((dev.current[2].clone(), dev.last_residual.clone()))
    for v, r in outs[1:]:
        assert torch.equal(v, outs[0][0]) and torch.equal(r, outs[0][1])


class _Elastic(Problem):
    def get_tensor_map(self):
        return M.LinearElastic(E=70e3, nu=0.3)

    def get_mass_map(self):
        return lambda u, x: torch.stack([x[0], -x[1], 1. + 0. * x[2]])


def test_single_pass_on_the_cylinder_mesh(gpu):
    # the reference's unstructured gmsh cylinder (fixture mesh): irregular incidences and slices
    d = golden('linear_elasticity_cylinder')
    mesh = Mesh(d['points'], d['cells'])
    bottom = lambda x: isclose(x[2], 0.)
    p = _Elastic(mesh, vec=3, dim=3, dirichlet_bc_info=[[bottom] * 3, [0, 1, 2], [lambda x: 0.] * 3])
    dev = p.device()
    dev.prepare()
    if dev.assembly_plan() is None:
        pytest.skip("rows of more than 32 blocks: the two-pass path assembles this mesh")
    u = 1e-3 * np.random.default_rng(2).standard_normal(dev.n)
    ut = torch.as_tensor(u, device='cuda')
    (v1, d1, r1), (v2, d2, r2) = _both(dev, ut, _lib.FEM_F64)
    assert float((v1 - v2).abs().max()) <= 1e-13 * float(v2.abs().max())
    assert float((r1 - r2).abs().max()) <= 1e-12 * float(r2.abs().max())
    lm, mu = OF.lame(70e3, 0.3)
    o = OF.Problem(mesh.points, mesh.cells, 3, mat=OF.Material('elastic', lmbda=lm, mu=mu))
    o.bc_dofs = dev.bc_dofs.cpu().numpy().astype(np.int64)
    dev.tangent(ut, _lib.FEM_F64)
    indptr, indices, data = dev.export_csr()
    assert rel(data, o.tangent(u).data) < 1e-12
