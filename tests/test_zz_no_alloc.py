"""fem_pcg_solve_noalloc (the Jacobi-PCG solve on the library's own device
scratch, for callers that cannot allocate) against fem_pcg_solve with a caller
workspace: the same kernels in the same order, so the iterates must agree bit
for bit.  Two systems in a row, the second smaller, on the FP64 and the mixed
FP32/FP64 path; systems above the scratch size are refused."""
import ctypes as C

import numpy as np
import pytest
import torch

from problems import box, isclose
from jax_fem_amd import _lib, materials as M
from jax_fem_amd.problem import Problem
from jax_fem_amd.solver import solver


class Cantilever(Problem):
    def get_tensor_map(self):
        return M.LinearElastic(E=70e3, nu=0.3)

    def get_surface_maps(self):
        return [lambda u, x: torch.tensor([0., 0., -1.], dtype=torch.float64)]


def _problem(n):
    mesh = box(*n, 2., 1., 1.)
    left = lambda p: isclose(p[0], 0.)
    return Cantilever(mesh, vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.] * 3],
                      location_fns=[lambda p: isclose(p[0], 2.)])


def _options(precision):
    o = _lib.FemPcgOptions()
    o.rtol, o.atol, o.maxiter, o.check_every = 1e-10, 0., 20000, 16
    o.precision, o.precond, o.reliable_delta, o.xcd_remap, o.timing = precision, 1, 0.01, 1, 0
    return o


def _both(p, key):
    """Solve the assembled system with both entry points on identical inputs."""
    solver(p, {'jax_solver': {'precision': key}})      # assembles the tangent in this precision
    dev = p.device()
    precision, _, vals, dinv = dev.current
    L = _lib.lib()
    ptr = lambda t: C.c_void_p(t.data_ptr())
    gen = torch.Generator(device='cuda').manual_seed(7)
    b = torch.randn(dev.n, dtype=torch.float64, device='cuda', generator=gen)
    b[dev.bc_dofs_long] = 0.
    u = torch.zeros(dev.n, dtype=torch.float64, device='cuda')
    op = dev.operator(u) if precision == _lib.FEM_F32 else None
    opt = _options(precision)
    status = torch.empty(128, dtype=torch.uint8).pin_memory()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    ws = torch.empty(int(L.fem_pcg_ws_bytes(dev.pref, precision)), dtype=torch.uint8, device='cuda')
    x_ws = torch.zeros_like(b)
    r_ws = _lib.FemPcgResult()
    rc = L.fem_pcg_solve(dev.pref, ptr(vals), ptr(dinv), C.byref(op) if op else None, ptr(b), ptr(x_ws),
                         C.byref(opt), ptr(ws), C.c_void_p(status.data_ptr()), C.byref(r_ws), stream)
    assert rc == _lib.FEM_OK, _lib.lib().fem_last_error()
    x_na = torch.zeros_like(b)
    r_na = _lib.FemPcgResult()
    rc = L.fem_pcg_solve_noalloc(dev.pref, ptr(vals), ptr(dinv), C.byref(op) if op else None, ptr(b), ptr(x_na),
                                 C.byref(opt), C.c_void_p(status.data_ptr()), C.byref(r_na), stream)
    assert rc == _lib.FEM_OK, _lib.lib().fem_last_error()
    torch.cuda.synchronize()
    return dev, x_ws, r_ws, x_na, r_na, b


@pytest.mark.gpu
@pytest.mark.parametrize('key', ['fp64', 'mixed'])
def test_noalloc_pcg_matches_workspace_pcg(gpu, key):
    for n in ((12, 6, 6), (6, 3, 3)):       # the second system is smaller: stale scratch must not matter
        dev, x_ws, r_ws, x_na, r_na, b = _both(_problem(n), key)
        print(f"noalloc {key} {n}: {dev.n} dofs, {r_na.iterations} iterations", flush=True)
        assert r_ws.status == _lib.FEM_OK and r_na.status == _lib.FEM_OK
        assert r_na.iterations == r_ws.iterations and r_na.replacements == r_ws.replacements
        assert torch.equal(x_na, x_ws)
        assert r_na.res_norm == r_ws.res_norm
        assert r_na.res_norm <= 1e-10 * r_na.b_norm
        assert float(torch.linalg.vector_norm(x_na)) > 0.


@pytest.mark.gpu
def test_noalloc_pcg_refuses_timing_and_oversize(gpu):
    p = _problem((4, 2, 2))
    solver(p, {'jax_solver': {}})
    dev = p.device()
    _, _, vals, dinv = dev.current
    L = _lib.lib()
    x = torch.zeros(dev.n, dtype=torch.float64, device='cuda')
    status = torch.empty(128, dtype=torch.uint8).pin_memory()
    opt = _options(_lib.FEM_F64)
    opt.timing = 1
    res = _lib.FemPcgResult()
    args = lambda pat: (pat, C.c_void_p(vals.data_ptr()), C.c_void_p(dinv.data_ptr()), None,
                        C.c_void_p(x.data_ptr()), C.c_void_p(x.data_ptr()), C.byref(opt),
                        C.c_void_p(status.data_ptr()), C.byref(res), None)
    assert L.fem_pcg_solve_noalloc(*args(dev.pref)) == _lib.FEM_ERR_BAD_ARG
    opt.timing = 0
    big = _lib.FemPattern.from_buffer_copy(dev.pat)
    big.num_rows = 1 << 24                      # sizes only: the call is refused before any launch
    big.num_slices = big.num_rows // 64
    assert L.fem_pcg_ws_bytes(C.byref(big), _lib.FEM_F64) > _lib.FEM_PCG_NOALLOC_SCRATCH_BYTES
    assert L.fem_pcg_solve_noalloc(*args(C.byref(big))) == _lib.FEM_ERR_CAPACITY
    assert b'scratch' in L.fem_last_error()
