"""Benchmark: seconds per Newton step and per adjoint gradient, mixed vs FP64.

Workload (N=1): linear-elasticity cantilever, HEX8 box 2x1x1 with 128x64x64
cells (545,025 nodes, 1,635,075 dofs), E = 70e3, nu = 0.3, clamped at x = 0,
uniform traction (0, 0, -1) on x = 2.  The Young's modulus is carried as a
per-quadrature-point field E(theta) = 70e3 * theta with theta = 1, so the
same problem also yields the adjoint gradient dC/dtheta of the compliance
C = f.u.  Data are synthetic (generated mesh); inputs are resident on the GPU.

A step is one Newton step of ``solver`` (residual, tangent assembly, BCs,
Jacobi-PCG to |r| <= 1e-10 |b|, update, new residual); for this linear problem
``solver`` performs exactly one.  ``value`` is the mixed-precision time
(FP32 matrix + FP32 PCG, FP64 iterate and FP64 residual replacement);
the plain FP64 path is timed beside it, and its tight solve (1e-12) is the
reference for the relative error.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'jax-fem_amd'))
sys.path.insert(0, ROOT)

METRIC = "s per Newton step & per adjoint gradient, mixed vs FP64, at rel.err ≤1e-8"
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_problem(nx, ny, nz):
    import numpy as np
    import torch
    from jax_fem_amd import materials
    from jax_fem_amd.generate_mesh import Mesh, box_mesh
    from jax_fem_amd.problem import Problem

    class Cantilever(Problem):
        def get_tensor_map(self):
            return materials.SIMPElastic(E0=70e3, Emin=0., nu=0.3, penal=1.)

        def get_surface_maps(self):
            return [lambda u, x: torch.tensor([0., 0., 1.], dtype=torch.float64)]

        def set_params(self, theta):
            self.internal_vars = [theta.reshape(-1, 1).expand(-1, self.fes[0].num_quads)]

    m = box_mesh(nx, ny, nz, 2., 1., 1.)
    mesh = Mesh(m.points, m.cells_dict['hexahedron'])
    left = lambda p: np.isclose(p[0], 0., atol=1e-5)
    right = lambda p: np.isclose(p[0], 2., atol=1e-5)
    p = Cantilever(mesh, vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.] * 3],
                   location_fns=[right])
    return p


def sell_bytes(dev, elem_bytes, vec=3):
    """Algorithmic bytes of one fused SpMV launch: SELL values + column indices once,
    z and p_old read once per row (gathers counted once), p_new and q written."""
    n = dev.n
    return dev.sell_blocks * (vec * vec * elem_bytes + 4) + n * elem_bytes * 4


def run(args):
    import numpy as np
    import torch
    from jax_fem_amd.solver import solver, ad_wrapper

    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    p = build_problem(*args.cells)
    theta = torch.ones(p.num_cells, dtype=torch.float64, device='cuda')
    p.set_params(theta)
    dev = p.device()
    dev.prepare()
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t0:.1f}s: {dev.dm.C} cells, {dev.n} dofs, {dev.nnzb * 9} nnz "
        f"({dev.sell_blocks} SELL blocks)")

    opts = {
        'fp64': {'jax_solver': {'precision': 'fp64', 'timing': True}},
        'mixed': {'jax_solver': {'precision': 'mixed', 'assembly': args.mixed_assembly, 'timing': True}},
    }
    # reference: FP64 path solved to rtol 1e-11 (about the attainable FP64 residual here)
    ref_opts = {'jax_solver': {'precision': 'fp64', 'tol': 1e-11, 'atol': 0.}}
    ref = None
    if args.only is None:
        ref = solver(p, ref_opts)[0].clone()
        log(f"reference FP64 solve (rtol 1e-11): {p._solve_stats['linear'][0]['iterations']} PCG its")

    def newton_step(name):
        torch.cuda.synchronize()
        t = time.perf_counter()
        sol = solver(p, opts[name])[0]
        torch.cuda.synchronize()
        return time.perf_counter() - t, sol, p._solve_stats

    res = {}
    paths = ('fp64', 'mixed') if args.only is None else (args.only,)
    for name in paths:
        for _ in range(args.warmup):
            newton_step(name)
        times, stats = [], []
        for _ in range(args.steps):
            dt, sol, st = newton_step(name)
            times.append(dt)
            stats.append(st)
        err = float(torch.linalg.vector_norm(sol - ref) / torch.linalg.vector_norm(ref)) if ref is not None \
            else float('nan')
        lin = stats[-1]['linear'][0]
        res[name] = dict(s_per_newton_step=statistics.median(times), times=times, rel_err=err,
                         newton_iterations=stats[-1]['newton_iterations'], pcg_iterations=lin['iterations'],
                         replacements=lin['replacements'], spmv_ms=lin['spmv_ms'], update_ms=lin['update_ms'],
                         timing=stats[-1]['timing'])
        log(f"{name}: {res[name]['s_per_newton_step']:.4f} s/step, rel err {err:.2e}, "
            f"{lin['iterations']} its, {lin['replacements']} repl, spmv {lin['spmv_ms']*1e3:.1f} us, "
            f"update {lin['update_ms']*1e3:.1f} us")

    if args.only is not None:
        if args.calib_copy_mib:
            copy_calibration(args.calib_copy_mib)
        print(json.dumps({name: {k: v for k, v in res[name].items()} for name in paths}), flush=True)
        return
    # adjoint gradient of the compliance C = f.u w.r.t. theta
    f = -dev.load
    for name in ('fp64', 'mixed'):
        fwd = ad_wrapper(p, opts[name], opts[name])
        tms = []
        for it in range(args.warmup + args.steps):
            th = theta.clone().requires_grad_(True)
            C = torch.dot(f, fwd(th)[0].reshape(-1))
            torch.cuda.synchronize()
            t = time.perf_counter()
            C.backward()
            torch.cuda.synchronize()
            if it >= args.warmup:
                tms.append(time.perf_counter() - t)
        g = th.grad
        res[name]['s_per_adjoint_gradient'] = statistics.median(tms)
        res[name]['adjoint_pcg_iterations'] = p._adjoint_stats['iterations']
        res[name]['grad'] = g.detach().clone()
        log(f"{name}: adjoint {res[name]['s_per_adjoint_gradient']:.4f} s, its {p._adjoint_stats['iterations']}")
    gerr = float(torch.linalg.vector_norm(res['mixed']['grad'] - res['fp64']['grad']) /
                 torch.linalg.vector_norm(res['fp64']['grad']))
    for name in ('fp64', 'mixed'):
        del res[name]['grad']

    mixed = res['mixed']
    spmv_bytes = sell_bytes(dev, 4)
    achieved = spmv_bytes / (mixed['spmv_ms'] * 1e-3) / 1e9 if mixed['spmv_ms'] else None
    fp64_bytes = sell_bytes(dev, 8)
    fp64_achieved = fp64_bytes / (res['fp64']['spmv_ms'] * 1e-3) / 1e9 if res['fp64']['spmv_ms'] else None
    out = {
        "metric": METRIC, "value": mixed['s_per_newton_step'], "unit": "s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": mixed['s_per_newton_step'] * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"HEX8 linear-elasticity cantilever 2x1x1, {args.cells[0]}x{args.cells[1]}x"
                               f"{args.cells[2]} cells, {dev.n} dofs, Jacobi-PCG rtol 1e-10",
                   "global_batch": 1, "seq_len": dev.n, "parallelism": "single GPU"},
        "fp64_s_per_newton_step": res['fp64']['s_per_newton_step'],
        "speedup_mixed_vs_fp64": res['fp64']['s_per_newton_step'] / mixed['s_per_newton_step'],
        "rel_err_mixed_vs_fp64_ref": mixed['rel_err'], "rel_err_fp64_vs_fp64_ref": res['fp64']['rel_err'],
        "s_per_adjoint_gradient": {"mixed": mixed['s_per_adjoint_gradient'],
                                   "fp64": res['fp64']['s_per_adjoint_gradient']},
        "adjoint_grad_rel_diff_mixed_vs_fp64": gerr,
        "pcg_iterations": {"mixed": mixed['pcg_iterations'], "fp64": res['fp64']['pcg_iterations'],
                           "mixed_replacements": mixed['replacements']},
        "roofline": {"bound": "hbm", "kernel": "k_pcg_spmv<3,float> (SELL-64 BSR3 FP32)",
                     "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS if achieved else None, "traffic": None,
                     "algorithmic_bytes": spmv_bytes, "avg_launch_ms": mixed['spmv_ms']},
        "fp64_spmv": {"achieved_GBs": fp64_achieved, "avg_launch_ms": res['fp64']['spmv_ms'],
                      "algorithmic_bytes": fp64_bytes},
        "detail": res,
    }
    if args.cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(dev, res['fp64']['pcg_iterations'])
    print(json.dumps(out), flush=True)


def copy_calibration(mib):
    """Plain streaming kernel of a known size (3 launches of an int32 bitwise_not:
    reads n*4 bytes, writes n*4 bytes): the reference against which hardware
    byte counters are scaled when profiling."""
    import torch
    n = mib * 2**20 // 4
    src = torch.randint(0, 2**30, (n,), dtype=torch.int32, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.bitwise_not(src, out=dst)
    torch.cuda.synchronize()
    log(f"copy calibration: {n * 4} bytes read + {n * 4} bytes written per launch")


def cpu_baseline(dev, gpu_iterations):
    """Oracle FP64 CPU restatement on a bounded sample, scaled to the workload."""
    import numpy as np
    from oracle import fem as OF
    nx, ny, nz = 32, 16, 16
    pts, cells = OF.box_mesh(nx, ny, nz, 2., 1., 1.)
    left = lambda p: p[0] < 1e-9
    o = OF.Problem(pts, cells, 3, mat=OF.Material('elastic', **dict(zip(('lmbda', 'mu'), OF.lame(70e3, 0.3)))),
                   bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.] * 3])
    u0 = np.zeros(o.n)
    t = time.perf_counter()
    A = o.tangent(u0)
    t_asm = time.perf_counter() - t
    b = np.random.default_rng(0).standard_normal(o.n)
    its = 50
    t = time.perf_counter()
    OF.pcg_jacobi(A, b, np.zeros(o.n), A.diagonal(), tol=0., atol=0., maxiter=its)
    t_it = (time.perf_counter() - t) / its
    C = len(cells)
    scale_cells = dev.dm.C / C
    scale_nnz = dev.nnzb * 9 / A.nnz
    value = t_asm * scale_cells + t_it * scale_nnz * gpu_iterations
    return {"value": value, "unit": "s per Newton step", "cores": 1, "kind": "port",
            "sample": f"oracle (numpy/scipy FP64) on a {nx}x{ny}x{nz} cantilever: assembly {t_asm:.2f}s "
                      f"scaled by cells x{scale_cells:.0f}, {its} Jacobi-PCG iterations ({t_it*1e3:.1f} ms each) "
                      f"scaled by nnz x{scale_nnz:.0f} and by the {gpu_iterations} iterations of the FP64 GPU solve"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--cells', type=int, nargs=3, default=[128, 64, 64])
    ap.add_argument('--mixed-assembly', choices=['fp32', 'fp64'], default='fp64')
    ap.add_argument('--no-cpu-baseline', dest='cpu_baseline', action='store_false')
    ap.add_argument('--only', choices=['fp64', 'mixed'], default=None,
                    help='time one path only (no adjoint, no summary line); for profiling')
    ap.add_argument('--calib-copy-mib', type=int, default=0,
                    help='with --only: also run a plain copy of this many MiB (counter calibration)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X (--gpus {args.gpus} requested); "
              "multi-GPU sharding is described in DESIGN.md but not implemented", file=sys.stderr)
        sys.exit(2)
    run(args)


if __name__ == '__main__':
    main()
