"""Benchmark: seconds per Newton step and per adjoint gradient, mixed vs FP64.

Workload (N=1): linear-elasticity cantilever, HEX8 box 2x1x1 with 128x64x64
cells (545,025 nodes, 1,635,075 dofs), E = 70e3, nu = 0.3, clamped at x = 0,
uniform traction (0, 0, -1) on x = 2.  The Young's modulus is carried as a
per-quadrature-point field E(theta) = 70e3 * theta with theta = 1, so the
same problem also yields the adjoint gradient dC/dtheta of the compliance
C = f.u.  Data are synthetic (generated mesh); inputs are resident on the GPU.

A step is one Newton step of ``solver`` (residual, tangent assembly, BCs,
linear solve to |r| <= 1e-10 |b|, update, new residual); for this linear
problem ``solver`` performs exactly one.  Four linear-solve paths are timed:
  amg_mixed  smoothed-aggregation AMG-CG ('amgx_solver', precision 'mixed'):
             FP64 outer CG on the FP64 matrix, V-cycle matrices stored in FP32
             (FP64 arithmetic); hierarchy setup inside the step.  ``value``.
  amg_mixed_l0fp16  the same with the level-0 smoother matrix stored as FP16
             D^-1 A (option level0='fp16'; reported, not the value).
  amg_fp64   the same with an FP64 hierarchy.
  mixed      Jacobi-PCG ('jax_solver', 'mixed'): FP32 matrix and PCG, FP64
             iterate and FP64 residual replacement.
  fp64       Jacobi-PCG in FP64: the plain FP64 path.
An FP64 Jacobi-PCG solve to rtol 1e-11 is the reference for the relative
errors.

``--workload`` selects one of the other configurations of the same path
instead (one JSON line each, same keys; not the default line):
  poisson      3D Poisson, HEX8 32^3 on the unit cube, f ~ N(0,1) per node
               (seed 0) as the source, u = 0 on x = 0; FP64 and mixed.
  neo_hookean  Neo-Hookean cube, HEX8 64^3, E = 10, nu = 0.3, bottom fixed,
               top u_z = 0.1 k/5 in 5 load steps (u_x, u_y free), Newton to
               |R| <= 1e-8 from the previous step's solution.
  simp         SIMP compliance gradient, HEX8 160x80x40 on 2x1x0.5, theta =
               0.5 + U(-0.05, 0.05) per cell (seed 0), E = 1e-3 + theta^3
               (1 - 1e-3), clamped x = 0, traction (0, 0, -1) on the 2x2
               faces at x = 2, y-centred, z at the bottom; adjoint vs central
               differences on 16 cells (seed 1).
  hex27        HEX27 linear elasticity, 48^3 cells on the unit cube, 27-point
               Gauss rule, E = exp(N(0, 0.25^2)) per cell (seed 0), nu = 0.3,
               clamped z = 0, body force (0, 0, -1); compliance gradient
               w.r.t. the E field.
  explicit     the residual-only path: central-difference explicit steps
               (one compute_residual each) on the default mesh, lumped mass.
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


TRAFFIC_PROFILE = os.path.join(ROOT, 'profiles', 'k_pcg_a_hbm_traffic_cantilever_128x64x64.md')
AMG_TRAFFIC_PROFILE = os.path.join(ROOT, 'profiles', 'k_sell_apply_cheb_hbm_traffic_cantilever_128x64x64.md')
F16_TRAFFIC_PROFILE = os.path.join(ROOT, 'profiles', 'k_sell_apply_cheb_f16_cantilever_128x64x64.md')


def measured_traffic(cells, path=TRAFFIC_PROFILE):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by a streaming
    kernel of known size) on the 128x64x64 workload, from the committed summary; {} otherwise."""
    if list(cells) != [128, 64, 64] or not os.path.exists(path):
        return {}
    for line in open(path):
        if line.startswith('traffic_bytes_per_launch:'):
            return {k: int(v) for k, v in (kv.split('=') for kv in line.split(':', 1)[1].split())}
    return {}


def cheb_bytes(dev, elem_bytes, vec=3):
    """Mean algorithmic bytes of the three timed level-0 Chebyshev steps of a V-cycle
    (k_sell_apply, CHEB; level-0 degree 2: the pre-smoothing step after k_cheb_first and both
    post-smoothing steps): the matrix blocks (values + one column index each, padding not
    counted) once; the FP64 vectors x (gathered, counted once), rhs and D^-1 read, d and the
    new x written, and d read by the two steps that continue a direction (the first
    post-smoothing step starts it from zero)."""
    return dev.nnzb * (vec * vec * elem_bytes + 4) + dev.n * 8 * (5 + 6 + 6) / 3


def assembly_bytes(dev):
    """Algorithmic bytes of one FP64 single-pass tangent assembly: the SELL values (padding not
    counted) and the inverse diagonal written once; connectivity, coordinates, the per-quad
    modulus field and the plan's index data (slice-local contribution ids, slot bounds) read once."""
    dm = dev.dm
    return (dev.nnzb * 72 + dev.n * 8 + dm.C * dm.N * 4 + dm.Nn * 24 + dm.C * dm.Q * 8 + dm.C * dm.N * dm.N * 4
            + (dev.nnzb + 1) * 4)


def sell_bytes(dev, elem_bytes, vec=3):
    """Algorithmic bytes of one SpMV launch of the PCG iteration: the matrix blocks (values +
    one column index each, padding not counted) once, p read once (gathers counted once),
    q written."""
    n = dev.n
    return dev.nnzb * (vec * vec * elem_bytes + 4) + n * elem_bytes * 2


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

    # timed steps run without per-launch events; one extra step with events gives the kernel averages.
    # Paths: smoothed-aggregation AMG-CG (amgx_solver) and Jacobi-PCG (jax_solver), each with the
    # FP32-storage ("mixed") and the FP64 variant; the AMG mixed path is the headline value.
    opts = {
        'amg_mixed': {'amgx_solver': {'precision': 'mixed'}},
        'amg_mixed_l0fp16': {'amgx_solver': {'precision': 'mixed', 'level0': 'fp16'}},
        'amg_fp64': {'amgx_solver': {'precision': 'fp64'}},
        'mixed': {'jax_solver': {'precision': 'mixed', 'assembly': args.mixed_assembly}},
        'fp64': {'jax_solver': {'precision': 'fp64'}},
    }
    event_opts = {k: _evented(v) for k, v in opts.items()}
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
    paths = tuple(opts) if args.only is None else (args.only,)
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
        solver(p, event_opts[name])
        lin = p._solve_stats['linear'][0]
        res[name] = dict(s_per_newton_step=statistics.median(times), times=times, rel_err=err,
                         newton_iterations=stats[-1]['newton_iterations'], iterations=lin['iterations'],
                         replacements=lin['replacements'], spmv_ms=lin['spmv_ms'], update_ms=lin['update_ms'],
                         timing=stats[-1]['timing'])
        tm = stats[-1]['timing']
        if name.startswith('amg') and getattr(dev, 'amg_timing', None):
            res[name]['amg_timing'] = dict(dev.amg_timing)
            log(f"{name}: AMG setup {dev.amg_timing['setup_s']*1e3:.1f} ms, solve {dev.amg_timing['solve_s']*1e3:.1f} "
                f"ms, levels {dev._amg.describe()}")
        log(f"{name}: {res[name]['s_per_newton_step']:.4f} s/step, rel err {err:.2e}, "
            f"{lin['iterations']} its, {lin['replacements']} repl, kernel {lin['spmv_ms']*1e3:.1f} us, "
            f"rest {lin['update_ms']*1e3:.1f} us; residuals {tm['local_assembly']*1e3:.1f} ms, "
            f"tangent {tm['global_matrix']*1e3:.1f} ms, linear {tm['linear']*1e3:.1f} ms")

    if args.only is not None:
        if args.calib_copy_mib:
            copy_calibration(args.calib_copy_mib)
        print(json.dumps({name: {k: v for k, v in res[name].items()} for name in paths}), flush=True)
        return
    # adjoint gradient of the compliance C = f.u w.r.t. theta
    f = -dev.load
    grads = {}
    for name in opts:
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
        res[name]['s_per_adjoint_gradient'] = statistics.median(tms)
        res[name]['adjoint_iterations'] = p._adjoint_stats['iterations']
        grads[name] = th.grad.detach().clone()
        log(f"{name}: adjoint {res[name]['s_per_adjoint_gradient']:.4f} s, its {p._adjoint_stats['iterations']}")
    # tangent assembly, FP64 standard form: the single-pass kernel against the two-pass path
    from jax_fem_amd import _lib
    u0 = torch.zeros(dev.n, dtype=torch.float64, device='cuda')

    def asm_ms(fused):
        for _ in range(2):
            dev.tangent(u0, _lib.FEM_F64, fused=fused)
        ts = []
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            dev.tangent(u0, _lib.FEM_F64, fused=fused)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return statistics.median(ts)
    asm = dict(single_pass_ms=asm_ms(True) if dev.fused_applies(_lib.FEM_F64) else None, two_pass_ms=asm_ms(False),
               algorithmic_bytes=assembly_bytes(dev))
    if asm['single_pass_ms']:
        asm['achieved_GBs'] = asm['algorithmic_bytes'] / (asm['single_pass_ms'] * 1e-3) / 1e9
        asm['frac'] = asm['achieved_GBs'] / PEAK_HBM_GBS
    log(f"assembly (FP64 tangent + Dirichlet rows): single pass {asm['single_pass_ms']} ms, two pass "
        f"{asm['two_pass_ms']:.3f} ms")
    # the reference's default Krylov method (BiCGSTAB + Jacobi, FP64) on the same system: its
    # iteration count sizes the CPU baseline's solve
    solver(p, {'jax_solver': {'precision': 'fp64', 'krylov': 'bicgstab'}})
    t = time.perf_counter()
    solver(p, {'jax_solver': {'precision': 'fp64', 'krylov': 'bicgstab'}})
    torch.cuda.synchronize()
    bicg = dict(s_per_newton_step=time.perf_counter() - t, iterations=p._solve_stats['linear'][0]['iterations'])
    log(f"bicgstab fp64: {bicg['s_per_newton_step']:.4f} s/step, {bicg['iterations']} its")
    gref = grads['fp64']
    gerr = {k: float(torch.linalg.vector_norm(g - gref) / torch.linalg.vector_norm(gref)) for k, g in grads.items()}

    head = res['amg_mixed']
    amg_traffic = measured_traffic(args.cells, AMG_TRAFFIC_PROFILE)
    traffic = measured_traffic(args.cells)
    cb = cheb_bytes(dev, 4)
    achieved = cb / (head['spmv_ms'] * 1e-3) / 1e9 if head['spmv_ms'] else None
    spmv_bytes = sell_bytes(dev, 4)
    pcg_achieved = spmv_bytes / (res['mixed']['spmv_ms'] * 1e-3) / 1e9 if res['mixed']['spmv_ms'] else None
    fp64_bytes = sell_bytes(dev, 8)
    fp64_achieved = fp64_bytes / (res['fp64']['spmv_ms'] * 1e-3) / 1e9 if res['fp64']['spmv_ms'] else None
    out = {
        "metric": METRIC, "value": head['s_per_newton_step'], "unit": "s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head['s_per_newton_step'] * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"HEX8 linear-elasticity cantilever 2x1x1, {args.cells[0]}x{args.cells[1]}x"
                               f"{args.cells[2]} cells, {dev.n} dofs, mixed precision (FP32 storage, FP64 "
                               f"arithmetic and outer CG), SA-AMG-CG rtol 1e-10",
                   "global_batch": 1, "seq_len": dev.n, "parallelism": "single GPU"},
        "fp64_s_per_newton_step": res['fp64']['s_per_newton_step'],
        "speedup_mixed_vs_fp64": res['fp64']['s_per_newton_step'] / head['s_per_newton_step'],
        "speedup_note": "value path (AMG-CG, FP32 hierarchy) vs the plain FP64 path (Jacobi-PCG, FP64): "
                        "mostly the change of algorithm; speedup_like_for_like isolates the precision",
        "speedup_like_for_like": {
            "amg_fp64_over_amg_mixed": res['amg_fp64']['s_per_newton_step'] / head['s_per_newton_step'],
            "jacobi_fp64_over_jacobi_mixed": res['fp64']['s_per_newton_step'] / res['mixed']['s_per_newton_step']},
        "paths_s_per_newton_step": {k: v['s_per_newton_step'] for k, v in res.items()},
        "paths_s_per_adjoint_gradient": {k: v['s_per_adjoint_gradient'] for k, v in res.items()},
        "paths_iterations": {k: v['iterations'] for k, v in res.items()},
        "bicgstab_fp64": bicg,
        "assembly": dict(asm, kernel="k_assemble_hex8<MatElastic<true>, double> (single-pass HEX8 tangent into the "
                                     "FP64 SELL values, Jacobi diagonal) + k_bc_rows, timed together with HIP events",
                         bytes_formula="nnzb*72 (values) + n*8 (D^-1) + C*32 (cells) + Nn*24 (points) + C*Q*8 "
                                       "(modulus field) + C*64*4 (slice-local contribution ids) + (nnzb+1)*4 "
                                       "(slot bounds)"),
        "rel_err_vs_fp64_ref": {k: v['rel_err'] for k, v in res.items()},
        "adjoint_grad_rel_diff_vs_fp64": gerr,
        "s_per_adjoint_gradient": {"mixed": head['s_per_adjoint_gradient'],
                                   "fp64": res['fp64']['s_per_adjoint_gradient']},
        "roofline": {"bound": "hbm",
                     "kernel": "k_sell_apply<3,CHEB,float,true> (level-0 Chebyshev steps of the V-cycle, the "
                               "mean of its three launches per cycle: SELL-64 3x3-block product on FP32 "
                               "difference-form storage with FP64 vectors and arithmetic, fused D^-1 "
                               "residual and update)",
                     "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS if achieved else None,
                     "traffic": amg_traffic.get('k_sell_apply_cheb_f32'),
                     "traffic_source": os.path.relpath(AMG_TRAFFIC_PROFILE, ROOT) if amg_traffic else None,
                     "algorithmic_bytes": cb, "avg_launch_ms": head['spmv_ms']},
        "level0_fp16": {"kernel": "k_sell_apply<3,CHEB,_Float16,true> (level-0 Chebyshev step on FP16 D^-1 A)",
                        "s_per_newton_step": res['amg_mixed_l0fp16']['s_per_newton_step'],
                        "iterations": res['amg_mixed_l0fp16']['iterations'],
                        "avg_launch_ms": res['amg_mixed_l0fp16']['spmv_ms'],
                        "algorithmic_bytes": cheb_bytes(dev, 2),
                        "traffic": measured_traffic(args.cells, F16_TRAFFIC_PROFILE).get('k_sell_apply_cheb_f16'),
                        "achieved_GBs": (cheb_bytes(dev, 2) / (res['amg_mixed_l0fp16']['spmv_ms'] * 1e-3) / 1e9
                                         if res['amg_mixed_l0fp16']['spmv_ms'] else None)},
        "pcg_spmv": {"kernel": "k_pcg_a<3,true,float> (Jacobi-PCG SpMV q = Ap, FP32 storage)",
                     "achieved_GBs": pcg_achieved, "frac": pcg_achieved / PEAK_HBM_GBS if pcg_achieved else None,
                     "avg_launch_ms": res['mixed']['spmv_ms'], "algorithmic_bytes": spmv_bytes,
                     "traffic": traffic.get('k_pcg_a_f32')},
        "fp64_spmv": {"kernel": "k_pcg_a<3,false,double>", "achieved_GBs": fp64_achieved,
                      "frac": fp64_achieved / PEAK_HBM_GBS if fp64_achieved else None,
                      "avg_launch_ms": res['fp64']['spmv_ms'], "algorithmic_bytes": fp64_bytes,
                      "traffic": traffic.get('k_pcg_a_f64')},
        "detail": res,
    }
    if args.cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, bicg['iterations'], res['fp64']['iterations'])
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


def _cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, bicgstab_its, pcg_its, reps=3, block=20):
    """The multi-threaded C restatement (oracle/cpu_fem.c) on the full workload, FP64, all
    threads OpenMP is given: the tangent assembly (COO-equivalent summation into CSR + zeroRows)
    timed whole, and blocks of `block` iterations of the reference's default solver (BiCGSTAB
    with Jacobi, solver.py:63-92) and of Jacobi-PCG on the full matrix, each the median of `reps`
    runs after one warm-up.  A Newton step = assembly + (iterations of the solve, counted by
    the GPU's FP64 run of the same algorithm on the same system) x time per iteration + one
    residual product."""
    import numpy as np
    from oracle import cpu as OC
    from oracle import fem as OF
    fe = p.fes[0]
    pts, cells = fe.points, fe.cells
    lm, mu = OF.lame(70e3, 0.3)
    left = np.flatnonzero(np.isclose(pts[:, 0], 0., atol=1e-5))
    bc = (3 * left[:, None] + np.arange(3)[None, :]).reshape(-1)
    b = -p.device().load.cpu().numpy()
    b[bc] = 0.
    A = OC.elastic_pattern(pts, cells)

    def med(fn):
        fn()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return statistics.median(ts)
    t_asm = med(lambda: OC.assemble_elastic(A, pts, cells, lm, mu, None, bc))
    x0 = np.zeros(A.n)
    y = np.empty(A.n)
    t_bi = med(lambda: OC.solve(A, b, x0, 'bicgstab', fixed=block)) / block
    t_pcg = med(lambda: OC.solve(A, b, x0, 'pcg', fixed=block)) / block
    L = OC.lib()
    t_res = med(lambda: L.cfem_spmv(A.n, OC._p(A.rowptr), OC._p(A.col), OC._p(A.vals), OC._p(b), OC._p(y)))
    threads = OC.threads()
    value = t_asm + bicgstab_its * t_bi + t_res
    return {"value": value, "unit": "s per Newton step", "cores": threads, "kind": "port",
            "sample": f"full workload ({len(cells)} cells, {A.vals.size} nnz), C restatement (oracle/cpu_fem.c, "
                      f"FP64, OpenMP {threads} threads, {_cpu_model()}): tangent assembly {t_asm:.3f} s, "
                      f"BiCGSTAB+Jacobi (the reference default) {t_bi * 1e3:.1f} ms per iteration x {bicgstab_its} "
                      f"iterations, residual product {t_res * 1e3:.1f} ms; each the median of {reps} runs "
                      f"({block}-iteration blocks)",
            "assembly_s": t_asm, "bicgstab_iteration_s": t_bi, "pcg_iteration_s": t_pcg, "residual_s": t_res,
            "iterations": {"bicgstab": bicgstab_its, "pcg": pcg_its},
            "pcg_value": t_asm + pcg_its * t_pcg + t_res,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "repeats": reps}


# ---------------------------------------------------------------- other workloads
def _emit(workload, value_s, dtype, extra):
    out = {"metric": METRIC, "value": value_s, "unit": "s", "n_gpus": 1, "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
           "config": {"workload": workload, "global_batch": 1, "parallelism": "single GPU"}}
    out.update(extra)
    print(json.dumps(out), flush=True)


def _rel(a, b):
    import torch
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def _timed_solve(p, opts, **kw):
    import torch
    from jax_fem_amd.solver import solver
    torch.cuda.synchronize()
    t = time.perf_counter()
    sol = solver(p, dict(opts, **kw))[0]
    torch.cuda.synchronize()
    return time.perf_counter() - t, sol.clone(), p._solve_stats


def _paths(args):
    if args.solver == 'amg':
        r = args.setup_reuse
        return {'fp64': {'amgx_solver': {'precision': 'fp64', 'setup_reuse': r}},
                'mixed': {'amgx_solver': {'precision': 'mixed', 'setup_reuse': r, 'outer': args.amg_outer,
                                          'level0': args.amg_level0}}}
    return {'fp64': {'jax_solver': {'precision': 'fp64'}},
            'mixed': {'jax_solver': {'precision': 'mixed', 'assembly': args.mixed_assembly}}}


def _solver_name(args):
    return 'SA-AMG-CG' if args.solver == 'amg' else 'Jacobi-PCG'


def _opt(o, **kw):
    """Solver options `o` with keys of its linear-solver dict replaced."""
    k = next(iter(o))
    return {k: dict(o[k], **kw)}


def _evented(o):
    """Options of the extra solve that times the dominant kernels with HIP events."""
    return _opt(o, timing=True)


def workload_poisson(args):
    """3D Poisson -lap u = f, HEX8 32^3, f nodal N(0,1) interpolated at the quadrature points."""
    import numpy as np
    import torch
    from jax_fem_amd.generate_mesh import Mesh, box_mesh
    from jax_fem_amd.problem import Problem

    class PoissonSource(Problem):
        def get_tensor_map(self):
            return lambda u_grad, f: u_grad

        def get_mass_map(self):
            return lambda u, x, f: -f.reshape(1)

    torch.cuda.set_device(0)
    n = 32
    m = box_mesh(n, n, n, 1., 1., 1.)
    mesh = Mesh(m.points, m.cells_dict['hexahedron'])
    left = lambda x: np.isclose(x[0], 0., atol=1e-5)
    p = PoissonSource(mesh, vec=1, dim=3, dirichlet_bc_info=[[left], [0], [lambda x: 0.]])
    f_node = np.random.default_rng(0).standard_normal(len(m.points))
    fe = p.fes[0]
    f_q = torch.as_tensor(np.asarray(fe.shape_vals) @ f_node[np.asarray(fe.cells)].T).T.contiguous()
    p.internal_vars = [f_q]
    ref = _timed_solve(p, {'jax_solver': {'precision': 'fp64', 'tol': 1e-12, 'atol': 0.}})[1]
    res = {}
    for name, o in _paths(args).items():
        for _ in range(args.warmup):
            _timed_solve(p, o)
        runs = [_timed_solve(p, o) for _ in range(args.steps)]
        res[name] = dict(s_per_newton_step=statistics.median(r[0] for r in runs), rel_err=_rel(runs[-1][1], ref),
                         pcg_iterations=runs[-1][2]['linear'][0]['iterations'], sol=runs[-1][1])
    res['mixed']['rel_err_vs_fp64'] = _rel(res['mixed']['sol'], res['fp64']['sol'])
    for name in res:
        del res[name]['sol']
        log(f"poisson {name}: {res[name]}")
    extra = {"fp64_s_per_newton_step": res['fp64']['s_per_newton_step'], "detail": res}
    if args.cpu_baseline:
        from oracle import fem as OF
        o = OF.Problem(m.points, m.cells_dict['hexahedron'], 1, mat=OF.Material('poisson', k=1.),
                       bc_info=[[left], [0], [lambda x: 0.]])
        o.add_mass(-np.einsum('qa,ca->cq', o.sv, f_node[o.cells])[..., None])
        t = time.perf_counter()
        u = o.newton(linear='pcg')
        dt = time.perf_counter() - t
        extra["cpu_baseline"] = {"value": dt, "unit": "s per Newton step", "cores": 1, "kind": "port",
                                 "sample": f"full workload: oracle FP64 assembly + Jacobi-PCG ({o.newton_its} its)",
                                 "rel_err_vs_gpu_fp64": float(np.linalg.norm(u - ref.cpu().numpy().reshape(-1)) /
                                                              np.linalg.norm(u))}
    _emit(f"HEX8 Poisson 32^3 unit cube, {p.fes[0].num_total_nodes} dofs, {_solver_name(args)} rtol 1e-10",
          res['mixed']['s_per_newton_step'], "f32+f64", extra)


def workload_neo_hookean(args):
    """Neo-Hookean cube 64^3, top face pulled to u_z = 0.1 in 5 load steps."""
    import numpy as np
    import torch
    from jax_fem_amd.generate_mesh import Mesh, box_mesh
    from jax_fem_amd.problem import Problem

    E, nu = 10., 0.3
    mu, kappa = E / (2. * (1. + nu)), E / (3. * (1. - 2. * nu))

    class NeoHookeanCube(Problem):
        # tensor map of the reference's hyperelasticity example (P = d psi / dF)
        def get_tensor_map(self):
            if args.tabulated:
                # the same law as a tabulated map (P written out: torch's second derivatives
                # of the energy through linalg.det are NaN at F = I)
                from jax_fem_amd import materials as M
                return M.Tabulated(M.NeoHookean(mu=mu, kappa=kappa))

            def psi(F):
                J = torch.linalg.det(F)
                I1 = torch.trace(F.T @ F)
                return mu / 2. * (J ** (-2. / 3.) * I1 - 3.) + kappa / 2. * (J - 1.) ** 2
            P = torch.func.grad(psi)
            return lambda u_grad: P(u_grad + torch.eye(3, dtype=u_grad.dtype, device=u_grad.device))

    torch.cuda.set_device(0)
    n = args.nh_cells
    m = box_mesh(n, n, n, 1., 1., 1.)
    mesh = Mesh(m.points, m.cells_dict['hexahedron'])
    bottom = lambda x: np.isclose(x[2], 0., atol=1e-5)
    top = lambda x: np.isclose(x[2], 1., atol=1e-5)

    def bc(k):
        disp = 0.1 * k / 5
        return [[bottom] * 3 + [top], [0, 1, 2, 2], [lambda x: 0.] * 3 + [lambda x: disp]]

    p = NeoHookeanCube(mesh, vec=3, dim=3, dirichlet_bc_info=bc(1))
    res, sols = {}, {}
    for name, o in _paths(args).items():
        o = dict(o, tol=1e-8, rel_tol=1e-14)
        for _ in range(args.warmup):     # untimed: one Newton solve of the first load step
            p.fes[0].update_Dirichlet_boundary_conditions(bc(1))
            _timed_solve(p, o)
        p.device().drop_amg_hierarchy()  # the timed sequence builds its own AMG hierarchies
        tot_t, tot_newton, tot_pcg = 0., 0, 0
        phases = {'local_assembly': 0., 'global_matrix': 0., 'linear': 0.}
        sol = None
        for k in range(1, 6):
            p.fes[0].update_Dirichlet_boundary_conditions(bc(k))
            kw = {} if sol is None else {'initial_guess': [sol]}
            dt, sol, st = _timed_solve(p, o, **kw)
            tot_t += dt
            tot_newton += st['newton_iterations']
            tot_pcg += sum(li['iterations'] for li in st['linear'])
            for key in phases:
                phases[key] += st['timing'][key]
            log(f"neo_hookean {name} step {k}: {st['newton_iterations']} Newton its, |R| = {st['res_norm']:.2e}, "
                f"{dt:.3f} s")
        sols[name] = sol
        res[name] = dict(s_total=tot_t, newton_iterations=tot_newton, s_per_newton_step=tot_t / tot_newton,
                         pcg_iterations=tot_pcg, final_residual=st['res_norm'],
                         phases_s={k: round(v, 4) for k, v in phases.items()})
    res['mixed']['rel_err_vs_fp64'] = _rel(sols['mixed'], sols['fp64'])
    log(f"neo_hookean: {res}")
    extra = {"fp64_s_per_newton_step": res['fp64']['s_per_newton_step'], "detail": res}
    if args.cpu_baseline:
        from oracle import fem as OF
        ns = 16
        ms = box_mesh(ns, ns, ns, 1., 1., 1.)
        o = OF.Problem(ms.points, ms.cells_dict['hexahedron'], 3, mat=OF.Material('neo_hookean', mu=mu, kappa=kappa))
        u = 0.01 * np.random.default_rng(0).standard_normal(o.n)
        extra["cpu_baseline"] = _cpu_port_sample(o, p.num_cells, p.device().nnzb * 9,
                                                 res['fp64']['pcg_iterations'] / res['fp64']['newton_iterations'],
                                                 "s per Newton step", u)
    _emit(f"HEX8 Neo-Hookean cube {n}^3, {p.fes[0].num_total_nodes * 3} dofs, 5 load steps, Newton |R| <= 1e-8, "
          f"{_solver_name(args)}" + (", tabulated law" if args.tabulated else ""),
          res['mixed']['s_per_newton_step'], "f32+f64", extra)


def _compliance_var(p, f, u):
    """Variational compliance 2 f.u - u.K u with u.K u summed from positive cell energies: its error is
    second order in the solve error, so central differences of it resolve tiny parameter steps."""
    import torch
    dev = p.device()
    dev.prepare()
    re = dev.cell_internal_force(u)                        # (C, N, V): K_e u_e
    ue = u.reshape(-1, dev.V)[dev.dm.cells.long()]
    return float(2. * torch.dot(f, u) - (re * ue).sum())


def workload_simp(args):
    """SIMP compliance gradient on the 160x80x40 cantilever; adjoint vs central differences."""
    import numpy as np
    import torch
    from jax_fem_amd.generate_mesh import Mesh, box_mesh
    from jax_fem_amd.problem import Problem
    from jax_fem_amd.solver import ad_wrapper

    nu, Emin, E0, penal = 0.3, 1e-3, 1.0, 3.0

    class SIMP(Problem):
        def get_tensor_map(self):
            def stress(u_grad, theta):
                E = Emin + (E0 - Emin) * theta[0] ** penal
                mu_, lm = E / (2. * (1. + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
                eps = 0.5 * (u_grad + u_grad.T)
                return lm * torch.trace(eps) * torch.eye(3, dtype=u_grad.dtype) + 2 * mu_ * eps
            return stress

        def get_surface_maps(self):
            return [lambda u, x: torch.tensor([0., 0., 1.], dtype=torch.float64)]

        def set_params(self, params):
            self.internal_vars = [params.reshape(-1, 1)[:, None, :].expand(-1, self.fes[0].num_quads, 1)]

    torch.cuda.set_device(0)
    Nx, Ny, Nz = args.simp_cells
    Lx, Ly, Lz = 2., 1., 0.5
    m = box_mesh(Nx, Ny, Nz, Lx, Ly, Lz)
    mesh = Mesh(m.points, m.cells_dict['hexahedron'])
    hy, hz = Ly / Ny, Lz / Nz
    left = lambda x: np.isclose(x[0], 0., atol=1e-5)
    patch = lambda x: np.isclose(x[0], Lx, atol=1e-5) & (np.abs(x[1] - Ly / 2) <= hy + 1e-6) & (x[2] <= 2 * hz + 1e-6)
    p = SIMP(mesh, vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda x: 0.] * 3], location_fns=[patch])
    assert len(p.boundary_inds_list[0]) == 4, len(p.boundary_inds_list[0])
    C = p.num_cells
    theta0 = torch.as_tensor(0.5 + np.random.default_rng(0).uniform(-0.05, 0.05, C), device='cuda')
    res, grads = {}, {}
    for name, o in _paths(args).items():
        fwd = ad_wrapper(p, o, o)
        fw, bw = [], []
        for it in range(args.warmup + args.steps):
            th = theta0.clone().requires_grad_(True)
            torch.cuda.synchronize()
            t = time.perf_counter()
            u = fwd(th)[0]
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            Cval = torch.dot(-p.device().load, u.reshape(-1))
            Cval.backward()
            torch.cuda.synchronize()
            if it >= args.warmup:
                fw.append(t1 - t)
                bw.append(time.perf_counter() - t1)
        grads[name] = th.grad.detach().clone()
        res[name] = dict(s_per_newton_step=statistics.median(fw), s_per_adjoint_gradient=statistics.median(bw),
                         compliance=float(Cval), adjoint_pcg_iterations=p._adjoint_stats['iterations'])
        log(f"simp {name}: {res[name]}")
    res['mixed']['grad_rel_diff_vs_fp64'] = _rel(grads['mixed'], grads['fp64'])
    # central differences of the variational compliance on 16 seeded cells (FP64 solves to rtol 1e-11)
    cells = np.random.default_rng(1).choice(C, 16, replace=False)
    h = args.fd_h
    fopt = {'jax_solver': {'precision': 'fp64', 'tol': 1e-11, 'atol': 0.}}
    f = -p.device().load
    fd = []
    for e in cells:
        vals = []
        for sgn in (1., -1.):
            th = theta0.clone()
            th[e] += sgn * h
            p.set_params(th)
            u = _timed_solve(p, fopt)[1].reshape(-1)
            vals.append(_compliance_var(p, f, u))
        fd.append((vals[0] - vals[1]) / (2 * h))
    fd = np.array(fd)
    adj = grads['fp64'][torch.as_tensor(cells, device='cuda')].cpu().numpy()
    fd_err = float(np.max(np.abs(adj - fd) / np.abs(fd)))
    log(f"simp FD (h={h}) vs adjoint on 16 cells: max rel diff {fd_err:.2e}")
    cpu = None
    if args.cpu_baseline:
        from oracle import fem as OF
        ms = box_mesh(Nx // 4, Ny // 4, Nz // 4, Lx, Ly, Lz)
        cs = ms.cells_dict['hexahedron']
        th = 0.5 + np.random.default_rng(0).uniform(-0.05, 0.05, len(cs))
        o = OF.Problem(ms.points, cs, 3, theta=np.repeat(th[:, None], 8, 1),
                       mat=OF.Material('simp', **dict(zip(('lmbda1', 'mu1'), OF.lame(1., nu))), Emin=Emin, E0=E0,
                                       p=penal))
        cpu = _cpu_port_sample(o, C, p.device().nnzb * 9, res['fp64']['adjoint_pcg_iterations'],
                               "s per adjoint gradient")
    _emit(f"HEX8 SIMP cantilever {Nx}x{Ny}x{Nz} on {Lx}x{Ly}x{Lz}, {p.fes[0].num_total_nodes * 3} dofs, "
          f"compliance gradient, {_solver_name(args)}", res['mixed']['s_per_adjoint_gradient'], "f32+f64",
          dict({"fp64_s_per_adjoint_gradient": res['fp64']['s_per_adjoint_gradient'],
                "fd_check": {"cells": cells.tolist(), "h": h, "max_rel_diff": fd_err, "fd": fd.tolist(),
                             "adjoint": adj.tolist()}, "detail": res}, **({"cpu_baseline": cpu} if cpu else {})))


def workload_hex27(args):
    """HEX27 elasticity 48^3 with a log-normal per-cell Young's modulus; gradient of compliance w.r.t. E."""
    import numpy as np
    import torch
    from jax_fem_amd.generate_mesh import Mesh, box_mesh_hex27
    from jax_fem_amd.problem import Problem
    from jax_fem_amd.solver import ad_wrapper

    nu = 0.3

    class Hex27Elastic(Problem):
        def get_tensor_map(self):
            def stress(u_grad, E):
                mu_, lm = E / (2. * (1. + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
                eps = 0.5 * (u_grad + u_grad.T)
                return lm * torch.trace(eps) * torch.eye(3, dtype=u_grad.dtype) + 2 * mu_ * eps
            return stress

        def get_mass_map(self):
            return lambda u, x, E: torch.tensor([0., 0., 1.], dtype=torch.float64)

        def set_params(self, params):
            self.internal_vars = [params.reshape(-1, 1).expand(-1, self.fes[0].num_quads)]

    torch.cuda.set_device(0)
    n = args.hex27_cells
    m = box_mesh_hex27(n, n, n, 1., 1., 1.)
    mesh = Mesh(m.points, m.cells_dict['hexahedron27'])
    bottom = lambda x: np.isclose(x[2], 0., atol=1e-5)
    t0 = time.perf_counter()
    p = Hex27Elastic(mesh, vec=3, dim=3, ele_type='HEX27', quadrature_order=4,
                     dirichlet_bc_info=[[bottom] * 3, [0, 1, 2], [lambda x: 0.] * 3])
    C = p.num_cells
    E = torch.as_tensor(np.exp(np.random.default_rng(0).normal(0., 0.25, C)), device='cuda')
    p.set_params(E)
    dev = p.device()
    dev.prepare()
    log(f"hex27 setup {time.perf_counter() - t0:.1f}s: {C} cells, {dev.n} dofs, {dev.nnzb * 9} nnz, "
        f"{dev.sell_blocks} SELL blocks ({dev.sell_blocks / dev.nnzb - 1:.1%} padding)")
    res, grads, sols = {}, {}, {}
    for name, o in _paths(args).items():
        # rtol 1e-11: at 1e-10 either solution is already ~1e-8 from the exact one on this mesh
        o = _opt(o, tol=1e-11, atol=0.)
        fwd = ad_wrapper(p, o, o)
        fw, bw = [], []
        for it in range(args.warmup + args.steps):
            e = E.clone().requires_grad_(True)
            torch.cuda.synchronize()
            t = time.perf_counter()
            u = fwd(e)[0]
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            Cval = torch.dot(-dev.load, u.reshape(-1))
            Cval.backward()
            torch.cuda.synchronize()
            if it >= args.warmup:
                fw.append(t1 - t)
                bw.append(time.perf_counter() - t1)
        grads[name] = e.grad.detach().clone()
        sols[name] = u.detach().clone()
        from jax_fem_amd.solver import solver
        solver(p, _evented(o))     # kernel averages (HIP events, Jacobi-PCG)
        lin = p._solve_stats['linear'][0]
        res[name] = dict(s_per_newton_step=statistics.median(fw), s_per_adjoint_gradient=statistics.median(bw),
                         pcg_iterations=lin['iterations'], spmv_ms=lin['spmv_ms'], update_ms=lin['update_ms'],
                         timing=p._solve_stats['timing'])
        log(f"hex27 {name}: {res[name]}")
    res['mixed']['rel_err_vs_fp64'] = _rel(sols['mixed'], sols['fp64'])
    res['mixed']['grad_rel_diff_vs_fp64'] = _rel(grads['mixed'], grads['fp64'])
    # compliance self-adjointness: dC/dE_e = -u_e^T K_e(1) u_e = -(energy of cell e) / E_e * 2 ... checked by sign
    assert bool((grads['fp64'] < 0).all()), "compliance gradient w.r.t. E must be negative"
    spmv_b = {nm: sell_bytes(dev, es) for nm, es in (('mixed', 4), ('fp64', 8))}
    cpu = None
    if args.cpu_baseline:
        from oracle import fem as OF
        ns = 6
        ms = box_mesh_hex27(ns, ns, ns, 1., 1., 1.)
        cs = ms.cells_dict['hexahedron27']
        Es = np.exp(np.random.default_rng(0).normal(0., 0.25, len(cs)))
        o = OF.Problem(ms.points, cs, 3, ele_type='HEX27', quadrature_order=4, theta=np.repeat(Es[:, None], 27, 1),
                       mat=OF.Material('simp', **dict(zip(('lmbda1', 'mu1'), OF.lame(1., nu))), Emin=0., E0=1., p=1.))
        cpu = _cpu_port_sample(o, C, dev.nnzb * 9, res['fp64']['pcg_iterations'], "s per Newton step")
    _emit(f"HEX27 elasticity {n}^3 unit cube, 27-pt Gauss, {dev.n} dofs, log-normal E field, {_solver_name(args)} rtol 1e-11",
          res['mixed']['s_per_newton_step'], "f32+f64",
          dict({"fp64_s_per_newton_step": res['fp64']['s_per_newton_step'],
                "spmv_GBs": {nm: spmv_b[nm] / (res[nm]['spmv_ms'] * 1e-3) / 1e9 for nm in spmv_b
                             if res[nm]['spmv_ms']},
                "detail": res}, **({"cpu_baseline": cpu} if cpu else {})))


def _cpu_port_sample(o, gpu_cells, gpu_nnz, gpu_iterations, what, u=None):
    """The oracle's FP64 tangent assembly and 50 Jacobi-PCG iterations on a small sample of the same
    problem family, on one core, scaled by the cell and nonzero ratios and the GPU's FP64 iteration count."""
    import numpy as np
    u = np.zeros(o.n) if u is None else u
    t = time.perf_counter()
    A = o.tangent(u)
    t_asm = time.perf_counter() - t
    b = np.random.default_rng(0).standard_normal(o.n)
    its = 50
    t = time.perf_counter()
    OF_pcg = __import__('oracle.fem', fromlist=['pcg_jacobi']).pcg_jacobi
    OF_pcg(A, b, np.zeros(o.n), A.diagonal(), tol=0., atol=0., maxiter=its)
    t_it = (time.perf_counter() - t) / its
    sc, sn = gpu_cells / len(o.cells), gpu_nnz / A.nnz
    return {"value": t_asm * sc + t_it * sn * gpu_iterations, "unit": what, "cores": 1, "kind": "port",
            "sample": f"oracle (numpy/scipy FP64) on {len(o.cells)} cells: tangent assembly {t_asm:.2f}s scaled by "
                      f"cells x{sc:.0f}, {its} Jacobi-PCG iterations ({t_it * 1e3:.1f} ms each) scaled by nnz "
                      f"x{sn:.0f} and by {gpu_iterations} FP64 GPU iterations"}


def workload_explicit(args):
    """The residual-only path (compute_residual, problem.py:462-475) under central-difference
    explicit steps, as in applications/explicit_dynamics/example.py:97-112, on the 128x64x64 HEX8
    cantilever mesh (2x1x1, E = 70e3, nu = 0.3, rho = 1, clamped x = 0, row-sum lumped mass),
    released from u_x = 1e-3 sin(pi x / 4).  A step is one compute_residual call and the update."""
    import numpy as np
    import torch
    from jax_fem_amd import materials
    from jax_fem_amd.generate_mesh import Mesh, box_mesh
    from jax_fem_amd.problem import Problem

    class Elastic(Problem):
        def get_tensor_map(self):
            return materials.LinearElastic(E=70e3, nu=0.3)

    nx, ny, nz = args.cells
    m = box_mesh(nx, ny, nz, 2., 1., 1.)
    left = lambda p: np.isclose(p[0], 0., atol=1e-5)
    p = Elastic(Mesh(m.points, m.cells_dict['hexahedron']), vec=3, dim=3,
                dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.] * 3])
    fe = p.fes[0]
    dev = p.device()
    dev.prepare()
    sv = torch.as_tensor(fe.shape_vals, device='cuda')
    cells = torch.as_tensor(fe.cells, device='cuda').long()
    em = torch.einsum('qi,eq->ei', sv, torch.as_tensor(fe.JxW, device='cuda'))
    nm = torch.zeros(fe.num_total_nodes, dtype=torch.float64, device='cuda').index_add_(0, cells.reshape(-1),
                                                                                          em.reshape(-1))
    mass = nm[:, None].repeat(1, 3)
    fixed = torch.zeros((fe.num_total_nodes, 3), dtype=torch.bool, device='cuda')
    fixed.view(-1)[dev.bc_dofs_long] = True
    lm, mu = 70e3 * 0.3 / (1.3 * 0.4), 70e3 / 2.6
    c = (lm + 2 * mu) ** 0.5
    dt = 0.5 * (1.0 / nz) / c
    x = torch.as_tensor(np.asarray(fe.points)[:, 0], device='cuda')
    u = torch.zeros((fe.num_total_nodes, 3), dtype=torch.float64, device='cuda')
    u[:, 0] = 1e-3 * torch.sin(np.pi * x / 4.)

    def acceleration(v):
        return torch.where(fixed, 0.0, -p.compute_residual([v])[0] / mass)

    up = torch.where(fixed, 0.0, u + 0.5 * dt ** 2 * acceleration(u))
    n_steps = max(args.steps, 1) * 20
    for it in range(args.warmup * 20 + n_steps):
        if it == args.warmup * 20:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
        un = torch.where(fixed, 0.0, 2.0 * u - up + dt ** 2 * acceleration(u))
        up, u = u, un
    torch.cuda.synchronize()
    per_step = (time.perf_counter() - t0) / n_steps
    energy = float((0.5 * mass * ((u - up) / dt) ** 2).sum())
    log(f"explicit: {per_step * 1e3:.3f} ms per step over {n_steps} steps, kinetic energy {energy:.3e}")
    _emit(f"HEX8 elasticity {nx}x{ny}x{nz} cells, {dev.n} dofs, explicit central differences "
          f"(compute_residual per step, FP64)", per_step, "f64",
          {"steps_timed": n_steps, "dt": dt, "kinetic_energy_end": energy})


WORKLOADS = {'poisson': workload_poisson, 'neo_hookean': workload_neo_hookean, 'simp': workload_simp,
             'hex27': workload_hex27, 'explicit': workload_explicit}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--workload', choices=['cantilever', 'poisson', 'neo_hookean', 'simp', 'hex27', 'explicit'],
                    default='cantilever')
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--cells', type=int, nargs=3, default=[128, 64, 64])
    ap.add_argument('--mixed-assembly', choices=['fp32', 'fp64', 'bf16x3'], default='fp64')
    ap.add_argument('--amg-level0', choices=['fp32', 'fp16'], default='fp32',
                    help="AMG mixed path: storage of the level-0 smoother matrix")
    ap.add_argument('--amg-outer', choices=['fp64', 'fp32'], default='fp64',
                    help="with --solver amg: operator of the mixed path's outer CG (fp32: difference form + "
                         "FP64 residual replacement)")
    ap.add_argument('--setup-reuse', type=int, default=0,
                    help="with --solver amg: solves that reuse one AMG hierarchy (Newton sequences)")
    ap.add_argument('--solver', choices=['pcg', 'amg'], default='pcg',
                    help="pcg: Jacobi-PCG ('jax_solver'); amg: smoothed-aggregation AMG-CG ('amgx_solver')")
    ap.add_argument('--no-cpu-baseline', dest='cpu_baseline', action='store_false')
    ap.add_argument('--only', choices=['fp64', 'mixed', 'amg_fp64', 'amg_mixed', 'amg_mixed_l0fp16'], default=None,
                    help='time one path only (no adjoint, no summary line); for profiling')
    ap.add_argument('--nh-cells', type=int, default=64)
    ap.add_argument('--tabulated', action='store_true',
                    help='neo_hookean: evaluate the law through the per-quad table (any tensor map)')
    ap.add_argument('--simp-cells', type=int, nargs=3, default=[160, 80, 40])
    ap.add_argument('--hex27-cells', type=int, default=48)
    ap.add_argument('--fd-h', type=float, default=1e-6)
    ap.add_argument('--calib-copy-mib', type=int, default=0,
                    help='with --only: also run a plain copy of this many MiB (counter calibration)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X (--gpus {args.gpus} requested); "
              "multi-GPU sharding is described in DESIGN.md but not implemented", file=sys.stderr)
        sys.exit(2)
    if args.workload == 'cantilever':
        run(args)
    else:
        WORKLOADS[args.workload](args)


if __name__ == '__main__':
    main()
