"""FP64 CPU oracle of jax-fem's assembly -> linear solve -> adjoint path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference algorithm in
numpy/scipy so the device library can be checked against it.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it, and
only as the checker (or as the timed CPU baseline); the product path
(jax-fem_amd/) never imports it.

What each function follows in the reference, and what pins it, is stated in its
docstring.  Pins (tests/test_oracle.py, CPU):
  basis       Kronecker delta at nodes, partition of unity, Gauss exactness,
              HEX8 points 0.5 +- sqrt(3)/6, face areas / normals.
  geometry    affine cubes (JxW = h^3/8, grad N closed form), sum JxW = volume,
              cylinder top-face area = 76.53668647 (FEniCSx fixture).
  kernels     tangent vs central differences of the per-cell residual
              (problem.py:189-214) for all four materials; Neo-Hookean P vs
              torch autograd of psi (test_hyper_elasticity.py:17-26); SIMP
              dP/dtheta vs differences; rigid-body null space; patch test.
  pattern     closed-form nnz; values vs a brute-force loop assembly.
  solve       FEniCSx fixture solutions (Poisson, elasticity cube, elasticity
              cylinder, Neo-Hookean cylinder + top traction); PCG and BiCGSTAB
              vs a direct solve.
  adjoint     central finite differences of the forward solve; compliance
              self-adjointness (lambda = u); sign of dC/dtheta; the reference's
              own printed ad_wrapper gradient drho[0, 0] = 0.002266954599447443
              (docs/source/learn/compute_gradients/example.ipynb cell 31),
              reproduced to 2e-12 at cell 0, quadrature point 0
              (tests/test_oracle_gradient.py).
  variables   MultiProblem (several variables on one mesh, problem.py:58-119):
              free thermal expansion and the prescribed-expansion temperature in
              closed form, and the single-variable oracle when uncoupled
              (tests/test_oracle_multivar.py).
  cpu_fem.c   (the CPU baseline, oracle/cpu.py) against the numpy oracle: CSR
              pattern bit for bit, tangent values, BiCGSTAB / PCG iterates
              (tests/test_oracle_cpu.py).
Parity unpinned: the JAX BiCGSTAB iterate sequence, and jax-fem's ad_wrapper
output at the benchmark configurations other than the notebook's (the adjoint
there is pinned by finite differences and closed forms).  The quadrature-point
order is pinned at HEX8 by the notebook gradient; for HEX27 it only matters for
per-quad parameter layouts.
"""
