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
              self-adjointness (lambda = u); sign of dC/dtheta.
Parity unpinned: the order of basix's HEX27 quadrature points (only matters for
per-quad parameter layouts), the JAX BiCGSTAB iterate sequence, and jax-fem's
ad_wrapper output at the benchmark configurations (no reference test covers
ad_wrapper; the adjoint is pinned by finite differences and closed forms only).
"""
