// Constitutive laws of the device path (the tensor_map of problem.py:189-214).
//
// Each law provides, per quadrature point:
//   prep     per-quad state from grad u (V x 3) and the per-quad parameter
//   stress   P(grad u)                                   (residual, problem.py:208-210)
//   tangent  K_ab += w * dP/dF contracted with grad N_a, grad N_b  (the jacfwd of
//            problem.py:262-266 written out in closed form)
//   jvp      dP = dP/dF : dG                             (matrix-free tangent product)
//   dparam   dP/dtheta for the per-quad parameter        (adjoint, solver.py:1386-1416)
// `cq` = cell * Q + quad indexes the per-quad parameter and state.
#pragma once
#include "fem_common.hpp"

namespace fem {

__device__ inline double param_at(const fem_material_t& m, int64_t cq) { return m.has_param ? m.param[cq] : 1.0; }

// P = k grad u  (Poisson / heat, tests/benchmarks/linear_poisson/test_linear_poisson.py:15-17)
struct MatPoisson {
  static constexpr int V = 1;
  static constexpr bool kNeedsU = false;
  template <typename T>
  struct St {
    T k;
  };
  template <typename T>
  __device__ static St<T> prep(const fem_material_t& m, int64_t cq, const T (&)[1][3]) {
    St<T> s;
    s.k = T(m.c[0]) * T(param_at(m, cq));
    return s;
  }
  template <typename T>
  __device__ static void stress(const St<T>& s, const T (&G)[1][3], T (&P)[1][3]) {
#pragma unroll
    for (int j = 0; j < 3; ++j) P[0][j] = s.k * G[0][j];
  }
  template <typename T>
  __device__ static void tangent(const St<T>& s, const T* ga, const T* gb, T w, T (&K)[1][1]) {
    K[0][0] += w * s.k * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
  }
  template <typename T>
  __device__ static void jvp(const St<T>& s, const T (&dG)[1][3], T (&dP)[1][3]) {
    stress(s, dG, dP);
  }
  template <typename T>
  __device__ static void dparam(const fem_material_t& m, int64_t, const St<T>&, const T (&G)[1][3],
                                T (&dP)[1][3]) {
#pragma unroll
    for (int j = 0; j < 3; ++j) dP[0][j] = m.has_param ? T(m.c[0]) * G[0][j] : T(0);
  }
};

// Small-strain isotropic elasticity, optionally with a SIMP-penalised modulus
// E(t) = Emin + (E0 - Emin) t^p (docs/source/learn/topology_optimization cell 9).
//   LINEAR_ELASTIC: lambda = c0, mu = c1
//   SIMP:           lambda = c0 E(t), mu = c1 E(t), E(t) = c2 + (c3 - c2) t^c4
template <bool kSimp>
struct MatElastic {
  static constexpr int V = 3;
  static constexpr bool kNeedsU = false;
  template <typename T>
  struct St {
    T lam, mu;
  };
  __device__ static double modulus(const fem_material_t& m, double t) {
    return m.c[2] + (m.c[3] - m.c[2]) * pow(t, m.c[4]);
  }
  __device__ static double dmodulus(const fem_material_t& m, double t) {
    return m.c[4] == 0.0 ? 0.0 : (m.c[3] - m.c[2]) * m.c[4] * pow(t, m.c[4] - 1.0);
  }
  template <typename T>
  __device__ static St<T> prep(const fem_material_t& m, int64_t cq, const T (&)[3][3]) {
    St<T> s;
    const double E = kSimp ? modulus(m, m.param[cq]) : 1.0;
    s.lam = T(m.c[0] * E);
    s.mu = T(m.c[1] * E);
    return s;
  }
  template <typename T>
  __device__ static void lin(T lam, T mu, const T (&G)[3][3], T (&P)[3][3]) {
    const T tr = G[0][0] + G[1][1] + G[2][2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) P[i][j] = mu * (G[i][j] + G[j][i]) + (i == j ? lam * tr : T(0));
  }
  template <typename T>
  __device__ static void stress(const St<T>& s, const T (&G)[3][3], T (&P)[3][3]) {
    lin(s.lam, s.mu, G, P);
  }
  template <typename T>
  __device__ static void tangent(const St<T>& s, const T* ga, const T* gb, T w, T (&K)[3][3]) {
    const T l = w * s.lam, m = w * s.mu;
    const T gg = m * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) K[i][k] += l * ga[i] * gb[k] + m * ga[k] * gb[i] + (i == k ? gg : T(0));
  }
  template <typename T>
  __device__ static void jvp(const St<T>& s, const T (&dG)[3][3], T (&dP)[3][3]) {
    lin(s.lam, s.mu, dG, dP);
  }
  template <typename T>
  __device__ static void dparam(const fem_material_t& m, int64_t cq, const St<T>&, const T (&G)[3][3],
                                T (&dP)[3][3]) {
    const double dE = kSimp ? dmodulus(m, m.param[cq]) : 0.0;
    lin(T(m.c[0] * dE), T(m.c[1] * dE), G, dP);
  }
};

// Compressible Neo-Hookean, psi = mu/2 (J^-2/3 I1 - 3) + kappa/2 (J - 1)^2
// (tests/benchmarks/hyperelasticity/test_hyper_elasticity.py:17-34),
// mu = c0 s, kappa = c1 s with s the per-quad parameter (or 1).
//   P = mu a (F - I1/3 H) + kappa J (J - 1) H,   a = J^-2/3, H = F^-T
struct MatNeoHookean {
  static constexpr int V = 3;
  static constexpr bool kNeedsU = true;
  template <typename T>
  struct St {
    T F[9], H[9];
    T J, I1, a, ma, k;  // a = J^-2/3, ma = mu * a, k = kappa
  };
  template <typename T>
  __device__ static St<T> prep(const fem_material_t& m, int64_t cq, const T (&G)[3][3]) {
    St<T> s;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) s.F[i * 3 + j] = G[i][j] + (i == j ? T(1) : T(0));
    s.J = det3(s.F);
    T Finv[9];
    inv3(s.F, s.J, Finv);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) s.H[i * 3 + j] = Finv[j * 3 + i];
    T I1 = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) I1 += s.F[i] * s.F[i];
    s.I1 = I1;
    const T sc = T(param_at(m, cq));
    s.a = pow(s.J, T(-2.0 / 3.0));
    s.ma = T(m.c[0]) * sc * s.a;
    s.k = T(m.c[1]) * sc;
    return s;
  }
  template <typename T>
  __device__ static void stress(const St<T>& s, const T (&)[3][3], T (&P)[3][3]) {
    const T c1 = s.ma * s.I1 / T(3), c2 = s.k * s.J * (s.J - T(1));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) P[i][j] = s.ma * s.F[i * 3 + j] + (c2 - c1) * s.H[i * 3 + j];
  }
  template <typename T>
  __device__ static void tangent(const St<T>& s, const T* ga, const T* gb, T w, T (&K)[3][3]) {
    T fa[3], ha[3], fb[3], hb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      fa[i] = s.F[i * 3 + 0] * ga[0] + s.F[i * 3 + 1] * ga[1] + s.F[i * 3 + 2] * ga[2];
      ha[i] = s.H[i * 3 + 0] * ga[0] + s.H[i * 3 + 1] * ga[1] + s.H[i * 3 + 2] * ga[2];
      fb[i] = s.F[i * 3 + 0] * gb[0] + s.F[i * 3 + 1] * gb[1] + s.F[i * 3 + 2] * gb[2];
      hb[i] = s.H[i * 3 + 0] * gb[0] + s.H[i * 3 + 1] * gb[1] + s.H[i * 3 + 2] * gb[2];
    }
    const T ma = w * s.ma, k = w * s.k;
    const T gg = ma * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
    const T c_fh = ma * T(-2.0 / 3.0);
    const T c_hh = ma * T(2.0 / 9.0) * s.I1 + k * (T(2) * s.J * s.J - s.J);
    const T c_tr = ma * s.I1 / T(3) - k * (s.J * s.J - s.J);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int kk = 0; kk < 3; ++kk)
        K[i][kk] += c_fh * (fa[i] * hb[kk] + ha[i] * fb[kk]) + c_hh * ha[i] * hb[kk] + c_tr * hb[i] * ha[kk] +
                    (i == kk ? gg : T(0));
  }
  template <typename T>
  __device__ static void jvp(const St<T>& s, const T (&dG)[3][3], T (&dP)[3][3]) {
    T HdG = 0, FdG = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        HdG += s.H[i * 3 + j] * dG[i][j];
        FdG += s.F[i * 3 + j] * dG[i][j];
      }
    // M = dG^T H, then H M
    T M[3][3];
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int j = 0; j < 3; ++j) M[l][j] = dG[0][l] * s.H[0 * 3 + j] + dG[1][l] * s.H[1 * 3 + j] + dG[2][l] * s.H[2 * 3 + j];
    const T c_f = s.ma * T(-2.0 / 3.0) * HdG;
    const T c_h = s.ma * (T(-2.0 / 3.0) * FdG + T(2.0 / 9.0) * s.I1 * HdG) + s.k * (T(2) * s.J * s.J - s.J) * HdG;
    const T c_t = s.ma * s.I1 / T(3) - s.k * (s.J * s.J - s.J);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const T HM = s.H[i * 3 + 0] * M[0][j] + s.H[i * 3 + 1] * M[1][j] + s.H[i * 3 + 2] * M[2][j];
        dP[i][j] = s.ma * dG[i][j] + c_f * s.F[i * 3 + j] + c_h * s.H[i * 3 + j] + c_t * HM;
      }
  }
  template <typename T>
  __device__ static void dparam(const fem_material_t& m, int64_t, const St<T>& s, const T (&G)[3][3],
                                T (&dP)[3][3]) {
    if (!m.has_param) {
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) dP[i][j] = T(0);
      return;
    }
    // P is linear in the scale s: dP/ds = P evaluated with mu = c0, kappa = c1
    const T ma = T(m.c[0]) * s.a, k = T(m.c[1]);
    const T c1 = ma * s.I1 / T(3), c2 = k * s.J * (s.J - T(1));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) dP[i][j] = ma * s.F[i * 3 + j] + (c2 - c1) * s.H[i * 3 + j];
  }
};

// Small-strain J2 (von Mises) perfect plasticity with radial return from the
// state of the last converged load step (tests/benchmarks/plasticity/
// test_plasticity.py:32-64):
//   eps = sym(G), sig_tr = lambda tr(eps - eps_old) I + 2 mu (eps - eps_old) + sig_old,
//   s = dev(sig_tr), q = sqrt(3/2 s:s), f = q - sig_y,
//   sig = sig_tr - max(f, 0) s / q.
// Consistent tangent (the derivative of that map, which the reference obtains
// by forward-mode AD): with beta = max(f, 0) / q and gam = 3 mu sig_y / q^3 on
// the plastic branch (beta = gam = 0 on the elastic one),
//   dsig = lam_t tr(deps) I + 2 mu_t deps - gam (s : deps) s,
//   lam_t = lambda + 2 mu beta / 3, mu_t = mu (1 - beta).
// c0 = lambda, c1 = mu, c2 = sig_y; state[cq] = (sig_old, eps_old), 3x3 row-major each.
struct MatJ2 {
  static constexpr int V = 3;
  static constexpr bool kNeedsU = true;
  template <typename T>
  struct St {
    T sig[9], s[9];
    T lam_t, mu_t, gam;
  };
  template <typename T>
  __device__ static St<T> prep(const fem_material_t& m, int64_t cq, const T (&G)[3][3]) {
    St<T> st;
    const double* so = m.state + cq * 18;
    const double lam = m.c[0], mu = m.c[1], sy = m.c[2];
    double de[9], tr = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) de[i * 3 + j] = 0.5 * (double(G[i][j]) + double(G[j][i])) - so[9 + i * 3 + j];
    tr = de[0] + de[4] + de[8];
    double sig[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) sig[k] = 2.0 * mu * de[k] + so[k];
    sig[0] += lam * tr;
    sig[4] += lam * tr;
    sig[8] += lam * tr;
    const double p = (sig[0] + sig[4] + sig[8]) / 3.0;
    double sd[9], ss = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      sd[k] = sig[k] - ((k == 0 || k == 4 || k == 8) ? p : 0.0);
      ss += sd[k] * sd[k];
    }
    const double q = ss > 0.0 ? sqrt(1.5 * ss) : 0.0;
    const double f = q - sy;
    double beta = 0.0, gam = 0.0;
    if (f > 0.0) {
      beta = f / q;
      gam = 3.0 * mu * sy / (q * q * q);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      st.sig[k] = T(sig[k] - beta * sd[k]);
      st.s[k] = T(sd[k]);
    }
    st.lam_t = T(lam + 2.0 * mu * beta / 3.0);
    st.mu_t = T(mu * (1.0 - beta));
    st.gam = T(gam);
    return st;
  }
  template <typename T>
  __device__ static void stress(const St<T>& st, const T (&)[3][3], T (&P)[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) P[i][j] = st.sig[i * 3 + j];
  }
  template <typename T>
  __device__ static void tangent(const St<T>& st, const T* ga, const T* gb, T w, T (&K)[3][3]) {
    T sa[3], sb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      sa[i] = st.s[i * 3 + 0] * ga[0] + st.s[i * 3 + 1] * ga[1] + st.s[i * 3 + 2] * ga[2];
      sb[i] = st.s[i * 3 + 0] * gb[0] + st.s[i * 3 + 1] * gb[1] + st.s[i * 3 + 2] * gb[2];
    }
    const T l = w * st.lam_t, mt = w * st.mu_t, g = w * st.gam;
    const T gg = mt * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        K[i][k] += l * ga[i] * gb[k] + mt * ga[k] * gb[i] + (i == k ? gg : T(0)) - g * sa[i] * sb[k];
  }
  template <typename T>
  __device__ static void jvp(const St<T>& st, const T (&dG)[3][3], T (&dP)[3][3]) {
    T de[9], sde = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        de[i * 3 + j] = T(0.5) * (dG[i][j] + dG[j][i]);
        sde += st.s[i * 3 + j] * de[i * 3 + j];
      }
    const T tr = de[0] + de[4] + de[8];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        dP[i][j] = T(2) * st.mu_t * de[i * 3 + j] - st.gam * sde * st.s[i * 3 + j] + (i == j ? st.lam_t * tr : T(0));
  }
  template <typename T>
  __device__ static void dparam(const fem_material_t&, int64_t, const St<T>&, const T (&)[3][3], T (&dP)[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) dP[i][j] = T(0);
  }
};

// Tabulated law: P and D = dP/d(grad u) come from a table filled at the
// current u before the launch (the value_and_jacfwd of an arbitrary tensor map,
// problem.py:262-266), dP/dtheta likewise for the adjoint.  The state is the
// quad point's table row; the contractions with grad N stay in the kernels.
template <int V_>
struct MatTabulated {
  static constexpr int V = V_;
  static constexpr bool kNeedsU = false;
  static constexpr int kStride = 6 * V + 9 * V * V;
  template <typename T>
  struct St {
    const double* s;  // P[V][3], D[V][3][V][3], dP/dtheta[V][3]
  };
  template <typename T>
  __device__ static St<T> prep(const fem_material_t& m, int64_t cq, const T (&)[V][3]) {
    return St<T>{m.state + cq * kStride};
  }
  template <typename T>
  __device__ static void stress(const St<T>& st, const T (&)[V][3], T (&P)[V][3]) {
#pragma unroll
    for (int i = 0; i < V; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) P[i][j] = T(st.s[i * 3 + j]);
  }
  template <typename T>
  __device__ static void tangent(const St<T>& st, const T* ga, const T* gb, T w, T (&K)[V][V]) {
    const double* D = st.s + 3 * V;
#pragma unroll
    for (int i = 0; i < V; ++i)
#pragma unroll
      for (int k = 0; k < V; ++k) {
        T acc = T(0);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          T t = T(0);
#pragma unroll
          for (int l = 0; l < 3; ++l) t += T(D[((i * 3 + j) * V + k) * 3 + l]) * gb[l];
          acc += ga[j] * t;
        }
        K[i][k] += w * acc;
      }
  }
  template <typename T>
  __device__ static void jvp(const St<T>& st, const T (&dG)[V][3], T (&dP)[V][3]) {
    const double* D = st.s + 3 * V;
#pragma unroll
    for (int i = 0; i < V; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        T t = T(0);
#pragma unroll
        for (int k = 0; k < V; ++k)
#pragma unroll
          for (int l = 0; l < 3; ++l) t += T(D[((i * 3 + j) * V + k) * 3 + l]) * dG[k][l];
        dP[i][j] = t;
      }
  }
  template <typename T>
  __device__ static void dparam(const fem_material_t& m, int64_t, const St<T>& st, const T (&)[V][3],
                                T (&dP)[V][3]) {
    const double* d = st.s + 3 * V + 9 * V * V;
#pragma unroll
    for (int i = 0; i < V; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) dP[i][j] = m.has_param ? T(d[i * 3 + j]) : T(0);
  }
};

}  // namespace fem
