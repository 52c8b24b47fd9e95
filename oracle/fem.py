"""FP64 CPU restatement of jax-fem's assembly, solve and adjoint (the oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Each function cites the
reference code it restates.  Everything is float64 numpy/scipy; per-cell work
is vectorised over chunks of cells.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import basis as B


# ----------------------------------------------------------------- meshes
def box_mesh(Nx, Ny, Nz, Lx, Ly, Lz):
    """HEX8 box, node/cell numbering of generate_mesh.py:151-189."""
    x, y, z = np.linspace(0, Lx, Nx + 1), np.linspace(0, Ly, Ny + 1), np.linspace(0, Lz, Nz + 1)
    X, Y, Z = np.meshgrid(x, y, z, indexing='ij')
    pts = np.stack([X, Y, Z], axis=3).reshape(-1, 3)
    ids = np.arange(len(pts)).reshape(Nx + 1, Ny + 1, Nz + 1)
    c = [ids[:-1, :-1, :-1], ids[1:, :-1, :-1], ids[1:, 1:, :-1], ids[:-1, 1:, :-1],
         ids[:-1, :-1, 1:], ids[1:, :-1, 1:], ids[1:, 1:, 1:], ids[:-1, 1:, 1:]]
    return pts, np.stack(c, axis=3).reshape(-1, 8)


# --------------------------------------------------------------- geometry
def shape_grads(points, cells, grads_ref, weights):
    """Physical grad N (C,Q,N,3) and JxW (C,Q) (fe.py:112-141)."""
    X = points[cells]                                         # (C,N,3)
    J = np.einsum('cai,qaj->cqij', X, grads_ref)              # dx_i/dxi_j
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)
    G = np.einsum('qak,cqkj->cqaj', grads_ref, Jinv)
    return G, det * weights[None, :]


def face_shape_grads(points, cells, boundary_inds, fgrads_ref, fweights, normals):
    """Face grad N (S,FQ,N,3) and Nanson scale (S,FQ) (fe.py:143-183)."""
    X = points[cells[boundary_inds[:, 0]]]                    # (S,N,3)
    f = boundary_inds[:, 1]
    J = np.einsum('sai,sqaj->sqij', X, fgrads_ref[f])
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)
    G = np.einsum('sqak,sqkj->sqaj', fgrads_ref[f], Jinv)
    nv = np.einsum('sk,sqkj->sqj', normals[f], Jinv)
    return G, np.linalg.norm(nv, axis=-1) * det * fweights[f]


def quad_points(points, cells, shape_vals):
    """(C,Q,3) (fe.py:185-197)."""
    return np.einsum('qa,cad->cqd', shape_vals, points[cells])


def face_quad_points(points, cells, boundary_inds, fvals):
    """(S,FQ,3) (fe.py:199-218)."""
    return np.einsum('sqa,sad->sqd', fvals[boundary_inds[:, 1]], points[cells[boundary_inds[:, 0]]])


def location_flags(fn, points):
    """fn(point) or fn(point, ind) evaluated point by point (fe.py:244-252)."""
    nargs = fn.__code__.co_argcount
    if nargs == 1:
        return np.array([bool(fn(p)) for p in points])
    return np.array([bool(fn(p, i)) for i, p in enumerate(points)])


def dirichlet(points, vec, bc_info):
    """Unique Dirichlet dofs and values, later sets winning (fe.py:220-258, solver.py:290-304)."""
    val = {}
    if bc_info is not None:
        for loc, comp, vf in zip(*bc_info):
            for n in np.argwhere(location_flags(loc, points)).reshape(-1):
                val[int(n) * vec + comp] = float(vf(points[n]))
    d = np.array(sorted(val), dtype=np.int64)
    return d, np.array([val[k] for k in d], dtype=float)


def boundary_inds(points, cells, face_inds, fn):
    """(S,2) selected (cell, face) pairs, row-major order (fe.py:271-322)."""
    flags = location_flags(fn, points)
    return np.argwhere(np.all(flags[cells[:, face_inds]], axis=-1))


# -------------------------------------------------------------- materials
def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


class Material:
    """Stress P(G) and its derivative A_iJkL = dP_iJ/dG_kL, batched over (...,V,3).

    kind: 'poisson' (k), 'elastic' (lmbda, mu), 'simp' (lmbda1, mu1, Emin, E0, p; per-quad
    theta), 'neo_hookean' (mu, kappa; optional per-quad scale), 'j2' (lmbda, mu, sig0; per-quad
    state (..., 18) = sigma_old then eps_old, 3x3 row-major, passed as theta), 'user' (fn: any
    map P = fn(G, theta) written with numpy operations that carry complex numbers; its
    derivatives are complex-step derivatives, Squire & Trapp, SIAM Review 40 (1998) 110-112:
    df/dx = Im f(x + i h) / h with h = 1e-30, exact to rounding for analytic maps, in place of
    the forward-mode AD of problem.py:262-266).
    """
    CSTEP = 1e-30

    def __init__(self, kind, **c):
        self.kind, self.c = kind, c

    def _lm(self, theta):
        c = self.c
        if self.kind == 'simp':
            E = c['Emin'] + (c['E0'] - c['Emin']) * theta ** c['p']
            return c['lmbda1'] * E, c['mu1'] * E
        return c['lmbda'], c['mu']

    def dE(self, theta):
        c = self.c
        return (c['E0'] - c['Emin']) * c['p'] * theta ** (c['p'] - 1)

    def _j2(self, G, state):
        """Radial return of test_plasticity.py:47-64: stress, deviator, beta = f+/q, gam = 3 mu sig0/q^3."""
        c = self.c
        st = np.asarray(state)
        sig_old = st[..., :9].reshape(st.shape[:-1] + (3, 3))
        eps_old = st[..., 9:].reshape(st.shape[:-1] + (3, 3))
        de = 0.5 * (G + np.swapaxes(G, -1, -2)) - eps_old
        tr = np.trace(de, axis1=-2, axis2=-1)[..., None, None]
        sig = c['lmbda'] * tr * np.eye(3) + 2. * c['mu'] * de + sig_old
        sd = sig - np.trace(sig, axis1=-2, axis2=-1)[..., None, None] / 3. * np.eye(3)
        q = np.sqrt(1.5 * np.sum(sd * sd, axis=(-2, -1)))
        f = q - c['sig0']
        plastic = f > 0
        qs = np.where(plastic, q, 1.)
        beta = np.where(plastic, f / qs, 0.)
        gam = np.where(plastic, 3. * c['mu'] * c['sig0'] / qs ** 3, 0.)
        return sig - beta[..., None, None] * sd, sd, beta, gam

    def stress(self, G, theta=None):
        if self.kind == 'user':
            return np.real(self.c['fn'](G, theta))
        if self.kind == 'j2':
            return self._j2(G, theta)[0]
        if self.kind == 'poisson':
            k = self.c['k'] * (1. if theta is None else np.asarray(theta)[..., None, None])
            return k * G
        if self.kind in ('elastic', 'simp'):
            lm, mu = self._lm(theta)
            lm = np.asarray(lm)[..., None, None]
            mu = np.asarray(mu)[..., None, None]
            tr = np.trace(G, axis1=-2, axis2=-1)[..., None, None]
            return lm * tr * np.eye(3) + mu * (G + np.swapaxes(G, -1, -2))
        s = 1. if theta is None else np.asarray(theta)[..., None, None]
        F = G + np.eye(3)
        J = np.linalg.det(F)[..., None, None]
        H = np.swapaxes(np.linalg.inv(F), -1, -2)
        I1 = np.sum(F * F, axis=(-2, -1))[..., None, None]
        mu, ka = self.c['mu'] * s, self.c['kappa'] * s
        return mu * J ** (-2. / 3.) * (F - I1 / 3. * H) + ka * J * (J - 1.) * H

    def tangent(self, G, theta=None):
        """A (...,V,3,V,3)."""
        if self.kind == 'user':
            G = np.asarray(G, dtype=float)
            V = G.shape[-2]
            A = np.empty(G.shape + (V, 3))
            for k in range(V):
                for l in range(3):
                    Gc = G.astype(complex)
                    Gc[..., k, l] += 1j * self.CSTEP
                    A[..., k, l] = np.imag(self.c['fn'](Gc, theta)) / self.CSTEP
            return A
        d = np.eye(3)
        if self.kind == 'j2':
            _, sd, beta, gam = self._j2(G, theta)
            c = self.c
            lt = (c['lmbda'] + 2. * c['mu'] * beta / 3.)[..., None, None, None, None]
            mt = (c['mu'] * (1. - beta))[..., None, None, None, None]
            return lt * np.einsum('ij,kl->ijkl', d, d) + mt * (np.einsum('ik,jl->ijkl', d, d) +
                                                             np.einsum('il,jk->ijkl', d, d)) - \
                gam[..., None, None, None, None] * np.einsum('...ij,...kl->...ijkl', sd, sd)
        if self.kind == 'poisson':
            k = self.c['k'] * (1. if theta is None else np.asarray(theta)[..., None, None, None, None])
            return k * np.broadcast_to(d[None, :, None, :], G.shape[:-2] + (1, 3, 1, 3))
        if self.kind in ('elastic', 'simp'):
            lm, mu = self._lm(theta)
            lm = np.asarray(lm)[..., None, None, None, None]
            mu = np.asarray(mu)[..., None, None, None, None]
            A = lm * np.einsum('ij,kl->ijkl', d, d) + mu * (np.einsum('ik,jl->ijkl', d, d) +
                                                         np.einsum('il,jk->ijkl', d, d))
            return np.broadcast_to(A, G.shape[:-2] + (3, 3, 3, 3))
        s = 1. if theta is None else np.asarray(theta)[..., None, None, None, None]
        F = G + np.eye(3)
        J = np.linalg.det(F)[..., None, None, None, None]
        H = np.swapaxes(np.linalg.inv(F), -1, -2)
        I1 = np.sum(F * F, axis=(-2, -1))[..., None, None, None, None]
        a = J ** (-2. / 3.)
        mu, ka = self.c['mu'] * s, self.c['kappa'] * s
        HF = np.einsum('...kl,...ij->...ijkl', H, F)
        FH = np.einsum('...kl,...ij->...ijkl', F, H)
        HH = np.einsum('...ij,...kl->...ijkl', H, H)
        HHt = np.einsum('...il,...kj->...ijkl', H, H)
        dd = np.einsum('ik,jl->ijkl', d, d)
        return mu * a * (dd - 2. / 3. * (HF + FH) + 2. / 9. * I1 * HH + I1 / 3. * HHt) + \
            ka * ((2. * J * J - J) * HH - (J * J - J) * HHt)

    def dstress_dtheta(self, G, theta):
        if self.kind == 'user':
            th = np.asarray(theta, dtype=float).astype(complex) + 1j * self.CSTEP
            return np.imag(self.c['fn'](np.asarray(G, dtype=float).astype(complex), th)) / self.CSTEP
        if self.kind == 'simp':
            c = self.c
            dE = self.dE(theta)[..., None, None]
            tr = np.trace(G, axis1=-2, axis2=-1)[..., None, None]
            return dE * (c['lmbda1'] * tr * np.eye(3) + c['mu1'] * (G + np.swapaxes(G, -1, -2)))
        if self.kind == 'neo_hookean':
            return self.stress(G, np.ones_like(theta))
        if self.kind == 'poisson':
            return self.c['k'] * G
        raise ValueError(self.kind)


# ------------------------------------------------------------- kernels
def cell_residual(mat, u, cells, G, JxW, theta=None, chunk=8192):
    """Per-cell internal force (C,N,V) of the laplace kernel (problem.py:189-214)."""
    C_, N = cells.shape
    V = u.shape[1]
    out = np.empty((C_, N, V))
    for s in range(0, C_, chunk):
        e = slice(s, min(C_, s + chunk))
        ue = u[cells[e]]                                      # (c,N,V)
        Gu = np.einsum('cav,cqaj->cqvj', ue, G[e])
        P = mat.stress(Gu, None if theta is None else theta[e])
        out[e] = np.einsum('cqvj,cqaj,cq->cav', P, G[e], JxW[e])
    return out


def cell_tangent(mat, u, cells, G, JxW, theta=None, chunk=2048):
    """Per-cell tangent (C, N*V, N*V), jac[a*V+i, b*V+k] = dR_ai/du_bk (problem.py:262-266)."""
    C_, N = cells.shape
    V = u.shape[1]
    out = np.empty((C_, N * V, N * V))
    for s in range(0, C_, chunk):
        e = slice(s, min(C_, s + chunk))
        Gu = np.einsum('cav,cqaj->cqvj', u[cells[e]], G[e])
        A = mat.tangent(Gu, None if theta is None else theta[e])   # (c,q,V,3,V,3)
        K = np.einsum('cqaj,cqijkl,cqbl,cq->caibk', G[e], A, G[e], JxW[e])
        out[e] = K.reshape(-1, N * V, N * V)
    return out


def cell_load(m, shape_vals, JxW):
    """Mass-map term (C,N,V) = sum_q m N_a JxW (problem.py:216-236) for u-independent m (C,Q,V)."""
    return np.einsum('cqv,qa,cq->cav', m, shape_vals, JxW)


def cell_mass_tangent(M, shape_vals, JxW):
    """Tangent (C, N*V, N*V) of a u-dependent mass term: the jacfwd of the mass kernel
    (problem.py:216-236, 262-266), sum_q N_a N_b dm_i/du_k JxW for M = dm/du (C,Q,V,V)."""
    C_, Q, V, _ = M.shape
    N = shape_vals.shape[1]
    K = np.einsum('cqik,qa,qb,cq->caibk', M, shape_vals, shape_vals, JxW)
    return K.reshape(C_, N * V, N * V)


def mass_map_derivative(fn, uq, xq, ivs, h=1e-30):
    """dm/du (C,Q,V,V) of a numpy mass map m(u, x, *ivs) by complex steps
    (Squire & Trapp 1998), the oracle's counterpart of forward-mode AD."""
    V = uq.shape[-1]
    M = np.empty(uq.shape + (V,))
    for k in range(V):
        uc = uq.astype(complex)
        uc[..., k] += 1j * h
        M[..., k] = np.imag(fn(uc, xq, *ivs)) / h
    return M


def face_load(t, fvals, nanson, bi):
    """Surface-map term (S,N,V) (problem.py:238-259)."""
    return np.einsum('sqv,sqa,sq->sav', t, fvals[bi[:, 1]], nanson)


def scatter(per_cell, cells, num_nodes):
    """Scatter-add (C,N,V) into (num_nodes,V) (problem.py:426-437)."""
    V = per_cell.shape[-1]
    out = np.zeros((num_nodes, V))
    np.add.at(out, cells.reshape(-1), per_cell.reshape(-1, V))
    return out


# ------------------------------------------------------- global matrix
def coo_pattern(cells, vec):
    """I, J of problem.py:86-97 (cells only)."""
    inds = (vec * cells[:, :, None] + np.arange(vec)[None, None, :]).reshape(len(cells), -1)
    I = np.repeat(inds[:, :, None], inds.shape[1], axis=2).reshape(-1)
    J = np.repeat(inds[:, None, :], inds.shape[1], axis=1).reshape(-1)
    return I, J


def csr_from_cells(Ke, cells, vec, n):
    """COO (I, J, V) -> CSR with duplicates summed, sorted columns (solver.py:472-478, 522-526)."""
    I, J = coo_pattern(cells, vec)
    A = sp.coo_matrix((Ke.reshape(-1), (I, J)), shape=(n, n)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A


def zero_rows(A, rows):
    """PETSc zeroRows(rows, diag=1) keeping the pattern (solver.py:527-528)."""
    A = A.tocsr(copy=True)
    for r in rows:
        s, e = A.indptr[r], A.indptr[r + 1]
        A.data[s:e] = 0.
        k = np.searchsorted(A.indices[s:e], r)
        A.data[s + k] = 1.
    return A


# -------------------------------------------------------------- solvers
def bicgstab_jax(A, b, x0, diag, tol=1e-10, atol=1e-10, maxiter=10000):
    """jax.scipy.sparse.linalg.bicgstab with M = x/diag(A) as called at solver.py:68-84.

    Restates JAX's algorithm (non-legacy tolerance: stop when |r|^2 <= max(tol^2|b|^2, atol^2));
    JAX is not pinned by the reference (environment.yml has no version).
    """
    M = lambda v: v / diag
    atol2 = max(tol ** 2 * (b @ b), atol ** 2)
    x = x0.copy()
    r = b - A @ x
    rhat = r.copy()
    alpha = rho = omega = 1.
    p = r.copy()
    q = r.copy()
    k = 0
    while r @ r > atol2 and k < maxiter:
        rho_ = rhat @ r
        beta = rho_ / rho * alpha / omega
        p = r + beta * (p - omega * q)
        phat = M(p)
        q = A @ phat
        alpha = rho_ / (rhat @ q)
        s = r - alpha * q
        if s @ s < atol2:
            x = x + alpha * phat
            r = s
            k += 1
            break
        shat = M(s)
        t = A @ shat
        omega = (t @ s) / (t @ t)
        x = x + alpha * phat + omega * shat
        r = s - omega * t
        rho = rho_
        k += 1
    return x, k


def pcg_jacobi(A, b, x0, diag, tol=1e-10, atol=1e-10, maxiter=10000):
    """Jacobi-preconditioned CG, same stopping rule as bicgstab_jax."""
    atol2 = max(tol ** 2 * (b @ b), atol ** 2)
    x = x0.copy()
    r = b - A @ x
    z = r / diag
    p = z.copy()
    rz = r @ z
    k = 0
    while r @ r > atol2 and k < maxiter:
        q = A @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        z = r / diag
        rz_new = r @ z
        p = z + rz_new / rz * p
        rz = rz_new
        k += 1
    return x, k


# ---------------------------------------------------------------- problem
class Problem:
    """Single-variable problem (problem.py:56-182 setup, restated).

    mat: Material; theta: per-quad parameter (C,Q) or None; mass (C,Q,V) and
    surfaces [(boundary_inds, t (S,FQ,V))] are u-independent load data.
    """

    def __init__(self, points, cells, vec, ele_type='HEX8', quadrature_order=None, bc_info=None,
                 mat=None, theta=None):
        self.points = np.asarray(points, float)
        self.cells = np.asarray(cells, np.int64)
        self.vec, self.ele_type = vec, ele_type
        self.sv, self.sg, self.w = B.shape_vals_and_grads(ele_type, quadrature_order)
        self.fv, self.fg, self.fw, self.fn, self.face_inds = B.face_shape_vals_and_grads(ele_type, quadrature_order)
        self.G, self.JxW = shape_grads(self.points, self.cells, self.sg, self.w)
        self.n = len(self.points) * vec
        self.bc_dofs, self.bc_vals = dirichlet(self.points, vec, bc_info)
        self.mat, self.theta = mat, theta
        self.load = np.zeros((len(self.points), vec))
        self.mass_fn, self.mass_ivs = None, ()
        self.kernels = []    # per-cell / per-face weak-form terms: (cell ids, fn, args)

    def quad_points(self):
        return quad_points(self.points, self.cells, self.sv)

    def add_mass(self, m):
        self.load += scatter(cell_load(m, self.sv, self.JxW), self.cells, len(self.points))

    def set_mass_map(self, fn, ivs=()):
        """A mass map m(u_q, x_q, *ivs) that depends on u (numpy, batched over (C,Q)), kept in
        the residual and differentiated by complex steps in the tangent."""
        self.mass_fn, self.mass_ivs = fn, tuple(ivs)

    def add_cell_kernel(self, fn, ivs=()):
        """get_universal_kernel (problem.py:275-283): fn(cell_sol (C, N*V), x (C,Q,3),
        shape_grads (C,Q,N,3), JxW (C,1,Q), v_grads_JxW (C,Q,N,1,3), *ivs) -> (C, N*V),
        numpy, batched over cells."""
        x = self.quad_points()
        vg = self.G[:, :, :, None, :] * self.JxW[:, :, None, None, None]
        self.kernels.append((np.arange(len(self.cells)), fn, (x, self.G, self.JxW[:, None, :], vg) + tuple(ivs)))

    def add_face_kernel(self, bi, fn, ivs=()):
        """get_universal_kernels_surface (problem.py:305-327): fn(cell_sol (S, N*V), x (S,FQ,3),
        face_shape_vals (S,FQ,N), face_shape_grads (S,FQ,N,3), nanson (S,1,FQ), *ivs) -> (S, N*V)."""
        g, ns = face_shape_grads(self.points, self.cells, bi, self.fg, self.fw, self.fn)
        x = face_quad_points(self.points, self.cells, bi, self.fv)
        self.kernels.append((bi[:, 0], fn, (x, self.fv[bi[:, 1]], g, ns[:, None, :]) + tuple(ivs)))

    def _kernel_terms(self, u, jac, h=1e-30):
        """Per-cell residual (C,N,V) and, with jac, tangent (C, N*V, N*V) of the kernels, the
        tangent by complex steps (Squire & Trapp 1998)."""
        C_, N = self.cells.shape
        V = self.vec
        r = np.zeros((C_, N * V))
        K = np.zeros((C_, N * V, N * V)) if jac else None
        for ids, fn, args in self.kernels:
            cs = u[self.cells[ids]].reshape(len(ids), N * V)
            np.add.at(r, ids, np.real(fn(cs, *args)))
            if jac:
                Kt = np.empty((len(ids), N * V, N * V))
                for k in range(N * V):
                    cc = cs.astype(complex)
                    cc[:, k] += 1j * h
                    Kt[:, :, k] = np.imag(fn(cc, *args)) / h
                np.add.at(K, ids, Kt)
        return r.reshape(C_, N, V), K

    def _mass_at(self, u):
        uq = np.einsum('qa,cav->cqv', self.sv, u[self.cells])
        return uq, self.quad_points()

    def add_surface(self, bi, t):
        _, ns = face_shape_grads(self.points, self.cells, bi, self.fg, self.fw, self.fn)
        self.load += scatter(face_load(t, self.fv, ns, bi), self.cells[bi[:, 0]], len(self.points))

    def residual(self, dofs, bc=True):
        u = dofs.reshape(-1, self.vec)
        r = scatter(cell_residual(self.mat, u, self.cells, self.G, self.JxW, self.theta), self.cells,
                    len(self.points)) + self.load
        if self.mass_fn is not None:
            uq, xq = self._mass_at(u)
            m = np.real(self.mass_fn(uq, xq, *self.mass_ivs))
            r = r + scatter(cell_load(m, self.sv, self.JxW), self.cells, len(self.points))
        if self.kernels:
            r = r + scatter(self._kernel_terms(u, False)[0], self.cells, len(self.points))
        r = r.reshape(-1)
        if bc:  # apply_bc_vec (solver.py:290-304)
            r[self.bc_dofs] = dofs[self.bc_dofs] - self.bc_vals
        return r

    def tangent(self, dofs, bc=True):
        Ke = cell_tangent(self.mat, dofs.reshape(-1, self.vec), self.cells, self.G, self.JxW, self.theta)
        if self.mass_fn is not None:
            uq, xq = self._mass_at(dofs.reshape(-1, self.vec))
            Ke = Ke + cell_mass_tangent(mass_map_derivative(self.mass_fn, uq, xq, self.mass_ivs), self.sv, self.JxW)
        if self.kernels:
            Ke = Ke + self._kernel_terms(dofs.reshape(-1, self.vec), True)[1]
        A = csr_from_cells(Ke, self.cells, self.vec, self.n)
        return zero_rows(A, self.bc_dofs) if bc else A

    def newton(self, dofs=None, linear='direct', tol=1e-6, rel_tol=1e-8, lin_tol=1e-10):
        """Newton loop of solver.py:1295-1333 with x0 of solver.py:402-409."""
        dofs = np.zeros(self.n) if dofs is None else dofs.copy()
        res = self.residual(dofs)
        r0 = rv = np.linalg.norm(res)
        rel = rv / r0 if r0 else np.nan
        its = []
        while rel > rel_tol and rv > tol:
            A = self.tangent(dofs)
            b = -res
            x0 = np.zeros(self.n)
            x0[self.bc_dofs] = self.bc_vals - dofs[self.bc_dofs]
            inc, k = self.linear_solve(A, b, x0, linear, lin_tol)
            its.append(k)
            dofs = dofs + inc
            res = self.residual(dofs)
            rv = np.linalg.norm(res)
            rel = rv / r0
        self.newton_its = its
        return dofs

    def newton_constrained(self, P, tol=1e-6, rel_tol=1e-8):
        """Newton on reduced dofs u = P u_r with the reduced system P^T A P (solver.py:1283-1339,
        get_A 545-551, x0 404-406), direct solves."""
        P = sp.csr_matrix(P)
        d = np.zeros(P.shape[1])
        res = P.T @ self.residual(P @ d)
        r0 = rv = np.linalg.norm(res)
        rel = rv / r0 if r0 else np.nan
        while rel > rel_tol and rv > tol:
            A = (P.T @ self.tangent(P @ d) @ P).tocsc()
            d = d + spla.spsolve(A, -res)
            res = P.T @ self.residual(P @ d)
            rv = np.linalg.norm(res)
            rel = rv / r0
        return P @ d

    @staticmethod
    def linear_solve(A, b, x0, linear, tol=1e-10):
        if linear == 'direct':
            return spla.spsolve(A.tocsc(), b), 0
        d = A.diagonal()
        if linear == 'bicgstab':
            return bicgstab_jax(A, b, x0, d, tol=tol, atol=tol)
        return pcg_jacobi(A, b, x0, d, tol=tol, atol=tol)

    def adjoint_theta(self, dofs, v):
        """-lambda^T dR/dtheta per (cell, quad) at u = dofs (solver.py:1362-1418): A^T lambda = v,
        then the parameter vjp of the BC-applied residual (BC rows do not depend on theta)."""
        A = self.tangent(dofs)
        lam = spla.spsolve(A.T.tocsc(), v)
        lam_f = lam.copy()
        lam_f[self.bc_dofs] = 0.
        u = dofs.reshape(-1, self.vec)
        L = lam_f.reshape(-1, self.vec)
        Gu = np.einsum('cav,cqaj->cqvj', u[self.cells], self.G)
        Gl = np.einsum('cav,cqaj->cqvj', L[self.cells], self.G)
        dP = self.mat.dstress_dtheta(Gu, self.theta)
        return -np.einsum('cqvj,cqvj,cq->cq', dP, Gl, self.JxW), lam


# ------------------------------------------------------ several variables
class MultiProblem:
    """Several variables on one mesh (problem.py:58-119), restated.

    Variable i has vecs[i] components; its dofs sit at offset[i] + vec_i * node + comp
    with offset[i] = sum_{j<i} num_nodes * vec_j (problem.py:82-90, 109-119).  A cell's
    dof vector is the concatenation over the variables of their (N, vec_i) blocks
    (find_ind, problem.py:86-92), and the COO pairs repeat it over rows and columns
    (problem.py:95-97).  The weak form comes from universal kernels (problem.py:275-301):
    fn(cell_sol (C, N*sum(vec)), x (C,Q,3), shape_grads (C,Q,N*nvars,3), JxW (C,nvars,Q),
    v_grads_JxW (C,Q,N*nvars,1,3), *ivs) -> (C, N*sum(vec)), numpy, batched over cells; the
    per-cell matrices are complex-step derivatives (Squire & Trapp 1998).
    """

    def __init__(self, points, cells, vecs, ele_type='HEX8', quadrature_order=None, bc_infos=None):
        self.points = np.asarray(points, float)
        self.cells = np.asarray(cells, np.int64)
        self.vecs = list(vecs)
        self.nv = len(self.vecs)
        self.sv, self.sg, self.w = B.shape_vals_and_grads(ele_type, quadrature_order)
        self.G, self.JxW = shape_grads(self.points, self.cells, self.sg, self.w)
        Nn = len(self.points)
        self.offset = np.cumsum([0] + [Nn * v for v in self.vecs[:-1]])
        self.n = Nn * sum(self.vecs)
        dofs, vals = [], []
        for i, info in enumerate(bc_infos or [None] * self.nv):
            d, v = dirichlet(self.points, self.vecs[i], info)
            dofs.append(d + self.offset[i])
            vals.append(v)
        self.bc_dofs = np.concatenate(dofs).astype(np.int64)
        self.bc_vals = np.concatenate(vals)
        self.kernels = []

    def cell_inds(self):
        """(C, N*sum(vec)) global dofs of each cell, variables concatenated (problem.py:86-92)."""
        return np.hstack([(v * self.cells[:, :, None] + np.arange(v)[None, None, :] + o).reshape(len(self.cells), -1)
                          for v, o in zip(self.vecs, self.offset)])

    def coo_pattern(self):
        inds = self.cell_inds()
        return (np.repeat(inds[:, :, None], inds.shape[1], axis=2).reshape(-1),
                np.repeat(inds[:, None, :], inds.shape[1], axis=1).reshape(-1))

    def add_cell_kernel(self, fn, ivs=()):
        x = quad_points(self.points, self.cells, self.sv)
        g = np.concatenate([self.G] * self.nv, axis=2)
        jxw = np.repeat(self.JxW[:, None, :], self.nv, axis=1)
        vg = g[:, :, :, None, :] * self.JxW[:, :, None, None, None]
        self.kernels.append((fn, (x, g, jxw, vg) + tuple(ivs)))

    def _terms(self, dofs, jac, h=1e-30):
        inds = self.cell_inds()
        cs = dofs[inds]
        r = np.zeros_like(cs)
        K = np.zeros((len(cs), cs.shape[1], cs.shape[1])) if jac else None
        for fn, args in self.kernels:
            r += np.real(fn(cs, *args))
            if jac:
                for k in range(cs.shape[1]):
                    cc = cs.astype(complex)
                    cc[:, k] += 1j * h
                    K[:, :, k] += np.imag(fn(cc, *args)) / h
        return inds, r, K

    def residual(self, dofs, bc=True):
        inds, r, _ = self._terms(dofs, False)
        out = np.zeros(self.n)
        np.add.at(out, inds.reshape(-1), r.reshape(-1))
        if bc:  # apply_bc_vec (solver.py:290-304)
            out[self.bc_dofs] = dofs[self.bc_dofs] - self.bc_vals
        return out

    def tangent(self, dofs, bc=True):
        inds, _, K = self._terms(dofs, True)
        I, J = self.coo_pattern()
        A = sp.coo_matrix((K.reshape(-1), (I, J)), shape=(self.n, self.n)).tocsr()
        A.sum_duplicates()
        A.sort_indices()
        return zero_rows(A, self.bc_dofs) if bc else A

    def newton(self, tol=1e-6, rel_tol=1e-8):
        """Newton loop of solver.py:1295-1333 (x0 of solver.py:402-409), direct solves."""
        dofs = np.zeros(self.n)
        res = self.residual(dofs)
        r0 = rv = np.linalg.norm(res)
        rel = rv / r0 if r0 else np.nan
        while rel > rel_tol and rv > tol:
            dofs = dofs + spla.spsolve(self.tangent(dofs).tocsc(), -res)
            res = self.residual(dofs)
            rv = np.linalg.norm(res)
            rel = rv / r0
        return dofs

    def sol_list(self, dofs):
        """[(num_nodes, vec_i)] (unflatten_fn_sol_list, problem.py:116-117)."""
        Nn = len(self.points)
        return [dofs[o:o + Nn * v].reshape(Nn, v) for v, o in zip(self.vecs, self.offset)]
