"""Oracle problems built from a device Problem's own definition.

Dirichlet data, body loads and tractions are formed by the oracle from the
problem's location functions and maps: the maps are evaluated (torch, CPU, one
quadrature point at a time) at the oracle's quadrature points, and the oracle
integrates and scatters them (oracle/fem.py add_mass / add_surface).  Nothing
is copied from the device state, so a wrong load weighting on the device fails
the comparison.
"""
import numpy as np
import torch

from oracle import fem as OF


def _eval_points(fn, V, u, x, ivs):
    """fn(u_q, x_q, *iv_q) -> (P, V) float64, point by point on the CPU."""
    out = np.empty((x.shape[0], V))
    ut = torch.as_tensor(u, dtype=torch.float64)
    xt = torch.as_tensor(x, dtype=torch.float64)
    it = [torch.as_tensor(np.asarray(iv), dtype=torch.float64) for iv in ivs]
    for i in range(x.shape[0]):
        r = fn(ut[i], xt[i], *[t[i] for t in it])
        if isinstance(r, (list, tuple)):
            r = torch.stack([torch.as_tensor(e, dtype=torch.float64) for e in r])
        out[i] = torch.as_tensor(r, dtype=torch.float64).reshape(V).numpy()
    return out


def _flat_iv(iv, lead):
    a = np.asarray(iv.detach().cpu() if torch.is_tensor(iv) else iv, dtype=np.float64)
    return a.reshape((int(np.prod(lead)),) + a.shape[len(lead):])


def oracle_problem(p, mat, theta=None, loads=True, mass=True):
    """OF.Problem on p's mesh with p's Dirichlet data and, with `loads`, p's u-independent
    surface maps and (with `mass`) its mass map integrated by the oracle."""
    fe = p.fes[0]
    V = fe.vec
    o = OF.Problem(fe.points, fe.cells, V, ele_type=fe.ele_type, quadrature_order=fe.quadrature_order,
                   bc_info=p.dirichlet_bc_info[0], mat=mat, theta=theta)
    if not loads:
        return o
    C_, Q = o.JxW.shape
    ivs = [_flat_iv(iv, (C_, Q)) for iv in (p.internal_vars or ())]
    if mass and hasattr(p, 'get_mass_map'):
        xq = o.quad_points().reshape(-1, 3)
        m = _eval_points(p.get_mass_map(), V, np.zeros((len(xq), V)), xq, ivs)
        o.add_mass(m.reshape(C_, Q, V))
    if hasattr(p, 'get_surface_maps'):
        for i, (loc, smap) in enumerate(zip(p.location_fns or [], p.get_surface_maps())):
            bi = OF.boundary_inds(o.points, o.cells, o.face_inds, loc)
            if len(bi) == 0:
                continue
            xq = OF.face_quad_points(o.points, o.cells, bi, o.fv)
            FQ = xq.shape[1]
            siv = p.internal_vars_surfaces[i] if i < len(p.internal_vars_surfaces) else ()
            sivs = [_flat_iv(iv, (len(bi), FQ)) for iv in siv]
            t = _eval_points(smap, V, np.zeros((len(bi) * FQ, V)), xq.reshape(-1, 3), sivs)
            o.add_surface(bi, t.reshape(len(bi), FQ, V))
    return o
