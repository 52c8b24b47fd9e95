"""The reference's explicit-dynamics application (applications/explicit_dynamics/
example.py:31-138) as a shared test problem: a 1 x 0.1 x 0.1 elastic waveguide,
HEX8 80x2x2, released from its first fixed-free longitudinal mode and integrated
by central differences over one period with row-sum lumped mass; the
displacement history is compared with u_x = A sin(kx) cos(c k t)."""
import numpy as np

LX, LY, LZ = 1.0, 0.1, 0.1
NX, NY, NZ = 80, 2, 2
E, NU, RHO = 1.0, 0.3, 1.0
AMPLITUDE = 1.0e-3
CFL = 0.5


def locations(isclose):
    left = lambda p: isclose(p[0], 0.0)
    y_sides = lambda p: isclose(p[1], 0.0) | isclose(p[1], LY)
    z_sides = lambda p: isclose(p[2], 0.0) | isclose(p[2], LZ)
    return [left, y_sides, z_sides]


def time_grid(lmbda, mu):
    c = np.sqrt((lmbda + 2.0 * mu) / RHO)
    k = np.pi / (2.0 * LX)
    period = 2.0 * np.pi / (c * k)
    dt_estimate = min(LX / NX, LY / NY, LZ / NZ) / c
    n = int(np.ceil(period / (CFL * dt_estimate)))
    return c, k, n, period / n


def run(acceleration, u0, fixed, dt, steps, xp):
    """Central differences with the second-order start u^-1 (example.py:103-112)."""
    up = xp.where(fixed, 0.0, u0 + 0.5 * dt ** 2 * acceleration(u0))
    u = u0
    hist = [u0]
    for _ in range(steps):
        un = xp.where(fixed, 0.0, 2.0 * u - up + dt ** 2 * acceleration(u))
        up, u = u, un
        hist.append(un)
    return hist


def wave_checks(hist_x, nodal_mass, points_x, c, k, dt):
    """(relative P-wave speed error, max normalised L2 error) as in example.py:116-137."""
    n = hist_x.shape[0] - 1
    times = np.arange(n + 1) * dt
    mode = np.sin(k * points_x)
    exact = AMPLITUDE * np.cos(c * k * times[:, None]) * mode[None, :]
    err = hist_x - exact
    l2 = np.max(np.sqrt(np.sum(nodal_mass[None, :] * err ** 2, axis=1))) / (
        AMPLITUDE * np.sqrt(np.sum(nodal_mass * mode ** 2)))
    modal = np.sum(nodal_mass[None, :] * hist_x * mode[None, :], axis=1)
    i = np.flatnonzero(modal[1:] <= 0.0)[0]
    alpha = modal[i] / (modal[i] - modal[i + 1])
    zero_time = times[i] + alpha * dt
    measured = (np.pi / (2.0 * zero_time)) / k
    return abs(measured / c - 1.0), l2
