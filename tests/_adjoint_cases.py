"""Shared states and objectives of the adjoint tests (CPU layout: ghost-padded
AoS float64 arrays of shape (m0, m1, m2, 5))."""
import math

import numpy as np

GAMMA = 1.4


def cons(w):
    # Euler::to_cons with the ideal-gas EOS, in the reference's operation order
    rho, u, v, ww, p = (float(x) for x in w)
    ke = 0.5 * rho * (u * u + v * v + ww * ww)
    rho_e = rho * ((p + GAMMA * 0.0) * (1.0 / rho - 0.0) / (GAMMA - 1.0) + 0.0)
    return np.array([rho, rho * u, rho * v, rho * ww, rho_e + ke])


def fill(o, fn):
    u = np.zeros(o.m + (5,))
    g = [o.ng if v > 1 else 0 for v in o.n]
    for i in range(o.n[0]):
        for j in range(o.n[1]):
            for k in range(o.n[2]):
                u[g[0] + i, g[1] + j, g[2] + k] = cons(fn(i, j, k))
    return u


RIEMANN = [(1.5, 0, 0, 0, 1.5), (0.5323, 1.206, 0, 0, 0.3), (0.138, 1.206, 1.206, 0, 0.029),
           (0.5323, 0, 1.206, 0, 0.3)]


def riemann(i, j, n, split=0.8):
    e, nn = (i + 0.5) / n > split, (j + 0.5) / n > split
    return RIEMANN[0] if e and nn else RIEMANN[1] if nn else RIEMANN[2] if not e else RIEMANN[3]


def smooth_state(o, seed):
    """Smooth data on which every characteristic field varies along every
    line.  WENO-Z weights (eps = 1e-40) are not differentiable where a field
    is exactly constant along a stencil: beta ~ h^2 under a perturbation h,
    so tau / beta jumps by O(1) and no finite difference converges there."""
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * np.pi, 10)
    tp = 2 * math.pi

    def f(i, j, k):
        x, y = (i + 0.5) / o.n[0], (j + 0.5) / o.n[1]
        return [1.0 + 0.2 * math.sin(tp * x + ph[0]) * math.cos(tp * y + ph[1]),
                0.3 + 0.1 * math.sin(tp * y + ph[2]) + 0.05 * math.cos(tp * x + ph[3]),
                -0.2 + 0.1 * math.cos(tp * x + ph[4]) * math.sin(tp * y + ph[5]),
                0.05 * math.sin(tp * (x + y) + ph[6]),
                1.0 + 0.1 * math.cos(tp * x + ph[7]) + 0.05 * math.sin(tp * y + ph[8])]

    return fill(o, f)


def pressure_window(o, u, lo=0.6, hi=0.8):
    """J = sum of p dx dy over cells with centre in [lo, hi]^2 and dJ/du."""
    g = [o.ng if v > 1 else 0 for v in o.n]
    dxdy = 1.0 / (o.n[0] * o.n[1])
    J = 0.0
    grad = np.zeros_like(u)
    for i in range(o.n[0]):
        for j in range(o.n[1]):
            x, y = (i + 0.5) / o.n[0], (j + 0.5) / o.n[1]
            if lo <= x <= hi and lo <= y <= hi:
                q = u[g[0] + i, g[1] + j, 0]
                rho, m, E = q[0], q[1:4], q[4]
                vel = m / rho
                J += (GAMMA - 1) * (E - 0.5 * np.dot(m, vel)) * dxdy
                grad[g[0] + i, g[1] + j, 0] = (GAMMA - 1) * dxdy * np.array(
                    [0.5 * np.dot(vel, vel), -vel[0], -vel[1], -vel[2], 1.0])
    return J, grad
