"""Pins the adjoint oracle (oracle/emref_adj.cpp): the Weno-mode residual
and SSP-RK3 step written over a scalar type, evaluated in double and on a
reverse-mode tape.

* The double instantiation is the reference arithmetic: it reproduces the
  reference's `weno 24 20` golden state hash and the restated solver's
  residuals bit for bit, boundaries included.
* The tape derivatives are checked against central finite differences of
  the double instantiation (dot-product form for the residual, per-cell
  gradients through SSP-RK3 steps with a frozen dt sequence), on states
  where no branch (limiter, fallback, HLLC case) switches inside the
  difference stencil.  A mistake in a derivative rule, a transposed operand
  or a dropped term fails these at the 1e-6 level.
"""
import math

import numpy as np
import pytest

from _oracle import OracleAdj, OracleSolver

GAMMA = 1.4


def _cons(w):
    # Euler::to_cons with the ideal-gas EOS, in the reference's operation order
    rho, u, v, ww, p = (float(x) for x in w)
    ke = 0.5 * rho * (u * u + v * v + ww * ww)
    rho_e = rho * ((p + GAMMA * 0.0) * (1.0 / rho - 0.0) / (GAMMA - 1.0) + 0.0)
    return np.array([rho, rho * u, rho * v, rho * ww, rho_e + ke])


def _fill(o, fn):
    u = np.zeros(o.m + (5,))
    g = [o.ng if v > 1 else 0 for v in o.n]
    for i in range(o.n[0]):
        for j in range(o.n[1]):
            for k in range(o.n[2]):
                u[g[0] + i, g[1] + j, g[2] + k] = _cons(fn(i, j, k))
    return u


def _fnv(a):
    h = 0xcbf29ce484222325
    for v in np.ascontiguousarray(a).view(np.uint64).ravel():
        h = ((h ^ int(v)) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_double_instantiation_reproduces_the_weno_golden_hash():
    # cpu_validate weno 24 20 -> hash 69cffca8bd6aac4d (analysis/cpu_round6.json:968-975)
    n = 24
    o = OracleAdj((n, n, n), (0, 0, 0), (1, 1, 1))

    tau, dx = 6.283185307179586, 1.0 / n

    def init(i, j, k):  # examples/cpu_validate.rs:30-39 (Mesh::x cell centres)
        x, y = 0.0 + (i + 0.5) * dx, 0.0 + (j + 0.5) * dx
        return [1.0 + 0.1 * math.sin(tau * (x - 0.3 * 0.0)) * math.cos(tau * (y - 0.2 * 0.0)), 0.3, 0.2, 0.1, 1.0]

    u = o.steps(_fill(o, init), [0.02 / 20] * 20)
    assert _fnv(o.interior(u)) == 0x69cffca8bd6aac4d


RIEMANN = [(1.5, 0, 0, 0, 1.5), (0.5323, 1.206, 0, 0, 0.3), (0.138, 1.206, 1.206, 0, 0.029),
           (0.5323, 0, 1.206, 0, 0.3)]


def _riemann(i, j, n, split=0.8):
    e, nn = (i + 0.5) / n > split, (j + 0.5) / n > split
    return RIEMANN[0] if e and nn else RIEMANN[1] if nn else RIEMANN[2] if not e else RIEMANN[3]


@pytest.mark.parametrize("bc", [
    [["outflow", "outflow"], ["outflow", "outflow"], ["periodic", "periodic"]],
    [["reflect", ("dirichlet", _cons((1.0, 0.1, 0.0, 0.0, 1.0)))], ["periodic", "periodic"], ["periodic"] * 2],
])
def test_double_residual_matches_the_restated_solver(bc):
    n = 20
    o = OracleAdj((n, n, 1), (0, 0, 0), (1, 1, 1), bc=bc)
    s = OracleSolver((n, n, 1), (0, 0, 0), (1, 1, 1))
    for d in range(3):
        for side in range(2):
            b = bc[d][side]
            s.set_bc(d, side, *((b,) if isinstance(b, str) else b))
    s.set_opts(mode="weno")
    u = _fill(o, lambda i, j, k: _riemann(i, j, n, 0.55))
    s.set_u(u)
    for stage_dt in (0.0, 2e-3, 5e-2):  # 5e-2 makes the limiter act
        ro, st = s.rhs(stage_dt)
        ra = o.rhs(stage_dt, u)
        assert np.array_equal(o.interior(ra), s.interior(ro))
    assert st[3] > 0, "the last residual should have limited faces"


def _smooth_state(o, seed):
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

    return _fill(o, f)


@pytest.mark.parametrize("bc", [
    [["periodic", "periodic"]] * 3,
    [["outflow", "reflect"], ["reflect", "outflow"], ["periodic", "periodic"]],
])
def test_tape_vjp_of_the_residual_matches_central_differences(bc):
    n = 12
    o = OracleAdj((n, n, 1), (0, 0, 0), (1, 1, 1), bc=bc)
    u = _smooth_state(o, 3)
    rng = np.random.default_rng(7)
    lam = np.zeros_like(u)
    dlt = np.zeros_like(u)
    o.interior(lam)[...] = rng.standard_normal(o.interior(lam).shape)
    o.interior(dlt)[...] = rng.standard_normal(o.interior(dlt).shape) * np.array([1, 1, 1, 0, 2])
    stage_dt = 1e-3
    vjp = o.vjp_rhs(stage_dt, u, lam)
    h = 1e-6
    jvp = (o.rhs(stage_dt, u + h * dlt) - o.rhs(stage_dt, u - h * dlt)) / (2 * h)
    a = np.sum(o.interior(jvp) * o.interior(lam))
    b = np.sum(o.interior(vjp) * o.interior(dlt))
    assert abs(a - b) <= 1e-6 * abs(b)


def _pressure_window(o, u, lo=0.6, hi=0.8):
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


def test_tape_gradient_through_steps_matches_central_differences():
    n = 12
    bc = [["outflow", "outflow"], ["outflow", "outflow"], ["periodic", "periodic"]]
    o = OracleAdj((n, n, 1), (0, 0, 0), (1, 1, 1), bc=bc)
    u0 = _smooth_state(o, 11)
    dts = [2e-3, 2.5e-3, 2e-3]
    uN = o.steps(u0, dts)
    _, dj = _pressure_window(o, uN, 0.3, 0.7)
    grad = o.gradient(u0, dts, dj)
    rng = np.random.default_rng(5)
    g = o.ng
    for _ in range(6):
        i, j, q = int(rng.integers(2, n - 2)), int(rng.integers(2, n - 2)), int(rng.integers(0, 5))
        h = 1e-6 * max(1.0, abs(u0[g + i, g + j, 0, q]))
        up, um = u0.copy(), u0.copy()
        up[g + i, g + j, 0, q] += h
        um[g + i, g + j, 0, q] -= h
        fd = (_pressure_window(o, o.steps(up, dts), 0.3, 0.7)[0] - _pressure_window(o, o.steps(um, dts), 0.3, 0.7)[0]) / (2 * h)
        ad = grad[g + i, g + j, 0, q]
        # relative to the entry, or to 1% of the largest entry where FD rounding dominates
        assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1e-2 * np.abs(grad).max()), (i, j, q, fd, ad)
