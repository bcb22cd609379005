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

from _adjoint_cases import cons as _cons
from _adjoint_cases import fill as _fill
from _adjoint_cases import pressure_window as _pressure_window
from _adjoint_cases import riemann as _riemann
from _adjoint_cases import smooth_state as _smooth_state
from _oracle import OracleAdj, OracleSolver

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


def test_tape_vjp_on_flat_stencils_takes_the_linear_weights():
    # A uniform state at rest makes every WENO-Z stencil flat to rounding.
    # With the linear-weight convention the adjoint residual inherits the
    # scheme's mirror symmetry: the response to a unit energy cotangent at one
    # cell is the same on both sides along each axis, with the normal
    # momentum negated.  Z-weight derivatives of rounding noise (the
    # unregularised tape) break this by O(1).
    n = 12
    o = OracleAdj((n, n, 1), (0, 0, 0), (1, 1, 1))
    u = _fill(o, lambda i, j, k: [1.0, 0.0, 0.0, 0.0, 1.0])
    g, c = o.ng, 5
    lam = np.zeros_like(u)
    lam[g + c, g + c, 0, 4] = 1.0
    v = o.interior(o.vjp_rhs(1e-3, u, lam))[:, :, 0]
    scale = np.abs(v).max()
    assert scale > 1.0
    for d in range(2):
        for off in (1, 2, 3):
            a = v[c + off, c] if d == 0 else v[c, c + off]
            b = (v[c - off, c] if d == 0 else v[c, c - off]).copy()
            b[1 + d] *= -1.0
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * scale)


def test_tape_vjp_on_partly_flat_stencils_holds_the_weights():
    # A 3x3 plateau of a second state at rest: stencils crossing its edges
    # have one or two sub-stencils flat to rounding (beta = 0 or a rounding
    # residue such as 3b - 4b + b) next to O(1) ones.  The Z-weight
    # derivative through a residue beta is rounding noise amplified by
    # ~1e-16 tau / sqrt(beta), which moves by O(1) when the plateau values
    # move by a few ulps (the raw tape changes by 107% here).  With the
    # weights held on such stencils the adjoint residual is as stable as the
    # residual itself.  Perturbations are kept to those under which the
    # forward residual does not move either: WENO-Z's own value jumps with
    # rounding on symmetric stencils (tau = |beta0 - beta2| a residue), and
    # there no derivative is defined.
    n, c = 16, 7
    o = OracleAdj((n, n, 1), (0, 0, 0), (1, 1, 1))

    def state(k):
        e = k * 2.0 ** -52
        b = [0.125 * (1 + e), 0.0, 0.0, 0.0, 0.1 * (1 + e)]

        def f(i, j, _):
            di, dj = min(abs(i - c), n - abs(i - c)), min(abs(j - c), n - abs(j - c))
            return b if max(di, dj) <= 1 else [1.0, 0.0, 0.0, 0.0, 1.0]

        return _fill(o, f)

    rng = np.random.default_rng(3)
    u0 = state(0)
    lam = np.zeros_like(u0)
    o.interior(lam)[...] = rng.standard_normal(o.interior(lam).shape)
    r0 = o.interior(o.rhs(1e-3, u0))
    v0 = o.interior(o.vjp_rhs(1e-3, u0, lam))
    stable = 0
    for k in range(1, 8):
        u = state(k)
        if np.abs(o.interior(o.rhs(1e-3, u)) - r0).max() > 1e-13 * np.abs(r0).max():
            continue
        stable += 1
        v = o.interior(o.vjp_rhs(1e-3, u, lam))
        assert np.abs(v - v0).max() <= 1e-12 * np.abs(v0).max(), k
    assert stable >= 2


def _weno_grad_exact(v):
    """d weno5z_interp / d v of the exact function (rational arithmetic on the
    given doubles, forward-mode duals; eps = 1e-40 as its double value)."""
    from fractions import Fraction as Fr

    n = 5

    class Du:
        def __init__(self, a, d):
            self.a, self.d = a, d

        def __add__(s, o):
            o = o if isinstance(o, Du) else Du(Fr(o), [Fr(0)] * n)
            return Du(s.a + o.a, [x + y for x, y in zip(s.d, o.d)])

        __radd__ = __add__

        def __sub__(s, o):
            o = o if isinstance(o, Du) else Du(Fr(o), [Fr(0)] * n)
            return Du(s.a - o.a, [x - y for x, y in zip(s.d, o.d)])

        def __rsub__(s, o):
            return Du(Fr(o), [Fr(0)] * n) - s

        def __mul__(s, o):
            if not isinstance(o, Du):
                o = Fr(o)
                return Du(s.a * o, [x * o for x in s.d])
            return Du(s.a * o.a, [x * o.a + s.a * y for x, y in zip(s.d, o.d)])

        __rmul__ = __mul__

        def __truediv__(s, o):
            if not isinstance(o, Du):
                o = Fr(o)
                return Du(s.a / o, [x / o for x in s.d])
            return Du(s.a / o.a, [(x * o.a - s.a * y) / (o.a * o.a) for x, y in zip(s.d, o.d)])

    x = [Du(Fr(float(v[k])), [Fr(int(j == k)) for j in range(n)]) for k in range(n)]
    q = [(3 * x[0] - 10 * x[1] + 15 * x[2]) / 8, (-1 * x[1] + 6 * x[2] + 3 * x[3]) / 8,
         (3 * x[2] + 6 * x[3] - 1 * x[4]) / 8]
    t = [x[0] - 2 * x[1] + x[2], x[1] - 2 * x[2] + x[3], x[2] - 2 * x[3] + x[4]]
    s = [x[0] - 4 * x[1] + 3 * x[2], x[1] - x[3], 3 * x[2] - 4 * x[3] + x[4]]
    b = [Fr(13, 12) * (t[k] * t[k]) + Fr(1, 4) * (s[k] * s[k]) for k in range(3)]
    d = b[0] - b[2]
    tau = d if d.a >= 0 else -1 * d
    eps = Fr(1e-40)
    dw = [Fr(1, 16), Fr(10, 16), Fr(5, 16)]
    a = [dw[k] * (1 + (tau / (b[k] + eps)) * (tau / (b[k] + eps))) for k in range(3)]
    out = (a[0] * q[0] + a[1] * q[1] + a[2] * q[2]) / (a[0] + a[1] + a[2])
    return np.array([float(g) for g in out.d])


def test_tape_weno_vjp_matches_exact_derivatives():
    # Exact derivatives of WENO5-Z (rational arithmetic on the same double
    # inputs) pin the tape's weight derivative.  On a nearly flat sub-stencil
    # (beta above the rounding floor but small) the chain rule through
    # out = N / S subtracts two nearly equal terms and leaves up to
    # eps |q| / sqrt(beta) of noise; the cancellation-free form is accurate.
    from _oracle import weno_vjp

    rng = np.random.default_rng(11)
    cases = [rng.uniform(-1, 2, 5) for _ in range(6)]
    # one sub-stencil varying by e (beta ~ e^2) next to O(1) jumps; whether
    # the direct form is noisy depends on how out rounds against q_k, and
    # several of these (e = 3e-11 .. 1e-9) leave it 1e-7 .. 3e-5 off
    for e in (1e-12, 3e-11, 1e-10, 3e-10, 1e-9, 3e-9):
        cases += [np.array([1.0, 1.0 + e, 1.0 + 3.3 * e, 1.2, 0.7]),
                  np.array([0.7, 1.2, 1.0 + 3.3 * e, 1.0 + e, 1.0]),
                  np.array([0.5, 0.8, 1.0, 1.0 + 2.1 * e, 1.0 + 3.9 * e])]
    for v in cases:
        got, ref = weno_vjp(v), _weno_grad_exact(v)
        assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max(), (v, got, ref)


@pytest.mark.parametrize("stage_dt", [0.8, 6.0])
def test_tape_vjp_with_the_limiter_acting_matches_central_differences(stage_dt):
    # Large stage dt makes the positivity limiter blend on 27 / 213 faces of
    # smooth data.  Its bisection resolves theta to 2^-30, so the residual
    # follows the admissibility boundary theta*(u): the tape differentiates
    # theta implicitly at the constraint that bounds it, and the
    # Lax-Friedrichs speed through its max.  Holding both constant is off by
    # 4e-3 .. 6e-2 here.
    n = 12
    o = OracleAdj((n, n, 1), (0, 0, 0), (1, 1, 1))
    u = _smooth_state(o, 3)
    s = OracleSolver((n, n, 1), (0, 0, 0), (1, 1, 1))
    s.set_opts(mode="weno")
    s.set_u(u)
    assert s.rhs(stage_dt)[1][3] > 20, "the limiter should act"
    rng = np.random.default_rng(7)
    lam = np.zeros_like(u)
    dlt = np.zeros_like(u)
    o.interior(lam)[...] = rng.standard_normal(o.interior(lam).shape)
    o.interior(dlt)[...] = rng.standard_normal(o.interior(dlt).shape) * np.array([1, 1, 1, 0, 2])
    b = np.sum(o.interior(o.vjp_rhs(stage_dt, u, lam)) * o.interior(dlt))
    h = 1e-6  # small enough that no face changes branch
    jvp = (o.rhs(stage_dt, u + h * dlt) - o.rhs(stage_dt, u - h * dlt)) / (2 * h)
    a = np.sum(o.interior(jvp) * o.interior(lam))
    assert abs(a - b) <= 2e-6 * abs(b)


def test_per_cell_fd_error_is_truncation_then_rounding():
    """Step-size study of the per-cell central difference on a reduced
    adjoint workload (the 2-D Riemann problem with 16 seeded density modes,
    32^2, 20 frozen-dt steps, J = integral of p over [0.6, 0.8]^2): for cells
    on the gradient's support the relative error |FD - tape| / |tape| falls
    like h^2 while truncation dominates, reaches its floor near h = 1e-6 rho,
    and grows again like 1/h below it (rounding of J over h).  So a per-cell
    error that is large at a fixed small h is the difference quotient's
    rounding floor, not an adjoint error: at the best h every cell agrees to a
    few 1e-6 or better."""
    import os
    import sys

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench  # the adjoint workload's initial state

    n, steps = 32, 20
    u0 = bench.adjoint_case(n)
    bc = [["outflow", "outflow"], ["outflow", "outflow"], ["periodic", "periodic"]]
    s = OracleSolver((n, n, 1), (0, 0, 0), (1, 1, 1))
    s.set_u(u0)
    s.set_opts(mode="weno", cfl=0.5)
    for d in range(2):
        for sd in range(2):
            s.set_bc(d, sd, "outflow")
    dts = []
    for _ in range(steps):
        dt = s.dt_cfl()
        dts.append(dt)
        s.step(dt)
    o = OracleAdj((n, n, 1), (0, 0, 0), (1, 1, 1), bc=bc)
    _, dj = _pressure_window(o, o.steps(u0, dts))
    g = o.gradient(u0, dts, dj)
    q0 = u0[6:6 + n, 6:6 + n, 0]
    vel = q0[..., 1:4] / q0[..., :1]
    dirn = np.concatenate([np.ones_like(q0[..., :1]), vel, 0.5 * np.sum(vel * vel, -1, keepdims=True)], -1)
    gr = np.sum(g[6:6 + n, 6:6 + n, 0] * dirn, -1)  # dJ/drho_0 at fixed velocity and pressure
    support = np.argwhere(np.abs(gr) >= 1e-2 * np.abs(gr).max())
    cells = support[np.random.default_rng(0x5EED).choice(len(support), size=8, replace=False)]
    hs = [1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9]
    best = []
    for i, j in cells:
        errs = []
        for h in hs:
            du = np.zeros_like(u0)
            du[6 + i, 6 + j, 0] = dirn[i, j] * h * q0[i, j, 0]
            jp = _pressure_window(o, o.steps(u0 + du, dts))[0]
            jm = _pressure_window(o, o.steps(u0 - du, dts))[0]
            errs.append(abs((jp - jm) / (2 * h * q0[i, j, 0]) - gr[i, j]) / abs(gr[i, j]))
        k = int(np.argmin(errs))
        best.append(errs[k])
        assert 2 <= k <= 5, (i, j, errs)  # the floor lies between h = 1e-5 and 1e-8
        assert errs[0] > 10 * errs[k] and errs[-1] > 3 * errs[k], (i, j, errs)  # truncation, then rounding
    assert max(best) < 5e-6 and float(np.median(best)) < 1e-6, best
