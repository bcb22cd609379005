"""Pins the oracle's generic scheme/solver (oracle/emref_gen.h) and its
two-phase Kapila five-equation system (oracle/emref_multi.h, restating
fluid/crates/compressible/src/multiphase.rs) before they are trusted as the
checker of the HIP FiveEq path.

* Golden vectors: the reference's `interface` probe lines (FiveEq, Hybrid
  mode, 1.75 M of 2.6 M faces on the shock path) are reproduced bit for bit
  by tests/test_oracle.py::test_cpu_validate_golden_hash; the same run is
  repeated here through the Python handle.
* The generic solver instantiated for Euler must equal the Euler oracle
  (which the Euler golden hashes pin) bit for bit, in every mode and with
  every boundary kind.
* Closed forms for FiveEq: to_prim inverts to_cons, L R = I of the
  eigenvectors, HLLC consistency F(w, w) = f(w), Wood's sound speed with the
  Kapila term, a uniform state is stationary, and the interface keeps
  pressure and velocity uniform (the property the reference's probe checks).
"""
import json
import os

import numpy as np
import pytest

from _oracle import FIVEEQ_WATER_AIR, OracleGen, OracleSolver

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cpu_validate_golden.json")))


def test_interface_probe_golden_hash_through_the_handle():
    case = next(c for c in GOLD["cases"] if c["args"] == ["interface", 24, 20])
    o = OracleGen("fiveeq", (24, 24, 24), (0, 0, 0), (1, 1, 1))
    o.init_interface()
    o.set_opts(mode="hybrid")
    for _ in range(20):
        o.step(1e-5)
    assert f"{o.hash():016x}" == case["hash"]
    st = o.stats()
    # the probe exercises the blend: shock-path faces but not all of them
    assert 0 < st[1] < st[0]


# ---------------------------------------------- generic Euler == Euler oracle

BCS = [
    [["outflow", "outflow"], ["outflow", "outflow"], ["periodic", "periodic"]],
    [["periodic"] * 2, ["reflect"] * 2, ["outflow", "reflect"]],
    [[("dirichlet", [1.2, 0.3, 0.1, 0.0, 2.6]), ("characteristic", [1.0, 0.0, 0.0, 0.0, 2.5])],
     [("characteristic", [1.0, 0.0, 0.0, 0.0, 2.5])] * 2, ["outflow"] * 2],
    [["reflect"] * 2, ["reflect"] * 2, ["reflect"] * 2],
]


def _set_bcs(o, bc):
    for d in range(3):
        for s in range(2):
            b = bc[d][s]
            if isinstance(b, str):
                o.set_bc(d, s, b)
            else:
                o.set_bc(d, s, b[0], b[1])


def _riemann_u(n, ng=6):
    import emergence_amd

    return emergence_amd.case_init(0, n, (0, 0, 0), (1, 1, 1), ng)


def _smooth_u(n, ng=6):
    import emergence_amd

    return emergence_amd.case_init(2, n, (0, 0, 0), (1, 1, 1), ng)


@pytest.mark.parametrize("mode", ["hybrid", "weno", "central"])
@pytest.mark.parametrize("bci", range(len(BCS)))
def test_generic_euler_residual_equals_the_euler_oracle(mode, bci):
    bc = BCS[bci]
    n = (20, 16, 1) if bci == 0 else (2, 3, 4) if bci == 3 else (9, 7, 6)
    u = _riemann_u(n) if bci == 0 else _smooth_u(n)
    a = OracleSolver(n, (0, 0, 0), (1, 1, 1))
    b = OracleGen("euler", n, (0, 0, 0), (1, 1, 1))
    for o in (a, b):
        o.set_u(u)
        o.set_opts(mode=mode)
        _set_bcs(o, bc)
    for sdt in (1e-3, 0.8):
        ra, sa = a.rhs(sdt)
        rb, sb = b.rhs(sdt)
        assert np.array_equal(ra, rb), (mode, bci, sdt)
        assert sa == sb
    assert a.dt_cfl() == b.dt_cfl()


def test_generic_euler_steps_equal_the_euler_oracle():
    n = (24, 20, 1)
    u = _riemann_u(n)
    a = OracleSolver(n, (0, 0, 0), (1, 1, 1))
    b = OracleGen("euler", n, (0, 0, 0), (1, 1, 1))
    for o in (a, b):
        o.set_u(u)
        o.set_opts(mode="hybrid")
        _set_bcs(o, BCS[0])
    assert a.run_to(0.05) == b.run_to(0.05)
    assert np.array_equal(a.get_u(), b.get_u())
    assert a.stats() == b.stats()


def test_generic_euler_reproduces_the_weno_golden_hash():
    case = next(c for c in GOLD["cases"] if c["args"] == ["weno", 24, 20])
    n = (24, 24, 24)
    o = OracleGen("euler", n, (0, 0, 0), (1, 1, 1))
    o.set_u(_smooth_u(n))
    o.set_opts(mode="weno")
    for _ in range(20):
        o.step(0.02 / 20)
    assert f"{o.hash():016x}" == case["hash"]


# ---------------------------------------------------------- FiveEq closed forms


def _five():
    return OracleGen("fiveeq", (4, 1, 1), (0, 0, 0), (1, 1, 1))


def _cons(o, w):
    """to_cons by hand (multiphase.rs:162-173) for the water/air pair."""
    g1, pi1, _, _, g2, pi2, _, _, _ = FIVEEQ_WATER_AIR
    a = w[6]
    d = a / (g1 - 1.0) + (1.0 - a) / (g2 - 1.0)
    b = g1 * pi1 * a / (g1 - 1.0) + g2 * pi2 * (1.0 - a) / (g2 - 1.0)
    rho = w[0] + w[1]
    return np.array([w[0], w[1], rho * w[2], rho * w[3], rho * w[4],
                     w[5] * d + b + 0.5 * rho * (w[2] ** 2 + w[3] ** 2 + w[4] ** 2), a])


STATES = [
    [500.0, 0.6, 0.3, -0.2, 0.1, 2e5, 0.5],
    [999.0, 1e-3, -1.0, 0.0, 0.0, 1e5, 0.999],
    [1e-3, 1.19, 5.0, 2.0, -3.0, 8e4, 1e-6],
]


@pytest.mark.parametrize("w", STATES)
def test_fiveeq_to_prim_inverts_to_cons(w):
    o = _five()
    u = _cons(o, np.array(w))
    back = o.to_prim(u)[0]
    np.testing.assert_allclose(back, w, rtol=1e-12, atol=1e-9)


def _wood_c(w):
    """Wood's mixture sound speed 1/(rho c^2) = sum alpha_k/(rho_k c_k^2) of
    stiffened gases, rho_k c_k^2 = gamma_k (p + p_inf,k)."""
    g1, pi1, _, _, g2, pi2, _, _, _ = FIVEEQ_WATER_AIR
    a = w[6]
    z1, z2 = g1 * (w[5] + pi1), g2 * (w[5] + pi2)
    inv = a / z1 + (1.0 - a) / z2
    return np.sqrt(1.0 / inv / (w[0] + w[1]))


@pytest.mark.parametrize("w", STATES)
@pytest.mark.parametrize("d", [0, 1, 2])
def test_fiveeq_eigenvectors_are_inverse(w, d):
    # L R = I; a wrong sign or a misplaced Kapila entry breaks it
    l, r = _five().eigen(np.array(w), d)
    np.testing.assert_allclose(l @ r, np.eye(7), atol=1e-10)


@pytest.mark.parametrize("w", STATES)
@pytest.mark.parametrize("d", [0, 1, 2])
def test_fiveeq_hllc_is_consistent_and_keeps_a_contact(w, d):
    o = _five()
    w = np.array(w)
    f, ss = o.riemann(w, w, d)
    np.testing.assert_allclose(f, o.phys_flux(w, d), rtol=1e-13, atol=1e-9)
    assert abs(ss - w[2 + d]) <= 1e-9 * (1 + abs(w[2 + d]))
    # an isolated material contact (same p and u, different phase content) is
    # resolved exactly: contact speed u, no pressure or velocity flux error
    w2 = w.copy()
    w2[6] = 1.0 - w[6]
    w2[0], w2[1] = w2[6] * 1000.0, (1.0 - w2[6]) * 1.2
    f, ss = o.riemann(w, w2, d)
    un = w[2 + d]
    assert abs(ss - un) <= 1e-9 * (1 + abs(un))
    up = w if un >= 0 else w2
    np.testing.assert_allclose(f, o.phys_flux(up, d), rtol=1e-10, atol=1e-6)


@pytest.mark.parametrize("w", STATES)
def test_fiveeq_sound_speed_is_woods(w):
    # the Kapila coefficient turns the model's mixture sound speed into Wood's
    # (multiphase.rs:8-12)
    o = _five()
    u = _cons(o, np.array(w))
    # max_speed enters dt_cfl: a 4-cell line of identical states at rest in
    # the x direction only, dx = 1/4, no dilatation -> dt = cfl / ((|u|+c)/dx)
    line = np.tile(u, (4 + 12, 1)).reshape(o.shape)
    o.set_u(line)
    o.set_opts(cfl=0.5)
    dt = o.dt_cfl()
    c = _wood_c(w)
    expect = 0.5 / ((abs(w[2]) + c) / 0.25)
    assert abs(dt - expect) <= 1e-12 * expect


def test_fiveeq_uniform_state_is_stationary_and_interface_keeps_equilibrium():
    # a moving material interface at uniform p and u: after steps p and u stay
    # uniform to round-off (the reference probe's max_relative_dp/max_du) and a
    # uniform state does not change at all
    n = (32, 1, 1)
    o = OracleGen("fiveeq", n, (0, 0, 0), (1, 1, 1))
    u = np.zeros(o.shape)
    for i in range(n[0]):
        a = 1e-6 + (1 - 2e-6) * (0.5 + 0.5 * np.tanh((i - 12.5) / 1.5))
        u[6 + i, 0, 0] = _cons(o, np.array([a * 1000.0, (1 - a) * 1.2, 50.0, 0.0, 0.0, 1e5, a]))
    o.set_u(u)
    o.set_opts(mode="hybrid")
    dt = o.dt_cfl()
    for _ in range(10):
        o.step(dt)
    w = o.to_prim(o.interior(o.get_u()))
    assert np.abs(w[:, 5] / 1e5 - 1).max() < 1e-10
    assert np.abs(w[:, 2] - 50.0).max() < 1e-9
    o2 = OracleGen("fiveeq", n, (0, 0, 0), (1, 1, 1))
    uni = np.zeros(o2.shape)
    uni[6:-6, 0, 0] = _cons(o2, np.array([500.0, 0.6, 3.0, 0.0, 0.0, 2e5, 0.5]))
    o2.set_u(uni)
    o2.set_opts(mode="weno")
    r, _ = o2.rhs(1e-4)
    assert np.abs(r).max() <= 1e-9 * np.abs(uni).max()


def test_fiveeq_dt_includes_the_dilatation_limit():
    # a bubbly mixture (alpha = 0.5: Wood's c ~ 23 m/s) under a square-wave
    # velocity of +-100 m/s: |Dp/Dt| / p = rho c^2 |div u| / p limits dt below
    # the acoustic CFL step (multiphase.rs:373-391, solver.rs:496-519);
    # expected value by hand: dt = min(cfl / rate, 0.5 / source)
    n = (16, 1, 1)
    o = OracleGen("fiveeq", n, (0, 0, 0), (1.0, 1, 1))
    u = np.zeros(o.shape)
    vel = [100.0 if (i // 2) % 2 == 0 else -100.0 for i in range(n[0])]
    ws = [np.array([500.0, 0.6, v, 0.0, 0.0, 1e5, 0.5]) for v in vel]
    for i, w in enumerate(ws):
        u[6 + i, 0, 0] = _cons(o, w)
    o.set_u(u)
    o.set_opts(cfl=0.5)
    dt = o.dt_cfl()
    dx = 1.0 / 16
    rate = max((abs(w[2]) + _wood_c(w)) / dx for w in ws)
    source = 0.0
    for i, w in enumerate(ws):
        va, vb = vel[(i - 1) % 16], vel[(i + 1) % 16]  # periodic neighbours
        rc2 = (w[0] + w[1]) * _wood_c(w) ** 2
        source = max(source, rc2 * abs(vb - va) / (2 * dx) / 1e5)
    expect = min(0.5 / rate, 0.5 / source)
    assert 0.5 / source < 0.5 / rate  # the dilatation limit binds
    assert abs(dt - expect) <= 1e-12 * expect
