"""Pins the oracle (the CPU restatement of the reference solver) before it is
trusted as the checker of the HIP path.

* Golden vectors: the reference's own recorded `cpu_validate` output lines
  (tests/golden/cpu_validate_golden.json) must be reproduced bit for bit.
* Closed forms / invariants for what the hashes leave open (Hybrid sensor,
  ghost kinds other than periodic and characteristic, the limiter trigger),
  restating the reference's unit tests (solver.rs:808-930, tests/cpu.rs).
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import _oracle
from _oracle import OracleSolver

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cpu_validate_golden.json")))


def _parse(line):
    out = {}
    for tok in line.split()[1:]:
        k, v = tok.split("=")
        out[k] = v
    return out


@pytest.mark.parametrize("case", GOLD["cases"], ids=lambda c: "-".join(str(a) for a in c["args"]))
def test_cpu_validate_golden_hash(case):
    _oracle.load()
    args = [str(a) for a in case["args"]]
    line = subprocess.run([_oracle.CPU_VALIDATE] + args, check=True, capture_output=True, text=True).stdout
    got = _parse(line.strip())
    assert got["hash"] == case["hash"], line
    for key in ("l1_rho", "mass_error", "max_dp", "max_relative_dp", "max_du"):
        if key in case:
            assert got[key] == case[key], line


def test_open_probe_initialiser_reproduces_the_golden_hash():
    # emref_init_open (the state the GPU open-probe test starts from) stepped
    # by the oracle gives the reference's recorded `open 24 60` line
    case = next(c for c in GOLD["cases"] if c["args"] == ["open", 24, 60])
    o = OracleSolver((24, 24, 24), (0, 0, 0), (1, 1, 1))
    o.init_open()
    far = [1.0, 0.0, 0.0, 0.0, 1.0 / (1.4 - 1.0)]
    for d in range(3):
        for sd in range(2):
            o.set_bc(d, sd, "characteristic", far)
    o.set_opts(mode="central")
    for _ in range(60):
        o.step(0.02 / 60)
    assert f"{o.hash():016x}" == case["hash"]


# ----------------------------------------------------------- closed forms


def test_weno5z_interp_is_exact_on_quadratics(oracle):
    # all three candidate stencils interpolate a quadratic exactly, so any
    # convex weighting does: value at x = 2.5 from samples at x = 0..4
    for a, b, c in [(1.0, 0.0, 0.0), (0.3, -1.2, 0.7), (2.0, 5.0, -3.0)]:
        f = lambda x: a + b * x + c * x * x  # noqa: E731
        v = [f(x) for x in range(5)]
        assert abs(oracle.emref_weno5z_interp(*v) - f(2.5)) < 1e-13 * (1 + abs(f(2.5)))


def test_weno5z_interp_picks_the_smooth_side_of_a_jump(oracle):
    # step between samples 3 and 4: only the leftmost stencil is smooth,
    # the interpolant must stay near the left plateau
    v = oracle.emref_weno5z_interp(1.0, 1.0, 1.0, 1.0, 10.0)
    assert abs(v - 1.0) < 1e-6


def _rf(oracle, wl, wr, d=0, eos=(1.4, 0, 0, 0)):
    import ctypes as C
    f = (C.c_double * 5)()
    s = C.c_double()
    oracle.emref_riemann_flux((C.c_double * 4)(*eos), (C.c_double * 5)(*wl), (C.c_double * 5)(*wr), d, f,
                              C.byref(s))
    return np.array(list(f)), s.value


def _euler_flux(w, d, g=1.4):
    rho, p = w[0], w[4]
    u = np.array(w[1:4])
    E = p / (g - 1) + 0.5 * rho * (u @ u)
    un = u[d]
    f = np.array([rho * un, rho * u[0] * un, rho * u[1] * un, rho * u[2] * un, (E + p) * un])
    f[1 + d] += p
    return f


@pytest.mark.parametrize("w", [[1.0, 0.3, -0.2, 0.1, 1.0], [0.5, 3.0, 0.1, 0.0, 0.2], [0.5, -3.0, 0.1, 0.0, 0.2]])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_hllc_is_consistent(oracle, w, d):
    # F(w, w) = f(w) in the subsonic and both supersonic branches
    f, s = _rf(oracle, w, w, d)
    np.testing.assert_allclose(f, _euler_flux(w, d), rtol=1e-14, atol=1e-14)
    assert abs(s - w[1 + d]) < 1e-14


def test_hllc_resolves_an_isolated_contact(oracle):
    # equal u and p, density jump: the exact flux is the upwind one
    wl, wr = [1.0, 0.2, 0.1, 0.0, 1.0], [0.125, 0.2, 0.1, 0.0, 1.0]
    f, s = _rf(oracle, wl, wr)
    np.testing.assert_allclose(f, _euler_flux(wl, 0), rtol=1e-13, atol=1e-14)
    assert abs(s - 0.2) < 1e-14


# ----------------------------------------------------------- ghosts


def _prim_to_cons(w, g=1.4):
    rho, u, v, ww, p = w
    return np.array([rho, rho * u, rho * v, rho * ww, p / (g - 1) + 0.5 * rho * (u * u + v * v + ww * ww)])


def _linear_field(s):
    u = np.full(s.m + (5,), np.nan)
    n = s.n
    g = [s.ng if v > 1 else 0 for v in n]
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                x = [(c + 0.5) / nn for c, nn in zip((i, j, k), n)]
                u[g[0] + i, g[1] + j, g[2] + k] = _prim_to_cons([1.0 + x[0] + 2 * x[1] + 3 * x[2], 0.3, -0.2, 0.1, 1.0])
    return u, g


@pytest.mark.parametrize("n", [(2, 3, 4), (8, 2, 1)])
@pytest.mark.parametrize("ng", [6, 10])
def test_deep_periodic_ghosts_match_the_torus(n, ng):
    # solver.rs:808-848
    s = OracleSolver(n, (0, 0, 0), (1, 1, 1), ng=ng)
    u, g = _linear_field(s)
    s.set_u(u)
    s.fill_ghosts()
    got = s.get_u()
    for i in range(s.m[0]):
        for j in range(s.m[1]):
            for k in range(s.m[2]):
                c = (i, j, k)
                src = tuple(g[d] + ((c[d] - g[d]) % n[d]) for d in range(3))
                assert np.array_equal(got[c], u[src])


@pytest.mark.parametrize("n", [(2, 3, 4), (8, 2, 1)])
@pytest.mark.parametrize("ng", [6, 10])
def test_short_reflecting_ghosts_match_mirror_images(n, ng):
    # solver.rs:850-880
    s = OracleSolver(n, (0, 0, 0), (1, 1, 1), ng=ng)
    for d in range(3):
        for side in range(2):
            s.set_bc(d, side, "reflect")
    u, g = _linear_field(s)
    s.set_u(u)
    s.fill_ghosts()
    got = s.get_u()
    for i in range(s.m[0]):
        for j in range(s.m[1]):
            for k in range(s.m[2]):
                c = (i, j, k)
                img, flip = [], []
                for d in range(3):
                    r = (c[d] - g[d]) % (2 * n[d])
                    img.append(g[d] + (r if r < n[d] else 2 * n[d] - 1 - r))
                    flip.append(r >= n[d])
                exp = u[tuple(img)].copy()
                for d in range(3):
                    if flip[d]:
                        exp[1 + d] *= -1.0
                assert np.array_equal(got[c], exp)


def test_outflow_and_dirichlet_ghosts():
    s = OracleSolver((8, 6, 1), (0, 0, 0), (1, 1, 1), ng=6)
    far = _prim_to_cons([2.0, 0.5, 0.0, 0.0, 3.0])
    s.set_bc(0, 0, "dirichlet", far)
    s.set_bc(0, 1, "outflow")
    s.set_bc(1, 0, "outflow")
    s.set_bc(1, 1, "outflow")
    u, g = _linear_field(s)
    s.set_u(u)
    s.fill_ghosts()
    got = s.get_u()
    for j in range(6):
        for gi in range(6):
            assert np.array_equal(got[gi, 6 + j, 0], far)
            assert np.array_equal(got[6 + 8 + gi, 6 + j, 0], u[6 + 7, 6 + j, 0])
    for i in range(8):
        for gi in range(6):
            assert np.array_equal(got[6 + i, gi, 0], u[6 + i, 6, 0])
            assert np.array_equal(got[6 + i, 6 + 6 + gi, 0], u[6 + i, 6 + 5, 0])


# ----------------------------------------------------------- invariants


def _smooth_init(n, lo=(0, 0, 0), hi=(1, 1, 1)):
    s = OracleSolver(n, lo, hi)
    u = np.zeros(s.m + (5,))
    g = [s.ng if v > 1 else 0 for v in n]
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                x = [(c + 0.5) / nn for c, nn in zip((i, j, k), n)]
                rho = 1.0 + 0.01 * sum(math.sin(2 * math.pi * v) for v in x)
                u[g[0] + i, g[1] + j, g[2] + k] = _prim_to_cons([rho, 0.3, 0.2, 0.1, 1.0])
    s.set_u(u)
    return s


@pytest.mark.parametrize("mode", ["central", "weno", "hybrid"])
def test_uniform_flow_is_preserved(mode):
    # tests/cpu.rs:266-298 (uniform states stay uniform on periodic axes)
    s = OracleSolver((8, 6, 4), (0, 0, 0), (1, 1, 1))
    u = np.zeros(s.m + (5,))
    u[...] = _prim_to_cons([1.3, 0.4, -0.2, 0.1, 0.9])
    s.set_u(u)
    s.set_opts(mode=mode)
    for _ in range(3):
        s.step(1e-3)
    got = s.interior(s.get_u())
    np.testing.assert_allclose(got, s.interior(u), rtol=1e-14, atol=1e-14)


def test_rk_updates_are_thread_count_independent(oracle):
    # tests/cpu.rs:92-132
    res = []
    for threads in (1, 3):
        oracle.emref_set_threads(threads)
        s = _smooth_init((48, 48, 1))
        s.set_opts(mode="weno")
        s.step(1e-4)
        s.step(1e-4)
        res.append(s.get_u())
    oracle.emref_set_threads(os.cpu_count() or 1)
    assert np.array_equal(res[0], res[1])


def test_strong_pressure_jump_triggers_the_limiter():
    # tests/cpu.rs:134-152
    s = OracleSolver((32, 1, 1), (0, 0, 0), (1, 1, 1))
    u = np.zeros(s.m + (5,))
    for i in range(32):
        x = (i + 0.5) / 32
        u[6 + i, 0, 0] = _prim_to_cons([1.0, 0.0, 0.0, 0.0, 1000.0 if x < 0.5 else 0.01])
    s.set_u(u)
    s.set_opts(mode="central", cfl=0.2)
    s.set_bc(0, 0, "outflow")
    s.set_bc(0, 1, "outflow")
    s.step(s.dt_cfl())
    assert s.stats()[3] > 0
    w = s.interior(s.get_u())[:, 0, 0]
    assert np.all(w[:, 0] > 0)
    p = 0.4 * (w[:, 4] - 0.5 * (w[:, 1] ** 2) / w[:, 0])
    assert np.all(p > 0)


def test_cfl_never_reads_ghosts():
    # tests/cpu.rs:26-52: NaN-poisoned ghosts leave dt unchanged
    s = _smooth_init((12, 10, 1))
    dt0 = s.dt_cfl()
    u = s.get_u()
    mask = np.ones(s.m, dtype=bool)
    mask[6:18, 6:16, 0] = False
    u[mask] = np.nan
    s.set_u(u)
    assert s.dt_cfl() == dt0


def test_conservation_to_round_off():
    s = _smooth_init((16, 12, 8))
    s.set_opts(mode="weno")
    t0 = s.totals()
    for _ in range(3):
        s.step(2e-3)
    t1 = s.totals()
    np.testing.assert_allclose(t1, t0, rtol=0, atol=1e-13)


def test_hybrid_sensor_limits():
    # theta = 0 on smooth data reproduces the central result; theta = 1
    # everywhere (Weno mode) on data with a jump equals the Hybrid result
    # where every face sees the jump
    s1 = _smooth_init((16, 16, 1))
    s1.set_opts(mode="hybrid", tol=1e6)
    s2 = _smooth_init((16, 16, 1))
    s2.set_opts(mode="central")
    r1, st1 = s1.rhs(1e-3)
    r2, st2 = s2.rhs(1e-3)
    assert st1[1] == 0
    assert np.array_equal(r1, r2)
    s3 = _smooth_init((16, 16, 1))
    s3.set_opts(mode="hybrid", tol=1e-30)
    s4 = _smooth_init((16, 16, 1))
    s4.set_opts(mode="weno")
    r3, st3 = s3.rhs(1e-3)
    r4, _ = s4.rhs(1e-3)
    assert st3[1] == st3[0]
    assert np.array_equal(r3, r4)


# ---- Hybrid blend weight, brute force on tiny lines (scheme.rs:94-99, 160-183, 268-276)

def _line_fluxes(w, mode, tol, eos=(1.4, 0.0, 0.0, 0.0), ng=6, d=0):
    import ctypes as C

    lib = _oracle.load()
    lib.emref_line_fluxes.argtypes = [C.POINTER(C.c_double)] * 2 + [C.c_int64, C.c_int, C.c_int64, C.c_int, C.c_int,
                                                                      C.c_double, C.c_double, C.POINTER(C.c_double),
                                                                      C.POINTER(C.c_uint64)]
    w = np.ascontiguousarray(w, dtype=np.float64)
    n = len(w) - 2 * ng
    out = np.empty((n + 1, 5))
    st = (C.c_uint64 * 4)()
    e = np.array(eos, dtype=np.float64)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    lib.emref_line_fluxes(dp(e), dp(w), len(w), d, ng, _oracle.MODE[mode], 0, tol, 0.0, dp(out), st)
    return out, list(st)


def _theta_by_hand(w, tol, p_inf=0.0, ng=6):
    """theta of every face, restated independently: sensor (ln rho, ln(p+p_inf))
    floored at 1e-300 with fmax semantics (NaN ignored), the two 5-point fourth
    differences over cells i-2..i+2 and i-1..i+3 of face i+1/2, max over both
    sensors skipping NaN, smoothstep of (r/tol - 1/2)/(1/2), 1 when r is not finite."""
    def fmax(a, b):
        return b if math.isnan(a) else a if math.isnan(b) else max(a, b)

    sens = [(math.log(fmax(row[0], 1e-300)), math.log(fmax(row[4] + p_inf, 1e-300))) for row in w]
    n = len(w) - 2 * ng
    th, rs = [], []
    for f in range(n + 1):
        i = ng - 1 + f
        r = 0.0
        for k in range(2):
            s = [sens[j][k] for j in range(i - 2, i + 4)]
            a = abs(s[0] - 4.0 * s[1] + 6.0 * s[2] - 4.0 * s[3] + s[4])
            b = abs(s[1] - 4.0 * s[2] + 6.0 * s[3] - 4.0 * s[4] + s[5])
            r = fmax(fmax(r, a), b)
        rs.append(r)
        if not math.isfinite(r):
            th.append(1.0)
            continue
        x = min(max((r / tol - 0.5) / 0.5, 0.0), 1.0)
        th.append(x * x * (3.0 - 2.0 * x))
    return th, rs


def _blend_expected(w, tol, eos=(1.4, 0.0, 0.0, 0.0)):
    th, rs = _theta_by_hand(w, tol, eos[1])
    fw, _ = _line_fluxes(w, "weno", tol, eos)
    fc, _ = _line_fluxes(w, "central", tol, eos)
    exp = np.empty_like(fw)
    for f, t in enumerate(th):
        if t == 0.0:
            exp[f] = fc[f]
        elif t == 1.0:
            exp[f] = fw[f]
        else:
            exp[f] = [t * fw[f][k] + (1.0 - t) * fc[f][k] for k in range(5)]
    return exp, th, rs


def _tiny_line(seed, n=7, ng=6, jump=0.0):
    rng = np.random.default_rng(seed)
    m = n + 2 * ng
    x = np.arange(m, dtype=np.float64)
    rho = 1.0 + 0.05 * np.sin(0.7 * x + rng.uniform(0, 6)) + jump * (x >= m // 2 + rng.integers(-2, 3))
    p = 1.0 + 0.04 * np.cos(0.9 * x + rng.uniform(0, 6)) + 0.01 * rng.standard_normal(m)
    w = np.stack([rho, 0.3 + 0.01 * rng.standard_normal(m), 0.1 * np.ones(m), -0.05 * np.ones(m), p], axis=1)
    return w


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("eos", [(1.4, 0.0, 0.0, 0.0), (1.4, 0.5, 0.0, 0.0)], ids=["ideal", "stiffened"])
def test_hybrid_theta_matches_brute_force_at_intermediate_ratios(seed, eos):
    w = _tiny_line(seed, jump=0.3 * seed)
    _, rs = _theta_by_hand(w, 1.0, eos[1])
    f_star = int(np.argmax(rs))
    r_star = rs[f_star]
    seen = set()
    # r/tol = 1/2 (theta 0), 3/4 (theta 1/2), 1 (theta 1) at the face with the
    # largest r, plus tolerances that put several faces strictly inside (0, 1)
    for tol in (2.0 * r_star, r_star / 0.75, r_star, float(np.median(rs)) / 0.7, r_star / 0.6):
        exp, th, _ = _blend_expected(w, tol, eos)
        got, st = _line_fluxes(w, "hybrid", tol, eos)
        assert np.array_equal(got, exp), (tol, th)
        assert st[0] == len(th) and st[1] == sum(1 for t in th if t > 0.0)
        seen.update("zero" if t == 0.0 else "one" if t == 1.0 else "mid" for t in th)
    assert {"zero", "one", "mid"} <= seen
    # the three exact points of the smoothstep on the face with the largest r
    assert _theta_by_hand(w, 2.0 * r_star, eos[1])[0][f_star] == 0.0
    assert _theta_by_hand(w, r_star / 0.75, eos[1])[0][f_star] == 0.5
    assert _theta_by_hand(w, r_star, eos[1])[0][f_star] == 1.0


def test_hybrid_theta_window_is_asymmetric_as_the_reference():
    # a sensor bump seen by cells i-2..i+3 of face i+1/2: faces whose window
    # reaches the bump switch, the next face over does not.  A window shifted
    # by one cell, or only one of the two fourth differences, fails here.
    ng, n = 6, 9
    w = np.tile([1.0, 0.2, 0.0, 0.0, 1.0], (n + 2 * ng, 1))
    bump = ng + 4
    w[bump, 0] = 1.5
    th, _ = _theta_by_hand(w, 1e-3)
    i_of = lambda f: ng - 1 + f
    for f, t in enumerate(th):
        inside = i_of(f) - 2 <= bump <= i_of(f) + 3
        assert (t == 1.0) == inside, (f, t)
    exp, _, _ = _blend_expected(w, 1e-3)
    got, st = _line_fluxes(w, "hybrid", 1e-3)
    assert np.array_equal(got, exp)
    assert st[1] == 6


def test_hybrid_theta_non_finite_sensor_gives_weno():
    # an infinite density makes r infinite on the faces whose windows hold it:
    # theta = 1 there (the shock path), whatever tol is.  A NaN density is
    # floored to 1e-300 by the fmax (NaN-ignoring) sensor, so r stays finite.
    ng = 6
    w = _tiny_line(7)
    w[ng + 3, 0] = np.inf
    th, rs = _theta_by_hand(w, 1e6)
    assert any(not math.isfinite(r) for r in rs)
    assert all(t == 1.0 for t, r in zip(th, rs) if not math.isfinite(r))
    got, st = _line_fluxes(w, "hybrid", 1e6)
    fw, _ = _line_fluxes(w, "weno", 1e6)
    for f, r in enumerate(rs):
        if not math.isfinite(r):
            assert np.array_equal(got[f], fw[f], equal_nan=True)
    assert st[1] == sum(1 for t in th if t > 0.0)
    w2 = _tiny_line(7)
    w2[ng + 3, 0] = np.nan
    _, rs2 = _theta_by_hand(w2, 1.0)
    assert all(math.isfinite(r) for r in rs2)
    exp2, th2, _ = _blend_expected(w2, max(rs2) / 0.8)
    got2, _ = _line_fluxes(w2, "hybrid", max(rs2) / 0.8)
    assert np.array_equal(got2, exp2, equal_nan=True)
