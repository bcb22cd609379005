"""Parity of the HIP two-phase Kapila five-equation path (libemhip.so,
em_fiveeq_* via the C ABI) with the oracle's restatement of
fluid/crates/compressible/src/multiphase.rs on identical inputs.

Bars:
* Weno and Central modes: bit-identical residuals (including the
  non-conservative (alpha + K) div u term), states after steps, dt (with the
  dilatation-rate limit) and counters, with every boundary kind;
* Hybrid mode: the sensor's ln is the device's, which differs from glibc's
  by an ulp on some inputs (glibc's own log is not correctly rounded on
  about 0.06% of the probe's sensor values), so the blend weight of a face
  in the blend band can differ in its last bits: identical counters, the
  per-stage residual within 1e-13 of the flux scale, the reference's
  `interface` probe state within 1e-12 per component of the oracle's (whose
  state reproduces the probe's recorded hash), and the probe's own physics
  check (pressure and velocity equilibrium across the moving interface) to
  round-off.
"""
import json
import os

import numpy as np
import pytest

from _oracle import FIVEEQ_WATER_AIR, OracleGen

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cpu_validate_golden.json")))
PHASES = (FIVEEQ_WATER_AIR[0:4], FIVEEQ_WATER_AIR[4:8])


@pytest.fixture(scope="module")
def em():
    import emergence_amd

    emergence_amd.load()
    return emergence_amd


def _cons(w):
    """to_cons (multiphase.rs:162-173) for the water/air pair, by hand."""
    g1, pi1, _, _, g2, pi2, _, _, _ = FIVEEQ_WATER_AIR
    a = w[..., 6]
    d = a / (g1 - 1.0) + (1.0 - a) / (g2 - 1.0)
    b = g1 * pi1 * a / (g1 - 1.0) + g2 * pi2 * (1.0 - a) / (g2 - 1.0)
    rho = w[..., 0] + w[..., 1]
    e = w[..., 5] * d + b + 0.5 * rho * (w[..., 2] ** 2 + w[..., 3] ** 2 + w[..., 4] ** 2)
    return np.stack([w[..., 0], w[..., 1], rho * w[..., 2], rho * w[..., 3], rho * w[..., 4], e, a], axis=-1)


def _drop(n, ng=6, vel=(0.1, 0.0, 0.0), seed=0, noise=0.05):
    """A water drop in air with a seeded pressure/velocity perturbation of
    relative size `noise` (so that every field varies and the limiter acts)."""
    rng = np.random.default_rng(seed)
    shape = tuple(v + 2 * ng if v > 1 else 1 for v in n) + (7,)
    u = np.zeros(shape)
    g = [ng if v > 1 else 0 for v in n]
    idx = np.meshgrid(*[np.arange(v) for v in n], indexing="ij")
    x = [(idx[d] + 0.5) / n[d] for d in range(3)]
    r = np.sqrt(sum((x[d] - 0.5) ** 2 for d in range(3) if n[d] > 1))
    alpha = np.clip(0.5 * (1.0 + np.tanh((0.25 - r) / (1.5 / max(n)))), 1e-6, 1 - 1e-6)
    w = np.zeros(tuple(n) + (7,))
    w[..., 0] = alpha * 1000.0
    w[..., 1] = (1.0 - alpha) * 1.2
    for d in range(3):
        w[..., 2 + d] = vel[d] * (1.0 + 4.0 * noise * rng.standard_normal(tuple(n))) if n[d] > 1 else 0.0
    w[..., 5] = 1e5 * (1.0 + noise * rng.standard_normal(tuple(n)))
    w[..., 6] = alpha
    u[g[0]:g[0] + n[0], g[1]:g[1] + n[1], g[2]:g[2] + n[2]] = _cons(w)
    return u


def _pair(em, n, u, mode, bc=None, positivity=True, hi=(1.0, 1.0, 1.0), cfl=0.5):
    g = em.HipFiveEq(n, (0, 0, 0), hi, u, phases=PHASES, bc=bc, mode=mode, positivity=positivity, cfl=cfl)
    o = OracleGen("fiveeq", n, (0, 0, 0), hi)
    o.set_u(u)
    o.set_opts(mode=mode, positivity=positivity, cfl=cfl)
    if bc:
        for d in range(3):
            for s in range(2):
                b = bc[d][s]
                if isinstance(b, str):
                    o.set_bc(d, s, b)
                else:
                    o.set_bc(d, s, b[0], b[1])
    return g, o


FAR = [float(v) for v in _cons(np.array([1e-3, 1.2, 0.0, 0.0, 0.0, 1e5, 1e-6]))]
BCS = [
    None,
    [["outflow", "outflow"], ["reflect", "reflect"], ["periodic", "periodic"]],
    [[("dirichlet", FAR), ("characteristic", FAR)], [("characteristic", FAR), "reflect"], ["outflow"] * 2],
]


@pytest.mark.parametrize("mode", ["weno", "central"])
@pytest.mark.parametrize("bci", range(len(BCS)))
def test_fiveeq_rhs_is_bit_identical(em, mode, bci):
    n = (24, 20, 1) if bci < 2 else (17, 12, 9)
    u = _drop(n, seed=bci)
    g, o = _pair(em, n, u, mode, BCS[bci])
    for sdt in (1e-6, 4e-4):  # the larger stage dt makes the limiter act
        rg, sg = g.rhs(sdt)
        ro, so = o.rhs(sdt)
        gi, oi = g.interior(rg), o.interior(ro)
        assert np.array_equal(gi, oi), (mode, bci, sdt, float(np.abs(gi - oi).max() / np.abs(oi).max()))
        assert sg == so


def test_fiveeq_limiter_and_fallback_counters_are_exercised(em):
    n = (24, 20, 1)
    g, o = _pair(em, n, _drop(n, seed=3), "weno", BCS[1])
    _, so = o.rhs(4e-4)
    _, sg = g.rhs(4e-4)
    assert sg == so
    assert so[3] > 0  # limited faces


@pytest.mark.parametrize("bci", range(len(BCS)))
def test_fiveeq_dt_is_bit_identical(em, bci):
    n = (24, 20, 1) if bci < 2 else (17, 12, 9)
    g, o = _pair(em, n, _drop(n, vel=(40.0, -30.0, 20.0), seed=bci), "weno", BCS[bci])
    assert g.dt_cfl() == o.dt_cfl()


@pytest.mark.parametrize("mode", ["weno", "central"])
def test_fiveeq_steps_are_bit_identical(em, mode):
    n = (20, 16, 12)
    g, o = _pair(em, n, _drop(n, seed=5), mode)
    for _ in range(4):
        dt = o.dt_cfl()
        assert g.dt_cfl() == dt
        o.step(dt)
        g.step(dt)
    gi, oi = g.interior(g.download()), o.interior(o.get_u())
    assert np.array_equal(gi, oi)
    assert g.counters() == o.stats()[1:]


def test_fiveeq_run_to_takes_the_oracles_steps(em):
    n = (32, 24, 1)
    g, o = _pair(em, n, _drop(n, vel=(30.0, 10.0, 0.0), seed=6), "weno", BCS[1])
    t_end = 1.2e-4
    ko = o.run_to(t_end)
    kg = g.run_to(t_end)
    assert kg == ko > 2
    assert g.t == o.t
    assert np.array_equal(g.interior(g.download()), o.interior(o.get_u()))


def _component_max(a):
    return np.abs(a).reshape(-1, 7).max(axis=0)


def _flux_scale(o, ui, n, hi=(1.0, 1.0, 1.0)):
    """Per component: sum over active axes of max |f_d(w)| / dx_d, the size of
    the face fluxes whose differences make the residual (multiphase.rs:319-331)."""
    w = o.to_prim(ui.reshape(-1, 7))
    q = ui.reshape(-1, 7)
    out = np.zeros(7)
    for d in range(3):
        if n[d] <= 1:
            continue
        ud = w[:, 2 + d]
        f = q * ud[:, None]
        f[:, 2 + d] += w[:, 5]
        f[:, 5] = (q[:, 5] + w[:, 5]) * ud
        f[:, 6] = w[:, 6] * ud
        out += np.abs(f).max(axis=0) * n[d] / hi[d]
    return out


@pytest.mark.parametrize("bci", [0, 2])
def test_fiveeq_hybrid_rhs_within_1e13_of_the_flux_scale(em, bci):
    """Hybrid residual against the oracle.  The sensor's ln differs from
    glibc's by an ulp on some inputs, which moves the blend weight of a face
    in the blend band by up to ~1e-10 and its flux by that fraction of the
    WCNS-central difference: a rounding-level change of the face flux.  The
    bar is therefore per component 1e-13 of the flux scale sum_d max|f_d|/dx_d
    (a wrong sensor or blend misses it by orders of magnitude), with the same
    faces on the shock path as the oracle."""
    n = (48, 40, 1) if bci < 2 else (32, 28, 20)
    # a smooth drop: faces far from the interface take the central flux,
    # faces across it the shock path, a band between them blends
    u = _drop(n, seed=bci, noise=0.0)
    g, o = _pair(em, n, u, "hybrid", BCS[bci])
    rg, sg = g.rhs(1e-6)
    ro, so = o.rhs(1e-6)
    assert sg == so  # same faces on the shock path and limited
    assert 0 < sg[1] < sg[0]  # the blend is exercised
    gi, oi = g.interior(rg), o.interior(ro)
    err = _component_max(gi - oi)
    scale = _flux_scale(o, o.interior(u), n)
    print("hybrid residual |GPU - oracle| / flux scale per component:",
          np.divide(err, scale, out=np.zeros_like(err), where=scale > 0))
    assert np.all(err <= 1e-13 * scale)


@pytest.mark.parametrize("case", [c for c in GOLD["cases"] if c["args"][0] == "interface" and c["args"][2] <= 20],
                         ids=lambda c: "-".join(str(a) for a in c["args"]))
def test_interface_probe_on_gpu(em, case):
    """The reference's `interface` probe (examples/cpu_validate.rs:100-146):
    3-D periodic, FiveEq (Kapila), Hybrid, dt = 1e-5."""
    _, n, steps = case["args"]
    o = OracleGen("fiveeq", (n, n, n), (0, 0, 0), (1, 1, 1))
    o.init_interface()  # initial state only
    u0 = o.get_u()
    g = em.HipFiveEq((n, n, n), (0, 0, 0), (1, 1, 1), u0, phases=PHASES, mode="hybrid")
    o.set_opts(mode="hybrid")
    for _ in range(steps):
        g.step(1e-5)
        o.step(1e-5)
    assert f"{o.hash():016x}" == case["hash"]  # the oracle's state is the reference's
    gi, oi = g.interior(g.download()), o.interior(o.get_u())
    assert g.counters() == o.stats()[1:]
    # per component, relative to the larger of the component's size and the
    # most the run could change it (steps x dt x flux scale): the momenta of
    # this probe stay at rounding level, so their size is no scale
    scale = np.maximum(_component_max(oi), steps * 1e-5 * _flux_scale(o, o.interior(u0), (n, n, n)))
    rel = _component_max(gi - oi) / scale
    print(f"interface {n} {steps}: GPU vs oracle per component", rel)
    assert rel.max() <= 1e-12
    w = o.to_prim(gi.reshape(-1, 7))
    assert np.abs(w[:, 5] / 1e5 - 1.0).max() < 1e-12  # pressure equilibrium
    assert np.abs(w[:, 2] - 0.1).max() < 1e-11  # velocity equilibrium
