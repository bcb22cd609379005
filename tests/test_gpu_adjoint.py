"""Parity of the hand-written HIP adjoint (libemhip.so via the C ABI) with the
tape oracle, and with central finite differences of the HIP forward run.

* (dR/du)^T lam of one residual: 1e-11 relative L-inf against the tape on
  generic smooth data (limiter idle, and blending on many faces with theta
  differentiated at its admissibility constraint) and on the piecewise constant
  Riemann data, whose WENO-Z stencils are flat or partly flat to rounding
  (both sides take the same convention there: linear weights when all
  sub-stencils are flat, the weights held when some are; both recompute the
  same forward values, so both take the same branches).
* dJ/du_0 through SSP-RK3 steps with checkpoint segments recomputed: 1e-10
  relative L-inf against the tape gradient with the same frozen dt sequence.
* dJ/drho_0 against central finite differences (h = 1e-6 relative) of the HIP
  forward run with the frozen dt sequence: 1e-6 relative plus the FD
  rounding floor (100 eps |J| / h), on smooth data
  where every characteristic field varies (WENO-Z with eps = 1e-40 is not
  differentiable where a field is exactly constant along a stencil).
* Repeated backward sweeps are bit-identical.
"""
import numpy as np
import pytest

from _adjoint_cases import fill, pressure_window, riemann, smooth_state
from _oracle import OracleAdj

pytestmark = pytest.mark.gpu

RIEMANN_BC = [["outflow", "outflow"], ["outflow", "outflow"], ["periodic", "periodic"]]
BCS = [
    [["periodic", "periodic"]] * 3,
    RIEMANN_BC,
    [["reflect", ("dirichlet", (1.0, 0.1, 0.0, 0.0, 2.6))], ["outflow", "reflect"], ["periodic", "periodic"]],
]


@pytest.fixture(scope="module")
def em():
    import emergence_amd

    emergence_amd.load()
    return emergence_amd


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _pair(em, n, bc, u, positivity=True):
    o = OracleAdj(n, (0, 0, 0), (1, 1, 1), bc=bc, positivity=positivity)
    g = em.HipSolver(n, (0, 0, 0), (1, 1, 1), u, bc=bc, mode="weno", positivity=positivity)
    return o, g


@pytest.mark.parametrize("bci", range(len(BCS)))
@pytest.mark.parametrize("stage_dt", [1e-3, 3e-2, 0.8])  # 0.8: the limiter blends on many faces
def test_adjoint_rhs_matches_the_tape(em, bci, stage_dt):
    print(f"adjoint rhs smooth 2-D bc={bci} stage_dt={stage_dt}", flush=True)
    n = (24, 20, 1)
    bc = BCS[bci]
    o0 = OracleAdj(n, (0, 0, 0), (1, 1, 1), bc=bc)
    u = smooth_state(o0, 4)
    o, g = _pair(em, n, bc, u)
    rng = np.random.default_rng(9)
    lam = np.zeros_like(u)
    o.interior(lam)[...] = rng.standard_normal(o.interior(lam).shape)
    ref = o.vjp_rhs(stage_dt, u, lam)
    got = g.adjoint_rhs(stage_dt, lam)
    assert _rel(o.interior(got), o.interior(ref)) < 1e-11


def test_adjoint_rhs_on_riemann_data(em):
    print("adjoint rhs riemann 2-D", flush=True)
    n = (32, 32, 1)
    o0 = OracleAdj(n, (0, 0, 0), (1, 1, 1), bc=RIEMANN_BC)
    u = fill(o0, lambda i, j, k: riemann(i, j, 32, 0.55))
    o, g = _pair(em, n, RIEMANN_BC, u)
    rng = np.random.default_rng(2)
    lam = np.zeros_like(u)
    o.interior(lam)[...] = rng.standard_normal(o.interior(lam).shape)
    for stage_dt in (1e-3, 5e-2):
        ref = o.vjp_rhs(stage_dt, u, lam)
        got = g.adjoint_rhs(stage_dt, lam)
        assert _rel(o.interior(got), o.interior(ref)) < 1e-11, stage_dt


def test_adjoint_rhs_3d(em):
    print("adjoint rhs smooth 3-D periodic", flush=True)
    n = (10, 8, 12)
    o0 = OracleAdj(n, (0, 0, 0), (1, 1, 1))
    rng = np.random.default_rng(1)
    u = fill(o0, lambda i, j, k: [1.0 + 0.1 * np.sin(i + 2 * j + 3 * k), 0.2 + 0.05 * np.cos(j - k),
                                  -0.1 + 0.05 * np.sin(i * k), 0.1 * np.cos(i + k), 1.0 + 0.05 * np.sin(2 * i - j)])
    o, g = _pair(em, n, [["periodic", "periodic"]] * 3, u)
    lam = np.zeros_like(u)
    o.interior(lam)[...] = rng.standard_normal(o.interior(lam).shape)
    ref = o.vjp_rhs(1e-3, u, lam)
    got = g.adjoint_rhs(1e-3, lam)
    assert _rel(o.interior(got), o.interior(ref)) < 1e-11


def test_backward_matches_the_tape_gradient(em):
    print("adjoint backward vs tape, 2-D riemann + smooth, 7 steps, checkpoints every 3", flush=True)
    n = (16, 16, 1)
    o0 = OracleAdj(n, (0, 0, 0), (1, 1, 1), bc=RIEMANN_BC)
    u0 = fill(o0, lambda i, j, k: [v * (1.0 + 0.01 * np.sin(3 * i + j)) if q == 0 else v
                                     for q, v in enumerate(riemann(i, j, 16, 0.55))])
    o, g = _pair(em, n, RIEMANN_BC, u0)
    dts = g.forward(7, 3)
    uN = g.download()
    assert np.array_equal(o.interior(uN), o.interior(o.steps(u0, dts)))
    _, dj = pressure_window(o, uN, 0.3, 0.7)
    got = g.backward(dj)
    ref = o.gradient(u0, dts, dj)
    assert _rel(o.interior(got), o.interior(ref)) < 1e-10
    # the state is left as after the forward run, and a second sweep repeats
    assert np.array_equal(g.download(), uN)
    assert np.array_equal(g.backward(dj), got)


def test_backward_matches_central_differences(em):
    print("adjoint backward vs central FD, 2-D smooth 48^2, 12 steps", flush=True)
    n = (48, 48, 1)
    bc = RIEMANN_BC
    o0 = OracleAdj(n, (0, 0, 0), (1, 1, 1), bc=bc)
    u0 = smooth_state(o0, 21)
    g = em.HipSolver(n, (0, 0, 0), (1, 1, 1), u0, bc=bc, mode="weno")
    dts = g.forward(12, 5)
    J0, dj = pressure_window(o0, g.download(), 0.4, 0.7)
    grad = g.backward(dj)
    gs = 6
    # dJ/drho_0 at fixed velocity and pressure: direction (1, u, v, w, |u|^2/2)
    rng = np.random.default_rng(17)
    checks = 0
    for _ in range(8):
        i, j = int(rng.integers(8, 40)), int(rng.integers(8, 40))
        q = u0[gs + i, gs + j, 0]
        vel = q[1:4] / q[0]
        dirn = np.array([1.0, vel[0], vel[1], vel[2], 0.5 * np.dot(vel, vel)])
        ad = float(np.dot(grad[gs + i, gs + j, 0], dirn))
        h = 1e-6 * q[0]
        Js = []
        for sgn in (1.0, -1.0):
            up = u0.copy()
            up[gs + i, gs + j, 0] += sgn * h * dirn
            f = em.HipSolver(n, (0, 0, 0), (1, 1, 1), up, bc=bc, mode="weno")
            f.forward(12, 12, dts)
            Js.append(pressure_window(o0, f.download(), 0.4, 0.7)[0])
        fd = (Js[0] - Js[1]) / (2 * h)
        scale = np.abs(grad).max()
        # FD rounding: the two runs' own rounding errors, ~100 eps |J| / h
        noise = 100 * np.finfo(float).eps * abs(J0) / h
        assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1e-2 * scale) + noise, (i, j, fd, ad)
        checks += 1
    assert checks == 8


@pytest.mark.parametrize("case", ["periodic-540", "reflect-dirichlet-300", "outflow-263"])
def test_adjoint_rhs_long_pencils(em, case):
    # pencils cut into tiles of at most 128 cells: 540 periodic cells (5
    # cyclic tiles of 108) across 4-cell pencils (one tile each, the window
    # wrapping several times); 300 cells between a reflecting and a Dirichlet
    # boundary and 263 outflow cells (edge tiles holding the ghost positions,
    # uneven splits)
    print(f"adjoint rhs long pencils {case}", flush=True)
    if case == "periodic-540":
        n, bc = (540, 4, 1), [["periodic", "periodic"]] * 3
    elif case == "reflect-dirichlet-300":
        n = (300, 4, 1)
        bc = [["reflect", ("dirichlet", (1.0, 0.1, 0.0, 0.0, 2.6))], ["periodic", "periodic"], ["periodic", "periodic"]]
    else:
        n, bc = (263, 5, 1), [["outflow", "outflow"], ["outflow", "reflect"], ["periodic", "periodic"]]
    o0 = OracleAdj(n, (0, 0, 0), (1, 1, 1), bc=bc)
    u = smooth_state(o0, 8)
    o, g = _pair(em, n, bc, u)
    rng = np.random.default_rng(12)
    lam = np.zeros_like(u)
    o.interior(lam)[...] = rng.standard_normal(o.interior(lam).shape)
    for stage_dt in (1e-3, 0.2):
        ref = o.vjp_rhs(stage_dt, u, lam)
        got = g.adjoint_rhs(stage_dt, lam)
        assert _rel(o.interior(got), o.interior(ref)) < 1e-11, stage_dt


@pytest.mark.parametrize("case", ["outflow-reflect-300", "periodic-263"])
def test_adjoint_results_do_not_depend_on_the_tiling(em, case):
    # the same VJP with tiles of at most 128 cells (the default), 37 and 10
    # cells, and with whole pencils: every cotangent sum runs in a fixed order
    # of global positions, so the bits are the same
    print(f"adjoint tiling independence {case}", flush=True)
    if case == "outflow-reflect-300":
        n, bc = (300, 12, 1), [["outflow", "reflect"], ["reflect", "outflow"], ["periodic", "periodic"]]
    else:
        n, bc = (263, 11, 1), [["periodic", "periodic"]] * 3
    o0 = OracleAdj(n, (0, 0, 0), (1, 1, 1), bc=bc)
    u = smooth_state(o0, 5)
    g = em.HipSolver(n, (0, 0, 0), (1, 1, 1), u, bc=bc, mode="weno")
    rng = np.random.default_rng(4)
    lam = np.zeros_like(u)
    o0.interior(lam)[...] = rng.standard_normal(o0.interior(lam).shape)
    ref = g.adjoint_rhs(0.2, lam)
    assert np.abs(o0.interior(ref)).max() > 0.0
    for cells in (37, 10, 300):
        g.set_adjoint_tile(cells)
        assert np.array_equal(g.adjoint_rhs(0.2, lam), ref), cells


def test_backward_over_kept_states_equals_the_checkpointed_sweep(em):
    # checkpoint_every = 0 keeps every step's u_n, u1 and u2 on the device:
    # the backward reverses the steps over them without recomputing, and the
    # gradient is the checkpointed sweep's bit for bit (with CFL dts recorded
    # on the device and with a frozen sequence); the tape bounds both
    print("adjoint backward over kept states vs checkpoints, 2-D riemann 24x20, 9 steps", flush=True)
    from emergence_amd import EmError

    n = (24, 20, 1)
    o = OracleAdj(n, (0, 0, 0), (1, 1, 1), bc=RIEMANN_BC)
    u0 = fill(o, lambda i, j, k: [v * (1.0 + 0.01 * np.sin(2 * i + 3 * j)) if q == 0 else v
                                    for q, v in enumerate(riemann(i, j, 24, 0.55))])
    g = em.HipSolver(n, (0, 0, 0), (1, 1, 1), u0, bc=RIEMANN_BC, mode="weno")
    dts = g.forward(9, 4)
    uN = g.download()
    _, dj = pressure_window(o, uN, 0.3, 0.7)
    ref = g.backward(dj)
    g.upload(u0)
    assert np.array_equal(g.forward(9, 0), dts)  # CFL dts recorded on the device
    assert np.array_equal(g.download(), uN)
    got = g.backward(dj)
    assert np.array_equal(got, ref)
    assert np.array_equal(g.download(), uN)
    assert np.array_equal(g.backward(dj), got)
    g.upload(u0)
    g.forward(9, 0, dts=dts)  # frozen sequence, kept buffers reused
    assert np.array_equal(g.backward(dj), ref)
    assert _rel(o.interior(ref), o.interior(o.gradient(u0, dts, dj))) < 1e-10
    # the device-array entry point: the same bits after either forward run
    dj_d = em.DeviceArray.from_host(g.to_soa(dj))
    out_d = em.DeviceArray(dj_d.n)
    assert np.array_equal(o.interior(g.from_soa(dj_d.to_host())), o.interior(dj))
    g.backward_device(dj_d, out_d)
    assert np.array_equal(g.from_soa(out_d.to_host()), ref)
    g.upload(u0)
    g.forward(9, 4)
    g.backward_device(dj_d, out_d)
    assert np.array_equal(g.from_soa(out_d.to_host()), ref)
    with pytest.raises(ValueError):
        g.backward_device(dj_d, em.DeviceArray(7))
    with pytest.raises(EmError):
        g.forward(9, -1)
