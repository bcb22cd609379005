"""The FP32 variant of the uniform-grid solver (em_solver_create_f32, the
reference's GpuSolver<f32>, compressible-cuda/src/lib.rs:26-48) against the
FP64 oracle on identical inputs.

The reference has no FP32 CPU solver: its own check of the f32 path compares
it with the f64 result (compressible-cuda/examples/gpu_validate.rs:7-8,
45-60; examples/gpu_amr_validate.rs:86-88).  The bars here, per test, are
set from FP32's unit roundoff (2^-24 = 6e-8) and the condition of each
quantity, and a wrong term (a dropped flux, a transposed axis, a double
constant where the kernels run in float) misses them by orders of
magnitude:
* per-stage residual on a smooth 3-D state: relative L-inf <= 2e-5 in
  Central mode (measured 8.7e-6) and <= 1e-4 in Weno and Hybrid mode
  (measured 3.4e-5: the WENO-Z weights amplify rounding), with the same
  shock-path, fallback and limiter counts as the oracle;
* a uniform flow gives a residual of exactly zero (every face flux equal);
* 20 Central steps of the smooth 3-D wave: relative L1 density difference
  <= 1e-6 (measured 3.4e-7), mass conserved to 1e-5;
* 20 Weno steps of the 2-D Riemann problem with the oracle's dt: relative
  L1 density difference <= 2e-4 (measured 5.6e-5).  Its piecewise-constant
  data has sub-stencils that are exactly flat next to a jump: there the
  reference's FP32 WENO-Z ratio tau / (beta + 1e-30) squares past FLT_MAX
  (kernels.cu:263 uses that epsilon), and the midpoint takes the first-order
  fallback, which FP64 does not.  So the FP32 run has more fallbacks, and
  its CFL dt (a max over cells at the discontinuities) drifts by up to 5%
  from FP64's while the L1 difference stays small; dt_cfl of the initial
  state agrees to 1e-6.
The FP32 handle refuses the adjoint entry points.
"""
import numpy as np
import pytest

from _oracle import OracleSolver

pytestmark = pytest.mark.gpu

RIEMANN_BC = [["outflow", "outflow"], ["outflow", "outflow"], ["periodic", "periodic"]]


@pytest.fixture(scope="module")
def em():
    import emergence_amd

    emergence_amd.load()
    return emergence_amd


def _pair(em, case, n, lo, hi, mode, bc=None, params=None, precise=None, cfl=0.5):
    u = em.case_init(case, n, lo, hi, 6, params)
    g = em.HipSolver(n, lo, hi, u, bc=bc, mode=mode, cfl=cfl, real="f32", precise=precise)
    o = OracleSolver(n, lo, hi, ng=6)
    o.set_u(u)
    o.set_opts(mode=mode, cfl=cfl)
    if bc:
        for d in range(3):
            for s in range(2):
                o.set_bc(d, s, bc[d][s])
    return g, o, u


def _interior(a, n, ng=6):
    g = [ng if v > 1 else 0 for v in n]
    return a[g[0]:g[0] + n[0], g[1]:g[1] + n[1], g[2]:g[2] + n[2]]


@pytest.mark.parametrize("precise", [False, True])
@pytest.mark.parametrize("mode", ["weno", "hybrid", "central"])
def test_f32_rhs_on_a_smooth_state(em, mode, precise):
    n = (28, 24, 20)
    g, o, _ = _pair(em, 1, n, (0, 0, 0), (2 * np.pi,) * 3, mode, params=[0.3, float(0x5EED)], precise=precise)
    dt = o.dt_cfl()
    rg, sg = g.rhs(dt)
    ro, so = o.rhs(dt)
    gi, oi = _interior(rg, n), _interior(ro, n)
    rel = np.abs(gi - oi).max() / np.abs(oi).max()
    print(f"f32 {mode} precise={precise}: residual relative L-inf {rel:.3e}; counters {sg} vs {so}")
    assert rel <= (2e-5 if mode == "central" else 1e-4)
    assert rel > 0.0  # the path really computes in FP32
    assert sg == so


def test_f32_uniform_flow_has_zero_residual(em):
    n = (16, 12, 10)
    u = em.case_init(1, n, (0, 0, 0), (1, 1, 1), 6, [0.3, 1.0])
    w = np.array([1.3, 0.4, -0.2, 0.1, 0.9])  # rho, u, v, w, p
    u[...] = [w[0], w[0] * w[1], w[0] * w[2], w[0] * w[3], w[4] / 0.4 + 0.5 * w[0] * (w[1:4] ** 2).sum()]
    for mode in ("weno", "hybrid", "central"):
        g = em.HipSolver(n, (0, 0, 0), (1, 1, 1), u, mode=mode, real="f32")
        r, _ = g.rhs(1e-3)
        assert np.array_equal(_interior(r, n), np.zeros_like(_interior(r, n))), mode


def test_f32_riemann_steps_track_the_oracle(em):
    n = (64, 64, 1)
    g, o, _ = _pair(em, 0, n, (0, 0, 0), (1, 1, 1), "weno", bc=RIEMANN_BC, params=[0.8, 0.8])
    ratios = []
    for _ in range(20):
        dt = o.dt_cfl()
        ratios.append(g.dt_cfl() / dt - 1.0)
        o.step(dt)
        g.step(dt)
    print("f32 Riemann dt_cfl relative differences:", " ".join(f"{r:.1e}" for r in ratios))
    assert abs(ratios[0]) <= 1e-6
    assert max(abs(r) for r in ratios) <= 0.05
    rg, ro = _interior(g.download(), n)[..., 0], _interior(o.get_u(), n)[..., 0]
    l1 = np.abs(rg - ro).sum() / np.abs(ro).sum()
    cg, co = g.counters(), o.stats()[1:]
    print(f"f32 Riemann 20 steps: relative L1 density {l1:.3e}, max {np.abs(rg - ro).max():.3e}; "
          f"counters {cg} vs {co}")
    assert l1 <= 2e-4
    assert cg[1] >= co[1]  # first-order fallbacks


def test_f32_smooth_steps_conserve_and_track(em):
    n = (24, 24, 24)
    g, o, u0 = _pair(em, 2, n, (0, 0, 0), (1, 1, 1), "central")
    dt = 0.02 / 20
    for _ in range(20):
        o.step(dt)
        g.step(dt)
    ug, uo = _interior(g.download(), n), _interior(o.get_u(), n)
    l1 = np.abs(ug[..., 0] - uo[..., 0]).sum() / np.abs(uo[..., 0]).sum()
    m0 = _interior(u0, n)[..., 0].sum()
    print(f"f32 smooth 20 steps: relative L1 density {l1:.3e}; mass drift {abs(ug[..., 0].sum() - m0) / m0:.3e}")
    assert l1 <= 1e-6
    assert abs(ug[..., 0].sum() - m0) <= 1e-5 * m0


def test_f32_run_steps_and_run_to(em):
    n = (32, 32, 32)
    u = em.case_init(1, n, (0, 0, 0), (2 * np.pi,) * 3, 6, [0.3, float(0x5EED)])
    g = em.HipSolver(n, (0, 0, 0), (2 * np.pi,) * 3, u, mode="weno", real="f32")
    g.autotune(1)
    ms = g.run_steps(3, timed=True)
    assert all(v > 0 for v in ms)
    k = g.run_to(g.t + 0.05)
    assert k >= 1
    assert np.isfinite(g.download()).all()


def test_f32_handle_refuses_the_adjoint(em):
    from emergence_amd import _lib

    n = (16, 16, 1)
    u = em.case_init(0, n, (0, 0, 0), (1, 1, 1), 6, [0.8, 0.8])
    g = em.HipSolver(n, (0, 0, 0), (1, 1, 1), u, mode="weno", bc=RIEMANN_BC, real="f32")
    rc = g.lib.em_solver_forward(g.h, 2, 1, None, None)
    assert rc == -1
    assert b"FP64" in _lib.load().em_last_error()
