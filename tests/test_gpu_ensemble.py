"""The ensemble (many independent uniform grids, one persistent launch per
pass) against the single-grid HIP solver, which is itself bit-identical to
the oracle:
* forward steps of every instance bit-identical to running it alone, in
  batched (after the tile autotuner) and per-instance launch modes;
* the ensemble adjoint (batched recompute, each instance's VJP launched on
  its own) bit-identical to each instance's own backward sweep with the same
  frozen dt sequence;
* dJ/dMach of the shock-bubble workload against central differences.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BC = [["outflow", "outflow"], ["outflow", "outflow"], ["periodic", "periodic"]]


@pytest.fixture(scope="module")
def em():
    import emergence_amd

    emergence_amd.load()
    return emergence_amd


def _instances(em, sizes, seed=1, near=False):
    # near: the shock starts two cells upstream of the bubble, so that J
    # (the bubble's mean pressure) depends on the Mach number within a few
    # steps
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        mach, r, dens = rng.uniform(1.1, 2.0), rng.uniform(0.1, 0.3), rng.uniform(0.1, 0.5)
        sx = 0.5 - r - 2.0 / n if near else 0.15
        u, du = em.shock_bubble_2d(n, mach, r, dens, shock_x=sx)
        out.append((n, mach, r, dens, sx, u, du))
    return out


def _interior(u, n, ng=6):
    return u[ng:ng + n, ng:ng + n, 0]


@pytest.mark.parametrize("batched", [True, False])
def test_ensemble_steps_equal_single_solvers(em, batched):
    print(f"ensemble forward vs single solvers, batched={batched}", flush=True)
    inst = _instances(em, [64, 96, 32, 128])
    ens = em.HipEnsemble([(n, n, 1) for n, *_ in inst], [x[5] for x in inst], bc=BC, batched=batched)
    if batched:
        # the autotuner picks a tile shape and leaves the states untouched
        cfg = ens.autotune(1)
        assert 0 <= cfg < 10
    ens.run_steps(6)
    for i, (n, _, _, _, _, u, _) in enumerate(inst):
        g = em.HipSolver((n, n, 1), (0, 0, 0), (1, 1, 1), u, bc=BC, mode="weno", cfl=0.5)
        g.run_steps(6)
        assert np.array_equal(_interior(ens.download(i), n), _interior(g.download(), n)), i
        assert ens.time(i) == g.t


def test_ensemble_adjoint_equals_single_backward(em):
    print("ensemble adjoint vs single backward", flush=True)
    inst = _instances(em, [48, 64, 32], seed=3)
    ens = em.HipEnsemble([(n, n, 1) for n, *_ in inst], [x[5] for x in inst], bc=BC)
    ens.forward(7, 3)
    djs, refs = [], []
    for i, (n, _, _, _, _, u, _) in enumerate(inst):
        dj = np.zeros_like(u)
        _interior(dj, n)[..., 4] = 1.0 / (n * n)
        djs.append(dj)
        g = em.HipSolver((n, n, 1), (0, 0, 0), (1, 1, 1), u, bc=BC, mode="weno", cfl=0.5)
        dts = g.forward(7, 3)
        assert np.array_equal(dts, ens.dts(i)), i
        assert np.array_equal(_interior(g.download(), n), _interior(ens.download(i), n)), i
        refs.append(g.backward(dj))
    got = ens.backward(djs)
    for i, (n, *_) in enumerate(inst):
        assert np.array_equal(_interior(got[i], n), _interior(refs[i], n)), i


def test_mach_gradient_matches_central_differences(em):
    print("ensemble dJ/dMach vs central FD", flush=True)
    inst = _instances(em, [32, 48], seed=7, near=True)
    steps = 10

    def bubble_mean_p(u, n, r):
        x = (np.arange(n) + 0.5) / n
        X, Y = np.meshgrid(x, x, indexing="ij")
        m = (X - 0.5) ** 2 + (Y - 0.5) ** 2 < r * r
        q = _interior(u, n)
        vel = q[..., 1:4] / q[..., :1]
        p = 0.4 * (q[..., 4] - 0.5 * q[..., 0] * np.sum(vel * vel, -1))
        g = np.zeros_like(u)
        gi = _interior(g, n)
        gi[..., 0] = 0.4 * 0.5 * np.sum(vel * vel, -1)
        gi[..., 1:4] = -0.4 * vel
        gi[..., 4] = 0.4
        gi *= (m / m.sum())[..., None]
        return float(p[m].mean()), g

    ens = em.HipEnsemble([(n, n, 1) for n, *_ in inst], [x[5] for x in inst], bc=BC)
    ens.forward(steps, 4)
    djs = [bubble_mean_p(ens.download(i), n, r)[1] for i, (n, _, r, *_) in enumerate(inst)]
    du0 = ens.backward(djs)
    for i, (n, mach, r, dens, sx, u, du) in enumerate(inst):
        ad = float(np.sum(du0[i] * du))
        dts = ens.dts(i)
        h = 1e-6
        js = []
        for s in (1.0, -1.0):
            up, _ = em.shock_bubble_2d(n, mach + s * h, r, dens, shock_x=sx)
            g = em.HipSolver((n, n, 1), (0, 0, 0), (1, 1, 1), up, bc=BC, mode="weno", cfl=0.5)
            g.forward(steps, steps, dts)
            js.append(bubble_mean_p(g.download(), n, r)[0])
        fd = (js[0] - js[1]) / (2 * h)
        assert abs(ad) > 1e-3, (i, ad)  # the shock has reached the bubble
        noise = 100 * np.finfo(float).eps * abs(js[0]) / h
        assert abs(fd - ad) <= 1e-5 * abs(ad) + noise, (i, fd, ad)


def _oracle_run(n, u, steps):
    from _oracle import OracleSolver

    o = OracleSolver((n, n, 1), (0, 0, 0), (1, 1, 1))
    o.set_u(u)
    o.set_opts(mode="weno", cfl=0.5)
    for d in range(2):
        for s in range(2):
            o.set_bc(d, s, "outflow")
    dts = []
    for _ in range(steps):
        dt = o.dt_cfl()
        dts.append(dt)
        o.step(dt)
    return o, np.array(dts)


def test_ensemble_against_the_oracle(em):
    """The ensemble workload (shock-bubble instances of uneven size, one with
    the shock two cells from the bubble) against the oracle: every instance's
    dt sequence and state after the batched forward run are bit-identical to
    the restated reference solver's, and the ensemble adjoint equals the
    oracle's tape gradient with the frozen dt sequence to 1e-10 (relative to
    its largest entry)."""
    from _oracle import OracleAdj

    print("ensemble vs oracle: forward bit-identical, adjoint vs tape", flush=True)
    inst = _instances(em, [96, 128, 64], seed=11, near=True)
    steps = 3
    ens = em.HipEnsemble([(n, n, 1) for n, *_ in inst], [x[5] for x in inst], bc=BC)
    ens.forward(steps, 2)
    djs = []
    for i, (n, _, r, _, _, u, _) in enumerate(inst):
        o, dts = _oracle_run(n, u, steps)
        assert np.array_equal(ens.dts(i), dts), i
        assert np.array_equal(_interior(ens.download(i), n), _interior(o.get_u(), n)), i
        dj = np.zeros_like(u)
        x = (np.arange(n) + 0.5) / n
        X, Y = np.meshgrid(x, x, indexing="ij")
        _interior(dj, n)[..., 4] = (((X - 0.5) ** 2 + (Y - 0.5) ** 2) < r * r) / (n * n)
        djs.append(dj)
    got = ens.backward(djs)
    for i in (0, 2):  # the tape gradient of the 128^2 instance takes ~20 s: forward only
        n, _, _, _, _, u, _ = inst[i]
        ref = OracleAdj((n, n, 1), (0, 0, 0), (1, 1, 1), bc=BC).gradient(u, ens.dts(i), djs[i])
        gi, ri = _interior(got[i], n), _interior(ref, n)
        err = np.abs(gi - ri).max() / np.abs(ri).max()
        print(f"instance {i} ({n}^2): ensemble adjoint vs tape {err:.2e}", flush=True)
        assert err < 1e-10, (i, err)
