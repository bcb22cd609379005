"""Parity of the HIP block AMR (libemhip.so via the C ABI) with the oracle's
restated CPU `Amr` (global SSP-RK3 step) on identical hierarchies and data.

Bars (precise variant, no FMA contraction):
* WCNS-everywhere and central modes: bit-identical residuals, restricted
  ancestors, dt, states after steps (with and without regrids), hierarchy
  edits (block ids, levels, leaf flags) and face counters;
* hybrid mode: the theta sensor takes logarithms, which the device library
  rounds differently from the host libm in the last place, so residuals are
  held to 1e-13 relative L-inf and states to 1e-12 after a few steps;
* wavelet details (logarithms again) to 1e-12 relative, and the regrid
  decisions they drive identical;
* batched (one launch per phase) and per-leaf launches bit-identical.
"""
import numpy as np
import pytest

from _oracle import OracleAmr

pytestmark = pytest.mark.gpu

RIEMANN_BC = [["outflow", "outflow"], ["outflow", "outflow"], ["periodic", "periodic"]]


@pytest.fixture(scope="module")
def em():
    import emergence_amd

    emergence_amd.load()
    return emergence_amd


def _from_oracle(em, o, **kw):
    lv, idx, leaf = o.blocks()
    return em.HipAmr(o.base, o.b, o.levels, o.lo, o.hi, blocks=(lv, idx, leaf, o.get_u()), **kw)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _same_tree(g, o):
    gl, gi, gf = g.blocks()
    ol, oi, of = o.blocks()
    assert np.array_equal(gl, ol) and np.array_equal(gi, oi) and np.array_equal(gf, of)


def _riemann(mode, levels=3, **kw):
    return OracleAmr("riemann2d", [0.55, 0.55], (4, 4, 0), 8, levels, (0, 0, 0), (1, 1, 1), eps=1e-3,
                     bc=RIEMANN_BC, mode=mode, **kw)


@pytest.mark.parametrize("mode", ["weno", "central"])
def test_rhs_bit_identical_2d(em, mode):
    print(f"amr rhs 2-D riemann mode={mode}", flush=True)
    o = _riemann(mode)
    lv, _, leaf = o.blocks()
    assert len(set(lv[leaf])) >= 2
    g = _from_oracle(em, o, bc=RIEMANN_BC, mode=mode)
    assert np.array_equal(g.download(), o.get_u())
    rg, sg = g.rhs(1e-3)
    ro, so = o.compute_rhs(1e-3)
    assert np.array_equal(rg, ro)
    assert sg == so
    # compute_rhs restricted the edge ancestors on both sides
    assert np.array_equal(g.download(), o.get_u())


def test_rhs_hybrid_2d(em):
    print("amr rhs 2-D riemann hybrid", flush=True)
    o = _riemann("hybrid")
    g = _from_oracle(em, o, bc=RIEMANN_BC, mode="hybrid")
    rg, sg = g.rhs(1e-3)
    ro, so = o.compute_rhs(1e-3)
    assert _rel(rg, ro) < 1e-13
    assert sg[0] == so[0]
    assert np.array_equal(g.download(), o.get_u())


def test_dt_and_steps_without_regrid(em):
    print("amr steps 2-D riemann weno, regrid off", flush=True)
    o = _riemann("weno")
    g = _from_oracle(em, o, bc=RIEMANN_BC, mode="weno")
    o.set_next_regrid(1 << 40)
    g.set_next_regrid(1 << 40)
    for _ in range(3):
        dt = o.dt_cfl()
        assert g.dt_cfl() == dt
        o.step(dt)
        g.step(dt)
    assert g.t == o.t
    assert np.array_equal(g.download(), o.get_u())
    _same_tree(g, o)


SB2D_BC = [["outflow", "outflow"], ["reflect", "reflect"], ["periodic", "periodic"]]


def test_steps_with_regrids_2d(em):
    # a Mach-3 shock crossing a light bubble: the hierarchy changes after
    # about 50 steps (regrid every 2 steps)
    print("amr steps 2-D shock-bubble weno, regrid every 2", flush=True)
    o = OracleAmr("shock_bubble", [3.0, 0.3, 0.6, 0.25, 0.0, 0.12, 0.1], (8, 4, 0), 8, 3, (0, 0, 0), (1, 0.5, 1),
                  eps=1e-2, bc=SB2D_BC, mode="weno", regrid_every=2)
    g = _from_oracle(em, o, bc=SB2D_BC, mode="weno", eps=1e-2, regrid_every=2)
    sizes = set()
    for k in range(56):
        dt = o.dt_cfl()
        assert g.dt_cfl() == dt, k
        o.step(dt)
        g.step(dt)
        _same_tree(g, o)
        sizes.add(o.nblocks)
        if k % 8 == 7:
            assert np.array_equal(g.download(), o.get_u()), k
    assert len(sizes) >= 2, "the hierarchy should have changed"
    assert np.array_equal(g.download(), o.get_u())
    assert g.counters() == o.stats()[1:]


def test_mass_coarsening_regrids(em):
    # Flattened data (deviations from the mean scaled by 1e-7) coarsen the 28
    # level-1 parents in one regrid (70 neighbouring pairs) and the 12 level-0
    # ones in the next: the restriction of each coarsening sees the earlier
    # coarsenings of its neighbours (their children gone from the index), as
    # in the reference's sequential loop.  Trees and data stay bit-identical.
    print("amr mass coarsening, 2-D riemann flattened", flush=True)
    o = _riemann("weno")
    _, _, leaf = o.blocks()
    u = o.get_u()
    ref = u[leaf].reshape(-1, 5).mean(0)
    o.set_u(ref + 1e-7 * (u - ref))
    g = _from_oracle(em, o, bc=RIEMANN_BC, mode="weno")
    for r in range(2):
        n0 = o.nblocks
        o.regrid()
        g.regrid()
        assert o.nblocks < n0, r
        _same_tree(g, o)
        assert np.array_equal(g.download(), o.get_u()), r
    for _ in range(2):
        dt = o.dt_cfl()
        assert g.dt_cfl() == dt
        o.step(dt)
        g.step(dt)
    assert np.array_equal(g.download(), o.get_u())


def test_graph_replay_across_regrids(em):
    # run_steps replays one captured step per hierarchy; when a regrid changes
    # the hierarchy the executable graph is updated in place (or instantiated
    # anew), and the replayed steps keep the oracle's bits
    print("amr graph replay across regrids, 2-D shock-bubble", flush=True)
    o = OracleAmr("shock_bubble", [3.0, 0.3, 0.6, 0.25, 0.0, 0.12, 0.1], (8, 4, 0), 8, 3, (0, 0, 0), (1, 0.5, 1),
                  eps=1e-2, bc=SB2D_BC, mode="weno", regrid_every=2)
    g = _from_oracle(em, o, bc=SB2D_BC, mode="weno", eps=1e-2, regrid_every=2)
    sizes = set()
    for _ in range(56):
        o.step(o.dt_cfl())
        sizes.add(o.nblocks)
    g.run_steps(56)
    assert len(sizes) >= 2, "the hierarchy should have changed"
    assert g.t == o.t
    _same_tree(g, o)
    assert np.array_equal(g.download(), o.get_u())


@pytest.mark.parametrize("b,base,nsteps", [(16, (4, 2, 0), 56), (32, (2, 1, 0), 12)])
def test_larger_blocks_2d(em, b, base, nsteps):
    # 16^2 and 32^2 blocks: their own flux-pass shapes, index shifts, ghost
    # slabs and restriction patches (2b + 8 cells).  With 16^2 blocks the
    # hierarchy changes after about 41 steps.
    print(f"amr 2-D shock-bubble weno, {b}^2 blocks, {nsteps} steps", flush=True)
    o = OracleAmr("shock_bubble", [3.0, 0.3, 0.6, 0.25, 0.0, 0.12, 0.1], base, b, 3, (0, 0, 0), (1, 0.5, 1),
                  eps=1e-2, bc=SB2D_BC, mode="weno", regrid_every=2)
    lv, _, leaf = o.blocks()
    assert len(set(lv[leaf])) >= 2
    g = _from_oracle(em, o, bc=SB2D_BC, mode="weno", eps=1e-2, regrid_every=2)
    rg, sg = g.rhs(1e-3)
    ro, so = o.compute_rhs(1e-3)
    assert np.array_equal(rg, ro)
    assert sg == so
    assert np.array_equal(g.download(), o.get_u())
    # rhs() adds its faces to the device counters, compute_rhs not to the
    # oracle's: compare what the steps add
    c0, s0 = g.counters(), o.stats()[1:]
    sizes = set()
    for k in range(nsteps):
        dt = o.dt_cfl()
        assert g.dt_cfl() == dt, k
        o.step(dt)
        g.step(dt)
        _same_tree(g, o)
        sizes.add(o.nblocks)
    assert np.array_equal(g.download(), o.get_u())
    assert [x - y for x, y in zip(g.counters(), c0)] == [x - y for x, y in zip(o.stats()[1:], s0)]
    if b == 16:
        assert len(sizes) >= 2, "the hierarchy should have changed"


def test_16cubed_blocks_3d(em):
    # 3-D shock-bubble on 16^3 blocks with leaves on two levels: residuals and
    # restricted ancestors for both flux-pass shapes of this block size, then
    # steps, bit-identical; per-leaf launches give the same bits
    print("amr 3-D shock-bubble weno, 16^3 blocks", flush=True)
    bc = [["outflow", "outflow"], ["reflect", "reflect"], ["reflect", "reflect"]]
    prm = [2.0, 0.3, 0.6, 0.25, 0.25, 0.15, 0.1]
    o = OracleAmr("shock_bubble", prm, (4, 2, 2), 16, 2, (0, 0, 0), (1, 0.5, 0.5), eps=0.5, bc=bc, mode="weno",
                  regrid_every=2)
    lv, _, leaf = o.blocks()
    assert len(set(lv[leaf])) == 2
    ro, so = o.compute_rhs(1e-3)
    for shape in range(2):
        g = _from_oracle(em, o, bc=bc, mode="weno", eps=0.5, regrid_every=2)
        g.set_shapes([shape] * 3)
        rg, sg = g.rhs(1e-3)
        assert np.array_equal(rg, ro), shape
        assert sg == so, shape
        assert np.array_equal(g.download(), o.get_u()), shape
    per = _from_oracle(em, o, bc=bc, mode="weno", eps=0.5, regrid_every=2, batched=False)
    for _ in range(3):
        dt = o.dt_cfl()
        assert g.dt_cfl() == dt
        o.step(dt)
        g.step(dt)
        per.step(dt)
    _same_tree(g, o)
    assert np.array_equal(g.download(), o.get_u())
    assert np.array_equal(per.download(), o.get_u())


def test_32cubed_blocks_3d(em):
    # 3-D shock-bubble on 32^3 blocks over two levels: the restriction's
    # sub-boxes (1 x 16 x 8 coarse cells, edge boxes of width 9 selecting
    # part of them) and the global-scratch ghost predictions of this block
    # size, bit-identical to the oracle through steps
    print("amr 3-D shock-bubble weno, 32^3 blocks", flush=True)
    bc = [["outflow", "outflow"], ["reflect", "reflect"], ["reflect", "reflect"]]
    prm = [2.0, 0.15, 0.3, 0.25, 0.25, 0.12, 0.1]  # shock and bubble in the first level-0 block
    o = OracleAmr("shock_bubble", prm, (2, 1, 1), 32, 2, (0, 0, 0), (1, 0.5, 0.5), eps=0.5, bc=bc, mode="weno",
                  regrid_every=2)
    lv, _, leaf = o.blocks()
    assert len(set(lv[leaf])) == 2
    ro, so = o.compute_rhs(1e-3)
    for shape in range(2):
        g = _from_oracle(em, o, bc=bc, mode="weno", eps=0.5, regrid_every=2)
        g.set_shapes([shape] * 3)
        rg, sg = g.rhs(1e-3)
        assert np.array_equal(rg, ro), shape
        assert sg == so, shape
        assert np.array_equal(g.download(), o.get_u()), shape
    for _ in range(2):
        dt = o.dt_cfl()
        assert g.dt_cfl() == dt
        o.step(dt)
        g.step(dt)
    _same_tree(g, o)
    assert np.array_equal(g.download(), o.get_u())


def test_details_and_regrid(em):
    print("amr details + regrid 2-D riemann", flush=True)
    o = _riemann("weno")
    g = _from_oracle(em, o, bc=RIEMANN_BC, mode="weno")
    o.restrict_all()
    do = o.details()
    dg = g.details()
    _, _, leaf = o.blocks()
    assert np.all(dg[~leaf] == 0.0)
    assert _rel(dg[leaf], do[leaf]) < 1e-12
    # move the discontinuities, then regrid both
    for _ in range(4):
        dt = o.dt_cfl()
        o.set_next_regrid(1 << 40)
        g.set_next_regrid(1 << 40)
        o.step(dt)
        g.step(dt)
    o.regrid()
    g.regrid()
    _same_tree(g, o)
    assert np.array_equal(g.download(), o.get_u())


def test_created_from_case_matches(em):
    print("amr create_case riemann2d", flush=True)
    o = _riemann("weno")
    g = em.HipAmr((4, 4, 0), 8, 3, (0, 0, 0), (1, 1, 1), case="riemann2d", params=[0.55, 0.55], eps=1e-3,
                  bc=RIEMANN_BC, mode="weno")
    _same_tree(g, o)
    assert np.array_equal(g.download(), o.get_u())


def test_batched_equals_per_leaf(em):
    print("amr batched vs per-leaf launches", flush=True)
    o = _riemann("hybrid")
    g1 = _from_oracle(em, o, bc=RIEMANN_BC, mode="hybrid")
    g2 = _from_oracle(em, o, bc=RIEMANN_BC, mode="hybrid", batched=False)
    for g in (g1, g2):
        g.set_next_regrid(1 << 40)
        for _ in range(2):
            g.step(2e-4)
    assert np.array_equal(g1.download(), g2.download())
    assert g1.counters() == g2.counters()


def test_periodic_conservation_and_run_steps(em):
    print("amr smooth periodic 2-D: conservation, run_steps", flush=True)
    o = OracleAmr("smooth", None, (4, 4, 0), 8, 2, (0, 0, 0), (1, 1, 1), eps=4e-6, mode="weno")
    lv, _, leaf = o.blocks()
    assert len(set(lv[leaf])) >= 2
    g1 = _from_oracle(em, o, mode="weno", eps=4e-6)
    g2 = _from_oracle(em, o, mode="weno", eps=4e-6)
    g3 = _from_oracle(em, o, mode="weno", eps=4e-6)
    g3.set_graph(False)

    def mass(g):
        lv, _, leaf = g.blocks()
        u = g.download()
        return sum(u[i, ..., 0].sum() * (1.0 / (32 * 2 ** lv[i])) ** 2 for i in range(len(lv)) if leaf[i])

    m0 = mass(g1)
    for _ in range(5):
        g1.step(g1.dt_cfl())
    g2.run_steps(5)
    g3.run_steps(5)
    assert g1.t == g2.t == g3.t
    assert np.array_equal(g1.download(), g2.download())
    assert np.array_equal(g1.download(), g3.download())
    assert abs(mass(g1) - m0) < 1e-13 * abs(m0)


def test_run_to_matches_oracle_loop(em):
    print("amr run_to vs oracle loop", flush=True)
    o = _riemann("weno", levels=2, regrid_every=3)
    g = _from_oracle(em, o, bc=RIEMANN_BC, mode="weno", regrid_every=3)
    t_end = 0.004
    n = 0
    while o.t < t_end * (1.0 - 1e-14):
        o.step(min(o.dt_cfl(), t_end - o.t))
        n += 1
    assert g.run_to(t_end) == n
    assert g.t == o.t
    _same_tree(g, o)
    assert np.array_equal(g.download(), o.get_u())


@pytest.mark.parametrize("mode", ["weno", "hybrid"])
def test_shock_bubble_3d(em, mode):
    print(f"amr 3-D shock-bubble mode={mode}", flush=True)
    bc = [["outflow", "outflow"], ["reflect", "reflect"], ["reflect", "reflect"]]
    prm = [2.0, 0.3, 0.6, 0.25, 0.25, 0.15, 0.1]
    o = OracleAmr("shock_bubble", prm, (4, 2, 2), 8, 2, (0, 0, 0), (1, 0.5, 0.5), eps=0.2, bc=bc, mode=mode,
                  regrid_every=2)
    lv, _, leaf = o.blocks()
    assert len(set(lv[leaf])) >= 2
    g = _from_oracle(em, o, bc=bc, mode=mode, eps=0.2, regrid_every=2)
    rg, _ = g.rhs(1e-3)
    ro, _ = o.compute_rhs(1e-3)
    if mode == "weno":
        assert np.array_equal(rg, ro)
    else:
        assert _rel(rg, ro) < 1e-13
    for _ in range(3):
        dt = o.dt_cfl()
        assert g.dt_cfl() == dt or mode != "weno"
        o.step(dt)
        g.step(dt)
    _same_tree(g, o)
    if mode == "weno":
        assert np.array_equal(g.download(), o.get_u())
    else:
        assert _rel(g.download(), o.get_u()) < 1e-12


def test_restriction_images_over_three_axes(em):
    # The shock starts next to the outflow x-low face and refines a layer of
    # blocks across the whole y-z section, so the fine patches of edge
    # ancestors reach past the x, y and z boundaries at once: their corner
    # cells are images of images (outflow, reflect, a Dirichlet state; the
    # momentum sign changes of several axes combined).  Residuals, the
    # restricted ancestors and states after steps stay bit-identical.
    print("amr restriction images over three axes, 3-D shock-bubble", flush=True)
    wall = (1.0, 0.0, 0.0, 0.0, 2.5)
    bc = [["outflow", ("dirichlet", wall)], ["reflect", ("dirichlet", wall)], ["reflect", "outflow"]]
    prm = [2.0, 0.06, 0.6, 0.25, 0.25, 0.15, 0.1]
    o = OracleAmr("shock_bubble", prm, (4, 2, 2), 8, 3, (0, 0, 0), (1, 0.5, 0.5), eps=0.2, bc=bc, mode="weno",
                  regrid_every=2)
    lv, idx, leaf = o.blocks()
    anc = (lv == 1) & ~leaf
    corner = anc & (idx[:, 0] == 0) & np.isin(idx[:, 1], (0, 3)) & np.isin(idx[:, 2], (0, 3))
    assert corner.sum() >= 1 and ((lv == 1) & leaf).sum() >= 1
    g = _from_oracle(em, o, bc=bc, mode="weno", eps=0.2, regrid_every=2)
    rg, sg = g.rhs(1e-3)
    ro, so = o.compute_rhs(1e-3)
    assert np.array_equal(rg, ro)
    assert sg == so
    assert np.array_equal(g.download(), o.get_u())
    for _ in range(3):
        dt = o.dt_cfl()
        assert g.dt_cfl() == dt
        o.step(dt)
        g.step(dt)
    _same_tree(g, o)
    assert np.array_equal(g.download(), o.get_u())


def test_characteristic_and_dirichlet_boundaries(em):
    print("amr characteristic / dirichlet boundaries", flush=True)
    far = (1.0, 0.1, 0.0, 0.0, 2.5)
    bc = [[("characteristic", far), ("dirichlet", (1.0, 0.0, 0.0, 0.0, 2.5))], ["reflect", "outflow"],
          ["periodic", "periodic"]]
    o = OracleAmr("vortex", None, (4, 4, 0), 8, 2, (-5, -5, 0), (5, 5, 1), eps=1e-3, bc=bc, mode="weno")
    g = _from_oracle(em, o, bc=bc, mode="weno", eps=1e-3)
    rg, sg = g.rhs(1e-3)
    ro, so = o.compute_rhs(1e-3)
    assert np.array_equal(rg, ro)
    assert sg == so


def test_nested_characteristic_corners(em):
    """Characteristic boundaries on two axes meeting at a refined corner
    (compressible-cuda/src/kernels.cu:722-786, amr.rs:450-504): the corner
    ancestor's restriction patch and the corner leaves' ghosts reach past both
    boundaries, whose images nest (the y image of an x image); residual,
    steps with regrids, trees and data bit-identical to the oracle."""
    print("amr nested characteristic corners", flush=True)
    far = (1.0, 0.0, 0.0, 0.0, 2.5)
    bc = [[("characteristic", far), "outflow"], [("characteristic", far), ("characteristic", far)],
          ["periodic", "periodic"]]
    # a Mach 1.5 shock at x = 0.6 and a light bubble in the (0, 0) corner
    prm = [1.5, 0.6, 0.1, 0.1, 0.0, 0.18, 0.2]
    o = OracleAmr("shock_bubble", prm, (4, 4, 0), 8, 2, (0, 0, 0), (1, 1, 1), eps=1e-2, bc=bc, mode="weno",
                  regrid_every=2)
    lv, idx, leaf = o.blocks()
    assert ((lv == 0) & ~leaf & (idx[:, 0] == 0) & (idx[:, 1] == 0)).sum() == 1  # the corner is refined
    g = _from_oracle(em, o, bc=bc, mode="weno", eps=1e-2, regrid_every=2)
    rg, sg = g.rhs(1e-3)
    ro, so = o.compute_rhs(1e-3)
    assert np.array_equal(rg, ro)
    assert sg == so
    for _ in range(3):
        dt = o.dt_cfl()
        assert g.dt_cfl() == dt
        o.step(dt)
        g.step(dt)
    _same_tree(g, o)
    assert np.array_equal(g.download(), o.get_u())


def test_pass_shapes_and_autotune_give_identical_results(em):
    # Every compiled shape of the batched flux pass (pencils per work item,
    # threads) gives the same bits: a cell's flux differences and their
    # accumulation do not depend on the grouping.  The autotuner picks one
    # of them and leaves the state and counters as they were.
    print("amr pass shapes and autotune, 3-D shock-bubble", flush=True)
    bc = [["outflow", "outflow"], ["reflect", "reflect"], ["reflect", "reflect"]]
    prm = [2.0, 0.3, 0.6, 0.25, 0.25, 0.15, 0.1]
    o = OracleAmr("shock_bubble", prm, (4, 2, 2), 8, 2, (0, 0, 0), (1, 0.5, 0.5), eps=0.2, bc=bc, mode="weno",
                  regrid_every=2)
    ref = None
    for shape in range(4):
        g = _from_oracle(em, o, bc=bc, mode="weno", eps=0.2, regrid_every=2)
        g.set_shapes([shape] * 3)
        g.run_steps(3)
        got = (g.download(), g.counters(), g.t)
        if ref is None:
            ref = got
        else:
            assert np.array_equal(got[0], ref[0]) and got[1:] == ref[1:], shape
    g = _from_oracle(em, o, bc=bc, mode="weno", eps=0.2, regrid_every=2)
    before = (g.download(), g.counters())
    shapes = g.autotune(2)
    assert all(0 <= s < 4 for s in shapes)
    assert np.array_equal(g.download(), before[0]) and g.counters() == before[1]
    g.run_steps(3)
    assert np.array_equal(g.download(), ref[0]) and g.counters() == ref[1]
