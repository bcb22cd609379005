"""Hierarchies of uneven block size (8, 16 and 32 cells a side) stepped
together (em_amr_group_*: one persistent flux-pass launch per axis and
stage over the leaves of every member) against the same hierarchies
stepped one at a time (em_amr_run_steps, itself bit-identical to the
oracle's CPU `Amr` in the AMR parity tests).

Bar: bit-identical states, identical trees through regrids, equal
counters and times.  The group kernel takes a different pencil count per
work item than the single-hierarchy pass, and every pencil's arithmetic is
the same, so nothing may change.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RIEMANN_BC = [["outflow", "outflow"], ["outflow", "outflow"], ["periodic", "periodic"]]
BUBBLE_BC = [["outflow", "outflow"], ["reflect", "reflect"], ["reflect", "reflect"]]


@pytest.fixture(scope="module")
def em():
    import emergence_amd

    emergence_amd.load()
    return emergence_amd


def _riemann(em, b, mode):
    base = (32 // b, 32 // b, 0)
    return em.HipAmr(base, b, 3, (0, 0, 0), (1, 1, 1), case="riemann2d", params=[0.55, 0.55], eps=1e-3,
                     regrid_every=4, bc=RIEMANN_BC, mode=mode)


def _bubble(em, b, mode):
    base = (32 // b, 16 // b, 16 // b)
    return em.HipAmr(base, b, 2, (0, 0, 0), (2, 1, 1), case="shock_bubble",
                     params=[1.22, 0.5, 1.0, 0.5, 0.5, 0.25, 0.138], eps=1e-2, regrid_every=3, bc=BUBBLE_BC,
                     mode=mode, cfl=0.4)


def _check_same(a, b):
    la, ia, fa = a.blocks()
    lb, ib, fb = b.blocks()
    assert np.array_equal(la, lb) and np.array_equal(ia, ib) and np.array_equal(fa, fb)
    assert np.array_equal(a.download(), b.download())
    assert a.counters() == b.counters()
    assert a.t == b.t


@pytest.mark.parametrize("mode", ["weno", "hybrid"])
def test_group_of_uneven_blocks_2d_equals_members_alone(em, mode):
    print(f"amr group 2-D riemann, blocks 8/16/32, mode={mode}", flush=True)
    solo = [_riemann(em, b, mode) for b in (8, 16, 32)]
    grp = [_riemann(em, b, mode) for b in (8, 16, 32)]
    g = em.HipAmrGroup(grp)
    for s in solo:
        s.run_steps(10)
    g.run_steps(10)
    for a, b in zip(grp, solo):
        _check_same(a, b)
    assert len({m.leaf_count() for m in grp}) >= 2


def test_group_of_uneven_blocks_3d_equals_members_alone(em):
    print("amr group 3-D shock-bubble, blocks 8/16", flush=True)
    solo = [_bubble(em, b, "hybrid") for b in (8, 16)]
    grp = [_bubble(em, b, "hybrid") for b in (8, 16)]
    g = em.HipAmrGroup(grp)
    for s in solo:
        s.run_steps(7)
    g.run_steps(4)
    g.run_steps(3)
    for a, b in zip(grp, solo):
        _check_same(a, b)


def test_group_rejects_mixed_modes(em):
    from emergence_amd import EmError

    a, b = _riemann(em, 8, "weno"), _riemann(em, 16, "hybrid")
    with pytest.raises(EmError):
        em.HipAmrGroup([a, b])
