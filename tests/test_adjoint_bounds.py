"""Host replay of the adjoint VJP kernel's index arithmetic (adj_group in
emergence_amd/csrc/em_adjoint.hip): for every tile of every pencil shape and
boundary combination the adjoint accepts, every LDS slot a phase reads or
writes lies inside the launch's LDS layout, every global position lies inside
the pencil, every slot read was written earlier in the same work item (no
stale LDS, so work items that share a workgroup cannot leak into each other),
and the scatter sub-phases of D3 and the two-step adds of C and D4 never let
two work items add to one slot between two barriers.

The ranges restate the kernel's own expressions (cited by phase); the replay
enumerates them exhaustively for the sizes of the C3 and C4 workloads, the
adjoint GPU tests and short axes.  It runs on the CPU only.
"""
import itertools

import pytest

PERIODIC, OUTFLOW, REFLECT, DIRICHLET = 0, 1, 2, 3
CAP = 128


def adj_tiles(n, klo, khi, cap=CAP):  # em_adjoint.hip adj_tiles
    if (klo == 0) != (khi == 0):
        return 1
    return max(1, (n + cap - 1) // cap)


def adj_tile_extent(n, ntile, klo, khi):  # adj_tile_extent
    cyc = klo == 0 and khi == 0
    return (n + ntile - 1) // ntile + (0 if cyc else (10 if ntile == 1 else 5))


def adj_lds_bytes(ne):  # adj_lds_bytes
    return 8 * (13 * (ne + 18) + 23 * (ne + 13))


def tile(n, ntile, t, cyc):  # adj_tile
    x0 = t * n // ntile
    x1 = (t + 1) * n // ntile
    e0 = -5 if (not cyc and t == 0) else x0
    e1 = n + 5 if (not cyc and t == ntile - 1) else x1
    return x0, x1, e0, e1


def bc_sources(n, g, kinds):
    """Interior positions bc_ghost loads for ghost position g (em_device.h)."""
    x, out = g, []
    for _ in range(64):
        if 0 <= x < n:
            return out + [x]
        side = 1 if x >= n else 0
        k = kinds[side]
        if k == REFLECT:
            x = -1 - x if side == 0 else 2 * n - 1 - x
            continue
        if k == PERIODIC:
            return out + [((x % n) + n) % n]
        if k == OUTFLOW:
            return out + [0 if side == 0 else n - 1]
        return out  # Dirichlet: the fixed state, no load
    raise AssertionError("reflection chain did not end")


def replay(n, klo, khi, cap=CAP, ne_launch=None):
    """Checks one pencil shape; returns the LDS bytes the launch needs."""
    cyc = klo == 0 and khi == 0
    ntile = adj_tiles(n, klo, khi, cap)
    ne_l = adj_tile_extent(n, ntile, klo, khi) if ne_launch is None else ne_launch
    NC, NM = ne_l + 18, ne_l + 13
    for t in range(ntile):
        x0, x1, e0, e1 = tile(n, ntile, t, cyc)
        ne = e1 - e0
        assert ne <= ne_l
        lo = 0 if cyc else max(0, 4 - e0)
        c_hi = ne + 18 if cyc else min(ne + 18, n + 14 - e0)
        m_hi = ne + 13 if cyc else min(ne + 13, n + 9 - e0)
        f_hi = ne + 9 if cyc else min(ne + 9, n + 5 - e0)
        v_lo, v_hi = max(4, lo), min(ne + 9, m_hi)
        # A: cells [lo, c_hi) loaded (sacc, sx zeroed over [0, NC))
        loaded = set(range(lo, c_hi))
        assert 0 <= lo and c_hi <= NC
        for c in loaded:
            x = c + e0 - 9
            if cyc:
                xs = [((x % n) + n) % n]
            elif 0 <= x < n:
                xs = [x]
            else:
                xs = bc_sources(n, x, (klo, khi))
                assert (klo, khi)[1 if x >= n else 0] == DIRICHLET or xs, (n, x)
            assert all(0 <= v < n for v in xs), (n, klo, khi, t, x, xs)
        # B0/B1/B2: midpoints [lo, m_hi) read cells m .. m+5 and write scc/smf
        mids = set(range(lo, m_hi))
        assert m_hi <= NM
        for m in mids:
            assert set(range(m, m + 6)) <= loaded, (n, klo, khi, t, m)
        # C: faces [lo, f_hi): midpoints f .. f+4 (written in B2), cells f+4, f+5,
        # LAM at global faces fg = f + e0 - 4 (right cell fg, left cell fg - 1)
        for f in range(lo, f_hi):
            assert set(range(f, f + 5)) <= mids
            assert {f + 4, f + 5} <= loaded
            fg = f + e0 - 4
            if not cyc:
                assert 0 <= fg <= n
                if fg < n:
                    assert 0 <= fg < n
                if fg > 0:
                    assert 0 <= fg - 1 < n
        # face slots written in C then read by D2: f in [lo, f_hi)
        # D2: midpoints [v_lo, v_hi) read faces m-4 .. m (the live ones) and
        # midpoint data of B1; D3/D4 read slot 12 written by D2 for the same m
        assert v_lo >= lo and v_hi <= m_hi
        for m in range(v_lo, v_hi):
            assert {m + 2, m + 3} <= loaded
        # D3: scatter to cells m .. m+5; D4: cells m+2, m+3; E: cells [9, ne+9)
        for m in range(v_lo, v_hi):
            assert m + 5 < NC
        for c in range(9, ne + 9):
            assert c in loaded and c < NC
        # E: interior outputs [x0, x1) and the ghost positions routed to them
        assert 0 <= x0 <= x1 <= n
        for x in range(x0, x1):
            c = x - e0 + 9
            assert 9 <= c < ne + 9
        if not cyc:
            for gi in range(10):
                gx = gi - 5 if gi < 5 else n + gi - 5
                if not (e0 <= gx < e1):
                    continue
                assert 9 <= gx - e0 + 9 < ne + 9
    return adj_lds_bytes(ne_l)


SHAPES = [2, 3, 5, 7, 9, 10, 11, 16, 37, 64, 96, 127, 128, 129, 200, 255, 256, 257, 384, 512, 562]
KINDS = [(PERIODIC, PERIODIC), (OUTFLOW, OUTFLOW), (REFLECT, REFLECT), (DIRICHLET, OUTFLOW), (OUTFLOW, PERIODIC),
         (PERIODIC, REFLECT), (REFLECT, DIRICHLET)]


@pytest.mark.parametrize("klo,khi", KINDS)
@pytest.mark.parametrize("cap", [CAP, 37, 10])
def test_adjoint_tile_indices_stay_in_bounds(klo, khi, cap):
    for n in SHAPES:
        if REFLECT in (klo, khi) and n < 5:
            continue  # the adjoint takes reflecting axes of n >= 5 (em_hip.h)
        lds = replay(n, klo, khi, cap)
        # split tiles take at most 48 KiB (three workgroups per CU); a whole
        # pencil beyond 160 KiB (gfx950's LDS per workgroup) is refused by the
        # host's adj_check before any launch: that happens above 544 cells
        if adj_tiles(n, klo, khi, cap) > 1:
            assert lds <= 48 * 1024, (n, lds)
        if lds > 160 * 1024:
            assert adj_tiles(n, klo, khi, cap) == 1 and n > 544, (n, lds)


def test_ensemble_launch_layout_fits_every_instance():
    # C4: 256 instances of sides 96..384 (multiples of 32), outflow x and y; a
    # batched launch would size LDS by the largest tile of all instances and
    # every instance's slots must fit that layout
    for klo, khi in [(OUTFLOW, OUTFLOW), (PERIODIC, PERIODIC)]:
        sides = list(range(96, 385, 32))
        nes = [adj_tile_extent(s, adj_tiles(s, klo, khi), klo, khi) for s in sides]
        ne_max = max(nes)
        for s in sides:
            replay(s, klo, khi, ne_launch=ne_max)


def _d3_targets(fa, un, t1, t2):
    """Rows of sacc (5..7 = sx 0..2) D3 adds to for field fa."""
    d1 = un if fa == 0 else 0 if fa == 1 else t1 if fa == 2 else t2 if fa == 3 else 6
    d2 = 4 if fa == 0 else 5 if fa == 1 else 7
    return (d1, d2) if fa in (0, 1, 4) else (d1,)


@pytest.mark.parametrize("D", [0, 1, 2])
def test_d3_scatter_sub_phases_are_conflict_free(D):
    un, t1, t2 = 1 + D, 1 + (D + 1) % 3, 1 + (D + 2) % 3
    nvm, NT = 70, 256
    for i0 in range(0, 5 * nvm, NT):
        items = [(it // nvm, it % nvm) for it in range(i0, min(i0 + NT, 5 * nvm))]
        for j in range(6):
            seen = set()
            for fa, m in items:
                for row in _d3_targets(fa, un, t1, t2):
                    key = (row, m + j)
                    assert key not in seen, (D, j, fa, m)
                    seen.add(key)


def test_two_step_adds_are_conflict_free():
    # C: a face adds its right state's cotangent to cell f+5, then (after a
    # barrier) its left state's to f+4; D4: midpoint m adds to m+3, then m+2.
    # Within one step the targets of distinct items are distinct.
    for first, second in [(5, 4), (3, 2)]:
        for base in range(0, 40):
            items = range(base, base + 256)
            assert len({i + first for i in items}) == len(items)
            assert len({i + second for i in items}) == len(items)
    for klo, khi in itertools.product([OUTFLOW, PERIODIC], repeat=2):
        replay(64, klo, khi)
