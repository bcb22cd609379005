"""Host mirror of the reference's GPU block AMR over the C ABI.

`HipAmr` has the method set of `GpuAmr<f64, 5, Euler>`
(fluid/crates/compressible-cuda/src/amr.rs:95-858): construction from a
hierarchy (`new(Amr, ..)`, here the block list of the CPU `Amr`) or from a
built-in initial condition (`Amr::new`, compressible/src/amr.rs:289-343);
`step`, `dt_cfl`, `run_to`, `download`, `t`, `counters`, `leaf_count`; plus
`rhs` (one residual evaluation), `details`, `regrid` and `run_steps`.

Block data is AoS float64 of shape (nblocks, b0, b1, b2, 5) with b_d = b on
active axes (base[d] > 0) and 1 otherwise.
"""
import ctypes as C

import numpy as np

from . import _lib
from .solver import BC, MODE

AMR_CASE = {"riemann2d": 0, "smooth": 1, "shock_bubble": 2, "vortex": 3, "poly": 4}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _bc_arrays(bc):
    kinds = (C.c_int32 * 6)()
    states = (C.c_double * 30)()
    bc = bc or [["periodic", "periodic"]] * 3
    for d in range(3):
        for s in range(2):
            b = bc[d][s]
            name, st = (b, None) if isinstance(b, str) else (b[0], b[1])
            kinds[2 * d + s] = BC[name]
            if st is not None:
                for q in range(5):
                    states[(2 * d + s) * 5 + q] = float(st[q])
    return kinds, states


class HipAmr:
    def __init__(self, base, b, levels, lo, hi, case=None, params=None, blocks=None, eps=1e-3, regrid_every=4,
                 ghosts=5, bc=None, mode="hybrid", cfl=0.4, positivity=True, tol=1e-3, eos=(1.4, 0.0, 0.0, 0.0),
                 precise=True, batched=True, real="f64"):
        """Either `case` (name of a built-in initial condition, with params) or
        `blocks` = (level[n], idx[n, 3], leaf[n], u[n, b0, b1, b2, 5]).
        real: "f64" (GpuAmr<f64>) or "f32" (GpuAmr<f32>: pool and arithmetic
        in FP32 on the device; arrays here stay float64)."""
        if real not in ("f64", "f32"):
            raise ValueError("real must be 'f64' or 'f32'")
        self.lib = _lib.load()
        self.real = real
        self.base = tuple(int(v) for v in base)
        self.b = int(b)
        self.levels = int(levels)
        self.lo, self.hi = tuple(lo), tuple(hi)
        self.bd = tuple(self.b if v > 0 else 1 for v in self.base)
        kinds, states = _bc_arrays(bc)
        h = C.c_void_p()
        cfg = (C.byref(h), (C.c_int64 * 3)(*self.base), self.b, self.levels, int(ghosts), float(eps),
               int(regrid_every), (C.c_double * 3)(*lo), (C.c_double * 3)(*hi), (C.c_double * 4)(*eos), kinds,
               states, MODE[mode], int(bool(positivity)), float(cfl), float(tol))
        if case is not None:
            prm = np.zeros(15)
            if params is not None:
                prm[:len(params)] = params
            create = self.lib.em_amr_create_case_f32 if real == "f32" else self.lib.em_amr_create_case
            _lib.check_amr(create(*cfg, AMR_CASE[case], _dp(prm)))
        else:
            level, idx, leaf, u = blocks
            level = np.ascontiguousarray(level, dtype=np.int32)
            idx = np.ascontiguousarray(idx, dtype=np.int64).reshape(-1, 3)
            leaf = np.ascontiguousarray(leaf, dtype=np.uint8)
            u = np.ascontiguousarray(u, dtype=np.float64)
            n = len(level)
            if u.size != n * int(np.prod(self.bd)) * 5:
                raise ValueError("block data has the wrong size")
            create = self.lib.em_amr_create_f32 if real == "f32" else self.lib.em_amr_create
            _lib.check_amr(create(*cfg, n, level.ctypes.data_as(C.POINTER(C.c_int32)),
                                  idx.ctypes.data_as(C.POINTER(C.c_int64)), leaf.ctypes.data_as(C.POINTER(C.c_uint8)),
                                  _dp(u)))
        self.h = h
        self.set_variant(precise)
        self.set_batched(batched)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.em_amr_destroy(self.h)
            self.h = None

    def set_variant(self, precise):
        _lib.check_amr(self.lib.em_amr_set_variant(self.h, int(bool(precise))))

    def set_batched(self, batched):
        _lib.check_amr(self.lib.em_amr_set_batched(self.h, int(bool(batched))))

    def set_graph(self, on):
        _lib.check_amr(self.lib.em_amr_set_graph(self.h, int(bool(on))))

    def set_shapes(self, shapes):
        """Pass shape per axis (see em_amr_set_shapes)."""
        arr = (C.c_int32 * 3)(*[int(v) for v in shapes])
        _lib.check_amr(self.lib.em_amr_set_shapes(self.h, arr))

    def autotune(self, repeats=3):
        """Times every pass shape per axis and keeps the fastest; returns them."""
        out = (C.c_int32 * 3)()
        _lib.check_amr(self.lib.em_amr_autotune(self.h, int(repeats), out))
        return list(out)

    def set_profile(self, on):
        _lib.check_amr(self.lib.em_amr_set_profile(self.h, int(bool(on))))

    def profile(self):
        """Accumulated ms of [restriction, ghosts, fluxes, reflux + update, regrid]."""
        out = (C.c_double * 5)()
        _lib.check_amr(self.lib.em_amr_profile(self.h, out))
        return list(out)

    def set_next_regrid(self, step):
        _lib.check_amr(self.lib.em_amr_set_next_regrid(self.h, int(step)))

    def counts(self):
        n, nl = C.c_int64(), C.c_int64()
        _lib.check_amr(self.lib.em_amr_nblocks(self.h, C.byref(n), C.byref(nl)))
        return n.value, nl.value

    @property
    def nblocks(self):
        return self.counts()[0]

    def leaf_count(self):
        return self.counts()[1]

    def blocks(self):
        n = self.nblocks
        lv = np.empty(n, np.int32)
        idx = np.empty((n, 3), np.int64)
        leaf = np.empty(n, np.uint8)
        _lib.check_amr(self.lib.em_amr_download(self.h, lv.ctypes.data_as(C.POINTER(C.c_int32)),
                                                idx.ctypes.data_as(C.POINTER(C.c_int64)),
                                                leaf.ctypes.data_as(C.POINTER(C.c_uint8)), None))
        return lv, idx, leaf.astype(bool)

    def download(self):
        u = np.empty((self.nblocks,) + self.bd + (5,))
        _lib.check_amr(self.lib.em_amr_download(self.h, None, None, None, _dp(u)))
        return u

    def upload(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.size != self.nblocks * int(np.prod(self.bd)) * 5:
            raise ValueError("block data has the wrong size")
        _lib.check_amr(self.lib.em_amr_upload(self.h, _dp(u)))

    def rhs(self, stage_dt):
        out = np.empty((self.nblocks,) + self.bd + (5,))
        st = (C.c_uint64 * 4)()
        _lib.check_amr(self.lib.em_amr_rhs(self.h, float(stage_dt), _dp(out), st))
        return out, list(st)

    def step(self, dt):
        _lib.check_amr(self.lib.em_amr_step(self.h, float(dt)))

    def run_steps(self, nsteps):
        _lib.check_amr(self.lib.em_amr_run_steps(self.h, int(nsteps)))

    def run_to(self, t_end):
        n = C.c_int64()
        _lib.check_amr(self.lib.em_amr_run_to(self.h, float(t_end), C.byref(n)))
        return n.value

    def dt_cfl(self):
        v = C.c_double()
        _lib.check_amr(self.lib.em_amr_dt_cfl(self.h, C.byref(v)))
        return v.value

    def details(self):
        out = np.empty(self.nblocks)
        _lib.check_amr(self.lib.em_amr_details(self.h, _dp(out)))
        return out

    def regrid(self):
        _lib.check_amr(self.lib.em_amr_regrid(self.h))

    def counters(self):
        out = (C.c_uint64 * 3)()
        _lib.check_amr(self.lib.em_amr_counters(self.h, out))
        return list(out)

    @property
    def t(self):
        v = C.c_double()
        _lib.check_amr(self.lib.em_amr_time(self.h, C.byref(v)))
        return v.value

    def synchronize(self):
        _lib.check_amr(self.lib.em_amr_synchronize(self.h))


class HipAmrGroup:
    """Hierarchies of uneven block size stepped together (em_amr_group_*):
    one flux-pass launch per axis and stage over the leaves of all members.
    The group borrows its members; keep them alive while it exists."""

    def __init__(self, members):
        self.lib = _lib.load()
        self.members = list(members)
        hs = (C.c_void_p * len(self.members))(*[m.h for m in self.members])
        h = C.c_void_p()
        _lib.check_amr(self.lib.em_amr_group_create(C.byref(h), hs, len(self.members)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.em_amr_group_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run_steps(self, nsteps):
        _lib.check_amr(self.lib.em_amr_group_run_steps(self.h, int(nsteps)))
