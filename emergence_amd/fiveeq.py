"""Host mirror of the reference's GPU solver object for the two-phase Kapila
five-equation system over the C ABI (em_fiveeq_*).

`HipFiveEq` has the method set of `GpuSolver<f64>` built with
`GpuSystem::FiveEq` (fluid/crates/compressible-cuda/src/lib.rs:50-82,
183-477): construction from a mesh, the two phases' equations of state, the
Kapila flag, boundary conditions, scheme options and a conserved state;
`step`, `dt_cfl`, `run_to`, `download`, `counters`, `synchronize`; plus `rhs`
(one residual evaluation, for parity checks).

State arrays: ghost-padded AoS float64 of shape (m0, m1, m2, 7), conserved
(a1, a2, rho u, rho v, rho w, E, alpha1).
"""
import ctypes as C

import numpy as np

from . import _lib
from .solver import BC, MODE

# water (stiffened gas) and air (ideal gas): the reference's `interface`
# probe materials (examples/cpu_validate.rs:101-112)
WATER_AIR = ((2.0, 1e6, 0.0, 0.0), (1.4, 0.0, 0.0, 0.0))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def padded_shape(n, ng):
    return tuple(int(v) + 2 * ng if v > 1 else 1 for v in n) + (7,)


class HipFiveEq:
    def __init__(self, n, lo, hi, u, ng=6, phases=WATER_AIR, kapila=True, bc=None, mode="hybrid",
                 positivity=True, cfl=0.5, tol=1e-3, t0=0.0, real="f64"):
        """bc: [[kind_lo, kind_hi]] * 3 with kinds as names, or (name, state7)
        tuples for Dirichlet / characteristic sides; default periodic.
        real: "f64" (GpuSolver<f64>) or "f32" (GpuSolver<f32>: state and
        arithmetic in FP32 on the device; arrays here stay float64)."""
        if real not in ("f64", "f32"):
            raise ValueError("real must be 'f64' or 'f32'")
        self.lib = _lib.load()
        self.n = tuple(int(v) for v in n)
        self.ng = int(ng)
        self.shape = padded_shape(self.n, self.ng)
        kinds = (C.c_int32 * 6)()
        states = (C.c_double * 42)()
        bc = bc or [["periodic", "periodic"]] * 3
        for d in range(3):
            for s in range(2):
                b = bc[d][s]
                name, st = (b, None) if isinstance(b, str) else (b[0], b[1])
                kinds[2 * d + s] = BC[name]
                if st is not None:
                    for q in range(7):
                        states[(2 * d + s) * 7 + q] = float(st[q])
        ph = (C.c_double * 8)(*[float(v) for p in phases for v in p])
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.shape)
        h = C.c_void_p()
        create = self.lib.em_fiveeq_create_f32 if real == "f32" else self.lib.em_fiveeq_create
        _lib.check_fiveeq(create(C.byref(h), (C.c_int64 * 3)(*self.n), (C.c_double * 3)(*lo), (C.c_double * 3)(*hi),
                                 self.ng, ph, int(bool(kapila)), kinds, states, MODE[mode], int(positivity), cfl, tol,
                                 t0, _dp(u)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.em_fiveeq_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def upload(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.shape)
        _lib.check_fiveeq(self.lib.em_fiveeq_upload(self.h, _dp(u)))

    def download(self):
        u = np.empty(self.shape, dtype=np.float64)
        _lib.check_fiveeq(self.lib.em_fiveeq_download(self.h, _dp(u)))
        return u

    def rhs(self, stage_dt):
        out = np.empty(self.shape, dtype=np.float64)
        st = (C.c_uint64 * 4)()
        _lib.check_fiveeq(self.lib.em_fiveeq_rhs(self.h, stage_dt, _dp(out), st))
        return out, list(st)

    def step(self, dt):
        _lib.check_fiveeq(self.lib.em_fiveeq_step(self.h, dt))

    def dt_cfl(self):
        dt = C.c_double()
        _lib.check_fiveeq(self.lib.em_fiveeq_dt_cfl(self.h, C.byref(dt)))
        return dt.value

    def run_to(self, t_end):
        k = C.c_int64()
        _lib.check_fiveeq(self.lib.em_fiveeq_run_to(self.h, t_end, C.byref(k)))
        return k.value

    def set_tiles(self, cfg):
        _lib.check_fiveeq(self.lib.em_fiveeq_set_tiles(self.h, (C.c_int32 * 3)(*cfg)))

    def autotune(self, repeats=3):
        out = (C.c_int32 * 3)()
        _lib.check_fiveeq(self.lib.em_fiveeq_autotune(self.h, repeats, out))
        return list(out)

    def counters(self):
        out = (C.c_uint64 * 3)()
        _lib.check_fiveeq(self.lib.em_fiveeq_counters(self.h, out))
        return list(out)

    @property
    def t(self):
        t = C.c_double()
        _lib.check_fiveeq(self.lib.em_fiveeq_time(self.h, C.byref(t)))
        return t.value

    def synchronize(self):
        _lib.check_fiveeq(self.lib.em_fiveeq_synchronize(self.h))

    def interior(self, a):
        g = [self.ng if v > 1 else 0 for v in self.n]
        return a[g[0]:g[0] + self.n[0], g[1]:g[1] + self.n[1], g[2]:g[2] + self.n[2]]
