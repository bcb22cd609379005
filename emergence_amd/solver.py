"""Host mirror of the reference's GPU solver object over the C ABI.

`HipSolver` has the method set of `GpuSolver<f64>`
(fluid/crates/compressible-cuda/src/lib.rs:183-477): `from_cpu`-style
construction from a mesh, boundary conditions, scheme options and a conserved
state; `step`, `dt_cfl`, `run_to`, `download`, `counters`, `synchronize`; plus
`rhs` (one residual evaluation, for parity checks) and `run_steps`.

State arrays are the reference CPU layout: ghost-padded AoS float64 of shape
(m0, m1, m2, 5) with m_d = n_d + 2 ng on active axes (solver.rs:40-77).
"""
import ctypes as C

import numpy as np

from . import _lib

BC = {"periodic": 0, "outflow": 1, "reflect": 2, "dirichlet": 3, "characteristic": 4}
MODE = {"hybrid": 0, "weno": 1, "central": 2}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def padded_shape(n, ng):
    return tuple(int(v) + 2 * ng if v > 1 else 1 for v in n) + (5,)


def case_init(case_id, n, lo, hi, ng=6, params=None):
    """Workload initialiser of the library (see em_case_init)."""
    lib = _lib.load()
    u = np.empty(padded_shape(n, ng), dtype=np.float64)
    prm = None if params is None else _dp(np.ascontiguousarray(params, dtype=np.float64))
    _lib.check(lib.em_case_init(case_id, (C.c_int64 * 3)(*n), (C.c_double * 3)(*lo), (C.c_double * 3)(*hi), ng,
                                prm, _dp(u)))
    return u


class HipSolver:
    def __init__(self, n, lo, hi, u, ng=6, eos=(1.4, 0.0, 0.0, 0.0), bc=None, mode="hybrid", positivity=True,
                 cfl=0.5, tol=1e-3, t0=0.0, precise=None, real="f64"):
        """bc: [[kind_lo, kind_hi]] * 3 with kinds as names, or (name, state5)
        tuples for Dirichlet / characteristic sides; default periodic.
        real: "f64" (the default) or "f32" (state and arithmetic in FP32 on
        the device, em_solver_create_f32).  precise: no FMA contraction;
        default True for f64, False for f32."""
        if real not in ("f64", "f32"):
            raise ValueError("real must be 'f64' or 'f32'")
        self.real = real
        if precise is None:
            precise = real == "f64"
        self.lib = _lib.load()
        self.n = tuple(int(v) for v in n)
        self.ng = int(ng)
        self.shape = padded_shape(self.n, self.ng)
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
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.shape)
        h = C.c_void_p()
        create = self.lib.em_solver_create if real == "f64" else self.lib.em_solver_create_f32
        _lib.check(create(C.byref(h), (C.c_int64 * 3)(*self.n), (C.c_double * 3)(*lo), (C.c_double * 3)(*hi),
                          self.ng, (C.c_double * 4)(*eos), kinds, states, MODE[mode], int(positivity), cfl, tol, t0,
                          _dp(u)))
        self.h = h
        self.set_variant(precise)

    def close(self):
        if getattr(self, "h", None):
            self.lib.em_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_variant(self, precise):
        _lib.check(self.lib.em_solver_set_variant(self.h, int(bool(precise))))

    def set_tiles(self, cfg):
        _lib.check(self.lib.em_solver_set_tiles(self.h, (C.c_int32 * 3)(*cfg)))

    def autotune(self, repeats=3):
        out = (C.c_int32 * 3)()
        _lib.check(self.lib.em_solver_autotune(self.h, repeats, out))
        return list(out)

    def upload(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.shape)
        _lib.check(self.lib.em_solver_upload(self.h, _dp(u)))

    def download(self):
        u = np.empty(self.shape, dtype=np.float64)
        _lib.check(self.lib.em_solver_download(self.h, _dp(u)))
        return u

    def rhs(self, stage_dt):
        out = np.empty(self.shape, dtype=np.float64)
        st = (C.c_uint64 * 4)()
        _lib.check(self.lib.em_solver_rhs(self.h, stage_dt, _dp(out), st))
        return out, list(st)

    def step(self, dt):
        _lib.check(self.lib.em_solver_step(self.h, dt))

    def dt_cfl(self):
        dt = C.c_double()
        _lib.check(self.lib.em_solver_dt_cfl(self.h, C.byref(dt)))
        return dt.value

    def run_to(self, t_end):
        k = C.c_int64()
        _lib.check(self.lib.em_solver_run_to(self.h, t_end, C.byref(k)))
        return k.value

    def run_steps(self, nsteps, timed=False):
        if not timed:
            _lib.check(self.lib.em_solver_run_steps(self.h, nsteps, None))
            return None
        ms = (C.c_double * 3)()
        _lib.check(self.lib.em_solver_run_steps(self.h, nsteps, ms))
        return list(ms)

    def counters(self):
        out = (C.c_uint64 * 3)()
        _lib.check(self.lib.em_solver_counters(self.h, out))
        return list(out)

    @property
    def t(self):
        t = C.c_double()
        _lib.check(self.lib.em_solver_time(self.h, C.byref(t)))
        return t.value

    def synchronize(self):
        _lib.check(self.lib.em_solver_synchronize(self.h))

    # ---- discrete adjoint (FluxMode::Weno)
    def set_adjoint_tile(self, cells):
        """Interior cells per VJP tile at most (>= 10); results do not depend on it."""
        _lib.check(self.lib.em_solver_set_adjoint_tile(self.h, int(cells)))

    def adjoint_rhs(self, stage_dt, lam):
        """(dR/du)^T lam of one residual evaluation at the current state."""
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(self.shape)
        out = np.empty(self.shape, dtype=np.float64)
        _lib.check(self.lib.em_solver_adjoint_rhs(self.h, float(stage_dt), _dp(lam), _dp(out)))
        return out

    def forward(self, nsteps, checkpoint_every, dts=None):
        """nsteps steps with checkpoints every checkpoint_every steps (0: every
        step's states kept on the device, nothing recomputed by backward);
        dts frozen (or CFL if None).  Returns the dt sequence used."""
        out = np.empty(int(nsteps), dtype=np.float64)
        d = None if dts is None else _dp(np.ascontiguousarray(dts, dtype=np.float64))
        _lib.check(self.lib.em_solver_forward(self.h, int(nsteps), int(checkpoint_every), d, _dp(out)))
        return out

    def backward(self, dj_final):
        """dJ/du_0 of the last forward run from dJ/du_N (CPU layout)."""
        dj = np.ascontiguousarray(dj_final, dtype=np.float64).reshape(self.shape)
        out = np.empty(self.shape, dtype=np.float64)
        _lib.check(self.lib.em_solver_backward(self.h, _dp(dj), _dp(out)))
        return out

    def backward_device(self, dj_final, dj_initial):
        """backward on device arrays (DeviceArray) of the interior cells,
        variable-major (to_soa / from_soa convert): nothing is converted
        on the host."""
        nv = 5 * int(np.prod(self.n))
        if dj_final.n != nv or dj_initial.n != nv:
            raise ValueError(f"device arrays must hold 5 x {nv // 5} values")
        _lib.check(self.lib.em_solver_backward_device(self.h, dj_final.ptr, dj_initial.ptr))

    def to_soa(self, a):
        """CPU layout -> interior cells, variable-major (5 x ncell)."""
        a = np.asarray(a, dtype=np.float64).reshape(self.shape)
        return np.ascontiguousarray(np.moveaxis(self.interior(a), -1, 0)).ravel()

    def from_soa(self, v):
        """Inverse of to_soa (ghost cells zero)."""
        out = np.zeros(self.shape, dtype=np.float64)
        self.interior(out)[...] = np.moveaxis(np.asarray(v, dtype=np.float64).reshape((5,) + tuple(self.n)), 0, -1)
        return out

    def interior(self, a):
        g = [self.ng if v > 1 else 0 for v in self.n]
        return a[g[0]:g[0] + self.n[0], g[1]:g[1] + self.n[1], g[2]:g[2] + self.n[2]]
