"""ctypes binding of the oracle (oracle/_build/libemref.so).

The oracle is the checker: a bit-faithful CPU restatement of the reference
solver.  Only tests, the smoke check and the bench's CPU-baseline leg use it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "_build", "libemref.so")
CPU_VALIDATE = os.path.join(ORACLE_DIR, "_build", "cpu_validate")

BC = {"periodic": 0, "outflow": 1, "reflect": 2, "dirichlet": 3, "characteristic": 4}
MODE = {"hybrid": 0, "weno": 1, "central": 2}
SHOCK = {"wcns": 0, "fdweno": 1}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not (os.path.exists(LIB) and os.path.exists(CPU_VALIDATE)):
        build()
    lib = C.CDLL(LIB)
    P = C.c_void_p
    D = C.POINTER(C.c_double)
    U = C.POINTER(C.c_uint64)
    I64 = C.POINTER(C.c_int64)
    lib.emref_create.restype = P
    lib.emref_create.argtypes = [I64, D, D, C.c_int64, D]
    lib.emref_destroy.argtypes = [P]
    lib.emref_len.restype = C.c_int64
    lib.emref_len.argtypes = [P]
    lib.emref_set_bc.argtypes = [P, C.c_int, C.c_int, C.c_int, D]
    lib.emref_set_opts.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double]
    lib.emref_set_u.argtypes = [P, D]
    lib.emref_get_u.argtypes = [P, D]
    lib.emref_fill_ghosts.argtypes = [P]
    lib.emref_rhs.argtypes = [P, C.c_double, D, U]
    lib.emref_step.argtypes = [P, C.c_double]
    lib.emref_dt_cfl.restype = C.c_double
    lib.emref_dt_cfl.argtypes = [P]
    lib.emref_run_to.restype = C.c_int64
    lib.emref_run_to.argtypes = [P, C.c_double]
    lib.emref_time.restype = C.c_double
    lib.emref_time.argtypes = [P]
    lib.emref_stats.argtypes = [P, U]
    lib.emref_totals.argtypes = [P, D]
    lib.emref_hash.restype = C.c_uint64
    lib.emref_hash.argtypes = [P]
    lib.emref_set_threads.argtypes = [C.c_int]
    lib.emref_max_threads.restype = C.c_int
    lib.emref_weno5z_interp.restype = C.c_double
    lib.emref_weno5z_interp.argtypes = [C.c_double] * 5
    lib.emref_riemann_flux.argtypes = [D, D, D, C.c_int, D, D]
    lib.emref_line_fluxes.argtypes = [D, D, C.c_int64, C.c_int, C.c_int64, C.c_int, C.c_int, C.c_double,
                                      C.c_double, D, U]
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleSolver:
    """Reference CPU solver (restated).  State arrays are ghost-padded AoS
    float64 of shape (m0, m1, m2, 5)."""

    def __init__(self, n, lo, hi, ng=6, eos=(1.4, 0.0, 0.0, 0.0)):
        self.lib = load()
        self.n = tuple(int(v) for v in n)
        self.ng = int(ng)
        nn = (C.c_int64 * 3)(*self.n)
        self.h = self.lib.emref_create(nn, (C.c_double * 3)(*lo), (C.c_double * 3)(*hi), ng,
                                       (C.c_double * 4)(*eos))
        self.m = tuple(v + 2 * ng if v > 1 else 1 for v in self.n)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.emref_destroy(self.h)
            self.h = None

    def set_bc(self, d, side, kind, state=None):
        st = None if state is None else _dp(np.ascontiguousarray(state, dtype=np.float64))
        self.lib.emref_set_bc(self.h, d, side, BC[kind], st)

    def set_opts(self, mode="hybrid", shock="wcns", tol=1e-3, positivity=True, cfl=0.5):
        self.lib.emref_set_opts(self.h, MODE[mode], SHOCK[shock], tol, int(positivity), cfl)

    def set_u(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.m + (5,))
        self.lib.emref_set_u(self.h, _dp(u))

    def get_u(self):
        u = np.empty(self.m + (5,), dtype=np.float64)
        self.lib.emref_get_u(self.h, _dp(u))
        return u

    def fill_ghosts(self):
        self.lib.emref_fill_ghosts(self.h)

    def rhs(self, stage_dt):
        out = np.empty(self.m + (5,), dtype=np.float64)
        st = (C.c_uint64 * 4)()
        self.lib.emref_rhs(self.h, stage_dt, _dp(out), st)
        return out, list(st)

    def step(self, dt):
        self.lib.emref_step(self.h, dt)

    def dt_cfl(self):
        return self.lib.emref_dt_cfl(self.h)

    def run_to(self, t_end):
        return self.lib.emref_run_to(self.h, t_end)

    @property
    def t(self):
        return self.lib.emref_time(self.h)

    def stats(self):
        st = (C.c_uint64 * 4)()
        self.lib.emref_stats(self.h, st)
        return list(st)

    def totals(self):
        out = (C.c_double * 5)()
        self.lib.emref_totals(self.h, out)
        return np.array(list(out))

    def hash(self):
        return self.lib.emref_hash(self.h)

    def interior(self, a):
        g = [self.ng if v > 1 else 0 for v in self.n]
        return a[g[0]:g[0] + self.n[0], g[1]:g[1] + self.n[1], g[2]:g[2] + self.n[2]]
