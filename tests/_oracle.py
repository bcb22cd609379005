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
    lib.emref_init_open.argtypes = [P]
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

    def init_open(self):
        """State of the reference's `open` probe (Gaussian pressure pulse)."""
        self.lib.emref_init_open(self.h)

    def interior(self, a):
        g = [self.ng if v > 1 else 0 for v in self.n]
        return a[g[0]:g[0] + self.n[0], g[1]:g[1] + self.n[1], g[2]:g[2] + self.n[2]]


def _load_gen(lib):
    if getattr(lib, "_gen_ready", False):
        return lib
    P = C.c_void_p
    D = C.POINTER(C.c_double)
    U = C.POINTER(C.c_uint64)
    I64 = C.POINTER(C.c_int64)
    lib.emref_gen_create.restype = P
    lib.emref_gen_create.argtypes = [C.c_int, D, I64, D, D, C.c_int64]
    lib.emref_gen_destroy.argtypes = [P]
    lib.emref_gen_nv.restype = C.c_int
    lib.emref_gen_nv.argtypes = [P]
    lib.emref_gen_set_bc.argtypes = [P, C.c_int, C.c_int, C.c_int, D]
    lib.emref_gen_set_opts.argtypes = [P, C.c_int, C.c_double, C.c_int, C.c_double]
    lib.emref_gen_set_u.argtypes = [P, D]
    lib.emref_gen_get_u.argtypes = [P, D]
    lib.emref_gen_rhs.argtypes = [P, C.c_double, D, U]
    lib.emref_gen_step.argtypes = [P, C.c_double]
    lib.emref_gen_dt_cfl.restype = C.c_double
    lib.emref_gen_dt_cfl.argtypes = [P]
    lib.emref_gen_run_to.restype = C.c_int64
    lib.emref_gen_run_to.argtypes = [P, C.c_double]
    lib.emref_gen_time.restype = C.c_double
    lib.emref_gen_time.argtypes = [P]
    lib.emref_gen_stats.argtypes = [P, U]
    lib.emref_gen_hash.restype = C.c_uint64
    lib.emref_gen_hash.argtypes = [P]
    lib.emref_gen_to_prim.argtypes = [P, D, D]
    lib.emref_gen_init_interface.argtypes = [P]
    lib.emref_gen_eigen.argtypes = [P, D, C.c_int, D, D]
    lib.emref_gen_riemann.argtypes = [P, D, D, C.c_int, D, D]
    lib.emref_gen_phys_flux.argtypes = [P, D, C.c_int, D]
    lib._gen_ready = True
    return lib


# the reference's `interface` probe materials (examples/cpu_validate.rs:101-112):
# water as a stiffened gas (gamma 2, p_inf 1e6) and air (ideal, gamma 1.4)
FIVEEQ_WATER_AIR = (2.0, 1e6, 0.0, 0.0, 1.4, 0.0, 0.0, 0.0, 1.0)


class OracleGen:
    """The oracle's generic solver (emref_gen.h) of the Euler system
    (sys="euler", params = gamma, p_inf, b, q) or of the Kapila five-equation
    two-phase system (sys="fiveeq", params = both phases' gamma, p_inf, b, q
    and the Kapila flag).  Ghost-padded AoS arrays of shape (m0, m1, m2, nv)."""

    def __init__(self, sys, n, lo, hi, params=None, ng=6):
        self.lib = _load_gen(load())
        self.sysid = {"euler": 0, "fiveeq": 1}[sys]
        if params is None:
            params = (1.4, 0.0, 0.0, 0.0) if self.sysid == 0 else FIVEEQ_WATER_AIR
        self.n = tuple(int(v) for v in n)
        self.ng = ng
        self.h = self.lib.emref_gen_create(self.sysid, _dp(np.asarray(params, dtype=np.float64)),
                                           (C.c_int64 * 3)(*self.n), (C.c_double * 3)(*lo), (C.c_double * 3)(*hi), ng)
        self.nv = self.lib.emref_gen_nv(self.h)
        self.shape = tuple(v + 2 * ng if v > 1 else 1 for v in self.n) + (self.nv,)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.emref_gen_destroy(self.h)
            self.h = None

    def set_bc(self, d, side, kind, state=None):
        st = None if state is None else _dp(np.asarray(state, dtype=np.float64))
        self.lib.emref_gen_set_bc(self.h, d, side, BC[kind], st)

    def set_opts(self, mode="hybrid", tol=1e-3, positivity=True, cfl=0.5):
        self.lib.emref_gen_set_opts(self.h, MODE[mode], tol, int(positivity), cfl)

    def set_u(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.shape)
        self.lib.emref_gen_set_u(self.h, _dp(u))

    def get_u(self):
        u = np.empty(self.shape)
        self.lib.emref_gen_get_u(self.h, _dp(u))
        return u

    def rhs(self, stage_dt):
        out = np.empty(self.shape)
        st = (C.c_uint64 * 4)()
        self.lib.emref_gen_rhs(self.h, stage_dt, _dp(out), st)
        return out, list(st)

    def step(self, dt):
        self.lib.emref_gen_step(self.h, dt)

    def dt_cfl(self):
        return self.lib.emref_gen_dt_cfl(self.h)

    def run_to(self, t_end):
        return self.lib.emref_gen_run_to(self.h, t_end)

    @property
    def t(self):
        return self.lib.emref_gen_time(self.h)

    def stats(self):
        st = (C.c_uint64 * 4)()
        self.lib.emref_gen_stats(self.h, st)
        return list(st)

    def hash(self):
        return self.lib.emref_gen_hash(self.h)

    def to_prim(self, u):
        """Primitive states of an array of conserved states (..., nv)."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, self.nv)
        w = np.empty_like(u)
        for r in range(u.shape[0]):
            self.lib.emref_gen_to_prim(self.h, _dp(u[r]), _dp(w[r]))
        return w

    def eigen(self, w, d):
        """(L, R) of the FiveEq primitive system at w along axis d."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        l, r = np.empty((7, 7)), np.empty((7, 7))
        self.lib.emref_gen_eigen(self.h, _dp(w), d, _dp(l), _dp(r))
        return l, r

    def riemann(self, wl, wr, d):
        """HLLC flux and contact speed of two FiveEq primitive states."""
        wl = np.ascontiguousarray(wl, dtype=np.float64)
        wr = np.ascontiguousarray(wr, dtype=np.float64)
        f, ss = np.empty(7), C.c_double()
        self.lib.emref_gen_riemann(self.h, _dp(wl), _dp(wr), d, _dp(f), C.byref(ss))
        return f, ss.value

    def phys_flux(self, w, d):
        w = np.ascontiguousarray(w, dtype=np.float64)
        f = np.empty(7)
        self.lib.emref_gen_phys_flux(self.h, _dp(w), d, _dp(f))
        return f

    def init_interface(self):
        """State of the reference's `interface` probe (FiveEq handles)."""
        self.lib.emref_gen_init_interface(self.h)

    def interior(self, a):
        g = [self.ng if v > 1 else 0 for v in self.n]
        return a[g[0]:g[0] + self.n[0], g[1]:g[1] + self.n[1], g[2]:g[2] + self.n[2]]


def _load_amr(lib):
    if getattr(lib, "_amr_ready", False):
        return lib
    P = C.c_void_p
    D = C.POINTER(C.c_double)
    U = C.POINTER(C.c_uint64)
    I64 = C.POINTER(C.c_int64)
    I32 = C.POINTER(C.c_int32)
    U8 = C.POINTER(C.c_uint8)
    lib.emref_amr_create.restype = P
    lib.emref_amr_create.argtypes = [C.c_int, D, I64, C.c_int64, C.c_int64, C.c_double, C.c_int64, C.c_int64, D, D,
                                     D, I32, D, C.c_int, C.c_double, C.c_int, C.c_double]
    lib.emref_amr_destroy.argtypes = [P]
    for n in ("emref_amr_nblocks", "emref_amr_bcells", "emref_amr_ng"):
        getattr(lib, n).restype = C.c_int64
        getattr(lib, n).argtypes = [P]
    lib.emref_amr_blocks.argtypes = [P, I32, I64, U8]
    lib.emref_amr_get_u.argtypes = [P, D]
    lib.emref_amr_set_u.argtypes = [P, D]
    lib.emref_amr_step.argtypes = [P, C.c_double]
    lib.emref_amr_dt_cfl.restype = C.c_double
    lib.emref_amr_dt_cfl.argtypes = [P]
    lib.emref_amr_time.restype = C.c_double
    lib.emref_amr_time.argtypes = [P]
    lib.emref_amr_restrict_all.argtypes = [P]
    lib.emref_amr_regrid.argtypes = [P]
    lib.emref_amr_set_next_regrid.argtypes = [P, C.c_int64]
    lib.emref_amr_compute_rhs.argtypes = [P, C.c_double, D, U]
    lib.emref_amr_fill_scratch.argtypes = [P, C.c_int64, D]
    lib.emref_amr_details.argtypes = [P, D]
    lib.emref_amr_totals.argtypes = [P, D]
    lib.emref_amr_stats.argtypes = [P, U]
    I8 = C.POINTER(C.c_int8)
    lib.emref_amr_flags_from_details.argtypes = [P, D, I8]
    lib.emref_amr_regrid_refine.argtypes = [P, I8]
    lib.emref_amr_regrid_coarsen.argtypes = [P, I8]
    lib.emref_amr_schedule_regrid.argtypes = [P, C.c_int, I64]
    lib._amr_ready = True
    return lib


AMR_CASE = {"riemann2d": 0, "smooth": 1, "shock_bubble": 2, "vortex": 3, "poly": 4}


class OracleAmr:
    """Reference CPU block AMR (restated), global SSP-RK3 step."""

    def __init__(self, case, params, base, b, levels, lo, hi, eps=1e-3, regrid_every=4, ghosts=5, bc=None,
                 mode="hybrid", cfl=0.4, positivity=True, tol=1e-3, eos=(1.4, 0.0, 0.0, 0.0)):
        self.lib = _load_amr(load())
        prm = np.zeros(16)
        if params is not None:
            prm[:len(params)] = params
        kinds = (C.c_int32 * 6)()
        states = (C.c_double * 30)()
        bc = bc or [["periodic", "periodic"]] * 3
        for d in range(3):
            for s in range(2):
                bb = bc[d][s]
                name, st = (bb, None) if isinstance(bb, str) else (bb[0], bb[1])
                kinds[2 * d + s] = BC[name]
                if st is not None:
                    for q in range(5):
                        states[(2 * d + s) * 5 + q] = st[q]
        self.bc = bc
        self.h = self.lib.emref_amr_create(AMR_CASE[case], _dp(prm), (C.c_int64 * 3)(*base), b, levels, eps,
                                           regrid_every, ghosts, (C.c_double * 3)(*lo), (C.c_double * 3)(*hi),
                                           (C.c_double * 4)(*eos), kinds, states, MODE[mode], cfl, int(positivity),
                                           tol)
        self.base, self.b, self.levels, self.lo, self.hi = tuple(base), b, levels, tuple(lo), tuple(hi)
        self.bd = tuple(b if v > 0 else 1 for v in base)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.emref_amr_destroy(self.h)
            self.h = None

    @property
    def nblocks(self):
        return self.lib.emref_amr_nblocks(self.h)

    @property
    def ng(self):
        return self.lib.emref_amr_ng(self.h)

    def blocks(self):
        n = self.nblocks
        lv = np.empty(n, np.int32)
        idx = np.empty((n, 3), np.int64)
        leaf = np.empty(n, np.uint8)
        self.lib.emref_amr_blocks(self.h, lv.ctypes.data_as(C.POINTER(C.c_int32)),
                                  idx.ctypes.data_as(C.POINTER(C.c_int64)), leaf.ctypes.data_as(C.POINTER(C.c_uint8)))
        return lv, idx, leaf.astype(bool)

    def get_u(self):
        u = np.empty((self.nblocks,) + self.bd + (5,))
        self.lib.emref_amr_get_u(self.h, _dp(u))
        return u

    def set_u(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self.lib.emref_amr_set_u(self.h, _dp(u))

    def step(self, dt):
        self.lib.emref_amr_step(self.h, dt)

    def dt_cfl(self):
        return self.lib.emref_amr_dt_cfl(self.h)

    @property
    def t(self):
        return self.lib.emref_amr_time(self.h)

    def restrict_all(self):
        self.lib.emref_amr_restrict_all(self.h)

    def regrid(self):
        self.lib.emref_amr_regrid(self.h)

    def set_next_regrid(self, n):
        self.lib.emref_amr_set_next_regrid(self.h, n)

    def compute_rhs(self, stage_dt):
        out = np.empty((self.nblocks,) + self.bd + (5,))
        st = (C.c_uint64 * 4)()
        self.lib.emref_amr_compute_rhs(self.h, stage_dt, _dp(out), st)
        return out, list(st)

    def fill_scratch(self, bid):
        ext = tuple(v + 2 * self.ng if v > 1 else 1 for v in self.bd)
        out = np.empty(ext + (5,))
        self.lib.emref_amr_fill_scratch(self.h, bid, _dp(out))
        return out

    def details(self):
        out = np.empty(self.nblocks)
        self.lib.emref_amr_details(self.h, _dp(out))
        return out

    def totals(self):
        out = (C.c_double * 5)()
        self.lib.emref_amr_totals(self.h, out)
        return np.array(list(out))

    def stats(self):
        st = (C.c_uint64 * 4)()
        self.lib.emref_amr_stats(self.h, st)
        return list(st)

    def flags_from_details(self, det):
        det = np.ascontiguousarray(det, dtype=np.float64)
        out = np.empty(self.nblocks, np.int8)
        self.lib.emref_amr_flags_from_details(self.h, _dp(det), out.ctypes.data_as(C.POINTER(C.c_int8)))
        return out

    def regrid_refine(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int8)
        assert f.shape == (self.nblocks,)
        self.lib.emref_amr_regrid_refine(self.h, f.ctypes.data_as(C.POINTER(C.c_int8)))

    def regrid_coarsen(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int8)
        assert f.shape == (self.nblocks,)
        self.lib.emref_amr_regrid_coarsen(self.h, f.ctypes.data_as(C.POINTER(C.c_int8)))

    def schedule_regrid(self, changed=None):
        """schedule_regrid(changed) (None: only read); -> (steps, interval, next)."""
        out = (C.c_int64 * 3)()
        self.lib.emref_amr_schedule_regrid(self.h, -1 if changed is None else int(bool(changed)), out)
        return tuple(out)

    def dx(self, level):
        return tuple((h - l) / (nb * self.b * 2 ** level) if nb > 0 else (h - l)
                     for l, h, nb in zip(self.lo, self.hi, self.base))


def _load_adj(lib):
    if getattr(lib, "_adj_ready", False):
        return lib
    P = C.c_void_p
    D = C.POINTER(C.c_double)
    lib.emadj_create.restype = P
    lib.emadj_create.argtypes = [C.POINTER(C.c_int64), D, D, C.c_int64, D, C.POINTER(C.c_int32), D, C.c_int32]
    lib.emadj_destroy.argtypes = [P]
    lib.emadj_rhs.argtypes = [P, C.c_double, D, D]
    lib.emadj_steps.argtypes = [P, D, C.c_int64, D]
    lib.emadj_vjp_rhs.argtypes = [P, C.c_double, D, D, D]
    lib.emadj_gradient.argtypes = [P, D, C.c_int64, D, D]
    lib.emadj_weno_vjp.argtypes = [D, C.c_double, D]
    lib._adj_ready = True
    return lib


def weno_vjp(v, ob=1.0):
    """Tape VJP of one WENO5-Z interpolation: (d out / d v)^T ob."""
    lib = _load_adj(load())
    vv = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty(5)
    lib.emadj_weno_vjp(_dp(vv), float(ob), _dp(out))
    return out


class OracleAdj:
    """Weno-mode residual and SSP-RK3 step of the restated reference solver,
    written over a scalar type: evaluated in double (bit-identical to
    OracleSolver) or on a reverse-mode tape (reference gradients).  Arrays are
    ghost-padded AoS float64 of shape (m0, m1, m2, 5)."""

    def __init__(self, n, lo, hi, ng=6, eos=(1.4, 0.0, 0.0, 0.0), bc=None, positivity=True):
        self.lib = _load_adj(load())
        self.n = tuple(int(v) for v in n)
        self.ng = int(ng)
        self.m = tuple(v + 2 * ng if v > 1 else 1 for v in self.n)
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
                        states[(2 * d + s) * 5 + q] = st[q]
        self.h = self.lib.emadj_create((C.c_int64 * 3)(*self.n), (C.c_double * 3)(*lo), (C.c_double * 3)(*hi), ng,
                                       (C.c_double * 4)(*eos), kinds, states, int(positivity))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.emadj_destroy(self.h)
            self.h = None

    def _arr(self, a):
        return np.ascontiguousarray(a, dtype=np.float64).reshape(self.m + (5,))

    def rhs(self, stage_dt, u):
        out = np.empty(self.m + (5,))
        self.lib.emadj_rhs(self.h, stage_dt, _dp(self._arr(u)), _dp(out))
        return out

    def steps(self, u, dts):
        u = self._arr(u).copy()
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        self.lib.emadj_steps(self.h, _dp(u), len(dts), _dp(dts))
        return u

    def vjp_rhs(self, stage_dt, u, lam):
        out = np.empty(self.m + (5,))
        self.lib.emadj_vjp_rhs(self.h, stage_dt, _dp(self._arr(u)), _dp(self._arr(lam)), _dp(out))
        return out

    def gradient(self, u0, dts, dj_du_final):
        g = self._arr(dj_du_final).copy()
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        self.lib.emadj_gradient(self.h, _dp(self._arr(u0)), len(dts), _dp(dts), _dp(g))
        return g

    def interior(self, a):
        g = [self.ng if v > 1 else 0 for v in self.n]
        return a[g[0]:g[0] + self.n[0], g[1]:g[1] + self.n[1], g[2]:g[2] + self.n[2]]
