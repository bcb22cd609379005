"""Ensemble of independent uniform-grid solvers over the C ABI.

`HipEnsemble` advances many instances (each its own mesh, domain, state, dt
and CFL rate) together: every axis pass of every instance is one persistent
launch, a step is one replayed HIP graph, and the discrete adjoint batches its
VJPs the same way.  Host arrays are per instance in the CPU layout of
`HipSolver`.

`shock_bubble_2d` builds the planar-shock / light-bubble initial state of the
ensemble workload and its derivative with respect to the shock Mach number.
"""
import ctypes as C

import numpy as np

from . import _lib
from .amr import _bc_arrays
from .solver import MODE, padded_shape

GAMMA = 1.4


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ptrs(arrays):
    return (C.POINTER(C.c_double) * len(arrays))(*[_dp(a) for a in arrays])


def post_shock(mach, g=GAMMA):
    """Rankine-Hugoniot state behind a shock of Mach `mach` moving into
    (rho 1, u 0, p 1), and its derivative with respect to the Mach number:
    (rho2, u2, p2), (drho2, du2, dp2)."""
    m2 = mach * mach
    den = (g - 1.0) * m2 + 2.0
    rho2 = (g + 1.0) * m2 / den
    p2 = 1.0 + 2.0 * g / (g + 1.0) * (m2 - 1.0)
    c1 = np.sqrt(g)
    u2 = mach * c1 * (1.0 - 1.0 / rho2)
    drho2 = (g + 1.0) * 2.0 * mach * den / den ** 2 - (g + 1.0) * m2 * 2.0 * (g - 1.0) * mach / den ** 2
    dp2 = 2.0 * g / (g + 1.0) * 2.0 * mach
    du2 = c1 * (1.0 - 1.0 / rho2) + mach * c1 * drho2 / rho2 ** 2
    return (rho2, u2, p2), (drho2, du2, dp2)


def shock_bubble_2d(n, mach, radius, density, shock_x=0.15, centre=(0.5, 0.5), ng=6):
    """2-D shock-bubble on [0,1]^2 with n x n cells: a shock of Mach `mach` at
    x = shock_x moving into (rho 1, p 1), a circular bubble of the given
    radius and density at `centre` (p 1).  Returns the conserved state (CPU
    layout) and its derivative with respect to `mach`."""
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    (rho2, u2, p2), (drho2, du2, dp2) = post_shock(mach)
    post = X < shock_x
    bub = (X - centre[0]) ** 2 + (Y - centre[1]) ** 2 < radius ** 2
    rho = np.where(post, rho2, np.where(bub, density, 1.0))
    vel = np.where(post, u2, 0.0)
    p = np.where(post, p2, 1.0)
    u = np.zeros(padded_shape((n, n, 1), ng))
    du = np.zeros_like(u)
    ui, dui = u[ng:ng + n, ng:ng + n, 0], du[ng:ng + n, ng:ng + n, 0]
    ui[..., 0] = rho
    ui[..., 1] = rho * vel
    ui[..., 4] = p / (GAMMA - 1.0) + 0.5 * rho * vel * vel
    dui[..., 0] = np.where(post, drho2, 0.0)
    dui[..., 1] = np.where(post, drho2 * u2 + rho2 * du2, 0.0)
    dui[..., 4] = np.where(post, dp2 / (GAMMA - 1.0) + 0.5 * drho2 * u2 * u2 + rho2 * u2 * du2, 0.0)
    return u, du


class HipEnsemble:
    def __init__(self, sizes, states, lo=None, hi=None, ng=6, eos=(GAMMA, 0.0, 0.0, 0.0), bc=None, mode="weno",
                 positivity=True, cfl=0.5, tol=1e-3, batched=True):
        """sizes: per-instance (n0, n1, n2); states: per-instance conserved
        arrays in the CPU layout; lo/hi per instance (default [0,1]^3)."""
        self.lib = _lib.load()
        self.sizes = [tuple(int(v) for v in s) for s in sizes]
        self.ng = int(ng)
        k = len(self.sizes)
        lo = lo or [(0.0, 0.0, 0.0)] * k
        hi = hi or [(1.0, 1.0, 1.0)] * k
        n_arr = np.ascontiguousarray(np.array(self.sizes, dtype=np.int64).ravel())
        lo_arr = np.ascontiguousarray(np.array(lo, dtype=np.float64).ravel())
        hi_arr = np.ascontiguousarray(np.array(hi, dtype=np.float64).ravel())
        self._u = [np.ascontiguousarray(u, dtype=np.float64).reshape(padded_shape(s, ng))
                   for u, s in zip(states, self.sizes)]
        kinds, st = _bc_arrays(bc)
        h = C.c_void_p()
        _lib.check(self.lib.em_ensemble_create(
            C.byref(h), k, n_arr.ctypes.data_as(C.POINTER(C.c_int64)), _dp(lo_arr), _dp(hi_arr), self.ng,
            (C.c_double * 4)(*eos), kinds, st, MODE[mode], int(bool(positivity)), float(cfl), float(tol),
            _ptrs(self._u)))
        self.h = h
        self.set_batched(batched)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.em_ensemble_destroy(self.h)
            self.h = None

    def __len__(self):
        return len(self.sizes)

    def set_batched(self, batched):
        _lib.check(self.lib.em_ensemble_set_batched(self.h, int(bool(batched))))

    def set_tiles(self, cfg):
        _lib.check(self.lib.em_ensemble_set_tiles(self.h, int(cfg)))

    def autotune(self, repeats=3):
        """Times every tile shape on all instances and keeps the fastest; returns it."""
        out = C.c_int32()
        _lib.check(self.lib.em_ensemble_autotune(self.h, int(repeats), C.byref(out)))
        return out.value

    def run_steps(self, nsteps):
        _lib.check(self.lib.em_ensemble_run_steps(self.h, int(nsteps)))

    def upload(self, i, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(padded_shape(self.sizes[i], self.ng))
        _lib.check(self.lib.em_ensemble_upload(self.h, int(i), _dp(u)))

    def download(self, i):
        u = np.empty(padded_shape(self.sizes[i], self.ng))
        _lib.check(self.lib.em_ensemble_download(self.h, int(i), _dp(u)))
        return u

    def time(self, i):
        t = C.c_double()
        _lib.check(self.lib.em_ensemble_time(self.h, int(i), C.byref(t)))
        return t.value

    def forward(self, nsteps, checkpoint_every):
        self.nsteps = int(nsteps)
        _lib.check(self.lib.em_ensemble_forward(self.h, self.nsteps, int(checkpoint_every)))

    def dts(self, i):
        out = np.empty(self.nsteps)
        _lib.check(self.lib.em_ensemble_dts(self.h, int(i), _dp(out)))
        return out

    def backward(self, dj_final):
        dj = [np.ascontiguousarray(d, dtype=np.float64).reshape(padded_shape(s, self.ng))
              for d, s in zip(dj_final, self.sizes)]
        out = [np.empty(padded_shape(s, self.ng)) for s in self.sizes]
        _lib.check(self.lib.em_ensemble_backward(self.h, _ptrs(dj), _ptrs(out)))
        return out
