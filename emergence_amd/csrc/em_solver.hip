// Host side of the uniform-grid solver: the GpuSolver object behind the C ABI
// (include/em_hip.h), its small kernels (CFL rate, device-resident dt, ghost
// padding for download), graph-replayed device-driven stepping, the tile
// autotuner and the workload initialisers.
//
// Reference: fluid/crates/compressible-cuda/src/lib.rs:155-477 (GpuSolver),
// fluid/crates/compressible/src/solver.rs (Solver: step, dt_cfl, run_to).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/em_hip.h"
#include "em_adjoint.h"
#include "em_pass.h"
#include "em_types.h"

namespace em {

static thread_local std::string g_last_error;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(EM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHK(x) hip_check((x), #x)

// ------------------------------------------------------------------ kernels

__global__ void k_set_dt(DevScalars* sc, double dt) { sc->dt = dt; }

// Zero fill on the solver's stream (a non-blocking stream does not order
// against work on the NULL stream, so no NULL-stream memset is used).
template <class Real>
__global__ void k_fill_zero(Real* p, long long n) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256ll) p[i] = Real(0);
}

__global__ void k_set_run(DevScalars* sc, double t_end) {
    sc->t_end = t_end;
    sc->active = 1;
    sc->err = 0;
    sc->steps = 0;
}

__global__ void k_zero_rate(DevScalars* sc) {
    if (threadIdx.x < kSlots) sc->rate[threadIdx.x][0] = 0ull;
}

// max over the rate slots (IEEE bits of non-negative doubles order like the
// values)
__device__ __forceinline__ unsigned long long rate_bits(const DevScalars* sc) {
    unsigned long long m = 0ull;
    for (int i = 0; i < kSlots; ++i) m = sc->rate[i][0] > m ? sc->rate[i][0] : m;
    return m;
}

// t += dt after a step (solver.rs:737)
__global__ void k_advance(DevScalars* sc) {
    if (sc->active == 0) return;
    sc->t = sc->t + sc->dt;
    sc->steps += 1;
}

// Records the step's dt at hist[steps] (the forward run's dt history, kept
// on the device so that steps can be replayed as a graph).
__global__ void k_record_dt(const DevScalars* sc, double* hist) {
    if (sc->active == 0) return;
    hist[sc->steps] = sc->dt;
}

// dt of the next step from the CFL rate of the current state.
//   run_to == 0: dt = cfl / rate (fixed number of steps)
//   run_to == 1: Solver::run_to rule (solver.rs:758-768): stop once
//                t >= t_end - 1e-10 |t_end|; dt = min(cfl / rate, t_end - t);
//                collapse if dt is not finite or <= 1e-12 max(|t_end|, 1e-300)
__global__ void k_dt_from_rate(DevScalars* sc, double cfl, int run_to) {
    if (sc->active == 0) return;
    const double rate = __longlong_as_double((long long)rate_bits(sc));
    double dt = cfl / rate;
    if (run_to) {
        const double te = sc->t_end;
        if (!(sc->t < te - 1e-10 * fabs(te))) {
            sc->active = 0;
            return;
        }
        dt = fmin(dt, te - sc->t);
        if (!(isfinite(dt) && dt > 1e-12 * fmax(fabs(te), 1e-300))) {
            sc->err |= 1;
            sc->active = 0;
            return;
        }
    } else if (!(isfinite(dt) && dt > 0.0)) {
        sc->err |= 1;
        sc->active = 0;
        return;
    }
    sc->dt = dt;
    for (int i = 0; i < kSlots; ++i) sc->rate[i][0] = 0ull;
}

// CFL rate max over interior cells: dt_cfl (solver.rs:478-523).  Wave
// reduction by shuffles, one atomicMax per workgroup on the IEEE bits of a
// non-negative double (a float rate is widened exactly).
template <class Real>
__global__ __launch_bounds__(256) void k_rate(const Real* __restrict__ U, long long ncell, EosT<Real> e, Real dx0,
                                              Real dx1, Real dx2, int a0, int a1, int a2, DevScalars* sc) {
    __shared__ double red[4];
    double r = 0.0;
    for (long long c = blockIdx.x * 256ll + threadIdx.x; c < ncell; c += (long long)gridDim.x * 256ll) {
        Real u[5], w[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) u[k] = U[k * ncell + c];
        to_prim(e, u, w);
        const Real cs = sound_speed(e, w[0], w[4]);
        Real ac = Real(-0.0);
        if (a0) ac += (fabs(w[1]) + cs) / dx0;
        if (a1) ac += (fabs(w[2]) + cs) / dx1;
        if (a2) ac += (fabs(w[3]) + cs) / dx2;
        r = fmax(r, (double)ac);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r = fmax(r, __shfl_xor(r, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        m = m > 0.0 ? m : 0.0;
        atomicMax(&sc->rate[blockIdx.x % kSlots][0], (unsigned long long)__double_as_longlong(m));
    }
}

// Interior SoA -> ghost-padded SoA (padded extents m[3], ghost offsets g[3]).
template <class Real>
__global__ void k_scatter(const Real* __restrict__ U, long long ncell, int n0, int n1, int n2,
                          Real* __restrict__ Pd, long long len, int m1, int m2, int g0, int g1, int g2) {
    const long long c = blockIdx.x * 256ll + threadIdx.x;
    if (c >= ncell) return;
    const int k = (int)(c % n2);
    const int j = (int)((c / n2) % n1);
    const int i = (int)(c / ((long long)n1 * n2));
    const long long q = ((long long)(i + g0) * m1 + (j + g1)) * m2 + (k + g2);
#pragma unroll
    for (int v = 0; v < 5; ++v) Pd[v * len + q] = U[v * ncell + c];
}

// Ghost sweep along axis D of a ghost-padded SoA array: one thread per
// pencil over the whole padded transverse plane, layers gi = 0..ng-1 low and
// high (solver.rs:278-348).  Sweeps run x, y, z in turn so edges and corners
// come out as in the reference.
template <int D, class Real>
__global__ void k_ghost_sweep(Real* __restrict__ Pd, long long len, int m0, int m1, int m2, int n, int ng,
                              EosT<Real> e, int k0, int k1, const Real* __restrict__ st) {
    const int mm[3] = {m0, m1, m2};
    constexpr int E0 = D == 0 ? 1 : 0;
    constexpr int E1 = D == 2 ? 1 : 2;
    const long long p = blockIdx.x * 128ll + threadIdx.x;
    if (p >= (long long)mm[E0] * mm[E1]) return;
    const int a = (int)(p / mm[E1]), b = (int)(p % mm[E1]);
    const long long sd[3] = {(long long)m1 * m2, (long long)m2, 1ll};
    const long long base = (long long)a * sd[E0] + (long long)b * sd[E1];
    const int kind[2] = {k0, k1};
    Real state[2][5];
#pragma unroll
    for (int q = 0; q < 10; ++q) state[q / 5][q % 5] = st[q];
    auto ld = [&](int x, Real* o) {
#pragma unroll
        for (int v = 0; v < 5; ++v) o[v] = Pd[v * len + base + (long long)(ng + x) * sd[D]];
    };
    for (int gi = 0; gi < ng; ++gi) {
        Real u[5];
        bc_ghost<D>(e, n, gi - ng, kind, state, ld, u);
        const long long lo = base + (long long)gi * sd[D];
#pragma unroll
        for (int v = 0; v < 5; ++v) Pd[v * len + lo] = u[v];
        bc_ghost<D>(e, n, n + gi, kind, state, ld, u);
        const long long hi = base + (long long)(n + ng + gi) * sd[D];
#pragma unroll
        for (int v = 0; v < 5; ++v) Pd[v * len + hi] = u[v];
    }
}

// ------------------------------------------------------------------ solver

// Mesh, boundary conditions, scheme options and tile shapes of one uniform
// grid: everything an axis pass needs besides its buffers.  Validates its
// arguments and allocates nothing (the allocation-free residual entry point
// builds one on the stack).
struct Grid {
    int64_t n[3];
    double lo[3], dx[3];
    int ng;
    Eos eos;
    int bc_kind[3][2];
    double bc_state[3][2][5];
    int mode, positivity;
    double cfl, tol;
    int cfg[3] = {6, 6, 6};
    long long ncell;
    long long stride[3];
    int nactive;
    int act[3];

    bool active(int d) const { return n[d] > 1; }
    int m(int d) const { return active(d) ? (int)(n[d] + 2 * ng) : 1; }
    int g(int d) const { return active(d) ? ng : 0; }
    long long len() const { return (long long)m(0) * m(1) * m(2); }

    Grid(const int64_t nn[3], const double lo_[3], const double hi[3], int ng_, const double e[4],
         const int32_t bk[6], const double bs[30], int mode_, int pos, double cfl_, double tol_) {
        for (int d = 0; d < 3; ++d) {
            if (nn[d] < 1) throw Error(EM_ERR_ARG, "mesh extents must be >= 1");
            if (nn[d] > (1 << 30)) throw Error(EM_ERR_ARG, "mesh extents must be <= 2^30");
            n[d] = nn[d];
            lo[d] = lo_[d];
            dx[d] = (hi[d] - lo_[d]) / (double)nn[d];  // Mesh::new, solver.rs:32
        }
        ng = ng_;
        eos = Eos{e[0], e[1], e[2], e[3]};
        mode = mode_;
        positivity = pos;
        cfl = cfl_;
        tol = tol_;
        if (mode < 0 || mode > 2) throw Error(EM_ERR_ARG, "mode must be 0 (Hybrid), 1 (Weno) or 2 (Central)");
        if (ng < (mode == MODE_CENTRAL ? 3 : 5))
            throw Error(EM_ERR_ARG, "ghost layers: need ng >= 5 for Hybrid/Weno, >= 3 for Central");
        nactive = 0;
        for (int d = 0; d < 3; ++d) {
            act[d] = active(d) ? 1 : 0;
            nactive += act[d];
            for (int sd = 0; sd < 2; ++sd) {
                bc_kind[d][sd] = bk[2 * d + sd];
                if (bc_kind[d][sd] < 0 || bc_kind[d][sd] > 4) throw Error(EM_ERR_ARG, "boundary kind out of range");
                for (int q = 0; q < 5; ++q) bc_state[d][sd][q] = bs ? bs[(2 * d + sd) * 5 + q] : 0.0;
            }
            if (active(d) && n[d] < 2 && (bc_kind[d][0] == BC_CHARACTERISTIC || bc_kind[d][1] == BC_CHARACTERISTIC))
                throw Error(EM_ERR_ARG, "characteristic boundaries need at least 2 cells");
        }
        if (nactive == 0) throw Error(EM_ERR_ARG, "mesh has no active axis");
        ncell = n[0] * n[1] * n[2];
        if (ncell > (1ll << 40)) throw Error(EM_ERR_ARG, "mesh has more than 2^40 cells");
        stride[0] = n[1] * n[2];
        stride[1] = n[2];
        stride[2] = 1;
    }

    PassArgs pass_args(int d, DevScalars* sc) const { return pass_args_t<double>(d, sc); }
    // The pass arguments in real type Real (mesh spacing, EOS, boundary
    // states and tolerance rounded to Real, as the reference's R::from_f64).
    template <class Real>
    PassArgsT<Real> pass_args_t(int d, DevScalars* sc) const {
        PassArgsT<Real> a{};
        a.sc = sc;
        for (int e = 0; e < 3; ++e) {
            a.n[e] = (int)n[e];
            a.stride[e] = stride[e];
            a.dx_all[e] = Real(dx[e]);
            a.active[e] = act[e];
        }
        a.ncell = ncell;
        a.eos = eos_as<Real>(eos);
        for (int sd = 0; sd < 2; ++sd) {
            a.bc_kind[sd] = bc_kind[d][sd];
            for (int q = 0; q < 5; ++q) a.bc_state[sd][q] = Real(bc_state[d][sd][q]);
        }
        a.inv_dx = Real(1.0 / dx[d]);
        a.dx_d = dx[d];
        a.ndim_pos = positivity ? (double)nactive : 0.0;
        a.tol = Real(tol);
        // transverse axes: the pencils of a workgroup are neighbours along the
        // innermost active transverse axis
        int t[2], nt = 0;
        for (int e = 0; e < 3; ++e)
            if (e != d) t[nt++] = e;
        if (active(t[1]) || !active(t[0])) {
            a.fast_axis = t[1];
            a.slow_axis = t[0];
        } else {
            a.fast_axis = t[0];
            a.slow_axis = t[1];
        }
        a.dfast = stride[d] == 1 ? 1 : 0;
        const TileCfg tc = kTileCfgs[cfg[d]];
        a.ntile = (int)((n[d] + tc.T - 1) / tc.T);
        a.ngroup = (int)((n[a.fast_axis] + tc.P - 1) / tc.P);
        return a;
    }

    uint64_t faces_per_stage() const {
        uint64_t f = 0;
        for (int d = 0; d < 3; ++d)
            if (active(d)) f += (uint64_t)(n[d] + 1) * (ncell / n[d]);
        return f;
    }
};

// Device state, stepping, CFL, graphs and autotuner of one uniform grid in
// real type Real: Real = double is the FP64 solver (the adjoint builds on
// it, below); Real = float is the FP32 variant (the reference's
// GpuSolver<f32>), whose host interface still takes and returns doubles.
template <class Real>
struct SolverCore : Grid {
    int precise = 1;  // reference arithmetic by default
    hipStream_t stream = nullptr;
    hipStream_t rstream = nullptr;  // the backward sweep's segment recompute (default priority)
    Real* U[3] = {nullptr, nullptr, nullptr};  // U^n, stage 1, stage 2
    Real* R = nullptr;
    Real* pad = nullptr;
    Real* bcdev = nullptr;  // [3][2][5] boundary states for the ghost sweep
    double* dt_hist = nullptr;  // device dt history of a CFL-driven forward run
    int64_t dt_hist_cap = 0;
    DevScalars* sc = nullptr;
    // per run mode: 0 run_steps, 1 run_to, 2 forward run recording dt
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
    int graph_variant[3] = {-1, -1, -1};
    int graph_cfg[3][3];
    bool rate_valid = false;

    SolverCore(const int64_t nn[3], const double lo_[3], const double hi[3], int ng_, const double e[4],
               const int32_t bk[6], const double bs[30], int mode_, int pos, double cfl_, double tol_, double t0)
        : Grid(nn, lo_, hi, ng_, e, bk, bs, mode_, pos, cfl_, tol_) {
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        for (int b = 0; b < 3; ++b) HIPCHK(hipMalloc(&U[b], sizeof(Real) * 5 * ncell));
        HIPCHK(hipMalloc(&R, sizeof(Real) * 5 * ncell));
        HIPCHK(hipMalloc(&sc, sizeof(DevScalars)));
        HIPCHK(hipMalloc(&bcdev, sizeof(Real) * 30));
        std::vector<Real> bsr(30);
        for (int q = 0; q < 30; ++q) bsr[q] = Real((&bc_state[0][0][0])[q]);
        HIPCHK(hipMemcpyAsync(bcdev, bsr.data(), sizeof(Real) * 30, hipMemcpyHostToDevice, stream));
        std::vector<DevScalars> h(1);
        std::memset(h.data(), 0, sizeof(DevScalars));
        h[0].t = t0;
        h[0].active = 1;
        HIPCHK(hipMemcpyAsync(sc, h.data(), sizeof(DevScalars), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        // default tile: 59 cells x 8 pencils, 512 threads (one midpoint per
        // thread, two workgroups per CU); the autotuner may pick another
        for (int d = 0; d < 3; ++d) cfg[d] = 6;
    }

    ~SolverCore() {
        if (dt_hist) (void)hipFree(dt_hist);
        for (auto& gx : graph)
            if (gx) (void)hipGraphExecDestroy(gx);
        for (auto* p : U)
            if (p) (void)hipFree(p);
        if (R) (void)hipFree(R);
        if (pad) (void)hipFree(pad);
        if (sc) (void)hipFree(sc);
        if (bcdev) (void)hipFree(bcdev);
        if (stream) (void)hipStreamDestroy(stream);
        if (rstream) (void)hipStreamDestroy(rstream);
    }

    void upload(const double* u) {
        std::vector<Real> h(5 * ncell);
        const int m1 = m(1), m2 = m(2);
        for (long long i = 0; i < n[0]; ++i)
            for (long long j = 0; j < n[1]; ++j)
                for (long long k = 0; k < n[2]; ++k) {
                    const long long q = ((i + g(0)) * m1 + (j + g(1))) * m2 + (k + g(2));
                    const long long c = (i * n[1] + j) * n[2] + k;
                    for (int v = 0; v < 5; ++v) h[v * ncell + c] = Real(u[q * 5 + v]);
                }
        HIPCHK(hipMemcpyAsync(U[0], h.data(), sizeof(Real) * 5 * ncell, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        rate_valid = false;
    }

    // Ghost-padded SoA image of U[0] in `pad` (ghosts by the reference's
    // axis-ordered sweeps).
    void fill_pad() {
        const long long L = len();
        if (!pad) {
            HIPCHK(hipMalloc(&pad, sizeof(Real) * 5 * L));
            const long long nz = 5 * L;
            hipLaunchKernelGGL(k_fill_zero<Real>, dim3((unsigned)std::min<long long>((nz + 255) / 256, 4096)), dim3(256), 0,
                               stream, pad, nz);
            HIPCHK(hipGetLastError());
        }
        const unsigned nb = (unsigned)((ncell + 255) / 256);
        hipLaunchKernelGGL(k_scatter<Real>, dim3(nb), dim3(256), 0, stream, U[0], ncell, (int)n[0], (int)n[1], (int)n[2],
                           pad, L, m(1), m(2), g(0), g(1), g(2));
        HIPCHK(hipGetLastError());
        for (int d = 0; d < 3; ++d) {
            if (!active(d)) continue;
            const int e0 = d == 0 ? 1 : 0, e1 = d == 2 ? 1 : 2;
            const long long np = (long long)m(e0) * m(e1);
            const unsigned gb = (unsigned)((np + 127) / 128);
            const Real* st = bcdev + d * 10;
            const EosT<Real> er = eos_as<Real>(eos);
            if (d == 0)
                hipLaunchKernelGGL((k_ghost_sweep<0, Real>), dim3(gb), dim3(128), 0, stream, pad, L, m(0), m(1), m(2),
                                   (int)n[0], ng, er, bc_kind[0][0], bc_kind[0][1], st);
            else if (d == 1)
                hipLaunchKernelGGL((k_ghost_sweep<1, Real>), dim3(gb), dim3(128), 0, stream, pad, L, m(0), m(1), m(2),
                                   (int)n[1], ng, er, bc_kind[1][0], bc_kind[1][1], st);
            else
                hipLaunchKernelGGL((k_ghost_sweep<2, Real>), dim3(gb), dim3(128), 0, stream, pad, L, m(0), m(1), m(2),
                                   (int)n[2], ng, er, bc_kind[2][0], bc_kind[2][1], st);
            HIPCHK(hipGetLastError());
        }
    }

    void download(double* u) {
        fill_pad();
        const long long L = len();
        std::vector<Real> h(5 * L);
        HIPCHK(hipMemcpyAsync(h.data(), pad, sizeof(Real) * 5 * L, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (long long q = 0; q < L; ++q)
            for (int v = 0; v < 5; ++v) u[q * 5 + v] = (double)h[v * L + q];
    }

    PassArgsT<Real> pass_args(int d) const { return Grid::pass_args_t<Real>(d, sc); }

    hipError_t launch(int d, const PassArgsT<Real>& a, hipStream_t st = nullptr) {
        if (!st) st = stream;
        return precise ? precise::launch_pass(d, mode, cfg[d], a, st) : fast::launch_pass(d, mode, cfg[d], a, st);
    }

    // One stage: residual of Uin over the active axes in order, fused with
    // the SSP-RK3 combination on the last axis (rk_stage 0: residual only).
    void stage(const Real* Uin, Real* Uout, int rk_stage, bool want_rate, int stage_dt_host, double stage_dt,
               hipEvent_t* ev, const Real* base = nullptr, hipStream_t st = nullptr) {
        if (!st) st = stream;
        int last = -1;
        for (int d = 0; d < 3; ++d)
            if (active(d)) last = d;
        bool first = true;
        for (int d = 0; d < 3; ++d) {
            if (!active(d)) continue;
            PassArgsT<Real> a = pass_args(d);
            a.U = Uin;
            a.R = R;
            a.U0 = base ? base : U[0];
            a.Uout = Uout;
            a.first = first ? 1 : 0;
            a.rk_stage = d == last ? rk_stage : 0;
            a.want_rate = (d == last && want_rate) ? 1 : 0;
            a.stage_dt_host = stage_dt_host;
            a.stage_dt = stage_dt;
            if (ev) HIPCHK(hipEventRecord(ev[2 * d], st));
            HIPCHK(launch(d, a, st));
            if (ev) HIPCHK(hipEventRecord(ev[2 * d + 1], st));
            first = false;
        }
    }

    // SSP-RK3 step with the dt already in sc->dt (solver.rs:659-738):
    // U1 = U + dt L(U); U2 = 3/4 U + 1/4 (U1 + dt L(U1));
    // U <- U/3 + 2/3 (U2 + dt L(U2)), the last stage also leaving the CFL
    // rate of the new state in sc->rate.
    void step_kernels(std::vector<hipEvent_t>* evs = nullptr) {
        hipEvent_t* e[3] = {nullptr, nullptr, nullptr};
        if (evs)
            for (int s = 0; s < 3; ++s) e[s] = evs->data() + 6 * s;
        stage(U[0], U[1], 1, false, 0, 0.0, e[0]);
        stage(U[1], U[2], 2, false, 0, 0.0, e[1]);
        hipLaunchKernelGGL(k_zero_rate, dim3(1), dim3(kSlots), 0, stream, sc);
        stage(U[2], U[0], 3, true, 0, 0.0, e[2]);
        hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, stream, sc);
        HIPCHK(hipGetLastError());
    }

    void reset_run(double t_end) {
        hipLaunchKernelGGL(k_set_run, dim3(1), dim3(1), 0, stream, sc, t_end);
        HIPCHK(hipGetLastError());
    }

    void compute_rate() {
        hipLaunchKernelGGL(k_zero_rate, dim3(1), dim3(kSlots), 0, stream, sc);
        long long nb = std::min<long long>((ncell + 255) / 256, 2048);
        hipLaunchKernelGGL(k_rate<Real>, dim3((unsigned)nb), dim3(256), 0, stream, U[0], ncell, eos_as<Real>(eos),
                           Real(dx[0]), Real(dx[1]), Real(dx[2]), act[0], act[1], act[2], sc);
        HIPCHK(hipGetLastError());
        rate_valid = true;
    }

    // Host view of the device scalars with the slots reduced.
    struct Scalars {
        double dt, t;
        int active, err;
        long long steps;
        uint64_t rate_bits;
        uint64_t counters[3];
    };
    Scalars read_scalars() {
        std::vector<DevScalars> h(1);
        HIPCHK(hipMemcpyAsync(h.data(), sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        Scalars r{h[0].dt, h[0].t, h[0].active, h[0].err, h[0].steps, 0, {0, 0, 0}};
        for (int i = 0; i < kSlots; ++i) {
            r.rate_bits = std::max<uint64_t>(r.rate_bits, h[0].rate[i][0]);
            for (int q = 0; q < 3; ++q) r.counters[q] += h[0].counters[q][i][0];
        }
        return r;
    }
    // Restores the counters (every slot) saved before a timing-only run.
    void restore_counters(const std::vector<DevScalars>& saved) {
        HIPCHK(hipMemcpyAsync(&sc->counters[0][0][0], &saved[0].counters[0][0][0], sizeof(saved[0].counters),
                              hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    std::vector<DevScalars> save_scalars() {
        std::vector<DevScalars> h(1);
        HIPCHK(hipMemcpyAsync(h.data(), sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return h;
    }

    void rhs(double stage_dt, double* out, uint64_t stats[4]) {
        Scalars before = read_scalars();
        reset_run(0.0);
        stage(U[0], nullptr, 0, false, 1, stage_dt, nullptr);
        std::vector<Real> h(5 * ncell);
        HIPCHK(hipMemcpyAsync(h.data(), R, sizeof(Real) * 5 * ncell, hipMemcpyDeviceToHost, stream));
        Scalars after = read_scalars();
        const long long L = len();
        std::fill(out, out + 5 * L, 0.0);
        const int m1 = m(1), m2 = m(2);
        for (long long i = 0; i < n[0]; ++i)
            for (long long j = 0; j < n[1]; ++j)
                for (long long k = 0; k < n[2]; ++k) {
                    const long long q = ((i + g(0)) * m1 + (j + g(1))) * m2 + (k + g(2));
                    const long long c = (i * n[1] + j) * n[2] + k;
                    for (int v = 0; v < 5; ++v) out[q * 5 + v] = (double)h[v * ncell + c];
                }
        if (stats) {
            stats[0] = faces_per_stage();
            for (int q = 0; q < 3; ++q) stats[1 + q] = after.counters[q] - before.counters[q];
        }
    }

    // one step with the given dt, queued on the stream (no host wait)
    void step_async(double dt) {
        reset_run(0.0);
        hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, stream, sc, dt);
        step_kernels();
    }
    void step(double dt) {
        step_async(dt);
        HIPCHK(hipStreamSynchronize(stream));
    }
    // One step of the backward sweep's recompute, dt given: the stages are
    // written straight into the store (u1, u2; unext = U^{n+1}, or null to
    // skip the last stage), with no CFL rate and no clock update (the sweep
    // restores the scalars at its end).
    void recompute_step(const Real* un, Real* u1, Real* u2, Real* unext, double dt, hipStream_t st) {
        hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, st, sc, dt);
        HIPCHK(hipGetLastError());
        stage(un, u1, 1, false, 0, 0.0, nullptr, un, st);
        stage(u1, u2, 2, false, 0, 0.0, nullptr, un, st);
        if (unext) stage(u2, unext, 3, false, 0, 0.0, nullptr, un, st);
    }

    double dt_cfl() {
        compute_rate();
        Scalars h = read_scalars();
        const double rate = __builtin_bit_cast(double, h.rate_bits);
        return cfl / rate;  // solver.rs:517
    }

    // One step driven by device-resident dt: dt kernel + stages + advance.
    // Mode 2 (forward run) also records dt in dt_hist[steps].
    void device_step(int run_mode) {
        hipLaunchKernelGGL(k_dt_from_rate, dim3(1), dim3(1), 0, stream, sc, cfl, run_mode == 1 ? 1 : 0);
        if (run_mode == 2) hipLaunchKernelGGL(k_record_dt, dim3(1), dim3(1), 0, stream, sc, dt_hist);
        step_kernels();
    }

    hipGraphExec_t get_graph(int run_mode) {
        bool same = graph[run_mode] && graph_variant[run_mode] == precise;
        for (int d = 0; d < 3 && same; ++d) same = graph_cfg[run_mode][d] == cfg[d];
        if (same) return graph[run_mode];
        if (graph[run_mode]) {
            HIPCHK(hipGraphExecDestroy(graph[run_mode]));
            graph[run_mode] = nullptr;
        }
        hipGraph_t gr;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        device_step(run_mode);
        HIPCHK(hipStreamEndCapture(stream, &gr));
        HIPCHK(hipGraphInstantiate(&graph[run_mode], gr, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(gr));
        graph_variant[run_mode] = precise;
        for (int d = 0; d < 3; ++d) graph_cfg[run_mode][d] = cfg[d];
        return graph[run_mode];
    }

    int64_t run_to(double t_end) {
        reset_run(t_end);
        compute_rate();
        hipGraphExec_t gx = get_graph(1);
        const int chunk = 16;
        for (;;) {
            for (int q = 0; q < chunk; ++q) HIPCHK(hipGraphLaunch(gx, stream));
            Scalars h = read_scalars();
            if (h.err & 1) throw Error(EM_ERR_COLLAPSE, "time step collapsed: the solution lost admissibility");
            if (!h.active) return h.steps;
        }
    }

    void run_steps(int64_t nsteps, double pass_ms[3]) {
        reset_run(0.0);
        compute_rate();
        if (!pass_ms) {
            hipGraphExec_t gx = get_graph(0);
            for (int64_t q = 0; q < nsteps; ++q) HIPCHK(hipGraphLaunch(gx, stream));
        } else {
            std::vector<hipEvent_t> ev(18 * (size_t)nsteps);
            for (auto& x : ev) HIPCHK(hipEventCreate(&x));
            for (int64_t q = 0; q < nsteps; ++q) {
                std::vector<hipEvent_t> step_ev(ev.begin() + 18 * q, ev.begin() + 18 * (q + 1));
                hipLaunchKernelGGL(k_dt_from_rate, dim3(1), dim3(1), 0, stream, sc, cfl, 0);
                step_kernels(&step_ev);
            }
            HIPCHK(hipStreamSynchronize(stream));
            for (int d = 0; d < 3; ++d) {
                double tot = 0.0;
                int cnt = 0;
                if (active(d)) {
                    for (int64_t q = 0; q < nsteps; ++q)
                        for (int s = 0; s < 3; ++s) {
                            float ms = 0.f;
                            HIPCHK(hipEventElapsedTime(&ms, ev[18 * q + 6 * s + 2 * d], ev[18 * q + 6 * s + 2 * d + 1]));
                            tot += ms;
                            ++cnt;
                        }
                }
                pass_ms[d] = cnt ? tot / cnt : 0.0;
            }
            for (auto& x : ev) (void)hipEventDestroy(x);
        }
        Scalars h = read_scalars();
        if (h.err & 1) throw Error(EM_ERR_COLLAPSE, "time step collapsed: the solution lost admissibility");
    }

    // Times every tile shape of every active axis on the residual pass of
    // the current state (limiter on, with the state's CFL dt) and keeps the
    // fastest; the state is untouched and the counters are restored.
    void autotune(int repeats, int32_t out[3]) {
        std::vector<DevScalars> saved = save_scalars();
        const double dt = dt_cfl();
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        reset_run(0.0);
        for (int d = 0; d < 3; ++d) {
            if (!active(d)) continue;
            double best = 1e300;
            int bestc = cfg[d];
            for (int c = 0; c < kNumTileCfgs; ++c) {
                if (kTileCfgs[c].T > 2 * n[d] && c != 0) continue;
                cfg[d] = c;
                PassArgsT<Real> a = pass_args(d);
                a.U = U[0];
                a.R = R;
                a.U0 = U[0];
                a.Uout = nullptr;
                a.first = 1;
                a.rk_stage = 0;
                a.stage_dt_host = 1;
                a.stage_dt = std::isfinite(dt) ? dt : 0.0;
                HIPCHK(launch(d, a));  // warm-up
                HIPCHK(hipEventRecord(e0, stream));
                for (int r = 0; r < repeats; ++r) HIPCHK(launch(d, a));
                HIPCHK(hipEventRecord(e1, stream));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    bestc = c;
                }
            }
            cfg[d] = bestc;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        restore_counters(saved);
        if (out)
            for (int d = 0; d < 3; ++d) out[d] = cfg[d];
    }

};

// The FP64 solver: SolverCore<double> plus the discrete adjoint.
struct Solver : SolverCore<double> {
    Solver(const int64_t nn[3], const double lo_[3], const double hi[3], int ng_, const double e[4],
           const int32_t bk[6], const double bs[30], int mode_, int pos, double cfl_, double tol_, double t0)
        : SolverCore<double>(nn, lo_, hi, ng_, e, bk, bs, mode_, pos, cfl_, tol_, t0) {}
    ~Solver() { free_adjoint(); }

    // ---------------------------------------------------------- adjoint
    // Checkpoints of the last forward run (every ck_every steps), its dt
    // sequence, and the device buffers of the backward sweep.
    std::vector<double*> ckpt;  // kept across forward runs; the last run uses nck_used
    int64_t nck_used = 0;
    std::vector<double> fwd_dts;
    int ck_every = 0;
    double* adjbuf = nullptr;  // A, B, Z, T: 4 x 5 ncell
    int adj_tile_cap = kAdjTileCells;  // interior cells per adjoint tile at most
    double* store = nullptr;   // u_n, u1, u2 of every step of one segment
    double* store2 = nullptr;  // a second segment (recomputed while the first is reversed), if it fits
    long long store_steps = 0;
    double* final_state = nullptr;
    // checkpoint_every = 0: u_n, u1, u2 of every step of the forward run,
    // kept in device memory, so that the backward recomputes nothing
    bool fwd_all = false;
    double* hist = nullptr;
    long long hist_steps = 0;

    void free_adjoint() {
        for (auto* p : ckpt) (void)hipFree(p);
        ckpt.clear();
        nck_used = 0;
        if (dt_hist) (void)hipFree(dt_hist);
        dt_hist = nullptr;
        dt_hist_cap = 0;
        if (adjbuf) (void)hipFree(adjbuf);
        if (store) (void)hipFree(store);
        if (store2) (void)hipFree(store2);
        if (final_state) (void)hipFree(final_state);
        if (hist) (void)hipFree(hist);
        adjbuf = store = store2 = final_state = hist = nullptr;
        store_steps = 0;
        hist_steps = 0;
        fwd_all = false;
    }

    void adj_check() {
        if (mode != MODE_WENO) throw Error(EM_ERR_ARG, "the adjoint is implemented for FluxMode::Weno (mode 1)");
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, dev));
        for (int d = 0; d < 3; ++d) {
            if (!active(d)) continue;
            for (int sd = 0; sd < 2; ++sd) {
                if (bc_kind[d][sd] == BC_CHARACTERISTIC)
                    throw Error(EM_ERR_ARG, "the adjoint does not support characteristic boundaries");
                if (bc_kind[d][sd] == BC_REFLECT && n[d] < 5)
                    throw Error(EM_ERR_ARG, "the adjoint needs at least 5 cells on a reflecting axis");
            }
            const int nt = adj_tiles((int)n[d], bc_kind[d][0], bc_kind[d][1], adj_tile_cap);
            const size_t need = adj_lds_bytes(adj_tile_extent((int)n[d], nt, bc_kind[d][0], bc_kind[d][1]));
            if (need > std::max(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin))
                throw Error(EM_ERR_ARG, "adjoint tile too long for LDS (whole pencils on an axis periodic on one "
                                        "side only, or a large tile size)");
        }
    }

    // out = (dL/du at state)^T lam (all active axes, limiter lambda of stage_dt);
    // combine != 0: the RK adjoint update of that stage fused into the last
    // axis (AdjArgs::combine; out is then scratch between the axes)
    void adj_apply(const double* state, const double* lam, double* out, double stage_dt, int combine = 0,
                   const double* cA = nullptr, const double* cB = nullptr, double* cZ = nullptr,
                   double* cBout = nullptr) {
        int last = -1;
        for (int d = 0; d < 3; ++d)
            if (active(d)) last = d;
        bool first = true;
        for (int d = 0; d < 3; ++d) {
            if (!active(d)) continue;
            AdjArgs a{};
            a.U = state;
            a.LAM = lam;
            a.MU = out;
            for (int e = 0; e < 3; ++e) {
                a.n[e] = (int)n[e];
                a.stride[e] = stride[e];
            }
            a.ncell = ncell;
            a.eos = eos;
            for (int sd = 0; sd < 2; ++sd) {
                a.bc_kind[sd] = bc_kind[d][sd];
                for (int q = 0; q < 5; ++q) a.bc_state[sd][q] = bc_state[d][sd][q];
            }
            a.inv_dx = 1.0 / dx[d];
            a.lam_lim = positivity ? (double)nactive * stage_dt / dx[d] : 0.0;  // as the forward pass
            a.first = first ? 1 : 0;
            a.ntile = adj_tiles((int)n[d], bc_kind[d][0], bc_kind[d][1], adj_tile_cap);
            a.ne = adj_tile_extent((int)n[d], a.ntile, bc_kind[d][0], bc_kind[d][1]);
            if (d == last && combine != 0) {
                a.combine = combine;
                a.cdt = stage_dt;
                a.cA = cA;
                a.cB = cB;
                a.cZ = cZ;
                a.cBout = cBout;
            }
            HIPCHK(launch_adj_pass(d, a, stream));
            first = false;
        }
    }

    void host_to_soa(const double* u, double* dev) {
        std::vector<double> h(5 * ncell);
        const int m1 = m(1), m2 = m(2);
        for (long long i = 0; i < n[0]; ++i)
            for (long long j = 0; j < n[1]; ++j)
                for (long long k = 0; k < n[2]; ++k) {
                    const long long q = ((i + g(0)) * m1 + (j + g(1))) * m2 + (k + g(2));
                    const long long c = (i * n[1] + j) * n[2] + k;
                    for (int v = 0; v < 5; ++v) h[v * ncell + c] = u[q * 5 + v];
                }
        HIPCHK(hipMemcpyAsync(dev, h.data(), sizeof(double) * 5 * ncell, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    void soa_to_host(const double* dev, double* u) {
        std::vector<double> h(5 * ncell);
        HIPCHK(hipMemcpyAsync(h.data(), dev, sizeof(double) * 5 * ncell, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        std::fill(u, u + 5 * len(), 0.0);
        const int m1 = m(1), m2 = m(2);
        for (long long i = 0; i < n[0]; ++i)
            for (long long j = 0; j < n[1]; ++j)
                for (long long k = 0; k < n[2]; ++k) {
                    const long long q = ((i + g(0)) * m1 + (j + g(1))) * m2 + (k + g(2));
                    const long long c = (i * n[1] + j) * n[2] + k;
                    for (int v = 0; v < 5; ++v) u[q * 5 + v] = h[v * ncell + c];
                }
    }

    void ensure_adjbuf() {
        if (!adjbuf) HIPCHK(hipMalloc(&adjbuf, sizeof(double) * 20 * ncell));
    }

    // One step reversed over its kept states st (u_n, u1, u2): T = L'(u2)^T A,
    // then L'(u1)^T B, then L'(u^n)^T B with the RK updates fused into each
    // stage's last axis; stage 2's new B goes to A's array (free once stage 3
    // has read it), stage 1 leaves the step's input cotangent in Zb, and A
    // and Zb are swapped, so A holds it on return.
    void reverse_step(const double* st, double dt, double*& A, double* B, double*& Zb, double* T) {
        const long long nv = 5 * ncell;
        adj_apply(st + 2 * nv, A, T, dt, 3, A, nullptr, Zb, B);
        adj_apply(st + nv, B, T, dt, 2, nullptr, B, Zb, A);
        adj_apply(st, A, T, dt, 1, nullptr, A, Zb, nullptr);
        std::swap(A, Zb);
    }

    // (dR/du)^T lam at the current state (one residual evaluation)
    void adjoint_rhs(double stage_dt, const double* lam, double* out) {
        adj_check();
        ensure_adjbuf();
        double* A = adjbuf;
        double* T = adjbuf + 15 * ncell;
        host_to_soa(lam, A);
        adj_apply(U[0], A, T, stage_dt);
        soa_to_host(T, out);
    }

    // Forward run of nsteps SSP-RK3 steps with checkpoints every k steps, or
    // with k = 0 every step's u_n and both stage states kept (3 states per
    // step in device memory; the backward then recomputes nothing); dt from
    // dts or, if null, the CFL dt of each step, computed and recorded on the
    // device (graph-replayed steps, no host wait per step).  Device buffers
    // of earlier runs (checkpoints, dt history, kept states, final state)
    // are reused, so a repeated run of the same length allocates nothing.
    void forward(int64_t nsteps, int k, const double* dts, double* dts_out) {
        if (nsteps < 1 || k < 0) throw Error(EM_ERR_ARG, "nsteps must be >= 1 and checkpoint_every >= 0");
        adj_check();
        if (dts)
            for (int64_t s = 0; s < nsteps; ++s)
                if (!(std::isfinite(dts[s]) && dts[s] > 0.0)) throw Error(EM_ERR_COLLAPSE, "time step collapsed");
        const size_t sz = sizeof(double) * 5 * ncell;
        const bool all = k == 0;
        nck_used = 0;
        fwd_all = false;
        if (all) {
            if (hist_steps < nsteps) {
                if (hist) HIPCHK(hipFree(hist));
                hist = nullptr;
                hist_steps = 0;
                if (hipMalloc(&hist, sz * 3 * (size_t)nsteps) != hipSuccess) {
                    (void)hipGetLastError();
                    hist = nullptr;
                    throw Error(EM_ERR_ARG, "keeping every stage state of " + std::to_string(nsteps) + " steps needs " +
                                                std::to_string(sz * 3 * (size_t)nsteps) +
                                                " bytes of device memory; use checkpoints");
                }
                hist_steps = nsteps;
            }
            k = (int)std::min<int64_t>(nsteps, INT32_MAX);
        } else {
            const int64_t nck = (nsteps + k - 1) / k;
            while ((int64_t)ckpt.size() < nck) {
                double* p;
                HIPCHK(hipMalloc(&p, sz));
                ckpt.push_back(p);
            }
        }
        ck_every = k;
        fwd_dts.assign((size_t)nsteps, 0.0);
        if (!final_state) HIPCHK(hipMalloc(&final_state, sz));
        const long long nv = 5 * ncell;
        // kept states: the stages of step s write straight into slots 3s + 1
        // and 3s + 2 and U^{s+1} into slot 3s + 3 (the last step into U[0]),
        // launched eagerly (the buffers change every step)
        auto kept_step = [&](int64_t s) {
            double* h = hist + 3 * s * nv;
            double* next = s + 1 < nsteps ? h + 3 * nv : U[0];
            stage(h, h + nv, 1, false, 0, 0.0, nullptr, h);
            stage(h + nv, h + 2 * nv, 2, false, 0, 0.0, nullptr, h);
            hipLaunchKernelGGL(k_zero_rate, dim3(1), dim3(kSlots), 0, stream, sc);
            stage(h + 2 * nv, next, 3, true, 0, 0.0, nullptr, h);
            hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, stream, sc);
            HIPCHK(hipGetLastError());
        };
        if (all) HIPCHK(hipMemcpyAsync(hist, U[0], sz, hipMemcpyDeviceToDevice, stream));
        if (dts) {
            for (int64_t s = 0; s < nsteps; ++s) {
                fwd_dts[s] = dts[s];
                if (all) {
                    reset_run(0.0);
                    hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, stream, sc, dts[s]);
                    kept_step(s);
                } else {
                    if (s % k == 0) HIPCHK(hipMemcpyAsync(ckpt[s / k], U[0], sz, hipMemcpyDeviceToDevice, stream));
                    step_async(dts[s]);
                }
            }
        } else {
            if (dt_hist_cap < nsteps) {
                if (dt_hist) HIPCHK(hipFree(dt_hist));
                HIPCHK(hipMalloc(&dt_hist, sizeof(double) * nsteps));
                dt_hist_cap = nsteps;
                if (graph[2]) {  // the captured step records into the old buffer
                    HIPCHK(hipGraphExecDestroy(graph[2]));
                    graph[2] = nullptr;
                }
            }
            reset_run(0.0);
            compute_rate();
            if (all) {
                // device_step(2) on the kept buffers
                for (int64_t s = 0; s < nsteps; ++s) {
                    hipLaunchKernelGGL(k_dt_from_rate, dim3(1), dim3(1), 0, stream, sc, cfl, 0);
                    hipLaunchKernelGGL(k_record_dt, dim3(1), dim3(1), 0, stream, sc, dt_hist);
                    kept_step(s);
                }
            } else {
                hipGraphExec_t gx = get_graph(2);
                for (int64_t s = 0; s < nsteps; ++s) {
                    if (s % k == 0) HIPCHK(hipMemcpyAsync(ckpt[s / k], U[0], sz, hipMemcpyDeviceToDevice, stream));
                    HIPCHK(hipGraphLaunch(gx, stream));
                }
            }
            HIPCHK(hipMemcpyAsync(fwd_dts.data(), dt_hist, sizeof(double) * nsteps, hipMemcpyDeviceToHost, stream));
            Scalars h = read_scalars();  // waits for the stream
            if (h.err & 1) throw Error(EM_ERR_COLLAPSE, "time step collapsed");
        }
        HIPCHK(hipMemcpyAsync(final_state, U[0], sz, hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        rate_valid = false;
        nck_used = all ? 1 : (nsteps + k - 1) / k;
        fwd_all = all;
        if (dts_out) std::copy(fwd_dts.begin(), fwd_dts.end(), dts_out);
    }

    // dJ/du_0 from dJ/du_N of the last forward run: per segment, the states
    // are recomputed from its checkpoint (u_n and both RK stage states of
    // every step kept), then the steps are reversed.  After a forward run
    // that kept every state (checkpoint_every = 0) the steps are reversed
    // over those, with nothing recomputed.
    // dev: dj_final and dj_initial are device arrays of the interior cells,
    // variable-major (5 x ncell, the solver's own layout), else host arrays
    // in the CPU layout.
    void backward(const double* dj_final, double* dj_initial, bool dev = false) {
        if (nck_used == 0) throw Error(EM_ERR_STATE, "backward needs a forward run with checkpoints first");
        ensure_adjbuf();
        const long long nv = 5 * ncell;
        const size_t sz = sizeof(double) * nv;
        auto load_in = [&](double* A) {
            if (dev)
                HIPCHK(hipMemcpyAsync(A, dj_final, sz, hipMemcpyDeviceToDevice, stream));
            else
                host_to_soa(dj_final, A);
        };
        auto store_out = [&](const double* A) {
            if (dev)
                HIPCHK(hipMemcpyAsync(dj_initial, A, sz, hipMemcpyDeviceToDevice, stream));
            else
                soa_to_host(A, dj_initial);
        };
        if (fwd_all) {
            std::vector<DevScalars> saved = save_scalars();
            double* A = adjbuf;
            double* B = adjbuf + nv;
            double* Zb = adjbuf + 2 * nv;
            double* T = adjbuf + 3 * nv;
            load_in(A);
            for (int64_t s = (int64_t)fwd_dts.size() - 1; s >= 0; --s)
                reverse_step(hist + 3 * s * nv, fwd_dts[s], A, B, Zb, T);
            store_out(A);
            HIPCHK(hipMemcpyAsync(U[0], final_state, sz, hipMemcpyDeviceToDevice, stream));
            HIPCHK(hipMemcpyAsync(sc, saved.data(), sizeof(DevScalars), hipMemcpyHostToDevice, stream));
            HIPCHK(hipStreamSynchronize(stream));
            rate_valid = false;
            return;
        }
        if (store_steps < ck_every) {
            if (store) HIPCHK(hipFree(store));
            if (store2) HIPCHK(hipFree(store2));
            store = store2 = nullptr;
            HIPCHK(hipMalloc(&store, sz * 3 * ck_every));
            // the second segment buffer only if it fits (else the recompute
            // waits for each sweep)
            if (hipMalloc(&store2, sz * 3 * ck_every) != hipSuccess) {
                (void)hipGetLastError();
                store2 = nullptr;
            }
            store_steps = ck_every;
        }
        if (!rstream) HIPCHK(hipStreamCreateWithFlags(&rstream, hipStreamNonBlocking));
        std::vector<DevScalars> saved = save_scalars();
        double* A = adjbuf;
        double* B = adjbuf + nv;
        double* Zb = adjbuf + 2 * nv;
        double* T = adjbuf + 3 * nv;
        load_in(A);
        reset_run(0.0);
        const int64_t nsteps = (int64_t)fwd_dts.size();
        // Segments are recomputed on rstream into store buffer (seg & 1) and
        // reversed on stream.  With two buffers the recompute of segment
        // seg - 1 runs while segment seg is reversed (it waits only for the
        // sweep of seg + 1, the last reader of its buffer); with one it
        // waits for the sweep of seg.  The sweep never touches sc or R, the
        // recompute nothing else.
        double* sbuf[2] = {store, store2 ? store2 : store};
        struct Events {  // destroyed when backward() returns, after its final synchronisation
            hipEvent_t e[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
            ~Events() {
                for (auto x : e)
                    if (x) (void)hipEventDestroy(x);
            }
        } evs;
        for (auto& x : evs.e) HIPCHK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
        hipEvent_t* ev_rec = evs.e;
        hipEvent_t* ev_done = evs.e + 2;
        hipEvent_t ev_start = evs.e[4];
        HIPCHK(hipEventRecord(ev_start, stream));
        HIPCHK(hipStreamWaitEvent(rstream, ev_start, 0));
        auto recompute = [&](int64_t seg) {
            const int64_t s0 = seg * ck_every, s1 = std::min<int64_t>(s0 + ck_every, nsteps);
            double* sto = sbuf[seg & 1];
            // store slots 3 (s - s0) .. +2: U^s and its two stage states
            HIPCHK(hipMemcpyAsync(sto, ckpt[seg], sz, hipMemcpyDeviceToDevice, rstream));
            for (int64_t s = s0; s < s1; ++s) {
                double* st = sto + 3 * (s - s0) * nv;
                recompute_step(st, st + nv, st + 2 * nv, s + 1 < s1 ? st + 3 * nv : nullptr, fwd_dts[s], rstream);
            }
            HIPCHK(hipEventRecord(ev_rec[seg & 1], rstream));
        };
        const int64_t last = nck_used - 1;
        recompute(last);
        for (int64_t seg = last; seg >= 0; --seg) {
            if (store2 && seg > 0) {
                if (seg + 1 <= last) HIPCHK(hipStreamWaitEvent(rstream, ev_done[(seg + 1) & 1], 0));
                recompute(seg - 1);
            }
            HIPCHK(hipStreamWaitEvent(stream, ev_rec[seg & 1], 0));
            const int64_t s0 = seg * ck_every, s1 = std::min<int64_t>(s0 + ck_every, nsteps);
            const double* sto = sbuf[seg & 1];
            for (int64_t s = s1 - 1; s >= s0; --s) reverse_step(sto + 3 * (s - s0) * nv, fwd_dts[s], A, B, Zb, T);
            HIPCHK(hipEventRecord(ev_done[seg & 1], stream));
            if (!store2 && seg > 0) {
                HIPCHK(hipStreamWaitEvent(rstream, ev_done[seg & 1], 0));
                recompute(seg - 1);
            }
        }
        store_out(A);
        HIPCHK(hipMemcpyAsync(U[0], final_state, sz, hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipMemcpyAsync(sc, saved.data(), sizeof(DevScalars), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        rate_valid = false;
    }
};

// ------------------------------------------------------------------ ensemble

// Small batched kernels over the instances of an ensemble (thread i or
// workgroup i serves instance i).
__global__ void k_ens_set_run(DevScalars* const* scs, int n) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    scs[i]->active = 1;
    scs[i]->err = 0;
    scs[i]->steps = 0;
}
__global__ void k_ens_zero_rate(DevScalars* const* scs) {
    if (threadIdx.x < kSlots) scs[blockIdx.x]->rate[threadIdx.x][0] = 0ull;
}
// dt = cfl / rate per instance, recorded in hist[step * n + i]; the step
// index lives on the device so a captured step can be replayed.
__global__ void k_ens_dt(DevScalars* const* scs, int n, double cfl, double* hist, const long long* step) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    DevScalars* sc = scs[i];
    if (sc->active == 0) return;
    const double dt = cfl / __longlong_as_double((long long)rate_bits(sc));
    if (!(isfinite(dt) && dt > 0.0)) {
        sc->err |= 1;
        sc->active = 0;
        return;
    }
    sc->dt = dt;
    if (hist) hist[*step * n + i] = dt;
}
__global__ void k_ens_set_dt(DevScalars* const* scs, int n, const double* hist, long long step) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < n) scs[i]->dt = hist[step * n + i];
}
__global__ void k_ens_advance(DevScalars* const* scs, int n, long long* step) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i == 0 && step) *step += 1;
    if (i >= n) return;
    DevScalars* sc = scs[i];
    if (sc->active == 0) return;
    sc->t = sc->t + sc->dt;
    sc->steps += 1;
}
struct CopyTask {
    const double* src;
    double* dst;
    long long nv;
};
__global__ void k_ens_copy(const CopyTask* __restrict__ t) {
    const CopyTask c = t[blockIdx.y];
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < c.nv; i += (long long)gridDim.x * 256ll) c.dst[i] = c.src[i];
}

// An ensemble of independent uniform-grid solvers (own mesh, domain, state,
// dt and CFL rate each) advanced together: every axis pass of every
// instance runs in one persistent launch (ens_pass), a step is replayed as
// one HIP graph; the adjoint launches each instance's VJP (adj_pass).
struct Ensemble {
    std::vector<Solver*> in;
    int n = 0;
    int mode, cfg = 6, precise = 1;
    bool batched = true;
    double cfl;
    hipStream_t stream = nullptr;
    DevScalars** d_scs = nullptr;
    PassArgs* d_args[3][3] = {};
    long long* d_prefix[3][3] = {};
    long long total[3][3] = {};
    long long* d_step = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_key = -1;
    std::vector<long long> off;  // instance offsets in concatenated 5 * ncell layouts
    long long tot_nv = 0, max_nv = 0;
    // adjoint
    int ck_every = 0;
    long long fwd_steps = 0;
    std::vector<double*> ckpt;  // kept across forward runs; the last run uses nck_used
    int64_t nck_used = 0;
    double* dt_hist = nullptr;
    long long dt_hist_cap = 0;
    std::vector<double> h_dts;
    double* work = nullptr;   // A, B, Z, T
    double* store = nullptr;  // 3 states per step of a segment
    int store_steps = 0;
    double* fin = nullptr;    // final states
    static constexpr int kSweepStreams = 4;
    hipStream_t sweep[kSweepStreams] = {};  // instance i's VJPs on sweep[i % 4] (default priority)

    Ensemble(int ninst, const int64_t* nn, const double* lo, const double* hi, int ng, const double e[4],
             const int32_t bk[6], const double bs[30], int mode_, int pos, double cfl_, double tol,
             const double* const* u) {
        if (ninst < 1) throw Error(EM_ERR_ARG, "an ensemble needs at least one instance");
        n = ninst;
        mode = mode_;
        cfl = cfl_;
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        try {
            for (int i = 0; i < n; ++i) {
                in.push_back(new Solver(nn + 3 * i, lo + 3 * i, hi + 3 * i, ng, e, bk, bs, mode_, pos, cfl_, tol, 0.0));
                in.back()->upload(u[i]);
            }
        } catch (...) {
            release();
            throw;
        }
        off.resize(n + 1);
        off[0] = 0;
        for (int i = 0; i < n; ++i) {
            const long long nv = 5 * in[i]->ncell;
            off[i + 1] = off[i] + nv;
            max_nv = std::max(max_nv, nv);
        }
        tot_nv = off[n];
        std::vector<DevScalars*> h(n);
        for (int i = 0; i < n; ++i) h[i] = in[i]->sc;
        HIPCHK(hipMalloc(&d_scs, sizeof(DevScalars*) * n));
        HIPCHK(hipMemcpyAsync(d_scs, h.data(), sizeof(DevScalars*) * n, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMalloc(&d_step, sizeof(long long)));
        build_args();
    }

    void release() {
        for (auto* p : in) delete p;
        in.clear();
        auto f = [](void* p) {
            if (p) (void)hipFree(p);
        };
        for (int sidx = 0; sidx < 3; ++sidx)
            for (int d = 0; d < 3; ++d) {
                f(d_args[sidx][d]);
                f(d_prefix[sidx][d]);
                d_args[sidx][d] = nullptr;
                d_prefix[sidx][d] = nullptr;
            }
        f(d_scs);
        f(d_step);
        f(d_copy);
        d_copy = nullptr;
        free_adjoint();
        d_scs = nullptr;
        d_step = nullptr;
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
        for (auto& x : sweep) {
            if (x) (void)hipStreamDestroy(x);
            x = nullptr;
        }
    }
    ~Ensemble() { release(); }

    void free_adjoint() {
        for (auto* p : ckpt) (void)hipFree(p);
        ckpt.clear();
        nck_used = 0;
        for (double** p : {&dt_hist, &work, &store, &fin})
            if (*p) {
                (void)hipFree(*p);
                *p = nullptr;
            }
        dt_hist_cap = 0;
        store_steps = 0;
    }

    // pass arguments of every instance for stage s (1..3) and axis d, as
    // Solver::stage builds them
    void build_args() {
        for (int i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) in[i]->cfg[d] = cfg;
        for (int st = 1; st <= 3; ++st)
            for (int d = 0; d < 3; ++d) {
                std::vector<PassArgs> a;
                std::vector<long long> pre(1, 0);
                for (int i = 0; i < n; ++i) {
                    Solver& S = *in[i];
                    if (!S.active(d)) {
                        pre.push_back(pre.back());
                        a.push_back(PassArgs{});
                        continue;
                    }
                    int first_ax = -1, last = -1;
                    for (int e = 0; e < 3; ++e)
                        if (S.active(e)) {
                            if (first_ax < 0) first_ax = e;
                            last = e;
                        }
                    PassArgs p = S.pass_args(d);
                    const double* Uin = st == 1 ? S.U[0] : st == 2 ? S.U[1] : S.U[2];
                    double* Uout = st == 1 ? S.U[1] : st == 2 ? S.U[2] : S.U[0];
                    p.U = Uin;
                    p.R = S.R;
                    p.U0 = S.U[0];
                    p.Uout = Uout;
                    p.first = d == first_ax ? 1 : 0;
                    p.rk_stage = d == last ? st : 0;
                    p.want_rate = (d == last && st == 3) ? 1 : 0;
                    p.stage_dt_host = 0;
                    p.stage_dt = 0.0;
                    a.push_back(p);
                    pre.push_back(pre.back() + (long long)p.ntile * p.ngroup * p.n[p.slow_axis]);
                }
                total[st - 1][d] = pre.back();
                if (d_args[st - 1][d]) HIPCHK(hipFree(d_args[st - 1][d]));
                if (d_prefix[st - 1][d]) HIPCHK(hipFree(d_prefix[st - 1][d]));
                HIPCHK(hipMalloc(&d_args[st - 1][d], sizeof(PassArgs) * n));
                HIPCHK(hipMalloc(&d_prefix[st - 1][d], sizeof(long long) * (n + 1)));
                HIPCHK(hipMemcpyAsync(d_args[st - 1][d], a.data(), sizeof(PassArgs) * n, hipMemcpyHostToDevice, stream));
                HIPCHK(hipMemcpyAsync(d_prefix[st - 1][d], pre.data(), sizeof(long long) * (n + 1),
                                      hipMemcpyHostToDevice, stream));
                HIPCHK(hipStreamSynchronize(stream));
            }
        graph_key = -1;
    }

    unsigned small_grid() const { return (unsigned)((n + 63) / 64); }

    // Times every tile shape on the stage-1 passes of all instances (a stage
    // writes only the scratch state U1 and the residual) and keeps the
    // fastest; the instances' states and counters are untouched.  Tile
    // shapes give identical results.
    int autotune(int repeats) {
        std::vector<std::vector<DevScalars>> saved;
        for (auto* S : in) saved.push_back(S->save_scalars());
        int maxn = 1;
        for (auto* S : in)
            for (int d = 0; d < 3; ++d)
                if (S->active(d)) maxn = std::max(maxn, (int)S->n[d]);
        double* hist = nullptr;
        HIPCHK(hipMalloc(&hist, sizeof(double) * n));
        start_run();
        hipLaunchKernelGGL(k_ens_dt, dim3(small_grid()), dim3(64), 0, stream, d_scs, n, cfl, hist, d_step);
        HIPCHK(hipGetLastError());
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        double best = 1e300;
        int bestc = cfg;
        for (int c = 0; c < kNumTileCfgs; ++c) {
            if (kTileCfgs[c].T > 2 * maxn && c != 0) continue;
            cfg = c;
            build_args();
            stage_launches(1);  // warm-up
            HIPCHK(hipEventRecord(e0, stream));
            for (int r = 0; r < repeats; ++r) stage_launches(1);
            HIPCHK(hipEventRecord(e1, stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                bestc = c;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIPCHK(hipStreamSynchronize(stream));
        (void)hipFree(hist);
        for (int i = 0; i < n; ++i)
            HIPCHK(hipMemcpyAsync(in[i]->sc, saved[i].data(), sizeof(DevScalars), hipMemcpyHostToDevice, stream));
        cfg = bestc;
        build_args();
        return cfg;
    }

    void stage_launches(int st, hipStream_t strm = nullptr) {
        hipStream_t stream = strm ? strm : this->stream;
        for (int d = 0; d < 3; ++d) {
            if (total[st - 1][d] == 0) continue;
            if (batched) {
                EnsPassArgs e{d_args[st - 1][d], d_prefix[st - 1][d], n, total[st - 1][d]};
                const unsigned grid = (unsigned)std::min<long long>(total[st - 1][d], 4096);
                HIPCHK(precise ? precise::launch_ens_pass(d, mode, cfg, e, grid, stream)
                               : fast::launch_ens_pass(d, mode, cfg, e, grid, stream));
            } else {
                for (int i = 0; i < n; ++i) {
                    Solver& S = *in[i];
                    if (!S.active(d)) continue;
                    PassArgs p = S.pass_args(d);
                    int first_ax = -1, last = -1;
                    for (int e2 = 0; e2 < 3; ++e2)
                        if (S.active(e2)) {
                            if (first_ax < 0) first_ax = e2;
                            last = e2;
                        }
                    p.U = st == 1 ? S.U[0] : st == 2 ? S.U[1] : S.U[2];
                    p.R = S.R;
                    p.U0 = S.U[0];
                    p.Uout = st == 1 ? S.U[1] : st == 2 ? S.U[2] : S.U[0];
                    p.first = d == first_ax ? 1 : 0;
                    p.rk_stage = d == last ? st : 0;
                    p.want_rate = (d == last && st == 3) ? 1 : 0;
                    HIPCHK(precise ? precise::launch_pass(d, mode, cfg, p, stream)
                                   : fast::launch_pass(d, mode, cfg, p, stream));
                }
            }
        }
    }

    // one step: dt (CFL; recorded in dt_hist when `record`) or the given
    // step's dt from the history, 3 stages
    void step_kernels(bool cfl_dt, long long given_step, hipStream_t strm = nullptr, bool record = false) {
        hipStream_t stream = strm ? strm : this->stream;
        if (cfl_dt)
            hipLaunchKernelGGL(k_ens_dt, dim3(small_grid()), dim3(64), 0, stream, d_scs, n, cfl,
                               record ? dt_hist : nullptr, d_step);
        else
            hipLaunchKernelGGL(k_ens_set_dt, dim3(small_grid()), dim3(64), 0, stream, d_scs, n, dt_hist, given_step);
        stage_launches(1, stream);
        stage_launches(2, stream);
        hipLaunchKernelGGL(k_ens_zero_rate, dim3(n), dim3(kSlots), 0, stream, d_scs);
        stage_launches(3, stream);
        hipLaunchKernelGGL(k_ens_advance, dim3(small_grid()), dim3(64), 0, stream, d_scs, n, d_step);
        HIPCHK(hipGetLastError());
    }

    void start_run() {
        hipLaunchKernelGGL(k_ens_set_run, dim3(small_grid()), dim3(64), 0, stream, d_scs, n);
        const long long zero = 0;
        HIPCHK(hipMemcpyAsync(d_step, &zero, sizeof(long long), hipMemcpyHostToDevice, stream));
        // CFL rate of the initial states (each solver on its own stream)
        HIPCHK(hipStreamSynchronize(stream));
        for (auto* S : in) S->compute_rate();
        for (auto* S : in) HIPCHK(hipStreamSynchronize(S->stream));
    }

    void check_err() {
        HIPCHK(hipStreamSynchronize(stream));
        for (int i = 0; i < n; ++i) {
            Solver::Scalars h = in[i]->read_scalars();
            if (h.err & 1) throw Error(EM_ERR_COLLAPSE, "time step collapsed in instance " + std::to_string(i));
        }
    }

    void run_steps(int64_t nsteps) {
        start_run();
        const int key = (batched ? 1 : 0) + 2 * precise + 4 * cfg;
        if (!graph || graph_key != key) {
            if (graph) HIPCHK(hipGraphExecDestroy(graph));
            graph = nullptr;
            hipGraph_t g;
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            step_kernels(true, 0);
            HIPCHK(hipStreamEndCapture(stream, &g));
            const hipError_t er = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(er);
            graph_key = key;
        }
        for (int64_t q = 0; q < nsteps; ++q) HIPCHK(hipGraphLaunch(graph, stream));
        check_err();
    }

    CopyTask* d_copy = nullptr;
    void copy_all(const std::vector<CopyTask>& tasks) {
        if (tasks.empty()) return;
        if (!d_copy) HIPCHK(hipMalloc(&d_copy, sizeof(CopyTask) * (n + 1)));
        HIPCHK(hipMemcpyAsync(d_copy, tasks.data(), sizeof(CopyTask) * tasks.size(), hipMemcpyHostToDevice, stream));
        const unsigned gx = (unsigned)std::min<long long>((max_nv + 255) / 256, 1024);
        hipLaunchKernelGGL(k_ens_copy, dim3(gx, (unsigned)tasks.size()), dim3(256), 0, stream, d_copy);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));  // the task table is rewritten next call
    }
    // the copies of cnt consecutive tasks of a device table, on stream strm
    void copy_table(const CopyTask* d_tab, int cnt, hipStream_t strm) {
        const unsigned gx = (unsigned)std::min<long long>((max_nv + 255) / 256, 1024);
        hipLaunchKernelGGL(k_ens_copy, dim3(gx, (unsigned)cnt), dim3(256), 0, strm, d_tab);
        HIPCHK(hipGetLastError());
    }
    // state buffer b (0: U^n, 1, 2: stage states) of every instance <-> a
    // concatenated buffer
    std::vector<CopyTask> gather_tasks(int b, double* dst) {
        std::vector<CopyTask> t;
        for (int i = 0; i < n; ++i) t.push_back(CopyTask{in[i]->U[b], dst + off[i], off[i + 1] - off[i]});
        return t;
    }
    std::vector<CopyTask> scatter_tasks(const double* src, int b) {
        std::vector<CopyTask> t;
        for (int i = 0; i < n; ++i) t.push_back(CopyTask{src + off[i], in[i]->U[b], off[i + 1] - off[i]});
        return t;
    }

    // forward with checkpoints every k steps; CFL dt per instance recorded.
    // Device buffers of earlier runs are reused (a repeated run of the same
    // length allocates nothing).
    void forward(int64_t nsteps, int k) {
        if (nsteps < 1 || k < 1) throw Error(EM_ERR_ARG, "nsteps and checkpoint_every must be >= 1");
        for (auto* S : in) S->adj_check();
        ck_every = k;
        fwd_steps = nsteps;
        if (dt_hist_cap < nsteps * n) {
            if (dt_hist) HIPCHK(hipFree(dt_hist));
            dt_hist = nullptr;
                HIPCHK(hipMalloc(&dt_hist, sizeof(double) * nsteps * n));
            dt_hist_cap = nsteps * n;
        }
        const int64_t nck = (nsteps + k - 1) / k;
        while ((int64_t)ckpt.size() < nck) {
            double* p;
            HIPCHK(hipMalloc(&p, sizeof(double) * tot_nv));
            ckpt.push_back(p);
        }
        nck_used = nck;
        start_run();
        for (int64_t s0 = 0; s0 < nsteps; s0 += k) {
            copy_all(gather_tasks(0, ckpt[s0 / k]));
            const int64_t s1 = std::min<int64_t>(s0 + k, nsteps);
            for (int64_t q = s0; q < s1; ++q) step_kernels(true, 0, nullptr, true);
        }
        check_err();
        if (!fin) HIPCHK(hipMalloc(&fin, sizeof(double) * tot_nv));
        copy_all(gather_tasks(0, fin));
        h_dts.resize((size_t)nsteps * n);
        HIPCHK(hipMemcpyAsync(h_dts.data(), dt_hist, sizeof(double) * nsteps * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    // dJ/du_0 of every instance from dJ/du_N (host arrays in each
    // instance's CPU layout)
    void backward(const double* const* dj_final, double* const* dj_initial) {
        if (nck_used == 0) throw Error(EM_ERR_STATE, "backward needs a forward run first");
        if (!work) HIPCHK(hipMalloc(&work, sizeof(double) * 4 * tot_nv));
        if (store_steps < ck_every) {
            if (store) HIPCHK(hipFree(store));
            store = nullptr;
            HIPCHK(hipMalloc(&store, sizeof(double) * 3 * tot_nv * ck_every));
            store_steps = ck_every;
        }
        std::vector<std::vector<DevScalars>> saved;
        for (auto* S : in) saved.push_back(S->save_scalars());
        // work arrays: A (the cotangent of the state after the step being
        // reversed) and Z swap roles every step; B and T keep theirs
        double* W[4] = {work, work + tot_nv, work + 2 * tot_nv, work + 3 * tot_nv};
        for (int i = 0; i < n; ++i) in[i]->host_to_soa(dj_final[i], W[0] + off[i]);
        const int64_t nsteps = fwd_steps;
        const int64_t nseg = nck_used;
        const int k = ck_every;
        // copy tables: per step slot j of a segment the gathers U^n -> slot
        // 3j, U1 -> 3j + 1, U2 -> 3j + 2 ([j][3][n]), then per segment the
        // scatter of its checkpoint into U^n ([seg][n])
        std::vector<CopyTask> cp((size_t)(k * 3 + nseg) * n);
        for (int j = 0; j < k; ++j)
            for (int g = 0; g < 3; ++g)
                for (int i = 0; i < n; ++i)
                    cp[((size_t)j * 3 + g) * n + i] =
                        CopyTask{in[i]->U[g], store + (3 * (size_t)j + g) * tot_nv + off[i], off[i + 1] - off[i]};
        const size_t scat0 = (size_t)k * 3 * n;
        for (int64_t sg = 0; sg < nseg; ++sg)
            for (int i = 0; i < n; ++i)
                cp[scat0 + (size_t)sg * n + i] = CopyTask{ckpt[sg] + off[i], in[i]->U[0], off[i + 1] - off[i]};
        struct Free {  // device table and events, released when backward() returns
            void* p = nullptr;
            std::vector<hipEvent_t> e;
            ~Free() {
                if (p) (void)hipFree(p);
                for (auto x : e)
                    if (x) (void)hipEventDestroy(x);
            }
        } fr;
        for (auto& x : sweep)
            if (!x) HIPCHK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
        // ev[0]: the recompute of a segment is done (the sweeps wait for it);
        // ev[1 + j]: sweep stream j is done with the segment (the next
        // recompute, which rewrites the store and the states, waits for all)
        hipEvent_t ev[1 + kSweepStreams];
        for (auto& x : ev) {
            HIPCHK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
            fr.e.push_back(x);
        }
        CopyTask* d_cp = nullptr;
        HIPCHK(hipMalloc(&d_cp, sizeof(CopyTask) * cp.size()));
        fr.p = d_cp;
        HIPCHK(hipMemcpyAsync(d_cp, cp.data(), sizeof(CopyTask) * cp.size(), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        // VJP arguments of one step ([stage 0..2][axis][instance]), with the
        // RK updates fused into each stage's last axis as in
        // Solver::backward (stage 2's new B into A's array; stage 1 leaves
        // the step's input cotangent in Z, which becomes A); each instance's
        // VJP is its own adj_pass launch, as its single-grid backward runs it,
        // on sweep stream i % 4: instances touch disjoint slices of the work
        // arrays, so different instances' launches may overlap
        int ia = 0, iz = 2;
        for (int64_t sg = nseg - 1; sg >= 0; --sg) {
            const int64_t s0 = sg * k, s1 = std::min<int64_t>(s0 + k, nsteps);
            // recompute: store slots 3 (q - s0) .. +2 hold U^q and its two
            // stage states
            if (sg < nseg - 1)
                for (int j = 0; j < kSweepStreams; ++j) HIPCHK(hipStreamWaitEvent(stream, ev[1 + j], 0));
            copy_table(d_cp + scat0 + (size_t)sg * n, n, stream);
            for (int64_t q = s0; q < s1; ++q) {
                const size_t j = (size_t)(q - s0);
                copy_table(d_cp + j * 3 * n, n, stream);
                step_kernels(false, q);
                copy_table(d_cp + (j * 3 + 1) * n, 2 * n, stream);
            }
            HIPCHK(hipEventRecord(ev[0], stream));
            for (int j = 0; j < kSweepStreams; ++j) HIPCHK(hipStreamWaitEvent(sweep[j], ev[0], 0));
            for (int64_t q = s1 - 1; q >= s0; --q) {
                const double* st = store + 3 * (q - s0) * tot_nv;
                double* A = W[ia];
                double* B = W[1];
                double* Zb = W[iz];
                double* T = W[3];
                for (int stg = 0; stg < 3; ++stg)
                    for (int d = 0; d < 3; ++d)
                        for (int i = 0; i < n; ++i) {
                            const Solver& S = *in[i];
                            if (!S.active(d)) continue;
                            AdjArgs a{};
                            const double dt = h_dts[(size_t)q * n + i];
                            // stage 3 linearises at u2, stage 2 at u1, stage 1 at u^n
                            a.U = st + (2 - stg) * tot_nv + off[i];
                            a.LAM = (stg == 1 ? B : A) + off[i];
                            a.MU = T + off[i];
                            for (int e = 0; e < 3; ++e) {
                                a.n[e] = (int)S.n[e];
                                a.stride[e] = S.stride[e];
                            }
                            a.ncell = S.ncell;
                            a.eos = S.eos;
                            for (int sd = 0; sd < 2; ++sd) {
                                a.bc_kind[sd] = S.bc_kind[d][sd];
                                for (int v = 0; v < 5; ++v) a.bc_state[sd][v] = S.bc_state[d][sd][v];
                            }
                            a.inv_dx = 1.0 / S.dx[d];
                            a.lam_lim = S.positivity ? (double)S.nactive * dt / S.dx[d] : 0.0;
                            int first_ax = -1, last_ax = -1;
                            for (int e = 0; e < 3; ++e)
                                if (S.active(e)) {
                                    if (first_ax < 0) first_ax = e;
                                    last_ax = e;
                                }
                            a.first = d == first_ax ? 1 : 0;
                            a.ntile = adj_tiles((int)S.n[d], S.bc_kind[d][0], S.bc_kind[d][1], S.adj_tile_cap);
                            a.ne = adj_tile_extent((int)S.n[d], a.ntile, S.bc_kind[d][0], S.bc_kind[d][1]);
                            if (d == last_ax) {
                                a.combine = 3 - stg;
                                a.cdt = dt;
                                a.cA = A + off[i];
                                a.cB = (stg == 1 ? B : A) + off[i];
                                a.cZ = Zb + off[i];
                                a.cBout = (stg == 0 ? B : A) + off[i];
                            }
                            HIPCHK(launch_adj_pass(d, a, sweep[i % kSweepStreams]));
                        }
                std::swap(ia, iz);
            }
            for (int j = 0; j < kSweepStreams; ++j) HIPCHK(hipEventRecord(ev[1 + j], sweep[j]));
        }
        for (int j = 0; j < kSweepStreams; ++j) HIPCHK(hipStreamWaitEvent(stream, ev[1 + j], 0));
        HIPCHK(hipStreamSynchronize(stream));  // soa_to_host reads on each instance's own stream
        for (int i = 0; i < n; ++i) in[i]->soa_to_host(W[ia] + off[i], dj_initial[i]);
        copy_all(scatter_tasks(fin, 0));
        for (int i = 0; i < n; ++i) {
            HIPCHK(hipMemcpyAsync(in[i]->sc, saved[i].data(), sizeof(DevScalars), hipMemcpyHostToDevice, stream));
        }
        HIPCHK(hipStreamSynchronize(stream));
    }
};

// ------------------------------------------------------------------ cases

static uint64_t splitmix64(uint64_t& s) {  // Steele, Lea & Flood 2014
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double unit(uint64_t& s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

static void prim_to_cons(const double w[5], double gam, double out[5]) {
    const double rho = w[0];
    const double ke = 0.5 * rho * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
    out[0] = rho;
    out[1] = rho * w[1];
    out[2] = rho * w[2];
    out[3] = rho * w[3];
    out[4] = rho * (w[4] * (1.0 / rho) / (gam - 1.0)) + ke;  // ideal gas rho_e, eos.rs:43-53
}

static void case_init(int id, const int64_t n[3], const double lo[3], const double hi[3], int ng, const double* prm,
                      double* u) {
    auto act = [&](int d) { return n[d] > 1; };
    const int64_t m[3] = {act(0) ? n[0] + 2 * ng : 1, act(1) ? n[1] + 2 * ng : 1, act(2) ? n[2] + 2 * ng : 1};
    const int64_t g[3] = {act(0) ? ng : 0, act(1) ? ng : 0, act(2) ? ng : 0};
    double dx[3];
    for (int d = 0; d < 3; ++d) dx[d] = (hi[d] - lo[d]) / (double)n[d];
    std::memset(u, 0, sizeof(double) * 5 * m[0] * m[1] * m[2]);
    const double gam = 1.4;
    // Fourier modes of case 1
    int km[8][3];
    double ph[8];
    if (id == 1) {
        uint64_t s = (uint64_t)(prm ? prm[1] : 0x5EED);
        for (int q = 0; q < 8; ++q) {
            for (;;) {
                int k[3];
                for (int d = 0; d < 3; ++d) k[d] = (int)(splitmix64(s) % 9) - 4;
                const int k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                if (k2 > 0 && k2 <= 16) {
                    for (int d = 0; d < 3; ++d) km[q][d] = k[d];
                    break;
                }
            }
            ph[q] = 6.283185307179586 * unit(s);
        }
    }
    for (int64_t i = 0; i < n[0]; ++i)
        for (int64_t j = 0; j < n[1]; ++j)
            for (int64_t k = 0; k < n[2]; ++k) {
                const double x[3] = {lo[0] + ((double)i + 0.5) * dx[0], lo[1] + ((double)j + 0.5) * dx[1],
                                     lo[2] + ((double)k + 0.5) * dx[2]};
                double w[5];
                if (id == 0) {
                    const double xs = prm ? prm[0] : 0.8, ys = prm ? prm[1] : 0.8;
                    const bool e = x[0] > xs, nn = x[1] > ys;
                    if (e && nn) {
                        w[0] = 1.5; w[1] = 0.0; w[2] = 0.0; w[4] = 1.5;
                    } else if (!e && nn) {
                        w[0] = 0.5323; w[1] = 1.206; w[2] = 0.0; w[4] = 0.3;
                    } else if (!e && !nn) {
                        w[0] = 0.138; w[1] = 1.206; w[2] = 1.206; w[4] = 0.029;
                    } else {
                        w[0] = 0.5323; w[1] = 0.0; w[2] = 1.206; w[4] = 0.3;
                    }
                    w[3] = 0.0;
                } else if (id == 1) {
                    const double U0 = prm ? prm[0] : 0.3;
                    double s = 0.0;
                    for (int q = 0; q < 8; ++q)
                        s += std::cos(km[q][0] * x[0] + km[q][1] * x[1] + km[q][2] * x[2] + ph[q]);
                    w[0] = 1.0 + 0.1 * s;
                    w[1] = U0 * std::sin(x[0]) * std::cos(x[1]) * std::cos(x[2]);
                    w[2] = -U0 * std::cos(x[0]) * std::sin(x[1]) * std::cos(x[2]);
                    w[3] = 0.0;
                    w[4] = 1.0 / gam;
                } else if (id == 2) {
                    const double TAU = 6.283185307179586;
                    w[0] = 1.0 + 0.1 * std::sin(TAU * x[0]) * std::cos(TAU * x[1]);
                    w[1] = 0.3; w[2] = 0.2; w[3] = 0.1; w[4] = 1.0;
                } else {
                    throw Error(EM_ERR_ARG, "unknown case id");
                }
                const int64_t q = ((i + g[0]) * m[1] + (j + g[1])) * m[2] + (k + g[2]);
                prim_to_cons(w, gam, u + 5 * q);
            }
}

// ------------------------------------------- allocation-free residual

// Scratch of em_rhs_noalloc: the device scalars its passes accumulate into
// (64 counter slots on separate 128-byte lines, the unused CFL-rate slots and
// the run header).  One copy per GPU, kept until the process ends; every call
// on a GPU reuses that GPU's copy, so calls that overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = EM_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static_assert(sizeof(DevScalars) <= kNoAllocScratchBytes, "scratch smaller than the device scalars");

constexpr int kMaxDevices = 64;
static DevScalars* g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per GPU
static std::mutex g_scratch_mu;

// Device address of the current GPU's scratch, looked up once per GPU with
// that GPU current.  Host code never passes `scratch` itself to a kernel: on
// the host that name is only a stand-in for the symbol.
static DevScalars* noalloc_scratch() {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(EM_ERR_ARG, "device ordinal out of range (at most 64 GPUs)");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(EM_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_ptr[dev] = (DevScalars*)p;
    }
    return g_scratch_ptr[dev];
}

// Puts the scratch in the state a residual pass expects: active run, counter
// slots zero (a kernel, not hipMemsetAsync, so callers may capture the call).
__global__ void k_noalloc_reset(DevScalars* sc) {
    const int t = threadIdx.x;
    if (t == 0) {
        sc->dt = 0.0;
        sc->t = 0.0;
        sc->t_end = 0.0;
        sc->active = 1;
        sc->err = 0;
        sc->steps = 0;
    }
    if (t < kSlots) {
        sc->rate[t][0] = 0ull;
        for (int q = 0; q < 3; ++q) sc->counters[q][t][0] = 0ull;
    }
}

// stats = [faces, shock-path faces, fallbacks, limited faces], slots reduced.
__global__ void k_noalloc_stats(const DevScalars* sc, unsigned long long faces, unsigned long long* out) {
    const int t = threadIdx.x;
    if (t == 0) out[0] = faces;
    if (t < 3) {
        unsigned long long c = 0;
        for (int i = 0; i < kSlots; ++i) c += sc->counters[t][i][0];
        out[1 + t] = c;
    }
}

static void rhs_noalloc(const Grid& G, int precise, double stage_dt, const double* u, double* rhs, uint64_t* stats,
                        hipStream_t st) {
    if (sizeof(DevScalars) > kNoAllocScratchBytes) throw Error(EM_ERR_ARG, "scratch too small for the device scalars");
    for (int d = 0; d < 3; ++d) {
        if (!G.active(d)) continue;
        const PassArgs a = G.pass_args(d, nullptr);
        const long long nblk = (long long)a.ntile * a.ngroup * a.n[a.slow_axis];
        if (nblk > 0x7fffffffll) throw Error(EM_ERR_ARG, "mesh too large for one axis-pass launch");
    }
    DevScalars* sc = noalloc_scratch();
    hipLaunchKernelGGL(k_noalloc_reset, dim3(1), dim3(kSlots), 0, st, sc);
    HIPCHK(hipGetLastError());
    bool first = true;
    for (int d = 0; d < 3; ++d) {
        if (!G.active(d)) continue;
        PassArgs a = G.pass_args(d, sc);
        a.U = u;
        a.R = rhs;
        a.U0 = u;
        a.Uout = nullptr;
        a.first = first ? 1 : 0;
        a.rk_stage = 0;
        a.want_rate = 0;
        a.stage_dt_host = 1;
        a.stage_dt = stage_dt;
        HIPCHK(precise ? precise::launch_pass(d, G.mode, G.cfg[d], a, st) : fast::launch_pass(d, G.mode, G.cfg[d], a, st));
        first = false;
    }
    if (stats) {
        hipLaunchKernelGGL(k_noalloc_stats, dim3(1), dim3(64), 0, st, sc, (unsigned long long)G.faces_per_stage(),
                           (unsigned long long*)stats);
        HIPCHK(hipGetLastError());
    }
}

}  // namespace em

// ------------------------------------------------------------------ C ABI

struct em_solver {
    em::Solver* s;                  // FP64 solver (null for an FP32 handle)
    em::SolverCore<float>* s32;     // FP32 solver (null for an FP64 handle)
};

struct em_ensemble {
    em::Ensemble* e;
};

template <class F>
static int guarded(F&& f) {
    try {
        f();
        em::g_last_error.clear();
        return EM_OK;
    } catch (const em::Error& e) {
        em::g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        em::g_last_error = e.what();
        return EM_ERR_STATE;
    }
}

#define NEED(h)                                                              \
    if (!(h) || (!(h)->s && !(h)->s32)) {                                    \
        em::g_last_error = "null solver handle";                             \
        return EM_ERR_ARG;                                                   \
    }

// entry points of the FP64 solver only (adjoint, ensemble)
#define NEED64(h)                                                            \
    NEED(h)                                                                  \
    if (!(h)->s) {                                                           \
        em::g_last_error = "this entry point takes an FP64 solver (em_solver_create)"; \
        return EM_ERR_ARG;                                                   \
    }

// f(core) on the handle's solver, FP64 or FP32, with the error convention
template <class F>
static int on_core(em_solver* h, F&& f) {
    if (h->s32) return guarded([&] { f(*h->s32); });
    return guarded([&] { f(static_cast<em::SolverCore<double>&>(*h->s)); });
}

extern "C" {

const char* em_last_error(void) { return em::g_last_error.c_str(); }

int em_solver_create(em_solver** out, const int64_t n[3], const double lo[3], const double hi[3], int32_t ng,
                     const double eos[4], const int32_t bc_kind[6], const double bc_state[30], int32_t mode,
                     int32_t positivity, double cfl, double tol, double t0, const double* u) {
    if (!out || !n || !lo || !hi || !eos || !bc_kind || !u) {
        em::g_last_error = "null argument";
        return EM_ERR_ARG;
    }
    return guarded([&] {
        auto* s = new em::Solver(n, lo, hi, ng, eos, bc_kind, bc_state, mode, positivity, cfl, tol, t0);
        try {
            s->upload(u);
        } catch (...) {
            delete s;
            throw;
        }
        *out = new em_solver{s, nullptr};
    });
}

int em_solver_create_f32(em_solver** out, const int64_t n[3], const double lo[3], const double hi[3], int32_t ng,
                         const double eos[4], const int32_t bc_kind[6], const double bc_state[30], int32_t mode,
                         int32_t positivity, double cfl, double tol, double t0, const double* u) {
    if (!out || !n || !lo || !hi || !eos || !bc_kind || !u) {
        em::g_last_error = "null argument";
        return EM_ERR_ARG;
    }
    return guarded([&] {
        auto* s = new em::SolverCore<float>(n, lo, hi, ng, eos, bc_kind, bc_state, mode, positivity, cfl, tol, t0);
        s->precise = 0;  // contracted, as the reference's kernels are compiled
        try {
            s->upload(u);
        } catch (...) {
            delete s;
            throw;
        }
        *out = new em_solver{nullptr, s};
    });
}

int em_solver_destroy(em_solver* h) {
    if (!h) return EM_OK;
    delete h->s;
    delete h->s32;
    delete h;
    return EM_OK;
}

int em_solver_set_variant(em_solver* h, int32_t precise) {
    NEED(h);
    return on_core(h, [&](auto& c) { c.precise = precise ? 1 : 0; });
}

int em_solver_set_tiles(em_solver* h, const int32_t cfg[3]) {
    NEED(h);
    for (int d = 0; d < 3; ++d)
        if (cfg[d] < 0 || cfg[d] >= em::kNumTileCfgs) {
            em::g_last_error = "tile configuration out of range";
            return EM_ERR_ARG;
        }
    return on_core(h, [&](auto& c) {
        for (int d = 0; d < 3; ++d) c.cfg[d] = cfg[d];
    });
}

int em_solver_autotune(em_solver* h, int32_t repeats, int32_t cfg_out[3]) {
    NEED(h);
    return on_core(h, [&](auto& c) { c.autotune(repeats > 0 ? repeats : 3, cfg_out); });
}

int em_solver_upload(em_solver* h, const double* u) {
    NEED(h);
    return on_core(h, [&](auto& c) { c.upload(u); });
}

int em_solver_download(em_solver* h, double* u) {
    NEED(h);
    return on_core(h, [&](auto& c) { c.download(u); });
}

int em_solver_rhs(em_solver* h, double stage_dt, double* rhs, uint64_t stats[4]) {
    NEED(h);
    return on_core(h, [&](auto& c) { c.rhs(stage_dt, rhs, stats); });
}

int em_solver_step(em_solver* h, double dt) {
    NEED(h);
    return on_core(h, [&](auto& c) { c.step(dt); });
}

int em_solver_dt_cfl(em_solver* h, double* dt) {
    NEED(h);
    return on_core(h, [&](auto& c) { *dt = c.dt_cfl(); });
}

int em_solver_run_to(em_solver* h, double t_end, int64_t* steps) {
    NEED(h);
    return on_core(h, [&](auto& c) {
        int64_t k = c.run_to(t_end);
        if (steps) *steps = k;
    });
}

int em_solver_run_steps(em_solver* h, int64_t nsteps, double pass_ms[3]) {
    NEED(h);
    return on_core(h, [&](auto& c) { c.run_steps(nsteps, pass_ms); });
}

int em_solver_counters(em_solver* h, uint64_t out[3]) {
    NEED(h);
    return on_core(h, [&](auto& c) {
        auto d = c.read_scalars();
        for (int q = 0; q < 3; ++q) out[q] = d.counters[q];
    });
}

int em_solver_time(em_solver* h, double* t) {
    NEED(h);
    return on_core(h, [&](auto& c) { *t = c.read_scalars().t; });
}

int em_solver_synchronize(em_solver* h) {
    NEED(h);
    return on_core(h, [&](auto& c) { em::hip_check(hipStreamSynchronize(c.stream), "hipStreamSynchronize"); });
}

int em_solver_adjoint_rhs(em_solver* h, double stage_dt, const double* lam, double* out) {
    NEED64(h);
    if (!lam || !out) {
        em::g_last_error = "null argument";
        return EM_ERR_ARG;
    }
    return guarded([&] { h->s->adjoint_rhs(stage_dt, lam, out); });
}

int em_solver_set_adjoint_tile(em_solver* h, int32_t cells) {
    NEED64(h);
    if (cells < 10) {
        em::g_last_error = "adjoint tile size must be at least 10 cells";
        return EM_ERR_ARG;
    }
    return guarded([&] {
        const int keep = h->s->adj_tile_cap;
        h->s->adj_tile_cap = cells;
        try {
            h->s->adj_check();
        } catch (...) {
            h->s->adj_tile_cap = keep;
            throw;
        }
    });
}

int em_solver_forward(em_solver* h, int64_t nsteps, int32_t checkpoint_every, const double* dts, double* dts_out) {
    NEED64(h);
    return guarded([&] { h->s->forward(nsteps, checkpoint_every, dts, dts_out); });
}

int em_solver_backward(em_solver* h, const double* dj_final, double* dj_initial) {
    NEED64(h);
    if (!dj_final || !dj_initial) {
        em::g_last_error = "null argument";
        return EM_ERR_ARG;
    }
    return guarded([&] { h->s->backward(dj_final, dj_initial); });
}

int em_solver_backward_device(em_solver* h, const double* dj_final, double* dj_initial) {
    NEED64(h);
    if (!dj_final || !dj_initial) {
        em::g_last_error = "null argument";
        return EM_ERR_ARG;
    }
    return guarded([&] { h->s->backward(dj_final, dj_initial, true); });
}

// ---------------------------------------------------------------- ensemble

int em_ensemble_create(em_ensemble** out, int32_t ninst, const int64_t* n, const double* lo, const double* hi,
                       int32_t ng, const double eos[4], const int32_t bc_kind[6], const double bc_state[30],
                       int32_t mode, int32_t positivity, double cfl, double tol, const double* const* u) {
    if (!out || !n || !lo || !hi || !eos || !bc_kind || !u) {
        em::g_last_error = "null argument";
        return EM_ERR_ARG;
    }
    return guarded([&] {
        auto* e = new em::Ensemble(ninst, n, lo, hi, ng, eos, bc_kind, bc_state, mode, positivity, cfl, tol, u);
        *out = new em_ensemble{e};
    });
}

int em_ensemble_destroy(em_ensemble* h) {
    if (!h) return EM_OK;
    delete h->e;
    delete h;
    return EM_OK;
}

#define ENEED(h)                                                             \
    if (!(h) || !(h)->e) {                                                   \
        em::g_last_error = "null ensemble handle";                           \
        return EM_ERR_ARG;                                                   \
    }

int em_ensemble_set_batched(em_ensemble* h, int32_t batched) {
    ENEED(h);
    h->e->batched = batched != 0;
    return EM_OK;
}

int em_ensemble_autotune(em_ensemble* h, int32_t repeats, int32_t* cfg_out) {
    ENEED(h);
    return guarded([&] {
        const int c = h->e->autotune(repeats > 0 ? repeats : 3);
        if (cfg_out) *cfg_out = c;
    });
}

int em_ensemble_set_tiles(em_ensemble* h, int32_t cfg) {
    ENEED(h);
    if (cfg < 0 || cfg >= em::kNumTileCfgs) {
        em::g_last_error = "tile configuration out of range";
        return EM_ERR_ARG;
    }
    return guarded([&] {
        h->e->cfg = cfg;
        h->e->build_args();
    });
}

int em_ensemble_run_steps(em_ensemble* h, int64_t nsteps) {
    ENEED(h);
    return guarded([&] { h->e->run_steps(nsteps); });
}

int em_ensemble_upload(em_ensemble* h, int32_t i, const double* u) {
    ENEED(h);
    if (i < 0 || i >= h->e->n || !u) {
        em::g_last_error = "instance index out of range or null buffer";
        return EM_ERR_ARG;
    }
    return guarded([&] { h->e->in[i]->upload(u); });
}

int em_ensemble_download(em_ensemble* h, int32_t i, double* u) {
    ENEED(h);
    if (i < 0 || i >= h->e->n || !u) {
        em::g_last_error = "instance index out of range or null buffer";
        return EM_ERR_ARG;
    }
    return guarded([&] { h->e->in[i]->download(u); });
}

int em_ensemble_time(em_ensemble* h, int32_t i, double* t) {
    ENEED(h);
    if (i < 0 || i >= h->e->n || !t) {
        em::g_last_error = "instance index out of range or null pointer";
        return EM_ERR_ARG;
    }
    return guarded([&] { *t = h->e->in[i]->read_scalars().t; });
}

int em_ensemble_forward(em_ensemble* h, int64_t nsteps, int32_t checkpoint_every) {
    ENEED(h);
    return guarded([&] { h->e->forward(nsteps, checkpoint_every); });
}

int em_ensemble_dts(em_ensemble* h, int32_t i, double* dts) {
    ENEED(h);
    if (i < 0 || i >= h->e->n || !dts) {
        em::g_last_error = "instance index out of range or null buffer";
        return EM_ERR_ARG;
    }
    return guarded([&] {
        if (h->e->h_dts.empty()) throw em::Error(EM_ERR_STATE, "no forward run");
        for (int64_t q = 0; q < h->e->fwd_steps; ++q) dts[q] = h->e->h_dts[(size_t)q * h->e->n + i];
    });
}

int em_ensemble_backward(em_ensemble* h, const double* const* dj_final, double* const* dj_initial) {
    ENEED(h);
    if (!dj_final || !dj_initial) {
        em::g_last_error = "null argument";
        return EM_ERR_ARG;
    }
    return guarded([&] { h->e->backward(dj_final, dj_initial); });
}

int em_case_init(int32_t case_id, const int64_t n[3], const double lo[3], const double hi[3], int32_t ng,
                 const double* params, double* u) {
    if (!n || !lo || !hi || !u) {
        em::g_last_error = "null argument";
        return EM_ERR_ARG;
    }
    return guarded([&] { em::case_init(case_id, n, lo, hi, ng, params, u); });
}

int em_rhs_noalloc(const int64_t n[3], const double lo[3], const double hi[3], const double eos[4],
                   const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity, double tol,
                   int32_t precise, double stage_dt, const double* u, double* rhs, uint64_t* stats, void* stream) {
    if (!n || !lo || !hi || !eos || !bc_kind || !u || !rhs) {
        em::g_last_error = "null argument";
        return EM_ERR_ARG;
    }
    if ((const void*)u == (const void*)rhs) {
        em::g_last_error = "u and rhs must not alias";
        return EM_ERR_ARG;
    }
    return guarded([&] {
        const em::Grid G(n, lo, hi, 6, eos, bc_kind, bc_state, mode, positivity, 0.5, tol);
        em::rhs_noalloc(G, precise, stage_dt, u, rhs, stats, (hipStream_t)stream);
    });
}

}  // extern "C"
