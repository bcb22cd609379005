// Fused per-axis pass of the compressible RHS on gfx950.
//
// One workgroup owns a tile of T cells along the pass axis D times P
// neighbouring pencils.  Phases (LDS-staged, one barrier between each):
//   1. load: coalesced SoA loads of the 5 conserved variables of the T+10
//      cells (5-cell halo on each side; ghost cells are produced from the
//      boundary condition of the pencil itself), conversion to primitives
//      into LDS (plus the hybrid sensor ln rho, ln(p + p_inf));
//   2. hybrid only: blend weight theta of every face;
//   3. WCNS midpoints: each of the T+5 midpoints of a pencil is computed
//      once (characteristic WENO5-Z interpolation + HLLC) into LDS;
//   4. faces: S6 midpoint-to-node combination, KEEP-PE central flux and the
//      theta blend, Hu-Adams-Shu positivity limiter;
//   5. cells: divergence -(F_{i+1/2} - F_{i-1/2})/dx accumulated in the
//      reference's axis order; on the last active axis fused with the
//      SSP-RK3 combination and the CFL max-rate of the updated state.
// Reference: fluid/crates/compressible/src/scheme.rs:132-281 (line_fluxes),
// 286-333 (midpoint_flux), 395-464 (positivity_limit);
// fluid/crates/compressible/src/solver.rs:97-137 (rk_combine), 351-475 (rhs),
// 259-348 (ghosts), 478-523 (dt_cfl).
//
// This file is compiled once per (variant, flux mode): EM_VARIANT=precise
// with -ffp-contract=off (reproduces the reference CPU arithmetic) or
// EM_VARIANT=fast with FMA contraction; EM_MODE = 0 hybrid, 1 weno,
// 2 central.  Kernels ask for 4 waves per SIMD (<= 128 VGPRs).
#include <hip/hip_runtime.h>

#include "em_pass.h"
#include "em_types.h"

#ifndef EM_VARIANT
#error "EM_VARIANT must name the arithmetic variant"
#endif
#ifndef EM_MODE
#error "EM_MODE must name the flux mode instantiated by this translation unit"
#endif
#if EM_MODE == 0
#define EM_LAUNCHER launch_pass_hybrid
#define EM_LAUNCHER32 launch_pass32_hybrid
#define EM_AMR_LAUNCHER launch_amr_hybrid
#define EM_AMR_LAUNCHER32 launch_amr32_hybrid
#define EM_AMR_MULTI_LAUNCHER launch_amr_multi_hybrid
#define EM_ENS_LAUNCHER launch_ens_hybrid
#elif EM_MODE == 1
#define EM_LAUNCHER launch_pass_weno
#define EM_LAUNCHER32 launch_pass32_weno
#define EM_AMR_LAUNCHER launch_amr_weno
#define EM_AMR_LAUNCHER32 launch_amr32_weno
#define EM_AMR_MULTI_LAUNCHER launch_amr_multi_weno
#define EM_ENS_LAUNCHER launch_ens_weno
#else
#define EM_LAUNCHER launch_pass_central
#define EM_LAUNCHER32 launch_pass32_central
#define EM_AMR_LAUNCHER launch_amr_central
#define EM_AMR_LAUNCHER32 launch_amr32_central
#define EM_AMR_MULTI_LAUNCHER launch_amr_multi_central
#define EM_ENS_LAUNCHER launch_ens_central
#endif
#ifndef EM_WAVES_PER_EU
#define EM_WAVES_PER_EU 4
#endif

#include "em_faces.h"
#include "em_amr_pass.h"

namespace em {
namespace EM_VARIANT {

// Conserved state of the virtual ghost cell at pencil coordinate g (outside
// [0, n)) of the pencil starting at cell `pen`.
template <int D, class Real>
__device__ __forceinline__ void ghost_state(const PassArgsT<Real>& a, long long pen, int g, Real u[5]) {
    const long long sD = a.stride[D];
    auto ld = [&](int x, Real* o) {
#pragma unroll
        for (int k = 0; k < 5; ++k) o[k] = a.U[k * a.ncell + pen + (long long)x * sD];
    };
    bc_ghost<D>(a.eos, a.n[D], g, a.bc_kind, a.bc_state, ld, u);
}

// LDS of one tile (declared by the kernels, passed to axis_tile)
template <class Real, int MODE, int T, int P, int NT>
struct TileLds {
    static constexpr bool HYB = MODE == MODE_HYBRID;
    Real sw[5][T + 10][P];
    Real sre[T + 10][P];  // rho_e of each cell
    Real sm[5][T + 5][P];
    Real ss[2][HYB ? T + 10 : 1][P];
    Real sth[HYB ? T + 1 : 1][P];
    double sred[NT / 64 + 1];  // per-wave CFL rate (in double for the slot max)
    unsigned int scnt[3];
};

// Tile `bid` of one axis pass described by `a` (every thread of the
// workgroup calls it; it returns uniformly).
template <class Real, int D, int MODE, int T, int P, int NT>
__device__ __forceinline__ void axis_tile(const PassArgsT<Real>& a, int bid, TileLds<Real, MODE, T, P, NT>& L) {
    constexpr int C = T + 10;  // cells held in LDS
    constexpr bool HYB = MODE == MODE_HYBRID;
    auto& sw = L.sw;
    auto& sre = L.sre;
    auto& sm = L.sm;
    auto& ss = L.ss;
    auto& sth = L.sth;
    auto& sred = L.sred;
    auto& scnt = L.scnt;

    if (a.sc->active == 0) return;
    const int tid = threadIdx.x;
    const int tile = bid % a.ntile;
    const int rest = bid / a.ntile;
    const int gF = rest % a.ngroup;
    const int sI = rest / a.ngroup;
    const int nD = a.n[D];
    const int nF = a.n[a.fast_axis];
    const int c0 = tile * T;
    const int Tv = min(T, nD - c0);
    const int p0 = gF * P;
    const int Pv = min(P, nF - p0);
    const long long sD = a.stride[D];
    const long long sF = a.stride[a.fast_axis];
    const long long base = (long long)sI * a.stride[a.slow_axis] + (long long)p0 * sF;
    const long long nc = a.ncell;
    const EosT<Real> e = a.eos;
    if (tid < 3) scnt[tid] = 0;

    // ---- 1. load + primitives
    for (int it = tid; it < C * P; it += NT) {
        int c, p;
        if (a.dfast) {
            c = it % C;
            p = it / C;
        } else {
            p = it % P;
            c = it / P;
        }
        if (p >= Pv || c >= Tv + 10) continue;
        const int g = c0 - 5 + c;
        const long long pen = base + (long long)p * sF;
        Real u[5], w[5];
        if (g >= 0 && g < nD) {
            const long long idx = pen + (long long)g * sD;
#pragma unroll
            for (int k = 0; k < 5; ++k) u[k] = a.U[k * nc + idx];
        } else {
            ghost_state<D>(a, pen, g, u);
        }
        to_prim(e, u, w);
#pragma unroll
        for (int k = 0; k < 5; ++k) sw[k][c][p] = w[k];
        sre[c][p] = eos_rho_e(e, w[0], w[4]);
        if constexpr (HYB) {
            // Euler::sensor, system.rs:409-413
            ss[0][c][p] = log(fmax(w[0], Real(1e-300)));
            ss[1][c][p] = log(fmax(w[4] + e.p_inf, Real(1e-300)));
        }
    }
    __syncthreads();

    unsigned int n_shock = 0, n_fb = 0, n_lim = 0;
    double lam = 0.0;
    if (a.ndim_pos > 0.0) {
        const double dt = a.stage_dt_host ? a.stage_dt : a.sc->dt;
        lam = a.ndim_pos * dt / a.dx_d;  // solver.rs:393
    }
    tile_faces<D, MODE, T, P, NT>(sw, sre, sm, ss, sth, Tv, Pv, c0, nD, e, Real(lam), a.tol, n_shock, n_fb, n_lim);

    // ---- 5. divergence, fused RK stage and CFL rate
    double rate = 0.0;
    const Real dt = Real(a.sc->dt);
    for (int it = tid; it < T * P; it += NT) {
        int c, p;
        if (a.dfast) {
            c = it % T;
            p = it / T;
        } else {
            p = it % P;
            c = it / P;
        }
        if (p >= Pv || c >= Tv) continue;
        const long long idx = base + (long long)p * sF + (long long)(c0 + c) * sD;
        Real r[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const Real prev = a.first ? Real(0.0) : a.R[k * nc + idx];
            r[k] = prev - (sm[k][c + 1][p] - sm[k][c][p]) * a.inv_dx;
        }
        if (a.rk_stage == 0) {
#pragma unroll
            for (int k = 0; k < 5; ++k) a.R[k * nc + idx] = r[k];
        } else {
            Real un_[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const Real u0 = a.U0[k * nc + idx];
                Real o;
                if (a.rk_stage == 1) {
                    o = u0 + dt * r[k];  // the stage input is U^n itself: no second load
                } else {
                    const Real us = a.U[k * nc + idx];
                    if (a.rk_stage == 2)
                        o = Real(0.75) * u0 + Real(0.25) * (us + dt * r[k]);
                    else
                        o = u0 / Real(3.0) + Real(2.0) / Real(3.0) * (us + dt * r[k]);
                }
                un_[k] = o;
                a.Uout[k * nc + idx] = o;
            }
            if (a.want_rate) {
                Real w[5];
                to_prim(e, un_, w);
                const Real c = sound_speed(e, w[0], w[4]);
                Real ac = Real(-0.0);
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    if (a.active[d]) ac += (fabs(w[1 + d]) + c) / a.dx_all[d];
                rate = fmax(rate, (double)ac);
            }
        }
    }

    // ---- counters and rate: wave reduction, one atomic per workgroup
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        n_shock += __shfl_xor(n_shock, off);
        n_fb += __shfl_xor(n_fb, off);
        n_lim += __shfl_xor(n_lim, off);
        rate = fmax(rate, __shfl_xor(rate, off));
    }
    const int lane = tid & 63, wid = tid >> 6;
    if (lane == 0) {
        if (n_shock) atomicAdd(&scnt[0], n_shock);
        if (n_fb) atomicAdd(&scnt[1], n_fb);
        if (n_lim) atomicAdd(&scnt[2], n_lim);
        sred[wid] = rate;
    }
    __syncthreads();
    if (tid == 0) {
        const int slot = bid % kSlots;
        if (scnt[0]) atomicAdd(&a.sc->counters[0][slot][0], (unsigned long long)scnt[0]);
        if (scnt[1]) atomicAdd(&a.sc->counters[1][slot][0], (unsigned long long)scnt[1]);
        if (scnt[2]) atomicAdd(&a.sc->counters[2][slot][0], (unsigned long long)scnt[2]);
        if (a.want_rate) {
            double rm = 0.0;
            for (int w = 0; w < (NT + 63) / 64; ++w) rm = fmax(rm, sred[w]);
            rm = rm > 0.0 ? rm : 0.0;  // IEEE bits of +x order like x (no -0.0)
            atomicMax(&a.sc->rate[slot][0], (unsigned long long)__double_as_longlong(rm));
        }
    }
}

template <int D, int MODE, int T, int P, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(EM_WAVES_PER_EU))) void axis_pass(PassArgs a) {
    __shared__ TileLds<double, MODE, T, P, NT> L;
    axis_tile<double, D, MODE, T, P, NT>(a, xcd_major((int)blockIdx.x, (int)gridDim.x), L);
}

// The FP32 variant of the same pass (the reference's GpuSolver<f32>): state,
// residual and arithmetic in float.
template <int D, int MODE, int T, int P, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(EM_WAVES_PER_EU))) void axis_pass_f32(
    PassArgsT<float> a) {
    __shared__ TileLds<float, MODE, T, P, NT> L;
    axis_tile<float, D, MODE, T, P, NT>(a, xcd_major((int)blockIdx.x, (int)gridDim.x), L);
}

// Ensemble pass: the tiles of every instance in one persistent launch.  Work
// item w belongs to the instance i with prefix[i] <= w < prefix[i+1] and is
// its tile w - prefix[i]; each instance brings its own arrays, extents,
// boundary conditions and device scalars (dt, CFL rate, counters).
template <int D, int MODE, int T, int P, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(EM_WAVES_PER_EU))) void ens_pass(EnsPassArgs e) {
    __shared__ TileLds<double, MODE, T, P, NT> L;
    const XcdStride xs = xcd_stride();
    for (long long t = xs.first;; t += xs.step) {
        const long long w = xs.item(t);
        if (w >= e.total) break;
        int lo = 0, hi = e.ninst;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (e.prefix[mid] <= w)
                lo = mid;
            else
                hi = mid;
        }
        const PassArgs& a = e.args[lo];  // read in place (global memory, cached)
        axis_tile<double, D, MODE, T, P, NT>(a, (int)(w - e.prefix[lo]), L);
        __syncthreads();
    }
}

template <int D, int MODE, int T, int P, int NT>
static hipError_t launch_one(const PassArgs& a, hipStream_t s) {
    const long long nblk = (long long)a.ntile * a.ngroup * a.n[a.slow_axis];
    hipLaunchKernelGGL((axis_pass<D, MODE, T, P, NT>), dim3((unsigned)nblk), dim3(NT), 0, s, a);
    return hipGetLastError();
}
template <int D, int MODE, int T, int P, int NT>
static hipError_t launch_one(const PassArgsT<float>& a, hipStream_t s) {
    const long long nblk = (long long)a.ntile * a.ngroup * a.n[a.slow_axis];
    hipLaunchKernelGGL((axis_pass_f32<D, MODE, T, P, NT>), dim3((unsigned)nblk), dim3(NT), 0, s, a);
    return hipGetLastError();
}

template <int D, int MODE, int T, int P, int NT>
static hipError_t launch_ens_one(const EnsPassArgs& e, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL((ens_pass<D, MODE, T, P, NT>), dim3(grid), dim3(NT), 0, s, e);
    return hipGetLastError();
}

// Tile table: must list the shapes of kTileCfgs (em_pass.h) in order.
template <int D, int MODE, class Args>
static hipError_t launch_cfg(int cfg, const Args& a, hipStream_t s) {
    switch (cfg) {
        case 0: return launch_one<D, MODE, 32, 16, 640>(a, s);
        case 1: return launch_one<D, MODE, 64, 8, 576>(a, s);
        case 2: return launch_one<D, MODE, 128, 4, 576>(a, s);
        case 3: return launch_one<D, MODE, 256, 2, 576>(a, s);
        case 4: return launch_one<D, MODE, 27, 8, 256>(a, s);
        case 5: return launch_one<D, MODE, 59, 4, 256>(a, s);
        case 6: return launch_one<D, MODE, 59, 8, 512>(a, s);
        case 7: return launch_one<D, MODE, 27, 16, 512>(a, s);
        case 8: return launch_one<D, MODE, 123, 4, 512>(a, s);
        case 9: return launch_one<D, MODE, 59, 16, 1024>(a, s);
        default: return hipErrorInvalidValue;
    }
}

template <int D, int MODE>
static hipError_t launch_ens_cfg(int cfg, const EnsPassArgs& e, unsigned grid, hipStream_t s) {
    switch (cfg) {
        case 0: return launch_ens_one<D, MODE, 32, 16, 640>(e, grid, s);
        case 1: return launch_ens_one<D, MODE, 64, 8, 576>(e, grid, s);
        case 2: return launch_ens_one<D, MODE, 128, 4, 576>(e, grid, s);
        case 3: return launch_ens_one<D, MODE, 256, 2, 576>(e, grid, s);
        case 4: return launch_ens_one<D, MODE, 27, 8, 256>(e, grid, s);
        case 5: return launch_ens_one<D, MODE, 59, 4, 256>(e, grid, s);
        case 6: return launch_ens_one<D, MODE, 59, 8, 512>(e, grid, s);
        case 7: return launch_ens_one<D, MODE, 27, 16, 512>(e, grid, s);
        case 8: return launch_ens_one<D, MODE, 123, 4, 512>(e, grid, s);
        case 9: return launch_ens_one<D, MODE, 59, 16, 1024>(e, grid, s);
        default: return hipErrorInvalidValue;
    }
}

// This translation unit instantiates one flux mode (EM_MODE).
hipError_t EM_LAUNCHER(int d, int cfg, const PassArgs& a, hipStream_t s) {
    switch (d) {
        case 0: return launch_cfg<0, EM_MODE>(cfg, a, s);
        case 1: return launch_cfg<1, EM_MODE>(cfg, a, s);
        case 2: return launch_cfg<2, EM_MODE>(cfg, a, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t EM_LAUNCHER32(int d, int cfg, const PassArgsT<float>& a, hipStream_t s) {
    switch (d) {
        case 0: return launch_cfg<0, EM_MODE>(cfg, a, s);
        case 1: return launch_cfg<1, EM_MODE>(cfg, a, s);
        case 2: return launch_cfg<2, EM_MODE>(cfg, a, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t EM_ENS_LAUNCHER(int d, int cfg, const EnsPassArgs& e, unsigned grid, hipStream_t s) {
    switch (d) {
        case 0: return launch_ens_cfg<0, EM_MODE>(cfg, e, grid, s);
        case 1: return launch_ens_cfg<1, EM_MODE>(cfg, e, grid, s);
        case 2: return launch_ens_cfg<2, EM_MODE>(cfg, e, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t EM_AMR_MULTI_LAUNCHER(int d, const AmrMultiArgs& m, unsigned grid, hipStream_t s) {
    switch (d) {
        case 0: return launch_amr_multi_d<0, EM_MODE>(m, grid, s);
        case 1: return launch_amr_multi_d<1, EM_MODE>(m, grid, s);
        case 2: return launch_amr_multi_d<2, EM_MODE>(m, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t EM_AMR_LAUNCHER(int d, int b, int shape, const AmrPassArgs& pa, unsigned grid, hipStream_t s) {
    switch (d) {
        case 0: return launch_amr_b<0, EM_MODE>(b, shape, pa, grid, s);
        case 1: return launch_amr_b<1, EM_MODE>(b, shape, pa, grid, s);
        case 2: return launch_amr_b<2, EM_MODE>(b, shape, pa, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t EM_AMR_LAUNCHER32(int d, int b, int shape, const AmrPassArgsT<float>& pa, unsigned grid, hipStream_t s) {
    switch (d) {
        case 0: return launch_amr_b<0, EM_MODE>(b, shape, pa, grid, s);
        case 1: return launch_amr_b<1, EM_MODE>(b, shape, pa, grid, s);
        case 2: return launch_amr_b<2, EM_MODE>(b, shape, pa, grid, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace EM_VARIANT
}  // namespace em
