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
#elif EM_MODE == 1
#define EM_LAUNCHER launch_pass_weno
#else
#define EM_LAUNCHER launch_pass_central
#endif
#ifndef EM_WAVES_PER_EU
#define EM_WAVES_PER_EU 4
#endif

namespace em {
namespace EM_VARIANT {

// Conserved state of the virtual ghost cell at pencil coordinate g (outside
// [0, n)) of the pencil starting at cell `pen`.
template <int D>
__device__ __forceinline__ void ghost_state(const PassArgs& a, long long pen, int g, double u[5]) {
    const long long sD = a.stride[D];
    auto ld = [&](int x, double* o) {
#pragma unroll
        for (int k = 0; k < 5; ++k) o[k] = a.U[k * a.ncell + pen + (long long)x * sD];
    };
    bc_ghost<D>(a.eos, a.n[D], g, a.bc_kind, a.bc_state, ld, u);
}

template <int D, int MODE, int T, int P, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(EM_WAVES_PER_EU))) void axis_pass(PassArgs a) {
    constexpr int C = T + 10;  // cells held in LDS
    constexpr int M = T + 5;   // midpoints
    constexpr int NF = T + 1;  // faces
    constexpr bool HYB = MODE == MODE_HYBRID;
    constexpr bool CEN = MODE == MODE_CENTRAL;
    constexpr int CS = HYB ? C : 1;
    constexpr int FS = HYB ? NF : 1;
    constexpr int FITER = (NF * P + NT - 1) / NT;
    __shared__ double sw[5][C][P];
    __shared__ double sm[5][M][P];
    __shared__ double ss[2][CS][P];
    __shared__ double sth[FS][P];
    __shared__ double sred[NT / 64 + 1];
    __shared__ unsigned int scnt[3];

    if (a.sc->active == 0) return;
    const int tid = threadIdx.x;
    const int bid = blockIdx.x;
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
    const Eos e = a.eos;
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
        double u[5], w[5];
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
        if constexpr (HYB) {
            // Euler::sensor, system.rs:409-413
            ss[0][c][p] = log(fmax(w[0], 1e-300));
            ss[1][c][p] = log(fmax(w[4] + e.p_inf, 1e-300));
        }
    }
    __syncthreads();

    // ---- 2. hybrid blend weights (scheme.rs:171-183)
    if constexpr (HYB) {
        for (int it = tid; it < NF * P; it += NT) {
            const int p = it % P, f = it / P;
            if (p >= Pv || f > Tv) continue;
            double r = 0.0;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double q[6];
#pragma unroll
                for (int j = 0; j < 6; ++j) q[j] = ss[k][f + 2 + j][p];
                r = fmax(fmax(r, d4(q[0], q[1], q[2], q[3], q[4])), d4(q[1], q[2], q[3], q[4], q[5]));
            }
            double x = (r / a.tol - 0.5) / 0.5;
            if (x < 0.0) x = 0.0;
            if (x > 1.0) x = 1.0;
            sth[f][p] = isfinite(r) ? x * x * (3.0 - 2.0 * x) : 1.0;
        }
        __syncthreads();
    }

    unsigned int n_shock = 0, n_fb = 0, n_lim = 0;

    // ---- 3. WCNS midpoints (midpoint m sits between cells c0-3+m and c0-2+m)
    if constexpr (!CEN) {
        for (int it = tid; it < M * P; it += NT) {
            const int p = it % P, m = it / P;
            if (p >= Pv || m >= Tv + 5) continue;
            if constexpr (HYB) {
                bool need = false;
                const int flo = max(0, m - 4), fhi = min(Tv, m);
                for (int f = flo; f <= fhi; ++f) need = need || sth[f][p] > 0.0;
                if (!need) continue;
            }
            double s[6][5];
#pragma unroll
            for (int j = 0; j < 6; ++j)
#pragma unroll
                for (int k = 0; k < 5; ++k) s[j][k] = sw[k][m + j][p];
            double f[5];
            const bool fb = midpoint_flux<D>(e, s, f);
#pragma unroll
            for (int k = 0; k < 5; ++k) sm[k][m][p] = f[k];
            if (fb) {
                const int j = c0 - 3 + m;  // global midpoint index, owned once per pencil
                const bool own = (j >= c0 && j < c0 + Tv) || (j < 0 && c0 == 0) || (j >= nD && c0 + Tv == nD);
                n_fb += own ? 1u : 0u;
            }
        }
        __syncthreads();
    }

    // ---- 4. face fluxes (face f between cells c0+f-1 and c0+f)
    double lam = 0.0;
    if (a.ndim_pos > 0.0) {
        const double dt = a.stage_dt_host ? a.stage_dt : a.sc->dt;
        lam = a.ndim_pos * dt / a.dx_d;  // solver.rs:393
    }
    double FR[FITER][5];
#pragma unroll
    for (int fi = 0; fi < FITER; ++fi) {
        const int it = tid + fi * NT;
        const int p = it % P, f = it / P;
        if (it >= NF * P || p >= Pv || f > Tv) continue;
        // a face shared by two tiles is counted by the tile on its right
        const unsigned own_face = (f < Tv || c0 + Tv == nD) ? 1u : 0u;
        double theta;
        if constexpr (MODE == MODE_WENO)
            theta = 1.0;
        else if constexpr (CEN)
            theta = 0.0;
        else
            theta = sth[f][p];
        double fo[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        double cen[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (theta < 1.0) {
            // KEEP-PE 6th-order central flux, scheme.rs:188-214
            const double a6[3] = {A6_0, A6_1, A6_2};
#pragma unroll
            for (int li = 0; li < 3; ++li) {
                const int lag = li + 1;
#pragma unroll
                for (int mm = 0; mm < lag; ++mm) {
                    double wa[5], wb[5], ft[5];
#pragma unroll
                    for (int k = 0; k < 5; ++k) {
                        wa[k] = sw[k][f + 4 - mm][p];
                        wb[k] = sw[k][f + 4 - mm + lag][p];
                    }
                    split_flux<D>(e, wa, wb, ft);
#pragma unroll
                    for (int k = 0; k < 5; ++k) cen[k] += 2.0 * a6[li] * ft[k];
                }
            }
        }
        if (theta == 0.0) {
#pragma unroll
            for (int k = 0; k < 5; ++k) fo[k] = cen[k];
        } else {
            n_shock += own_face;
            // F = sum_s S6[s] (F_{i+s} + F_{i-s}), order s=0, i+1, i-1, i+2, i-2
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                double v = 0.0;
                v += S6_0 * sm[k][f + 2][p];
                v += S6_1 * sm[k][f + 3][p];
                v += S6_1 * sm[k][f + 1][p];
                v += S6_2 * sm[k][f + 4][p];
                v += S6_2 * sm[k][f][p];
                fo[k] = v;
            }
            if (theta < 1.0) {
#pragma unroll
                for (int k = 0; k < 5; ++k) fo[k] = theta * fo[k] + (1.0 - theta) * cen[k];
            }
        }
        if (lam > 0.0) {
            // positivity limiter, scheme.rs:401-444
            double wl[5], wr[5], ul[5], ur[5], fl[5], fr[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                wl[k] = sw[k][f + 4][p];
                wr[k] = sw[k][f + 5][p];
            }
            to_cons(e, wl, ul);
            to_cons(e, wr, ur);
            phys_flux<D>(wl, ul, fl);
            phys_flux<D>(wr, ur, fr);
            const double tl = 2.0 * lam;
            auto ok = [&](const double* fx) {
                double qa[5], qb[5], pa[5], pb[5];
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    qa[k] = ul[k] - tl * (fx[k] - fl[k]);
                    qb[k] = ur[k] + tl * (fx[k] - fr[k]);
                }
                to_prim(e, qa, pa);
                to_prim(e, qb, pb);
                return admissible(e, pa[0], pa[4]) && admissible(e, pb[0], pb[4]);
            };
            bool fin = true;
#pragma unroll
            for (int k = 0; k < 5; ++k) fin = fin && isfinite(fo[k]);
            if (!(fin && ok(fo))) {
                constexpr int un = 1 + D;
                const double al = fmax(fabs(wl[un]) + sound_speed(e, wl[0], wl[4]),
                                       fabs(wr[un]) + sound_speed(e, wr[0], wr[4]));
                double flf[5], fh[5];
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    flf[k] = 0.5 * (fl[k] + fr[k]) - 0.5 * al * (ur[k] - ul[k]);
                    fh[k] = fo[k];
                }
                n_lim += own_face;
                double lo = 0.0, hi = 1.0;
                for (int itb = 0; itb < 30; ++itb) {
                    const double mid = 0.5 * (lo + hi);
                    double bl[5];
#pragma unroll
                    for (int k = 0; k < 5; ++k) bl[k] = isfinite(fh[k]) ? mid * fh[k] + (1.0 - mid) * flf[k] : flf[k];
                    if (ok(bl))
                        lo = mid;
                    else
                        hi = mid;
                }
#pragma unroll
                for (int k = 0; k < 5; ++k) fo[k] = isfinite(fh[k]) ? lo * fh[k] + (1.0 - lo) * flf[k] : flf[k];
            }
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) FR[fi][k] = fo[k];
    }
    __syncthreads();  // every midpoint read is done: reuse sm for face fluxes
#pragma unroll
    for (int fi = 0; fi < FITER; ++fi) {
        const int it = tid + fi * NT;
        const int p = it % P, f = it / P;
        if (it >= NF * P || p >= Pv || f > Tv) continue;
#pragma unroll
        for (int k = 0; k < 5; ++k) sm[k][f][p] = FR[fi][k];
    }
    __syncthreads();

    // ---- 5. divergence, fused RK stage and CFL rate
    double rate = 0.0;
    const double dt = a.sc->dt;
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
        double r[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const double prev = a.first ? 0.0 : a.R[k * nc + idx];
            r[k] = prev - (sm[k][c + 1][p] - sm[k][c][p]) * a.inv_dx;
        }
        if (a.rk_stage == 0) {
#pragma unroll
            for (int k = 0; k < 5; ++k) a.R[k * nc + idx] = r[k];
        } else {
            double un_[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const double u0 = a.U0[k * nc + idx];
                const double us = a.U[k * nc + idx];
                double o;
                if (a.rk_stage == 1)
                    o = u0 + dt * r[k];
                else if (a.rk_stage == 2)
                    o = 0.75 * u0 + 0.25 * (us + dt * r[k]);
                else
                    o = u0 / 3.0 + 2.0 / 3.0 * (us + dt * r[k]);
                un_[k] = o;
                a.Uout[k * nc + idx] = o;
            }
            if (a.want_rate) {
                double w[5];
                to_prim(e, un_, w);
                const double c = sound_speed(e, w[0], w[4]);
                double ac = -0.0;
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    if (a.active[d]) ac += (fabs(w[1 + d]) + c) / a.dx_all[d];
                rate = fmax(rate, ac);
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
static hipError_t launch_one(const PassArgs& a, hipStream_t s) {
    const long long nblk = (long long)a.ntile * a.ngroup * a.n[a.slow_axis];
    hipLaunchKernelGGL((axis_pass<D, MODE, T, P, NT>), dim3((unsigned)nblk), dim3(NT), 0, s, a);
    return hipGetLastError();
}

// Tile table: must list the shapes of kTileCfgs (em_pass.h) in order.
template <int D, int MODE>
static hipError_t launch_cfg(int cfg, const PassArgs& a, hipStream_t s) {
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

}  // namespace EM_VARIANT
}  // namespace em
