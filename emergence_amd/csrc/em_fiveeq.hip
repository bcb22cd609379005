// Two-phase Kapila five-equation system (FiveEq, seven variables) on a
// uniform grid: the reference's GpuSolver with GpuSystem::FiveEq
// (fluid/crates/compressible-cuda/src/lib.rs:50-82, 155-477), with the
// arithmetic of the CPU Solver<7, FiveEq> (fluid/crates/compressible/src/
// solver.rs, scheme.rs, multiphase.rs), which is its parity target.
//
// Kernels:
//   fe_pass<D, MODE, T, P, NT>  the whole per-axis residual of a tile of T
//       cells x P pencils in one launch, LDS-staged like the Euler pass:
//       load + primitives (+ sensor), hybrid blend weight, WCNS midpoints
//       (flux and contact speed, each once), faces (S6 combination, KEEP-PE
//       central flux, blend, positivity limiter, face velocity), cells
//       (divergence, the non-conservative (alpha + K) div u term of the
//       volume fraction, fused SSP-RK3 stage on the last axis);
//   k_fe_rate   CFL rate (sum of (|u_d| + c)/dx_d) and the dilatation-rate
//       limit |Dp/Dt|/(p + p_inf,min) of every interior cell
//       (solver.rs:478-523, multiphase.rs:373-391), wave-reduced into slots.
// Built without FMA contraction: results are bit-identical to the oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/em_hip.h"
#include "em_fiveeq.h"
#include "em_types.h"

namespace em {
namespace fe {

static thread_local std::string g_err;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(EM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define FECHK(x) ::em::fe::hip_check((x), #x)

// Device scalars of one FiveEq solver; slot accumulators on separate
// 128-byte lines, reduced on read.
struct alignas(256) FeScalars {
    double dt;
    int active;
    int pad_;
    unsigned long long rate[kSlots][16];
    unsigned long long source[kSlots][16];
    unsigned long long counters[3][kSlots][16];
};

// Pass arguments in the real type Real of the state (EOS parameters,
// boundary states, 1/dx and the sensor tolerance rounded to Real, as the
// reference's R::from_f64); the limiter's lambda is formed in double from
// ndim_pos, the device dt and dx_d, then rounded.
template <class Real>
struct FePassArgsT {
    const Real* U;  // stage input (7 planes of ncell)
    Real* R;        // residual accumulator
    const Real* U0;
    Real* Uout;
    FeScalars* sc;
    int n[3];
    long long stride[3];
    long long ncell;
    SysT<Real> sys;
    int bc_kind[2];
    Real bc_state[2][7];
    Real inv_dx, tol;
    double ndim_pos, dx_d;
    int first, rk_stage;
    int fast_axis, slow_axis, dfast, ntile, ngroup;
    int stage_dt_host;
    double stage_dt;
};
using FePassArgs = FePassArgsT<double>;

template <class Real, int MODE, int T, int P>
struct FeLds {
    static constexpr bool HYB = MODE == MODE_HYBRID;
    Real sw[7][T + 10][P];
    Real sre[T + 10][P];  // rho e(w) of each tile cell (central pair fluxes, limiter)
    Real sm[8][T + 5][P];  // midpoints: flux + contact speed; then faces: flux + face velocity
    Real ss[3][HYB ? T + 10 : 1][P];
    Real sth[HYB ? T + 1 : 1][P];
    unsigned int scnt[3];
};

#ifndef EM_FE_WAVES_PER_EU
#define EM_FE_WAVES_PER_EU 2
#endif

template <class Real, int D, int MODE, int T, int P, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(EM_FE_WAVES_PER_EU))) void fe_pass(FePassArgsT<Real> a) {
    constexpr int C = T + 10;
    constexpr int M = T + 5;
    constexpr int NF = T + 1;
    constexpr bool HYB = MODE == MODE_HYBRID;
    constexpr bool CEN = MODE == MODE_CENTRAL;
    constexpr int un = IU + D;
    __shared__ FeLds<Real, MODE, T, P> L;
    auto& sw = L.sw;
    auto& sre = L.sre;
    auto& sm = L.sm;
    auto& ss = L.ss;
    auto& sth = L.sth;
    if (a.sc->active == 0) return;
    const int tid = threadIdx.x;
    const int bid = xcd_major((int)blockIdx.x, (int)gridDim.x);
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
    const SysT<Real>& sys = a.sys;
    if (tid < 3) L.scnt[tid] = 0;

    // ---- 1. load + primitives (+ sensor ln rho, ln(p + p_inf,mix), alpha)
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
        Real u[7], w[7];
        if (g >= 0 && g < nD) {
            const long long idx = pen + (long long)g * sD;
#pragma unroll
            for (int k = 0; k < 7; ++k) u[k] = a.U[k * nc + idx];
        } else {
            auto ld = [&](int x, Real* o) {
#pragma unroll
                for (int k = 0; k < 7; ++k) o[k] = a.U[k * nc + pen + (long long)x * sD];
            };
            bc_ghost<D>(sys, nD, g, a.bc_kind, a.bc_state, ld, u);
        }
        to_prim(sys, u, w);
#pragma unroll
        for (int k = 0; k < 7; ++k) sw[k][c][p] = w[k];
        sre[c][p] = rho_e(sys, w);
        if constexpr (HYB) {
            // FiveEq::sensor, multiphase.rs:385-390
            ss[0][c][p] = log(fmax(mix_rho(w), Real(1e-300)));
            ss[1][c][p] = log(fmax(w[IP] + p_inf_mix(sys, w), Real(1e-300)));
            ss[2][c][p] = w[IA];
        }
    }
    __syncthreads();

    double lam_d = 0.0;
    if (a.ndim_pos > 0.0) {
        const double dt = a.stage_dt_host ? a.stage_dt : a.sc->dt;
        lam_d = a.ndim_pos * dt / a.dx_d;  // solver.rs:393
    }
    const Real lam = Real(lam_d);
    unsigned n_shock = 0, n_fb = 0, n_lim = 0;

    // ---- 2. hybrid blend weights (scheme.rs:171-183)
    if constexpr (HYB) {
        for (int it = tid; it < NF * P; it += NT) {
            const int p = it % P, f = it / P;
            if (p >= Pv || f > Tv) continue;
            Real r = Real(0.0);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Real q[6];
#pragma unroll
                for (int j = 0; j < 6; ++j) q[j] = ss[k][f + 2 + j][p];
                r = fmax(fmax(r, d4f(q[0], q[1], q[2], q[3], q[4])), d4f(q[1], q[2], q[3], q[4], q[5]));
            }
            Real x = (r / a.tol - Real(0.5)) / Real(0.5);
            if (x < Real(0.0)) x = Real(0.0);
            if (x > Real(1.0)) x = Real(1.0);
            sth[f][p] = isfinite(r) ? x * x * (Real(3.0) - Real(2.0) * x) : Real(1.0);
        }
        __syncthreads();
    }

    // ---- 3. WCNS midpoints (midpoint m between tile cells m+2 and m+3)
    if constexpr (!CEN) {
        for (int it = tid; it < M * P; it += NT) {
            const int p = it % P, m = it / P;
            if (p >= Pv || m >= Tv + 5) continue;
            if constexpr (HYB) {
                bool need = false;
                const int flo = max(0, m - 4), fhi = min(Tv, m);
                for (int f = flo; f <= fhi; ++f) need = need || sth[f][p] > Real(0.0);
                if (!need) continue;
            }
            Real f[7], us;
            const bool fb = midpoint_flux<D>(sys, [&](int j, int k) { return sw[k][m + j][p]; }, f, us);
#pragma unroll
            for (int k = 0; k < 7; ++k) sm[k][m][p] = f[k];
            sm[7][m][p] = us;
            if (fb) {
                const int j = c0 - 3 + m;  // global midpoint index, counted once per pencil
                const bool own = (j >= c0 && j < c0 + Tv) || (j < 0 && c0 == 0) || (j >= nD && c0 + Tv == nD);
                n_fb += own ? 1u : 0u;
            }
        }
        __syncthreads();
    }

    // ---- 4. faces (face f between tile cells f+4 and f+5)
    constexpr int FITER = (NF * P + NT - 1) / NT;
    Real FR[FITER][8];
#pragma unroll
    for (int fi = 0; fi < FITER; ++fi) {
        const int it = tid + fi * NT;
        const int p = it % P, f = it / P;
        if (it >= NF * P || p >= Pv || f > Tv) continue;
        const unsigned own_face = (f < Tv || c0 + Tv == nD) ? 1u : 0u;
        Real theta;
        if constexpr (MODE == MODE_WENO)
            theta = Real(1.0);
        else if constexpr (CEN)
            theta = Real(0.0);
        else
            theta = sth[f][p];
        Real fo[7], cen[7];
        Real uo = Real(0.0), uc = Real(0.0);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            fo[k] = Real(0.0);
            cen[k] = Real(0.0);
        }
        if (theta < Real(1.0)) {
            // KEEP-PE 6th-order central flux and face velocity, scheme.rs:188-214
            const Real a6[3] = {Real(0.75), -Real(0.15), Real(1.0) / Real(60.0)};
#pragma unroll
            for (int li = 0; li < 3; ++li) {
                const int lag = li + 1;
#pragma unroll
                for (int mm = 0; mm < lag; ++mm) {
                    Real wa[7], wb[7], ft[7];
#pragma unroll
                    for (int k = 0; k < 7; ++k) {
                        wa[k] = sw[k][f + 4 - mm][p];
                        wb[k] = sw[k][f + 4 - mm + lag][p];
                    }
                    split_flux_re<D>(wa, wb, sre[f + 4 - mm][p], sre[f + 4 - mm + lag][p], ft);
#pragma unroll
                    for (int k = 0; k < 7; ++k) cen[k] += Real(2.0) * a6[li] * ft[k];
                    uc += a6[li] * (wa[un] + wb[un]);
                }
            }
        }
        if (theta == Real(0.0)) {
#pragma unroll
            for (int k = 0; k < 7; ++k) fo[k] = cen[k];
            uo = uc;
        } else {
            n_shock += own_face;
            // F = sum_s S6[s] (F_{i+s} + F_{i-s}), order s=0, i+1, i-1, i+2, i-2
            const Real s0 = Real(75.0) / Real(64.0) - Real(25.0) / Real(384.0) + Real(3.0) / Real(640.0);
            const Real s1 = -Real(25.0) / Real(384.0) + Real(3.0) / Real(640.0);
            const Real s2 = Real(3.0) / Real(640.0);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                Real v = Real(0.0);
                v += s0 * sm[k][f + 2][p];
                v += s1 * sm[k][f + 3][p];
                v += s1 * sm[k][f + 1][p];
                v += s2 * sm[k][f + 4][p];
                v += s2 * sm[k][f][p];
                if (k < 7)
                    fo[k] = v;
                else
                    uo = v;
            }
            if (theta < Real(1.0)) {
#pragma unroll
                for (int k = 0; k < 7; ++k) fo[k] = theta * fo[k] + (Real(1.0) - theta) * cen[k];
                uo = theta * uo + (Real(1.0) - theta) * uc;
            }
        }
        if (lam > Real(0.0)) {
            // positivity limiter, scheme.rs:401-450
            Real wl[7], wr[7], ul[7], ur[7], fl[7], fr[7];
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                wl[k] = sw[k][f + 4][p];
                wr[k] = sw[k][f + 5][p];
            }
            to_cons_re(wl, sre[f + 4][p], ul);
            to_cons_re(wr, sre[f + 5][p], ur);
            phys_flux<D>(wl, ul, fl);
            phys_flux<D>(wr, ur, fr);
            const Real tl = Real(2.0) * lam;
            auto ok = [&](const Real* fx) {
                Real qa[7], qb[7], pa[7], pb[7];
#pragma unroll
                for (int k = 0; k < 7; ++k) {
                    qa[k] = ul[k] - tl * (fx[k] - fl[k]);
                    qb[k] = ur[k] + tl * (fx[k] - fr[k]);
                }
                to_prim(sys, qa, pa);
                to_prim(sys, qb, pb);
                return admissible(sys, pa) && admissible(sys, pb);
            };
            bool fin = true;
#pragma unroll
            for (int k = 0; k < 7; ++k) fin = fin && isfinite(fo[k]);
            if (!(fin && ok(fo))) {
                const Real al = fmax(fabs(wl[un]) + sound_speed(sys, wl), fabs(wr[un]) + sound_speed(sys, wr));
                Real flf[7], fh[7];
#pragma unroll
                for (int k = 0; k < 7; ++k) {
                    flf[k] = Real(0.5) * (fl[k] + fr[k]) - Real(0.5) * al * (ur[k] - ul[k]);
                    fh[k] = fo[k];
                }
                n_lim += own_face;
                Real lo = Real(0.0), hi = Real(1.0);
                for (int itb = 0; itb < 30; ++itb) {
                    const Real mid = Real(0.5) * (lo + hi);
                    Real bl[7];
#pragma unroll
                    for (int k = 0; k < 7; ++k) bl[k] = isfinite(fh[k]) ? mid * fh[k] + (Real(1.0) - mid) * flf[k] : flf[k];
                    if (ok(bl))
                        lo = mid;
                    else
                        hi = mid;
                }
#pragma unroll
                for (int k = 0; k < 7; ++k) fo[k] = isfinite(fh[k]) ? lo * fh[k] + (Real(1.0) - lo) * flf[k] : flf[k];
                const Real um = Real(0.5) * (wl[un] + wr[un]);
                uo = isfinite(uo) ? lo * uo + (Real(1.0) - lo) * um : um;
            }
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) FR[fi][k] = fo[k];
        FR[fi][7] = uo;
    }
    __syncthreads();  // every midpoint read is done: reuse sm for the faces
#pragma unroll
    for (int fi = 0; fi < FITER; ++fi) {
        const int it = tid + fi * NT;
        const int p = it % P, f = it / P;
        if (it >= NF * P || p >= Pv || f > Tv) continue;
#pragma unroll
        for (int k = 0; k < 8; ++k) sm[k][f][p] = FR[fi][k];
    }
    __syncthreads();

    // ---- 5. cells: divergence, non-conservative term, fused RK stage
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
        Real r[7], w[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const Real prev = a.first ? Real(0.0) : a.R[k * nc + idx];
            r[k] = prev - (sm[k][c + 1][p] - sm[k][c][p]) * a.inv_dx;
            w[k] = sw[k][c + 5][p];
        }
        // solver.rs:440-442: + c(w_i) (U_{i+1/2} - U_{i-1/2}) / dx
        r[IA] += nc_coef(sys, w) * (sm[7][c + 1][p] - sm[7][c][p]) * a.inv_dx;
        if (a.rk_stage == 0) {
#pragma unroll
            for (int k = 0; k < 7; ++k) a.R[k * nc + idx] = r[k];
        } else {
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const Real u0 = a.U0[k * nc + idx];
                Real o;
                if (a.rk_stage == 1) {
                    o = u0 + dt * r[k];
                } else {
                    const Real us = a.U[k * nc + idx];
                    // solver.rs:113-116 and 716-726: dest = U^n, old = the stage
                    if (a.rk_stage == 2)
                        o = Real(0.75) * u0 + Real(0.25) * (us + dt * r[k]);
                    else
                        o = u0 / Real(3.0) + Real(2.0) / Real(3.0) * (us + dt * r[k]);
                }
                a.Uout[k * nc + idx] = o;
            }
        }
    }

#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        n_shock += __shfl_xor(n_shock, off);
        n_fb += __shfl_xor(n_fb, off);
        n_lim += __shfl_xor(n_lim, off);
    }
    if ((tid & 63) == 0) {
        if (n_shock) atomicAdd(&L.scnt[0], n_shock);
        if (n_fb) atomicAdd(&L.scnt[1], n_fb);
        if (n_lim) atomicAdd(&L.scnt[2], n_lim);
    }
    __syncthreads();
    if (tid == 0) {
        const int slot = bid % kSlots;
        for (int q = 0; q < 3; ++q)
            if (L.scnt[q]) atomicAdd(&a.sc->counters[q][slot][0], (unsigned long long)L.scnt[q]);
    }
}

// Conserved neighbour of interior cell x (pencil coordinate) along d through
// the boundary condition when x is a boundary cell (solver.rs:592-640).
template <int D, class Real, class LD>
__device__ __forceinline__ void neighbor_conserved(const SysT<Real>& s, int n, int x, bool high, const int kind[2],
                                                   const Real (*state)[7], LD ld, Real* out) {
    const bool boundary = high ? x == n - 1 : x == 0;
    if (!boundary) {
        ld(high ? x + 1 : x - 1, out);
        return;
    }
    const int side = high ? 1 : 0;
    switch (kind[side]) {
        case 0: ld(high ? 0 : n - 1, out); return;
        case 1: ld(x, out); return;
        case 3:
#pragma unroll
            for (int q = 0; q < 7; ++q) out[q] = state[side][q];
            return;
        case 2:
            ld(x, out);
            out[IU + D] *= -Real(1.0);
            return;
        default: {
            Real ub[7], ub1[7];
            ld(x, ub);
            ld(high ? x - 1 : x + 1, ub1);
            characteristic_ghost<D>(s, ub, ub1, state[side], side, Real(1.0), out);
            return;
        }
    }
}

template <class Real>
struct FeRateArgsT {
    const Real* U;
    FeScalars* sc;
    long long ncell;
    int n[3];
    long long stride[3];
    Real dx[3];
    int act[3];
    SysT<Real> sys;
    int bc_kind[3][2];
    Real bc_state[3][2][7];
};

template <int D, class Real>
__device__ __forceinline__ Real divu_axis(const FeRateArgsT<Real>& a, long long cell, int x) {
    auto ld = [&](int y, Real* o) {
        const long long q = cell + (long long)(y - x) * a.stride[D];
#pragma unroll
        for (int k = 0; k < 7; ++k) o[k] = a.U[k * a.ncell + q];
    };
    Real lo[7], hi[7];
    neighbor_conserved<D>(a.sys, a.n[D], x, false, a.bc_kind[D], a.bc_state[D], ld, lo);
    neighbor_conserved<D>(a.sys, a.n[D], x, true, a.bc_kind[D], a.bc_state[D], ld, hi);
    const Real va = lo[IU + D] / (lo[0] + lo[1]);  // velocity_conserved, multiphase.rs:158-161
    const Real vb = hi[IU + D] / (hi[0] + hi[1]);
    return fabs(vb - va) / (Real(2.0) * a.dx[D]);
}

// CFL rate and dilatation-rate limit of every interior cell (solver.rs:478-
// 523): rate = max sum_d (|u_d| + c)/dx_d, source = max dilatation_rate(w,
// sum_d |v_{+d} - v_{-d}| / (2 dx_d)); one atomicMax per workgroup per slot.
template <class Real>
__global__ __launch_bounds__(256) void k_fe_rate(FeRateArgsT<Real> a) {
    double rate = 0.0, source = 0.0;  // maxima of Real values, exact in double
    for (long long cell = blockIdx.x * 256ll + threadIdx.x; cell < a.ncell; cell += (long long)gridDim.x * 256ll) {
        Real u[7], w[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) u[k] = a.U[k * a.ncell + cell];
        to_prim(a.sys, u, w);
        const Real c = sound_speed(a.sys, w);
        Real ac = -Real(0.0);
#pragma unroll
        for (int d = 0; d < 3; ++d)
            if (a.act[d]) ac += (fabs(w[IU + d]) + c) / a.dx[d];
        rate = fmax(rate, (double)ac);
        const int x0 = (int)(cell / a.stride[0]);
        const int x1 = (int)((cell / a.stride[1]) % a.n[1]);
        const int x2 = (int)(cell % a.n[2]);
        Real divu = Real(0.0);
        if (a.act[0]) divu += divu_axis<0>(a, cell, x0);
        if (a.act[1]) divu += divu_axis<1>(a, cell, x1);
        if (a.act[2]) divu += divu_axis<2>(a, cell, x2);
        source = fmax(source, (double)dilatation_rate(a.sys, w, divu));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        rate = fmax(rate, __shfl_xor(rate, off));
        source = fmax(source, __shfl_xor(source, off));
    }
    __shared__ double sr[4], sq[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        sr[wid] = rate;
        sq[wid] = source;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double rm = 0.0, qm = 0.0;
        for (int i = 0; i < 4; ++i) {
            rm = fmax(rm, sr[i]);
            qm = fmax(qm, sq[i]);
        }
        rm = rm > 0.0 ? rm : 0.0;  // IEEE bits of +x order like x
        qm = qm > 0.0 ? qm : 0.0;
        const int slot = blockIdx.x % kSlots;
        atomicMax(&a.sc->rate[slot][0], (unsigned long long)__double_as_longlong(rm));
        atomicMax(&a.sc->source[slot][0], (unsigned long long)__double_as_longlong(qm));
    }
}

__global__ void k_fe_zero_rate(FeScalars* sc) {
    if (threadIdx.x < kSlots) {
        sc->rate[threadIdx.x][0] = 0ull;
        sc->source[threadIdx.x][0] = 0ull;
    }
}
__global__ void k_fe_set_dt(FeScalars* sc, double dt) { sc->dt = dt; }

// Tile shapes (cells along the axis x pencils, threads)
struct FeTile {
    int T, P, NT;
};
constexpr FeTile kFeTiles[] = {{27, 8, 256}, {59, 4, 256}, {59, 8, 512}, {27, 16, 512}, {123, 4, 512}};
constexpr int kNumFeTiles = 5;

template <int D, int MODE, int T, int P, int NT, class Real>
static hipError_t launch_fe(const FePassArgsT<Real>& a, hipStream_t st) {
    const long long nblk = (long long)a.ntile * a.ngroup * a.n[a.slow_axis];
    hipLaunchKernelGGL((fe_pass<Real, D, MODE, T, P, NT>), dim3((unsigned)nblk), dim3(NT), 0, st, a);
    return hipGetLastError();
}
template <int D, int MODE, class Real>
static hipError_t launch_fe_cfg(int cfg, const FePassArgsT<Real>& a, hipStream_t st) {
    switch (cfg) {
        case 0: return launch_fe<D, MODE, 27, 8, 256>(a, st);
        case 1: return launch_fe<D, MODE, 59, 4, 256>(a, st);
        case 2: return launch_fe<D, MODE, 59, 8, 512>(a, st);
        case 3: return launch_fe<D, MODE, 27, 16, 512>(a, st);
        case 4: return launch_fe<D, MODE, 123, 4, 512>(a, st);
        default: return hipErrorInvalidValue;
    }
}
template <int D, class Real>
static hipError_t launch_fe_mode(int mode, int cfg, const FePassArgsT<Real>& a, hipStream_t st) {
    switch (mode) {
        case MODE_HYBRID: return launch_fe_cfg<D, MODE_HYBRID>(cfg, a, st);
        case MODE_WENO: return launch_fe_cfg<D, MODE_WENO>(cfg, a, st);
        case MODE_CENTRAL: return launch_fe_cfg<D, MODE_CENTRAL>(cfg, a, st);
        default: return hipErrorInvalidValue;
    }
}
template <class Real>
static hipError_t launch_fe_pass(int d, int mode, int cfg, const FePassArgsT<Real>& a, hipStream_t st) {
    switch (d) {
        case 0: return launch_fe_mode<0>(mode, cfg, a, st);
        case 1: return launch_fe_mode<1>(mode, cfg, a, st);
        case 2: return launch_fe_mode<2>(mode, cfg, a, st);
        default: return hipErrorInvalidValue;
    }
}

// The system in real type Real (the reference's R::from_f64 of each phase)
template <class Real>
static SysT<Real> sys_as(const Sys& s) {
    SysT<Real> r;
    for (int k = 0; k < 2; ++k)
        r.e[k] = PhaseT<Real>{Real(s.e[k].gamma), Real(s.e[k].p_inf), Real(s.e[k].b), Real(s.e[k].q)};
    r.kapila = s.kapila;
    return r;
}

// Host object: mesh, boundaries, options, device state in real type Real
// (double: the reference's GpuSolver<f64>; float: GpuSolver<f32>).  The
// host side (mesh, options, time, dt) stays double.
template <class Real>
struct FeCore {
    int64_t n[3];
    double lo[3], dx[3];
    int ng;
    Sys sys;
    int bc_kind[3][2];
    double bc_state[3][2][7];
    int mode, positivity;
    double cfl, tol, t = 0.0;
    int cfg[3] = {0, 0, 0};
    long long ncell;
    long long stride[3];
    int nactive = 0, act[3];
    hipStream_t stream = nullptr;
    Real* U[3] = {nullptr, nullptr, nullptr};
    Real* R = nullptr;
    FeScalars* sc = nullptr;
    uint64_t counters_host[3] = {0, 0, 0};

    bool active(int d) const { return n[d] > 1; }
    int m(int d) const { return active(d) ? (int)(n[d] + 2 * ng) : 1; }
    int g(int d) const { return active(d) ? ng : 0; }

    FeCore(const int64_t nn[3], const double lo_[3], const double hi[3], int ng_, const double ph[8], int kapila,
             const int32_t bk[6], const double bs[42], int mode_, int pos, double cfl_, double tol_, double t0) {
        for (int d = 0; d < 3; ++d) {
            if (nn[d] < 1 || nn[d] > (1 << 30)) throw Error(EM_ERR_ARG, "mesh extents must be in [1, 2^30]");
            n[d] = nn[d];
            lo[d] = lo_[d];
            dx[d] = (hi[d] - lo_[d]) / (double)nn[d];  // Mesh::new, solver.rs:32
        }
        ng = ng_;
        for (int k = 0; k < 2; ++k) sys.e[k] = Phase{ph[4 * k], ph[4 * k + 1], ph[4 * k + 2], ph[4 * k + 3]};
        sys.kapila = kapila ? 1 : 0;
        mode = mode_;
        positivity = pos;
        cfl = cfl_;
        tol = tol_;
        t = t0;
        if (mode < 0 || mode > 2) throw Error(EM_ERR_ARG, "mode must be 0 (Hybrid), 1 (Weno) or 2 (Central)");
        if (ng < (mode == MODE_CENTRAL ? 3 : 5))
            throw Error(EM_ERR_ARG, "ghost layers: need ng >= 5 for Hybrid/Weno, >= 3 for Central");
        for (int d = 0; d < 3; ++d) {
            act[d] = active(d) ? 1 : 0;
            nactive += act[d];
            for (int sd = 0; sd < 2; ++sd) {
                bc_kind[d][sd] = bk[2 * d + sd];
                if (bc_kind[d][sd] < 0 || bc_kind[d][sd] > 4) throw Error(EM_ERR_ARG, "boundary kind out of range");
                for (int q = 0; q < 7; ++q) bc_state[d][sd][q] = bs ? bs[(2 * d + sd) * 7 + q] : 0.0;
            }
        }
        if (nactive == 0) throw Error(EM_ERR_ARG, "mesh has no active axis");
        ncell = n[0] * n[1] * n[2];
        stride[0] = n[1] * n[2];
        stride[1] = n[2];
        stride[2] = 1;
        for (int d = 0; d < 3; ++d) cfg[d] = stride[d] == 1 ? 1 : 0;
        FECHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        for (auto& b : U) FECHK(hipMalloc(&b, sizeof(Real) * 7 * ncell));
        FECHK(hipMalloc(&R, sizeof(Real) * 7 * ncell));
        FECHK(hipMalloc(&sc, sizeof(FeScalars)));
        std::vector<FeScalars> h(1);
        std::memset(h.data(), 0, sizeof(FeScalars));
        h[0].active = 1;
        FECHK(hipMemcpyAsync(sc, h.data(), sizeof(FeScalars), hipMemcpyHostToDevice, stream));
        FECHK(hipStreamSynchronize(stream));
    }
    ~FeCore() {
        for (auto* b : U)
            if (b) (void)hipFree(b);
        if (R) (void)hipFree(R);
        if (sc) (void)hipFree(sc);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void upload(const double* u) {
        std::vector<Real> h(7 * ncell);
        const int m1 = m(1), m2 = m(2);
        for (long long i = 0; i < n[0]; ++i)
            for (long long j = 0; j < n[1]; ++j)
                for (long long k = 0; k < n[2]; ++k) {
                    const long long q = ((i + g(0)) * m1 + (j + g(1))) * m2 + (k + g(2));
                    const long long c = (i * n[1] + j) * n[2] + k;
                    for (int v = 0; v < 7; ++v) h[v * ncell + c] = Real(u[q * 7 + v]);
                }
        FECHK(hipMemcpyAsync(U[0], h.data(), sizeof(Real) * 7 * ncell, hipMemcpyHostToDevice, stream));
        FECHK(hipStreamSynchronize(stream));
    }
    // ghost-padded AoS with the interior filled and zero ghosts
    void to_host(const Real* dev, double* u) {
        std::vector<Real> h(7 * ncell);
        FECHK(hipMemcpyAsync(h.data(), dev, sizeof(Real) * 7 * ncell, hipMemcpyDeviceToHost, stream));
        FECHK(hipStreamSynchronize(stream));
        const long long L = (long long)m(0) * m(1) * m(2);
        std::fill(u, u + 7 * L, 0.0);
        const int m1 = m(1), m2 = m(2);
        for (long long i = 0; i < n[0]; ++i)
            for (long long j = 0; j < n[1]; ++j)
                for (long long k = 0; k < n[2]; ++k) {
                    const long long q = ((i + g(0)) * m1 + (j + g(1))) * m2 + (k + g(2));
                    const long long c = (i * n[1] + j) * n[2] + k;
                    for (int v = 0; v < 7; ++v) u[q * 7 + v] = (double)h[v * ncell + c];
                }
    }

    FePassArgsT<Real> pass_args(int d) const {
        FePassArgsT<Real> a{};
        a.sc = sc;
        for (int e = 0; e < 3; ++e) {
            a.n[e] = (int)n[e];
            a.stride[e] = stride[e];
        }
        a.ncell = ncell;
        a.sys = sys_as<Real>(sys);
        for (int sd = 0; sd < 2; ++sd) {
            a.bc_kind[sd] = bc_kind[d][sd];
            for (int q = 0; q < 7; ++q) a.bc_state[sd][q] = Real(bc_state[d][sd][q]);
        }
        a.inv_dx = Real(1.0 / dx[d]);
        a.dx_d = dx[d];
        a.ndim_pos = positivity ? (double)nactive : 0.0;
        a.tol = Real(tol);
        int tt[2], nt = 0;
        for (int e = 0; e < 3; ++e)
            if (e != d) tt[nt++] = e;
        if (active(tt[1]) || !active(tt[0])) {
            a.fast_axis = tt[1];
            a.slow_axis = tt[0];
        } else {
            a.fast_axis = tt[0];
            a.slow_axis = tt[1];
        }
        a.dfast = stride[d] == 1 ? 1 : 0;
        const FeTile tc = kFeTiles[cfg[d]];
        a.ntile = (int)((n[d] + tc.T - 1) / tc.T);
        a.ngroup = (int)((n[a.fast_axis] + tc.P - 1) / tc.P);
        const long long nblk = (long long)a.ntile * a.ngroup * a.n[a.slow_axis];
        if (nblk > 0x7fffffffll) throw Error(EM_ERR_ARG, "mesh too large for one axis-pass launch");
        return a;
    }

    // residual of Uin over the active axes in order; fused RK stage on the last
    void stage(const Real* Uin, Real* Uout, int rk_stage, int stage_dt_host, double stage_dt) {
        int last = -1;
        for (int d = 0; d < 3; ++d)
            if (active(d)) last = d;
        bool first = true;
        for (int d = 0; d < 3; ++d) {
            if (!active(d)) continue;
            FePassArgsT<Real> a = pass_args(d);
            a.U = Uin;
            a.R = R;
            a.U0 = U[0];
            a.Uout = Uout;
            a.first = first ? 1 : 0;
            a.rk_stage = d == last ? rk_stage : 0;
            a.stage_dt_host = stage_dt_host;
            a.stage_dt = stage_dt;
            FECHK(launch_fe_pass(d, mode, cfg[d], a, stream));
            first = false;
        }
    }

    void read_counters(uint64_t out[3]) {
        std::vector<FeScalars> h(1);
        FECHK(hipMemcpyAsync(h.data(), sc, sizeof(FeScalars), hipMemcpyDeviceToHost, stream));
        FECHK(hipStreamSynchronize(stream));
        for (int q = 0; q < 3; ++q) {
            out[q] = 0;
            for (int i = 0; i < kSlots; ++i) out[q] += h[0].counters[q][i][0];
        }
    }

    uint64_t faces_per_stage() const {
        uint64_t f = 0;
        for (int d = 0; d < 3; ++d)
            if (active(d)) f += (uint64_t)(n[d] + 1) * (ncell / n[d]);
        return f;
    }

    void rhs(double stage_dt, double* out, uint64_t stats[4]) {
        uint64_t before[3], after[3];
        read_counters(before);
        stage(U[0], nullptr, 0, 1, stage_dt);
        read_counters(after);
        to_host(R, out);
        if (stats) {
            stats[0] = faces_per_stage();
            for (int q = 0; q < 3; ++q) stats[1 + q] = after[q] - before[q];
        }
    }

    // SSP-RK3 step (solver.rs:659-738), dt from the host
    void step(double dt) {
        hipLaunchKernelGGL(k_fe_set_dt, dim3(1), dim3(1), 0, stream, sc, dt);
        FECHK(hipGetLastError());
        stage(U[0], U[1], 1, 0, 0.0);
        stage(U[1], U[2], 2, 0, 0.0);
        stage(U[2], U[0], 3, 0, 0.0);
        FECHK(hipStreamSynchronize(stream));
        t += dt;
    }

    double dt_cfl() {
        hipLaunchKernelGGL(k_fe_zero_rate, dim3(1), dim3(kSlots), 0, stream, sc);
        FECHK(hipGetLastError());
        FeRateArgsT<Real> a{};
        a.U = U[0];
        a.sc = sc;
        a.ncell = ncell;
        a.sys = sys_as<Real>(sys);
        for (int d = 0; d < 3; ++d) {
            a.n[d] = (int)n[d];
            a.stride[d] = stride[d];
            a.dx[d] = Real(dx[d]);
            a.act[d] = act[d];
            for (int sd = 0; sd < 2; ++sd) {
                a.bc_kind[d][sd] = bc_kind[d][sd];
                for (int q = 0; q < 7; ++q) a.bc_state[d][sd][q] = Real(bc_state[d][sd][q]);
            }
        }
        const long long nb = std::min<long long>((ncell + 255) / 256, 2048);
        hipLaunchKernelGGL(k_fe_rate<Real>, dim3((unsigned)nb), dim3(256), 0, stream, a);
        FECHK(hipGetLastError());
        std::vector<FeScalars> h(1);
        FECHK(hipMemcpyAsync(h.data(), sc, sizeof(FeScalars), hipMemcpyDeviceToHost, stream));
        FECHK(hipStreamSynchronize(stream));
        uint64_t rb = 0, qb = 0;
        for (int i = 0; i < kSlots; ++i) {
            rb = std::max<uint64_t>(rb, h[0].rate[i][0]);
            qb = std::max<uint64_t>(qb, h[0].source[i][0]);
        }
        const double rate = __builtin_bit_cast(double, rb), source = __builtin_bit_cast(double, qb);
        double dt = cfl / rate;  // solver.rs:514-519
        if (source > 0.0) dt = std::fmin(dt, 0.5 / source);
        return dt;
    }

    // Times every tile shape of each active axis on the current state (one
    // residual pass at the CFL stage dt, after a warm-up pass) and keeps the
    // fastest; the device scalars (counters) are restored afterwards.
    void autotune(int repeats, int32_t out[3]) {
        std::vector<FeScalars> saved(1);
        FECHK(hipMemcpyAsync(saved.data(), sc, sizeof(FeScalars), hipMemcpyDeviceToHost, stream));
        FECHK(hipStreamSynchronize(stream));
        const double dt = dt_cfl();
        hipEvent_t e0, e1;
        FECHK(hipEventCreate(&e0));
        FECHK(hipEventCreate(&e1));
        for (int d = 0; d < 3; ++d) {
            if (!active(d)) continue;
            double best = 1e300;
            int bestc = cfg[d];
            for (int c = 0; c < kNumFeTiles; ++c) {
                if (kFeTiles[c].T > 2 * n[d] && c != 0) continue;
                cfg[d] = c;
                FePassArgsT<Real> a = pass_args(d);
                a.U = U[0];
                a.R = R;
                a.U0 = U[0];
                a.Uout = nullptr;
                a.first = 1;
                a.rk_stage = 0;
                a.stage_dt_host = 1;
                a.stage_dt = std::isfinite(dt) ? dt : 0.0;
                FECHK(launch_fe_pass(d, mode, c, a, stream));  // warm-up
                FECHK(hipEventRecord(e0, stream));
                for (int r = 0; r < repeats; ++r) FECHK(launch_fe_pass(d, mode, c, a, stream));
                FECHK(hipEventRecord(e1, stream));
                FECHK(hipEventSynchronize(e1));
                float ms = 0.f;
                FECHK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    bestc = c;
                }
            }
            cfg[d] = bestc;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        FECHK(hipMemcpyAsync(sc, saved.data(), sizeof(FeScalars), hipMemcpyHostToDevice, stream));
        FECHK(hipStreamSynchronize(stream));
        if (out)
            for (int d = 0; d < 3; ++d) out[d] = cfg[d];
    }

    // solver.rs:758-768
    int64_t run_to(double t_end) {
        int64_t steps = 0;
        while (t < t_end - 1e-10 * std::fabs(t_end)) {
            const double dt = std::fmin(dt_cfl(), t_end - t);
            if (!(std::isfinite(dt) && dt > 1e-12 * std::fmax(std::fabs(t_end), 1e-300)))
                throw Error(EM_ERR_COLLAPSE, "time step collapsed");
            step(dt);
            ++steps;
        }
        return steps;
    }
};

using FeSolver = FeCore<double>;

}  // namespace fe
}  // namespace em

// ------------------------------------------------------------------ C ABI

struct em_fiveeq {
    em::fe::FeCore<double>* s;   // FP64 solver (null for an FP32 handle)
    em::fe::FeCore<float>* s32;  // FP32 solver (null for an FP64 handle)
};

template <class F>
static int fe_guarded(F&& f) {
    try {
        f();
        em::fe::g_err.clear();
        return EM_OK;
    } catch (const em::fe::Error& e) {
        em::fe::g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        em::fe::g_err = e.what();
        return EM_ERR_STATE;
    }
}

#define FENEED(h)                                     \
    if (!(h) || (!(h)->s && !(h)->s32)) {             \
        em::fe::g_err = "null FiveEq handle";         \
        return EM_ERR_ARG;                            \
    }

// f(core) on the handle's solver, whichever real type it holds
template <class F>
static int fe_on(em_fiveeq* h, F&& f) {
    return fe_guarded([&] {
        if (h->s)
            f(*h->s);
        else
            f(*h->s32);
    });
}

template <class Real>
static int fe_make(em_fiveeq** out, const int64_t n[3], const double lo[3], const double hi[3], int32_t ng,
                   const double phases[8], int32_t kapila, const int32_t bc_kind[6], const double bc_state[42],
                   int32_t mode, int32_t positivity, double cfl, double tol, double t0, const double* u) {
    if (!out || !n || !lo || !hi || !phases || !bc_kind || !u) {
        em::fe::g_err = "null argument";
        return EM_ERR_ARG;
    }
    return fe_guarded([&] {
        auto* s = new em::fe::FeCore<Real>(n, lo, hi, ng, phases, kapila, bc_kind, bc_state, mode, positivity, cfl,
                                           tol, t0);
        try {
            s->upload(u);
        } catch (...) {
            delete s;
            throw;
        }
        if constexpr (sizeof(Real) == 8)
            *out = new em_fiveeq{s, nullptr};
        else
            *out = new em_fiveeq{nullptr, s};
    });
}

extern "C" {

const char* em_fiveeq_last_error(void) { return em::fe::g_err.c_str(); }

int em_fiveeq_create(em_fiveeq** out, const int64_t n[3], const double lo[3], const double hi[3], int32_t ng,
                     const double phases[8], int32_t kapila, const int32_t bc_kind[6], const double bc_state[42],
                     int32_t mode, int32_t positivity, double cfl, double tol, double t0, const double* u) {
    return fe_make<double>(out, n, lo, hi, ng, phases, kapila, bc_kind, bc_state, mode, positivity, cfl, tol, t0, u);
}

int em_fiveeq_create_f32(em_fiveeq** out, const int64_t n[3], const double lo[3], const double hi[3], int32_t ng,
                         const double phases[8], int32_t kapila, const int32_t bc_kind[6], const double bc_state[42],
                         int32_t mode, int32_t positivity, double cfl, double tol, double t0, const double* u) {
    return fe_make<float>(out, n, lo, hi, ng, phases, kapila, bc_kind, bc_state, mode, positivity, cfl, tol, t0, u);
}

int em_fiveeq_destroy(em_fiveeq* h) {
    if (!h) return EM_OK;
    delete h->s;
    delete h->s32;
    delete h;
    return EM_OK;
}

int em_fiveeq_upload(em_fiveeq* h, const double* u) {
    FENEED(h);
    return fe_on(h, [&](auto& c) { c.upload(u); });
}

int em_fiveeq_download(em_fiveeq* h, double* u) {
    FENEED(h);
    return fe_on(h, [&](auto& c) { c.to_host(c.U[0], u); });
}

int em_fiveeq_rhs(em_fiveeq* h, double stage_dt, double* rhs, uint64_t stats[4]) {
    FENEED(h);
    return fe_on(h, [&](auto& c) { c.rhs(stage_dt, rhs, stats); });
}

int em_fiveeq_step(em_fiveeq* h, double dt) {
    FENEED(h);
    return fe_on(h, [&](auto& c) { c.step(dt); });
}

int em_fiveeq_dt_cfl(em_fiveeq* h, double* dt) {
    FENEED(h);
    return fe_on(h, [&](auto& c) { *dt = c.dt_cfl(); });
}

int em_fiveeq_run_to(em_fiveeq* h, double t_end, int64_t* steps) {
    FENEED(h);
    return fe_on(h, [&](auto& c) {
        const int64_t k = c.run_to(t_end);
        if (steps) *steps = k;
    });
}

int em_fiveeq_set_tiles(em_fiveeq* h, const int32_t cfg[3]) {
    FENEED(h);
    for (int d = 0; d < 3; ++d)
        if (cfg[d] < 0 || cfg[d] >= em::fe::kNumFeTiles) {
            em::fe::g_err = "tile configuration out of range";
            return EM_ERR_ARG;
        }
    return fe_on(h, [&](auto& c) {
        for (int d = 0; d < 3; ++d) c.cfg[d] = cfg[d];
    });
}

int em_fiveeq_autotune(em_fiveeq* h, int32_t repeats, int32_t cfg_out[3]) {
    FENEED(h);
    if (repeats < 1) {
        em::fe::g_err = "repeats must be >= 1";
        return EM_ERR_ARG;
    }
    return fe_on(h, [&](auto& c) { c.autotune(repeats, cfg_out); });
}

int em_fiveeq_counters(em_fiveeq* h, uint64_t out[3]) {
    FENEED(h);
    return fe_on(h, [&](auto& c) { c.read_counters(out); });
}

int em_fiveeq_time(em_fiveeq* h, double* t) {
    FENEED(h);
    return fe_on(h, [&](auto& c) { *t = c.t; });
}

int em_fiveeq_synchronize(em_fiveeq* h) {
    FENEED(h);
    return fe_on(h, [&](auto& c) { FECHK(hipStreamSynchronize(c.stream)); });
}

}  // extern "C"
