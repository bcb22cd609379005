// Device-side physics of the Euler system (ideal / NASG gas) for the
// gfx950 kernels.  Each function restates the reference formula it cites
// (paths relative to the reference repository root) in the same evaluation
// order, so that a build without FMA contraction reproduces the reference
// CPU arithmetic bit for bit; the contracted build differs by rounding only.
//
// Every function is generic over the real type R (double, or float for the
// FP32 variant, as the reference's `Real` trait, compressible-cuda/src/
// lib.rs:26-48): constants are written R(c), so a float instantiation
// computes in float throughout and the double one is the original
// arithmetic.
//
// Characteristic matrices are used in their sparse form (8 non-zeros in L,
// 9 in R): skipping the exact-zero products of the reference's dense 5x5
// loops changes at most the sign of a zero.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace em {

template <class R>
struct EosT {
    R gamma, p_inf, b, q;
};
using Eos = EosT<double>;

// EOS parameters in another real type (the reference's R::from_f64)
template <class R>
__host__ __device__ inline EosT<R> eos_as(const Eos& e) {
    return EosT<R>{R(e.gamma), R(e.p_inf), R(e.b), R(e.q)};
}

// fluid/crates/compressible/src/eos.rs:36-40
template <class R>
__device__ __forceinline__ R eos_pressure(const EosT<R>& e, R rho, R en) {
    R v = R(1.0) / rho;
    return (e.gamma - R(1.0)) * (en - e.q) / (v - e.b) - e.gamma * e.p_inf;
}
// eos.rs:43-53 (rho_e = rho * energy)
template <class R>
__device__ __forceinline__ R eos_rho_e(const EosT<R>& e, R rho, R p) {
    R v = R(1.0) / rho;
    return rho * ((p + e.gamma * e.p_inf) * (v - e.b) / (e.gamma - R(1.0)) + e.q);
}
// eos.rs:55-59 and system.rs:219-222 (sqrt(max(c^2, 0)))
template <class R>
__device__ __forceinline__ R sound_speed(const EosT<R>& e, R rho, R p) {
    R v = R(1.0) / rho;
    R c2 = e.gamma * (p + e.p_inf) * v * v / (v - e.b);
    return sqrt(fmax(c2, R(0.0)));
}
// system.rs:404-407 with eos.rs:76-79
template <class R>
__device__ __forceinline__ bool admissible(const EosT<R>& e, R rho, R p) {
    return isfinite(rho) && isfinite(p) && rho > R(0.0) && R(1.0) / rho > e.b && p + e.p_inf > R(0.0);
}

// system.rs:197-203
template <class R>
__device__ __forceinline__ void to_prim(const EosT<R>& e, const R u[5], R w[5]) {
    R rho = u[0];
    R vx = u[1] / rho, vy = u[2] / rho, vz = u[3] / rho;
    R en = u[4] / rho - R(0.5) * (vx * vx + vy * vy + vz * vz);
    w[0] = rho;
    w[1] = vx;
    w[2] = vy;
    w[3] = vz;
    w[4] = eos_pressure(e, rho, en);
}
// system.rs:210-217
template <class R>
__device__ __forceinline__ void to_cons(const EosT<R>& e, const R w[5], R u[5]) {
    R rho = w[0];
    R ke = R(0.5) * rho * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
    u[0] = rho;
    u[1] = rho * w[1];
    u[2] = rho * w[2];
    u[3] = rho * w[3];
    u[4] = eos_rho_e(e, rho, w[4]) + ke;
}
// to_cons with the cell's rho_e = eos_rho_e(e, w[0], w[4]) already at hand
// (the same expressions, system.rs:210-217): a cell's rho_e is computed once
// per tile and reused by every face that needs its conserved state.
template <class R>
__device__ __forceinline__ void to_cons_re(const R w[5], R re, R u[5]) {
    R rho = w[0];
    R ke = R(0.5) * rho * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
    u[0] = rho;
    u[1] = rho * w[1];
    u[2] = rho * w[2];
    u[3] = rho * w[3];
    u[4] = re + ke;
}
// Physical flux from (w, q = to_cons(w)): system.rs:336-348 and the closure
// of riemann_flux (system.rs:268-279) compute the same expressions.
template <int D, class R>
__device__ __forceinline__ void phys_flux(const R w[5], const R q[5], R f[5]) {
    constexpr int un = 1 + D;
    R u = w[un];
    f[0] = q[0] * u;
    f[1] = q[1] * u;
    f[2] = q[2] * u;
    f[3] = q[3] * u;
    f[un] += w[4];
    f[4] = (q[4] + w[4]) * u;
}

// HLLC with Davis wave-speed bounds: system.rs:260-312.  Returns the flux.
template <int D, class R>
__device__ __forceinline__ void hllc(const EosT<R>& e, const R wl[5], const R wr[5], R f[5]) {
    constexpr int un = 1 + D;
    R cl = sound_speed(e, wl[0], wl[4]), cr = sound_speed(e, wr[0], wr[4]);
    R ul = wl[un], ur = wr[un];
    R sl = fmin(ul - cl, ur - cr);
    R sr = fmax(ul + cl, ur + cr);
    R ql[5], qr[5];
    to_cons(e, wl, ql);
    to_cons(e, wr, qr);
    if (sl >= R(0.0)) {
        phys_flux<D>(wl, ql, f);
        return;
    }
    if (sr <= R(0.0)) {
        phys_flux<D>(wr, qr, f);
        return;
    }
    R rl = wl[0], rr = wr[0], pl = wl[4], pr = wr[4];
    R ss = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur));
    // one side is selected by the contact speed; its star state and flux
    // (value selects keep both register arrays out of scratch memory)
    const bool left = ss >= R(0.0);
    R w[5], q[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        w[k] = left ? wl[k] : wr[k];
        q[k] = left ? ql[k] : qr[k];
    }
    R s = left ? sl : sr;
    R rho = w[0], u = w[un];
    R fac = rho * (s - u) / (s - ss);
    R qs[5];
    qs[0] = fac;
    qs[1] = fac * w[1];
    qs[2] = fac * w[2];
    qs[3] = fac * w[3];
    qs[un] = fac * ss;
    qs[4] = fac * (q[4] / rho + (ss - u) * (ss + w[4] / (rho * (s - u))));
    R fw[5];
    phys_flux<D>(w, q, fw);
#pragma unroll
    for (int k = 0; k < 5; ++k) f[k] = fw[k] + s * (qs[k] - q[k]);
}

// KEEP-PE two-point flux: system.rs:317-334
template <int D, class R>
__device__ __forceinline__ void split_flux(const EosT<R>& e, const R a[5], const R b[5], R f[5]) {
    constexpr int un = 1 + D;
    R rho = R(0.5) * (a[0] + b[0]);
    R u = R(0.5) * (a[un] + b[un]);
    R c = rho * u;
    f[0] = c;
    f[1] = c * R(0.5) * (a[1] + b[1]);
    f[2] = c * R(0.5) * (a[2] + b[2]);
    f[3] = c * R(0.5) * (a[3] + b[3]);
    f[un] += R(0.5) * (a[4] + b[4]);
    R kin = c * R(0.5) * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
    R rhoe = R(0.5) * (eos_rho_e(e, a[0], a[4]) + eos_rho_e(e, b[0], b[4]));
    R pw = R(0.5) * (a[un] * b[4] + b[un] * a[4]);
    f[4] = kin + u * rhoe + pw;
}

// split_flux with the two cells' rho_e given (each computed once per tile)
template <int D, class R>
__device__ __forceinline__ void split_flux_re(const R a[5], const R b[5], R rea, R reb,
                                              R f[5]) {
    constexpr int un = 1 + D;
    R rho = R(0.5) * (a[0] + b[0]);
    R u = R(0.5) * (a[un] + b[un]);
    R c = rho * u;
    f[0] = c;
    f[1] = c * R(0.5) * (a[1] + b[1]);
    f[2] = c * R(0.5) * (a[2] + b[2]);
    f[3] = c * R(0.5) * (a[3] + b[3]);
    f[un] += R(0.5) * (a[4] + b[4]);
    R kin = c * R(0.5) * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
    R rhoe = R(0.5) * (rea + reb);
    R pw = R(0.5) * (a[un] * b[4] + b[un] * a[4]);
    f[4] = kin + u * rhoe + pw;
}

// WENO5-Z interpolation of point values to the midpoint x_{2+1/2}:
// scheme.rs:69-86 (EPS = 1e-40, scheme.rs:55).  Division by 8 is an exact
// power-of-two scaling, written as a multiplication.  In FP32 the epsilon is
// 1e-30, as in the reference's f32 kernels (compressible-cuda/src/
// kernels.cu:263): with 1e-40 (a float denormal), a stencil with one exactly
// flat sub-stencil squares tau / 1e-40 past FLT_MAX, and the midpoint falls
// back to first order.
template <class R>
__device__ __forceinline__ R weno5z_interp(R v0, R v1, R v2, R v3, R v4) {
    R q0 = (R(3.0) * v0 - R(10.0) * v1 + R(15.0) * v2) * R(0.125);
    R q1 = (-v1 + R(6.0) * v2 + R(3.0) * v3) * R(0.125);
    R q2 = (R(3.0) * v2 + R(6.0) * v3 - v4) * R(0.125);
    R t0 = v0 - R(2.0) * v1 + v2, s0 = v0 - R(4.0) * v1 + R(3.0) * v2;
    R t1 = v1 - R(2.0) * v2 + v3, s1 = v1 - v3;
    R t2 = v2 - R(2.0) * v3 + v4, s2 = R(3.0) * v2 - R(4.0) * v3 + v4;
    const R c13 = R(13.0) / R(12.0);
    R b0 = c13 * (t0 * t0) + R(0.25) * (s0 * s0);
    R b1 = c13 * (t1 * t1) + R(0.25) * (s1 * s1);
    R b2 = c13 * (t2 * t2) + R(0.25) * (s2 * s2);
    R tau = fabs(b0 - b2);
    const R eps = sizeof(R) == 4 ? R(1e-30) : R(1e-40);
    R r0 = tau / (b0 + eps), r1 = tau / (b1 + eps), r2 = tau / (b2 + eps);
    R a0 = R(0.0625) * (R(1.0) + r0 * r0);
    R a1 = R(0.625) * (R(1.0) + r1 * r1);
    R a2 = R(0.3125) * (R(1.0) + r2 * r2);
    return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}

// WCNS midpoint flux: characteristic WENO5-Z interpolation of the six
// primitive states s[0..5] = w_{j-2..j+3} to the midpoint j+1/2, first-order
// fallback on an inadmissible state, HLLC there.  scheme.rs:286-333, with the
// eigenvectors of system.rs:232-258 applied in sparse form.
// Returns true on the first-order fallback.
template <int D, class R>
__device__ __forceinline__ bool midpoint_flux(const EosT<R>& e, const R (&s)[6][5], R f[5]) {
    constexpr int un = 1 + D;
    constexpr int t1 = 1 + (D + 1) % 3;
    constexpr int t2 = 1 + (D + 2) % 3;
    R rho = R(0.5) * (s[2][0] + s[3][0]);
    R avun = R(0.5) * (s[2][un] + s[3][un]);
    R avp = R(0.5) * (s[2][4] + s[3][4]);
    (void)avun;
    R c = sound_speed(e, rho, avp);
    // L rows (waves u-c, entropy, shear t1, shear t2, u+c)
    R l0un = -rho / (R(2.0) * c);
    R l04 = R(1.0) / (R(2.0) * c * c);
    R l14 = -R(1.0) / (c * c);
    R l4un = rho / (R(2.0) * c);
    R l44 = R(1.0) / (R(2.0) * c * c);
    R ch[6][5];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        ch[k][0] = l0un * s[k][un] + l04 * s[k][4];
        ch[k][1] = s[k][0] + l14 * s[k][4];
        ch[k][2] = s[k][t1];
        ch[k][3] = s[k][t2];
        ch[k][4] = l4un * s[k][un] + l44 * s[k][4];
    }
    R cl[5], cr[5];
#pragma unroll
    for (int a = 0; a < 5; ++a) {
        cl[a] = weno5z_interp(ch[0][a], ch[1][a], ch[2][a], ch[3][a], ch[4][a]);
        cr[a] = weno5z_interp(ch[5][a], ch[4][a], ch[3][a], ch[2][a], ch[1][a]);
    }
    // R columns: r[0][0]=r[0][1]=r[0][4]=1, r[un][0]=-c/rho, r[un][4]=c/rho,
    // r[t1][2]=r[t2][3]=1, r[4][0]=r[4][4]=c*c
    R run0 = -c / rho, run4 = c / rho, r40 = c * c;
    R wl[5], wr[5];
    wl[0] = cl[0] + cl[1] + cl[4];
    wr[0] = cr[0] + cr[1] + cr[4];
    wl[un] = run0 * cl[0] + run4 * cl[4];
    wr[un] = run0 * cr[0] + run4 * cr[4];
    wl[t1] = cl[2];
    wr[t1] = cr[2];
    wl[t2] = cl[3];
    wr[t2] = cr[3];
    wl[4] = r40 * cl[0] + r40 * cl[4];
    wr[4] = r40 * cr[0] + r40 * cr[4];
    bool fb = false;
    if (!admissible(e, wl[0], wl[4]) || !admissible(e, wr[0], wr[4])) {
        fb = true;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            wl[k] = s[2][k];
            wr[k] = s[3][k];
        }
    }
    hllc<D>(e, wl, wr, f);
    return fb;
}

// Flux-form weights of the 6th-order midpoint-to-node difference:
// scheme.rs:92
__device__ constexpr double S6_0 = 75.0 / 64.0 - 25.0 / 384.0 + 3.0 / 640.0;
__device__ constexpr double S6_1 = -25.0 / 384.0 + 3.0 / 640.0;
__device__ constexpr double S6_2 = 3.0 / 640.0;
// 6th-order split-form coefficients: scheme.rs:89
__device__ constexpr double A6_0 = 0.75;
__device__ constexpr double A6_1 = -0.15;
__device__ constexpr double A6_2 = 1.0 / 60.0;

// Fourth difference of a 5-value window: scheme.rs:97-99
template <class R>
__device__ __forceinline__ R d4(R q0, R q1, R q2, R q3, R q4) {
    return fabs(q0 - R(4.0) * q1 + R(6.0) * q2 - R(4.0) * q3 + q4);
}

// Conserved ghost state of the characteristic (non-reflecting) boundary:
// boundary.rs:57-95, for side (0 low, 1 high), boundary cell u_b, its
// interior neighbour u_b1, far-field u_far, layer depth m.
template <int D, class R>
__device__ __forceinline__ void characteristic_ghost(const EosT<R>& e, const R ub[5], const R ub1[5],
                                                     const R uf[5], int side, R m, R out[5]) {
    constexpr int un = 1 + D;
    constexpr int t1 = 1 + (D + 1) % 3;
    constexpr int t2 = 1 + (D + 2) % 3;
    R wi[5], wf[5], w1[5];
    to_prim(e, ub, wi);
    to_prim(e, uf, wf);
    to_prim(e, ub1, w1);
    R rho = wi[0], c = sound_speed(e, wi[0], wi[4]);
    R L[5][5], Rm[5][5];
#pragma unroll
    for (int a = 0; a < 5; ++a)
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            L[a][k] = R(0.0);
            Rm[a][k] = R(0.0);
        }
    L[0][un] = -rho / (R(2.0) * c);
    L[0][4] = R(1.0) / (R(2.0) * c * c);
    L[1][0] = R(1.0);
    L[1][4] = -R(1.0) / (c * c);
    L[2][t1] = R(1.0);
    L[3][t2] = R(1.0);
    L[4][un] = rho / (R(2.0) * c);
    L[4][4] = R(1.0) / (R(2.0) * c * c);
    Rm[0][0] = R(1.0);
    Rm[un][0] = -c / rho;
    Rm[4][0] = c * c;
    Rm[0][1] = R(1.0);
    Rm[t1][2] = R(1.0);
    Rm[t2][3] = R(1.0);
    Rm[0][4] = R(1.0);
    Rm[un][4] = c / rho;
    Rm[4][4] = c * c;
    R u = wi[un];
    R lam[5] = {u - c, u, u, u, u + c};
    R outward = side == 1 ? R(1.0) : -R(1.0);
    R dw[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) dw[k] = wi[k] - wf[k] + m * (wi[k] - w1[k]);
    R wg[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) wg[k] = wf[k];
#pragma unroll
    for (int kf = 0; kf < 5; ++kf) {
        if (lam[kf] * outward > R(0.0)) {
            R amp = -R(0.0);
#pragma unroll
            for (int q = 0; q < 5; ++q) amp += L[kf][q] * dw[q];
#pragma unroll
            for (int q = 0; q < 5; ++q) wg[q] += Rm[q][kf] * amp;
        }
    }
    to_cons(e, wg, out);
}

// Conserved state of the ghost cell at pencil coordinate g (outside [0, n))
// of a pencil whose interior cell x is read by ld(x, out).  Follows
// Solver::fill_ghosts (solver.rs:278-348) and reflected_ghost
// (boundary.rs:102-145): periodic wraps (also around axes shorter than the
// ghost depth), outflow copies the boundary cell, Dirichlet gives the fixed
// state, characteristic extrapolates outgoing waves, reflect mirrors with the
// normal momentum negated (and resolves deep images through the opposite
// side's condition).  Every ghost depends only on its own pencil, so the
// value is the same as after the reference's axis-ordered sweeps.
// Kinds: 0 periodic, 1 outflow, 2 reflect, 3 Dirichlet, 4 characteristic.
template <int D, class R, class LD>
__device__ __forceinline__ void bc_ghost(const EosT<R>& e, int n, int g, const int kind[2], const R (*state)[5],
                                         LD ld, R u[5]) {
    int x = g;
    bool flip = false;
    for (int it = 0; it < 64; ++it) {
        if (x >= 0 && x < n) {
            ld(x, u);
            break;
        }
        const int side = x >= n ? 1 : 0;
        const int k = kind[side];
        if (k == 2) {
            x = side == 0 ? -1 - x : 2 * n - 1 - x;
            flip = !flip;
            continue;
        }
        if (k == 0) {
            ld(((x % n) + n) % n, u);
        } else if (k == 1) {
            ld(side == 0 ? 0 : n - 1, u);
        } else if (k == 3) {
#pragma unroll
            for (int q = 0; q < 5; ++q) u[q] = state[side][q];
        } else {
            R ub[5], ub1[5];
            ld(side == 0 ? 0 : n - 1, ub);
            ld(side == 0 ? 1 : n - 2, ub1);
            characteristic_ghost<D>(e, ub, ub1, state[side], side, side == 0 ? R(-x) : R(x - n + 1), u);
        }
        break;
    }
    if (flip) u[1 + D] *= -R(1.0);
}

}  // namespace em
