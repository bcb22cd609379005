// Device-side physics of the two-phase Kapila five-equation system (FiveEq,
// seven variables) for the gfx950 kernels.  Each function restates the
// reference formula it cites, fluid/crates/compressible/src/multiphase.rs,
// in the same evaluation order; built without FMA contraction it reproduces
// the reference CPU arithmetic bit for bit.
//
// Conserved (a1, a2, rho u, rho v, rho w, E, alpha1); primitive
// (a1, a2, u, v, w, p, alpha1) with a_k = alpha_k rho_k (multiphase.rs:37-40).
// Each phase has a Noble-Abel stiffened-gas EOS (gamma, p_inf, b, q).
//
// The WCNS midpoint uses the characteristic matrices in sparse form (13
// non-zeros in L, 17 in R): skipping the exact-zero products of the
// reference's dense 7x7 loops changes at most the sign of a zero.  The
// characteristic boundary keeps the dense form (it is off the hot path).
#pragma once
#include <hip/hip_runtime.h>

namespace em {
namespace fe {

constexpr int NV = 7;
constexpr int IA = 6, IP = 5, IU = 2;

template <class R>
struct PhaseT {
    R gamma, p_inf, b, q;
};
template <class R>
struct SysT {
    PhaseT<R> e[2];
    int kapila;
};
using Phase = PhaseT<double>;
using Sys = SysT<double>;

// multiphase.rs:63-75: D, B of rho e = p D + B
template <class R>
__device__ __forceinline__ void db(const SysT<R>& s, const R* w, R& d, R& b) {
    const R al0 = w[IA], al1 = R(1.0) - w[IA];
    d = R(0.0);
    b = R(0.0);
    {
        const PhaseT<R>& e = s.e[0];
        const R vol = al0 - e.b * w[0];
        d += vol / (e.gamma - R(1.0));
        b += e.gamma * e.p_inf * vol / (e.gamma - R(1.0)) + e.q * w[0];
    }
    {
        const PhaseT<R>& e = s.e[1];
        const R vol = al1 - e.b * w[1];
        d += vol / (e.gamma - R(1.0));
        b += e.gamma * e.p_inf * vol / (e.gamma - R(1.0)) + e.q * w[1];
    }
}
// 77-81
template <class R>
__device__ __forceinline__ R rho_e(const SysT<R>& s, const R* w) {
    R d, b;
    db(s, w, d, b);
    return w[IP] * d + b;
}
// 83-90: phase rho_k c_k^2
template <class R>
__device__ __forceinline__ R rho_c2_phase(const SysT<R>& s, const R* w, int k) {
    const R al = k == 0 ? w[IA] : R(1.0) - w[IA];
    const PhaseT<R>& e = s.e[k];
    const R rho_k = (al > R(1e-12) && w[k] > R(0.0)) ? w[k] / al : R(0.0);
    return e.gamma * (w[IP] + e.p_inf) / fmax(R(1.0) - e.b * rho_k, R(1e-12));
}
// f64::clamp (a NaN stays NaN)
template <class R>
__device__ __forceinline__ R clamp01(R x) {
    if (x < R(0.0)) x = R(0.0);
    if (x > R(1.0)) x = R(1.0);
    return x;
}
// 92-104: Kapila coefficient K
template <class R>
__device__ __forceinline__ R kapila_k(const SysT<R>& s, const R* w) {
    if (!s.kapila) return R(0.0);
    const R a = clamp01(w[IA]);
    const R al0 = a, al1 = R(1.0) - a;
    const R z1 = rho_c2_phase(s, w, 0), z2 = rho_c2_phase(s, w, 1);
    const R den = al1 * z1 + al0 * z2;
    if (den <= R(0.0)) return R(0.0);
    return al0 * al1 * (z2 - z1) / den;
}
// 120-124
template <class R>
__device__ __forceinline__ R drhoe_dalpha(const SysT<R>& s, R p) {
    const PhaseT<R> &e1 = s.e[0], &e2 = s.e[1];
    return (p + e1.gamma * e1.p_inf) / (e1.gamma - R(1.0)) - (p + e2.gamma * e2.p_inf) / (e2.gamma - R(1.0));
}
// 106-118: mixture rho c^2 (Wood's with K)
template <class R>
__device__ __forceinline__ R rho_c2(const SysT<R>& s, const R* w) {
    R d, b;
    db(s, w, d, b);
    const R p = w[IP];
    R num = (w[IP] * d + b) + p;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const PhaseT<R>& e = s.e[k];
        num += w[k] * (e.b * (p + e.gamma * e.p_inf) / (e.gamma - R(1.0)) - e.q);
    }
    num += kapila_k(s, w) * drhoe_dalpha(s, p);
    return num / d;
}
template <class R>
__device__ __forceinline__ R mix_rho(const R* w) { return w[0] + w[1]; }  // 126-130
// 175-177
template <class R>
__device__ __forceinline__ R sound_speed(const SysT<R>& s, const R* w) {
    return sqrt(fmax(rho_c2(s, w), R(0.0)) / mix_rho(w));
}
// 132-136
template <class R>
__device__ __forceinline__ R p_inf_mix(const SysT<R>& s, const R* w) {
    const R a = clamp01(w[IA]);
    return a * s.e[0].p_inf + (R(1.0) - a) * s.e[1].p_inf;
}

// 140-156
template <class R>
__device__ __forceinline__ void to_prim(const SysT<R>& s, const R* u, R* w) {
    const R r = u[0] + u[1];
    const R vx = u[2] / r, vy = u[3] / r, vz = u[4] / r;
    w[0] = u[0];
    w[1] = u[1];
    w[2] = vx;
    w[3] = vy;
    w[4] = vz;
    w[IA] = u[6];
    const R rhoe = u[5] - R(0.5) * r * (vx * vx + vy * vy + vz * vz);
    R d, b;
    db(s, w, d, b);
    w[IP] = (rhoe - b) / d;
}
// 162-173
template <class R>
__device__ __forceinline__ void to_cons(const SysT<R>& s, const R* w, R* u) {
    const R r = w[0] + w[1];
    u[0] = w[0];
    u[1] = w[1];
    u[2] = r * w[2];
    u[3] = r * w[3];
    u[4] = r * w[4];
    u[5] = rho_e(s, w) + R(0.5) * r * (w[2] * w[2] + w[3] * w[3] + w[4] * w[4]);
    u[6] = w[IA];
}
// to_cons with rho e(w) already evaluated (the same value rho_e returns)
template <class R>
__device__ __forceinline__ void to_cons_re(const R* w, R re, R* u) {
    const R r = w[0] + w[1];
    u[0] = w[0];
    u[1] = w[1];
    u[2] = r * w[2];
    u[3] = r * w[3];
    u[4] = r * w[4];
    u[5] = re + R(0.5) * r * (w[2] * w[2] + w[3] * w[3] + w[4] * w[4]);
    u[6] = w[IA];
}
// 319-331, from w and q = to_cons(w)
template <int D, class R>
__device__ __forceinline__ void phys_flux(const R* w, const R* q, R* f) {
    constexpr int un = IU + D;
    const R u = w[un];
#pragma unroll
    for (int k = 0; k < 5; ++k) f[k] = q[k] * u;
    f[un] += w[IP];
    f[5] = (q[5] + w[IP]) * u;
    f[6] = w[IA] * u;
}
// 358-363
template <class R>
__device__ __forceinline__ bool admissible(const SysT<R>& s, const R* w) {
#pragma unroll
    for (int k = 0; k < 7; ++k)
        if (!isfinite(w[k])) return false;
    return mix_rho(w) > R(0.0) && (w[IA] >= -R(0.05) && w[IA] <= R(1.05)) && rho_c2(s, w) > R(0.0);
}
// 369-371
template <class R>
__device__ __forceinline__ R nc_coef(const SysT<R>& s, const R* w) { return w[IA] + kapila_k(s, w); }
// 373-391
template <class R>
__device__ __forceinline__ R dilatation_rate(const SysT<R>& s, const R* w, R divu) {
    const R a = w[IA];
    R pmin_inf = R(__builtin_inf());
    if (a > R(1e-6)) pmin_inf = fmin(pmin_inf, s.e[0].p_inf);
    if (R(1.0) - a > R(1e-6)) pmin_inf = fmin(pmin_inf, s.e[1].p_inf);
    const R floor_ = R(1e-6) * fmax(s.e[0].p_inf, s.e[1].p_inf);
    const R head = fmax(fmax(w[IP] + pmin_inf, floor_), R(1e-300));
    return fmax(rho_c2(s, w), R(0.0)) * fabs(divu) / head;
}

// HLLC with Davis speeds: 265-310.  Returns the flux; us = contact speed
// (or the upwind velocity in the supersonic branches).
template <int D, class R>
__device__ __forceinline__ void hllc(const SysT<R>& s, const R* wl, const R* wr, R* f, R& us) {
    constexpr int un = IU + D;
    const R cl = sound_speed(s, wl), cr = sound_speed(s, wr);
    const R ul = wl[un], ur = wr[un];
    const R sl = fmin(ul - cl, ur - cr);
    const R sr = fmax(ul + cl, ur + cr);
    R ql[7], qr[7];
    to_cons(s, wl, ql);
    to_cons(s, wr, qr);
    if (sl >= R(0.0)) {
        phys_flux<D>(wl, ql, f);
        us = ul;
        return;
    }
    if (sr <= R(0.0)) {
        phys_flux<D>(wr, qr, f);
        us = ur;
        return;
    }
    const R rl = mix_rho(wl), rr = mix_rho(wr);
    const R pl = wl[IP], pr = wr[IP];
    const R ss = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur));
    // the side selected by the contact speed (value selects keep both
    // register arrays out of scratch memory)
    const bool left = ss >= R(0.0);
    R w[7], q[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        w[k] = left ? wl[k] : wr[k];
        q[k] = left ? ql[k] : qr[k];
    }
    const R sw = left ? sl : sr;
    const R rho = mix_rho(w), u = w[un];
    const R fac = (sw - u) / (sw - ss);
    R qs[7];
    qs[0] = w[0] * fac;
    qs[1] = w[1] * fac;
#pragma unroll
    for (int k = IU; k < IU + 3; ++k) qs[k] = rho * fac * w[k];
    qs[un] = rho * fac * ss;
    qs[5] = rho * fac * (q[5] / rho + (ss - u) * (ss + w[IP] / (rho * (sw - u))));
    qs[6] = w[IA] * fac;
    R fw[7];
    phys_flux<D>(w, q, fw);
#pragma unroll
    for (int k = 0; k < 7; ++k) f[k] = fw[k] + sw * (qs[k] - q[k]);
    us = ss;
}

// KEEP-PE pair flux: 312-331, with rho e of the two states given
template <int D, class R>
__device__ __forceinline__ void split_flux_re(const R* a, const R* b, R rea, R reb, R* f) {
    constexpr int un = IU + D;
    const R u = R(0.5) * (a[un] + b[un]);
    const R rho = R(0.5) * (mix_rho(a) + mix_rho(b));
    const R c = rho * u;
    f[0] = R(0.5) * (a[0] + b[0]) * u;
    f[1] = R(0.5) * (a[1] + b[1]) * u;
#pragma unroll
    for (int k = IU; k < IU + 3; ++k) f[k] = c * R(0.5) * (a[k] + b[k]);
    f[un] += R(0.5) * (a[IP] + b[IP]);
    const R kin = c * R(0.5) * (a[2] * b[2] + a[3] * b[3] + a[4] * b[4]);
    const R rhoe = R(0.5) * (rea + reb);
    const R pw = R(0.5) * (a[un] * b[IP] + b[un] * a[IP]);
    f[5] = kin + u * rhoe + pw;
    f[6] = R(0.5) * (a[IA] + b[IA]) * u;
}

// Fourth difference of a 5-value window: scheme.rs:97-99
template <class R>
__device__ __forceinline__ R d4f(R q0, R q1, R q2, R q3, R q4) {
    return fabs(q0 - R(4.0) * q1 + R(6.0) * q2 - R(4.0) * q3 + q4);
}

// WENO5-Z interpolation (scheme.rs:69-86); same arithmetic as the Euler
// path's weno5z_interp
template <class R>
__device__ __forceinline__ R weno5z(R v0, R v1, R v2, R v3, R v4) {
    const R q0 = (R(3.0) * v0 - R(10.0) * v1 + R(15.0) * v2) * R(0.125);
    const R q1 = (-v1 + R(6.0) * v2 + R(3.0) * v3) * R(0.125);
    const R q2 = (R(3.0) * v2 + R(6.0) * v3 - v4) * R(0.125);
    const R t0 = v0 - R(2.0) * v1 + v2, s0 = v0 - R(4.0) * v1 + R(3.0) * v2;
    const R t1 = v1 - R(2.0) * v2 + v3, s1 = v1 - v3;
    const R t2 = v2 - R(2.0) * v3 + v4, s2 = R(3.0) * v2 - R(4.0) * v3 + v4;
    const R c13 = R(13.0) / R(12.0);
    const R b0 = c13 * (t0 * t0) + R(0.25) * (s0 * s0);
    const R b1 = c13 * (t1 * t1) + R(0.25) * (s1 * s1);
    const R b2 = c13 * (t2 * t2) + R(0.25) * (s2 * s2);
    const R tau = fabs(b0 - b2);
    const R eps = sizeof(R) == 4 ? R(1e-30) : R(1e-40);  // the reference's float epsilon, kernels.cu:263
    const R r0 = tau / (b0 + eps), r1 = tau / (b1 + eps), r2 = tau / (b2 + eps);
    const R a0 = R(0.0625) * (R(1.0) + r0 * r0);
    const R a1 = R(0.625) * (R(1.0) + r1 * r1);
    const R a2 = R(0.3125) * (R(1.0) + r2 * r2);
    return (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}

// WCNS midpoint j+1/2 from the six primitive states w_{j-2..j+3}, read
// through ld(j, k) (component k of stencil state j, j = 0..5): scheme.rs:
// 286-333 with FiveEq::eigen (multiphase.rs:219-256) in sparse form.
// Returns true on the first-order fallback; f = HLLC flux, us its contact
// speed.  The characteristic fields are taken one at a time, each stencil
// value read when it is needed and each interpolated value folded into the
// primitive states at once (in the reference's column order), so that only
// one field's stencil is live: this keeps the pass within 256 registers.
template <int D, class R, class LD>
__device__ __forceinline__ bool midpoint_flux(const SysT<R>& s, LD&& ld, R* f, R& us) {
    constexpr int un = IU + D;
    constexpr int t1 = IU + (D + 1) % 3;
    constexpr int t2 = IU + (D + 2) % 3;
    R avg[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) avg[k] = R(0.5) * (ld(2, k) + ld(3, k));
    const R rho = mix_rho(avg);
    const R c = sound_speed(s, avg);
    const R kk = kapila_k(s, avg);
    const R c2 = c * c;
    // L rows (fields a1, a2, u-c, t1, t2, u+c, alpha)
    const R l0p = -avg[0] / rho / c2;
    const R l1p = -avg[1] / rho / c2;
    const R l2u = -rho / (R(2.0) * c);
    const R l2p = R(1.0) / (R(2.0) * c2);
    const R l5u = rho / (R(2.0) * c);
    const R l5p = R(1.0) / (R(2.0) * c2);
    const R l6p = kk / rho / c2;
    // R: r[0][0]=r[1][1]=r[t1][3]=r[t2][4]=r[IA][6]=1; columns 2 (sign -1)
    // and 5 (sign +1): r[0]=a1/rho, r[1]=a2/rho, r[un]=sgn c/rho, r[IP]=c^2,
    // r[IA]=-K/rho
    const R r0 = avg[0] / rho, r1 = avg[1] / rho;
    const R rum = -R(1.0) * c / rho, rup = R(1.0) * c / rho;
    const R ra = -kk / rho;
    // w = R c, each row summed over columns a = 0..6 from 0.0 (zero terms
    // skipped)
    R wl[7], wr[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        wl[k] = R(0.0);
        wr[k] = R(0.0);
    }
#pragma unroll
    for (int a = 0; a < 7; ++a) {
        // c_a = L[a] . w_j, summed over columns k = 0..6 from 0.0
        R ch[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            if (a == 0) ch[j] = R(0.0) + R(1.0) * ld(j, 0) + l0p * ld(j, IP);
            if (a == 1) ch[j] = R(0.0) + R(1.0) * ld(j, 1) + l1p * ld(j, IP);
            if (a == 2) ch[j] = R(0.0) + l2u * ld(j, un) + l2p * ld(j, IP);
            if (a == 3) ch[j] = R(0.0) + R(1.0) * ld(j, t1);
            if (a == 4) ch[j] = R(0.0) + R(1.0) * ld(j, t2);
            if (a == 5) ch[j] = R(0.0) + l5u * ld(j, un) + l5p * ld(j, IP);
            if (a == 6) ch[j] = R(0.0) + l6p * ld(j, IP) + R(1.0) * ld(j, IA);
        }
        const R cl = weno5z(ch[0], ch[1], ch[2], ch[3], ch[4]);
        const R cr = weno5z(ch[5], ch[4], ch[3], ch[2], ch[1]);
        if (a == 0) {
            wl[0] = wl[0] + R(1.0) * cl;
            wr[0] = wr[0] + R(1.0) * cr;
        }
        if (a == 1) {
            wl[1] = wl[1] + R(1.0) * cl;
            wr[1] = wr[1] + R(1.0) * cr;
        }
        if (a == 2 || a == 5) {
            const R ru = a == 2 ? rum : rup;
            wl[0] = wl[0] + r0 * cl;
            wr[0] = wr[0] + r0 * cr;
            wl[1] = wl[1] + r1 * cl;
            wr[1] = wr[1] + r1 * cr;
            wl[un] = wl[un] + ru * cl;
            wr[un] = wr[un] + ru * cr;
            wl[IP] = wl[IP] + c2 * cl;
            wr[IP] = wr[IP] + c2 * cr;
            wl[IA] = wl[IA] + ra * cl;
            wr[IA] = wr[IA] + ra * cr;
        }
        if (a == 3) {
            wl[t1] = wl[t1] + R(1.0) * cl;
            wr[t1] = wr[t1] + R(1.0) * cr;
        }
        if (a == 4) {
            wl[t2] = wl[t2] + R(1.0) * cl;
            wr[t2] = wr[t2] + R(1.0) * cr;
        }
        if (a == 6) {
            wl[IA] = wl[IA] + R(1.0) * cl;
            wr[IA] = wr[IA] + R(1.0) * cr;
        }
    }
    bool fb = false;
    if (!admissible(s, wl) || !admissible(s, wr)) {
        fb = true;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            wl[k] = ld(2, k);
            wr[k] = ld(3, k);
        }
    }
    hllc<D>(s, wl, wr, f, us);
    return fb;
}

// Dense eigenvectors of FiveEq (multiphase.rs:219-256) at w along D
template <int D, class R>
__device__ __forceinline__ void eigen_dense(const SysT<R>& s, const R* w, R (&l)[7][7], R (&r)[7][7]) {
    constexpr int un = IU + D;
    constexpr int t1 = IU + (D + 1) % 3;
    constexpr int t2 = IU + (D + 2) % 3;
    const R rho = mix_rho(w);
    const R c = sound_speed(s, w);
    const R kk = kapila_k(s, w);
#pragma unroll
    for (int a = 0; a < 7; ++a)
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            l[a][k] = R(0.0);
            r[a][k] = R(0.0);
        }
    const R c2 = c * c;
    l[0][0] = R(1.0);
    l[0][IP] = -w[0] / rho / c2;
    l[1][1] = R(1.0);
    l[1][IP] = -w[1] / rho / c2;
    l[2][un] = -rho / (R(2.0) * c);
    l[2][IP] = R(1.0) / (R(2.0) * c2);
    l[3][t1] = R(1.0);
    l[4][t2] = R(1.0);
    l[5][un] = rho / (R(2.0) * c);
    l[5][IP] = R(1.0) / (R(2.0) * c2);
    l[6][IA] = R(1.0);
    l[6][IP] = kk / rho / c2;
    r[0][0] = R(1.0);
    r[1][1] = R(1.0);
    r[t1][3] = R(1.0);
    r[t2][4] = R(1.0);
    r[IA][6] = R(1.0);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int col = q == 0 ? 2 : 5;
        const R sgn = q == 0 ? -R(1.0) : R(1.0);
        r[0][col] = w[0] / rho;
        r[1][col] = w[1] / rho;
        r[un][col] = sgn * c / rho;
        r[IP][col] = c2;
        r[IA][col] = -kk / rho;
    }
}

// Characteristic (non-reflecting) ghost: boundary.rs:57-95, dense form
template <int D, class R>
__device__ __forceinline__ void characteristic_ghost(const SysT<R>& s, const R* ub, const R* ub1,
                                                     const R* uf, int side, R m, R* out) {
    R wi[7], wf[7], w1[7];
    to_prim(s, ub, wi);
    to_prim(s, uf, wf);
    to_prim(s, ub1, w1);
    R L[7][7], Rm[7][7];
    eigen_dense<D>(s, wi, L, Rm);
    const R u = wi[IU + D], c = sound_speed(s, wi);
    const R lam[7] = {u, u, u - c, u, u, u + c, u};  // multiphase.rs:258-263
    const R outward = side == 1 ? R(1.0) : -R(1.0);
    R dw[7], wg[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        dw[k] = wi[k] - wf[k] + m * (wi[k] - w1[k]);
        wg[k] = wf[k];
    }
    for (int kf = 0; kf < 7; ++kf) {
        if (lam[kf] * outward > R(0.0)) {
            R amp = -R(0.0);
#pragma unroll
            for (int q = 0; q < 7; ++q) amp += L[kf][q] * dw[q];
#pragma unroll
            for (int q = 0; q < 7; ++q) wg[q] += Rm[q][kf] * amp;
        }
    }
    to_cons(s, wg, out);
}

// Conserved state of the ghost at pencil coordinate g (outside [0, n)) of a
// pencil whose interior cell x is read by ld(x, out): the ghost fill of
// solver.rs:278-348 and boundary.rs:102-145 (the normal momentum is
// conserved index 2 + D).  Kinds: 0 periodic, 1 outflow, 2 reflect,
// 3 Dirichlet, 4 characteristic.
template <int D, class R, class LD>
__device__ __forceinline__ void bc_ghost(const SysT<R>& s, int n, int g, const int kind[2], const R (*state)[7],
                                         LD ld, R* u) {
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
            for (int q = 0; q < 7; ++q) u[q] = state[side][q];
        } else {
            R ub[7], ub1[7];
            ld(side == 0 ? 0 : n - 1, ub);
            ld(side == 0 ? 1 : n - 2, ub1);
            characteristic_ghost<D>(s, ub, ub1, state[side], side, side == 0 ? R(-x) : R(x - n + 1), u);
        }
        break;
    }
    if (flip) u[IU + D] *= -R(1.0);
}

}  // namespace fe
}  // namespace em
