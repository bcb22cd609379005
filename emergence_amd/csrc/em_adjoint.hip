// Hand-written discrete adjoint of the Weno-mode residual (WCNS midpoints:
// characteristic projection at the face-average state, WENO5-Z
// interpolation, back-projection, first-order fallback, HLLC; S6
// midpoint-to-node combination; positivity limiter; divergence) for the
// uniform-grid solver.  The reference has no adjoint; the conventions are
// those of the tape oracle it is checked against:
//   * branches follow the forward evaluation (HLLC case, fallback, limiter
//     trigger), min/max differentiate the operand they return (the first on
//     a tie), d|x|/dx = sign(x) with sign(0) = 0;
//   * the limiter's blend weight theta follows the admissibility boundary
//     its bisection resolves (implicit derivative at the constraint that
//     bounds it), its Lax-Friedrichs speed the max of the two cells';
//   * dt is given (frozen), so the CFL reduction is not differentiated;
//   * a WENO-Z stencil flat to rounding takes the linear weights' derivative,
//     a partly flat one holds its weights, and the weight derivative is
//     formed without cancellation (see weno_vjp).
// Forward values are recomputed with the forward pass's own device
// functions and operation order (no FMA contraction), so every branch
// decision matches the forward run bit for bit.
//
// One workgroup (256 threads, three per CU) owns one tile of a pencil along
// the pass axis (at most kAdjTileCells interior cells, with the ghost
// positions of a non-periodic edge) and the halo its cotangents need, in LDS:
//   A   load U, primitives (boundary ghosts through bc_ghost);
//   B1  WENO5-Z reconstructions per (midpoint, characteristic field);
//   B2  back-projection, fallback and HLLC per midpoint: forward fluxes;
//   C   face fluxes and limiter state, face cotangents
//       Fbar_f = (lam_f - lam_{f-1}) / dx, the limiter's Lax-Friedrichs part
//       and blend weight scattered to its two cells;
//   D2  per midpoint: cotangent from its faces (S6 weights), HLLC and
//       back-projection VJPs (the reconstructions of B1 are reused);
//   D3  per (midpoint, field): WENO5-Z VJPs of both sides, projection VJP,
//       scattered to the six stencil cells in six conflict-free sub-phases;
//   D4  per midpoint: the face-average state's cotangent (or the fallback's);
//   E   primitive -> conserved cotangents, ghost cotangents routed back to
//       the interior cells they were copied from, accumulated into MU.
// Every accumulation has a fixed order that does not depend on the tiling,
// so results are bit-reproducible across launch shapes.
#include <hip/hip_runtime.h>

#include "em_adjoint.h"
#include "em_device.h"

namespace em {
namespace adj {

__device__ __forceinline__ double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// c = sqrt(max(c2, 0)), c2 = gamma (p + p_inf) v^2 / (v - b), v = 1 / rho
__device__ __forceinline__ void sound_speed_vjp(const Eos& e, double rho, double p, double cb, double& rhob,
                                                double& pb) {
    const double v = 1.0 / rho;
    const double c2 = e.gamma * (p + e.p_inf) * v * v / (v - e.b);
    if (!(c2 > 0.0)) return;
    const double c = sqrt(c2);
    const double c2b = cb / (2.0 * c);
    const double A = e.gamma * (p + e.p_inf);
    const double vb = c2b * A * v * (v - 2.0 * e.b) / ((v - e.b) * (v - e.b));
    rhob += vb * (-v * v);
    pb += c2b * e.gamma * v * v / (v - e.b);
}

// q = to_cons(w)
__device__ __forceinline__ void to_cons_vjp(const Eos& e, const double w[5], const double qb[5], double wb[5]) {
    const double rho = w[0];
    const double drhoe_drho = -e.b * (w[4] + e.gamma * e.p_inf) / (e.gamma - 1.0) + e.q;
    const double drhoe_dp = (1.0 - rho * e.b) / (e.gamma - 1.0);
    wb[0] += qb[0] + qb[1] * w[1] + qb[2] * w[2] + qb[3] * w[3] +
             qb[4] * (drhoe_drho + 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]));
#pragma unroll
    for (int k = 1; k < 4; ++k) wb[k] += qb[k] * rho + qb[4] * rho * w[k];
    wb[4] += qb[4] * drhoe_dp;
}

// f = phys_flux<D>(w, q)
template <int D>
__device__ __forceinline__ void phys_flux_vjp(const double w[5], const double q[5], const double fb[5], double wb[5],
                                              double qb[5]) {
    constexpr int un = 1 + D;
    const double u = w[un];
#pragma unroll
    for (int k = 0; k < 4; ++k) qb[k] += fb[k] * u;
    wb[4] += fb[un];
    qb[4] += fb[4] * u;
    wb[4] += fb[4] * u;
    wb[un] += fb[0] * q[0] + fb[1] * q[1] + fb[2] * q[2] + fb[3] * q[3] + fb[4] * (q[4] + w[4]);
}

// w = to_prim(u), written with the primitives only (the cotangent needs
// rho, the velocity, p and the specific internal energy)
__device__ __forceinline__ void to_prim_vjp(const Eos& e, const double w[5], const double wb[5], double ub[5]) {
    const double rho = w[0];
    const double v = 1.0 / rho;
    const double pg = w[4] + e.gamma * e.p_inf;
    const double en = pg * (v - e.b) / (e.gamma - 1.0) + e.q;
    const double pb = wb[4];
    const double enb = pb * (e.gamma - 1.0) / (v - e.b);
    const double vb = -pb * pg / (v - e.b);
    double rhob = wb[0] + vb * (-v * v);
    const double ke = 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
    ub[4] = enb / rho;
    rhob -= enb * (en + ke) / rho;
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        const double velb = wb[k] - enb * w[k];
        ub[k] = velb / rho;
        rhob -= velb * w[k] / rho;
    }
    ub[0] = rhob;
}

// out = weno5z_interp(v0..v4).  Where a smoothness indicator is at the
// rounding floor (beta_k <= 1e-26 max v^2) the Z-weights are not
// differentiable: tau / (beta_k + eps) is a ratio involving rounding errors
// and its derivative is noise amplified by up to 1e40.  Convention (shared
// with the tape oracle): all three at the floor -> the linear weights'
// derivative (the smooth limit of WENO-Z); some at the floor -> the weights
// held at their computed values.
__device__ __forceinline__ void weno_vjp(const double v[5], double ob, double vb[5]) {
    const double q0 = (3.0 * v[0] - 10.0 * v[1] + 15.0 * v[2]) * 0.125;
    const double q1 = (-v[1] + 6.0 * v[2] + 3.0 * v[3]) * 0.125;
    const double q2 = (3.0 * v[2] + 6.0 * v[3] - v[4]) * 0.125;
    const double t0 = v[0] - 2.0 * v[1] + v[2], s0 = v[0] - 4.0 * v[1] + 3.0 * v[2];
    const double t1 = v[1] - 2.0 * v[2] + v[3], s1 = v[1] - v[3];
    const double t2 = v[2] - 2.0 * v[3] + v[4], s2 = 3.0 * v[2] - 4.0 * v[3] + v[4];
    const double c13 = 13.0 / 12.0;
    const double b0 = c13 * (t0 * t0) + 0.25 * (s0 * s0);
    const double b1 = c13 * (t1 * t1) + 0.25 * (s1 * s1);
    const double b2 = c13 * (t2 * t2) + 0.25 * (s2 * s2);
    const double mv = fmax(fmax(fmax(fabs(v[0]), fabs(v[1])), fmax(fabs(v[2]), fabs(v[3]))), fabs(v[4]));
    const double flat = 1e-26 * mv * mv;
    if (b0 <= flat && b1 <= flat && b2 <= flat) {
        const double q0b = 0.0625 * ob, q1b = 0.625 * ob, q2b = 0.3125 * ob;
        vb[0] += 0.375 * q0b;
        vb[1] += -1.25 * q0b - 0.125 * q1b;
        vb[2] += 1.875 * q0b + 0.75 * q1b + 0.375 * q2b;
        vb[3] += 0.375 * q1b + 0.75 * q2b;
        vb[4] += -0.125 * q2b;
        return;
    }
    // derivative values only (no branch depends on them): reciprocals
    // instead of divisions
    const double tau = fabs(b0 - b2);
    const double i0 = 1.0 / (b0 + 1e-40), i1 = 1.0 / (b1 + 1e-40), i2 = 1.0 / (b2 + 1e-40);
    const double r0 = tau * i0, r1 = tau * i1, r2 = tau * i2;
    const double a0 = 0.0625 * (1.0 + r0 * r0);
    const double a1 = 0.625 * (1.0 + r1 * r1);
    const double a2 = 0.3125 * (1.0 + r2 * r2);
    const double iS = 1.0 / (a0 + a1 + a2);
    const double Nb = ob * iS;
    const double q0b = a0 * Nb, q1b = a1 * Nb, q2b = a2 * Nb;
    if (b0 <= flat || b1 <= flat || b2 <= flat) {
        vb[0] += 0.375 * q0b;
        vb[1] += -1.25 * q0b - 0.125 * q1b;
        vb[2] += 1.875 * q0b + 0.75 * q1b + 0.375 * q2b;
        vb[3] += 0.375 * q1b + 0.75 * q2b;
        vb[4] += -0.125 * q2b;
        return;
    }
    // d out / d a_k = (q_k - out) / S, with q_k - out = sum_j w_j (q_k - q_j)
    // formed without cancellation: on a nearly flat sub-stencil (w_k -> 1)
    // the direct form q_k / S - out / S loses every digit and leaves
    // eps |q| / sqrt(beta_k) of noise in the beta_k derivative.
    const double d0 = (a1 * (q0 - q1) + a2 * (q0 - q2)) * iS;
    const double d1 = (a0 * (q1 - q0) + a2 * (q1 - q2)) * iS;
    const double d2 = (a0 * (q2 - q0) + a1 * (q2 - q1)) * iS;
    const double r0b = d0 * Nb * 0.0625 * 2.0 * r0;
    const double r1b = d1 * Nb * 0.625 * 2.0 * r1;
    const double r2b = d2 * Nb * 0.3125 * 2.0 * r2;
    const double taub = r0b * i0 + r1b * i1 + r2b * i2;
    const double sg = sgn(b0 - b2);
    const double b0b = -r0b * r0 * i0 + sg * taub;
    const double b1b = -r1b * r1 * i1;
    const double b2b = -r2b * r2 * i2 - sg * taub;
    const double t0b = b0b * 2.0 * c13 * t0, s0b = b0b * 0.5 * s0;
    const double t1b = b1b * 2.0 * c13 * t1, s1b = b1b * 0.5 * s1;
    const double t2b = b2b * 2.0 * c13 * t2, s2b = b2b * 0.5 * s2;
    vb[0] += 0.375 * q0b + t0b + s0b;
    vb[1] += -1.25 * q0b - 0.125 * q1b - 2.0 * t0b - 4.0 * s0b + t1b + s1b;
    vb[2] += 1.875 * q0b + 0.75 * q1b + 0.375 * q2b + t0b + 3.0 * s0b - 2.0 * t1b + t2b + 3.0 * s2b;
    vb[3] += 0.375 * q1b + 0.75 * q2b + t1b - s1b - 2.0 * t2b - 4.0 * s2b;
    vb[4] += -0.125 * q2b + t2b + s2b;
}

// Both WENO5-Z VJPs of a midpoint's characteristic field, left (v0..v4)
// and right (v5, v4, .., v1), as straight-line code that the compiler can
// interleave: the same values as weno_vjp on each side (its three cases
// become selects), chb[k] += the cotangent of v_k.  Only where both sides
// are flat to rounding (often whole waves, e.g. a velocity component that is
// zero) is the general form skipped.
struct WenoTerms {
    double q0, q1, q2, t0, s0, t1, s1, t2, s2, b0, b1, b2;
    bool all_flat, some_flat;
};
__device__ __forceinline__ WenoTerms weno_terms(double v0, double v1, double v2, double v3, double v4) {
    WenoTerms w;
    w.q0 = (3.0 * v0 - 10.0 * v1 + 15.0 * v2) * 0.125;
    w.q1 = (-v1 + 6.0 * v2 + 3.0 * v3) * 0.125;
    w.q2 = (3.0 * v2 + 6.0 * v3 - v4) * 0.125;
    w.t0 = v0 - 2.0 * v1 + v2;
    w.s0 = v0 - 4.0 * v1 + 3.0 * v2;
    w.t1 = v1 - 2.0 * v2 + v3;
    w.s1 = v1 - v3;
    w.t2 = v2 - 2.0 * v3 + v4;
    w.s2 = 3.0 * v2 - 4.0 * v3 + v4;
    const double c13 = 13.0 / 12.0;
    w.b0 = c13 * (w.t0 * w.t0) + 0.25 * (w.s0 * w.s0);
    w.b1 = c13 * (w.t1 * w.t1) + 0.25 * (w.s1 * w.s1);
    w.b2 = c13 * (w.t2 * w.t2) + 0.25 * (w.s2 * w.s2);
    const double mv = fmax(fmax(fmax(fabs(v0), fabs(v1)), fmax(fabs(v2), fabs(v3))), fabs(v4));
    const double flat = 1e-26 * mv * mv;
    w.all_flat = w.b0 <= flat && w.b1 <= flat && w.b2 <= flat;
    w.some_flat = w.b0 <= flat || w.b1 <= flat || w.b2 <= flat;
    return w;
}
// cotangents of v0..v4 from (q0b, q1b, q2b) and the beta parts (t*b, s*b)
__device__ __forceinline__ void weno_back(double q0b, double q1b, double q2b, double t0b, double s0b, double t1b,
                                          double s1b, double t2b, double s2b, double g[5]) {
    g[0] = 0.375 * q0b + t0b + s0b;
    g[1] = -1.25 * q0b - 0.125 * q1b - 2.0 * t0b - 4.0 * s0b + t1b + s1b;
    g[2] = 1.875 * q0b + 0.75 * q1b + 0.375 * q2b + t0b + 3.0 * s0b - 2.0 * t1b + t2b + 3.0 * s2b;
    g[3] = 0.375 * q1b + 0.75 * q2b + t1b - s1b - 2.0 * t2b - 4.0 * s2b;
    g[4] = -0.125 * q2b + t2b + s2b;
}
// weno_vjp's general and held-weight cases of one side (the caller selects
// the linear case)
__device__ __forceinline__ void weno_general(const WenoTerms& w, double ob, double g[5]) {
    const double c13 = 13.0 / 12.0;
    const double tau = fabs(w.b0 - w.b2);
    const double i0 = 1.0 / (w.b0 + 1e-40), i1 = 1.0 / (w.b1 + 1e-40), i2 = 1.0 / (w.b2 + 1e-40);
    const double r0 = tau * i0, r1 = tau * i1, r2 = tau * i2;
    const double a0 = 0.0625 * (1.0 + r0 * r0);
    const double a1 = 0.625 * (1.0 + r1 * r1);
    const double a2 = 0.3125 * (1.0 + r2 * r2);
    const double iS = 1.0 / (a0 + a1 + a2);
    const double Nb = ob * iS;
    const double q0b = a0 * Nb, q1b = a1 * Nb, q2b = a2 * Nb;
    const double d0 = (a1 * (w.q0 - w.q1) + a2 * (w.q0 - w.q2)) * iS;
    const double d1 = (a0 * (w.q1 - w.q0) + a2 * (w.q1 - w.q2)) * iS;
    const double d2 = (a0 * (w.q2 - w.q0) + a1 * (w.q2 - w.q1)) * iS;
    const double r0b = d0 * Nb * 0.0625 * 2.0 * r0;
    const double r1b = d1 * Nb * 0.625 * 2.0 * r1;
    const double r2b = d2 * Nb * 0.3125 * 2.0 * r2;
    const double taub = r0b * i0 + r1b * i1 + r2b * i2;
    const double sg = sgn(w.b0 - w.b2);
    const double b0b = -r0b * r0 * i0 + sg * taub;
    const double b1b = -r1b * r1 * i1;
    const double b2b = -r2b * r2 * i2 - sg * taub;
    const bool held = w.some_flat;  // weights held: no beta parts
    const double t0b = held ? 0.0 : b0b * 2.0 * c13 * w.t0, s0b = held ? 0.0 : b0b * 0.5 * w.s0;
    const double t1b = held ? 0.0 : b1b * 2.0 * c13 * w.t1, s1b = held ? 0.0 : b1b * 0.5 * w.s1;
    const double t2b = held ? 0.0 : b2b * 2.0 * c13 * w.t2, s2b = held ? 0.0 : b2b * 0.5 * w.s2;
    weno_back(q0b, q1b, q2b, t0b, s0b, t1b, s1b, t2b, s2b, g);
}
__device__ __forceinline__ void weno_vjp_pair(const double v[6], double obl, double obr, double chb[6]) {
    const WenoTerms L = weno_terms(v[0], v[1], v[2], v[3], v[4]);
    const WenoTerms R = weno_terms(v[5], v[4], v[3], v[2], v[1]);
    double gl[5], gr[5];
    weno_back(0.0625 * obl, 0.625 * obl, 0.3125 * obl, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, gl);
    weno_back(0.0625 * obr, 0.625 * obr, 0.3125 * obr, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, gr);
    if (!(L.all_flat && R.all_flat)) {
        double hl[5], hr[5];
        weno_general(L, obl, hl);
        weno_general(R, obr, hr);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            gl[k] = L.all_flat ? gl[k] : hl[k];
            gr[k] = R.all_flat ? gr[k] : hr[k];
        }
    }
    chb[0] = gl[0];
#pragma unroll
    for (int k = 1; k < 5; ++k) chb[k] = gl[k] + gr[5 - k];
    chb[5] = gr[0];
}

// f = hllc<D>(wl, wr) with the Davis speeds (system.rs:260-312)
template <int D>
__device__ void hllc_vjp(const Eos& e, const double wl[5], const double wr[5], const double fb[5], double wlb[5],
                         double wrb[5]) {
    constexpr int un = 1 + D;
    const double cl = sound_speed(e, wl[0], wl[4]), cr = sound_speed(e, wr[0], wr[4]);
    const double ul = wl[un], ur = wr[un];
    const double a1 = ul - cl, a2 = ur - cr;
    const double b1 = ul + cl, b2 = ur + cr;
    const bool sl_right = a2 < a1;  // fmin returns its second operand only when smaller
    const bool sr_right = b2 > b1;
    const double sl = fmin(a1, a2);
    const double sr = fmax(b1, b2);
    double ql[5], qr[5];
    to_cons(e, wl, ql);
    to_cons(e, wr, qr);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        wlb[k] = 0.0;
        wrb[k] = 0.0;
    }
    if (sl >= 0.0 || sr <= 0.0) {
        const bool lft = sl >= 0.0;
        const double* w = lft ? wl : wr;
        const double* q = lft ? ql : qr;
        double wb[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, qb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        phys_flux_vjp<D>(w, q, fb, wb, qb);
        to_cons_vjp(e, w, qb, wb);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            if (lft)
                wlb[k] = wb[k];
            else
                wrb[k] = wb[k];
        }
        return;
    }
    const double rl = wl[0], rr = wr[0], pl = wl[4], pr = wr[4];
    const double num = pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur);
    const double den = rl * (sl - ul) - rr * (sr - ur);
    const double ss = num / den;
    const bool left = ss >= 0.0;
    double w[5], q[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        w[k] = left ? wl[k] : wr[k];
        q[k] = left ? ql[k] : qr[k];
    }
    const double s = left ? sl : sr;
    const double rho = w[0], u = w[un];
    const double G = s - ss;
    const double H = rho * (s - u);
    const double fac = H / G;
    const double Z = w[4] / H;
    const double Y = ss + Z;
    const double X = q[4] / rho + (ss - u) * Y;
    double qs[5];
    qs[0] = fac;
    qs[1] = fac * w[1];
    qs[2] = fac * w[2];
    qs[3] = fac * w[3];
    qs[un] = fac * ss;
    qs[4] = fac * X;
    // f = phys_flux(w, q) + s (qs - q)
    double wb[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, qb[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, qsb[5];
    double sb = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        sb += fb[k] * (qs[k] - q[k]);
        qsb[k] = s * fb[k];
        qb[k] -= s * fb[k];
    }
    phys_flux_vjp<D>(w, q, fb, wb, qb);
    double facb = qsb[0] + qsb[un] * ss + qsb[4] * X;
#pragma unroll
    for (int k = 1; k < 4; ++k)
        if (k != un) {
            facb += qsb[k] * w[k];
            wb[k] += qsb[k] * fac;
        }
    double ssb = qsb[un] * fac;
    const double Xb = qsb[4] * fac;
    double rhob = 0.0, ub = 0.0, pb = 0.0;
    qb[4] += Xb / rho;
    rhob -= Xb * q[4] / (rho * rho);
    ssb += Xb * Y;
    ub -= Xb * Y;
    const double Yb = Xb * (ss - u);
    ssb += Yb;
    pb += Yb / H;
    const double Hb = -Yb * Z / H;
    rhob += Hb * (s - u);
    sb += Hb * rho;
    ub -= Hb * rho;
    // fac = rho (s - u) / (s - ss)
    rhob += facb * (s - u) / G;
    sb += facb * (rho - fac) / G;
    ub -= facb * rho / G;
    ssb += facb * fac / G;
    wb[0] += rhob;
    wb[un] += ub;
    wb[4] += pb;
    to_cons_vjp(e, w, qb, wb);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (left)
            wlb[k] += wb[k];
        else
            wrb[k] += wb[k];
    }
    double slb = left ? sb : 0.0, srb = left ? 0.0 : sb;
    // ss = num / den
    const double numb = ssb / den, denb = -ssb * ss / den;
    wrb[4] += numb;
    wlb[4] -= numb;
    wlb[0] += numb * ul * (sl - ul) + denb * (sl - ul);
    wlb[un] += numb * rl * (sl - 2.0 * ul) - denb * rl;
    slb += numb * rl * ul + denb * rl;
    wrb[0] += -numb * ur * (sr - ur) - denb * (sr - ur);
    wrb[un] += -numb * rr * (sr - 2.0 * ur) + denb * rr;
    srb += -numb * rr * ur - denb * rr;
    // Davis speeds
    double clb = 0.0, crb = 0.0;
    if (sl_right) {
        wrb[un] += slb;
        crb -= slb;
    } else {
        wlb[un] += slb;
        clb -= slb;
    }
    if (sr_right) {
        wrb[un] += srb;
        crb += srb;
    } else {
        wlb[un] += srb;
        clb += srb;
    }
    sound_speed_vjp(e, wl[0], wl[4], clb, wlb[0], wlb[4]);
    sound_speed_vjp(e, wr[0], wr[4], crb, wrb[0], wrb[4]);
}

// The limiter's admissibility constraint that fails at the bisection's
// upper bound hi, for the test states ul - tl (fx - fl) (0..2) and
// ur + tl (fx - fr) (3..5): rho > 0, 1 / rho > b, p + p_inf > 0, the first
// failing in that order; -1 when a test state is not finite or none fails.
__device__ __forceinline__ int limiter_constraint(const Eos& e, const double ul[5], const double ur[5],
                                                  const double fl[5], const double fr[5], const double fo[5],
                                                  const double flf[5], double tl, double hi) {
#pragma unroll 1
    for (int side = 0; side < 2; ++side) {
        double st[5], ws[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const double fx = isfinite(fo[k]) ? hi * fo[k] + (1.0 - hi) * flf[k] : flf[k];
            st[k] = side == 0 ? ul[k] - tl * (fx - fl[k]) : ur[k] + tl * (fx - fr[k]);
        }
        to_prim(e, st, ws);
        if (!isfinite(ws[0]) || !isfinite(ws[4])) return -1;
        if (!(ws[0] > 0.0)) return 3 * side;
        if (!(1.0 / ws[0] > e.b)) return 3 * side + 1;
        if (!(ws[4] + e.p_inf > 0.0)) return 3 * side + 2;
    }
    return -1;
}

// Characteristic projection at a midpoint (rows of L at the face-average
// state, sparse): the coefficients of system.rs:232-258 as the forward
// midpoint flux computes them.
struct Eig {
    double rho, avp, c, l0un, l04, l14, l4un, l44;
};
__device__ __forceinline__ Eig eig_at(const Eos& e, double rho2, double rho3, double p2, double p3) {
    Eig g;
    g.rho = 0.5 * (rho2 + rho3);
    g.avp = 0.5 * (p2 + p3);
    g.c = sound_speed(e, g.rho, g.avp);
    g.l0un = -g.rho / (2.0 * g.c);
    g.l04 = 1.0 / (2.0 * g.c * g.c);
    g.l14 = -1.0 / (g.c * g.c);
    g.l4un = g.rho / (2.0 * g.c);
    g.l44 = 1.0 / (2.0 * g.c * g.c);
    return g;
}
// ------------------------------------------------------------------ kernel

// Output range of tile t of a pencil: interior cells [x0, x1) split evenly,
// plus the ghost positions of a non-periodic edge (their cotangents go back
// to the interior cells they were copied from, which lie in the same tile).
struct AdjTile {
    int x0, x1, e0, e1;
};
__device__ __forceinline__ AdjTile adj_tile(int n, int ntile, int t, bool cyc) {
    AdjTile r;
    r.x0 = (int)((long long)t * n / ntile);
    r.x1 = (int)((long long)(t + 1) * n / ntile);
    r.e0 = (!cyc && t == 0) ? -5 : r.x0;
    r.e1 = (!cyc && t == ntile - 1) ? n + 5 : r.x1;
    return r;
}

__device__ __forceinline__ int wrap(int x, int n) { return ((x % n) + n) % n; }

// Work item w (pencil w / ntile, tile w % ntile) of one axis VJP described
// by `a`; every thread of the workgroup calls it and it returns uniformly.
//
// Positions are global along the pencil (cell X, midpoint j between cells j
// and j + 1, face f between cells f - 1 and f); the tile's output positions
// are [e0, e1).  Local slots: cell X at X - e0 + 9, midpoint j at j - e0 + 7,
// face f at f - e0 + 4, so that midpoint slot m reads cells m .. m + 5 and
// averages cells m + 2, m + 3; face slot f combines midpoint slots f .. f + 4
// and its limiter reads cells f + 4, f + 5; midpoint slot m takes the
// cotangents of face slots m - 4 .. m.  The window: cells [e0 - 9, e1 + 9),
// forward midpoints [e0 - 7, e1 + 6), faces [e0 - 4, e1 + 5), VJP midpoints
// [e0 - 3, e1 + 2), clipped to the pencil's positions unless the axis is
// periodic, where positions wrap (the forward's periodic ghosts are the
// cyclic neighbours, so a position's values and branches are those of the
// cell it wraps to).
template <int D>
__device__ __forceinline__ void adj_group(const AdjArgs& a, long long w, double* lds) {
    constexpr int un = 1 + D;
    constexpr int E0 = D == 0 ? 1 : 0;
    constexpr int E1 = D == 2 ? 1 : 2;
    const int n = a.n[D];
    const int NC = a.ne + 18, NM = a.ne + 13;  // slot extents (the launch's largest tile)
    const int NT = blockDim.x;
    const int tid = threadIdx.x;
    double* sw = lds;              // [5][NC] primitives
    double* sacc = sw + 5 * NC;    // [5][NC] cotangents (primitive, then conserved)
    double* sx = sacc + 5 * NC;    // [3][NC] cotangent parts of fields 1 and 4 (see D3)
    double* smf = sx + 3 * NC;     // [5][NM] midpoint fluxes, then face cotangents
    double* scc = smf + 5 * NM;    // [18][NM] cl, cr (B1); clb, crb, partials, fallback flag (D2, D3);
                                   // 13..17: rows of L (B0)
    auto WC = [&](double* arr, int q, int c) -> double& { return arr[q * NC + c]; };
    auto WM = [&](double* arr, int q, int m) -> double& { return arr[q * NM + m]; };
    const long long npen = a.ncell / n;
    const long long pen = w / a.ntile;
    const int t = (int)(w - pen * a.ntile);
    if (pen >= npen) return;
    const bool cyc = a.bc_kind[0] == 0 && a.bc_kind[1] == 0;
    const AdjTile tl_ = adj_tile(n, a.ntile, t, cyc);
    const int e0 = tl_.e0, ne = tl_.e1 - tl_.e0;
    // slot ranges (non-periodic: cells [-5, n + 5), midpoints [-3, n + 2),
    // faces [0, n])
    const int lo = cyc ? 0 : max(0, 4 - e0);
    const int c_hi = cyc ? ne + 18 : min(ne + 18, n + 14 - e0);
    const int m_hi = cyc ? ne + 13 : min(ne + 13, n + 9 - e0);
    const int f_hi = cyc ? ne + 9 : min(ne + 9, n + 5 - e0);
    const int v_lo = max(4, lo), v_hi = min(ne + 9, m_hi);  // VJP midpoints
    const Eos e = a.eos;
    const long long base = (pen / a.n[E1]) * a.stride[E0] + (pen % a.n[E1]) * a.stride[E1];
    const long long sd = a.stride[D];
    // ---- A. load (boundary ghosts through bc_ghost, periodic positions wrapped)
    for (int c = tid; c < NC; c += NT) {
#pragma unroll
        for (int q = 0; q < 5; ++q) WC(sacc, q, c) = 0.0;
#pragma unroll
        for (int q = 0; q < 3; ++q) WC(sx, q, c) = 0.0;
        if (c < lo || c >= c_hi) continue;
        auto ld = [&](int x, double* o) {
#pragma unroll
            for (int q = 0; q < 5; ++q) o[q] = a.U[q * a.ncell + base + (long long)x * sd];
        };
        const int x = c + e0 - 9;
        double u[5], wv[5];
        if (cyc)
            ld(wrap(x, n), u);
        else if (x >= 0 && x < n)
            ld(x, u);
        else
            bc_ghost<D>(e, n, x, a.bc_kind, a.bc_state, ld, u);
        to_prim(e, u, wv);
#pragma unroll
        for (int q = 0; q < 5; ++q) WC(sw, q, c) = wv[q];
    }
    __syncthreads();
    constexpr int t1 = 1 + (D + 1) % 3;
    constexpr int t2 = 1 + (D + 2) % 3;
    // characteristic field fa over cells m .. m+5: row fa of L (sparse) at
    // the midpoint's face-average state, whose nonzero coefficients are
    // (la, lb) from the rows stored in B0 (13 l0un, 14 l04, 15 l14, 16 l4un,
    // 17 l44):
    //   fa = 0: l0un u_n + l04 p,  1: rho + l14 p,  2: u_t1,  3: u_t2,
    //   fa = 4: l4un u_n + l44 p
    // Rows and coefficients are selected, not branched on, so that the
    // twelve loads (rows ra and 4 of the six cells) issue together; a[] and
    // b[] keep them (rows u_n and p for fields 0 and 4: the coefficient sums
    // of D3 reuse them).
    auto stencil = [&](int fa, int m, double v[6], double ar[6], double br[6], double& la, double& lb) {
        const int ra = (fa == 0 || fa == 4) ? un : fa == 1 ? 0 : fa == 2 ? t1 : t2;
        const double lav = WM(scc, fa == 0 ? 13 : 16, m);
        const double lbv = WM(scc, fa == 0 ? 14 : fa == 1 ? 15 : 17, m);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            ar[k] = WC(sw, ra, m + k);
            br[k] = WC(sw, 4, m + k);
        }
        la = (fa == 0 || fa == 4) ? lav : 0.0;
        lb = (fa == 2 || fa == 3) ? 0.0 : lbv;
        const double ca = fa == 1 ? 1.0 : la;  // 1.0 * rho is rho exactly
        const bool single = fa == 2 || fa == 3;
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = single ? ar[k] : ca * ar[k] + lb * br[k];
    };
    auto eig_of = [&](int m) -> Eig {
        return eig_at(e, WC(sw, 0, m + 2), WC(sw, 0, m + 3), WC(sw, 4, m + 2), WC(sw, 4, m + 3));
    };
    // left / right reconstructions of midpoint m: back-projection w = R c,
    // first-order fallback on an inadmissible state (scheme.rs:310-331)
    auto recon = [&](const Eig& g, int m, double* wl, double* wr) -> bool {
        const double run0 = -g.c / g.rho, run4 = g.c / g.rho, r40 = g.c * g.c;
        double cl[5], cr[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            cl[q] = WM(scc, q, m);
            cr[q] = WM(scc, 5 + q, m);
        }
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
        if (!admissible(e, wl[0], wl[4]) || !admissible(e, wr[0], wr[4])) {
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                wl[q] = WC(sw, q, m + 2);
                wr[q] = WC(sw, q, m + 3);
            }
            return true;
        }
        return false;
    };
    // ---- B. forward midpoint fluxes, the same operations as the forward
    // pass (midpoint_flux).
    // B0: the rows of L per midpoint (once: divisions and a square root)
    for (int m = lo + tid; m < m_hi; m += NT) {
        const Eig g = eig_of(m);
        WM(scc, 13, m) = g.l0un;
        WM(scc, 14, m) = g.l04;
        WM(scc, 15, m) = g.l14;
        WM(scc, 16, m) = g.l4un;
        WM(scc, 17, m) = g.l44;
    }
    __syncthreads();
    // B1: WENO5-Z reconstructions, one work item per (field, midpoint),
    // field-major so that a wave mostly holds one field
    const int nfm = m_hi - lo;
    for (int it = tid; it < 5 * nfm; it += NT) {
        const int fa = it / nfm, m = lo + it - fa * nfm;
        double v[6], ar[6], br[6], la, lb;
        stencil(fa, m, v, ar, br, la, lb);
        WM(scc, fa, m) = weno5z_interp(v[0], v[1], v[2], v[3], v[4]);
        WM(scc, 5 + fa, m) = weno5z_interp(v[5], v[4], v[3], v[2], v[1]);
    }
    __syncthreads();
    // B2: back-projection, fallback, HLLC per midpoint
    for (int m = lo + tid; m < m_hi; m += NT) {
        const Eig g = eig_of(m);
        double wl[5], wr[5], fx[5];
        recon(g, m, wl, wr);
        hllc<D>(e, wl, wr, fx);
#pragma unroll
        for (int q = 0; q < 5; ++q) WM(smf, q, m) = fx[q];
    }
    __syncthreads();
    // ---- C. faces in chunks of increasing f: a chunk only overwrites the
    // midpoint slots of its own faces, which later chunks do not read
    for (int f0 = lo; f0 < f_hi; f0 += NT) {
        const int f = f0 + tid;
        const bool live = f < f_hi;
        double fhb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        double wlb[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, wrb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (live) {
            double fo[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                double v = 0.0;
                v += S6_0 * WM(smf, k, f + 2);
                v += S6_1 * WM(smf, k, f + 3);
                v += S6_1 * WM(smf, k, f + 1);
                v += S6_2 * WM(smf, k, f + 4);
                v += S6_2 * WM(smf, k, f);
                fo[k] = v;
            }
            // face cotangent from the divergence: Fbar = (lam_f - lam_{f-1}) / dx
            // (lam = 0 outside the pencil, cyclic on a periodic axis)
            const int fg = f + e0 - 4;
            const bool has_r = cyc || fg < n, has_l = cyc || fg > 0;
            const long long xr = cyc ? wrap(fg, n) : fg, xl = cyc ? wrap(fg - 1, n) : fg - 1;
            double Fb[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const double lr = has_r ? a.LAM[k * a.ncell + base + xr * sd] : 0.0;
                const double ll = has_l ? a.LAM[k * a.ncell + base + xl * sd] : 0.0;
                Fb[k] = (lr - ll) * a.inv_dx;
            }
            bool limited = false;
            double lo_t = 1.0, hi_t = 1.0, al = 0.0;
            double wl[5], wr[5], ul[5], ur[5], fl[5], fr[5], flf[5];
            if (a.lam_lim > 0.0) {
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    wl[k] = WC(sw, k, f + 4);
                    wr[k] = WC(sw, k, f + 5);
                }
                to_cons(e, wl, ul);
                to_cons(e, wr, ur);
                phys_flux<D>(wl, ul, fl);
                phys_flux<D>(wr, ur, fr);
                const double tl = 2.0 * a.lam_lim;
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
                    limited = true;
                    al = fmax(fabs(wl[un]) + sound_speed(e, wl[0], wl[4]), fabs(wr[un]) + sound_speed(e, wr[0], wr[4]));
#pragma unroll
                    for (int k = 0; k < 5; ++k) flf[k] = 0.5 * (fl[k] + fr[k]) - 0.5 * al * (ur[k] - ul[k]);
                    double l0 = 0.0, h0 = 1.0;
                    for (int itb = 0; itb < 30; ++itb) {
                        const double mid = 0.5 * (l0 + h0);
                        double bl[5];
#pragma unroll
                        for (int k = 0; k < 5; ++k) bl[k] = isfinite(fo[k]) ? mid * fo[k] + (1.0 - mid) * flf[k] : flf[k];
                        if (ok(bl))
                            l0 = mid;
                        else
                            h0 = mid;
                    }
                    lo_t = l0;
                    hi_t = h0;
                }
            }
            if (limited) {
                // out = lo fh + (1 - lo) flf per finite component, flf otherwise
                const double tl = 2.0 * a.lam_lim;
                double flfb[5], flb[5], frb[5], ulb[5], urb[5];
                double thb = 0.0;
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    const bool fk = isfinite(fo[k]);
                    fhb[k] = fk ? lo_t * Fb[k] : 0.0;
                    flfb[k] = fk ? (1.0 - lo_t) * Fb[k] : Fb[k];
                    if (fk) thb += Fb[k] * (fo[k] - flf[k]);
                    flb[k] = 0.0;
                    frb[k] = 0.0;
                    ulb[k] = 0.0;
                    urb[k] = 0.0;
                }
                // theta: the bisection resolves theta* (the admissibility
                // boundary) to 2^-30, so theta follows theta*(u) and is
                // differentiated implicitly at the constraint g that fails
                // at the upper bound: d theta = -dg / (dg/dtheta).
                // theta = 0 (nothing admissible above it) stays constant.
                const int which = lo_t > 0.0 ? limiter_constraint(e, ul, ur, fl, fr, fo, flf, tl, hi_t) : -1;
                if (which >= 0) {
                    const bool left = which < 3;  // test state ul - tl (fx - fl), else ur + tl (fx - fr)
                    const double sg = left ? -tl : tl;
                    double st[5], ws[5], gw[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gs[5];
#pragma unroll
                    for (int k = 0; k < 5; ++k) {
                        const double fx = isfinite(fo[k]) ? lo_t * fo[k] + (1.0 - lo_t) * flf[k] : flf[k];
                        st[k] = left ? ul[k] - tl * (fx - fl[k]) : ur[k] + tl * (fx - fr[k]);
                    }
                    to_prim(e, st, ws);
                    const int c = which % 3;
                    if (c == 0)
                        gw[0] = 1.0;  // rho
                    else if (c == 1)
                        gw[0] = -1.0 / (ws[0] * ws[0]);  // 1 / rho - b
                    else
                        gw[4] = 1.0;  // p + p_inf
                    to_prim_vjp(e, ws, gw, gs);
                    double gth = 0.0;
#pragma unroll
                    for (int k = 0; k < 5; ++k)
                        if (isfinite(fo[k])) gth += gs[k] * (sg * (fo[k] - flf[k]));
                    if (isfinite(gth) && gth != 0.0) {
                        const double gb = -thb / gth;
#pragma unroll
                        for (int k = 0; k < 5; ++k) {
                            const double sb = gb * gs[k];  // cotangent of the test state
                            const double fxb = sg * sb;
                            if (isfinite(fo[k])) {
                                fhb[k] += lo_t * fxb;
                                flfb[k] += (1.0 - lo_t) * fxb;
                            } else {
                                flfb[k] += fxb;
                            }
                            if (left) {
                                ulb[k] += sb;
                                flb[k] += tl * sb;
                            } else {
                                urb[k] += sb;
                                frb[k] -= tl * sb;
                            }
                        }
                    }
                }
                // flf = (fl + fr) / 2 - al (ur - ul) / 2; al = max of the two
                // cells' |u_n| + c (the operand max returns, the first on a tie)
                double alb = 0.0;
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    flb[k] += 0.5 * flfb[k];
                    frb[k] += 0.5 * flfb[k];
                    urb[k] += -0.5 * al * flfb[k];
                    ulb[k] += 0.5 * al * flfb[k];
                    alb -= 0.5 * flfb[k] * (ur[k] - ul[k]);
                }
                const double spl = fabs(wl[un]) + sound_speed(e, wl[0], wl[4]);
                const double spr = fabs(wr[un]) + sound_speed(e, wr[0], wr[4]);
                if (spr > spl) {
                    wrb[un] += alb * sgn(wr[un]);
                    sound_speed_vjp(e, wr[0], wr[4], alb, wrb[0], wrb[4]);
                } else {
                    wlb[un] += alb * sgn(wl[un]);
                    sound_speed_vjp(e, wl[0], wl[4], alb, wlb[0], wlb[4]);
                }
                phys_flux_vjp<D>(wl, ul, flb, wlb, ulb);
                phys_flux_vjp<D>(wr, ur, frb, wrb, urb);
                to_cons_vjp(e, wl, ulb, wlb);
                to_cons_vjp(e, wr, urb, wrb);
            } else {
#pragma unroll
                for (int k = 0; k < 5; ++k) fhb[k] = Fb[k];
            }
        }
        // a cell takes the right state of face c-5 before the left state of
        // face c-4: increasing face order within a chunk as across chunks,
        // so the sums do not depend on the tiling
        __syncthreads();
        if (live) {
#pragma unroll
            for (int k = 0; k < 5; ++k) WM(smf, k, f) = fhb[k];
#pragma unroll
            for (int k = 0; k < 5; ++k) WC(sacc, k, f + 5) += wrb[k];
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int k = 0; k < 5; ++k) WC(sacc, k, f + 4) += wlb[k];
        }
        __syncthreads();
    }
    // ---- D2. per VJP midpoint: cotangent from its five faces (S6 weights),
    // the HLLC VJP and the back-projection VJP -> clb, crb (scc 0..9) and the
    // eigenvector partials (scc 10, 11); on the fallback the state
    // cotangents of cells m+2, m+3 instead (flag scc 12)
    for (int m = v_lo + tid; m < v_hi; m += NT) {
        double mb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        const double w6[5] = {S6_2, S6_1, S6_0, S6_1, S6_2};
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const int f = m - 4 + j;  // faces f = m-4 .. m use midpoint m with weight S6[|m - f - 2|]
            if (f < lo || f >= f_hi) continue;
#pragma unroll
            for (int q = 0; q < 5; ++q) mb[q] += w6[4 - j] * WM(smf, q, f);
        }
        const Eig g = eig_of(m);
        double wl[5], wr[5], wlb[5], wrb[5];
        const bool fb = recon(g, m, wl, wr);
        hllc_vjp<D>(e, wl, wr, mb, wlb, wrb);
        if (fb) {
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                WM(scc, q, m) = wlb[q];
                WM(scc, 5 + q, m) = wrb[q];
            }
            WM(scc, 10, m) = 0.0;
            WM(scc, 11, m) = 0.0;
            WM(scc, 12, m) = 1.0;
            continue;
        }
        double cl[5], cr[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            cl[q] = WM(scc, q, m);
            cr[q] = WM(scc, 5 + q, m);
        }
        const double run0 = -g.c / g.rho, run4 = g.c / g.rho, r40 = g.c * g.c;
        const double run0b = wlb[un] * cl[0] + wrb[un] * cr[0];
        const double run4b = wlb[un] * cl[4] + wrb[un] * cr[4];
        const double r40b = wlb[4] * (cl[0] + cl[4]) + wrb[4] * (cr[0] + cr[4]);
        WM(scc, 0, m) = wlb[0] + run0 * wlb[un] + r40 * wlb[4];
        WM(scc, 5, m) = wrb[0] + run0 * wrb[un] + r40 * wrb[4];
        WM(scc, 1, m) = wlb[0];
        WM(scc, 6, m) = wrb[0];
        WM(scc, 2, m) = wlb[t1];
        WM(scc, 7, m) = wrb[t1];
        WM(scc, 3, m) = wlb[t2];
        WM(scc, 8, m) = wrb[t2];
        WM(scc, 4, m) = wlb[0] + run4 * wlb[un] + r40 * wlb[4];
        WM(scc, 9, m) = wrb[0] + run4 * wrb[un] + r40 * wrb[4];
        // run0 = -c / rho, run4 = c / rho, r40 = c^2 as functions of rho, c
        WM(scc, 10, m) = run0b * g.c / (g.rho * g.rho) - run4b * g.c / (g.rho * g.rho);
        WM(scc, 11, m) = -run0b / g.rho + run4b / g.rho + 2.0 * g.c * r40b;
        WM(scc, 12, m) = 0.0;
    }
    __syncthreads();
    // ---- D3. per (field, VJP midpoint): the WENO5-Z VJPs of both sides and
    // the projection ch = L s (the coefficient sums left in scc 0..9 for D4),
    // scattered to the six stencil cells in six conflict-free sub-phases.
    // Fields 1 and 4 keep their shares of the pressure (and field 4 of the
    // normal velocity) cotangent apart in sx, so that no two fields add to
    // one slot in a sub-phase.  A cell takes midpoints c-5 .. c in
    // increasing order, within a chunk as across chunks, so the sums do not
    // depend on the tiling.  Work items run field-major, so that a wave
    // mostly holds one field of neighbouring midpoints: flat stencils (e.g.
    // every stencil of a velocity component that is zero) then take
    // weno_vjp's short path for whole waves.
    const int nvm = v_hi - v_lo;
    for (int i0 = 0; i0 < 5 * nvm; i0 += NT) {
        const int it = i0 + tid;
        const bool in = it < 5 * nvm;
        const int fa = in ? it / nvm : 0, m = in ? v_lo + it - fa * nvm : v_lo;
        const bool live = in && WM(scc, 12, m) == 0.0;
        double chb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        double la = 0.0, lb = 0.0;
        if (live) {
            double v[6], ar[6], br[6];
            stencil(fa, m, v, ar, br, la, lb);
            weno_vjp_pair(v, WM(scc, fa, m), WM(scc, 5 + fa, m), chb);
            // sums for the coefficient cotangents: d la = sum chb u_n (fields
            // 0, 4: ar is row u_n), d lb = sum chb p (fields 0, 1, 4); the
            // other fields' sums are not read
            double sa = 0.0, sbp = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                sa += chb[k] * ar[k];
                sbp += chb[k] * br[k];
            }
            WM(scc, fa, m) = sa;
            WM(scc, 5 + fa, m) = sbp;
        }
        // targets as rows of sacc (rows 5, 6, 7 are sx 0, 1, 2):
        //   fa = 0: u_n (la), p (lb);  1: rho (1), sx 0 (lb);  2: u_t1 (1);
        //   fa = 3: u_t2 (1);  4: sx 1 (la), sx 2 (lb)
        const int d1 = fa == 0 ? un : fa == 1 ? 0 : fa == 2 ? t1 : fa == 3 ? t2 : 6;
        const int d2 = fa == 0 ? 4 : fa == 1 ? 5 : 7;
        const bool two = fa == 0 || fa == 1 || fa == 4;
        const double k1 = (fa == 0 || fa == 4) ? la : 1.0;  // 1.0 * g is g exactly
        for (int j = 5; j >= 0; --j) {
            if (live) {
                const double gj = chb[j];
                const int c = m + j;
                sacc[d1 * NC + c] += k1 * gj;
                if (two) sacc[d2 * NC + c] += lb * gj;
            }
            __syncthreads();
        }
    }
    // ---- D4. per VJP midpoint: the average state's cotangent (rho and p of
    // cells m+2, m+3 through rho, c), or the fallback's state cotangents;
    // a cell takes midpoint c-3 (as m+3) before c-2 (as m+2)
    for (int m0 = v_lo; m0 < v_hi; m0 += NT) {
        const int m = m0 + tid;
        const bool live = m < v_hi;
        double b2[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, b3[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (live) {
            if (WM(scc, 12, m) != 0.0) {
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    b2[q] = WM(scc, q, m);
                    b3[q] = WM(scc, 5 + q, m);
                }
            } else {
                const Eig g = eig_of(m);
                // L coefficients as functions of rho and c: l0un = -rho/(2c),
                // l04 = l44 = 1/(2c^2), l14 = -1/c^2, l4un = rho/(2c)
                const double l0unb = WM(scc, 0, m), l04b = WM(scc, 5, m);
                const double l14b = WM(scc, 6, m);
                const double l4unb = WM(scc, 4, m), l44b = WM(scc, 9, m);
                const double c3 = g.c * g.c * g.c;
                double rhob = WM(scc, 10, m) - l0unb / (2.0 * g.c) + l4unb / (2.0 * g.c);
                double cb = WM(scc, 11, m) + l0unb * g.rho / (2.0 * g.c * g.c) - l4unb * g.rho / (2.0 * g.c * g.c) -
                            (l04b + l44b) / c3 + 2.0 * l14b / c3;
                double pb = 0.0;
                sound_speed_vjp(e, g.rho, g.avp, cb, rhob, pb);
                b2[0] = b3[0] = 0.5 * rhob;
                b2[4] = b3[4] = 0.5 * pb;
            }
        }
        if (live) {
#pragma unroll
            for (int q = 0; q < 5; ++q) WC(sacc, q, m + 3) += b3[q];
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int q = 0; q < 5; ++q) WC(sacc, q, m + 2) += b2[q];
        }
        __syncthreads();
    }
    // ---- E. conserved cotangents of the tile's output positions
    for (int c = 9 + tid; c < ne + 9; c += NT) {
        double wv[5], wb[5], ub[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            wv[q] = WC(sw, q, c);
            wb[q] = WC(sacc, q, c);
        }
        wb[un] += WC(sx, 1, c);
        wb[4] += WC(sx, 0, c) + WC(sx, 2, c);
        to_prim_vjp(e, wv, wb, ub);
#pragma unroll
        for (int q = 0; q < 5; ++q) WC(sacc, q, c) = ub[q];
    }
    __syncthreads();
    // interior cells: own cotangent plus the ghosts copied from them (a
    // non-periodic edge tile holds its edge's ghosts)
    const int x0 = tl_.x0, nx = tl_.x1 - tl_.x0;
    for (int it = tid; it < nx; it += NT) {
        const int x = x0 + it;
        const int c = x - e0 + 9;
        double tot[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) tot[q] = WC(sacc, q, c);
        if (!cyc && (x < 5 || x >= n - 5)) {
            for (int gi = 0; gi < 10; ++gi) {
                const int gx = gi < 5 ? gi - 5 : n + gi - 5;
                if (gx < e0 || gx >= tl_.e1) continue;
                const int side = gx < 0 ? 0 : 1;
                const int kind = a.bc_kind[side];
                int src = -1;
                double flip = 1.0;
                if (kind == 0)
                    src = wrap(gx, n);
                else if (kind == 1)
                    src = side == 0 ? 0 : n - 1;
                else if (kind == 2) {
                    src = side == 0 ? -1 - gx : 2 * n - 1 - gx;
                    flip = -1.0;
                }
                if (src != x) continue;
#pragma unroll
                for (int q = 0; q < 5; ++q) tot[q] += (q == un ? flip : 1.0) * WC(sacc, q, gx - e0 + 9);
            }
        }
        const long long cell = base + (long long)x * sd;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const long long ix = q * a.ncell + cell;
            const double T = (a.first ? 0.0 : a.MU[ix]) + tot[q];
            if (a.combine == 0) {
                a.MU[ix] = T;
            } else if (a.combine == 3) {
                const double av = a.cA[ix];
                a.cBout[ix] = (2.0 / 3.0) * av + (2.0 / 3.0) * a.cdt * T;
                a.cZ[ix] = av / 3.0;
            } else if (a.combine == 2) {
                const double b = a.cB[ix];
                a.cZ[ix] = a.cZ[ix] + 0.75 * b;
                a.cBout[ix] = 0.25 * b + 0.25 * a.cdt * T;
            } else {
                a.cZ[ix] = a.cZ[ix] + a.cB[ix] + a.cdt * T;
            }
        }
    }
}

// Three workgroups per CU (LDS of a tile <= 45 KB): at most 168 VGPRs, so
// that their 12 waves are resident together.  A tile's phases keep only a
// few of its waves busy (about 140 midpoints for 256 threads), so tiles in
// flight per CU, not LDS per tile, set the speed: 128-cell tiles with 256
// threads ran the 512^2 backward 12% faster than whole pencils, and 10-25%
// faster than 200-256-cell tiles with 384 threads.
template <int D>
__global__ __launch_bounds__(kAdjThreads) __attribute__((amdgpu_waves_per_eu(3))) void adj_pass(AdjArgs a) {
    extern __shared__ double lds[];
    // neighbouring work items (tiles of neighbouring pencils) on one XCD:
    // workgroups are dealt round-robin to the 8 XCDs, each with its own L2
    const long long G = gridDim.x, b = blockIdx.x;
    const long long q = G / 8, r = G % 8, x = b % 8, k = b / 8;
    const long long w = x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
    adj_group<D>(a, w, lds);
}

}  // namespace adj

template <class K>
static hipError_t allow_lds(K kernel, size_t lds) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds);
}

// Tiles per pencil: pencils longer than cap cells (kAdjTileCells unless set
// otherwise; at least 10, so that an even split leaves every tile the 5
// cells a reflecting edge routes its ghosts to) are split evenly; an axis
// periodic on one side only keeps whole pencils (its ghosts are copies of
// the far end).
int adj_tiles(int n, int kind_lo, int kind_hi, int cap) {
    if ((kind_lo == 0) != (kind_hi == 0)) return 1;
    return std::max(1, (n + cap - 1) / cap);
}

// Output positions of a pencil's longest tile: its interior cells plus the
// ghost positions of a non-periodic edge.
int adj_tile_extent(int n, int ntile, int kind_lo, int kind_hi) {
    const bool cyc = kind_lo == 0 && kind_hi == 0;
    return (n + ntile - 1) / ntile + (cyc ? 0 : (ntile == 1 ? 10 : 5));
}

// LDS of a tile with ne output positions: 13 arrays over its ne + 18 cells
// (primitives, cotangents, field shares) and 23 over its ne + 13 forward
// midpoints (fluxes / face cotangents, reconstructions / cotangents /
// partials / flag, rows of L)
size_t adj_lds_bytes(int ne) {
    return sizeof(double) * (13 * (size_t)(ne + 18) + 23 * (size_t)(ne + 13));
}

hipError_t launch_adj_pass(int d, const AdjArgs& a, hipStream_t s) {
    const int n = a.n[d];
    const long long npen = a.ncell / n;
    const unsigned grid = (unsigned)(npen * a.ntile);
    const size_t lds = adj_lds_bytes(a.ne);
    hipError_t er;
    if (d == 0) {
        if ((er = allow_lds(adj::adj_pass<0>, lds)) != hipSuccess) return er;
        hipLaunchKernelGGL(adj::adj_pass<0>, dim3(grid), dim3(kAdjThreads), lds, s, a);
    } else if (d == 1) {
        if ((er = allow_lds(adj::adj_pass<1>, lds)) != hipSuccess) return er;
        hipLaunchKernelGGL(adj::adj_pass<1>, dim3(grid), dim3(kAdjThreads), lds, s, a);
    } else {
        if ((er = allow_lds(adj::adj_pass<2>, lds)) != hipSuccess) return er;
        hipLaunchKernelGGL(adj::adj_pass<2>, dim3(grid), dim3(kAdjThreads), lds, s, a);
    }
    return hipGetLastError();
}

}  // namespace em
