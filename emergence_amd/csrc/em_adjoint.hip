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
// One workgroup owns whole pencils along the pass axis (interior cells plus
// the 5 virtual ghost layers on each side) in LDS:
//   A  load U, primitives (boundary ghosts through bc_ghost);
//   B  forward midpoint fluxes;
//   C  face fluxes and limiter state, face cotangents
//      Fbar_f = (lam_f - lam_{f-1}) / dx, the limiter's Lax-Friedrichs part
//      scattered to its two cells;
//   D  midpoint cotangents (S6 weights) and the midpoint VJP, scattered to
//      the six stencil cells in six conflict-free sub-phases;
//   E  primitive -> conserved cotangents, ghost cotangents routed back to
//      the interior cells they were copied from, accumulated into MU.
// Every accumulation has a fixed order that does not depend on the pencils
// per workgroup, so results are bit-reproducible across launch shapes.
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
    const double tau = fabs(b0 - b2);
    const double e0 = b0 + 1e-40, e1 = b1 + 1e-40, e2 = b2 + 1e-40;
    const double r0 = tau / e0, r1 = tau / e1, r2 = tau / e2;
    const double a0 = 0.0625 * (1.0 + r0 * r0);
    const double a1 = 0.625 * (1.0 + r1 * r1);
    const double a2 = 0.3125 * (1.0 + r2 * r2);
    const double S = a0 + a1 + a2;
    const double Nb = ob / S;
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
    const double d0 = (a1 * (q0 - q1) + a2 * (q0 - q2)) / S;
    const double d1 = (a0 * (q1 - q0) + a2 * (q1 - q2)) / S;
    const double d2 = (a0 * (q2 - q0) + a1 * (q2 - q1)) / S;
    const double r0b = d0 * Nb * 0.0625 * 2.0 * r0;
    const double r1b = d1 * Nb * 0.625 * 2.0 * r1;
    const double r2b = d2 * Nb * 0.3125 * 2.0 * r2;
    const double taub = r0b / e0 + r1b / e1 + r2b / e2;
    const double sg = sgn(b0 - b2);
    const double b0b = -r0b * r0 / e0 + sg * taub;
    const double b1b = -r1b * r1 / e1;
    const double b2b = -r2b * r2 / e2 - sg * taub;
    const double t0b = b0b * 2.0 * c13 * t0, s0b = b0b * 0.5 * s0;
    const double t1b = b1b * 2.0 * c13 * t1, s1b = b1b * 0.5 * s1;
    const double t2b = b2b * 2.0 * c13 * t2, s2b = b2b * 0.5 * s2;
    vb[0] += 0.375 * q0b + t0b + s0b;
    vb[1] += -1.25 * q0b - 0.125 * q1b - 2.0 * t0b - 4.0 * s0b + t1b + s1b;
    vb[2] += 1.875 * q0b + 0.75 * q1b + 0.375 * q2b + t0b + 3.0 * s0b - 2.0 * t1b + t2b + 3.0 * s2b;
    vb[3] += 0.375 * q1b + 0.75 * q2b + t1b - s1b - 2.0 * t2b - 4.0 * s2b;
    vb[4] += -0.125 * q2b + t2b + s2b;
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

// fbar of the WCNS midpoint flux of the six states s[0..5] -> sb (added).
// The characteristic values are recomputed per field instead of being kept
// (they are cheap and registers are not).
template <int D>
__device__ void midpoint_vjp(const Eos& e, const double (&s)[6][5], const double fbar[5], double (&sb)[6][5]) {
    constexpr int un = 1 + D;
    constexpr int t1 = 1 + (D + 1) % 3;
    constexpr int t2 = 1 + (D + 2) % 3;
    const double rho = 0.5 * (s[2][0] + s[3][0]);
    const double avp = 0.5 * (s[2][4] + s[3][4]);
    const double c = sound_speed(e, rho, avp);
    const double l0un = -rho / (2.0 * c);
    const double l04 = 1.0 / (2.0 * c * c);
    const double l14 = -1.0 / (c * c);
    const double l4un = rho / (2.0 * c);
    const double l44 = 1.0 / (2.0 * c * c);
    // characteristic field a of state k (rows of L, sparse)
    auto chv = [&](int a, int k) -> double {
        switch (a) {
            case 0: return l0un * s[k][un] + l04 * s[k][4];
            case 1: return s[k][0] + l14 * s[k][4];
            case 2: return s[k][t1];
            case 3: return s[k][t2];
            default: return l4un * s[k][un] + l44 * s[k][4];
        }
    };
    double cl[5], cr[5];
#pragma unroll 1
    for (int a = 0; a < 5; ++a) {
        cl[a] = weno5z_interp(chv(a, 0), chv(a, 1), chv(a, 2), chv(a, 3), chv(a, 4));
        cr[a] = weno5z_interp(chv(a, 5), chv(a, 4), chv(a, 3), chv(a, 2), chv(a, 1));
    }
    const double run0 = -c / rho, run4 = c / rho, r40 = c * c;
    double wl[5], wr[5];
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
    const bool fb = !admissible(e, wl[0], wl[4]) || !admissible(e, wr[0], wr[4]);
    if (fb) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            wl[k] = s[2][k];
            wr[k] = s[3][k];
        }
    }
    double wlb[5], wrb[5];
    hllc_vjp<D>(e, wl, wr, fbar, wlb, wrb);
    if (fb) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            sb[2][k] += wlb[k];
            sb[3][k] += wrb[k];
        }
        return;
    }
    // back-projection w = R c
    double clb[5], crb[5];
    clb[0] = wlb[0] + run0 * wlb[un] + r40 * wlb[4];
    crb[0] = wrb[0] + run0 * wrb[un] + r40 * wrb[4];
    clb[1] = wlb[0];
    crb[1] = wrb[0];
    clb[2] = wlb[t1];
    crb[2] = wrb[t1];
    clb[3] = wlb[t2];
    crb[3] = wrb[t2];
    clb[4] = wlb[0] + run4 * wlb[un] + r40 * wlb[4];
    crb[4] = wrb[0] + run4 * wrb[un] + r40 * wrb[4];
    const double run0b = wlb[un] * cl[0] + wrb[un] * cr[0];
    const double run4b = wlb[un] * cl[4] + wrb[un] * cr[4];
    const double r40b = wlb[4] * (cl[0] + cl[4]) + wrb[4] * (cr[0] + cr[4]);
    // per characteristic field: WENO5-Z (right side mirrored), then the
    // projection ch = L s
    double l0unb = 0.0, l04b = 0.0, l14b = 0.0, l4unb = 0.0, l44b = 0.0;
#pragma unroll 1
    for (int a = 0; a < 5; ++a) {
        double v[5] = {chv(a, 0), chv(a, 1), chv(a, 2), chv(a, 3), chv(a, 4)};
        double chb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        weno_vjp(v, clb[a], chb);
        double vr[5] = {chv(a, 5), v[4], v[3], v[2], v[1]};
        double vrb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        weno_vjp(vr, crb[a], vrb);
#pragma unroll
        for (int k = 0; k < 5; ++k) chb[5 - k] += vrb[k];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double g = chb[k];
            if (a == 0) {
                sb[k][un] += l0un * g;
                sb[k][4] += l04 * g;
                l0unb += g * s[k][un];
                l04b += g * s[k][4];
            } else if (a == 1) {
                sb[k][0] += g;
                sb[k][4] += l14 * g;
                l14b += g * s[k][4];
            } else if (a == 2) {
                sb[k][t1] += g;
            } else if (a == 3) {
                sb[k][t2] += g;
            } else {
                sb[k][un] += l4un * g;
                sb[k][4] += l44 * g;
                l4unb += g * s[k][un];
                l44b += g * s[k][4];
            }
        }
    }
    // eigenvector entries as functions of the average density and c
    double rhob = -l0unb / (2.0 * c) + l4unb / (2.0 * c) + run0b * c / (rho * rho) - run4b * c / (rho * rho);
    const double c3 = c * c * c;
    const double cb = l0unb * rho / (2.0 * c * c) - l4unb * rho / (2.0 * c * c) - (l04b + l44b) / c3 +
                      2.0 * l14b / c3 - run0b / rho + run4b / rho + 2.0 * c * r40b;
    double pb = 0.0;
    sound_speed_vjp(e, rho, avp, cb, rhob, pb);
    sb[2][0] += 0.5 * rhob;
    sb[3][0] += 0.5 * rhob;
    sb[2][4] += 0.5 * pb;
    sb[3][4] += 0.5 * pb;
}

// ------------------------------------------------------------------ kernel

// Pencil group `grp` (P pencils) of one axis VJP described by `a`; every
// thread of the workgroup calls it and it returns uniformly.
template <int D>
__device__ __forceinline__ void adj_group(const AdjArgs& a, long long grp, double* lds) {
    constexpr int un = 1 + D;
    constexpr int E0 = D == 0 ? 1 : 0;
    constexpr int E1 = D == 2 ? 1 : 2;
    const int n = a.n[D];
    const int C = n + 10, M = n + 5, NF = n + 1;
    const int P = a.P;
    const int NT = blockDim.x;
    const int tid = threadIdx.x;
    double* sw = lds;                 // [5][C][P] primitives
    double* sacc = sw + 5 * C * P;    // [5][C][P] cotangents (primitive, then conserved)
    double* smf = sacc + 5 * C * P;   // [5][M][P] midpoint fluxes, then face cotangents
    auto W = [&](double* arr, int q, int c, int p, int ext) -> double& { return arr[(q * ext + c) * P + p]; };
    const long long npen = a.ncell / n;
    const long long pen0 = grp * P;
    const int Pv = (int)min((long long)P, npen - pen0);
    if (Pv <= 0) return;
    const Eos e = a.eos;
    // pencil p -> base offset (transverse coordinates t0 along E0, t1 along E1)
    auto base_of = [&](int p) -> long long {
        const long long pi = pen0 + p;
        const long long t0 = pi / a.n[E1], t1 = pi % a.n[E1];
        return t0 * a.stride[E0] + t1 * a.stride[E1];
    };
    const long long sd = a.stride[D];
    // ---- A. load
    for (int it = tid; it < C * P; it += NT) {
        int c, p;
        if (sd == 1) {
            c = it % C;
            p = it / C;
        } else {
            p = it % P;
            c = it / P;
        }
        if (p >= Pv) continue;
        const long long base = base_of(p);
        auto ld = [&](int x, double* o) {
#pragma unroll
            for (int q = 0; q < 5; ++q) o[q] = a.U[q * a.ncell + base + (long long)x * sd];
        };
        const int x = c - 5;
        double u[5], w[5];
        if (x >= 0 && x < n)
            ld(x, u);
        else
            bc_ghost<D>(e, n, x, a.bc_kind, a.bc_state, ld, u);
        to_prim(e, u, w);
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            W(sw, q, c, p, C) = w[q];
            W(sacc, q, c, p, C) = 0.0;
        }
    }
    __syncthreads();
    // ---- B. forward midpoint fluxes (midpoint m between cells m+2 and m+3)
    for (int it = tid; it < M * P; it += NT) {
        const int p = it % P, m = it / P;
        if (p >= Pv) continue;
        double s[6][5];
#pragma unroll
        for (int j = 0; j < 6; ++j)
#pragma unroll
            for (int q = 0; q < 5; ++q) s[j][q] = W(sw, q, m + j, p, C);
        double f[5];
        midpoint_flux<D>(e, s, f);
#pragma unroll
        for (int q = 0; q < 5; ++q) W(smf, q, m, p, M) = f[q];
    }
    __syncthreads();
    // ---- C. faces in chunks of increasing f: a chunk only overwrites the
    // midpoint slots of its own faces, which later chunks do not read
    const int fchunk = NT / P;
    for (int f0 = 0; f0 < NF; f0 += fchunk) {
        const int p = tid % P, f = f0 + tid / P;
        const bool live = tid < fchunk * P && p < Pv && f < NF;
        double fhb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        double wlb[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, wrb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (live) {
            double fo[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                double v = 0.0;
                v += S6_0 * W(smf, k, f + 2, p, M);
                v += S6_1 * W(smf, k, f + 3, p, M);
                v += S6_1 * W(smf, k, f + 1, p, M);
                v += S6_2 * W(smf, k, f + 4, p, M);
                v += S6_2 * W(smf, k, f, p, M);
                fo[k] = v;
            }
            // face cotangent from the divergence
            const long long base = base_of(p);
            double Fb[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const double lr = f < n ? a.LAM[k * a.ncell + base + (long long)f * sd] : 0.0;
                const double ll = f > 0 ? a.LAM[k * a.ncell + base + (long long)(f - 1) * sd] : 0.0;
                Fb[k] = (lr - ll) * a.inv_dx;
            }
            bool limited = false;
            double lo = 1.0, hi = 1.0, al = 0.0;
            double wl[5], wr[5], ul[5], ur[5], fl[5], fr[5], flf[5];
            if (a.lam_lim > 0.0) {
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    wl[k] = W(sw, k, f + 4, p, C);
                    wr[k] = W(sw, k, f + 5, p, C);
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
                    lo = l0;
                    hi = h0;
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
                    fhb[k] = fk ? lo * Fb[k] : 0.0;
                    flfb[k] = fk ? (1.0 - lo) * Fb[k] : Fb[k];
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
                const int which = lo > 0.0 ? limiter_constraint(e, ul, ur, fl, fr, fo, flf, tl, hi) : -1;
                if (which >= 0) {
                    const bool left = which < 3;  // test state ul - tl (fx - fl), else ur + tl (fx - fr)
                    const double sg = left ? -tl : tl;
                    double st[5], ws[5], gw[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gs[5];
#pragma unroll
                    for (int k = 0; k < 5; ++k) {
                        const double fx = isfinite(fo[k]) ? lo * fo[k] + (1.0 - lo) * flf[k] : flf[k];
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
                                fhb[k] += lo * fxb;
                                flfb[k] += (1.0 - lo) * fxb;
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
        // so the sums do not depend on the chunking (on P)
        __syncthreads();
        if (live) {
#pragma unroll
            for (int k = 0; k < 5; ++k) W(smf, k, f, p, M) = fhb[k];
#pragma unroll
            for (int k = 0; k < 5; ++k) W(sacc, k, f + 5, p, C) += wrb[k];
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int k = 0; k < 5; ++k) W(sacc, k, f + 4, p, C) += wlb[k];
        }
        __syncthreads();
    }
    // ---- D. midpoint cotangents and VJPs (chunks; six scatter sub-phases)
    for (int m0 = 0; m0 < M * P; m0 += NT) {
        const int it = m0 + tid;
        const int p = it % P, m = it / P;
        const bool live = it < M * P && p < Pv;
        double sb[6][5];
#pragma unroll
        for (int j = 0; j < 6; ++j)
#pragma unroll
            for (int q = 0; q < 5; ++q) sb[j][q] = 0.0;
        if (live) {
            // faces f = m-4 .. m use midpoint m with weight S6[|m - f - 2|]
            double mb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            const double w6[5] = {S6_2, S6_1, S6_0, S6_1, S6_2};
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                const int f = m - 4 + j;
                if (f < 0 || f > n) continue;
#pragma unroll
                for (int q = 0; q < 5; ++q) mb[q] += w6[4 - j] * W(smf, q, f, p, M);
            }
            double s[6][5];
#pragma unroll
            for (int j = 0; j < 6; ++j)
#pragma unroll
                for (int q = 0; q < 5; ++q) s[j][q] = W(sw, q, m + j, p, C);
            midpoint_vjp<D>(e, s, mb, sb);
        }
        // cell c takes midpoints c-5 .. c in increasing order (j from 5 down
        // to 0), within a chunk as across chunks: independent of P
        for (int j = 5; j >= 0; --j) {
            if (live) {
#pragma unroll
                for (int q = 0; q < 5; ++q) W(sacc, q, m + j, p, C) += sb[j][q];
            }
            __syncthreads();
        }
    }
    // ---- E. conserved cotangents of every loaded cell
    for (int it = tid; it < C * P; it += NT) {
        const int p = it % P, c = it / P;
        if (p >= Pv) continue;
        double w[5], wb[5], ub[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            w[q] = W(sw, q, c, p, C);
            wb[q] = W(sacc, q, c, p, C);
        }
        to_prim_vjp(e, w, wb, ub);
#pragma unroll
        for (int q = 0; q < 5; ++q) W(sacc, q, c, p, C) = ub[q];
    }
    __syncthreads();
    // interior cells: own cotangent plus the ghosts copied from them
    for (int it = tid; it < n * P; it += NT) {
        int x, p;
        if (sd == 1) {
            x = it % n;
            p = it / n;
        } else {
            p = it % P;
            x = it / P;
        }
        if (p >= Pv) continue;
        double tot[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) tot[q] = W(sacc, q, x + 5, p, C);
        for (int gi = 0; gi < 10; ++gi) {
            const int gx = gi < 5 ? gi - 5 : n + gi - 5;
            const int side = gx < 0 ? 0 : 1;
            const int kind = a.bc_kind[side];
            int src = -1;
            double flip = 1.0;
            if (kind == 0)
                src = ((gx % n) + n) % n;
            else if (kind == 1)
                src = side == 0 ? 0 : n - 1;
            else if (kind == 2) {
                src = side == 0 ? -1 - gx : 2 * n - 1 - gx;
                flip = -1.0;
            }
            if (src != x) continue;
#pragma unroll
            for (int q = 0; q < 5; ++q) tot[q] += (q == un ? flip : 1.0) * W(sacc, q, gx + 5, p, C);
        }
        const long long cell = base_of(p) + (long long)x * sd;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const long long ix = q * a.ncell + cell;
            a.MU[ix] = (a.first ? 0.0 : a.MU[ix]) + tot[q];
        }
    }
}

template <int D>
__global__ __launch_bounds__(256) void adj_pass(AdjArgs a) {
    extern __shared__ double lds[];
    adj_group<D>(a, blockIdx.x, lds);
}

// Ensemble VJP: the pencil groups of every instance in one persistent
// launch (instance i owns work items prefix[i] .. prefix[i+1]-1).
template <int D>
__global__ __launch_bounds__(256) void ens_adj_pass(EnsAdjArgs e) {
    extern __shared__ double lds[];
    for (long long w = blockIdx.x; w < e.total; w += gridDim.x) {
        int lo = 0, hi = e.ninst;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (e.prefix[mid] <= w)
                lo = mid;
            else
                hi = mid;
        }
        adj_group<D>(e.args[lo], w - e.prefix[lo], lds);
        __syncthreads();
    }
}

}  // namespace adj

hipError_t launch_ens_adj_pass(int d, const EnsAdjArgs& e, size_t lds, unsigned grid, hipStream_t s) {
    if (d == 0)
        hipLaunchKernelGGL(adj::ens_adj_pass<0>, dim3(grid), dim3(256), lds, s, e);
    else if (d == 1)
        hipLaunchKernelGGL(adj::ens_adj_pass<1>, dim3(grid), dim3(256), lds, s, e);
    else
        hipLaunchKernelGGL(adj::ens_adj_pass<2>, dim3(grid), dim3(256), lds, s, e);
    return hipGetLastError();
}

int adj_pencils_per_group(int n) {
    const long long per = 8ll * (15ll * (n + 10));  // bytes of LDS per pencil (upper bound)
    int P = (int)std::max(1ll, std::min(8ll, (64ll << 10) / per));
    return P;
}

size_t adj_lds_bytes(int n, int P) { return sizeof(double) * (size_t)P * (10 * (size_t)(n + 10) + 5 * (size_t)(n + 5)); }

hipError_t launch_adj_pass(int d, const AdjArgs& a, hipStream_t s) {
    const int n = a.n[d];
    const long long npen = a.ncell / n;
    const unsigned grid = (unsigned)((npen + a.P - 1) / a.P);
    const size_t lds = adj_lds_bytes(n, a.P);
    if (d == 0)
        hipLaunchKernelGGL(adj::adj_pass<0>, dim3(grid), dim3(256), lds, s, a);
    else if (d == 1)
        hipLaunchKernelGGL(adj::adj_pass<1>, dim3(grid), dim3(256), lds, s, a);
    else
        hipLaunchKernelGGL(adj::adj_pass<2>, dim3(grid), dim3(256), lds, s, a);
    return hipGetLastError();
}

// SSP-RK3 adjoint combinations (elementwise over 5 * ncell values), for the
// forward u1 = u0 + dt L(u0); u2 = 3/4 u0 + 1/4 (u1 + dt L(u1));
// u3 = u0 / 3 + 2/3 (u2 + dt L(u2)), with T = L'(u_s)^T (cotangent):
//   stage 3: B = 2/3 A + 2/3 dt T,  Z = A / 3
//   stage 2: Z += 3/4 B,  B = 1/4 B + 1/4 dt T
//   stage 1: Z += B + dt T
__global__ void k_adj_combine(int stage, long long nv, double dt, const double* __restrict__ A,
                              const double* __restrict__ T, double* __restrict__ B, double* __restrict__ Z) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < nv; i += (long long)gridDim.x * 256ll) {
        if (stage == 3) {
            B[i] = (2.0 / 3.0) * A[i] + (2.0 / 3.0) * dt * T[i];
            Z[i] = A[i] / 3.0;
        } else if (stage == 2) {
            const double b = B[i];
            Z[i] = Z[i] + 0.75 * b;
            B[i] = 0.25 * b + 0.25 * dt * T[i];
        } else {
            Z[i] = Z[i] + B[i] + dt * T[i];
        }
    }
}

// Ensemble form: instance blockIdx.y with its own arrays and the dt of the
// step from the device dt history; stage 1 leaves the step's input
// cotangent in A (A is free once stage 3 has read it).
__global__ void k_ens_adj_combine(int stage, const AdjCombineTask* __restrict__ tasks, const double* __restrict__ dts) {
    const AdjCombineTask t = tasks[blockIdx.y];
    const double dt = dts[t.dt_index];
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < t.nv; i += (long long)gridDim.x * 256ll) {
        if (stage == 3) {
            t.B[i] = (2.0 / 3.0) * t.A[i] + (2.0 / 3.0) * dt * t.T[i];
            t.Z[i] = t.A[i] / 3.0;
        } else if (stage == 2) {
            const double b = t.B[i];
            t.Z[i] = t.Z[i] + 0.75 * b;
            t.B[i] = 0.25 * b + 0.25 * dt * t.T[i];
        } else {
            t.A[i] = t.Z[i] + t.B[i] + dt * t.T[i];  // the cotangent of the step's input
        }
    }
}

hipError_t launch_ens_adj_combine(int stage, const AdjCombineTask* tasks, int ninst, const double* dts,
                                  long long maxnv, hipStream_t s) {
    const unsigned gx = (unsigned)std::min<long long>((maxnv + 255) / 256, 1024);
    hipLaunchKernelGGL(k_ens_adj_combine, dim3(gx, (unsigned)ninst), dim3(256), 0, s, stage, tasks, dts);
    return hipGetLastError();
}

hipError_t launch_adj_combine(int stage, long long nv, double dt, const double* A, const double* T, double* B,
                              double* Z, hipStream_t s) {
    const unsigned g = (unsigned)std::min<long long>((nv + 255) / 256, 8192);
    hipLaunchKernelGGL(k_adj_combine, dim3(g), dim3(256), 0, s, stage, nv, dt, A, T, B, Z);
    return hipGetLastError();
}

}  // namespace em
