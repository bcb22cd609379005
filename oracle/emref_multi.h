// Oracle systems on the generic scheme and solver of emref_gen.h: the
// two-phase Kapila five-equation system (FiveEq, N = 7) and the Euler system
// in the same generic form.  Test infrastructure only: the product library
// never links it.
//
// Restated from fluid/crates/compressible/src/multiphase.rs:1-392 (FiveEq)
// and system.rs:190-414 (Euler through the System<5> trait).  Expressions
// keep the reference's evaluation order; no contraction, no fast-math.
#pragma once
#include <cmath>
#include <limits>

#include "emref.h"
#include "emref_gen.h"

namespace emref {
namespace gen {

// multiphase.rs:29-392.  Conserved (a1, a2, rho u, rho v, rho w, E, alpha1);
// primitive (a1, a2, u, v, w, p, alpha1), a_k = alpha_k rho_k.
struct FiveEq {
    static constexpr int N = 7;
    static constexpr int IA = 6, IP = 5, IU = 2;
    using V = Vn<7>;
    using M = Mn<7>;
    Nasg eos[2];
    bool kapila = true;

    // 63-75: D, B of rho e = p D + B
    void db(const V& w, double& d, double& b) const {
        const double al[2] = {w[IA], 1.0 - w[IA]};
        d = 0.0;
        b = 0.0;
        for (int k = 0; k < 2; ++k) {
            const Nasg& e = eos[k];
            const double vol = al[k] - e.b * w[k];
            d += vol / (e.gamma - 1.0);
            b += e.gamma * e.p_inf * vol / (e.gamma - 1.0) + e.q * w[k];
        }
    }
    double rho_e(const V& w) const {  // 77-81
        double d, b;
        db(w, d, b);
        return w[IP] * d + b;
    }
    double rho_c2_phase(const V& w, int k) const {  // 83-90
        const double al[2] = {w[IA], 1.0 - w[IA]};
        const Nasg& e = eos[k];
        const double rho_k = (al[k] > 1e-12 && w[k] > 0.0) ? w[k] / al[k] : 0.0;
        return e.gamma * (w[IP] + e.p_inf) / fmax_(1.0 - e.b * rho_k, 1e-12);
    }
    double kapila_k(const V& w) const {  // 92-104
        if (!kapila) return 0.0;
        const double a = clamp_(w[IA], 0.0, 1.0);
        const double al[2] = {a, 1.0 - clamp_(w[IA], 0.0, 1.0)};
        const double z1 = rho_c2_phase(w, 0), z2 = rho_c2_phase(w, 1);
        const double den = al[1] * z1 + al[0] * z2;
        if (den <= 0.0) return 0.0;
        return al[0] * al[1] * (z2 - z1) / den;
    }
    double drhoe_dalpha(double p) const {  // 120-124
        const Nasg &e1 = eos[0], &e2 = eos[1];
        return (p + e1.gamma * e1.p_inf) / (e1.gamma - 1.0) - (p + e2.gamma * e2.p_inf) / (e2.gamma - 1.0);
    }
    double rho_c2(const V& w) const {  // 106-118
        double d, b;
        db(w, d, b);
        const double p = w[IP];
        double num = rho_e(w) + p;
        for (int k = 0; k < 2; ++k) {
            const Nasg& e = eos[k];
            num += w[k] * (e.b * (p + e.gamma * e.p_inf) / (e.gamma - 1.0) - e.q);
        }
        num += kapila_k(w) * drhoe_dalpha(p);
        return num / d;
    }
    static double rho(const V& w) { return w[0] + w[1]; }  // 126-130
    double p_inf_mix(const V& w) const {                   // 132-136
        const double a = clamp_(w[IA], 0.0, 1.0);
        return a * eos[0].p_inf + (1.0 - a) * eos[1].p_inf;
    }

    V to_prim(const V& u) const {  // 140-156
        const double r = u[0] + u[1];
        const double vx = u[2] / r, vy = u[3] / r, vz = u[4] / r;
        V w{};
        w[0] = u[0];
        w[1] = u[1];
        w[2] = vx;
        w[3] = vy;
        w[4] = vz;
        w[IA] = u[6];
        const double rhoe = u[5] - 0.5 * r * (vx * vx + vy * vy + vz * vz);
        double d, b;
        db(w, d, b);
        w[IP] = (rhoe - b) / d;
        return w;
    }
    double velocity_conserved(const V& u, int d) const { return u[2 + d] / (u[0] + u[1]); }  // 158-161
    V to_cons(const V& w) const {                                                           // 162-173
        const double r = w[0] + w[1];
        V u{};
        u[0] = w[0];
        u[1] = w[1];
        u[2] = r * w[2];
        u[3] = r * w[3];
        u[4] = r * w[4];
        u[5] = rho_e(w) + 0.5 * r * (w[2] * w[2] + w[3] * w[3] + w[4] * w[4]);
        u[6] = w[IA];
        return u;
    }
    double sound_speed(const V& w) const { return std::sqrt(fmax_(rho_c2(w), 0.0) / rho(w)); }  // 175-177
    double max_speed(const V& w, int d) const {  // trait default, system.rs:185-187
        return std::fabs(w[IU + d]) + sound_speed(w);
    }
    int vel_index() const { return IU; }
    int momentum_index(int d) const { return 2 + d; }

    void eigen(const V& w, int d, M& l, M& r) const {  // 219-256
        const double rh = rho(w);
        const double c = sound_speed(w);
        const double kk = kapila_k(w);
        const int un = IU + d;
        const int t1 = IU + (d + 1) % 3, t2 = IU + (d + 2) % 3;
        for (auto& row : l) row.fill(0.0);
        for (auto& row : r) row.fill(0.0);
        const double c2 = c * c;
        l[0][0] = 1.0;
        l[0][IP] = -w[0] / rh / c2;
        l[1][1] = 1.0;
        l[1][IP] = -w[1] / rh / c2;
        l[2][un] = -rh / (2.0 * c);
        l[2][IP] = 1.0 / (2.0 * c2);
        l[3][t1] = 1.0;
        l[4][t2] = 1.0;
        l[5][un] = rh / (2.0 * c);
        l[5][IP] = 1.0 / (2.0 * c2);
        l[6][IA] = 1.0;
        l[6][IP] = kk / rh / c2;
        r[0][0] = 1.0;
        r[1][1] = 1.0;
        r[t1][3] = 1.0;
        r[t2][4] = 1.0;
        r[IA][6] = 1.0;
        const int cols[2] = {2, 5};
        const double sgns[2] = {-1.0, 1.0};
        for (int q = 0; q < 2; ++q) {
            const int col = cols[q];
            const double sgn = sgns[q];
            r[0][col] = w[0] / rh;
            r[1][col] = w[1] / rh;
            r[un][col] = sgn * c / rh;
            r[IP][col] = c2;
            r[IA][col] = -kk / rh;
        }
    }
    V eigenvalues(const V& w, int d) const {  // 258-263
        const double u = w[IU + d], c = sound_speed(w);
        return V{u, u, u - c, u, u, u + c, u};
    }
    V flux(const V& w, int d) const {  // 319-331
        const int un = IU + d;
        const V q = to_cons(w);
        const double u = w[un];
        V f{};
        for (int k = 0; k < 5; ++k) f[k] = q[k] * u;
        f[un] += w[IP];
        f[5] = (q[5] + w[IP]) * u;
        f[6] = w[IA] * u;
        return f;
    }
    V riemann_flux(const V& wl, const V& wr, int d, double* s_star) const {  // 265-310 (HLLC)
        const int un = IU + d;
        const double cl = sound_speed(wl), cr = sound_speed(wr);
        const double ul = wl[un], ur = wr[un];
        const double sl = fmin_(ul - cl, ur - cr);
        const double sr = fmax_(ul + cl, ur + cr);
        const V ql = to_cons(wl), qr = to_cons(wr);
        if (sl >= 0.0) {
            if (s_star) *s_star = ul;
            return flux(wl, d);
        }
        if (sr <= 0.0) {
            if (s_star) *s_star = ur;
            return flux(wr, d);
        }
        const double rl = rho(wl), rr = rho(wr);
        const double pl = wl[IP], pr = wr[IP];
        const double ss = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur));
        auto star = [&](const V& w, const V& q, double s) {
            const double rh = rho(w), u = w[un];
            const double fac = (s - u) / (s - ss);
            V qs{};
            qs[0] = w[0] * fac;
            qs[1] = w[1] * fac;
            for (int k = IU; k < IU + 3; ++k) qs[k] = rh * fac * w[k];
            qs[un] = rh * fac * ss;
            qs[5] = rh * fac * (q[5] / rh + (ss - u) * (ss + w[IP] / (rh * (s - u))));
            qs[6] = w[IA] * fac;
            return qs;
        };
        V f{};
        if (ss >= 0.0) {
            const V fl = flux(wl, d), qs = star(wl, ql, sl);
            for (int k = 0; k < 7; ++k) f[k] = fl[k] + sl * (qs[k] - ql[k]);
        } else {
            const V fr = flux(wr, d), qs = star(wr, qr, sr);
            for (int k = 0; k < 7; ++k) f[k] = fr[k] + sr * (qs[k] - qr[k]);
        }
        if (s_star) *s_star = ss;
        return f;
    }
    V split_flux(const V& a, const V& b, int d) const {  // 312-331 (KEEP-PE pair flux)
        const int un = IU + d;
        const double u = 0.5 * (a[un] + b[un]);
        const double rh = 0.5 * (rho(a) + rho(b));
        const double c = rh * u;
        V f{};
        f[0] = 0.5 * (a[0] + b[0]) * u;
        f[1] = 0.5 * (a[1] + b[1]) * u;
        for (int k = IU; k < IU + 3; ++k) f[k] = c * 0.5 * (a[k] + b[k]);
        f[un] += 0.5 * (a[IP] + b[IP]);
        const double kin = c * 0.5 * (a[2] * b[2] + a[3] * b[3] + a[4] * b[4]);
        const double rhoe = 0.5 * (rho_e(a) + rho_e(b));
        const double pw = 0.5 * (a[un] * b[IP] + b[un] * a[IP]);
        f[5] = kin + u * rhoe + pw;
        f[6] = 0.5 * (a[IA] + b[IA]) * u;
        return f;
    }
    bool admissible(const V& w) const {  // 358-363
        for (int k = 0; k < 7; ++k)
            if (!std::isfinite(w[k])) return false;
        return rho(w) > 0.0 && (w[IA] >= -0.05 && w[IA] <= 1.05) && rho_c2(w) > 0.0;
    }
    int nc_var() const { return 6; }                                   // 365-367
    double nc_coef(const V& w) const { return w[IA] + kapila_k(w); }  // 369-371
    double dilatation_rate(const V& w, double divu) const {           // 373-391
        const double a = w[IA];
        double pmin_inf = std::numeric_limits<double>::infinity();
        const double als[2] = {a, 1.0 - a};
        for (int k = 0; k < 2; ++k)
            if (als[k] > 1e-6) pmin_inf = fmin_(pmin_inf, eos[k].p_inf);
        const double floor = 1e-6 * fmax_(eos[0].p_inf, eos[1].p_inf);
        const double head = fmax_(fmax_(w[IP] + pmin_inf, floor), 1e-300);
        return fmax_(rho_c2(w), 0.0) * std::fabs(divu) / head;
    }
    int sensor(const V& w, double* out) const {  // 385-390
        out[0] = std::log(fmax_(rho(w), 1e-300));
        out[1] = std::log(fmax_(w[IP] + p_inf_mix(w), 1e-300));
        out[2] = w[IA];
        return 3;
    }
};

// Euler (system.rs:190-414) through the generic trait surface
struct EulerG {
    static constexpr int N = 5;
    using V = Vn<5>;
    using M = Mn<5>;
    Euler e;
    V to_prim(const V& u) const { return e.to_prim(u); }
    V to_cons(const V& w) const { return e.to_cons(w); }
    double sound_speed(const V& w) const { return e.sound_speed(w); }
    double max_speed(const V& w, int d) const { return e.max_speed(w, d); }
    void eigen(const V& w, int d, M& l, M& r) const { e.eigen(w, d, l, r); }
    V eigenvalues(const V& w, int d) const { return e.eigenvalues(w, d); }
    V riemann_flux(const V& wl, const V& wr, int d, double* ss) const { return e.riemann_flux(wl, wr, d, ss); }
    V split_flux(const V& a, const V& b, int d) const { return e.split_flux(a, b, d); }
    V flux(const V& w, int d) const { return e.flux(w, d); }
    bool admissible(const V& w) const { return e.admissible(w); }
    int vel_index() const { return 1; }
    int momentum_index(int d) const { return 1 + d; }
    int nc_var() const { return -1; }
    double nc_coef(const V&) const { return 0.0; }
    double dilatation_rate(const V&, double) const { return 0.0; }
    double velocity_conserved(const V& u, int d) const { return to_prim(u)[1 + d]; }
    int sensor(const V& w, double* out) const {  // system.rs:409-413
        out[0] = std::log(fmax_(w[0], 1e-300));
        out[1] = std::log(fmax_(w[4] + e.eos.p_inf, 1e-300));
        return 2;
    }
};

// Initial state of the reference's `interface` probe
// (examples/cpu_validate.rs:100-126): a water drop (alpha = 1) in air,
// smeared over one cell by tanh, moving at u = 0.1 in p = 1e5.
inline void interface_init(SolverG<FiveEq>& s) {
    const Mesh& m = s.mesh;
    const double nn = (double)m.n[0];
    for (size_t i = 0; i < m.n[0]; ++i)
        for (size_t j = 0; j < m.n[1]; ++j)
            for (size_t k = 0; k < m.n[2]; ++k) {
                const double x[3] = {m.x(i, 0), m.x(j, 1), m.x(k, 2)};
                const double r = std::sqrt((x[0] - 0.5) * (x[0] - 0.5) + (x[1] - 0.5) * (x[1] - 0.5) +
                                           (x[2] - 0.5) * (x[2] - 0.5));
                const double alpha = clamp_(0.5 * (1.0 + std::tanh((0.25 - r) / (1.0 / nn))), 1e-6, 1.0 - 1e-6);
                const Vn<7> w{alpha * 1000.0, (1.0 - alpha) * 1.2, 0.1, 0.0, 0.0, 1e5, alpha};
                s.u[m.idx(i + m.g(0), j + m.g(1), k + m.g(2))] = s.sys.to_cons(w);
            }
}

}  // namespace gen
}  // namespace emref
