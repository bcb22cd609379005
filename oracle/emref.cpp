// Oracle implementation (test infrastructure only; see emref.h).
// Every function cites the reference lines it restates.  Expressions keep the
// reference's evaluation order; `-0.0` seeds reproduce Rust's float `Sum`.
#include "emref.h"

#include <cmath>
#include <cstring>
#include <stdexcept>

namespace emref {

static inline double rmax(double a, double b) { return std::fmax(a, b); }  // f64::max
static inline double rmin(double a, double b) { return std::fmin(a, b); }  // f64::min
static inline double rclamp(double x, double lo, double hi) {               // f64::clamp
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

// ---------------------------------------------------------------- eos.rs:36-79
double Nasg::pressure(double rho, double e) const {
    double v = 1.0 / rho;
    return (gamma - 1.0) * (e - q) / (v - b) - gamma * p_inf;
}
double Nasg::energy(double rho, double p) const {
    double v = 1.0 / rho;
    return (p + gamma * p_inf) * (v - b) / (gamma - 1.0) + q;
}
double Nasg::rho_e(double rho, double p) const { return rho * energy(rho, p); }
double Nasg::sound_speed_sq(double rho, double p) const {
    double v = 1.0 / rho;
    return gamma * (p + p_inf) * v * v / (v - b);
}
bool Nasg::admissible(double rho, double p) const {
    return rho > 0.0 && 1.0 / rho > b && p + p_inf > 0.0;
}

// ------------------------------------------------------------- system.rs:197-407
V Euler::to_prim(const V& u) const {  // 197-203
    double rho = u[0];
    double vx = u[1] / rho, vy = u[2] / rho, vz = u[3] / rho;
    double e = u[4] / rho - 0.5 * (vx * vx + vy * vy + vz * vz);
    return V{rho, vx, vy, vz, eos.pressure(rho, e)};
}
V Euler::to_cons(const V& w) const {  // 210-217
    double rho = w[0];
    double ke = 0.5 * rho * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
    return V{rho, rho * w[1], rho * w[2], rho * w[3], eos.rho_e(rho, w[4]) + ke};
}
double Euler::sound_speed(const V& w) const {  // 219-222
    return std::sqrt(rmax(eos.sound_speed_sq(w[0], w[4]), 0.0));
}
double Euler::max_speed(const V& w, int d) const {  // trait default 185-187
    return std::fabs(w[1 + d]) + sound_speed(w);
}
void Euler::eigen(const V& w, int d, Mat& l, Mat& r) const {  // 232-258
    double rho = w[0], c = sound_speed(w);
    int un = 1 + d;
    int t1 = 1 + (d + 1) % 3, t2 = 1 + (d + 2) % 3;
    for (auto& row : l) row.fill(0.0);
    for (auto& row : r) row.fill(0.0);
    l[0][un] = -rho / (2.0 * c);
    l[0][4] = 1.0 / (2.0 * c * c);
    l[1][0] = 1.0;
    l[1][4] = -1.0 / (c * c);
    l[2][t1] = 1.0;
    l[3][t2] = 1.0;
    l[4][un] = rho / (2.0 * c);
    l[4][4] = 1.0 / (2.0 * c * c);
    r[0][0] = 1.0;
    r[un][0] = -c / rho;
    r[4][0] = c * c;
    r[0][1] = 1.0;
    r[t1][2] = 1.0;
    r[t2][3] = 1.0;
    r[0][4] = 1.0;
    r[un][4] = c / rho;
    r[4][4] = c * c;
}

static inline V phys_flux_wq(const V& w, const V& q, int un) {  // closure in 268-279
    double u = w[un];
    V f{};
    f[0] = q[0] * u;
    f[1] = q[1] * u;
    f[2] = q[2] * u;
    f[3] = q[3] * u;
    f[un] += w[4];
    f[4] = (q[4] + w[4]) * u;
    return f;
}

V Euler::riemann_flux(const V& wl, const V& wr, int d, double* s_star) const {  // 260-312
    int un = 1 + d;
    double cl = sound_speed(wl), cr = sound_speed(wr);
    double ul = wl[un], ur = wr[un];
    double sl = rmin(ul - cl, ur - cr);
    double sr = rmax(ul + cl, ur + cr);
    V ql = to_cons(wl), qr = to_cons(wr);
    if (sl >= 0.0) {
        if (s_star) *s_star = ul;
        return phys_flux_wq(wl, ql, un);
    }
    if (sr <= 0.0) {
        if (s_star) *s_star = ur;
        return phys_flux_wq(wr, qr, un);
    }
    double rl = wl[0], rr = wr[0], pl = wl[4], pr = wr[4];
    double ss = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur));
    auto star = [&](const V& w, const V& q, double s) {
        double rho = w[0], u = w[un];
        double fac = rho * (s - u) / (s - ss);
        V qs{};
        qs[0] = fac;
        for (int k = 1; k < 4; ++k) qs[k] = fac * w[k];
        qs[un] = fac * ss;
        qs[4] = fac * (q[4] / rho + (ss - u) * (ss + w[4] / (rho * (s - u))));
        return qs;
    };
    V f{};
    if (ss >= 0.0) {
        V fl = phys_flux_wq(wl, ql, un), qs = star(wl, ql, sl);
        for (int k = 0; k < 5; ++k) f[k] = fl[k] + sl * (qs[k] - ql[k]);
    } else {
        V fr = phys_flux_wq(wr, qr, un), qs = star(wr, qr, sr);
        for (int k = 0; k < 5; ++k) f[k] = fr[k] + sr * (qs[k] - qr[k]);
    }
    if (s_star) *s_star = ss;
    return f;
}

V Euler::split_flux(const V& a, const V& b, int d) const {  // 317-334
    int un = 1 + d;
    double rho = 0.5 * (a[0] + b[0]);
    double u = 0.5 * (a[un] + b[un]);
    double c = rho * u;
    V f{};
    f[0] = c;
    for (int k = 1; k < 4; ++k) f[k] = c * 0.5 * (a[k] + b[k]);
    f[un] += 0.5 * (a[4] + b[4]);
    double kin = c * 0.5 * (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
    double rhoe = 0.5 * (eos.rho_e(a[0], a[4]) + eos.rho_e(b[0], b[4]));
    double pw = 0.5 * (a[un] * b[4] + b[un] * a[4]);
    f[4] = kin + u * rhoe + pw;
    return f;
}

V Euler::flux(const V& w, int d) const {  // 336-348
    int un = 1 + d;
    V q = to_cons(w);
    double u = w[un];
    V f{};
    for (int k = 0; k < 4; ++k) f[k] = q[k] * u;
    f[un] += w[4];
    f[4] = (q[4] + w[4]) * u;
    return f;
}

V Euler::eigenvalues(const V& w, int d) const {  // 379-384
    double u = w[1 + d], c = sound_speed(w);
    return V{u - c, u, u, u, u + c};
}

Mat Euler::dudw(const V& w) const {  // 386-402
    double rho = w[0], p = w[4];
    Mat j{};
    for (auto& row : j) row.fill(0.0);
    j[0][0] = 1.0;
    for (int k = 1; k < 4; ++k) {
        j[k][0] = w[k];
        j[k][k] = rho;
        j[4][k] = rho * w[k];
    }
    double drhoe_drho = -eos.b * (p + eos.gamma * eos.p_inf) / (eos.gamma - 1.0) + eos.q;
    double drhoe_dp = (1.0 - eos.b * rho) / (eos.gamma - 1.0);
    j[4][0] = drhoe_drho + 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
    j[4][4] = drhoe_dp;
    return j;
}

bool Euler::admissible(const V& w) const {  // 404-407
    return std::isfinite(w[0]) && std::isfinite(w[4]) && eos.admissible(w[0], w[4]);
}

// system.rs:14-28
static Mat matmul(const Mat& a, const Mat& b) {
    Mat c;
    for (auto& row : c) row.fill(0.0);
    for (int i = 0; i < NV; ++i)
        for (int k = 0; k < NV; ++k) {
            double aik = a[i][k];
            if (aik != 0.0)
                for (int j = 0; j < NV; ++j) c[i][j] += aik * b[k][j];
        }
    return c;
}

// system.rs:31-61 (Gauss-Jordan, partial pivoting; ties pick the LAST row,
// as Iterator::max_by does)
static Mat invert(const Mat& a) {
    Mat m = a, inv;
    for (auto& row : inv) row.fill(0.0);
    for (int i = 0; i < NV; ++i) inv[i][i] = 1.0;
    for (int c = 0; c < NV; ++c) {
        int piv = c;
        for (int x = c + 1; x < NV; ++x)
            if (!(std::fabs(m[x][c]) < std::fabs(m[piv][c]))) piv = x;  // >= keeps the last max
        std::swap(m[c], m[piv]);
        std::swap(inv[c], inv[piv]);
        double d = 1.0 / m[c][c];
        for (int j = 0; j < NV; ++j) {
            m[c][j] *= d;
            inv[c][j] *= d;
        }
        for (int r = 0; r < NV; ++r) {
            if (r != c) {
                double f = m[r][c];
                if (f != 0.0)
                    for (int j = 0; j < NV; ++j) {
                        m[r][j] -= f * m[c][j];
                        inv[r][j] -= f * inv[c][j];
                    }
            }
        }
    }
    return inv;
}

// ------------------------------------------------------------ scheme.rs:55-99
static const double EPS = 1e-40;

static inline double weno_z_combine(const double q[3], const double dw[3], double v0, double v1,
                                    double v2, double v3, double v4) {  // 77-86
    double t0 = v0 - 2.0 * v1 + v2, s0 = v0 - 4.0 * v1 + 3.0 * v2;
    double t1 = v1 - 2.0 * v2 + v3, s1 = v1 - v3;
    double t2 = v2 - 2.0 * v3 + v4, s2 = 3.0 * v2 - 4.0 * v3 + v4;
    double b0 = 13.0 / 12.0 * (t0 * t0) + 0.25 * (s0 * s0);
    double b1 = 13.0 / 12.0 * (t1 * t1) + 0.25 * (s1 * s1);
    double b2 = 13.0 / 12.0 * (t2 * t2) + 0.25 * (s2 * s2);
    double tau = std::fabs(b0 - b2);
    double r0 = tau / (b0 + EPS), r1 = tau / (b1 + EPS), r2 = tau / (b2 + EPS);
    double a0 = dw[0] * (1.0 + r0 * r0);
    double a1 = dw[1] * (1.0 + r1 * r1);
    double a2 = dw[2] * (1.0 + r2 * r2);
    return (a0 * q[0] + a1 * q[1] + a2 * q[2]) / (a0 + a1 + a2);
}

double weno5z(double v0, double v1, double v2, double v3, double v4) {  // 60-65
    double q[3] = {(2.0 * v0 - 7.0 * v1 + 11.0 * v2) / 6.0, (-v1 + 5.0 * v2 + 2.0 * v3) / 6.0,
                   (2.0 * v2 + 5.0 * v3 - v4) / 6.0};
    static const double d[3] = {0.1, 0.6, 0.3};
    return weno_z_combine(q, d, v0, v1, v2, v3, v4);
}

double weno5z_interp(double v0, double v1, double v2, double v3, double v4) {  // 69-74
    double q[3] = {(3.0 * v0 - 10.0 * v1 + 15.0 * v2) / 8.0, (-v1 + 6.0 * v2 + 3.0 * v3) / 8.0,
                   (3.0 * v2 + 6.0 * v3 - v4) / 8.0};
    static const double d[3] = {1.0 / 16.0, 10.0 / 16.0, 5.0 / 16.0};
    return weno_z_combine(q, d, v0, v1, v2, v3, v4);
}

static const double A6[3] = {0.75, -0.15, 1.0 / 60.0};  // 89
static const double S6[3] = {75.0 / 64.0 - 25.0 / 384.0 + 3.0 / 640.0, -25.0 / 384.0 + 3.0 / 640.0,
                             3.0 / 640.0};  // 92

static inline double d4(double q0, double q1, double q2, double q3, double q4) {  // 97-99
    return std::fabs(q0 - 4.0 * q1 + 6.0 * q2 - 4.0 * q3 + q4);
}

// scheme.rs:286-333: characteristic WENO5-Z interpolation to midpoint j+1/2
// and the HLLC flux there.  Returns true on the first-order fallback.
static bool midpoint_flux(const Euler& sys, const V* w, size_t j, int d, V& fv) {
    V avg;
    for (int k = 0; k < NV; ++k) avg[k] = 0.5 * (w[j][k] + w[j + 1][k]);
    Mat l, r;
    sys.eigen(avg, d, l, r);
    double ch[6][NV];
    for (int s = 0; s < 6; ++s) {
        const V& ws = w[j - 2 + s];
        for (int a = 0; a < NV; ++a) {
            double acc = 0.0;
            for (int k = 0; k < NV; ++k) acc += l[a][k] * ws[k];
            ch[s][a] = acc;
        }
    }
    double cl[NV], cr[NV];
    for (int a = 0; a < NV; ++a) {
        cl[a] = weno5z_interp(ch[0][a], ch[1][a], ch[2][a], ch[3][a], ch[4][a]);
        cr[a] = weno5z_interp(ch[5][a], ch[4][a], ch[3][a], ch[2][a], ch[1][a]);
    }
    V wl, wr;
    for (int k = 0; k < NV; ++k) {
        double sl = 0.0, sr = 0.0;
        for (int a = 0; a < NV; ++a) {
            sl += r[k][a] * cl[a];
            sr += r[k][a] * cr[a];
        }
        wl[k] = sl;
        wr[k] = sr;
    }
    bool fb = false;
    if (!sys.admissible(wl) || !sys.admissible(wr)) {
        fb = true;
        wl = w[j];
        wr = w[j + 1];
    }
    fv = sys.riemann_flux(wl, wr, d, nullptr);
    return fb;
}

// scheme.rs:337-384
static V fd_weno_flux(const Euler& sys, const V* w, size_t i, int d) {
    V avg;
    for (int k = 0; k < NV; ++k) avg[k] = 0.5 * (w[i][k] + w[i + 1][k]);
    Mat l, r;
    sys.eigen(avg, d, l, r);
    Mat jw = sys.dudw(avg);
    Mat lc = matmul(l, invert(jw));
    Mat rc = matmul(jw, r);
    double alpha[NV] = {0, 0, 0, 0, 0};
    double gs[6][NV], vs[6][NV];
    for (int s = 0; s < 6; ++s) {
        const V& ws = w[i - 2 + s];
        V q = sys.to_cons(ws);
        V fx = sys.flux(ws, d);
        for (int a = 0; a < NV; ++a) {
            double g = 0.0, v = 0.0;
            for (int k = 0; k < NV; ++k) {
                g += lc[a][k] * fx[k];
                v += lc[a][k] * q[k];
            }
            gs[s][a] = g;
            vs[s][a] = v;
        }
        V lam = sys.eigenvalues(ws, d);
        for (int a = 0; a < NV; ++a) alpha[a] = rmax(alpha[a], std::fabs(lam[a]));
    }
    double fc[NV];
    for (int a = 0; a < NV; ++a) {
        auto fp = [&](int s) { return 0.5 * (gs[s][a] + alpha[a] * vs[s][a]); };
        auto fm = [&](int s) { return 0.5 * (gs[s][a] - alpha[a] * vs[s][a]); };
        fc[a] = weno5z(fp(0), fp(1), fp(2), fp(3), fp(4)) + weno5z(fm(5), fm(4), fm(3), fm(2), fm(1));
    }
    V fl{};
    fl.fill(0.0);
    for (int k = 0; k < NV; ++k)
        for (int a = 0; a < NV; ++a) fl[k] += rc[k][a] * fc[a];
    bool finite = true;
    for (int k = 0; k < NV; ++k) finite = finite && std::isfinite(fl[k]);
    if (!finite) fl = sys.riemann_flux(w[i], w[i + 1], d, nullptr);
    return fl;
}

// scheme.rs:395-464 (Euler has no non-conservative face terms, so only the
// flux is limited; the face velocity blend of 445-450 feeds no Euler term)
static void positivity_limit(const Euler& sys, const V* w, size_t len, int d, size_t ng, double lambda,
                             V* out, Stats& st) {
    size_t n = len - 2 * ng;
    V previous_cons = sys.to_cons(w[ng - 1]);
    V previous_flux = sys.flux(w[ng - 1], d);
    for (size_t f = 0; f <= n; ++f) {
        size_t i = ng - 1 + f;
        V ul = previous_cons, fl_phys = previous_flux;
        V ur = sys.to_cons(w[i + 1]);
        V fr_phys = sys.flux(w[i + 1], d);
        previous_cons = ur;
        previous_flux = fr_phys;
        auto ok = [&](const V& fx) {
            V a, b;
            for (int k = 0; k < NV; ++k) {
                a[k] = ul[k] - 2.0 * lambda * (fx[k] - fl_phys[k]);
                b[k] = ur[k] + 2.0 * lambda * (fx[k] - fr_phys[k]);
            }
            return sys.admissible(sys.to_prim(a)) && sys.admissible(sys.to_prim(b));
        };
        V fh = out[f];
        bool finite = true;
        for (int k = 0; k < NV; ++k) finite = finite && std::isfinite(fh[k]);
        if (finite && ok(fh)) continue;
        double alpha = rmax(sys.max_speed(w[i], d), sys.max_speed(w[i + 1], d));
        V flf;
        for (int k = 0; k < NV; ++k) flf[k] = 0.5 * (fl_phys[k] + fr_phys[k]) - 0.5 * alpha * (ur[k] - ul[k]);
        st.limited_faces += 1;
        auto blend = [&](double t) {
            V v;
            for (int k = 0; k < NV; ++k) v[k] = std::isfinite(fh[k]) ? t * fh[k] + (1.0 - t) * flf[k] : flf[k];
            return v;
        };
        double lo = 0.0, hi = 1.0;
        for (int it = 0; it < 30; ++it) {
            double mid = 0.5 * (lo + hi);
            if (ok(blend(mid)))
                lo = mid;
            else
                hi = mid;
        }
        out[f] = blend(lo);
    }
}

// scheme.rs:132-281
void line_fluxes(const Euler& sys, const V* w, size_t len, int d, size_t ng, const SchemeOpts& o, V* out,
                 Stats& st) {
    size_t n = len - 2 * ng;
    // Hybrid sensor: ln rho, ln(p + p_inf) (system.rs:409-413)
    std::vector<std::array<double, 2>> sens;
    if (o.mode == FluxMode::Hybrid) {
        sens.resize(len);
        for (size_t c = 0; c < len; ++c) {
            sens[c][0] = std::log(rmax(w[c][0], 1e-300));
            sens[c][1] = std::log(rmax(w[c][4] + sys.eos.p_inf, 1e-300));
        }
    }
    // Midpoints (memoised, index j = midpoint j+1/2) and central pair fluxes
    // (the reference's ring cache, 155-203) computed once per line: the same
    // values in the same accumulation order as recomputing them.
    std::vector<V> mid;
    std::vector<signed char> have;
    if (o.mode != FluxMode::Central && o.shock == ShockScheme::Wcns) {
        mid.resize(len);
        have.assign(len, 0);
    }
    std::vector<V> pair[3];
    std::vector<signed char> pair_have[3];
    for (int lg = 0; lg < 3; ++lg) {
        pair[lg].resize(len);
        pair_have[lg].assign(len, 0);
    }
    for (size_t f = 0; f <= n; ++f) {
        size_t i = ng - 1 + f;
        double theta;
        if (o.mode == FluxMode::Central) {
            theta = 0.0;
        } else if (o.mode == FluxMode::Weno) {
            theta = 1.0;
        } else {
            double r = 0.0;
            for (int k = 0; k < 2; ++k) {
                auto s = [&](size_t j) { return sens[j][k]; };
                r = rmax(rmax(r, d4(s(i - 2), s(i - 1), s(i), s(i + 1), s(i + 2))),
                         d4(s(i - 1), s(i), s(i + 1), s(i + 2), s(i + 3)));
            }
            double x = rclamp((r / o.tol - 0.5) / 0.5, 0.0, 1.0);
            theta = std::isfinite(r) ? x * x * (3.0 - 2.0 * x) : 1.0;
        }
        st.faces += 1;
        V central{};
        central.fill(0.0);
        if (theta < 1.0) {
            for (int li = 0; li < 3; ++li) {
                double a = A6[li];
                size_t lag = li + 1;
                for (size_t m = 0; m < lag; ++m) {
                    size_t left = i - m;
                    if (!pair_have[li][left]) {
                        pair[li][left] = sys.split_flux(w[left], w[left + lag], d);
                        pair_have[li][left] = 1;
                    }
                    const V& ft = pair[li][left];
                    for (int k = 0; k < NV; ++k) central[k] += 2.0 * a * ft[k];
                }
            }
            if (theta == 0.0) {
                out[f] = central;
                continue;
            }
        }
        st.weno_faces += 1;
        if (o.shock == ShockScheme::Wcns) {
            V fl{};
            fl.fill(0.0);
            for (int s = 0; s < 3; ++s) {
                double c = S6[s];
                size_t js[2] = {i + s, i - s};
                int nj = s == 0 ? 1 : 2;
                for (int q = 0; q < nj; ++q) {
                    size_t j = js[q];
                    if (!have[j]) {
                        bool fb = midpoint_flux(sys, w, j, d, mid[j]);
                        st.fallback_faces += fb ? 1 : 0;
                        have[j] = 1;
                    }
                    for (int k = 0; k < NV; ++k) fl[k] += c * mid[j][k];
                }
            }
            out[f] = fl;
        } else {
            out[f] = fd_weno_flux(sys, w, i, d);
        }
        if (theta < 1.0) {
            for (int k = 0; k < NV; ++k) out[f][k] = theta * out[f][k] + (1.0 - theta) * central[k];
        }
    }
    if (o.lambda > 0.0) positivity_limit(sys, w, len, d, ng, o.lambda, out, st);
}

// ------------------------------------------------------------ boundary.rs:48-145
CharacteristicStencil::CharacteristicStencil(const Euler& sys, const V& u_b, const V& u_b1, const V& u_far,
                                             int d, int side) {
    wi = sys.to_prim(u_b);
    wf = sys.to_prim(u_far);
    w1 = sys.to_prim(u_b1);
    sys.eigen(wi, d, l, r);
    V lam = sys.eigenvalues(wi, d);
    double outward = side == 1 ? 1.0 : -1.0;
    for (int k = 0; k < NV; ++k) outgoing[k] = lam[k] * outward > 0.0;
}

V CharacteristicStencil::conserved(const Euler& sys, double m) const {
    V dw;
    for (int k = 0; k < NV; ++k) dw[k] = wi[k] - wf[k] + m * (wi[k] - w1[k]);
    V wg = wf;
    for (int kf = 0; kf < NV; ++kf) {
        if (outgoing[kf]) {
            double amp = -0.0;  // Rust float Sum folds from -0.0
            for (int q = 0; q < NV; ++q) amp += l[kf][q] * dw[q];
            for (int q = 0; q < NV; ++q) wg[q] += r[q][kf] * amp;
        }
    }
    return sys.to_cons(wg);
}

// ---------------------------------------------------------------- solver.rs
Mesh::Mesh(const size_t nn[3], const double l[3], const double h[3], size_t ghosts) {  // 31-34
    for (int d = 0; d < 3; ++d) {
        n[d] = nn[d];
        lo[d] = l[d];
        dx[d] = (h[d] - l[d]) / (double)nn[d];
    }
    ng = ghosts;
}

void Mesh::periodic_ghost_sources(int d, size_t gi, size_t& lo_src, size_t& hi_src) const {  // 62-70
    size_t nn = n[d];
    if (nn >= ng) {
        lo_src = nn + gi;
        hi_src = ng + gi;
    } else {
        lo_src = ng + (gi + nn - ng % nn) % nn;
        hi_src = ng + gi % nn;
    }
}

Solver::Solver(const Euler& s, const Mesh& m) : sys(s), mesh(m) {  // 199-244
    u.assign(mesh.len(), V{0, 0, 0, 0, 0});
    rhs_buf.assign(mesh.len(), V{0, 0, 0, 0, 0});
    stage.assign(mesh.len(), V{0, 0, 0, 0, 0});
}

// boundary.rs:102-145
static V reflected_ghost(const Euler& sys, const std::vector<V>& states, const size_t* at_tab, long x, size_t n,
                         size_t ng, int d, const Bc bc[2], const CharacteristicStencil* const ch[2]) {
    bool flip = false;
    V v;
    for (;;) {
        if (x >= 0 && x < (long)n) {
            v = states[at_tab[ng + (size_t)x]];
            break;
        }
        int side = x >= (long)n ? 1 : 0;
        BcKind k = bc[side].kind;
        if (k == BcKind::Reflect) {
            x = side == 0 ? -1 - x : 2 * (long)n - 1 - x;
            flip = !flip;
        } else if (k == BcKind::Periodic) {
            long r = ((x % (long)n) + (long)n) % (long)n;
            v = states[at_tab[ng + (size_t)r]];
            break;
        } else if (k == BcKind::Outflow) {
            v = states[at_tab[ng + (side == 0 ? 0 : n - 1)]];
            break;
        } else if (k == BcKind::Dirichlet) {
            v = bc[side].state;
            break;
        } else {
            long layer = side == 0 ? -x : x - (long)n + 1;
            v = ch[side]->conserved(sys, (double)layer);
            break;
        }
    }
    if (flip) v[1 + d] *= -1.0;
    return v;
}

// solver.rs:259-348 (fill_ghosts_with + fill_ghosts_impl<DEEP>)
void Solver::fill_ghosts(std::vector<V>& uu) const {
    const Mesh& m = mesh;
    bool deep = false;
    for (int d = 0; d < 3; ++d)
        if (m.active(d) && m.n[d] < m.ng && (bc[d][0].kind == BcKind::Reflect || bc[d][1].kind == BcKind::Reflect))
            deep = true;
    for (int d = 0; d < 3; ++d) {
        if (!m.active(d)) continue;
        size_t n = m.n[d], ng = m.ng;
        int others[2], no = 0;
        for (int e = 0; e < 3; ++e)
            if (e != d) others[no++] = e;
        int e0 = others[0], e1 = others[1];
        std::vector<size_t> at_tab(m.m(d));
        for (size_t a = 0; a < m.m(e0); ++a) {
            for (size_t b = 0; b < m.m(e1); ++b) {
                for (size_t c = 0; c < m.m(d); ++c) {
                    size_t ijk[3];
                    ijk[d] = c;
                    ijk[e0] = a;
                    ijk[e1] = b;
                    at_tab[c] = m.idx(ijk[0], ijk[1], ijk[2]);
                }
                auto at = [&](size_t c) { return at_tab[c]; };
                CharacteristicStencil* chs[2] = {nullptr, nullptr};
                alignas(CharacteristicStencil) unsigned char buf0[sizeof(CharacteristicStencil)];
                alignas(CharacteristicStencil) unsigned char buf1[sizeof(CharacteristicStencil)];
                if (bc[d][0].kind == BcKind::Characteristic)
                    chs[0] = new (buf0) CharacteristicStencil(sys, uu[at(ng)], uu[at(ng + 1)], bc[d][0].state, d, 0);
                if (bc[d][1].kind == BcKind::Characteristic)
                    chs[1] = new (buf1)
                        CharacteristicStencil(sys, uu[at(n + ng - 1)], uu[at(n + ng - 2)], bc[d][1].state, d, 1);
                const CharacteristicStencil* const chc[2] = {chs[0], chs[1]};
                for (size_t gi = 0; gi < ng; ++gi) {
                    size_t lo_dst = at(gi), hi_dst = at(n + ng + gi);
                    size_t lo_s, hi_s;
                    m.periodic_ghost_sources(d, gi, lo_s, hi_s);
                    size_t lo_src = at(lo_s), hi_src = at(hi_s);
                    size_t lo_ref = at(2 * ng - 1 - gi), hi_ref = at(n + ng - 1 - gi);
                    V v;
                    switch (bc[d][0].kind) {
                        case BcKind::Periodic: v = uu[lo_src]; break;
                        case BcKind::Outflow: v = uu[at(ng)]; break;
                        case BcKind::Dirichlet: v = bc[d][0].state; break;
                        case BcKind::Characteristic: v = chc[0]->conserved(sys, (double)(ng - gi)); break;
                        case BcKind::Reflect:
                            if (deep && n < ng) {
                                v = reflected_ghost(sys, uu, at_tab.data(), (long)gi - (long)ng, n, ng, d, bc[d], chc);
                            } else {
                                v = uu[lo_ref];
                                v[1 + d] *= -1.0;
                            }
                            break;
                    }
                    uu[lo_dst] = v;
                    switch (bc[d][1].kind) {
                        case BcKind::Periodic: v = uu[hi_src]; break;
                        case BcKind::Outflow: v = uu[at(n + ng - 1)]; break;
                        case BcKind::Dirichlet: v = bc[d][1].state; break;
                        case BcKind::Characteristic: v = chc[1]->conserved(sys, (double)(gi + 1)); break;
                        case BcKind::Reflect:
                            if (deep && n < ng) {
                                v = reflected_ghost(sys, uu, at_tab.data(), (long)(n + gi), n, ng, d, bc[d], chc);
                            } else {
                                v = uu[hi_ref];
                                v[1 + d] *= -1.0;
                            }
                            break;
                    }
                    uu[hi_dst] = v;
                }
            }
        }
    }
}

// solver.rs:351-475 (Euler: no non-conservative, viscous or capillary terms)
Stats Solver::rhs(std::vector<V>& uu) {
    fill_ghosts(uu);
    std::vector<V>& out = rhs_buf;
    for (auto& v : out) v.fill(0.0);
    const Mesh& m = mesh;
    Stats total;
    double ndim = 0.0;
    for (int e = 0; e < 3; ++e) ndim += m.active(e) ? 1.0 : 0.0;
    for (int d = 0; d < 3; ++d) {
        if (!m.active(d)) continue;
        int others[2], no = 0;
        for (int e = 0; e < 3; ++e)
            if (e != d) others[no++] = e;
        int e0 = others[0], e1 = others[1];
        SchemeOpts opts{mode, shock, tol, positivity ? ndim * stage_dt / m.dx[d] : 0.0};
        long np = (long)(m.n[e0] * m.n[e1]);
        size_t len = m.m(d);
        double inv = 1.0 / m.dx[d];
        uint64_t sf = 0, swf = 0, sfb = 0, slim = 0;
#pragma omp parallel reduction(+ : sf, swf, sfb, slim)
        {
            std::vector<V> w(len), fl(m.n[d] + 1);
            std::vector<size_t> at_tab(len);
#pragma omp for schedule(static)
            for (long p = 0; p < np; ++p) {
                size_t a = (size_t)p / m.n[e1], b = (size_t)p % m.n[e1];
                for (size_t c = 0; c < len; ++c) {
                    size_t ijk[3];
                    ijk[d] = c;
                    ijk[e0] = a + m.g(e0);
                    ijk[e1] = b + m.g(e1);
                    at_tab[c] = m.idx(ijk[0], ijk[1], ijk[2]);
                    w[c] = sys.to_prim(uu[at_tab[c]]);
                }
                Stats s;
                line_fluxes(sys, w.data(), len, d, m.ng, opts, fl.data(), s);
                for (size_t c = 0; c < m.n[d]; ++c) {
                    V& o = out[at_tab[c + m.ng]];
                    for (int k = 0; k < NV; ++k) o[k] -= (fl[c + 1][k] - fl[c][k]) * inv;
                }
                sf += s.faces;
                swf += s.weno_faces;
                sfb += s.fallback_faces;
                slim += s.limited_faces;
            }
        }
        Stats st;
        st.faces = sf;
        st.weno_faces = swf;
        st.fallback_faces = sfb;
        st.limited_faces = slim;
        total.add(st);
    }
    return total;
}

// solver.rs:478-523 (Euler: no dilatation source, no viscous limit)
double Solver::dt_cfl() const {
    const Mesh& m = mesh;
    double rate = 0.0;
#pragma omp parallel for reduction(max : rate) schedule(static)
    for (long i = 0; i < (long)m.n[0]; ++i) {
        double rr = 0.0;
        for (size_t j = 0; j < m.n[1]; ++j)
            for (size_t k = 0; k < m.n[2]; ++k) {
                V w = sys.to_prim(u[m.idx(i + m.g(0), j + m.g(1), k + m.g(2))]);
                double acoustic = -0.0;
                for (int d = 0; d < 3; ++d)
                    if (m.active(d)) acoustic += sys.max_speed(w, d) / m.dx[d];
                rr = rmax(rr, acoustic);
            }
        rate = rmax(rate, rr);
    }
    return cfl / rate;
}

// solver.rs:97-137
template <int STAGE>
static void rk_combine(const Mesh& m, std::vector<V>& dest, const std::vector<V>& old, const std::vector<V>& res,
                       double dt) {
    long ni = (long)m.n[0];
#pragma omp parallel for schedule(static)
    for (long ii = 0; ii < ni; ++ii) {
        size_t i = (size_t)ii + m.g(0);
        for (size_t j = m.g(1); j < m.g(1) + m.n[1]; ++j)
            for (size_t k = m.g(2); k < m.g(2) + m.n[2]; ++k) {
                size_t q = m.idx(i, j, k);
                for (int v = 0; v < NV; ++v) {
                    if (STAGE == 1)
                        dest[q][v] = old[q][v] + dt * res[q][v];
                    else if (STAGE == 2)
                        dest[q][v] = 0.75 * old[q][v] + 0.25 * (dest[q][v] + dt * res[q][v]);
                    else
                        dest[q][v] = dest[q][v] / 3.0 + 2.0 / 3.0 * (old[q][v] + dt * res[q][v]);
                }
            }
    }
}

// solver.rs:659-738 (the plain explicit path)
void Solver::step(double dt) {
    stage_dt = dt;
    Stats s1 = rhs(u);
    stage = u;
    rk_combine<1>(mesh, stage, u, rhs_buf, dt);
    Stats s2 = rhs(stage);
    rk_combine<2>(mesh, stage, u, rhs_buf, dt);
    Stats s3 = rhs(stage);
    rk_combine<3>(mesh, u, stage, rhs_buf, dt);
    stats.add(s1);
    stats.add(s2);
    stats.add(s3);
    t += dt;
}

// solver.rs:758-768
size_t Solver::run_to(double t_end) {
    size_t steps = 0;
    while (t < t_end - 1e-10 * std::fabs(t_end)) {
        double dt = rmin(dt_cfl(), t_end - t);
        if (!(std::isfinite(dt) && dt > 1e-12 * rmax(std::fabs(t_end), 1e-300)))
            throw std::runtime_error("time step collapsed: the solution lost admissibility");
        step(dt);
        ++steps;
    }
    return steps;
}

// solver.rs:784-795
V Solver::totals() const {
    const Mesh& m = mesh;
    double vol = 1.0;
    for (int d = 0; d < 3; ++d)
        if (m.active(d)) vol *= m.dx[d];
    V s{0, 0, 0, 0, 0};
    for (size_t i = 0; i < m.n[0]; ++i)
        for (size_t j = 0; j < m.n[1]; ++j)
            for (size_t k = 0; k < m.n[2]; ++k) {
                const V& q = u[m.idx(i + m.g(0), j + m.g(1), k + m.g(2))];
                for (int v = 0; v < NV; ++v) s[v] += q[v] * vol;
            }
    return s;
}

uint64_t state_hash(const Solver& s) {
    uint64_t h = 0xcbf29ce484222325ULL;
    const Mesh& m = s.mesh;
    for (size_t i = 0; i < m.n[0]; ++i)
        for (size_t j = 0; j < m.n[1]; ++j)
            for (size_t k = 0; k < m.n[2]; ++k) {
                const V& q = s.u[m.idx(i + m.g(0), j + m.g(1), k + m.g(2))];
                for (int v = 0; v < NV; ++v) {
                    uint64_t bits;
                    std::memcpy(&bits, &q[v], 8);
                    h = (h ^ bits) * 0x100000001b3ULL;
                }
            }
    return h;
}

}  // namespace emref
