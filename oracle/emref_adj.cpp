// Oracle for the discrete adjoint: the Weno-mode (WCNS + HLLC + positivity
// limiter) residual and SSP-RK3 step of the reference CPU solver, written
// once over a scalar type T and instantiated
//   * for double: the same arithmetic, operation for operation, as
//     emref.cpp (checked bit-for-bit against it and against the reference's
//     `weno` golden state hash), and
//   * for a reverse-mode tape type `ad`: the tape-based reference gradient
//     the hand-written GPU adjoint is checked against.
// Test infrastructure only.  The reference has no adjoint; the derivative
// conventions it would need are fixed here and followed by the GPU adjoint:
//   * branches (HLLC wave-speed cases, first-order fallback, limiter
//     trigger) follow the forward evaluation;
//   * min/max take the derivative of the operand they return, the first on
//     a tie (as f64::min / f64::max evaluate);
//   * d|x|/dx = sign(x) with sign(0) = 0;
//   * the limiter's blend weight theta is differentiated implicitly at the
//     admissibility constraint that bounds it (the bisection resolves it to
//     2^-30), its Lax-Friedrichs speed alpha through the max;
//   * the time-step sequence is given (frozen), not recomputed from the CFL
//     condition;
//   * WENO-Z smoothness indicators at the rounding floor (below 1e-26
//     max v^2) make the Z-weights non-differentiable: all three at the
//     floor -> the derivative of the linear weights; some -> the weights
//     held at their computed values.  Elsewhere the weight derivative
//     (q_k - out) / S is formed without cancellation.
// Restated from fluid/crates/compressible/src/{eos.rs, system.rs,
// scheme.rs, solver.rs} (Weno mode, Euler; boundaries periodic, outflow,
// reflect with n >= ng, Dirichlet).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "emref.h"

namespace emadj {

using emref::BcKind;
using emref::Mesh;

// ------------------------------------------------------------------ tape
struct Tape {
    std::vector<int> a, b;
    std::vector<double> da, db;
    int push(int ia, double pa, int ib, double pb) {
        a.push_back(ia);
        b.push_back(ib);
        da.push_back(pa);
        db.push_back(pb);
        return (int)a.size() - 1;
    }
    void clear() {
        a.clear();
        b.clear();
        da.clear();
        db.clear();
    }
    // adjoint sweep: adj has one entry per node, outputs seeded
    void reverse(std::vector<double>& adj) const {
        for (long i = (long)a.size() - 1; i >= 0; --i) {
            const double g = adj[i];
            if (g == 0.0) continue;
            if (a[i] >= 0) adj[a[i]] += da[i] * g;
            if (b[i] >= 0) adj[b[i]] += db[i] * g;
        }
    }
};
static thread_local Tape* g_tape = nullptr;

struct ad {
    double v;
    int i;
    ad() : v(0.0), i(-1) {}
    ad(double x) : v(x), i(-1) {}
    ad(double x, int idx) : v(x), i(idx) {}
    static ad input(double x) { return ad(x, g_tape->push(-1, 0.0, -1, 0.0)); }
};
static inline ad mk(double v, int ia, double pa, int ib, double pb) {
    if (ia < 0 && ib < 0) return ad(v);
    return ad(v, g_tape->push(ia, pa, ib, pb));
}
static inline ad operator+(const ad& x, const ad& y) { return mk(x.v + y.v, x.i, 1.0, y.i, 1.0); }
static inline ad operator-(const ad& x, const ad& y) { return mk(x.v - y.v, x.i, 1.0, y.i, -1.0); }
static inline ad operator*(const ad& x, const ad& y) { return mk(x.v * y.v, x.i, y.v, y.i, x.v); }
static inline ad operator/(const ad& x, const ad& y) {
    const double q = x.v / y.v;
    return mk(q, x.i, 1.0 / y.v, y.i, -q / y.v);
}
static inline ad operator-(const ad& x) { return mk(-x.v, x.i, -1.0, -1, 0.0); }
static inline ad& operator+=(ad& x, const ad& y) { return x = x + y; }
static inline ad& operator-=(ad& x, const ad& y) { return x = x - y; }
static inline ad& operator*=(ad& x, const ad& y) { return x = x * y; }
static inline bool operator<(const ad& x, const ad& y) { return x.v < y.v; }
static inline bool operator>(const ad& x, const ad& y) { return x.v > y.v; }
static inline bool operator<=(const ad& x, const ad& y) { return x.v <= y.v; }
static inline bool operator>=(const ad& x, const ad& y) { return x.v >= y.v; }
static inline bool operator==(const ad& x, const ad& y) { return x.v == y.v; }
static inline bool operator!=(const ad& x, const ad& y) { return x.v != y.v; }
static inline ad sqrt(const ad& x) {
    const double s = std::sqrt(x.v);
    return mk(s, x.i, 0.5 / s, -1, 0.0);
}
static inline ad fabs(const ad& x) {
    const double sg = x.v > 0.0 ? 1.0 : (x.v < 0.0 ? -1.0 : 0.0);
    return mk(std::fabs(x.v), x.i, sg, -1, 0.0);
}
// f64::max / f64::min: the non-NaN operand; on a tie the first
static inline ad fmax(const ad& x, const ad& y) {
    if (std::isnan(x.v)) return y;
    if (std::isnan(y.v)) return x;
    return y.v > x.v ? y : x;
}
static inline ad fmin(const ad& x, const ad& y) {
    if (std::isnan(x.v)) return y;
    if (std::isnan(y.v)) return x;
    return y.v < x.v ? y : x;
}
static inline bool isfinite(const ad& x) { return std::isfinite(x.v); }
static inline double val(const ad& x) { return x.v; }
static inline double val(double x) { return x; }

// ------------------------------------------------------------------ scheme over T
template <class T>
using VT = std::array<T, 5>;
template <class T>
using MatT = std::array<VT<T>, 5>;

struct Ctx {
    emref::Nasg eos;
    Mesh mesh;
    emref::Bc bc[3][2];
    bool positivity = true;
};

template <class T>
static inline T rmax(const T& a, const T& b) {
    using std::fmax;
    return fmax(a, b);
}
template <class T>
static inline T rmin(const T& a, const T& b) {
    using std::fmin;
    return fmin(a, b);
}

// eos.rs:36-79
template <class T>
static T pressure(const emref::Nasg& e, const T& rho, const T& en) {
    T v = 1.0 / rho;
    return (e.gamma - 1.0) * (en - e.q) / (v - e.b) - e.gamma * e.p_inf;
}
template <class T>
static T rho_e(const emref::Nasg& e, const T& rho, const T& p) {
    T v = 1.0 / rho;
    return rho * ((p + e.gamma * e.p_inf) * (v - e.b) / (e.gamma - 1.0) + e.q);
}
template <class T>
static T sound_speed_sq(const emref::Nasg& e, const T& rho, const T& p) {
    T v = 1.0 / rho;
    return e.gamma * (p + e.p_inf) * v * v / (v - e.b);
}

// system.rs:197-222, 185-187, 404-407
template <class T>
static VT<T> to_prim(const emref::Nasg& e, const VT<T>& u) {
    T rho = u[0];
    T vx = u[1] / rho, vy = u[2] / rho, vz = u[3] / rho;
    T en = u[4] / rho - 0.5 * (vx * vx + vy * vy + vz * vz);
    return VT<T>{rho, vx, vy, vz, pressure(e, rho, en)};
}
template <class T>
static VT<T> to_cons(const emref::Nasg& e, const VT<T>& w) {
    T rho = w[0];
    T ke = 0.5 * rho * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
    return VT<T>{rho, rho * w[1], rho * w[2], rho * w[3], rho_e(e, rho, w[4]) + ke};
}
template <class T>
static T sound_speed(const emref::Nasg& e, const VT<T>& w) {
    using std::sqrt;
    return sqrt(rmax<T>(sound_speed_sq(e, w[0], w[4]), T(0.0)));
}
template <class T>
static T max_speed(const emref::Nasg& e, const VT<T>& w, int d) {
    using std::fabs;
    return fabs(w[1 + d]) + sound_speed(e, w);
}
template <class T>
static bool admissible(const emref::Nasg& e, const VT<T>& w) {
    using std::isfinite;
    return isfinite(w[0]) && isfinite(w[4]) && w[0] > 0.0 && 1.0 / val(w[0]) > e.b && w[4] + e.p_inf > 0.0;
}

// system.rs:232-258 (dense, as the reference)
template <class T>
static void eigen(const emref::Nasg& e, const VT<T>& w, int d, MatT<T>& l, MatT<T>& r) {
    T rho = w[0], c = sound_speed(e, w);
    int un = 1 + d;
    int t1 = 1 + (d + 1) % 3, t2 = 1 + (d + 2) % 3;
    for (auto& row : l) row.fill(T(0.0));
    for (auto& row : r) row.fill(T(0.0));
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

template <class T>
static VT<T> phys_flux_wq(const VT<T>& w, const VT<T>& q, int un) {
    T u = w[un];
    VT<T> f{};
    f[0] = q[0] * u;
    f[1] = q[1] * u;
    f[2] = q[2] * u;
    f[3] = q[3] * u;
    f[un] += w[4];
    f[4] = (q[4] + w[4]) * u;
    return f;
}

// system.rs:260-312
template <class T>
static VT<T> riemann_flux(const emref::Nasg& e, const VT<T>& wl, const VT<T>& wr, int d) {
    int un = 1 + d;
    T cl = sound_speed(e, wl), cr = sound_speed(e, wr);
    T ul = wl[un], ur = wr[un];
    T sl = rmin<T>(ul - cl, ur - cr);
    T sr = rmax<T>(ul + cl, ur + cr);
    VT<T> ql = to_cons(e, wl), qr = to_cons(e, wr);
    if (sl >= 0.0) return phys_flux_wq(wl, ql, un);
    if (sr <= 0.0) return phys_flux_wq(wr, qr, un);
    T rl = wl[0], rr = wr[0], pl = wl[4], pr = wr[4];
    T ss = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur));
    auto star = [&](const VT<T>& w, const VT<T>& q, const T& s) {
        T rho = w[0], u = w[un];
        T fac = rho * (s - u) / (s - ss);
        VT<T> qs{};
        qs[0] = fac;
        for (int k = 1; k < 4; ++k) qs[k] = fac * w[k];
        qs[un] = fac * ss;
        qs[4] = fac * (q[4] / rho + (ss - u) * (ss + w[4] / (rho * (s - u))));
        return qs;
    };
    VT<T> f{};
    if (ss >= 0.0) {
        VT<T> fl = phys_flux_wq(wl, ql, un), qs = star(wl, ql, sl);
        for (int k = 0; k < 5; ++k) f[k] = fl[k] + sl * (qs[k] - ql[k]);
    } else {
        VT<T> fr = phys_flux_wq(wr, qr, un), qs = star(wr, qr, sr);
        for (int k = 0; k < 5; ++k) f[k] = fr[k] + sr * (qs[k] - qr[k]);
    }
    return f;
}

// system.rs:336-348
template <class T>
static VT<T> flux(const emref::Nasg& e, const VT<T>& w, int d) {
    int un = 1 + d;
    VT<T> q = to_cons(e, w);
    T u = w[un];
    VT<T> f{};
    for (int k = 0; k < 4; ++k) f[k] = q[k] * u;
    f[un] += w[4];
    f[4] = (q[4] + w[4]) * u;
    return f;
}

// scheme.rs:69-86
template <class T>
static T weno5z_interp(const T& v0, const T& v1, const T& v2, const T& v3, const T& v4) {
    using std::fabs;
    T q[3] = {(3.0 * v0 - 10.0 * v1 + 15.0 * v2) / 8.0, (-v1 + 6.0 * v2 + 3.0 * v3) / 8.0,
              (3.0 * v2 + 6.0 * v3 - v4) / 8.0};
    static const double dw[3] = {1.0 / 16.0, 10.0 / 16.0, 5.0 / 16.0};
    T t0 = v0 - 2.0 * v1 + v2, s0 = v0 - 4.0 * v1 + 3.0 * v2;
    T t1 = v1 - 2.0 * v2 + v3, s1 = v1 - v3;
    T t2 = v2 - 2.0 * v3 + v4, s2 = 3.0 * v2 - 4.0 * v3 + v4;
    T b0 = 13.0 / 12.0 * (t0 * t0) + 0.25 * (s0 * s0);
    T b1 = 13.0 / 12.0 * (t1 * t1) + 0.25 * (s1 * s1);
    T b2 = 13.0 / 12.0 * (t2 * t2) + 0.25 * (s2 * s2);
    T tau = fabs(b0 - b2);
    const double EPS = 1e-40;
    T r0 = tau / (b0 + EPS), r1 = tau / (b1 + EPS), r2 = tau / (b2 + EPS);
    T a0 = dw[0] * (1.0 + r0 * r0);
    T a1 = dw[1] * (1.0 + r1 * r1);
    T a2 = dw[2] * (1.0 + r2 * r2);
    T out = (a0 * q[0] + a1 * q[1] + a2 * q[2]) / (a0 + a1 + a2);
    if constexpr (std::is_same<T, ad>::value) {
        // A smoothness indicator at the rounding floor (below 1e-26 max v^2)
        // makes the Z-weights non-differentiable (tau / (beta + eps) is a
        // ratio involving rounding errors).  All three at the floor: the
        // derivative of the linear weights (the smooth limit); some: the
        // weights held at their computed values.  The value stays the
        // forward one.
        const double mv = std::fmax(std::fmax(std::fmax(std::fabs(v0.v), std::fabs(v1.v)),
                                              std::fmax(std::fabs(v2.v), std::fabs(v3.v))),
                                    std::fabs(v4.v));
        const double flat = 1e-26 * mv * mv;
        const int nflat = (b0.v <= flat) + (b1.v <= flat) + (b2.v <= flat);
        if (nflat == 3) {
            ad t = mk(out.v, q[0].i, dw[0], q[1].i, dw[1]);
            return mk(out.v, t.i, 1.0, q[2].i, dw[2]);
        }
        const double S = (a0 + a1 + a2).v;
        ad t = mk(out.v, q[0].i, a0.v / S, q[1].i, a1.v / S);
        ad wq = mk(out.v, t.i, 1.0, q[2].i, a2.v / S);
        if (nflat > 0) return wq;
        // out = sum_k a_k q_k / S with d out / d a_k = (q_k - out) / S, where
        // q_k - out = sum_j a_j (q_k - q_j) / S is formed without
        // cancellation (the chain rule through N / S would subtract two
        // nearly equal terms and leave eps |q| / sqrt(beta_k) of noise on a
        // nearly flat sub-stencil).
        const double qv[3] = {q[0].v, q[1].v, q[2].v}, av[3] = {a0.v, a1.v, a2.v};
        double dk[3];
        for (int k = 0; k < 3; ++k) {
            double acc = 0.0;
            for (int j = 0; j < 3; ++j)
                if (j != k) acc += av[j] * (qv[k] - qv[j]);
            dk[k] = acc / S / S;
        }
        ad u = mk(out.v, wq.i, 1.0, a0.i, dk[0]);
        u = mk(out.v, u.i, 1.0, a1.i, dk[1]);
        return mk(out.v, u.i, 1.0, a2.i, dk[2]);
    }
    return out;
}

// scheme.rs:286-333
template <class T>
static bool midpoint_flux(const emref::Nasg& e, const VT<T>* w, size_t j, int d, VT<T>& fv) {
    VT<T> avg;
    for (int k = 0; k < 5; ++k) avg[k] = 0.5 * (w[j][k] + w[j + 1][k]);
    MatT<T> l, r;
    eigen(e, avg, d, l, r);
    T ch[6][5];
    for (int s = 0; s < 6; ++s) {
        const VT<T>& ws = w[j - 2 + s];
        for (int a = 0; a < 5; ++a) {
            T acc = 0.0;
            for (int k = 0; k < 5; ++k) acc += l[a][k] * ws[k];
            ch[s][a] = acc;
        }
    }
    T cl[5], cr[5];
    for (int a = 0; a < 5; ++a) {
        cl[a] = weno5z_interp(ch[0][a], ch[1][a], ch[2][a], ch[3][a], ch[4][a]);
        cr[a] = weno5z_interp(ch[5][a], ch[4][a], ch[3][a], ch[2][a], ch[1][a]);
    }
    VT<T> wl, wr;
    for (int k = 0; k < 5; ++k) {
        T sl = 0.0, sr = 0.0;
        for (int a = 0; a < 5; ++a) {
            sl += r[k][a] * cl[a];
            sr += r[k][a] * cr[a];
        }
        wl[k] = sl;
        wr[k] = sr;
    }
    bool fb = false;
    if (!admissible(e, wl) || !admissible(e, wr)) {
        fb = true;
        wl = w[j];
        wr = w[j + 1];
    }
    fv = riemann_flux(e, wl, wr, d);
    return fb;
}

// Derivative of the limiter's blend weight.  The bisection returns theta
// within 2^-30 below the largest admissible theta*, so on any scale a
// perturbation can resolve theta follows theta*(data): it is differentiated
// implicitly at the admissibility constraint g (rho, covolume 1/rho - b or
// p + p_inf of one of the two test states) that fails at the bisection's
// upper bound, d theta = -dg / (dg/dtheta).  theta = 0 (nothing admissible
// above it) stays constant.
static double tape_local_derivative(int out, int in) {
    // d out / d in over the nodes in..out (created after `in`)
    const Tape& t = *g_tape;
    std::vector<double> adj((size_t)(out - in + 1), 0.0);
    adj[(size_t)(out - in)] = 1.0;
    for (int i = out; i > in; --i) {
        const double g = adj[(size_t)(i - in)];
        if (g == 0.0) continue;
        if (t.a[i] >= in) adj[(size_t)(t.a[i] - in)] += t.da[i] * g;
        if (t.b[i] >= in) adj[(size_t)(t.b[i] - in)] += t.db[i] * g;
    }
    return adj[0];
}

template <class T>
static T limiter_theta(const emref::Nasg& e, double lambda, const VT<T>& ul, const VT<T>& ur, const VT<T>& fl,
                       const VT<T>& fr, const VT<T>& fh, const VT<T>& flf, double lo, double hi) {
    if constexpr (!std::is_same<T, ad>::value) {
        return lo;
    } else {
        if (!(lo > 0.0)) return ad(lo);
        // the constraint that fails at hi (first in a fixed order)
        int which = -1;
        for (int side = 0; side < 2 && which < 0; ++side) {
            VT<double> st;
            for (int k = 0; k < 5; ++k) {
                const double fx = std::isfinite(fh[k].v) ? hi * fh[k].v + (1.0 - hi) * flf[k].v : flf[k].v;
                st[k] = side == 0 ? ul[k].v - 2.0 * lambda * (fx - fl[k].v) : ur[k].v + 2.0 * lambda * (fx - fr[k].v);
            }
            const VT<double> ws = to_prim(e, st);
            if (!std::isfinite(ws[0]) || !std::isfinite(ws[4])) return ad(lo);
            if (!(ws[0] > 0.0))
                which = 3 * side;
            else if (!(1.0 / ws[0] > e.b))
                which = 3 * side + 1;
            else if (!(ws[4] + e.p_inf > 0.0))
                which = 3 * side + 2;
        }
        if (which < 0) return ad(lo);
        const ad tin = ad::input(lo);
        VT<ad> st;
        for (int k = 0; k < 5; ++k) {
            const ad fx = std::isfinite(fh[k].v) ? tin * fh[k] + (1.0 - tin) * flf[k] : flf[k];
            st[k] = which < 3 ? ul[k] - 2.0 * lambda * (fx - fl[k]) : ur[k] + 2.0 * lambda * (fx - fr[k]);
        }
        const VT<ad> ws = to_prim(e, st);
        const ad g = which % 3 == 0 ? ws[0] : which % 3 == 1 ? 1.0 / ws[0] - e.b : ws[4] + e.p_inf;
        const double gth = g.i > tin.i ? tape_local_derivative(g.i, tin.i) : 0.0;
        if (!std::isfinite(gth) || gth == 0.0) return ad(lo);
        return mk(lo, g.i, -1.0 / gth, -1, 0.0);
    }
}

// scheme.rs:395-464; theta differentiated implicitly (limiter_theta), alpha
// through its max (the operand it returns)
template <class T>
static void positivity_limit(const emref::Nasg& e, const VT<T>* w, size_t len, int d, size_t ng, double lambda,
                             VT<T>* out) {
    size_t n = len - 2 * ng;
    VT<T> previous_cons = to_cons(e, w[ng - 1]);
    VT<T> previous_flux = flux(e, w[ng - 1], d);
    for (size_t f = 0; f <= n; ++f) {
        size_t i = ng - 1 + f;
        VT<T> ul = previous_cons, fl_phys = previous_flux;
        VT<T> ur = to_cons(e, w[i + 1]);
        VT<T> fr_phys = flux(e, w[i + 1], d);
        previous_cons = ur;
        previous_flux = fr_phys;
        auto ok = [&](const VT<T>& fx) {
            VT<T> a, b;
            for (int k = 0; k < 5; ++k) {
                a[k] = ul[k] - 2.0 * lambda * (fx[k] - fl_phys[k]);
                b[k] = ur[k] + 2.0 * lambda * (fx[k] - fr_phys[k]);
            }
            return admissible(e, to_prim(e, a)) && admissible(e, to_prim(e, b));
        };
        VT<T> fh = out[f];
        bool finite = true;
        for (int k = 0; k < 5; ++k) finite = finite && std::isfinite(val(fh[k]));
        if (finite && ok(fh)) continue;
        const T alpha = rmax<T>(max_speed(e, w[i], d), max_speed(e, w[i + 1], d));
        VT<T> flf;
        for (int k = 0; k < 5; ++k) flf[k] = 0.5 * (fl_phys[k] + fr_phys[k]) - 0.5 * alpha * (ur[k] - ul[k]);
        auto blend = [&](double t) {
            VT<T> v;
            for (int k = 0; k < 5; ++k) v[k] = std::isfinite(val(fh[k])) ? t * fh[k] + (1.0 - t) * flf[k] : flf[k];
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
        const T th = limiter_theta<T>(e, lambda, ul, ur, fl_phys, fr_phys, fh, flf, lo, hi);
        for (int k = 0; k < 5; ++k) out[f][k] = std::isfinite(val(fh[k])) ? th * fh[k] + (1.0 - th) * flf[k] : flf[k];
    }
}

static const double S6[3] = {75.0 / 64.0 - 25.0 / 384.0 + 3.0 / 640.0, -25.0 / 384.0 + 3.0 / 640.0,
                             3.0 / 640.0};  // scheme.rs:92

// scheme.rs:132-281, FluxMode::Weno with the WCNS shock scheme
template <class T>
static void line_fluxes(const emref::Nasg& e, const VT<T>* w, size_t len, int d, size_t ng, double lambda,
                        VT<T>* out, std::vector<VT<T>>& mid, std::vector<signed char>& have) {
    size_t n = len - 2 * ng;
    mid.resize(len);
    have.assign(len, 0);
    for (size_t f = 0; f <= n; ++f) {
        size_t i = ng - 1 + f;
        VT<T> fl{};
        fl.fill(T(0.0));
        for (int s = 0; s < 3; ++s) {
            double c = S6[s];
            size_t js[2] = {i + s, i - s};
            int nj = s == 0 ? 1 : 2;
            for (int q = 0; q < nj; ++q) {
                size_t j = js[q];
                if (!have[j]) {
                    midpoint_flux(e, w, j, d, mid[j]);
                    have[j] = 1;
                }
                for (int k = 0; k < 5; ++k) fl[k] += c * mid[j][k];
            }
        }
        out[f] = fl;
    }
    if (lambda > 0.0) positivity_limit(e, w, len, d, ng, lambda, out);
}

// solver.rs:259-348 without characteristic boundaries and deep images
template <class T>
static void fill_ghosts(const Ctx& cx, std::vector<VT<T>>& uu) {
    const Mesh& m = cx.mesh;
    for (int d = 0; d < 3; ++d) {
        if (!m.active(d)) continue;
        size_t n = m.n[d], ng = m.ng;
        int others[2], no = 0;
        for (int e = 0; e < 3; ++e)
            if (e != d) others[no++] = e;
        int e0 = others[0], e1 = others[1];
        std::vector<size_t> at_tab(m.m(d));
        for (size_t a = 0; a < m.m(e0); ++a)
            for (size_t b = 0; b < m.m(e1); ++b) {
                for (size_t c = 0; c < m.m(d); ++c) {
                    size_t ijk[3];
                    ijk[d] = c;
                    ijk[e0] = a;
                    ijk[e1] = b;
                    at_tab[c] = m.idx(ijk[0], ijk[1], ijk[2]);
                }
                auto at = [&](size_t c) { return at_tab[c]; };
                for (size_t gi = 0; gi < ng; ++gi) {
                    size_t lo_dst = at(gi), hi_dst = at(n + ng + gi);
                    size_t lo_s, hi_s;
                    m.periodic_ghost_sources(d, gi, lo_s, hi_s);
                    size_t lo_src = at(lo_s), hi_src = at(hi_s);
                    size_t lo_ref = at(2 * ng - 1 - gi), hi_ref = at(n + ng - 1 - gi);
                    VT<T> v;
                    for (int side = 0; side < 2; ++side) {
                        const emref::Bc& bc = cx.bc[d][side];
                        switch (bc.kind) {
                            case BcKind::Periodic: v = uu[side ? hi_src : lo_src]; break;
                            case BcKind::Outflow: v = uu[at(side ? n + ng - 1 : ng)]; break;
                            case BcKind::Dirichlet:
                                for (int q = 0; q < 5; ++q) v[q] = bc.state[q];
                                break;
                            case BcKind::Reflect:
                                v = uu[side ? hi_ref : lo_ref];
                                v[1 + d] *= -1.0;
                                break;
                            default: throw std::runtime_error("boundary kind not supported by the adjoint oracle");
                        }
                        uu[side ? hi_dst : lo_dst] = v;
                    }
                }
            }
    }
}

// solver.rs:351-475: residual of uu (ghosts refreshed) into out
template <class T>
static void rhs(const Ctx& cx, std::vector<VT<T>>& uu, double stage_dt, std::vector<VT<T>>& out) {
    fill_ghosts(cx, uu);
    const Mesh& m = cx.mesh;
    out.assign(m.len(), VT<T>{});
    for (auto& v : out) v.fill(T(0.0));
    double ndim = 0.0;
    for (int e = 0; e < 3; ++e) ndim += m.active(e) ? 1.0 : 0.0;
    std::vector<VT<T>> mid;
    std::vector<signed char> have;
    for (int d = 0; d < 3; ++d) {
        if (!m.active(d)) continue;
        int others[2], no = 0;
        for (int e = 0; e < 3; ++e)
            if (e != d) others[no++] = e;
        int e0 = others[0], e1 = others[1];
        double lambda = cx.positivity ? ndim * stage_dt / m.dx[d] : 0.0;
        size_t len = m.m(d);
        double inv = 1.0 / m.dx[d];
        std::vector<VT<T>> w(len), fl(m.n[d] + 1);
        std::vector<size_t> at_tab(len);
        for (size_t a = 0; a < m.n[e0]; ++a)
            for (size_t b = 0; b < m.n[e1]; ++b) {
                for (size_t c = 0; c < len; ++c) {
                    size_t ijk[3];
                    ijk[d] = c;
                    ijk[e0] = a + m.g(e0);
                    ijk[e1] = b + m.g(e1);
                    at_tab[c] = m.idx(ijk[0], ijk[1], ijk[2]);
                    w[c] = to_prim(cx.eos, uu[at_tab[c]]);
                }
                line_fluxes(cx.eos, w.data(), len, d, m.ng, lambda, fl.data(), mid, have);
                for (size_t c = 0; c < m.n[d]; ++c) {
                    VT<T>& o = out[at_tab[c + m.ng]];
                    for (int k = 0; k < 5; ++k) o[k] -= (fl[c + 1][k] - fl[c][k]) * inv;
                }
            }
    }
}

// solver.rs:97-137, 659-738
template <class T>
static void step(const Ctx& cx, std::vector<VT<T>>& u, double dt) {
    const Mesh& m = cx.mesh;
    std::vector<VT<T>> r, stage;
    auto interior = [&](auto&& fn) {
        for (size_t i = m.g(0); i < m.g(0) + m.n[0]; ++i)
            for (size_t j = m.g(1); j < m.g(1) + m.n[1]; ++j)
                for (size_t k = m.g(2); k < m.g(2) + m.n[2]; ++k) fn(m.idx(i, j, k));
    };
    rhs(cx, u, dt, r);
    stage = u;
    interior([&](size_t q) {
        for (int v = 0; v < 5; ++v) stage[q][v] = u[q][v] + dt * r[q][v];
    });
    rhs(cx, stage, dt, r);
    interior([&](size_t q) {
        for (int v = 0; v < 5; ++v) stage[q][v] = 0.75 * u[q][v] + 0.25 * (stage[q][v] + dt * r[q][v]);
    });
    rhs(cx, stage, dt, r);
    interior([&](size_t q) {
        for (int v = 0; v < 5; ++v) u[q][v] = u[q][v] / 3.0 + 2.0 / 3.0 * (stage[q][v] + dt * r[q][v]);
    });
}

}  // namespace emadj

// ------------------------------------------------------------------ C API
using namespace emadj;

static void load_u(const Ctx& cx, const double* u, std::vector<VT<double>>& out) {
    out.resize(cx.mesh.len());
    std::memcpy(out.data(), u, cx.mesh.len() * 5 * sizeof(double));
}

extern "C" {

void* emadj_create(const int64_t n[3], const double lo[3], const double hi[3], int64_t ng, const double eos[4],
                   const int32_t bc_kind[6], const double bc_state[30], int32_t positivity) {
    Ctx* cx = new Ctx();
    size_t nn[3] = {(size_t)n[0], (size_t)n[1], (size_t)n[2]};
    cx->mesh = Mesh(nn, lo, hi, (size_t)ng);
    cx->eos = emref::Nasg{eos[0], eos[1], eos[2], eos[3]};
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s) {
            cx->bc[d][s].kind = (BcKind)bc_kind[2 * d + s];
            for (int q = 0; q < 5; ++q) cx->bc[d][s].state[q] = bc_state ? bc_state[(2 * d + s) * 5 + q] : 0.0;
        }
    cx->positivity = positivity != 0;
    return cx;
}

void emadj_destroy(void* h) { delete (Ctx*)h; }

// double instantiation: residual and steps (pins the templated code)
void emadj_rhs(void* h, double stage_dt, const double* u, double* out) {
    const Ctx& cx = *(Ctx*)h;
    std::vector<VT<double>> uu, r;
    load_u(cx, u, uu);
    rhs(cx, uu, stage_dt, r);
    std::memcpy(out, r.data(), r.size() * 5 * sizeof(double));
}

void emadj_steps(void* h, double* u, int64_t nsteps, const double* dts) {
    const Ctx& cx = *(Ctx*)h;
    std::vector<VT<double>> uu;
    load_u(cx, u, uu);
    for (int64_t s = 0; s < nsteps; ++s) step(cx, uu, dts[s]);
    std::memcpy(u, uu.data(), uu.size() * 5 * sizeof(double));
}

// tape: vb = (d weno5z_interp / d v)^T ob for one stencil v[0..4]
void emadj_weno_vjp(const double* v, double ob, double* vb) {
    Tape tape;
    g_tape = &tape;
    ad x[5];
    for (int k = 0; k < 5; ++k) x[k] = ad::input(v[k]);
    ad o = weno5z_interp<ad>(x[0], x[1], x[2], x[3], x[4]);
    std::vector<double> adj(tape.a.size(), 0.0);
    if (o.i >= 0) adj[o.i] = ob;
    tape.reverse(adj);
    for (int k = 0; k < 5; ++k) vb[k] = adj[x[k].i];
    g_tape = nullptr;
}

// tape: (d rhs / d u)^T lam at u (interior cells; ghosts of u are recomputed)
void emadj_vjp_rhs(void* h, double stage_dt, const double* u, const double* lam, double* out) {
    const Ctx& cx = *(Ctx*)h;
    const Mesh& m = cx.mesh;
    Tape tape;
    g_tape = &tape;
    std::vector<VT<ad>> uu(m.len()), r;
    std::vector<int> in_idx(m.len() * 5, -1);
    for (size_t c = 0; c < m.len(); ++c)
        for (int q = 0; q < 5; ++q) uu[c][q] = ad(u[c * 5 + q]);
    for (size_t i = m.g(0); i < m.g(0) + m.n[0]; ++i)
        for (size_t j = m.g(1); j < m.g(1) + m.n[1]; ++j)
            for (size_t k = m.g(2); k < m.g(2) + m.n[2]; ++k) {
                const size_t c = m.idx(i, j, k);
                for (int q = 0; q < 5; ++q) {
                    uu[c][q] = ad::input(u[c * 5 + q]);
                    in_idx[c * 5 + q] = uu[c][q].i;
                }
            }
    rhs(cx, uu, stage_dt, r);
    std::vector<double> adj(tape.a.size(), 0.0);
    for (size_t c = 0; c < m.len(); ++c)
        for (int q = 0; q < 5; ++q)
            if (r[c][q].i >= 0) adj[r[c][q].i] += lam[c * 5 + q];
    tape.reverse(adj);
    for (size_t c = 0; c < m.len() * 5; ++c) out[c] = in_idx[c] >= 0 ? adj[in_idx[c]] : 0.0;
    g_tape = nullptr;
}

// tape gradient of J(u_N) through nsteps steps with the given dt sequence:
// dj_du holds dJ/du_N on entry and dJ/du_0 on exit (interior cells; the
// forward states are kept, each step is taped on its own)
void emadj_gradient(void* h, const double* u0, int64_t nsteps, const double* dts, double* dj_du) {
    const Ctx& cx = *(Ctx*)h;
    const Mesh& m = cx.mesh;
    std::vector<std::vector<VT<double>>> states((size_t)nsteps + 1);
    load_u(cx, u0, states[0]);
    for (int64_t s = 0; s < nsteps; ++s) {
        states[s + 1] = states[s];
        step(cx, states[s + 1], dts[s]);
    }
    std::vector<double> bar(dj_du, dj_du + m.len() * 5);
    for (int64_t s = nsteps - 1; s >= 0; --s) {
        Tape tape;
        g_tape = &tape;
        std::vector<VT<ad>> uu(m.len());
        std::vector<int> in_idx(m.len() * 5, -1);
        for (size_t c = 0; c < m.len(); ++c)
            for (int q = 0; q < 5; ++q) uu[c][q] = ad(states[s][c][q]);
        for (size_t i = m.g(0); i < m.g(0) + m.n[0]; ++i)
            for (size_t j = m.g(1); j < m.g(1) + m.n[1]; ++j)
                for (size_t k = m.g(2); k < m.g(2) + m.n[2]; ++k) {
                    const size_t c = m.idx(i, j, k);
                    for (int q = 0; q < 5; ++q) {
                        uu[c][q] = ad::input(states[s][c][q]);
                        in_idx[c * 5 + q] = uu[c][q].i;
                    }
                }
        step(cx, uu, dts[s]);
        std::vector<double> adj(tape.a.size(), 0.0);
        for (size_t i = m.g(0); i < m.g(0) + m.n[0]; ++i)
            for (size_t j = m.g(1); j < m.g(1) + m.n[1]; ++j)
                for (size_t k = m.g(2); k < m.g(2) + m.n[2]; ++k) {
                    const size_t c = m.idx(i, j, k);
                    for (int q = 0; q < 5; ++q)
                        if (uu[c][q].i >= 0) adj[uu[c][q].i] += bar[c * 5 + q];
                }
        tape.reverse(adj);
        for (size_t c = 0; c < m.len() * 5; ++c) bar[c] = in_idx[c] >= 0 ? adj[in_idx[c]] : 0.0;
        g_tape = nullptr;
    }
    std::memcpy(dj_du, bar.data(), bar.size() * sizeof(double));
}

}  // extern "C"
