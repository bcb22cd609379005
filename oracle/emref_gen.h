// Oracle, generic over the physics system: the reference's `line_fluxes`,
// positivity limiter, ghost fill, residual (with the non-conservative
// c(w) div u term), CFL time step (with the dilatation-rate limit) and
// SSP-RK3 step, written once for any System<N> the way the reference writes
// them (generic over `const N: usize, S: System<N>`).  Test infrastructure
// only: the product library never links it.
//
// Restated from (paths relative to the reference repository root):
//   fluid/crates/compressible/src/scheme.rs   129-281 (line_fluxes),
//                                             283-333 (midpoint_flux),
//                                             386-464 (positivity_limit)
//   fluid/crates/compressible/src/solver.rs   259-348 (ghosts), 350-475
//                                             (rhs), 477-523 (dt_cfl),
//                                             565-640 (neighbour states),
//                                             659-738 (step)
//   fluid/crates/compressible/src/boundary.rs 48-145 (characteristic and
//                                             deep reflected ghosts)
// FD-WENO, viscous, capillary, scalar and all-speed hooks are not restated
// (no configuration here uses them).
//
// A system type S provides: static constexpr int N; to_prim, to_cons,
// sound_speed, max_speed, eigen (dense rows/columns), eigenvalues,
// riemann_flux (returning the contact speed), split_flux, flux, admissible,
// sensor (count of values), vel_index, momentum_index, nc_var (-1 if none),
// nc_coef, dilatation_rate, velocity_conserved.  Same IEEE discipline as the
// Euler oracle: no contraction, reference evaluation order, -0.0 seeds for
// Rust's float Sum.
#pragma once
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

#include "emref.h"

namespace emref {
namespace gen {

template <int N>
using Vn = std::array<double, N>;
template <int N>
using Mn = std::array<Vn<N>, N>;

inline double fmax_(double a, double b) { return std::fmax(a, b); }  // f64::max
inline double fmin_(double a, double b) { return std::fmin(a, b); }  // f64::min
inline double clamp_(double x, double lo, double hi) {               // f64::clamp
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

constexpr double kA6[3] = {0.75, -0.15, 1.0 / 60.0};  // scheme.rs:89
constexpr double kS6[3] = {75.0 / 64.0 - 25.0 / 384.0 + 3.0 / 640.0, -25.0 / 384.0 + 3.0 / 640.0,
                           3.0 / 640.0};  // scheme.rs:92

inline double d4g(double q0, double q1, double q2, double q3, double q4) {  // scheme.rs:97-99
    return std::fabs(q0 - 4.0 * q1 + 6.0 * q2 - 4.0 * q3 + q4);
}

// scheme.rs:286-333; returns true on the first-order fallback
template <class S>
bool midpoint_flux(const S& sys, const Vn<S::N>* w, size_t j, int d, Vn<S::N>& fv, double& us) {
    constexpr int N = S::N;
    Vn<N> avg;
    for (int k = 0; k < N; ++k) avg[k] = 0.5 * (w[j][k] + w[j + 1][k]);
    Mn<N> l, r;
    sys.eigen(avg, d, l, r);
    double ch[6][N];
    for (int s = 0; s < 6; ++s) {
        const Vn<N>& ws = w[j - 2 + s];
        for (int a = 0; a < N; ++a) {
            double acc = 0.0;
            for (int k = 0; k < N; ++k) acc += l[a][k] * ws[k];
            ch[s][a] = acc;
        }
    }
    double cl[N], cr[N];
    for (int a = 0; a < N; ++a) {
        cl[a] = weno5z_interp(ch[0][a], ch[1][a], ch[2][a], ch[3][a], ch[4][a]);
        cr[a] = weno5z_interp(ch[5][a], ch[4][a], ch[3][a], ch[2][a], ch[1][a]);
    }
    Vn<N> wl, wr;
    for (int k = 0; k < N; ++k) {
        double sl = 0.0, sr = 0.0;
        for (int a = 0; a < N; ++a) {
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
    fv = sys.riemann_flux(wl, wr, d, &us);
    return fb;
}

// scheme.rs:395-464 (the face velocity follows the blend, 445-450)
template <class S>
void positivity_limit(const S& sys, const Vn<S::N>* w, size_t len, int d, size_t ng, double lambda, Vn<S::N>* out,
                      double* uface, Stats& st) {
    constexpr int N = S::N;
    const size_t n = len - 2 * ng;
    Vn<N> previous_cons = sys.to_cons(w[ng - 1]);
    Vn<N> previous_flux = sys.flux(w[ng - 1], d);
    for (size_t f = 0; f <= n; ++f) {
        const size_t i = ng - 1 + f;
        const Vn<N> ul = previous_cons, fl_phys = previous_flux;
        const Vn<N> ur = sys.to_cons(w[i + 1]);
        const Vn<N> fr_phys = sys.flux(w[i + 1], d);
        previous_cons = ur;
        previous_flux = fr_phys;
        auto ok = [&](const Vn<N>& fx) {
            Vn<N> a, b;
            for (int k = 0; k < N; ++k) {
                a[k] = ul[k] - 2.0 * lambda * (fx[k] - fl_phys[k]);
                b[k] = ur[k] + 2.0 * lambda * (fx[k] - fr_phys[k]);
            }
            return sys.admissible(sys.to_prim(a)) && sys.admissible(sys.to_prim(b));
        };
        const Vn<N> fh = out[f];
        bool finite = true;
        for (int k = 0; k < N; ++k) finite = finite && std::isfinite(fh[k]);
        if (finite && ok(fh)) continue;
        const double alpha = fmax_(sys.max_speed(w[i], d), sys.max_speed(w[i + 1], d));
        Vn<N> flf;
        for (int k = 0; k < N; ++k) flf[k] = 0.5 * (fl_phys[k] + fr_phys[k]) - 0.5 * alpha * (ur[k] - ul[k]);
        st.limited_faces += 1;
        auto blend = [&](double t) {
            Vn<N> v;
            for (int k = 0; k < N; ++k) v[k] = std::isfinite(fh[k]) ? t * fh[k] + (1.0 - t) * flf[k] : flf[k];
            return v;
        };
        double lo = 0.0, hi = 1.0;
        for (int it = 0; it < 30; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (ok(blend(mid)))
                lo = mid;
            else
                hi = mid;
        }
        out[f] = blend(lo);
        const int un = sys.vel_index() + d;
        const double um = 0.5 * (w[i][un] + w[i + 1][un]);
        uface[f] = std::isfinite(uface[f]) ? lo * uface[f] + (1.0 - lo) * um : um;
    }
}

// scheme.rs:129-281 (WCNS shock path): n+1 face fluxes and face normal
// velocities of one line of primitive states (length n + 2 ng)
template <class S>
void line_fluxes(const S& sys, const Vn<S::N>* w, size_t len, int d, size_t ng, const SchemeOpts& o, Vn<S::N>* out,
                 double* uface, Stats& st) {
    constexpr int N = S::N;
    if (o.shock != ShockScheme::Wcns) throw std::runtime_error("generic oracle: only the WCNS shock scheme");
    const size_t n = len - 2 * ng;
    const int un = sys.vel_index() + d;
    int ns = 0;
    std::vector<std::array<double, 8>> sens;
    if (o.mode == FluxMode::Hybrid) {
        sens.resize(len);
        for (size_t c = 0; c < len; ++c) ns = sys.sensor(w[c], sens[c].data());
    }
    std::vector<Vn<N>> mid;
    std::vector<double> mid_us;
    std::vector<signed char> have;
    if (o.mode != FluxMode::Central) {
        mid.resize(len);
        mid_us.resize(len);
        have.assign(len, 0);
    }
    // central pair fluxes, computed once per (lag, left cell): the reference's
    // ring cache (155-203) reuses the same values in the same order
    std::vector<Vn<N>> pair[3];
    std::vector<signed char> pair_have[3];
    for (int lg = 0; lg < 3; ++lg) {
        pair[lg].resize(len);
        pair_have[lg].assign(len, 0);
    }
    for (size_t f = 0; f <= n; ++f) {
        const size_t i = ng - 1 + f;
        double theta;
        if (o.mode == FluxMode::Central) {
            theta = 0.0;
        } else if (o.mode == FluxMode::Weno) {
            theta = 1.0;
        } else {
            double r = 0.0;
            for (int k = 0; k < ns; ++k) {
                auto s = [&](size_t j) { return sens[j][k]; };
                r = fmax_(fmax_(r, d4g(s(i - 2), s(i - 1), s(i), s(i + 1), s(i + 2))),
                          d4g(s(i - 1), s(i), s(i + 1), s(i + 2), s(i + 3)));
            }
            const double x = clamp_((r / o.tol - 0.5) / 0.5, 0.0, 1.0);
            theta = std::isfinite(r) ? x * x * (3.0 - 2.0 * x) : 1.0;
        }
        st.faces += 1;
        Vn<N> central;
        central.fill(0.0);
        double uf_c = 0.0;
        if (theta < 1.0) {
            for (int li = 0; li < 3; ++li) {
                const double a = kA6[li];
                const size_t lag = li + 1;
                for (size_t m = 0; m < lag; ++m) {
                    const size_t left = i - m;
                    if (!pair_have[li][left]) {
                        pair[li][left] = sys.split_flux(w[left], w[left + lag], d);
                        pair_have[li][left] = 1;
                    }
                    const Vn<N>& ft = pair[li][left];
                    for (int k = 0; k < N; ++k) central[k] += 2.0 * a * ft[k];
                    uf_c += a * (w[left][un] + w[left + lag][un]);
                }
            }
            if (theta == 0.0) {
                out[f] = central;
                uface[f] = uf_c;
                continue;
            }
        }
        st.weno_faces += 1;
        Vn<N> fl;
        fl.fill(0.0);
        double uf = 0.0;
        for (int s = 0; s < 3; ++s) {
            const double c = kS6[s];
            const size_t js[2] = {i + s, i - s};
            const int nj = s == 0 ? 1 : 2;
            for (int q = 0; q < nj; ++q) {
                const size_t j = js[q];
                if (!have[j]) {
                    const bool fb = midpoint_flux(sys, w, j, d, mid[j], mid_us[j]);
                    st.fallback_faces += fb ? 1 : 0;
                    have[j] = 1;
                }
                for (int k = 0; k < N; ++k) fl[k] += c * mid[j][k];
                uf += c * mid_us[j];
            }
        }
        out[f] = fl;
        uface[f] = uf;
        if (theta < 1.0) {
            for (int k = 0; k < N; ++k) out[f][k] = theta * out[f][k] + (1.0 - theta) * central[k];
            uface[f] = theta * uface[f] + (1.0 - theta) * uf_c;
        }
    }
    if (o.lambda > 0.0) positivity_limit(sys, w, len, d, ng, o.lambda, out, uface, st);
}

template <int N>
struct BcG {
    BcKind kind = BcKind::Periodic;
    Vn<N> state{};  // conserved far/inflow state (Dirichlet, Characteristic)
};

// boundary.rs:48-96
template <class S>
struct CharStencilG {
    static constexpr int N = S::N;
    Vn<N> wi, wf, w1;
    Mn<N> l, r;
    bool outgoing[N];
    CharStencilG(const S& sys, const Vn<N>& u_b, const Vn<N>& u_b1, const Vn<N>& u_far, int d, int side) {
        wi = sys.to_prim(u_b);
        wf = sys.to_prim(u_far);
        w1 = sys.to_prim(u_b1);
        sys.eigen(wi, d, l, r);
        const Vn<N> lam = sys.eigenvalues(wi, d);
        const double outward = side == 1 ? 1.0 : -1.0;
        for (int k = 0; k < N; ++k) outgoing[k] = lam[k] * outward > 0.0;
    }
    Vn<N> conserved(const S& sys, double m) const {
        Vn<N> dw;
        for (int k = 0; k < N; ++k) dw[k] = wi[k] - wf[k] + m * (wi[k] - w1[k]);
        Vn<N> wg = wf;
        for (int kf = 0; kf < N; ++kf) {
            if (outgoing[kf]) {
                double amp = -0.0;  // Rust float Sum folds from -0.0
                for (int q = 0; q < N; ++q) amp += l[kf][q] * dw[q];
                for (int q = 0; q < N; ++q) wg[q] += r[q][kf] * amp;
            }
        }
        return sys.to_cons(wg);
    }
};

// solver.rs:139-795, the plain explicit path over any System<N>
template <class S>
struct SolverG {
    static constexpr int N = S::N;
    using V = Vn<N>;
    S sys;
    Mesh mesh;
    std::vector<V> u;
    BcG<N> bc[3][2];
    double cfl = 0.5;
    FluxMode mode = FluxMode::Hybrid;
    double tol = 1e-3;
    double t = 0.0;
    bool positivity = true;
    Stats stats;
    double stage_dt = 0.0;
    std::vector<V> rhs_buf, stage;

    SolverG(const S& s, const Mesh& m) : sys(s), mesh(m) {
        V z;
        z.fill(0.0);
        u.assign(mesh.len(), z);
        rhs_buf.assign(mesh.len(), z);
        stage.assign(mesh.len(), z);
    }

    // boundary.rs:102-145
    V reflected_ghost(const std::vector<V>& st, const size_t* at_tab, long x, size_t n, size_t ng, int d,
                      const CharStencilG<S>* const ch[2]) const {
        bool flip = false;
        V v;
        for (;;) {
            if (x >= 0 && x < (long)n) {
                v = st[at_tab[ng + (size_t)x]];
                break;
            }
            const int side = x >= (long)n ? 1 : 0;
            const BcKind k = bc[d][side].kind;
            if (k == BcKind::Reflect) {
                x = side == 0 ? -1 - x : 2 * (long)n - 1 - x;
                flip = !flip;
            } else if (k == BcKind::Periodic) {
                const long r = ((x % (long)n) + (long)n) % (long)n;
                v = st[at_tab[ng + (size_t)r]];
                break;
            } else if (k == BcKind::Outflow) {
                v = st[at_tab[ng + (side == 0 ? 0 : n - 1)]];
                break;
            } else if (k == BcKind::Dirichlet) {
                v = bc[d][side].state;
                break;
            } else {
                const long layer = side == 0 ? -x : x - (long)n + 1;
                v = ch[side]->conserved(sys, (double)layer);
                break;
            }
        }
        if (flip) v[sys.momentum_index(d)] *= -1.0;
        return v;
    }

    // solver.rs:259-348
    void fill_ghosts(std::vector<V>& uu) const {
        const Mesh& m = mesh;
        bool deep = false;
        for (int d = 0; d < 3; ++d)
            if (m.active(d) && m.n[d] < m.ng &&
                (bc[d][0].kind == BcKind::Reflect || bc[d][1].kind == BcKind::Reflect))
                deep = true;
        for (int d = 0; d < 3; ++d) {
            if (!m.active(d)) continue;
            const size_t n = m.n[d], ng = m.ng;
            int others[2], no = 0;
            for (int e = 0; e < 3; ++e)
                if (e != d) others[no++] = e;
            const int e0 = others[0], e1 = others[1];
            std::vector<size_t> at_tab(m.m(d));
            const int mi = sys.momentum_index(d);
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
                    std::vector<CharStencilG<S>> chv;
                    chv.reserve(2);
                    const CharStencilG<S>* ch[2] = {nullptr, nullptr};
                    if (bc[d][0].kind == BcKind::Characteristic) {
                        chv.emplace_back(sys, uu[at(ng)], uu[at(ng + 1)], bc[d][0].state, d, 0);
                        ch[0] = &chv.back();
                    }
                    if (bc[d][1].kind == BcKind::Characteristic) {
                        chv.emplace_back(sys, uu[at(n + ng - 1)], uu[at(n + ng - 2)], bc[d][1].state, d, 1);
                        ch[1] = &chv.back();
                    }
                    for (size_t gi = 0; gi < ng; ++gi) {
                        const size_t lo_dst = at(gi), hi_dst = at(n + ng + gi);
                        size_t lo_s, hi_s;
                        m.periodic_ghost_sources(d, gi, lo_s, hi_s);
                        const size_t lo_src = at(lo_s), hi_src = at(hi_s);
                        const size_t lo_ref = at(2 * ng - 1 - gi), hi_ref = at(n + ng - 1 - gi);
                        V v;
                        switch (bc[d][0].kind) {
                            case BcKind::Periodic: v = uu[lo_src]; break;
                            case BcKind::Outflow: v = uu[at(ng)]; break;
                            case BcKind::Dirichlet: v = bc[d][0].state; break;
                            case BcKind::Characteristic: v = ch[0]->conserved(sys, (double)(ng - gi)); break;
                            case BcKind::Reflect:
                                if (deep && n < ng) {
                                    v = reflected_ghost(uu, at_tab.data(), (long)gi - (long)ng, n, ng, d, ch);
                                } else {
                                    v = uu[lo_ref];
                                    v[mi] *= -1.0;
                                }
                                break;
                        }
                        uu[lo_dst] = v;
                        switch (bc[d][1].kind) {
                            case BcKind::Periodic: v = uu[hi_src]; break;
                            case BcKind::Outflow: v = uu[at(n + ng - 1)]; break;
                            case BcKind::Dirichlet: v = bc[d][1].state; break;
                            case BcKind::Characteristic: v = ch[1]->conserved(sys, (double)(gi + 1)); break;
                            case BcKind::Reflect:
                                if (deep && n < ng) {
                                    v = reflected_ghost(uu, at_tab.data(), (long)(n + gi), n, ng, d, ch);
                                } else {
                                    v = uu[hi_ref];
                                    v[mi] *= -1.0;
                                }
                                break;
                        }
                        uu[hi_dst] = v;
                    }
                }
            }
        }
    }

    // solver.rs:350-475: -div F plus c(w_i) (U_{i+1/2} - U_{i-1/2}) / dx_d
    // for the non-conservative variable, axes in order
    Stats rhs(std::vector<V>& uu) {
        fill_ghosts(uu);
        std::vector<V>& out = rhs_buf;
        for (auto& v : out) v.fill(0.0);
        const Mesh& m = mesh;
        Stats total;
        double ndim = 0.0;
        for (int e = 0; e < 3; ++e) ndim += m.active(e) ? 1.0 : 0.0;
        const int nc = sys.nc_var();
        for (int d = 0; d < 3; ++d) {
            if (!m.active(d)) continue;
            int others[2], no = 0;
            for (int e = 0; e < 3; ++e)
                if (e != d) others[no++] = e;
            const int e0 = others[0], e1 = others[1];
            const SchemeOpts opts{mode, ShockScheme::Wcns, tol, positivity ? ndim * stage_dt / m.dx[d] : 0.0};
            const long np = (long)(m.n[e0] * m.n[e1]);
            const size_t len = m.m(d);
            const double inv = 1.0 / m.dx[d];
            uint64_t sf = 0, swf = 0, sfb = 0, slim = 0;
#pragma omp parallel reduction(+ : sf, swf, sfb, slim)
            {
                std::vector<V> w(len), fl(m.n[d] + 1);
                std::vector<double> uf(m.n[d] + 1);
                std::vector<size_t> at_tab(len);
#pragma omp for schedule(static)
                for (long p = 0; p < np; ++p) {
                    const size_t a = (size_t)p / m.n[e1], b = (size_t)p % m.n[e1];
                    for (size_t c = 0; c < len; ++c) {
                        size_t ijk[3];
                        ijk[d] = c;
                        ijk[e0] = a + m.g(e0);
                        ijk[e1] = b + m.g(e1);
                        at_tab[c] = m.idx(ijk[0], ijk[1], ijk[2]);
                        w[c] = sys.to_prim(uu[at_tab[c]]);
                    }
                    Stats s;
                    line_fluxes(sys, w.data(), len, d, m.ng, opts, fl.data(), uf.data(), s);
                    for (size_t c = 0; c < m.n[d]; ++c) {
                        V& o = out[at_tab[c + m.ng]];
                        for (int k = 0; k < N; ++k) o[k] -= (fl[c + 1][k] - fl[c][k]) * inv;
                        if (nc >= 0) o[nc] += sys.nc_coef(w[c + m.ng]) * (uf[c + 1] - uf[c]) * inv;
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

    // solver.rs:592-640: the conserved neighbour of interior cell c along d
    // through the boundary condition when c is a boundary cell
    V neighbor_conserved(const size_t c[3], int d, bool high) const {
        const Mesh& m = mesh;
        const bool boundary = high ? c[d] == m.g(d) + m.n[d] - 1 : c[d] == m.g(d);
        size_t q[3] = {c[0], c[1], c[2]};
        if (!boundary) {
            if (high)
                q[d] += 1;
            else
                q[d] -= 1;
            return u[m.idx(q[0], q[1], q[2])];
        }
        const V center = u[m.idx(c[0], c[1], c[2])];
        const BcG<N>& b = bc[d][high ? 1 : 0];
        switch (b.kind) {
            case BcKind::Periodic:
                q[d] = high ? m.g(d) : m.g(d) + m.n[d] - 1;
                return u[m.idx(q[0], q[1], q[2])];
            case BcKind::Outflow: return center;
            case BcKind::Dirichlet: return b.state;
            case BcKind::Reflect: {
                V v = center;
                v[sys.momentum_index(d)] *= -1.0;
                return v;
            }
            default: {
                if (high)
                    q[d] -= 1;
                else
                    q[d] += 1;
                return CharStencilG<S>(sys, center, u[m.idx(q[0], q[1], q[2])], b.state, d, high ? 1 : 0)
                    .conserved(sys, 1.0);
            }
        }
    }

    // solver.rs:565-584
    void neighbor_velocities(const size_t c[3], int d, double& va, double& vb) const {
        const Mesh& m = mesh;
        if (c[d] > m.g(d) && c[d] < m.g(d) + m.n[d] - 1) {
            const size_t q = m.idx(c[0], c[1], c[2]);
            const size_t stride = d == 0 ? m.m(1) * m.m(2) : d == 1 ? m.m(2) : 1;
            va = sys.velocity_conserved(u[q - stride], d);
            vb = sys.velocity_conserved(u[q + stride], d);
        } else {
            va = sys.velocity_conserved(neighbor_conserved(c, d, false), d);
            vb = sys.velocity_conserved(neighbor_conserved(c, d, true), d);
        }
    }

    // solver.rs:478-523 (no viscous limit)
    double dt_cfl() const {
        const Mesh& m = mesh;
        const bool need_source = sys.nc_var() >= 0;
        double rate = 0.0, source = 0.0;
#pragma omp parallel for reduction(max : rate, source) schedule(static)
        for (long i = 0; i < (long)m.n[0]; ++i) {
            double rr = 0.0, ss = 0.0;
            for (size_t j = 0; j < m.n[1]; ++j)
                for (size_t k = 0; k < m.n[2]; ++k) {
                    const size_t c[3] = {(size_t)i + m.g(0), j + m.g(1), k + m.g(2)};
                    const V w = sys.to_prim(u[m.idx(c[0], c[1], c[2])]);
                    double acoustic = -0.0;
                    for (int d = 0; d < 3; ++d)
                        if (m.active(d)) acoustic += sys.max_speed(w, d) / m.dx[d];
                    rr = fmax_(rr, acoustic);
                    if (need_source) {
                        double divu = 0.0;
                        for (int d = 0; d < 3; ++d) {
                            if (!m.active(d)) continue;
                            double va, vb;
                            neighbor_velocities(c, d, va, vb);
                            divu += std::fabs(vb - va) / (2.0 * m.dx[d]);
                        }
                        ss = fmax_(ss, sys.dilatation_rate(w, divu));
                    }
                }
            rate = fmax_(rate, rr);
            source = fmax_(source, ss);
        }
        double dt = cfl / rate;
        if (source > 0.0) dt = fmin_(dt, 0.5 / source);
        return dt;
    }

    // solver.rs:97-137
    template <int STAGE>
    void rk_combine(std::vector<V>& dest, const std::vector<V>& old, const std::vector<V>& res, double dt) const {
        const Mesh& m = mesh;
#pragma omp parallel for schedule(static)
        for (long ii = 0; ii < (long)m.n[0]; ++ii) {
            const size_t i = (size_t)ii + m.g(0);
            for (size_t j = m.g(1); j < m.g(1) + m.n[1]; ++j)
                for (size_t k = m.g(2); k < m.g(2) + m.n[2]; ++k) {
                    const size_t q = m.idx(i, j, k);
                    for (int v = 0; v < N; ++v) {
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

    // solver.rs:659-738
    void step(double dt) {
        stage_dt = dt;
        const Stats s1 = rhs(u);
        stage = u;
        rk_combine<1>(stage, u, rhs_buf, dt);
        const Stats s2 = rhs(stage);
        rk_combine<2>(stage, u, rhs_buf, dt);
        const Stats s3 = rhs(stage);
        rk_combine<3>(u, stage, rhs_buf, dt);
        stats.add(s1);
        stats.add(s2);
        stats.add(s3);
        t += dt;
    }

    // solver.rs:758-768
    size_t run_to(double t_end) {
        size_t steps = 0;
        while (t < t_end - 1e-10 * std::fabs(t_end)) {
            const double dt = fmin_(dt_cfl(), t_end - t);
            if (!(std::isfinite(dt) && dt > 1e-12 * fmax_(std::fabs(t_end), 1e-300)))
                throw std::runtime_error("time step collapsed: the solution lost admissibility");
            step(dt);
            ++steps;
        }
        return steps;
    }

    // FNV-1a 64 over interior conserved values (examples/cpu_validate.rs:7-22)
    uint64_t hash() const {
        const Mesh& m = mesh;
        uint64_t h = 0xcbf29ce484222325ull;
        for (size_t i = 0; i < m.n[0]; ++i)
            for (size_t j = 0; j < m.n[1]; ++j)
                for (size_t k = 0; k < m.n[2]; ++k)
                    for (double v : u[m.idx(i + m.g(0), j + m.g(1), k + m.g(2))]) {
                        uint64_t bits;
                        std::memcpy(&bits, &v, 8);
                        h = (h ^ bits) * 0x100000001b3ull;
                    }
        return h;
    }
};

}  // namespace gen
}  // namespace emref
