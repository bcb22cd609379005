// C ABI of the oracle for the Python tests and the bench's CPU-baseline leg
// (ctypes).  Test infrastructure only: the product library never links it.
// Layout of state arrays: the reference's ghost-padded AoS, cell index
// (i*m1 + j)*m2 + k, five conserved values per cell (solver.rs:72-74).
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>

#include "emref.h"

using namespace emref;

extern "C" {

void* emref_create(const int64_t n[3], const double lo[3], const double hi[3], int64_t ng, const double eos[4]) {
    size_t nn[3] = {(size_t)n[0], (size_t)n[1], (size_t)n[2]};
    Mesh m(nn, lo, hi, (size_t)ng);
    Euler sys;
    sys.eos = Nasg{eos[0], eos[1], eos[2], eos[3]};
    return new Solver(sys, m);
}

void emref_destroy(void* h) { delete (Solver*)h; }

int64_t emref_len(void* h) { return (int64_t)((Solver*)h)->mesh.len(); }

void emref_set_bc(void* h, int d, int side, int kind, const double* state) {
    Solver* s = (Solver*)h;
    s->bc[d][side].kind = (BcKind)kind;
    for (int k = 0; k < NV; ++k) s->bc[d][side].state[k] = state ? state[k] : 0.0;
}

void emref_set_opts(void* h, int mode, int shock, double tol, int positivity, double cfl) {
    Solver* s = (Solver*)h;
    s->mode = (FluxMode)mode;
    s->shock = (ShockScheme)shock;
    s->tol = tol;
    s->positivity = positivity != 0;
    s->cfl = cfl;
}

void emref_set_u(void* h, const double* u) {
    Solver* s = (Solver*)h;
    std::memcpy(s->u.data(), u, sizeof(V) * s->u.size());
}

void emref_get_u(void* h, double* u) {
    Solver* s = (Solver*)h;
    std::memcpy(u, s->u.data(), sizeof(V) * s->u.size());
}

// Initial state of the reference's `open` probe (examples/cpu_validate.rs:
// 68-80): a Gaussian pressure pulse of 1e-4 at the centre of [0,1]^3 at rest,
// rho = 1 + dp/1.4, p = 1 + dp, dp = 1e-4 exp(-r^2/0.04).
void emref_init_open(void* h) {
    Solver* s = (Solver*)h;
    const Mesh& m = s->mesh;
    for (size_t i = 0; i < m.n[0]; ++i)
        for (size_t j = 0; j < m.n[1]; ++j)
            for (size_t k = 0; k < m.n[2]; ++k) {
                const double xs[3] = {m.x(i, 0), m.x(j, 1), m.x(k, 2)};
                double r2 = -0.0;  // Rust's f64 Sum starts from -0.0
                for (int d = 0; d < 3; ++d) r2 += (xs[d] - 0.5) * (xs[d] - 0.5);
                const double dp = 1e-4 * std::exp(-r2 / 0.04);
                s->u[m.idx(i + m.g(0), j + m.g(1), k + m.g(2))] = s->sys.to_cons(V{1.0 + dp / 1.4, 0.0, 0.0, 0.0, 1.0 + dp});
            }
}

void emref_fill_ghosts(void* h) {
    Solver* s = (Solver*)h;
    s->fill_ghosts(s->u);
}

// Per-stage residual of the current state (ghosts refreshed first), with the
// limiter's lambda taken from stage_dt (solver.rs:392-393).
void emref_rhs(void* h, double stage_dt, double* out, uint64_t stats[4]) {
    Solver* s = (Solver*)h;
    s->stage_dt = stage_dt;
    Stats st = s->rhs(s->u);
    std::memcpy(out, s->rhs_buf.data(), sizeof(V) * s->rhs_buf.size());
    if (stats) {
        stats[0] = st.faces;
        stats[1] = st.weno_faces;
        stats[2] = st.fallback_faces;
        stats[3] = st.limited_faces;
    }
}

void emref_step(void* h, double dt) { ((Solver*)h)->step(dt); }

double emref_dt_cfl(void* h) { return ((Solver*)h)->dt_cfl(); }

int64_t emref_run_to(void* h, double t_end) {
    try {
        return (int64_t)((Solver*)h)->run_to(t_end);
    } catch (const std::exception&) {
        return -1;
    }
}

double emref_time(void* h) { return ((Solver*)h)->t; }

void emref_stats(void* h, uint64_t out[4]) {
    Stats& st = ((Solver*)h)->stats;
    out[0] = st.faces;
    out[1] = st.weno_faces;
    out[2] = st.fallback_faces;
    out[3] = st.limited_faces;
}

void emref_totals(void* h, double out[5]) {
    V t = ((Solver*)h)->totals();
    for (int k = 0; k < NV; ++k) out[k] = t[k];
}

uint64_t emref_hash(void* h) { return state_hash(*(Solver*)h); }

void emref_set_threads(int n) { omp_set_num_threads(n); }
int emref_max_threads(void) { return omp_get_max_threads(); }

// ---- point-level helpers for unit tests of the scheme's building blocks
double emref_weno5z_interp(double v0, double v1, double v2, double v3, double v4) {
    return weno5z_interp(v0, v1, v2, v3, v4);
}

void emref_riemann_flux(const double eos[4], const double wl[5], const double wr[5], int d, double f[5],
                        double* s_star) {
    Euler sys;
    sys.eos = Nasg{eos[0], eos[1], eos[2], eos[3]};
    V a, b;
    for (int k = 0; k < NV; ++k) {
        a[k] = wl[k];
        b[k] = wr[k];
    }
    V r = sys.riemann_flux(a, b, d, s_star);
    for (int k = 0; k < NV; ++k) f[k] = r[k];
}

// One line of primitive states (len x 5) -> n+1 face fluxes ((len-2ng+1) x 5).
void emref_line_fluxes(const double eos[4], const double* w, int64_t len, int d, int64_t ng, int mode, int shock,
                       double tol, double lambda, double* out, uint64_t stats[4]) {
    Euler sys;
    sys.eos = Nasg{eos[0], eos[1], eos[2], eos[3]};
    SchemeOpts o{(FluxMode)mode, (ShockScheme)shock, tol, lambda};
    Stats st;
    line_fluxes(sys, (const V*)w, (size_t)len, d, (size_t)ng, o, (V*)out, st);
    if (stats) {
        stats[0] = st.faces;
        stats[1] = st.weno_faces;
        stats[2] = st.fallback_faces;
        stats[3] = st.limited_faces;
    }
}

}  // extern "C"
