// C ABI of the generic oracle solver (emref_gen.h) over the Euler and FiveEq
// systems (emref_multi.h), for the Python tests.  The generic Euler must give
// the Euler oracle's results, which the reference's golden hashes pin, so
// the generic code is pinned too.  Test infrastructure only.
#include <cstdint>
#include <cstring>

#include "emref_multi.h"


using namespace emref;
using namespace emref::gen;

namespace {

// One handle: a generic solver of either system (sys 0 Euler, 1 FiveEq)
struct GenHandle {
    int sys;
    SolverG<EulerG>* e = nullptr;
    SolverG<FiveEq>* f = nullptr;
    ~GenHandle() {
        delete e;
        delete f;
    }
    int nv() const { return sys == 0 ? 5 : 7; }
    template <class F>
    auto with(F&& fn) {
        if (sys == 0) return fn(*e);
        return fn(*f);
    }
};

}  // namespace

extern "C" {

// sys 0: Euler, params = {gamma, p_inf, b, q}
// sys 1: FiveEq, params = {gamma1, p_inf1, b1, q1, gamma2, p_inf2, b2, q2, kapila}
void* emref_gen_create(int sys, const double* params, const int64_t n[3], const double lo[3], const double hi[3],
                       int64_t ng) {
    const size_t nn[3] = {(size_t)n[0], (size_t)n[1], (size_t)n[2]};
    Mesh m(nn, lo, hi, (size_t)ng);
    GenHandle* h = new GenHandle{sys};
    if (sys == 0) {
        EulerG s;
        s.e.eos = Nasg{params[0], params[1], params[2], params[3]};
        h->e = new SolverG<EulerG>(s, m);
    } else {
        FiveEq s;
        s.eos[0] = Nasg{params[0], params[1], params[2], params[3]};
        s.eos[1] = Nasg{params[4], params[5], params[6], params[7]};
        s.kapila = params[8] != 0.0;
        h->f = new SolverG<FiveEq>(s, m);
    }
    return h;
}

void emref_gen_destroy(void* h) { delete (GenHandle*)h; }
int emref_gen_nv(void* h) { return ((GenHandle*)h)->nv(); }

void emref_gen_set_bc(void* hh, int d, int side, int kind, const double* state) {
    GenHandle* h = (GenHandle*)hh;
    h->with([&](auto& s) {
        s.bc[d][side].kind = (BcKind)kind;
        for (int k = 0; k < h->nv(); ++k) s.bc[d][side].state[k] = state ? state[k] : 0.0;
        return 0;
    });
}

void emref_gen_set_opts(void* hh, int mode, double tol, int positivity, double cfl) {
    ((GenHandle*)hh)->with([&](auto& s) {
        s.mode = (FluxMode)mode;
        s.tol = tol;
        s.positivity = positivity != 0;
        s.cfl = cfl;
        return 0;
    });
}

void emref_gen_set_u(void* hh, const double* u) {
    ((GenHandle*)hh)->with([&](auto& s) {
        std::memcpy((void*)s.u.data(), u, sizeof(s.u[0]) * s.u.size());
        return 0;
    });
}

void emref_gen_get_u(void* hh, double* u) {
    ((GenHandle*)hh)->with([&](auto& s) {
        std::memcpy(u, (const void*)s.u.data(), sizeof(s.u[0]) * s.u.size());
        return 0;
    });
}

// Per-stage residual of the current state (ghosts refreshed), limiter lambda
// from stage_dt; out in the ghost-padded layout with zero ghosts.
void emref_gen_rhs(void* hh, double stage_dt, double* out, uint64_t stats[4]) {
    ((GenHandle*)hh)->with([&](auto& s) {
        s.stage_dt = stage_dt;
        const Stats st = s.rhs(s.u);
        std::memcpy(out, (const void*)s.rhs_buf.data(), sizeof(s.rhs_buf[0]) * s.rhs_buf.size());
        if (stats) {
            stats[0] = st.faces;
            stats[1] = st.weno_faces;
            stats[2] = st.fallback_faces;
            stats[3] = st.limited_faces;
        }
        return 0;
    });
}

void emref_gen_step(void* hh, double dt) {
    ((GenHandle*)hh)->with([&](auto& s) {
        s.step(dt);
        return 0;
    });
}

double emref_gen_dt_cfl(void* hh) {
    return ((GenHandle*)hh)->with([&](auto& s) { return s.dt_cfl(); });
}

int64_t emref_gen_run_to(void* hh, double t_end) {
    return ((GenHandle*)hh)->with([&](auto& s) -> int64_t {
        try {
            return (int64_t)s.run_to(t_end);
        } catch (const std::exception&) {
            return -1;
        }
    });
}

double emref_gen_time(void* hh) {
    return ((GenHandle*)hh)->with([&](auto& s) { return s.t; });
}

void emref_gen_stats(void* hh, uint64_t out[4]) {
    ((GenHandle*)hh)->with([&](auto& s) {
        out[0] = s.stats.faces;
        out[1] = s.stats.weno_faces;
        out[2] = s.stats.fallback_faces;
        out[3] = s.stats.limited_faces;
        return 0;
    });
}

uint64_t emref_gen_hash(void* hh) {
    return ((GenHandle*)hh)->with([&](auto& s) { return s.hash(); });
}

// Conserved <-> primitive of one state (FiveEq handles), for the tests
void emref_gen_to_prim(void* hh, const double* u, double* w) {
    GenHandle* h = (GenHandle*)hh;
    if (h->sys == 0) {
        Vn<5> uu;
        std::memcpy(uu.data(), u, sizeof(uu));
        const Vn<5> ww = h->e->sys.to_prim(uu);
        std::memcpy(w, ww.data(), sizeof(ww));
    } else {
        Vn<7> uu;
        std::memcpy(uu.data(), u, sizeof(uu));
        const Vn<7> ww = h->f->sys.to_prim(uu);
        std::memcpy(w, ww.data(), sizeof(ww));
    }
}

// Eigenvectors (rows of L, columns of R, row-major nv x nv), HLLC flux with
// its contact speed, and the physical flux of primitive states (FiveEq
// handles), for the closed-form tests
void emref_gen_eigen(void* hh, const double* w, int d, double* l, double* r) {
    GenHandle* h = (GenHandle*)hh;
    Vn<7> ww;
    std::memcpy(ww.data(), w, sizeof(ww));
    Mn<7> ll, rr;
    h->f->sys.eigen(ww, d, ll, rr);
    for (int a = 0; a < 7; ++a)
        for (int b = 0; b < 7; ++b) {
            l[a * 7 + b] = ll[a][b];
            r[a * 7 + b] = rr[a][b];
        }
}

void emref_gen_riemann(void* hh, const double* wl, const double* wr, int d, double* f, double* ss) {
    GenHandle* h = (GenHandle*)hh;
    Vn<7> a, b;
    std::memcpy(a.data(), wl, sizeof(a));
    std::memcpy(b.data(), wr, sizeof(b));
    const Vn<7> ff = h->f->sys.riemann_flux(a, b, d, ss);
    std::memcpy(f, ff.data(), sizeof(ff));
}

void emref_gen_phys_flux(void* hh, const double* w, int d, double* f) {
    GenHandle* h = (GenHandle*)hh;
    Vn<7> a;
    std::memcpy(a.data(), w, sizeof(a));
    const Vn<7> ff = h->f->sys.flux(a, d);
    std::memcpy(f, ff.data(), sizeof(ff));
}

// Initial state of the reference's `interface` probe: a water drop
// (alpha = 1) in air, smeared over one cell by tanh, moving at u = 0.1 in
// p = 1e5.
void emref_gen_init_interface(void* hh) {
    GenHandle* h = (GenHandle*)hh;
    if (h->sys == 1) interface_init(*h->f);
}

}  // extern "C"
