// C ABI of the oracle AMR (test infrastructure only).  Block data is exported
// block by block in id order: per block bcells x 5 conserved values, cell
// index (i*b1 + j)*b2 + k (amr.rs:278-281).
#include <cmath>
#include <cstring>
#include <map>

#include "emref_amr.h"

using namespace emref;

// Blocks protected by the last regrid_refine of a handle, kept until its
// next regrid_coarsen (or destroy).
static std::map<void*, std::set<AmrKey>> g_protected;

namespace {

struct CaseCtx {
    int id;
    double p[16];
};

// Primitive state at a point for the built-in workloads.
//  0 four-quadrant 2-D Riemann (examples/amr_validate.rs:53-60), split p[0], p[1]
//  1 smooth wave 1 + 0.1 sin(2 pi x) cos(2 pi y), velocity (0.3, 0.2, 0.1), p = 1
//  2 shock-bubble: Mach p[0] shock at x = p[1] moving to +x into (rho 1, p 1),
//    sphere centre (p[2], p[3], p[4]) radius p[5] density p[6] (gamma 1.4)
//  3 isentropic vortex (examples/amr_validate.rs:44-51)
//  4 polynomial state at rest: rho = 1 + sum_d c_d x_d^5 / 5, p = 1 + sum_d c_d x_d^4 / 4
//    with c = (p[0], p[1], p[2]) (exactness checks of prediction / restriction)
V init_case(const double x[3], const void* vctx) {
    const CaseCtx* c = (const CaseCtx*)vctx;
    const double* p = c->p;
    switch (c->id) {
        case 0: {
            bool e = x[0] > p[0], n = x[1] > p[1];
            if (e && n) return V{1.5, 0.0, 0.0, 0.0, 1.5};
            if (!e && n) return V{0.5323, 1.206, 0.0, 0.0, 0.3};
            if (!e && !n) return V{0.138, 1.206, 1.206, 0.0, 0.029};
            return V{0.5323, 0.0, 1.206, 0.0, 0.3};
        }
        case 1: {
            const double TAU = 6.283185307179586;
            return V{1.0 + 0.1 * std::sin(TAU * x[0]) * std::cos(TAU * x[1]), 0.3, 0.2, 0.1, 1.0};
        }
        case 2: {
            const double g = 1.4, M = p[0];
            double r2 = 0.0;
            for (int d = 0; d < 3; ++d) r2 += (x[d] - p[2 + d]) * (x[d] - p[2 + d]);
            if (x[0] < p[1]) {
                // Rankine-Hugoniot post-shock state for a shock of Mach M into rest (rho 1, p 1)
                double rho2 = (g + 1.0) * M * M / ((g - 1.0) * M * M + 2.0);
                double p2 = 1.0 + 2.0 * g / (g + 1.0) * (M * M - 1.0);
                double c1 = std::sqrt(g);
                double u2 = M * c1 * (1.0 - 1.0 / rho2);
                return V{rho2, u2, 0.0, 0.0, p2};
            }
            if (r2 < p[5] * p[5]) return V{p[6], 0.0, 0.0, 0.0, 1.0};
            return V{1.0, 0.0, 0.0, 0.0, 1.0};
        }
        case 3: {
            const double gam = 1.4, beta = 5.0, pi = 3.141592653589793;
            double r2 = x[0] * x[0] + x[1] * x[1];
            double f = beta / (2.0 * pi) * std::exp(0.5 * (1.0 - r2));
            double tt = 1.0 - (gam - 1.0) * beta * beta / (8.0 * gam * pi * pi) * std::exp(1.0 - r2);
            double rho = std::pow(tt, 1.0 / (gam - 1.0));
            return V{rho, 1.0 - f * x[1], 1.0 + f * x[0], 0.0, rho * tt};
        }
        default: {
            double rho = 1.0, pr = 1.0;
            for (int d = 0; d < 3; ++d) {
                double x2 = x[d] * x[d];
                rho += p[d] * x2 * x2 * x[d] / 5.0;
                pr += p[d] * x2 * x2 / 4.0;
            }
            return V{rho, 0.0, 0.0, 0.0, pr};
        }
    }
}

}  // namespace

extern "C" {

void* emref_amr_create(int case_id, const double* params, const int64_t base[3], int64_t b, int64_t levels,
                       double eps, int64_t regrid_every, int64_t ghosts, const double lo[3], const double hi[3],
                       const double eos[4], const int32_t bck[6], const double bcs[30], int mode, double cfl,
                       int positivity, double tol) {
    CaseCtx ctx;
    ctx.id = case_id;
    for (int i = 0; i < 16; ++i) ctx.p[i] = params ? params[i] : 0.0;
    AmrConfig c;
    c.b = (size_t)b;
    c.levels = (size_t)levels;
    c.eps = eps;
    c.regrid_every = (size_t)regrid_every;
    for (int d = 0; d < 3; ++d) c.base_blocks[d] = (size_t)base[d];
    c.ghosts = (size_t)ghosts;
    Euler sys;
    sys.eos = Nasg{eos[0], eos[1], eos[2], eos[3]};
    Bc bc[3][2];
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s) {
            bc[d][s].kind = (BcKind)bck[2 * d + s];
            for (int q = 0; q < NV; ++q) bc[d][s].state[q] = bcs ? bcs[(2 * d + s) * 5 + q] : 0.0;
        }
    Amr* a = new Amr(sys, c, lo, hi, init_case, &ctx, bc);
    a->mode = (FluxMode)mode;
    a->cfl = cfl;
    a->positivity = positivity != 0;
    a->tol = tol;
    return a;
}

void emref_amr_destroy(void* h) {
    g_protected.erase(h);
    delete (Amr*)h;
}
int64_t emref_amr_nblocks(void* h) { return (int64_t)((Amr*)h)->blocks.size(); }
int64_t emref_amr_bcells(void* h) { return (int64_t)((Amr*)h)->bcells(); }
int64_t emref_amr_ng(void* h) { return (int64_t)((Amr*)h)->ng; }

void emref_amr_blocks(void* h, int32_t* level, int64_t* idx, uint8_t* leaf) {
    Amr* a = (Amr*)h;
    for (size_t i = 0; i < a->blocks.size(); ++i) {
        level[i] = (int32_t)a->blocks[i].level;
        for (int d = 0; d < 3; ++d) idx[3 * i + d] = a->blocks[i].idx[d];
        leaf[i] = a->blocks[i].leaf ? 1 : 0;
    }
}

void emref_amr_get_u(void* h, double* u) {
    Amr* a = (Amr*)h;
    size_t nc = a->bcells();
    for (size_t i = 0; i < a->blocks.size(); ++i) std::memcpy(u + i * nc * 5, a->blocks[i].u.data(), nc * 5 * 8);
}

void emref_amr_set_u(void* h, const double* u) {
    Amr* a = (Amr*)h;
    size_t nc = a->bcells();
    for (size_t i = 0; i < a->blocks.size(); ++i) std::memcpy(a->blocks[i].u.data(), u + i * nc * 5, nc * 5 * 8);
}

void emref_amr_step(void* h, double dt) { ((Amr*)h)->step(dt); }
double emref_amr_dt_cfl(void* h) { return ((Amr*)h)->dt_cfl(); }
double emref_amr_time(void* h) { return ((Amr*)h)->t; }
void emref_amr_restrict_all(void* h) { ((Amr*)h)->restrict_all(); }
void emref_amr_regrid(void* h) { ((Amr*)h)->regrid(); }
void emref_amr_set_next_regrid(void* h, int64_t next) { ((Amr*)h)->next_regrid = (size_t)next; }

// Residual of every leaf (restriction, ghosts, fluxes, reflux) with the
// limiter's lambda from stage_dt; non-leaf blocks are written as zero.
void emref_amr_compute_rhs(void* h, double stage_dt, double* rhs, uint64_t stats[4]) {
    Amr* a = (Amr*)h;
    a->stage_dt = stage_dt;
    Stats st = a->compute_rhs();
    size_t nc = a->bcells();
    for (size_t i = 0; i < a->blocks.size(); ++i) {
        if (a->blocks[i].leaf)
            std::memcpy(rhs + i * nc * 5, a->blocks[i].rhs.data(), nc * 5 * 8);
        else
            std::memset(rhs + i * nc * 5, 0, nc * 5 * 8);
    }
    if (stats) {
        stats[0] = st.faces;
        stats[1] = st.weno_faces;
        stats[2] = st.fallback_faces;
        stats[3] = st.limited_faces;
    }
}

// Ghost-extended box of block id (fill_scratch), (b+2ng)^D x 5.
void emref_amr_fill_scratch(void* h, int64_t id, double* out) {
    Amr* a = (Amr*)h;
    std::vector<V> s = a->fill_scratch((size_t)id);
    std::memcpy(out, s.data(), s.size() * 5 * 8);
}

void emref_amr_details(void* h, double* det) {
    Amr* a = (Amr*)h;
    for (size_t i = 0; i < a->blocks.size(); ++i) det[i] = a->blocks[i].leaf ? a->block_detail(i) : 0.0;
}

void emref_amr_totals(void* h, double out[5]) {
    V t = ((Amr*)h)->totals();
    for (int q = 0; q < NV; ++q) out[q] = t[q];
}

void emref_amr_stats(void* h, uint64_t out[4]) {
    Stats& st = ((Amr*)h)->stats;
    out[0] = st.faces;
    out[1] = st.weno_faces;
    out[2] = st.fallback_faces;
    out[3] = st.limited_faces;
}

// Regrid rules with caller-given flags (amr.rs:877-892, 1051-1106,
// 1545-1548), for tests that pin them directly.

void emref_amr_flags_from_details(void* h, const double* det, int8_t* flags) {
    Amr* a = (Amr*)h;
    std::vector<double> d(det, det + a->blocks.size());
    std::vector<int8_t> f = a->flags_from_details(d);
    std::memcpy(flags, f.data(), f.size());
}

void emref_amr_regrid_refine(void* h, const int8_t* flags) {
    Amr* a = (Amr*)h;
    std::vector<int8_t> f(flags, flags + a->blocks.size());
    g_protected[h] = a->regrid_refine(f);
}

void emref_amr_regrid_coarsen(void* h, const int8_t* flags) {
    Amr* a = (Amr*)h;
    std::vector<int8_t> f(flags, flags + a->blocks.size());
    a->regrid_coarsen(f, g_protected[h]);
    g_protected.erase(h);
}

// schedule_regrid(changed); returns steps, regrid interval, next regrid step
void emref_amr_schedule_regrid(void* h, int changed, int64_t out[3]) {
    Amr* a = (Amr*)h;
    if (changed >= 0) a->schedule_regrid(changed != 0);
    out[0] = (int64_t)a->steps;
    out[1] = (int64_t)a->regrid_interval;
    out[2] = (int64_t)a->next_regrid;
}

}  // extern "C"

