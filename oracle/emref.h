// Oracle: a CPU restatement of the reference's compressible finite-volume
// solver (Euler system, ideal/NASG equation of state), used ONLY as the
// checker of the HIP path (tests, smoke check, CPU baseline leg of the bench).
// It is test infrastructure: nothing in the product library links it.
//
// Restated from (paths relative to the reference repository root):
//   fluid/crates/compressible/src/eos.rs        (Nasg)
//   fluid/crates/compressible/src/system.rs     (Euler: to_prim, to_cons,
//                                                eigen, HLLC, KEEP-PE, dudw)
//   fluid/crates/compressible/src/scheme.rs     (line_fluxes: Hybrid / Weno /
//                                                Central, WCNS, FD-WENO,
//                                                positivity limiter)
//   fluid/crates/compressible/src/solver.rs     (Mesh, Bc, ghosts, rhs,
//                                                dt_cfl, SSP-RK3 step, run_to)
//   fluid/crates/compressible/src/boundary.rs   (characteristic ghosts, deep
//                                                reflections)
//
// Arithmetic discipline: IEEE binary64, no FMA contraction, no fast-math,
// every expression evaluated in the reference's left-to-right order, so the
// state after N steps is bit-identical to the reference CPU solver.  This is
// checked against the FNV-1a state hashes the reference recorded for its
// `cpu_validate` probe (tests/test_oracle_golden.py).
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace emref {

constexpr int NV = 5;
using V = std::array<double, NV>;
using Mat = std::array<V, NV>;

// eos.rs:12-80
struct Nasg {
    double gamma = 1.4, p_inf = 0.0, b = 0.0, q = 0.0;
    double pressure(double rho, double e) const;
    double energy(double rho, double p) const;
    double rho_e(double rho, double p) const;
    double sound_speed_sq(double rho, double p) const;
    bool admissible(double rho, double p) const;
};

// system.rs:190-414 (Euler as System<5>)
struct Euler {
    Nasg eos;
    V to_prim(const V& u) const;
    V to_cons(const V& w) const;
    double sound_speed(const V& w) const;
    double max_speed(const V& w, int d) const;
    void eigen(const V& w, int d, Mat& l, Mat& r) const;
    V riemann_flux(const V& wl, const V& wr, int d, double* s_star) const;
    V split_flux(const V& a, const V& b, int d) const;
    V flux(const V& w, int d) const;
    Mat dudw(const V& w) const;
    V eigenvalues(const V& w, int d) const;
    bool admissible(const V& w) const;
};

// scheme.rs:39-53
enum class FluxMode { Hybrid = 0, Weno = 1, Central = 2 };
enum class ShockScheme { Wcns = 0, FdWeno = 1 };

// scheme.rs:101-117
struct Stats {
    uint64_t faces = 0, weno_faces = 0, fallback_faces = 0, limited_faces = 0;
    void add(const Stats& o) {
        faces += o.faces; weno_faces += o.weno_faces;
        fallback_faces += o.fallback_faces; limited_faces += o.limited_faces;
    }
};

struct SchemeOpts {
    FluxMode mode;
    ShockScheme shock;
    double tol;
    double lambda;  // D*dt/dx_d of the stage, 0 disables the limiter
};

// scheme.rs:57-86
double weno5z(double v0, double v1, double v2, double v3, double v4);
double weno5z_interp(double v0, double v1, double v2, double v3, double v4);

// scheme.rs:132-281: face fluxes of the n+1 faces of one line of primitive
// states w (length n + 2 ng).  `out` receives n+1 fluxes.
void line_fluxes(const Euler& sys, const V* w, size_t len, int d, size_t ng,
                 const SchemeOpts& o, V* out, Stats& st);

// solver.rs:8-20
enum class BcKind { Periodic = 0, Outflow = 1, Reflect = 2, Dirichlet = 3, Characteristic = 4 };
struct Bc {
    BcKind kind = BcKind::Periodic;
    V state{};  // conserved far/inflow state (Dirichlet, Characteristic)
};

// solver.rs:22-86
struct Mesh {
    size_t n[3];
    double lo[3];
    double dx[3];
    size_t ng = 6;
    Mesh() = default;
    Mesh(const size_t nn[3], const double l[3], const double h[3], size_t ghosts = 6);
    bool active(int d) const { return n[d] > 1; }
    size_t m(int d) const { return active(d) ? n[d] + 2 * ng : 1; }
    size_t g(int d) const { return active(d) ? ng : 0; }
    size_t idx(size_t i, size_t j, size_t k) const { return (i * m(1) + j) * m(2) + k; }
    size_t len() const { return m(0) * m(1) * m(2); }
    size_t cells() const { return n[0] * n[1] * n[2]; }
    double x(size_t c, int d) const { return lo[d] + ((double)c + 0.5) * dx[d]; }
    void periodic_ghost_sources(int d, size_t gi, size_t& lo_src, size_t& hi_src) const;
};

// solver.rs:139-795 (Euler, no viscous / all-speed / scalar hooks)
struct Solver {
    Euler sys;
    Mesh mesh;
    std::vector<V> u;      // conserved, ghost-padded
    Bc bc[3][2];
    double cfl = 0.5;
    FluxMode mode = FluxMode::Hybrid;
    ShockScheme shock = ShockScheme::Wcns;
    double tol = 1e-3;
    double t = 0.0;
    bool positivity = true;
    Stats stats;
    double stage_dt = 0.0;
    std::vector<V> rhs_buf, stage;

    Solver(const Euler& s, const Mesh& m);
    void fill_ghosts(std::vector<V>& v) const;
    // Residual -div F of `v` (ghosts of `v` refreshed) into rhs_buf.
    Stats rhs(std::vector<V>& v);
    double dt_cfl() const;
    void step(double dt);
    size_t run_to(double t_end);
    V totals() const;
};

// boundary.rs:48-96
struct CharacteristicStencil {
    V wi, wf, w1;
    Mat l, r;
    bool outgoing[NV];
    CharacteristicStencil(const Euler& sys, const V& u_b, const V& u_b1, const V& u_far, int d, int side);
    V conserved(const Euler& sys, double m) const;
};

// FNV-1a 64 over interior conserved values in (i, j, k) order
// (examples/cpu_validate.rs:7-22).
uint64_t state_hash(const Solver& s);

}  // namespace emref
