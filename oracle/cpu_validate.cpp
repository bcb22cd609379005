// Restatement of the reference's CPU known-answer probe
// (fluid/crates/compressible/examples/cpu_validate.rs:23-149): `smooth`,
// `weno`, `fdweno`, `interface` (two-phase FiveEq, Hybrid) and `open` cases;
// prints l1_rho / mass_error (or max_dp, or max_relative_dp / max_du) and the
// FNV-1a state hash in the reference's output format.
// Test infrastructure: run by tests/test_oracle.py.
//   cpu_validate [smooth|weno|fdweno|interface|open] [N=32] [steps=40]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "emref.h"
#include "emref_multi.h"

using namespace emref;

// Rust's `{:.9e}` (LowerExp with precision): mantissa, 'e', exponent without
// sign padding, e.g. 3.622120917e-8.
static std::string rust_exp9(double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.9e", v);
    std::string s(buf);
    size_t e = s.find('e');
    std::string mant = s.substr(0, e);
    int ex = std::atoi(s.c_str() + e + 1);
    return mant + "e" + std::to_string(ex);
}

int main(int argc, char** argv) {
    std::string cs = argc > 1 ? argv[1] : "smooth";
    size_t n = argc > 2 ? (size_t)std::atol(argv[2]) : 32;
    size_t steps = argc > 3 ? (size_t)std::atol(argv[3]) : 40;
    size_t nn[3] = {n, n, n};
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    Mesh mesh(nn, lo, hi);
    Euler sys;
    sys.eos = Nasg{1.4, 0.0, 0.0, 0.0};
    const double TAU = 6.283185307179586;
    if (cs == "smooth" || cs == "weno" || cs == "fdweno") {
        auto exact = [&](double x0, double x1, double t) {
            return 1.0 + 0.1 * std::sin(TAU * (x0 - 0.3 * t)) * std::cos(TAU * (x1 - 0.2 * t));
        };
        Solver s(sys, mesh);
        for (size_t i = 0; i < n; ++i)
            for (size_t j = 0; j < n; ++j)
                for (size_t k = 0; k < n; ++k) {
                    double x0 = mesh.x(i, 0), x1 = mesh.x(j, 1);
                    V w{exact(x0, x1, 0.0), 0.3, 0.2, 0.1, 1.0};
                    s.u[mesh.idx(i + mesh.g(0), j + mesh.g(1), k + mesh.g(2))] = sys.to_cons(w);
                }
        s.mode = cs == "smooth" ? FluxMode::Central : FluxMode::Weno;
        if (cs == "fdweno") s.shock = ShockScheme::FdWeno;
        double dt = 0.02 / (double)steps;
        auto t0 = std::chrono::steady_clock::now();
        for (size_t q = 0; q < steps; ++q) s.step(dt);
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        double err = 0.0, mass = 0.0;
        for (size_t i = 0; i < n; ++i)
            for (size_t j = 0; j < n; ++j)
                for (size_t k = 0; k < n; ++k) {
                    V w = sys.to_prim(s.u[mesh.idx(i + mesh.g(0), j + mesh.g(1), k + mesh.g(2))]);
                    if (!sys.admissible(w)) {
                        std::fprintf(stderr, "inadmissible state\n");
                        return 2;
                    }
                    err += std::fabs(w[0] - exact(mesh.x(i, 0), mesh.x(j, 1), s.t));
                    mass += w[0];
                }
        double nc = (double)(n * n * n);
        std::printf("%s n=%zu steps=%zu wall_s=%.9f l1_rho=%s mass_error=%s hash=%016llx\n", cs.c_str(), n, steps, el,
                    rust_exp9(err / nc).c_str(), rust_exp9(std::fabs(mass / nc - 1.0)).c_str(),
                    (unsigned long long)state_hash(s));
    } else if (cs == "open") {
        V far = sys.to_cons(V{1.0, 0.0, 0.0, 0.0, 1.0});
        Solver s(sys, mesh);
        for (size_t i = 0; i < n; ++i)
            for (size_t j = 0; j < n; ++j)
                for (size_t k = 0; k < n; ++k) {
                    double xs[3] = {mesh.x(i, 0), mesh.x(j, 1), mesh.x(k, 2)};
                    double r2 = -0.0;
                    for (int d = 0; d < 3; ++d) r2 += (xs[d] - 0.5) * (xs[d] - 0.5);
                    double dp = 1e-4 * std::exp(-r2 / 0.04);
                    s.u[mesh.idx(i + mesh.g(0), j + mesh.g(1), k + mesh.g(2))] =
                        sys.to_cons(V{1.0 + dp / 1.4, 0.0, 0.0, 0.0, 1.0 + dp});
                }
        for (int d = 0; d < 3; ++d)
            for (int sd = 0; sd < 2; ++sd) s.bc[d][sd] = Bc{BcKind::Characteristic, far};
        s.mode = FluxMode::Central;
        double dt = 0.02 / (double)steps;
        auto t0 = std::chrono::steady_clock::now();
        for (size_t q = 0; q < steps; ++q) s.step(dt);
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        double pressure = 0.0;
        for (size_t i = 0; i < n; ++i)
            for (size_t j = 0; j < n; ++j)
                for (size_t k = 0; k < n; ++k) {
                    V w = sys.to_prim(s.u[mesh.idx(i + mesh.g(0), j + mesh.g(1), k + mesh.g(2))]);
                    if (!sys.admissible(w)) return 2;
                    pressure = std::fmax(pressure, std::fabs(w[4] - 1.0));
                }
        std::printf("open n=%zu steps=%zu wall_s=%.9f max_dp=%s hash=%016llx\n", n, steps, el,
                    rust_exp9(pressure).c_str(), (unsigned long long)state_hash(s));
    } else if (cs == "interface") {
        // examples/cpu_validate.rs:100-146
        gen::FiveEq fe;
        fe.eos[0] = Nasg{2.0, 1e6, 0.0, 0.0};
        fe.eos[1] = Nasg{1.4, 0.0, 0.0, 0.0};
        fe.kapila = true;
        gen::SolverG<gen::FiveEq> s(fe, mesh);
        gen::interface_init(s);
        s.mode = FluxMode::Hybrid;
        const double dt = 1e-5;
        auto t0 = std::chrono::steady_clock::now();
        for (size_t q = 0; q < steps; ++q) s.step(dt);
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        double dp = 0.0, du = 0.0;
        for (size_t i = 0; i < n; ++i)
            for (size_t j = 0; j < n; ++j)
                for (size_t k = 0; k < n; ++k) {
                    gen::Vn<7> w = fe.to_prim(s.u[mesh.idx(i + mesh.g(0), j + mesh.g(1), k + mesh.g(2))]);
                    if (!fe.admissible(w)) return 2;
                    dp = std::fmax(dp, std::fabs(w[5] / 1e5 - 1.0));
                    du = std::fmax(du, std::fabs(w[2] - 0.1));
                }
        std::printf("interface n=%zu steps=%zu wall_s=%.9f max_relative_dp=%s max_du=%s hash=%016llx\n", n, steps, el,
                    rust_exp9(dp).c_str(), rust_exp9(du).c_str(), (unsigned long long)s.hash());
    } else {
        std::fprintf(stderr, "smooth, weno, fdweno, interface or open expected\n");
        return 1;
    }
    return 0;
}
