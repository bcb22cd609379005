// Oracle: CPU restatement of the reference's block-structured AMR with a
// global SSP-RK3 step (fluid/crates/compressible/src/amr.rs, subcycling and
// relaxation left out).  Test infrastructure only, like emref.h.
#pragma once
#include <cstdint>
#include <set>
#include <vector>

#include "emref.h"

namespace emref {

// amr.rs:50-63
struct AmrConfig {
    size_t b = 16;
    size_t levels = 3;
    double eps = 1e-3;
    size_t regrid_every = 4;
    size_t base_blocks[3] = {2, 2, 0};
    size_t ghosts = 5;
};

struct AmrKey {
    int level;
    int64_t idx[3];
    bool operator<(const AmrKey& o) const {
        if (level != o.level) return level < o.level;
        for (int d = 0; d < 3; ++d)
            if (idx[d] != o.idx[d]) return idx[d] < o.idx[d];
        return false;
    }
    bool operator==(const AmrKey& o) const {
        return level == o.level && idx[0] == o.idx[0] && idx[1] == o.idx[1] && idx[2] == o.idx[2];
    }
};

// amr.rs:67-139 (dense per-level block index)
struct BlockIndex {
    int64_t nb0[3] = {1, 1, 1};
    bool active[3] = {false, false, false};
    std::vector<std::vector<int32_t>> levels;
    void init(const int64_t n0[3], const bool act[3], size_t nl);
    int64_t nb(size_t l, int d) const { return active[d] ? (nb0[d] << l) : 1; }
    long slot(int l, const int64_t idx[3]) const;
    int get(int l, const int64_t idx[3]) const;  // -1 when absent
    void insert(int l, const int64_t idx[3], int id);
    void remove(int l, const int64_t idx[3]);
};

// amr.rs:193-204
struct AmrBlock {
    size_t level = 0;
    int64_t idx[3] = {0, 0, 0};
    bool leaf = true;
    std::vector<V> u, u0, rhs;
    std::vector<V> face[3][2];
};

using AmrInit = V (*)(const double x[3], const void* ctx);

struct Amr {
    Euler sys;
    AmrConfig cfg;
    double lo[3], hi[3];
    Bc bc[3][2];
    FluxMode mode = FluxMode::Hybrid;
    ShockScheme shock = ShockScheme::Wcns;
    double tol = 1e-3;
    double cfl = 0.4;
    bool positivity = true;
    double t = 0.0;
    size_t steps = 0;
    Stats stats;
    std::vector<AmrBlock> blocks;
    BlockIndex map;
    size_t ng;
    double wp[2][6];
    double wr[6];
    double stage_dt = 0.0;
    size_t regrid_interval, next_regrid;

    Amr(const Euler& s, const AmrConfig& c, const double lo_[3], const double hi_[3], AmrInit init, const void* ctx,
        const Bc bcs[3][2]);

    bool active(int d) const;
    size_t bd(int d) const { return cfg.base_blocks[d] == 0 ? 1 : cfg.b; }
    std::vector<int> dims() const;
    int64_t n_level(size_t l, int d) const;
    double dx(size_t l, int d) const;
    size_t bcells() const { return bd(0) * bd(1) * bd(2); }
    size_t cidx(const size_t c[3]) const { return (c[0] * bd(1) + c[1]) * bd(2) + c[2]; }
    void x(size_t l, const int64_t g[3], double out[3]) const;

    AmrBlock new_block(size_t level, const int64_t idx[3]) const;
    void add_block(size_t level, const int64_t idx[3], AmrInit init, const void* ctx);
    void block_key(const int64_t g[3], int64_t out[3]) const;
    V sample(size_t l, const int64_t g[3]) const;
    void bc_box(size_t l, const int64_t g0[3], const size_t sz[3], std::vector<V>& val, std::vector<char>& ok) const;
    std::vector<V> gather_box(size_t l, const int64_t g0[3], const size_t sz[3]) const;
    std::vector<V> predict_box(size_t l, const int64_t fmin[3], const int64_t fmax[3], size_t sz_out[3]) const;
    V predict(size_t l, const int64_t g[3]) const;
    std::vector<V> restrict_box(size_t l, const int64_t idx[3], const size_t lo_[3], const size_t hi_[3]) const;
    void restrict_all();
    double block_detail(size_t id) const;
    std::vector<int8_t> flags() const;
    std::vector<int8_t> flags_from_details(const std::vector<double>& det) const;
    std::vector<std::array<int64_t, 3>> children(const int64_t idx[3]) const;
    std::vector<std::array<int64_t, 3>> neighbours(size_t l, const int64_t idx[3]) const;
    void refine_block(size_t id, AmrInit init, const void* ctx);
    void coarsen(size_t l, const int64_t idx[3]);
    std::vector<size_t> leaves() const;
    void regrid();
    std::set<AmrKey> regrid_refine(const std::vector<int8_t>& flags);
    void regrid_coarsen(const std::vector<int8_t>& flags, const std::set<AmrKey>& prot);
    std::vector<V> fill_scratch(size_t id) const;
    void block_rhs(size_t id, std::vector<V>& rhs, std::vector<V> face[3][2], Stats& st) const;
    Stats compute_rhs();
    void reflux();
    double dt_cfl() const;
    void step(double dt);
    bool regrid_if_due();
    void schedule_regrid(bool changed);
    V totals() const;
    size_t cells() const;
};

}  // namespace emref
