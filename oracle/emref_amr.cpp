// Oracle: restatement of fluid/crates/compressible/src/amr.rs (global step).
// Test infrastructure only.  Every function cites the reference lines it
// follows; loops keep the reference's iteration and accumulation order.
#include "emref_amr.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace emref {

static inline int64_t div_euclid(int64_t a, int64_t b) {
    int64_t q = a / b;
    if (a % b < 0) q = b > 0 ? q - 1 : q + 1;
    return q;
}
static inline int64_t rem_euclid(int64_t a, int64_t b) {
    int64_t r = a % b;
    return r < 0 ? r + (b < 0 ? -b : b) : r;
}

// amr.rs:41-47
static void lagrange6(double x0, double x, double out[6]) {
    for (int k = 0; k < 6; ++k) {
        double xk = x0 + (double)k;
        double p = 1.0;  // Iterator::product folds from 1.0
        for (int j = 0; j < 6; ++j) {
            if (j == k) continue;
            p *= (x - (x0 + (double)j)) / (xk - (x0 + (double)j));
        }
        out[k] = p;
    }
}

// ---------------------------------------------------------------- BlockIndex
void BlockIndex::init(const int64_t n0[3], const bool act[3], size_t nl) {
    for (int d = 0; d < 3; ++d) {
        nb0[d] = n0[d];
        active[d] = act[d];
    }
    levels.clear();
    for (size_t l = 0; l < nl; ++l) {
        size_t n = 1;
        for (int d = 0; d < 3; ++d) n *= (size_t)nb(l, d);
        levels.emplace_back(n, -1);
    }
}
long BlockIndex::slot(int l, const int64_t idx[3]) const {
    if (l < 0 || l >= (int)levels.size()) return -1;
    long s = 0;
    for (int d = 0; d < 3; ++d) {
        int64_t n = nb(l, d);
        if (idx[d] < 0 || idx[d] >= n) return -1;
        s = s * n + idx[d];
    }
    return s;
}
int BlockIndex::get(int l, const int64_t idx[3]) const {
    long s = slot(l, idx);
    return s < 0 ? -1 : levels[l][s];
}
void BlockIndex::insert(int l, const int64_t idx[3], int id) {
    long s = slot(l, idx);
    if (s >= 0) levels[l][s] = id;
}
void BlockIndex::remove(int l, const int64_t idx[3]) {
    long s = slot(l, idx);
    if (s >= 0) levels[l][s] = -1;
}

// ---------------------------------------------------------------- geometry
bool Amr::active(int d) const {  // amr.rs:250-252
    return cfg.base_blocks[d] > 0 && (hi[d] - lo[d]) > 0.0 && bd(d) > 1;
}
std::vector<int> Amr::dims() const {
    std::vector<int> v;
    for (int d = 0; d < 3; ++d)
        if (active(d)) v.push_back(d);
    return v;
}
int64_t Amr::n_level(size_t l, int d) const {  // amr.rs:265-271
    return active(d) ? (int64_t)(cfg.base_blocks[d] * cfg.b) * ((int64_t)1 << l) : 1;
}
double Amr::dx(size_t l, int d) const { return (hi[d] - lo[d]) / (double)n_level(l, d); }
void Amr::x(size_t l, const int64_t g[3], double out[3]) const {  // amr.rs:283-285
    for (int d = 0; d < 3; ++d)
        out[d] = active(d) ? lo[d] + ((double)g[d] + 0.5) * dx(l, d) : 0.5 * (lo[d] + hi[d]);
}

// amr.rs:289-343
Amr::Amr(const Euler& s, const AmrConfig& c, const double lo_[3], const double hi_[3], AmrInit init, const void* ctx,
         const Bc bcs[3][2])
    : sys(s), cfg(c) {
    for (int d = 0; d < 3; ++d) {
        lo[d] = lo_[d];
        hi[d] = hi_[d];
        for (int sd = 0; sd < 2; ++sd) bc[d][sd] = bcs[d][sd];
    }
    ng = std::max<size_t>(cfg.ghosts, 5);
    lagrange6(-3.0, -0.25, wp[0]);
    lagrange6(-2.0, 0.25, wp[1]);
    const double w[6] = {3.0 / 256.0, -25.0 / 256.0, 150.0 / 256.0, 150.0 / 256.0, -25.0 / 256.0, 3.0 / 256.0};
    for (int k = 0; k < 6; ++k) wr[k] = w[k];
    regrid_interval = cfg.regrid_every;
    next_regrid = cfg.regrid_every;
    int64_t nb[3];
    bool act[3];
    for (int d = 0; d < 3; ++d) {
        nb[d] = (int64_t)std::max<size_t>(cfg.base_blocks[d], 1);
        act[d] = active(d);
    }
    map.init(nb, act, cfg.levels);
    for (int64_t i = 0; i < nb[0]; ++i)
        for (int64_t j = 0; j < nb[1]; ++j)
            for (int64_t k = 0; k < nb[2]; ++k) {
                int64_t idx[3] = {i, j, k};
                add_block(0, idx, init, ctx);
            }
    for (size_t it = 0; it < cfg.levels; ++it) {
        restrict_all();
        std::vector<int8_t> fl = flags();
        std::vector<AmrKey> keys;
        for (size_t i = 0; i < fl.size(); ++i)
            if (fl[i] > 0 && blocks[i].leaf) {
                AmrKey k{(int)blocks[i].level, {blocks[i].idx[0], blocks[i].idx[1], blocks[i].idx[2]}};
                keys.push_back(k);
            }
        if (keys.empty()) break;
        for (const AmrKey& k : keys) {
            int id = map.get(k.level, k.idx);
            if (id >= 0) refine_block((size_t)id, init, ctx);
        }
    }
}

AmrBlock Amr::new_block(size_t level, const int64_t idx[3]) const {  // amr.rs:345-359
    AmrBlock b;
    b.level = level;
    for (int d = 0; d < 3; ++d) b.idx[d] = idx[d];
    b.leaf = true;
    b.u.assign(bcells(), V{0, 0, 0, 0, 0});
    b.rhs.assign(bcells(), V{0, 0, 0, 0, 0});
    for (int d = 0; d < 3; ++d) {
        size_t f = 1;
        for (int e = 0; e < 3; ++e)
            if (e != d) f *= bd(e);
        b.face[d][0].assign(f, V{0, 0, 0, 0, 0});
        b.face[d][1].assign(f, V{0, 0, 0, 0, 0});
    }
    return b;
}

void Amr::add_block(size_t level, const int64_t idx[3], AmrInit init, const void* ctx) {  // amr.rs:361-374
    AmrBlock b = new_block(level, idx);
    size_t b0 = bd(0), b1 = bd(1), b2 = bd(2);
    for (size_t i = 0; i < b0; ++i)
        for (size_t j = 0; j < b1; ++j)
            for (size_t k = 0; k < b2; ++k) {
                int64_t g[3] = {idx[0] * (int64_t)b0 + (int64_t)i, idx[1] * (int64_t)b1 + (int64_t)j,
                                idx[2] * (int64_t)b2 + (int64_t)k};
                double xx[3];
                x(level, g, xx);
                b.u[(i * b1 + j) * b2 + k] = sys.to_cons(init(xx, ctx));
            }
    map.insert((int)level, idx, (int)blocks.size());
    blocks.push_back(std::move(b));
}

void Amr::block_key(const int64_t g[3], int64_t out[3]) const {  // amr.rs:376-379
    for (int d = 0; d < 3; ++d) out[d] = active(d) ? div_euclid(g[d], (int64_t)bd(d)) : 0;
}

// amr.rs:383-439
V Amr::sample(size_t l, const int64_t g_in[3]) const {
    int64_t g[3] = {g_in[0], g_in[1], g_in[2]};
    bool flip[3] = {false, false, false};
    for (int d = 2; d >= 0; --d) {
        if (!active(d)) {
            g[d] = 0;
            continue;
        }
        int64_t n = n_level(l, d);
        if (g[d] < 0 || g[d] >= n) {
            int side = g[d] >= n ? 1 : 0;
            switch (bc[d][side].kind) {
                case BcKind::Periodic: g[d] = rem_euclid(g[d], n); break;
                case BcKind::Outflow: g[d] = std::clamp<int64_t>(g[d], 0, n - 1); break;
                case BcKind::Reflect:
                    g[d] = side == 0 ? -1 - g[d] : 2 * n - 1 - g[d];
                    flip[d] = true;
                    break;
                case BcKind::Dirichlet: return bc[d][side].state;
                case BcKind::Characteristic: {
                    int64_t gi[3] = {g[0], g[1], g[2]};
                    gi[d] = std::clamp<int64_t>(g[d], 0, n - 1);
                    int64_t g1[3] = {gi[0], gi[1], gi[2]};
                    g1[d] += side == 1 ? -1 : 1;
                    double m = (double)std::llabs(g[d] - gi[d]);
                    V ub = sample(l, gi), ub1 = sample(l, g1);
                    V v = CharacteristicStencil(sys, ub, ub1, bc[d][side].state, d, side).conserved(sys, m);
                    for (int e = 0; e < 3; ++e)
                        if (flip[e]) v[1 + e] *= -1.0;
                    return v;
                }
            }
        }
    }
    int64_t key[3];
    block_key(g, key);
    int id = map.get((int)l, key);
    V v;
    if (id >= 0) {
        size_t c[3];
        for (int d = 0; d < 3; ++d) c[d] = active(d) ? (size_t)rem_euclid(g[d], (int64_t)bd(d)) : 0;
        v = blocks[id].u[cidx(c)];
    } else {
        v = predict(l, g);
    }
    for (int d = 0; d < 3; ++d)
        if (flip[d]) v[1 + d] *= -1.0;
    return v;
}

// amr.rs:450-504
void Amr::bc_box(size_t l, const int64_t g0[3], const size_t sz[3], std::vector<V>& val,
                 std::vector<char>& ok) const {
    std::vector<int> dm = dims();
    int64_t n[3];
    for (int d = 0; d < 3; ++d) n[d] = n_level(l, d);
    auto at = [&](const int64_t c[3]) -> long {
        for (int d = 0; d < 3; ++d)
            if (c[d] < 0 || c[d] >= (int64_t)sz[d]) return -1;
        return ((long)c[0] * (long)sz[1] + c[1]) * (long)sz[2] + c[2];
    };
    auto outside = [&](const int64_t g[3], int d) {
        return (g[d] < 0 || g[d] >= n[d]) && bc[d][g[d] >= n[d] ? 1 : 0].kind != BcKind::Periodic;
    };
    for (size_t di = 0; di < dm.size(); ++di) {
        int d = dm[di];
        for (size_t i = 0; i < sz[0]; ++i)
            for (size_t j = 0; j < sz[1]; ++j)
                for (size_t k = 0; k < sz[2]; ++k) {
                    int64_t c[3] = {(int64_t)i, (int64_t)j, (int64_t)k};
                    int64_t g[3] = {g0[0] + c[0], g0[1] + c[1], g0[2] + c[2]};
                    if (!outside(g, d)) continue;
                    bool later = false;
                    for (size_t ej = di + 1; ej < dm.size(); ++ej)
                        if (outside(g, dm[ej])) later = true;
                    if (later) continue;
                    long q = at(c);
                    int side = g[d] >= n[d] ? 1 : 0;
                    auto local = [&](int64_t gg) -> long {
                        int64_t s[3] = {c[0], c[1], c[2]};
                        s[d] = gg - g0[d];
                        long p = at(s);
                        return (p >= 0 && ok[p]) ? p : -1;
                    };
                    int64_t gi = std::clamp<int64_t>(g[d], 0, n[d] - 1);
                    bool have = false;
                    V v{};
                    switch (bc[d][side].kind) {
                        case BcKind::Periodic: break;
                        case BcKind::Dirichlet:
                            v = bc[d][side].state;
                            have = true;
                            break;
                        case BcKind::Outflow: {
                            long p = local(gi);
                            if (p >= 0) {
                                v = val[p];
                                have = true;
                            }
                            break;
                        }
                        case BcKind::Reflect: {
                            long p = local(side == 0 ? -1 - g[d] : 2 * n[d] - 1 - g[d]);
                            if (p >= 0) {
                                v = val[p];
                                v[1 + d] *= -1.0;
                                have = true;
                            }
                            break;
                        }
                        case BcKind::Characteristic: {
                            int64_t g1 = gi + (side == 1 ? -1 : 1);
                            long pb = local(gi), p1 = local(g1);
                            if (pb >= 0 && p1 >= 0) {
                                v = CharacteristicStencil(sys, val[pb], val[p1], bc[d][side].state, d, side)
                                        .conserved(sys, (double)std::llabs(g[d] - gi));
                                have = true;
                            }
                            break;
                        }
                    }
                    if (have) {
                        val[q] = v;
                        ok[q] = 1;
                    }
                }
    }
}

// amr.rs:508-543
std::vector<V> Amr::gather_box(size_t l, const int64_t g0[3], const size_t sz[3]) const {
    int64_t n[3];
    for (int d = 0; d < 3; ++d) n[d] = n_level(l, d);
    std::vector<int> dm = dims();
    size_t len = sz[0] * sz[1] * sz[2];
    std::vector<V> val(len, V{0, 0, 0, 0, 0});
    std::vector<char> ok(len, 1);
    bool any = false;
    for (size_t i = 0; i < sz[0]; ++i)
        for (size_t j = 0; j < sz[1]; ++j)
            for (size_t k = 0; k < sz[2]; ++k) {
                int64_t g[3] = {g0[0] + (int64_t)i, g0[1] + (int64_t)j, g0[2] + (int64_t)k};
                size_t q = (i * sz[1] + j) * sz[2] + k;
                bool out = false;
                for (int d : dm)
                    if ((g[d] < 0 || g[d] >= n[d]) && bc[d][g[d] >= n[d] ? 1 : 0].kind != BcKind::Periodic) out = true;
                if (out) {
                    ok[q] = 0;
                    any = true;
                } else {
                    val[q] = sample(l, g);
                }
            }
    if (any) {
        bc_box(l, g0, sz, val, ok);
        for (size_t i = 0; i < sz[0]; ++i)
            for (size_t j = 0; j < sz[1]; ++j)
                for (size_t k = 0; k < sz[2]; ++k) {
                    size_t q = (i * sz[1] + j) * sz[2] + k;
                    if (!ok[q]) {
                        int64_t g[3] = {g0[0] + (int64_t)i, g0[1] + (int64_t)j, g0[2] + (int64_t)k};
                        val[q] = sample(l, g);
                    }
                }
    }
    return val;
}

// amr.rs:548-586
std::vector<V> Amr::predict_box(size_t l, const int64_t fmin[3], const int64_t fmax[3], size_t sz_out[3]) const {
    std::vector<int> dm = dims();
    bool act[3];
    for (int d = 0; d < 3; ++d) act[d] = active(d);
    int64_t cmin[3], cmax[3];
    size_t sz[3];
    for (int d = 0; d < 3; ++d) {
        cmin[d] = act[d] ? div_euclid(fmin[d], 2) - 3 : 0;
        cmax[d] = act[d] ? div_euclid(fmax[d], 2) + 3 : 0;
        sz[d] = (size_t)(cmax[d] - cmin[d] + 1);
    }
    std::vector<V> arr = gather_box(l - 1, cmin, sz);
    for (int d : dm) {
        size_t nf = (size_t)(fmax[d] - fmin[d] + 1);
        size_t nsz[3] = {sz[0], sz[1], sz[2]};
        nsz[d] = nf;
        std::vector<V> na(nsz[0] * nsz[1] * nsz[2]);
        for (size_t i = 0; i < nsz[0]; ++i)
            for (size_t j = 0; j < nsz[1]; ++j)
                for (size_t k = 0; k < nsz[2]; ++k) {
                    size_t c[3] = {i, j, k};
                    int64_t f = fmin[d] + (int64_t)c[d];
                    int par = (int)rem_euclid(f, 2);
                    int64_t base = div_euclid(f, 2) + (par == 1 ? -2 : -3) - cmin[d];
                    V acc{0, 0, 0, 0, 0};
                    for (int kk = 0; kk < 6; ++kk) {
                        size_t src[3] = {c[0], c[1], c[2]};
                        src[d] = (size_t)(base + kk);
                        const V& v = arr[(src[0] * sz[1] + src[1]) * sz[2] + src[2]];
                        double wgt = wp[par][kk];
                        for (int q = 0; q < NV; ++q) acc[q] += wgt * v[q];
                    }
                    na[(i * nsz[1] + j) * nsz[2] + k] = acc;
                }
        arr.swap(na);
        for (int e = 0; e < 3; ++e) sz[e] = nsz[e];
    }
    for (int e = 0; e < 3; ++e) sz_out[e] = sz[e];
    return arr;
}

// amr.rs:589-612
V Amr::predict(size_t l, const int64_t g[3]) const {
    std::vector<int> dm = dims();
    V acc{0, 0, 0, 0, 0};
    size_t nd = dm.size();
    size_t total = 1;
    for (size_t i = 0; i < nd; ++i) total *= 6;
    for (size_t s = 0; s < total; ++s) {
        double w = 1.0;
        int64_t c[3] = {0, 0, 0};
        size_t rem = s;
        for (int d : dm) {
            size_t k = rem % 6;
            rem /= 6;
            int odd = (int)rem_euclid(g[d], 2);
            int64_t base = div_euclid(g[d], 2) + (odd == 1 ? -2 : -3);
            c[d] = base + (int64_t)k;
            w *= wp[odd][k];
        }
        V v = sample(l - 1, c);
        for (int q = 0; q < NV; ++q) acc[q] += w * v[q];
    }
    return acc;
}

// amr.rs:622-730
std::vector<V> Amr::restrict_box(size_t l, const int64_t idx[3], const size_t lo_[3], const size_t hi_[3]) const {
    std::vector<int> dm = dims();
    size_t bd0[3] = {bd(0), bd(1), bd(2)};
    size_t bdx[3];
    bool act[3];
    int64_t c0[3], f0[3], nf[3];
    size_t sz[3];
    for (int d = 0; d < 3; ++d) {
        bdx[d] = hi_[d] - lo_[d];
        act[d] = active(d);
        c0[d] = idx[d] * (int64_t)bd0[d] + (int64_t)lo_[d];
        f0[d] = act[d] ? 2 * c0[d] - 4 : 0;
        sz[d] = act[d] ? 2 * bdx[d] + 8 : 1;
        nf[d] = n_level(l + 1, d);
    }
    size_t len = sz[0] * sz[1] * sz[2];
    std::vector<V> val;
    std::vector<char> ok;
    val.reserve(len);
    ok.reserve(len);
    bool any_out = false;
    for (size_t i = 0; i < sz[0]; ++i)
        for (size_t j = 0; j < sz[1]; ++j)
            for (size_t k = 0; k < sz[2]; ++k) {
                int64_t g[3] = {f0[0] + (int64_t)i, f0[1] + (int64_t)j, f0[2] + (int64_t)k};
                int64_t gw[3] = {g[0], g[1], g[2]};
                bool outside = false;
                for (int d = 0; d < 3; ++d)
                    if (act[d] && (g[d] < 0 || g[d] >= nf[d])) {
                        if (bc[d][g[d] >= nf[d] ? 1 : 0].kind == BcKind::Periodic)
                            gw[d] = rem_euclid(g[d], nf[d]);
                        else
                            outside = true;
                    }
                if (outside) {
                    val.push_back(V{0, 0, 0, 0, 0});
                    ok.push_back(0);
                    any_out = true;
                    continue;
                }
                int64_t key[3];
                block_key(gw, key);
                int bid = map.get((int)l + 1, key);
                if (bid >= 0) {
                    size_t c[3];
                    for (int d = 0; d < 3; ++d) c[d] = act[d] ? (size_t)rem_euclid(gw[d], (int64_t)bd0[d]) : 0;
                    val.push_back(blocks[bid].u[cidx(c)]);
                    ok.push_back(1);
                } else {
                    val.push_back(V{0, 0, 0, 0, 0});
                    ok.push_back(0);
                }
            }
    if (any_out) bc_box(l + 1, f0, sz, val, ok);
    const int64_t shifts[5] = {0, -1, 1, -2, 2};
    double wts[5][6];
    for (int k = 0; k < 6; ++k) wts[0][k] = wr[k];
    for (int m = 1; m < 5; ++m) lagrange6(-2.5 + (double)shifts[m], 0.0, wts[m]);
    for (int d : dm) {
        size_t nsz[3] = {sz[0], sz[1], sz[2]};
        nsz[d] = bdx[d];
        std::vector<V> nv(nsz[0] * nsz[1] * nsz[2], V{0, 0, 0, 0, 0});
        std::vector<char> nok(nsz[0] * nsz[1] * nsz[2], 0);
        for (size_t i = 0; i < nsz[0]; ++i)
            for (size_t j = 0; j < nsz[1]; ++j)
                for (size_t k = 0; k < nsz[2]; ++k) {
                    size_t c[3] = {i, j, k};
                    auto at_src = [&](size_t pos) {
                        size_t s2[3] = {c[0], c[1], c[2]};
                        s2[d] = pos;
                        return (s2[0] * sz[1] + s2[1]) * sz[2] + s2[2];
                    };
                    int64_t base = 2 * (int64_t)c[d] + 2;
                    bool done = false;
                    size_t o = (i * nsz[1] + j) * nsz[2] + k;
                    for (int m = 0; m < 5; ++m) {
                        int64_t st = base + shifts[m];
                        if (st < 0 || st + 5 >= (int64_t)sz[d]) continue;
                        bool all = true;
                        for (int kk = 0; kk < 6; ++kk) all = all && ok[at_src((size_t)(st + kk))];
                        if (all) {
                            V acc{0, 0, 0, 0, 0};
                            for (int kk = 0; kk < 6; ++kk) {
                                const V& v = val[at_src((size_t)(st + kk))];
                                for (int q = 0; q < NV; ++q) acc[q] += wts[m][kk] * v[q];
                            }
                            nv[o] = acc;
                            nok[o] = 1;
                            done = true;
                            break;
                        }
                    }
                    if (!done) {
                        size_t pa = at_src((size_t)(base + 2)), pb = at_src((size_t)(base + 3));
                        if (ok[pa] && ok[pb]) {
                            V acc;
                            for (int q = 0; q < NV; ++q) acc[q] = 0.5 * (val[pa][q] + val[pb][q]);
                            nv[o] = acc;
                            nok[o] = 1;
                        }
                    }
                }
        val.swap(nv);
        ok.swap(nok);
        for (int e = 0; e < 3; ++e) sz[e] = nsz[e];
    }
    return val;
}

// amr.rs:737-816
void Amr::restrict_all() {
    if (cfg.levels < 2) return;
    std::vector<int> dm = dims();
    auto isact = [&](int d) { return std::find(dm.begin(), dm.end(), d) != dm.end(); };
    size_t bdv[3] = {bd(0), bd(1), bd(2)};
    for (long ll = (long)cfg.levels - 2; ll >= 0; --ll) {
        size_t l = (size_t)ll;
        std::vector<size_t> ids;
        for (size_t i = 0; i < blocks.size(); ++i) {
            const AmrBlock& b = blocks[i];
            if (b.leaf || b.level != l) continue;
            bool edge = false;
            for (auto& nk : neighbours(l, b.idx)) {
                int j = map.get((int)l, nk.data());
                if (j < 0 || blocks[j].leaf) edge = true;
            }
            if (edge) ids.push_back(i);
        }
        struct Box {
            size_t lo[3], hi[3];
        };
        std::vector<std::vector<Box>> boxes(ids.size());
        for (size_t q = 0; q < ids.size(); ++q) {
            const AmrBlock& b = blocks[ids[q]];
            int64_t nbk[3];
            for (int d = 0; d < 3; ++d) nbk[d] = isact(d) ? (int64_t)cfg.base_blocks[d] * ((int64_t)1 << l) : 1;
            std::vector<Box> bx;
            auto offs = [&](int d) { return isact(d) ? std::vector<int64_t>{-1, 0, 1} : std::vector<int64_t>{0}; };
            for (int64_t o0 : offs(0))
                for (int64_t o1 : offs(1))
                    for (int64_t o2 : offs(2)) {
                        int64_t o[3] = {o0, o1, o2};
                        if (o0 == 0 && o1 == 0 && o2 == 0) continue;
                        int64_t nk[3];
                        bool skip = false;
                        for (int d = 0; d < 3; ++d) {
                            nk[d] = b.idx[d] + o[d];
                            if (nk[d] < 0 || nk[d] >= nbk[d]) {
                                if (bc[d][nk[d] >= nbk[d] ? 1 : 0].kind == BcKind::Periodic)
                                    nk[d] = rem_euclid(nk[d], nbk[d]);
                                else
                                    skip = true;
                            }
                        }
                        if (skip) continue;
                        int j = map.get((int)l, nk);
                        if (j >= 0 && !blocks[j].leaf) continue;
                        size_t w = std::min<size_t>(ng + 3, std::max(bdv[0], std::max(bdv[1], bdv[2])));
                        Box x;
                        for (int d = 0; d < 3; ++d) {
                            x.lo[d] = o[d] == 1 ? (bdv[d] >= w ? bdv[d] - w : 0) : 0;
                            x.hi[d] = o[d] == -1 ? std::min(w, bdv[d]) : bdv[d];
                        }
                        bx.push_back(x);
                    }
            size_t vol = 0;
            for (auto& x : bx) vol += (x.hi[0] - x.lo[0]) * (x.hi[1] - x.lo[1]) * (x.hi[2] - x.lo[2]);
            if (2 * vol > bdv[0] * bdv[1] * bdv[2]) {
                Box all;
                for (int d = 0; d < 3; ++d) {
                    all.lo[d] = 0;
                    all.hi[d] = bdv[d];
                }
                bx.assign(1, all);
            }
            boxes[q] = bx;
        }
        std::vector<std::vector<std::vector<V>>> vals(ids.size());
#pragma omp parallel for schedule(dynamic)
        for (long q = 0; q < (long)ids.size(); ++q) {
            const AmrBlock& b = blocks[ids[q]];
            for (auto& x : boxes[q]) vals[q].push_back(restrict_box(l, b.idx, x.lo, x.hi));
        }
        for (size_t q = 0; q < ids.size(); ++q) {
            AmrBlock& b = blocks[ids[q]];
            for (size_t bi = 0; bi < boxes[q].size(); ++bi) {
                const Box& x = boxes[q][bi];
                size_t it = 0;
                for (size_t i = x.lo[0]; i < x.hi[0]; ++i)
                    for (size_t j = x.lo[1]; j < x.hi[1]; ++j)
                        for (size_t k = x.lo[2]; k < x.hi[2]; ++k) {
                            size_t c[3] = {i, j, k};
                            b.u[cidx(c)] = vals[q][bi][it++];
                        }
            }
        }
    }
}

// amr.rs:823-867
double Amr::block_detail(size_t id) const {
    std::vector<int> dm = dims();
    size_t bdv[3] = {bd(0), bd(1), bd(2)};
    size_t ext[3], gof[3];
    for (int d = 0; d < 3; ++d) {
        bool a = active(d);
        ext[d] = a ? bdv[d] + 2 * ng : 1;
        gof[d] = a ? ng : 0;
    }
    std::vector<V> scr = fill_scratch(id);
    std::vector<std::array<double, 2>> sens(scr.size());
    for (size_t i = 0; i < scr.size(); ++i) {
        V w = sys.to_prim(scr[i]);
        sens[i][0] = std::log(std::fmax(w[0], 1e-300));
        sens[i][1] = std::log(std::fmax(w[4] + sys.eos.p_inf, 1e-300));
    }
    auto at = [&](const size_t c[3]) { return (c[0] * ext[1] + c[1]) * ext[2] + c[2]; };
    double mx = 0.0;
    const int64_t offs[6] = {-5, -3, -1, 1, 3, 5};
    for (size_t i = 0; i < bdv[0]; ++i)
        for (size_t j = 0; j < bdv[1]; ++j)
            for (size_t k = 0; k < bdv[2]; ++k) {
                size_t c[3] = {i + gof[0], j + gof[1], k + gof[2]};
                const auto& s0 = sens[at(c)];
                for (int d : dm) {
                    double pred[2] = {0.0, 0.0};
                    for (int m = 0; m < 6; ++m) {
                        size_t cc[3] = {c[0], c[1], c[2]};
                        cc[d] = (size_t)((int64_t)cc[d] + offs[m]);
                        const auto& sv = sens[at(cc)];
                        for (int q = 0; q < 2; ++q) pred[q] += wr[m] * sv[q];
                    }
                    for (int q = 0; q < 2; ++q) {
                        double dd = std::fabs(s0[q] - pred[q]);
                        mx = std::isfinite(dd) ? std::fmax(mx, dd) : INFINITY;
                    }
                }
            }
    return mx;
}

std::vector<int8_t> Amr::flags() const {  // amr.rs:870-874
    std::vector<double> det(blocks.size(), 0.0);
#pragma omp parallel for schedule(dynamic)
    for (long i = 0; i < (long)blocks.size(); ++i) det[i] = blocks[i].leaf ? block_detail((size_t)i) : 0.0;
    return flags_from_details(det);
}

std::vector<int8_t> Amr::flags_from_details(const std::vector<double>& det) const {  // amr.rs:877-892
    std::vector<int8_t> f(blocks.size(), 0);
    for (size_t i = 0; i < blocks.size(); ++i) {
        const AmrBlock& b = blocks[i];
        if (!b.leaf)
            f[i] = 0;
        else if (det[i] > cfg.eps && b.level + 1 < cfg.levels)
            f[i] = 1;
        else if (det[i] < cfg.eps / 128.0)
            f[i] = -1;
        else
            f[i] = 0;
    }
    return f;
}

std::vector<std::array<int64_t, 3>> Amr::children(const int64_t idx[3]) const {  // amr.rs:894-916
    std::vector<std::array<int64_t, 3>> out{{0, 0, 0}};
    for (int d = 0; d < 3; ++d) {
        std::vector<std::array<int64_t, 3>> next;
        for (auto c : out) {
            if (active(d)) {
                for (int64_t o = 0; o < 2; ++o) {
                    auto cc = c;
                    cc[d] = 2 * idx[d] + o;
                    next.push_back(cc);
                }
            } else {
                auto cc = c;
                cc[d] = 0;
                next.push_back(cc);
            }
        }
        out.swap(next);
    }
    return out;
}

std::vector<std::array<int64_t, 3>> Amr::neighbours(size_t l, const int64_t idx[3]) const {  // amr.rs:919-949
    int64_t nb[3];
    for (int d = 0; d < 3; ++d) nb[d] = active(d) ? (int64_t)cfg.base_blocks[d] * ((int64_t)1 << l) : 1;
    std::vector<std::array<int64_t, 3>> out;
    auto r = [&](int d) { return active(d) ? std::pair<int, int>{-1, 1} : std::pair<int, int>{0, 0}; };
    for (int a = r(0).first; a <= r(0).second; ++a)
        for (int b = r(1).first; b <= r(1).second; ++b)
            for (int c = r(2).first; c <= r(2).second; ++c) {
                if (a == 0 && b == 0 && c == 0) continue;
                std::array<int64_t, 3> k = {idx[0] + a, idx[1] + b, idx[2] + c};
                bool inside = true;
                for (int d = 0; d < 3; ++d)
                    if (k[d] < 0 || k[d] >= nb[d]) {
                        if (bc[d][k[d] >= nb[d] ? 1 : 0].kind == BcKind::Periodic)
                            k[d] = rem_euclid(k[d], nb[d]);
                        else
                            inside = false;
                    }
                if (inside) out.push_back(k);
            }
    return out;
}

// amr.rs:954-1011
void Amr::refine_block(size_t id0, AmrInit init, const void* ctx) {
    size_t l = blocks[id0].level;
    int64_t idx[3] = {blocks[id0].idx[0], blocks[id0].idx[1], blocks[id0].idx[2]};
    if (!blocks[id0].leaf) return;
    if (l > 0) {
        for (auto& nk : neighbours(l, idx)) {
            if (map.get((int)l, nk.data()) < 0) {
                for (long m = (long)l - 1; m >= 0; --m) {
                    int64_t pk[3];
                    for (int d = 0; d < 3; ++d) pk[d] = active(d) ? (nk[d] >> (l - (size_t)m)) : 0;
                    int pid = map.get((int)m, pk);
                    if (pid >= 0) {
                        refine_block((size_t)pid, init, ctx);
                        break;
                    }
                }
            }
        }
    }
    int idi = map.get((int)l, idx);
    if (idi < 0) return;
    size_t id = (size_t)idi;
    if (!blocks[id].leaf) return;
    auto kids = children(idx);
    size_t b0 = bd(0), b1 = bd(1), b2 = bd(2);
    size_t bdv[3] = {b0, b1, b2};
    std::vector<AmrBlock> newb(kids.size());
#pragma omp parallel for schedule(dynamic)
    for (long q = 0; q < (long)kids.size(); ++q) {
        const auto& c = kids[q];
        AmrBlock b = new_block(l + 1, c.data());
        if (init) {
            for (size_t i = 0; i < b0; ++i)
                for (size_t j = 0; j < b1; ++j)
                    for (size_t k = 0; k < b2; ++k) {
                        int64_t g[3] = {c[0] * (int64_t)b0 + (int64_t)i, c[1] * (int64_t)b1 + (int64_t)j,
                                        c[2] * (int64_t)b2 + (int64_t)k};
                        double xx[3];
                        x(l + 1, g, xx);
                        b.u[(i * b1 + j) * b2 + k] = sys.to_cons(init(xx, ctx));
                    }
        } else {
            int64_t fmin[3], fmax[3];
            for (int d = 0; d < 3; ++d) {
                fmin[d] = c[d] * (int64_t)bdv[d];
                fmax[d] = c[d] * (int64_t)bdv[d] + (int64_t)bdv[d] - 1;
            }
            size_t szo[3];
            b.u = predict_box(l + 1, fmin, fmax, szo);
        }
        newb[q] = std::move(b);
    }
    blocks[id].leaf = false;
    for (auto& b : newb) {
        map.insert((int)b.level, b.idx, (int)blocks.size());
        blocks.push_back(std::move(b));
    }
}

// amr.rs:1014-1032
void Amr::coarsen(size_t l, const int64_t idx[3]) {
    auto kids = children(idx);
    int pid = map.get((int)l, idx);
    size_t lo0[3] = {0, 0, 0}, hi0[3] = {bd(0), bd(1), bd(2)};
    std::vector<V> u = restrict_box(l, idx, lo0, hi0);
    blocks[pid].u = u;
    blocks[pid].leaf = true;
    std::vector<size_t> ids;
    for (auto& c : kids) ids.push_back((size_t)map.get((int)l + 1, c.data()));
    std::sort(ids.begin(), ids.end(), [](size_t a, size_t b) { return a > b; });
    for (size_t i : ids) {
        AmrBlock& b = blocks[i];
        map.remove((int)b.level, b.idx);
        if (i != blocks.size() - 1) std::swap(blocks[i], blocks.back());
        blocks.pop_back();
        if (i < blocks.size()) {
            AmrBlock& m = blocks[i];
            map.insert((int)m.level, m.idx, (int)i);
        }
    }
}

std::vector<size_t> Amr::leaves() const {
    std::vector<size_t> v;
    for (size_t i = 0; i < blocks.size(); ++i)
        if (blocks[i].leaf) v.push_back(i);
    return v;
}

// amr.rs:1039-1046
void Amr::regrid() {
    restrict_all();
    std::vector<int8_t> f = flags();
    std::set<AmrKey> prot = regrid_refine(f);
    restrict_all();
    f = flags();
    regrid_coarsen(f, prot);
}

// amr.rs:1051-1077
std::set<AmrKey> Amr::regrid_refine(const std::vector<int8_t>& fl) {
    std::vector<AmrKey> want;
    for (size_t i = 0; i < fl.size(); ++i) {
        if (fl[i] > 0) {
            const AmrBlock& b = blocks[i];
            want.push_back(AmrKey{(int)b.level, {b.idx[0], b.idx[1], b.idx[2]}});
            for (auto& nk : neighbours(b.level, b.idx)) {
                int j = map.get((int)b.level, nk.data());
                if (j >= 0 && blocks[j].leaf) want.push_back(AmrKey{(int)b.level, {nk[0], nk[1], nk[2]}});
            }
        }
    }
    std::sort(want.begin(), want.end());
    want.erase(std::unique(want.begin(), want.end()), want.end());
    std::set<AmrKey> prot(want.begin(), want.end());
    for (const AmrKey& k : want) {
        int id = map.get(k.level, k.idx);
        if (id >= 0 && (size_t)k.level + 1 < cfg.levels) refine_block((size_t)id, nullptr, nullptr);
    }
    return prot;
}

// amr.rs:1081-1106
void Amr::regrid_coarsen(const std::vector<int8_t>& fl, const std::set<AmrKey>& prot) {
    size_t nkids = (size_t)1 << dims().size();
    std::vector<AmrKey> cand;
    for (size_t bi = 0; bi < blocks.size(); ++bi) {
        const AmrBlock& b = blocks[bi];
        if (b.leaf) continue;
        auto kids = children(b.idx);
        bool ok = kids.size() == nkids;
        for (auto& c : kids) {
            if (!ok) break;
            int j = map.get((int)b.level + 1, c.data());
            AmrKey ck{(int)b.level + 1, {c[0], c[1], c[2]}};
            ok = j >= 0 && blocks[j].leaf && fl[j] < 0 && prot.find(ck) == prot.end();
        }
        if (ok) cand.push_back(AmrKey{(int)b.level, {b.idx[0], b.idx[1], b.idx[2]}});
    }
    std::stable_sort(cand.begin(), cand.end(), [](const AmrKey& a, const AmrKey& b) { return b.level < a.level; });
    for (const AmrKey& k : cand) {
        size_t l = (size_t)k.level;
        bool ok = true;
        for (auto& nk : neighbours(l, k.idx)) {
            for (auto& c : children(nk.data())) {
                int j = map.get((int)l + 1, c.data());
                if (j >= 0 && !blocks[j].leaf) ok = false;
            }
        }
        if (ok && map.get((int)l, k.idx) >= 0) coarsen(l, k.idx);
    }
}

// amr.rs:1113-1221
std::vector<V> Amr::fill_scratch(size_t id) const {
    const AmrBlock& b = blocks[id];
    size_t l = b.level;
    size_t bdv[3] = {bd(0), bd(1), bd(2)};
    bool act[3];
    size_t ext[3], gof[3];
    for (int d = 0; d < 3; ++d) {
        act[d] = active(d);
        ext[d] = act[d] ? bdv[d] + 2 * ng : 1;
        gof[d] = act[d] ? ng : 0;
    }
    auto at = [&](const size_t c[3]) { return (c[0] * ext[1] + c[1]) * ext[2] + c[2]; };
    std::vector<V> out(ext[0] * ext[1] * ext[2], V{0, 0, 0, 0, 0});
    std::vector<char> filled(out.size(), 1);
    for (size_t i = 0; i < bdv[0]; ++i)
        for (size_t j = 0; j < bdv[1]; ++j)
            for (size_t k = 0; k < bdv[2]; ++k) {
                size_t c[3] = {i + gof[0], j + gof[1], k + gof[2]};
                size_t lc[3] = {i, j, k};
                out[at(c)] = b.u[cidx(lc)];
            }
    int64_t nbk[3];
    for (int d = 0; d < 3; ++d) nbk[d] = act[d] ? (int64_t)cfg.base_blocks[d] * ((int64_t)1 << l) : 1;
    auto rng = [&](int64_t o, int d, size_t& s, size_t& e) {
        if (!act[d]) {
            s = 0;
            e = 1;
            return;
        }
        if (o == -1) {
            s = 0;
            e = ng;
        } else if (o == 0) {
            s = ng;
            e = ng + bdv[d];
        } else {
            s = ng + bdv[d];
            e = 2 * ng + bdv[d];
        }
    };
    auto offs = [&](int d) { return act[d] ? std::vector<int64_t>{-1, 0, 1} : std::vector<int64_t>{0}; };
    auto gl = [&](const size_t c[3], int64_t out3[3]) {
        for (int d = 0; d < 3; ++d) out3[d] = b.idx[d] * (int64_t)bdv[d] + (int64_t)c[d] - (int64_t)gof[d];
    };
    for (int64_t o0 : offs(0))
        for (int64_t o1 : offs(1))
            for (int64_t o2 : offs(2)) {
                int64_t o[3] = {o0, o1, o2};
                if (o0 == 0 && o1 == 0 && o2 == 0) continue;
                size_t rs[3], re[3];
                for (int d = 0; d < 3; ++d) rng(o[d], d, rs[d], re[d]);
                int64_t nk[3];
                bool inside = true;
                for (int d = 0; d < 3; ++d) {
                    nk[d] = b.idx[d] + o[d];
                    if (nk[d] < 0 || nk[d] >= nbk[d]) {
                        if (bc[d][nk[d] >= nbk[d] ? 1 : 0].kind == BcKind::Periodic)
                            nk[d] = rem_euclid(nk[d], nbk[d]);
                        else
                            inside = false;
                    }
                }
                if (!inside) {
                    for (size_t i = rs[0]; i < re[0]; ++i)
                        for (size_t j = rs[1]; j < re[1]; ++j)
                            for (size_t k = rs[2]; k < re[2]; ++k) {
                                size_t c[3] = {i, j, k};
                                filled[at(c)] = 0;
                            }
                    continue;
                }
                int nid = map.get((int)l, nk);
                if (nid >= 0) {
                    const AmrBlock& nb = blocks[nid];
                    for (size_t i = rs[0]; i < re[0]; ++i)
                        for (size_t j = rs[1]; j < re[1]; ++j)
                            for (size_t k = rs[2]; k < re[2]; ++k) {
                                size_t c[3] = {i, j, k};
                                size_t lc[3];
                                for (int d = 0; d < 3; ++d)
                                    lc[d] = act[d] ? (size_t)((int64_t)c[d] - (int64_t)gof[d] - o[d] * (int64_t)bdv[d]) : 0;
                                out[at(c)] = nb.u[cidx(lc)];
                            }
                    continue;
                }
                size_t cs[3] = {rs[0], rs[1], rs[2]}, ce[3] = {re[0] - 1, re[1] - 1, re[2] - 1};
                int64_t fmin[3], fmax[3];
                gl(cs, fmin);
                gl(ce, fmax);
                size_t sz[3];
                std::vector<V> arr = predict_box(l, fmin, fmax, sz);
                for (size_t i = rs[0]; i < re[0]; ++i)
                    for (size_t j = rs[1]; j < re[1]; ++j)
                        for (size_t k = rs[2]; k < re[2]; ++k) {
                            size_t c[3] = {i, j, k};
                            int64_t g[3];
                            gl(c, g);
                            size_t li[3];
                            for (int d = 0; d < 3; ++d) li[d] = act[d] ? (size_t)(g[d] - fmin[d]) : 0;
                            out[at(c)] = arr[(li[0] * sz[1] + li[1]) * sz[2] + li[2]];
                        }
            }
    bool anyf = false;
    for (char f : filled) anyf = anyf || !f;
    if (anyf) {
        int64_t g0[3];
        for (int d = 0; d < 3; ++d) g0[d] = b.idx[d] * (int64_t)bdv[d] - (int64_t)gof[d];
        bc_box(l, g0, ext, out, filled);
        for (size_t i = 0; i < ext[0]; ++i)
            for (size_t j = 0; j < ext[1]; ++j)
                for (size_t k = 0; k < ext[2]; ++k) {
                    size_t c[3] = {i, j, k};
                    if (!filled[at(c)]) {
                        int64_t g[3] = {g0[0] + (int64_t)i, g0[1] + (int64_t)j, g0[2] + (int64_t)k};
                        out[at(c)] = sample(l, g);
                    }
                }
    }
    return out;
}

// amr.rs:1224-1299 (Euler: no non-conservative terms)
void Amr::block_rhs(size_t id, std::vector<V>& rhs, std::vector<V> face[3][2], Stats& st) const {
    std::vector<int> dm = dims();
    size_t bdv[3] = {bd(0), bd(1), bd(2)};
    size_t ext[3], gof[3];
    for (int d = 0; d < 3; ++d) {
        bool a = active(d);
        ext[d] = a ? bdv[d] + 2 * ng : 1;
        gof[d] = a ? ng : 0;
    }
    double ndim = (double)dm.size();
    const AmrBlock& b = blocks[id];
    size_t l = b.level;
    std::vector<V> scr = fill_scratch(id);
    std::vector<V> w(scr.size());
    for (size_t i = 0; i < scr.size(); ++i) w[i] = sys.to_prim(scr[i]);
    rhs.assign(bcells(), V{0, 0, 0, 0, 0});
    for (int d = 0; d < 3; ++d) {
        face[d][0].clear();
        face[d][1].clear();
    }
    for (int d : dm) {
        int others[2], no = 0;
        for (int e = 0; e < 3; ++e)
            if (e != d) others[no++] = e;
        int e0 = others[0], e1 = others[1];
        double dxl = dx(l, d);
        SchemeOpts opts{mode, shock, tol, positivity ? ndim * stage_dt / dxl : 0.0};
        std::vector<V> fl(bdv[d] + 1), line(ext[d]);
        std::vector<V> f_lo, f_hi;
        for (size_t a = 0; a < bdv[e0]; ++a)
            for (size_t bb = 0; bb < bdv[e1]; ++bb) {
                auto at = [&](size_t cc) {
                    size_t c[3];
                    c[d] = cc;
                    c[e0] = a + gof[e0];
                    c[e1] = bb + gof[e1];
                    return (c[0] * ext[1] + c[1]) * ext[2] + c[2];
                };
                for (size_t cc = 0; cc < ext[d]; ++cc) line[cc] = w[at(cc)];
                line_fluxes(sys, line.data(), ext[d], d, ng, opts, fl.data(), st);
                for (size_t cc = 0; cc < bdv[d]; ++cc) {
                    size_t c[3];
                    c[d] = cc;
                    c[e0] = a;
                    c[e1] = bb;
                    V& o = rhs[cidx(c)];
                    for (int q = 0; q < NV; ++q) o[q] -= (fl[cc + 1][q] - fl[cc][q]) / dxl;
                }
                f_lo.push_back(fl[0]);
                f_hi.push_back(fl[bdv[d]]);
            }
        face[d][0] = f_lo;
        face[d][1] = f_hi;
    }
}

// amr.rs:1302-1323
Stats Amr::compute_rhs() {
    restrict_all();
    std::vector<size_t> lv = leaves();
    std::vector<Stats> sts(lv.size());
#pragma omp parallel for schedule(dynamic)
    for (long q = 0; q < (long)lv.size(); ++q) {
        AmrBlock& b = blocks[lv[q]];
        std::vector<V> rhs;
        std::vector<V> face[3][2];
        block_rhs(lv[q], rhs, face, sts[q]);
        b.rhs.swap(rhs);
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s) b.face[d][s].swap(face[d][s]);
    }
    Stats total;
    for (auto& s : sts) total.add(s);
    reflux();
    return total;
}

// amr.rs:1327-1406
void Amr::reflux() {
    std::vector<int> dm = dims();
    auto isact = [&](int d) { return std::find(dm.begin(), dm.end(), d) != dm.end(); };
    size_t bdv[3] = {bd(0), bd(1), bd(2)};
    struct Corr {
        size_t id, cell;
        V dr;
    };
    std::vector<Corr> corr;
    for (size_t id = 0; id < blocks.size(); ++id) {
        const AmrBlock& b = blocks[id];
        size_t l = b.level;
        if (!b.leaf || l + 1 >= cfg.levels) continue;
        for (int d : dm) {
            int others[2], no = 0;
            for (int e = 0; e < 3; ++e)
                if (e != d) others[no++] = e;
            int e0 = others[0], e1 = others[1];
            double dxl = dx(l, d);
            int64_t nbk = (int64_t)cfg.base_blocks[d] * ((int64_t)1 << l);
            for (int side = 0; side < 2; ++side) {
                int64_t nk[3] = {b.idx[0], b.idx[1], b.idx[2]};
                nk[d] += side == 1 ? 1 : -1;
                if (nk[d] < 0 || nk[d] >= nbk) {
                    if (bc[d][side].kind != BcKind::Periodic) continue;
                    nk[d] = rem_euclid(nk[d], nbk);
                }
                int nj = map.get((int)l, nk);
                if (!(nj >= 0 && !blocks[nj].leaf)) continue;
                for (size_t a = 0; a < bdv[e0]; ++a)
                    for (size_t bb = 0; bb < bdv[e1]; ++bb) {
                        V sum{0, 0, 0, 0, 0};
                        double cnt = 0.0;
                        int fa_n = isact(e0) ? 2 : 1, fb_n = isact(e1) ? 2 : 1;
                        for (int fa = 0; fa < fa_n; ++fa)
                            for (int fb = 0; fb < fb_n; ++fb) {
                                int64_t ga = 2 * (b.idx[e0] * (int64_t)bdv[e0] + (int64_t)a) + fa;
                                int64_t gb = 2 * (b.idx[e1] * (int64_t)bdv[e1] + (int64_t)bb) + fb;
                                int64_t ck[3] = {nk[0], nk[1], nk[2]};
                                ck[d] = 2 * nk[d] + (side == 1 ? 0 : 1);
                                ck[e0] = isact(e0) ? div_euclid(ga, (int64_t)bdv[e0]) : 0;
                                ck[e1] = isact(e1) ? div_euclid(gb, (int64_t)bdv[e1]) : 0;
                                int cid = map.get((int)l + 1, ck);
                                if (cid < 0 || !blocks[cid].leaf) continue;
                                size_t la = isact(e0) ? (size_t)rem_euclid(ga, (int64_t)bdv[e0]) : 0;
                                size_t lb = isact(e1) ? (size_t)rem_euclid(gb, (int64_t)bdv[e1]) : 0;
                                const V& fv = blocks[cid].face[d][1 - side][la * bdv[e1] + lb];
                                for (int q = 0; q < NV; ++q) sum[q] += fv[q];
                                cnt += 1.0;
                            }
                        if (cnt == 0.0) continue;
                        const V& fc = b.face[d][side][a * bdv[e1] + bb];
                        size_t c[3];
                        c[d] = side == 1 ? bdv[d] - 1 : 0;
                        c[e0] = a;
                        c[e1] = bb;
                        V dr;
                        for (int q = 0; q < NV; ++q) {
                            double df = sum[q] / cnt - fc[q];
                            dr[q] = side == 1 ? -df / dxl : df / dxl;
                        }
                        corr.push_back(Corr{id, cidx(c), dr});
                    }
            }
        }
    }
    for (auto& c : corr)
        for (int q = 0; q < NV; ++q) blocks[c.id].rhs[c.cell][q] += c.dr[q];
}

// amr.rs:1409-1430 (Euler: no dilatation source)
double Amr::dt_cfl() const {
    std::vector<int> dm = dims();
    double rate = 0.0;
    for (const AmrBlock& b : blocks) {
        if (!b.leaf) continue;
        double r = 0.0;
        for (const V& u : b.u) {
            V w = sys.to_prim(u);
            double s = -0.0;
            for (int d : dm) s += sys.max_speed(w, d) / dx(b.level, d);
            r = std::fmax(r, s);
        }
        rate = std::fmax(rate, r);
    }
    return cfl / rate;
}

// amr.rs:1470-1512
void Amr::step(double dt) {
    stage_dt = dt;
    for (auto& b : blocks)
        if (b.leaf) b.u0 = b.u;
    Stats s1 = compute_rhs();
    for (auto& b : blocks)
        if (b.leaf)
            for (size_t c = 0; c < b.u.size(); ++c)
                for (int q = 0; q < NV; ++q) b.u[c][q] += dt * b.rhs[c][q];
    Stats s2 = compute_rhs();
    for (auto& b : blocks)
        if (b.leaf)
            for (size_t c = 0; c < b.u.size(); ++c)
                for (int q = 0; q < NV; ++q) b.u[c][q] = 0.75 * b.u0[c][q] + 0.25 * (b.u[c][q] + dt * b.rhs[c][q]);
    Stats s3 = compute_rhs();
    for (auto& b : blocks)
        if (b.leaf)
            for (size_t c = 0; c < b.u.size(); ++c)
                for (int q = 0; q < NV; ++q) b.u[c][q] = b.u0[c][q] / 3.0 + 2.0 / 3.0 * (b.u[c][q] + dt * b.rhs[c][q]);
    stats.add(s1);
    stats.add(s2);
    stats.add(s3);
    t += dt;
    steps += 1;
    regrid_if_due();
}

// amr.rs:1525-1548
bool Amr::regrid_if_due() {
    if (steps < next_regrid) return false;
    std::vector<std::tuple<size_t, int64_t, int64_t, int64_t, bool>> before, after;
    for (auto& b : blocks) before.emplace_back(b.level, b.idx[0], b.idx[1], b.idx[2], b.leaf);
    regrid();
    for (auto& b : blocks) after.emplace_back(b.level, b.idx[0], b.idx[1], b.idx[2], b.leaf);
    bool changed = before != after;
    schedule_regrid(changed);
    return true;
}

void Amr::schedule_regrid(bool changed) {
    regrid_interval = changed ? cfg.regrid_every : std::min(2 * regrid_interval, 2 * cfg.regrid_every);
    next_regrid = steps + regrid_interval;
}

V Amr::totals() const {  // amr.rs:1992-2004
    std::vector<int> dm = dims();
    V s{0, 0, 0, 0, 0};
    for (const AmrBlock& b : blocks) {
        if (!b.leaf) continue;
        double vol = 1.0;
        for (int d : dm) vol *= dx(b.level, d);
        for (const V& u : b.u)
            for (int q = 0; q < NV; ++q) s[q] += u[q] * vol;
    }
    return s;
}

size_t Amr::cells() const {
    size_t n = 0;
    for (auto& b : blocks)
        if (b.leaf) n += bcells();
    return n;
}

}  // namespace emref
