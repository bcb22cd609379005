// Device kernels of the block-AMR pool other than the flux pass:
// restriction of ancestors, ghost slabs (copy / boundary / prediction),
// prediction of refined children, refluxing, SSP-RK3 update, CFL rate,
// wavelet details and slot moves.  Compiled without FMA contraction so the
// operators reproduce the reference's arithmetic (amr.rs).
#pragma once
#include <hip/hip_runtime.h>

#include "em_amr_types.h"

namespace em {

// Weights of the separable 6th-order operators (amr.rs:41-47, 305-308,
// 675-676): prediction by fine-cell parity, centred and shifted restriction.
template <class Real>
struct AmrWeightsT {
    Real wp[2][6];
    Real wr[6];
    Real wts[5][6];  // restriction stencils for shifts 0, -1, 1, -2, 2
};

// Index arithmetic is 32-bit throughout: global cell indices of one level
// stay below 2^31 for any hierarchy that fits in device memory.
__device__ __forceinline__ int rem_e(int a, int b) {
    const int r = a % b;
    return r < 0 ? r + (b < 0 ? -b : b) : r;
}

template <class Real>
__device__ __forceinline__ int amr_n_level(const AmrDevT<Real>& a, int l, int d) {
    return a.act[d] ? (a.nb0[d] * a.bd[d]) << l : 1;
}

// Block id holding global cell g (inside the domain) at level l, -1 if
// absent; *cell receives the cell index inside the block.
template <class Real>
__device__ __forceinline__ int amr_find(const AmrDevT<Real>& a, int l, const int g[3], int* cell) {
    int s = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int nb = a.act[d] ? a.nb0[d] << l : 1;
        const int k = g[d] >> a.lb[d];
        if (k < 0 || k >= nb) return -1;
        s = s * nb + k;
    }
    *cell = ((g[0] & (a.bd[0] - 1)) * a.bd[1] + (g[1] & (a.bd[1] - 1))) * a.bd[2] + (g[2] & (a.bd[2] - 1));
    return a.map[a.map_off[l] + s];
}

__device__ __forceinline__ void unflat(int i, const int sz[3], int c[3]) {
    const int q = i / sz[2];
    c[2] = i - q * sz[2];
    c[0] = q / sz[1];
    c[1] = q - c[0] * sz[1];
}

// Block ids under a box of level-l cells g0 .. g0+sz-1 (per axis, before any
// periodic wrap), resolved once per workgroup so that a gather reads U with
// one dependent load instead of two (index, then data).  ids[] gets the id,
// -1 for an absent block, -2 outside a non-periodic boundary.  Returns false
// (and resolves nothing) when the box spans more than kBoxBlocks blocks.
// Every thread of the workgroup calls it; it ends with a barrier.
constexpr int kBoxBlocks = 512;
struct BoxBlocks {
    int b0[3], nbk[3];
};
template <class Real>
__device__ __forceinline__ bool box_blocks(const AmrDevT<Real>& a, int l, const int g0[3], const int sz[3], int* ids, BoxBlocks& bb) {
    int total = 1;
    for (int d = 0; d < 3; ++d) {
        if (a.act[d]) {
            bb.b0[d] = g0[d] >> a.lb[d];  // floor division (arithmetic shift)
            bb.nbk[d] = ((g0[d] + sz[d] - 1) >> a.lb[d]) - bb.b0[d] + 1;
        } else {
            bb.b0[d] = 0;
            bb.nbk[d] = 1;
        }
        total *= bb.nbk[d];
    }
    if (total > kBoxBlocks) {
        __syncthreads();
        return false;
    }
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        int c[3];
        unflat(i, bb.nbk, c);
        int s = 0, id = 0;
        for (int d = 0; d < 3; ++d) {
            const int nb = a.act[d] ? a.nb0[d] << l : 1;
            int k = bb.b0[d] + c[d];
            if (k < 0 || k >= nb) {
                if (a.act[d] && a.bc_kind[d][k >= nb ? 1 : 0] == BC_PERIODIC)
                    k = rem_e(k, nb);
                else
                    id = -2;
            }
            s = s * nb + k;
        }
        ids[i] = id == -2 ? -2 : a.map[a.map_off[l] + s];
    }
    __syncthreads();
    return true;
}
// entry of ids[] for the (unwrapped) cell g, and its cell index in the block
template <class Real>
__device__ __forceinline__ int box_block_of(const AmrDevT<Real>& a, const BoxBlocks& bb, const int g[3], int* cell) {
    int bi = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) bi = bi * bb.nbk[d] + ((a.act[d] ? g[d] >> a.lb[d] : 0) - bb.b0[d]);
    *cell = ((g[0] & (a.bd[0] - 1)) * a.bd[1] + (g[1] & (a.bd[1] - 1))) * a.bd[2] + (g[2] & (a.bd[2] - 1));
    return bi;
}

// bc_box of amr.rs:450-504 on a box held in scratch (values val[q*cap+i],
// validity ok[i]): cells outside a non-periodic boundary are filled from
// cells of the same box, one axis sweep at a time; called by one workgroup.
template <class Real>
__device__ void box_bc(const AmrDevT<Real>& a, int l, const int g0[3], const int sz[3], Real* val, Real* ok,
                       long long cap) {
    const int n3[3] = {amr_n_level(a, l, 0), amr_n_level(a, l, 1), amr_n_level(a, l, 2)};
    const int tot = sz[0] * sz[1] * sz[2];
    auto outside = [&](const int g[3], int d) {
        return a.act[d] && (g[d] < 0 || g[d] >= n3[d]) && a.bc_kind[d][g[d] >= n3[d] ? 1 : 0] != BC_PERIODIC;
    };
    for (int d = 0; d < 3; ++d) {
        if (!a.act[d]) continue;
        for (int i = threadIdx.x; i < tot; i += blockDim.x) {
            int c[3];
            unflat(i, sz, c);
            const int g[3] = {g0[0] + c[0], g0[1] + c[1], g0[2] + c[2]};
            if (!outside(g, d)) continue;
            bool later = false;
            for (int e = d + 1; e < 3; ++e) later = later || outside(g, e);
            if (later) continue;
            const int side = g[d] >= n3[d] ? 1 : 0;
            auto local = [&](int gg) -> int {
                const int cc = gg - g0[d];
                if (cc < 0 || cc >= sz[d]) return -1;
                int s3[3] = {c[0], c[1], c[2]};
                s3[d] = cc;
                const int p = (s3[0] * sz[1] + s3[1]) * sz[2] + s3[2];
                return ok[p] != Real(0.0) ? p : -1;
            };
            const int gi = g[d] < 0 ? 0 : (g[d] > n3[d] - 1 ? n3[d] - 1 : g[d]);
            const int kind = a.bc_kind[d][side];
            Real v[5];
            bool have = false;
            if (kind == BC_DIRICHLET) {
                for (int q = 0; q < 5; ++q) v[q] = a.bc_state[d][side][q];
                have = true;
            } else if (kind == BC_OUTFLOW) {
                const int p = local(gi);
                if (p >= 0) {
                    for (int q = 0; q < 5; ++q) v[q] = val[q * cap + p];
                    have = true;
                }
            } else if (kind == BC_REFLECT) {
                const int p = local(side == 0 ? -1 - g[d] : 2 * n3[d] - 1 - g[d]);
                if (p >= 0) {
                    for (int q = 0; q < 5; ++q) v[q] = val[q * cap + p];
                    v[1 + d] *= -Real(1.0);
                    have = true;
                }
            } else if (kind == BC_CHARACTERISTIC) {
                const int pb = local(gi), p1 = local(gi + (side == 1 ? -1 : 1));
                if (pb >= 0 && p1 >= 0) {
                    Real ub[5], ub1[5];
                    for (int q = 0; q < 5; ++q) {
                        ub[q] = val[q * cap + pb];
                        ub1[q] = val[q * cap + p1];
                    }
                    const Real m = Real(g[d] > gi ? g[d] - gi : gi - g[d]);
                    if (d == 0)
                        characteristic_ghost<0>(a.eos, ub, ub1, a.bc_state[d][side], side, m, v);
                    else if (d == 1)
                        characteristic_ghost<1>(a.eos, ub, ub1, a.bc_state[d][side], side, m, v);
                    else
                        characteristic_ghost<2>(a.eos, ub, ub1, a.bc_state[d][side], side, m, v);
                    have = true;
                }
            }
            if (have) {
                for (int q = 0; q < 5; ++q) val[q * cap + i] = v[q];
                ok[i] = Real(1.0);
            }
        }
        __syncthreads();
    }
}

// predict_box of amr.rs:548-586 by one workgroup: the level-(lf-1) patch is
// gathered into A / ok (gather_box: blocks of that level, periodic wrap,
// bc_box for cells outside the domain), then one 6-point pass per active
// axis maps coarse to fine indices, writing Bf, then A, then Bf.  The last
// pass hands every cell of the result (box fmin..fmax of level lf) to
// fin(i, c, v) instead: its row-major index and coordinates in the box and
// its five values.  A and ok hold capA cells, Bf capB; the buffers may be
// global scratch or LDS.  Ends with a barrier.
template <class Real, class Fin>
__device__ __forceinline__ void predict_box(const AmrDevT<Real>& a, const AmrWeightsT<Real>& W, int lf, const int fmin[3],
                                            const int fmax[3], Real* A, Real* ok, long long capA, Real* Bf,
                                            long long capB, Fin&& fin) {
    __shared__ int any_out;
    const int lc = lf - 1;
    int cmin[3], sz[3];
    for (int d = 0; d < 3; ++d) {
        cmin[d] = a.act[d] ? (fmin[d] >> 1) - 3 : 0;  // floor division by 2
        const int cmax = a.act[d] ? (fmax[d] >> 1) + 3 : 0;
        sz[d] = cmax - cmin[d] + 1;
    }
    const int n3[3] = {amr_n_level(a, lc, 0), amr_n_level(a, lc, 1), amr_n_level(a, lc, 2)};
    if (threadIdx.x == 0) any_out = 0;
    __shared__ int bids[kBoxBlocks];
    BoxBlocks bb;
    const bool cached = box_blocks(a, lc, cmin, sz, bids, bb);  // synchronises
    const int tot = sz[0] * sz[1] * sz[2];
    // four cells per thread and round, their loads issued together (a cell
    // with no block reads U[0] and keeps zero)
    constexpr int kBatch = 4;
    for (int i0 = threadIdx.x; i0 < tot && cached; i0 += kBatch * blockDim.x) {
        long long src[kBatch];
        int id[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            const int i = i0 + k * blockDim.x;
            id[k] = -3;  // beyond the box
            src[k] = 0;
            if (i >= tot) continue;
            int c[3];
            unflat(i, sz, c);
            int g[3] = {cmin[0] + c[0], cmin[1] + c[1], cmin[2] + c[2]};
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (!a.act[d]) g[d] = 0;
            int cell;
            id[k] = bids[box_block_of(a, bb, g, &cell)];
            if (id[k] >= 0) src[k] = (long long)id[k] * a.bcells + cell;
        }
        Real v[kBatch][5];
#pragma unroll
        for (int k = 0; k < kBatch; ++k)
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const Real x = a.U[q * a.vs + src[k]];
                v[k][q] = id[k] >= 0 ? x : Real(0.0);
            }
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            const int i = i0 + k * blockDim.x;
            if (id[k] == -3) continue;
            if (id[k] == -2) {
                any_out = 1;
            } else if (id[k] < 0) {
                atomicOr(&a.sc->err, 2);  // the hierarchy is 2:1 balanced (see below)
            }
#pragma unroll
            for (int q = 0; q < 5; ++q) A[q * capA + i] = v[k][q];
            ok[i] = id[k] == -2 ? Real(0.0) : Real(1.0);
        }
    }
    for (int i = threadIdx.x; i < tot && !cached; i += blockDim.x) {
        int c[3];
        unflat(i, sz, c);
        int g[3] = {cmin[0] + c[0], cmin[1] + c[1], cmin[2] + c[2]};
        bool out = false;
        for (int d = 0; d < 3; ++d) {
            if (!a.act[d]) {
                g[d] = 0;
                continue;
            }
            if (g[d] < 0 || g[d] >= n3[d]) {
                if (a.bc_kind[d][g[d] >= n3[d] ? 1 : 0] == BC_PERIODIC)
                    g[d] = rem_e(g[d], n3[d]);
                else
                    out = true;
            }
        }
        if (out) {
            ok[i] = Real(0.0);
            any_out = 1;
            for (int q = 0; q < 5; ++q) A[q * capA + i] = Real(0.0);
            continue;
        }
        int cell;
        const int id = amr_find(a, lc, g, &cell);
        if (id < 0) {
            // the hierarchy is 2:1 balanced, so the level below a leaf's
            // neighbourhood exists; report otherwise
            atomicOr(&a.sc->err, 2);
            for (int q = 0; q < 5; ++q) A[q * capA + i] = Real(0.0);
        } else {
            for (int q = 0; q < 5; ++q) A[q * capA + i] = a.U[amr_u(a, q, id, cell)];
        }
        ok[i] = Real(1.0);
    }
    __syncthreads();
    if (any_out) {
        box_bc(a, lc, cmin, sz, A, ok, capA);
        for (int i = threadIdx.x; i < tot; i += blockDim.x)
            if (ok[i] == Real(0.0)) atomicOr(&a.sc->err, 4);
        __syncthreads();
    }
    Real* in = A;
    Real* out = Bf;
    long long icap = capA, ocap = capB;
    const int dlast = a.act[2] ? 2 : a.act[1] ? 1 : 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {  // unrolled: arrays indexed by d stay in registers
        if (!a.act[d]) continue;
        int nsz[3] = {sz[0], sz[1], sz[2]};
        nsz[d] = fmax[d] - fmin[d] + 1;
        const int ntot = nsz[0] * nsz[1] * nsz[2];
        for (int i = threadIdx.x; i < ntot; i += blockDim.x) {
            int c[3];
            unflat(i, nsz, c);
            const int f = fmin[d] + c[d];
            const int par = f & 1;
            const int base = (f >> 1) + (par == 1 ? -2 : -3) - cmin[d];
            Real acc[5] = {Real(0.0), Real(0.0), Real(0.0), Real(0.0), Real(0.0)};
            int s3[3] = {c[0], c[1], c[2]};
            for (int kk = 0; kk < 6; ++kk) {
                s3[d] = base + kk;
                const int src = (s3[0] * sz[1] + s3[1]) * sz[2] + s3[2];
                const Real wgt = W.wp[par][kk];
#pragma unroll
                for (int q = 0; q < 5; ++q) acc[q] += wgt * in[q * icap + src];
            }
            if (d == dlast) {
                fin(i, c, acc);
            } else {
#pragma unroll
                for (int q = 0; q < 5; ++q) out[q * ocap + i] = acc[q];
            }
        }
        __syncthreads();
        Real* tmp = in;
        in = out;
        out = tmp;
        const long long tc = icap;
        icap = ocap;
        ocap = tc;
        for (int e = 0; e < 3; ++e) sz[e] = nsz[e];
    }
}

// predict_box with the buffers in global scratch (12 arrays of cap doubles);
// the result is left in scratch, in the buffer returned.
template <class Real>
__device__ Real* predict_region(const AmrDevT<Real>& a, const AmrWeightsT<Real>& W, int lf, const int fmin[3], const int fmax[3],
                                  Real* scr, long long cap) {
    const int nact = a.act[0] + a.act[1] + a.act[2];
    Real* res = (nact & 1) ? scr + 5 * cap : scr;  // the last pass's output buffer
    predict_box(a, W, lf, fmin, fmax, scr, scr + 10 * cap, cap, scr + 5 * cap, cap,
                [&](int i, const int*, const Real* v) {
#pragma unroll
                    for (int q = 0; q < 5; ++q) res[q * cap + i] = v[q];
                });
    return res;
}

// A box of level-l cells g0 .. g0+sz-1 read in place: at box cell c, the
// value and validity that gathering the box into scratch (blocks of level l,
// periodic wrap) and then filling it with box_bc would leave there, for boxes
// that reach past no characteristic boundary.  A cell with no coordinate
// outside a non-periodic boundary is its block's data (invalid, zero, where
// the block is absent).  A cell outside is box_bc's image along the last axis
// it lies outside of, of a source cell that box_bc has filled along the
// earlier axes first: so the source chain is walked from the last outside
// axis down.  Outflow moves to the boundary cell, reflect to the mirror cell
// and negates that axis's momentum, a Dirichlet state ends the chain; each
// axis negates a different component, so the sign changes commute.  Along an
// axis the step depends on that coordinate alone, so it is tabulated per axis
// (map, built by patch_map).
constexpr int kPatchLen = 72;  // 2b + 8 for the largest block side, 32
constexpr int kMapInvalid = 1 << 12, kMapFlip = 1 << 13, kMapDirichlet = 1 << 14;  // low bits: position or side
struct FinePatch {
    int f0[3], sz[3], n3[3];
    int l;
    bool cached;  // bids/bb hold the box's block ids (box_blocks)
    const int* bids;
    const int* map;  // [3][kPatchLen]
    const int* xs;   // blocks treated as absent (nx of them)
    int nx;
    BoxBlocks bb;
};

// map[d][p] for every box position p along every axis; every thread of the
// workgroup calls it, a barrier must follow before the map is read
template <class Real>
__device__ __forceinline__ void patch_map(const AmrDevT<Real>& a, const FinePatch& P, int* map) {
    for (int i = threadIdx.x; i < 3 * kPatchLen; i += blockDim.x) {
        const int d = i / kPatchLen, p = i - d * kPatchLen;
        // per-axis fields by selection, not by a runtime index (which would
        // place the arrays in scratch)
        auto pick = [d](const int* v) { return d == 0 ? v[0] : d == 1 ? v[1] : v[2]; };
        const int n3 = pick(P.n3), f0 = pick(P.f0), psz = pick(P.sz), act = pick(a.act);
        const int g = f0 + p;
        int e = p;
        if (act && p < psz && (g < 0 || g >= n3)) {
            const int side = g >= n3 ? 1 : 0;
            const int k2[3] = {a.bc_kind[0][side], a.bc_kind[1][side], a.bc_kind[2][side]};
            const int kind = pick(k2);
            // box_bc's source position: the boundary cell (outflow) or the
            // mirror cell (reflect), which lies inside the domain because a
            // patch spans -4 .. n3 + 3 and n3 >= 16
            int src = 0;
            bool moved = false;
            if (kind == BC_DIRICHLET) {
                e = kMapDirichlet | side;
            } else if (kind == BC_OUTFLOW) {
                src = side == 0 ? 0 : n3 - 1;
                moved = true;
            } else if (kind == BC_REFLECT) {
                src = side == 0 ? -1 - g : 2 * n3 - 1 - g;
                moved = true;
            } else if (kind != BC_PERIODIC) {
                e = kMapInvalid;  // characteristic boxes take the general path
            }
            if (moved) {
                const int ps = src - f0;
                e = ps >= 0 && ps < psz ? ps | (kind == BC_REFLECT ? kMapFlip : 0) : kMapInvalid;
            }
        }
        map[i] = e;
    }
}

// block id (-1: absent) and cell index of a box cell with no outside axis
template <class Real>
__device__ __forceinline__ int patch_block(const AmrDevT<Real>& a, const FinePatch& P, const int c[3], int* cell) {
    int g[3] = {P.f0[0] + c[0], P.f0[1] + c[1], P.f0[2] + c[2]};
    if (P.cached) return P.bids[box_block_of(a, P.bb, g, cell)];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (a.act[d]) g[d] = rem_e(g[d], P.n3[d]);  // out of range only across a periodic boundary
    int id = amr_find(a, P.l, g, cell);
    for (int k = 0; k < P.nx; ++k)
        if (id == P.xs[k]) id = -1;
    return id;
}

// value of a resolved box cell (zero where invalid)
template <class Real>
__device__ __forceinline__ void patch_load(const AmrDevT<Real>& a, int r, long long base, unsigned neg, Real v[5]) {
    if (r == -1) {
#pragma unroll
        for (int q = 0; q < 5; ++q) v[q] = a.U[q * a.vs + base];
    } else if (r >= 0) {
#pragma unroll
        for (int q = 0; q < 5; ++q) v[q] = a.bc_state[r >> 1][r & 1][q];
    } else {
#pragma unroll
        for (int q = 0; q < 5; ++q) v[q] = Real(0.0);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if ((neg >> d) & 1) v[1 + d] *= -Real(1.0);
}

// Box cells along axis D at fixed other coordinates c, resolved for a load
// (needs P.cached).  A box cell's source walks the axes from the last to
// the first through the per-axis maps: the first axis whose entry is a
// Dirichlet side or invalid decides the cell (a Dirichlet state, or zero and
// invalid); reflecting entries met before it negate that axis's momentum;
// with no deciding axis the cell is read from its block at the mapped
// position (invalid where that block is absent).  The other axes' part of
// the walk (those above D come before D, those below after it) and of the
// block and cell index is computed once per row.
template <int D, class Real>
struct PatchRow {
    int r_hi = -1, r_lo = -1;  // status set by the axes above / below D
    unsigned neg_hi = 0, neg_lo = 0;
    int bi = 0, cell = 0;      // block-index and cell-index terms of the other axes
    __device__ __forceinline__ PatchRow(const AmrDevT<Real>& a, const FinePatch& P, const int c[3]) {
#pragma unroll
        for (int d = 2; d >= 0; --d) {
            if (d == D) continue;
            const int e = P.map[d * kPatchLen + c[d]];
            int& r = d > D ? r_hi : r_lo;
            unsigned& ng = d > D ? neg_hi : neg_lo;
            if (r == -1) {
                if (e & kMapDirichlet)
                    r = 2 * d + (e & 1);
                else if (e & kMapInvalid)
                    r = -2;
                else if (e & kMapFlip)
                    ng |= 1u << d;
            }
            const int g = P.f0[d] + (e & 0xfff);
            bi += term_b(a, P, d, g);
            cell += term_c(a, d, g);
        }
    }
    __device__ __forceinline__ static int term_b(const AmrDevT<Real>& a, const FinePatch& P, int d, int g) {
        const int k = (a.act[d] ? g >> a.lb[d] : 0) - P.bb.b0[d];
        return d == 0 ? k * P.bb.nbk[1] * P.bb.nbk[2] : d == 1 ? k * P.bb.nbk[2] : k;
    }
    __device__ __forceinline__ static int term_c(const AmrDevT<Real>& a, int d, int g) {
        const int k = g & (a.bd[d] - 1);
        return d == 0 ? k * a.bd[1] * a.bd[2] : d == 1 ? k * a.bd[2] : k;
    }
    // box cell at position p along D: -1 (block cell, U index base + q * vs),
    // -2 (invalid) or the Dirichlet boundary 2 d + side; neg collects the negated axes
    __device__ __forceinline__ int resolve(const AmrDevT<Real>& a, const FinePatch& P, int p, long long& base,
                                           unsigned& neg) const {
        int r = r_hi;
        neg = neg_hi;
        base = 0;
        if (r != -1) return r;
        const int e = P.map[D * kPatchLen + p];
        if (e & kMapDirichlet) return 2 * D + (e & 1);
        if (e & kMapInvalid) return -2;
        if (e & kMapFlip) neg |= 1u << D;
        neg |= neg_lo;
        if (r_lo != -1) return r_lo;
        const int g = P.f0[D] + (e & 0xfff);
        const int id = P.bids[bi + term_b(a, P, D, g)];
        if (id < 0) return -2;
        base = (long long)id * a.bcells + cell + term_c(a, D, g);
        return -1;
    }
};

// First restriction pass, along axis D, reading the fine patch in place:
// output cell c of the (patch with b cells along D) box takes the first of
// the shifted 6-point stencils whose cells are all valid, else the average
// of its two children, else zero and invalid.
// The output box (extents nsz) starts at off: coarse cell off[D] along D,
// patch position off[e] along the other axes.
template <int D, class Real>
__device__ __forceinline__ void restrict_first_pass(const AmrDevT<Real>& a, const AmrWeightsT<Real>& W, const FinePatch& P,
                                                    const int nsz[3], const int off[3], Real* out, Real* ook,
                                                    long long cap) {
    const int ntot = nsz[0] * nsz[1] * nsz[2];
    for (int i = threadIdx.x; i < ntot; i += blockDim.x) {
        int c[3];
        unflat(i, nsz, c);
#pragma unroll
        for (int e = 0; e < 3; ++e) c[e] += off[e];
        const PatchRow<D, Real> row(a, P, c);
        const int base = 2 * c[D] + 2;
        unsigned okw = 0;  // bit j: patch cell base - 2 + j is valid
#pragma unroll
        for (int j = 0; j < 10; ++j) {
            const int p = base - 2 + j;
            long long bj;
            unsigned nj;
            if (p >= 0 && p < P.sz[D] && row.resolve(a, P, p, bj, nj) != -2) okw |= 1u << j;
        }
        bool done = false;
        Real acc[5] = {Real(0.0), Real(0.0), Real(0.0), Real(0.0), Real(0.0)};
#pragma unroll 1
        for (int m = 0; m < 5 && !done; ++m) {
            const int sh = (m & 1) ? -((m + 1) >> 1) : (m >> 1);  // shifts 0, -1, 1, -2, 2
            const int st = base + sh;
            if (st < 0 || st + 5 >= P.sz[D]) continue;
            if (((okw >> (sh + 2)) & 0x3fu) != 0x3fu) continue;
            // resolve the six cells, then issue their loads together
            long long bs[6];
            int rr[6];
            unsigned ng[6];
#pragma unroll
            for (int kk = 0; kk < 6; ++kk) rr[kk] = row.resolve(a, P, st + kk, bs[kk], ng[kk]);
            Real v[6][5];
#pragma unroll
            for (int kk = 0; kk < 6; ++kk) patch_load(a, rr[kk], bs[kk], ng[kk], v[kk]);
#pragma unroll
            for (int kk = 0; kk < 6; ++kk)
#pragma unroll
                for (int q = 0; q < 5; ++q) acc[q] += W.wts[m][kk] * v[kk][q];
            done = true;
        }
        Real okv = done ? Real(1.0) : Real(0.0);
        if (!done && ((okw >> 4) & 3u) == 3u) {
            Real va[5], vb[5];
            long long ba, bb;
            unsigned na, nb;
            const int ra = row.resolve(a, P, base + 2, ba, na);
            const int rb = row.resolve(a, P, base + 3, bb, nb);
            patch_load(a, ra, ba, na, va);
            patch_load(a, rb, bb, nb, vb);
#pragma unroll
            for (int q = 0; q < 5; ++q) acc[q] = Real(0.5) * (va[q] + vb[q]);
            okv = Real(1.0);
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) out[q * cap + i] = acc[q];
        ook[i] = okv;
    }
}

// One output cell c of a restriction pass along d over a box held in a
// buffer (values in[q * cap + p], validity iok[p], extents sz, coarse cell
// c[d] over box positions 2 c[d] .. 2 c[d] + 9): the first shifted 6-point
// stencil whose cells are all valid, else the average of the two children,
// else zero; returns the validity (1 or 0).
template <class Real>
__device__ __forceinline__ Real restrict_cell(const AmrWeightsT<Real>& W, int d, const int c[3], const int sz[3],
                                                const Real* in, const Real* iok, long long cap, Real acc[5]) {
    const int shifts[5] = {0, -1, 1, -2, 2};
    int s3[3] = {c[0], c[1], c[2]};
    auto at_src = [&](int pos) {
        s3[d] = pos;
        return (s3[0] * sz[1] + s3[1]) * sz[2] + s3[2];
    };
    const int base = 2 * c[d] + 2;
    bool done = false;
#pragma unroll
    for (int q = 0; q < 5; ++q) acc[q] = Real(0.0);
    for (int m = 0; m < 5 && !done; ++m) {
        const int st = base + shifts[m];
        if (st < 0 || st + 5 >= sz[d]) continue;
        bool all = true;
        for (int kk = 0; kk < 6; ++kk) all = all && iok[at_src(st + kk)] != Real(0.0);
        if (all) {
            for (int kk = 0; kk < 6; ++kk) {
                const int sidx = at_src(st + kk);
#pragma unroll
                for (int q = 0; q < 5; ++q) acc[q] += W.wts[m][kk] * in[q * cap + sidx];
            }
            done = true;
        }
    }
    Real okv = done ? Real(1.0) : Real(0.0);
    if (!done) {
        const int pa = at_src(base + 2), pb = at_src(base + 3);
        if (iok[pa] != Real(0.0) && iok[pb] != Real(0.0)) {
#pragma unroll
            for (int q = 0; q < 5; ++q) acc[q] = Real(0.5) * (in[q * cap + pa] + in[q * cap + pb]);
            okv = Real(1.0);
        }
    }
    return okv;
}

// Whether the write mask of k_amr_restrict (bit 27: the whole block; bit
// (o0+1)*9 + (o1+1)*3 + (o2+1): the edge box of width w towards offset o)
// selects any cell of the block box lo .. hi (inclusive; lo == hi: one cell).
template <class Real>
__device__ __forceinline__ bool restrict_selects(const AmrDevT<Real>& a, int mask, int w, const int lo[3], const int hi[3]) {
    if ((mask >> 27) & 1) return true;
    // per axis, offset -1 holds cells c < w, 0 all, 1 cells c >= b - w
    for (int o0 = -1; o0 <= 1; ++o0)
        for (int o1 = -1; o1 <= 1; ++o1)
            for (int o2 = -1; o2 <= 1; ++o2) {
                if (!((mask >> ((o0 + 1) * 9 + (o1 + 1) * 3 + (o2 + 1))) & 1)) continue;
                const int o[3] = {o0, o1, o2};
                bool meets = true;
                for (int d = 0; d < 3; ++d) {
                    const int blo = o[d] == 1 ? (a.bd[d] >= w ? a.bd[d] - w : 0) : 0;
                    const int bhi = o[d] == -1 ? (w < a.bd[d] ? w : a.bd[d]) : a.bd[d];
                    meets = meets && hi[d] >= blo && lo[d] < bhi;
                }
                if (meets) return true;
            }
    return false;
}

// restrict_box of amr.rs:622-730 for a whole ancestor block (level l) by
// one workgroup: over the (2b+8)^D fine patch with its validity mask (level
// l+1 blocks, periodic wrap, bc_box), one pass per axis with the centred
// 6-point stencil, shifted where the fine data ends, the child average as a
// last resort.  Cells left invalid are written as zero, as the reference does.
// The first pass reads the patch in place (FinePatch); a patch reaching past
// a characteristic boundary is gathered into scratch and filled by box_bc
// first.  Later passes read the previous pass from scratch.  Only cells
// selected by `mask` are stored (see k_amr_restrict); a coarse cell's
// stencils stay inside its own box's fine patch, so its value is the same
// whichever box the reference restricts it in.  The nx blocks xs are treated
// as absent from the index (a coarsening that removed them came first).
template <class Real>
__device__ __forceinline__ void restrict_block(const AmrDevT<Real>& a, const AmrWeightsT<Real>& W, int l, int slot,
                                               const int idx[3], int mask, int w, Real* scr, long long cap,
                                               const int* xs, int nx) {
    Real* A = scr;
    Real* Bf = scr + 5 * cap;
    Real* okA = scr + 10 * cap;
    Real* okB = scr + 11 * cap;
    int sz[3];
    FinePatch P;
    P.l = l + 1;
    bool general = false;
    int d0 = 3;  // first active axis
    for (int d = 0; d < 3; ++d) {
        P.f0[d] = a.act[d] ? 2 * idx[d] * a.bd[d] - 4 : 0;
        P.sz[d] = sz[d] = a.act[d] ? 2 * a.bd[d] + 8 : 1;
        P.n3[d] = amr_n_level(a, P.l, d);
        if (a.act[d]) {
            d0 = d < d0 ? d : d0;
            general = general || P.sz[d] > kPatchLen ||
                      (P.f0[d] < 0 && a.bc_kind[d][0] == BC_CHARACTERISTIC) ||
                      (P.f0[d] + P.sz[d] > P.n3[d] && a.bc_kind[d][1] == BC_CHARACTERISTIC);
        }
    }
    __shared__ int bids[kBoxBlocks];
    __shared__ int pmap[3 * kPatchLen];
    P.map = pmap;
    P.xs = xs;
    P.nx = nx;
    if (!general) patch_map(a, P, pmap);
    P.cached = box_blocks(a, P.l, P.f0, P.sz, bids, P.bb);  // synchronises
    P.bids = bids;
    if (P.cached && nx > 0) {
        const int nbid = P.bb.nbk[0] * P.bb.nbk[1] * P.bb.nbk[2];
        for (int i = threadIdx.x; i < nbid; i += blockDim.x)
            for (int k = 0; k < nx; ++k)
                if (bids[i] == xs[k]) bids[i] = -1;
        __syncthreads();
    }
    general = general || !P.cached;  // the in-place reads use the block-id cache
    Real* in = A;
    Real* out = Bf;
    Real* iok = okA;
    Real* ook = okB;
    int dstart = d0;
    if (general) {
        const int tot = sz[0] * sz[1] * sz[2];
        for (int i = threadIdx.x; i < tot; i += blockDim.x) {
            int c[3];
            unflat(i, sz, c);
            Real v[5];
            bool ok = false;
            bool inside = true;
            for (int d = 0; d < 3; ++d) {
                const int g = P.f0[d] + c[d];
                inside = inside && !(a.act[d] && (g < 0 || g >= P.n3[d]) &&
                                     a.bc_kind[d][g >= P.n3[d] ? 1 : 0] != BC_PERIODIC);
            }
            if (inside) {
                int cell;
                const int id = patch_block(a, P, c, &cell);
                ok = id >= 0;
#pragma unroll
                for (int q = 0; q < 5; ++q) v[q] = ok ? a.U[amr_u(a, q, id, cell)] : Real(0.0);
            } else {
#pragma unroll
                for (int q = 0; q < 5; ++q) v[q] = Real(0.0);
            }
#pragma unroll
            for (int q = 0; q < 5; ++q) A[q * cap + i] = v[q];
            okA[i] = ok ? Real(1.0) : Real(0.0);
        }
        __syncthreads();
        box_bc(a, P.l, P.f0, sz, A, okA, cap);
    } else if (d0 < 3) {
        int nsz[3] = {sz[0], sz[1], sz[2]};
        nsz[d0] = a.bd[d0];
        const int zero3[3] = {0, 0, 0};
        if (d0 == 0)
            restrict_first_pass<0>(a, W, P, nsz, zero3, Bf, okB, cap);
        else if (d0 == 1)
            restrict_first_pass<1>(a, W, P, nsz, zero3, Bf, okB, cap);
        else
            restrict_first_pass<2>(a, W, P, nsz, zero3, Bf, okB, cap);
        __syncthreads();
        in = Bf;
        out = A;
        iok = okB;
        ook = okA;
        for (int e = 0; e < 3; ++e) sz[e] = nsz[e];
        dstart = d0 + 1;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {  // unrolled: arrays indexed by d stay in registers
        if (d < dstart || !a.act[d]) continue;
        int nsz[3] = {sz[0], sz[1], sz[2]};
        nsz[d] = a.bd[d];
        const int ntot = nsz[0] * nsz[1] * nsz[2];
        for (int i = threadIdx.x; i < ntot; i += blockDim.x) {
            int c[3];
            unflat(i, nsz, c);
            Real acc[5];
            const Real okv = restrict_cell(W, d, c, sz, in, iok, cap, acc);
#pragma unroll
            for (int q = 0; q < 5; ++q) out[q * cap + i] = acc[q];
            ook[i] = okv;
        }
        __syncthreads();
        Real* t = in;
        in = out;
        out = t;
        t = iok;
        iok = ook;
        ook = t;
        for (int e = 0; e < 3; ++e) sz[e] = nsz[e];
    }
    for (int i = threadIdx.x; i < a.bcells; i += blockDim.x) {
        int c[3];
        unflat(i, a.bd, c);
        if (restrict_selects(a, mask, w, c, c))
#pragma unroll
            for (int q = 0; q < 5; ++q) a.U[amr_u(a, q, slot, i)] = in[q * cap + i];
    }
    __syncthreads();
}

// Restriction of one sub-box of an ancestor block (k_amr_restrict_sub).
// The coarse block is cut into boxes of S[e] cells per axis (nsub[e] of
// them); the stencils of box cells cs .. cs + S - 1 reach patch positions
// 2 cs .. 2 cs + 2 S + 7 along every active axis (F = 2 S + 8 of them), so
// in 3-D the first pass (along x, read in place from the fine blocks)
// produces an S0 x F1 x F2 box, the second an S0 x S1 x F2 box, both held
// in LDS, and the last pass stores into the block.  The fine patch, its validity and every stencil are those of
// restrict_block, in whole-block patch coordinates: each cell gets the same
// value, whichever box computes it.
struct RestrictSub {
    int S[3], nsub[3], nsubt;
    int cap1, cap2;  // cells of the first and second pass outputs
};

template <class Real>
__device__ __forceinline__ void restrict_subbox(const AmrDevT<Real>& a, const AmrWeightsT<Real>& W, int l, int slot,
                                                const int idx[3], int mask, int w, const RestrictSub& rs, int sb,
                                                Real* T1, Real* ok1, Real* T2, Real* ok2, int* bids,
                                                int* pmap) {
    int cs[3], ce[3];
    unflat(sb, rs.nsub, cs);
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        cs[e] *= rs.S[e];
        ce[e] = cs[e] + rs.S[e] - 1;
    }
    if (!restrict_selects(a, mask, w, cs, ce)) return;  // the same for every thread
    FinePatch P;
    P.l = l + 1;
    int d0 = 3;
    for (int d = 0; d < 3; ++d) {
        P.f0[d] = a.act[d] ? 2 * idx[d] * a.bd[d] - 4 : 0;
        P.sz[d] = a.act[d] ? 2 * a.bd[d] + 8 : 1;
        P.n3[d] = amr_n_level(a, P.l, d);
        if (a.act[d] && d < d0) d0 = d;
    }
    P.map = pmap;
    P.xs = nullptr;
    P.nx = 0;
    patch_map(a, P, pmap);
    P.cached = box_blocks(a, P.l, P.f0, P.sz, bids, P.bb);  // synchronises
    P.bids = bids;
    if (!P.cached || d0 == 3) {  // excluded by the launch (blocks of 4 cells or more)
        if (threadIdx.x == 0) atomicOr(&a.sc->err, 8);
        __syncthreads();
        return;
    }
    int sz[3], off[3];
    for (int e = 0; e < 3; ++e) {
        sz[e] = !a.act[e] ? 1 : e == d0 ? rs.S[e] : 2 * rs.S[e] + 8;
        off[e] = !a.act[e] ? 0 : e == d0 ? cs[e] : 2 * cs[e];
    }
    if (d0 == 0)
        restrict_first_pass<0>(a, W, P, sz, off, T1, ok1, rs.cap1);
    else if (d0 == 1)
        restrict_first_pass<1>(a, W, P, sz, off, T1, ok1, rs.cap1);
    else
        restrict_first_pass<2>(a, W, P, sz, off, T1, ok1, rs.cap1);
    __syncthreads();
    const int dlast = a.act[2] ? 2 : a.act[1] ? 1 : 0;
    auto store = [&](int i, const int c[3], const Real* v) {
        int g[3];
#pragma unroll
        for (int e = 0; e < 3; ++e) g[e] = cs[e] + c[e];
        if (restrict_selects(a, mask, w, g, g)) {
            const int cell = (g[0] * a.bd[1] + g[1]) * a.bd[2] + g[2];
#pragma unroll
            for (int q = 0; q < 5; ++q) a.U[amr_u(a, q, slot, cell)] = v[q];
        }
    };
    if (dlast == d0) {  // one active axis: the first pass was the last
        for (int i = threadIdx.x; i < sz[0] * sz[1] * sz[2]; i += blockDim.x) {
            int c[3];
            unflat(i, sz, c);
            Real v[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) v[q] = T1[q * rs.cap1 + i];
            store(i, c, v);
        }
        __syncthreads();
        return;
    }
    const Real* in = T1;
    const Real* iok = ok1;
    long long icap = rs.cap1;
#pragma unroll
    for (int d = 1; d < 3; ++d) {  // unrolled: arrays indexed by d stay in registers
        if (d <= d0 || !a.act[d]) continue;
        int nsz[3] = {sz[0], sz[1], sz[2]};
        nsz[d] = rs.S[d];
        const int ntot = nsz[0] * nsz[1] * nsz[2];
        for (int i = threadIdx.x; i < ntot; i += blockDim.x) {
            int c[3];
            unflat(i, nsz, c);
            Real acc[5];
            const Real okv = restrict_cell(W, d, c, sz, in, iok, icap, acc);
            if (d == dlast) {
                store(i, c, acc);
            } else {
#pragma unroll
                for (int q = 0; q < 5; ++q) T2[q * rs.cap2 + i] = acc[q];
                ok2[i] = okv;
            }
        }
        __syncthreads();
        in = T2;
        iok = ok2;
        icap = rs.cap2;
        for (int e = 0; e < 3; ++e) sz[e] = nsz[e];
    }
}

}  // namespace em
