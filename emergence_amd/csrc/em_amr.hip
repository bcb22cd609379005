// Block-structured AMR on the MI355X: the device-resident block pool, its
// kernels (restriction, ghost slabs, batched flux passes, refluxing, SSP-RK3,
// CFL rate, wavelet details) and the host hierarchy that edits the tree.
//
// The host hierarchy mirrors the reference's tree management
// (fluid/crates/compressible/src/amr.rs: neighbours 919-949, refine_block
// 954-1011 with 2:1 balance, coarsen 1014-1032, regrid_refine/coarsen
// 1051-1106, flags_from_details 877-892, schedule_regrid 1545-1548) and keeps
// block ids in the reference's order (children appended, coarsened children
// swap-removed), so a hierarchy handed over by the reference's `Amr` (its
// `layout()`, amr.rs:1552-1637) evolves identically.  All block data lives on
// the device; a regrid moves only metadata to the host (the per-leaf
// details) and new blocks are predicted / restricted on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <functional>
#include <tuple>
#include <unordered_map>
#include <cstring>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/em_hip.h"
#include "em_amr_kernels.h"
#include "em_amr_types.h"
#include "em_pass.h"
#include "em_types.h"

namespace em {

extern thread_local std::string g_last_error_amr;

struct AmrError : std::runtime_error {
    int code;
    AmrError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

static void hchk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw AmrError(EM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HCHK(x) hchk((x), #x)

// ------------------------------------------------------------------ kernels

__global__ void k_amr_set_run(DevScalars* sc) {
    sc->active = 1;
    sc->err = 0;
}
__global__ void k_amr_set_dt(DevScalars* sc, double dt) { sc->dt = dt; }
__global__ void k_amr_zero_rate(DevScalars* sc) {
    if (threadIdx.x < kSlots) sc->rate[threadIdx.x][0] = 0ull;
}
// dt = cfl / rate on the device (run_steps); collapse flagged in err bit 0
__global__ void k_amr_dt_from_rate(DevScalars* sc, double cfl) {
    unsigned long long m = 0ull;
    for (int i = 0; i < kSlots; ++i) m = sc->rate[i][0] > m ? sc->rate[i][0] : m;
    const double dt = cfl / __longlong_as_double((long long)m);
    if (!(isfinite(dt) && dt > 0.0)) sc->err |= 1;
    sc->dt = dt;
}
__global__ void k_amr_advance(DevScalars* sc) { sc->t = sc->t + sc->dt; }

// dense index updates (block coordinates -> id), applied in stream order
struct MapEdits {
    long long off[8];
    int val[8];
    int n;
};
__global__ void k_amr_map_edit(int* map, MapEdits e) {
    if (threadIdx.x < e.n) map[e.off[threadIdx.x]] = e.val[threadIdx.x];
}
// a list of index edits with distinct entries
__global__ void k_amr_map_set(int* map, const long long* __restrict__ off, const int* __restrict__ val, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) map[off[i]] = val[i];
}

// pool slots src[i] -> tmp block i, and tmp block i -> slot dst[i] (a
// permutation of blocks applied through a staging buffer)
template <class Real>
__global__ void k_amr_blocks_out(AmrDevT<Real> a, const int* __restrict__ src, int n, Real* __restrict__ tmp) {
    const long long tot = (long long)n * a.bcells;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < tot; i += (long long)gridDim.x * 256ll) {
        const long long b = i / a.bcells, c = i % a.bcells;
        for (int q = 0; q < 5; ++q) tmp[(b * 5 + q) * a.bcells + c] = a.U[amr_u(a, q, src[b], c)];
    }
}
template <class Real>
__global__ void k_amr_blocks_in(AmrDevT<Real> a, const int* __restrict__ dst, int n, const Real* __restrict__ tmp) {
    const long long tot = (long long)n * a.bcells;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < tot; i += (long long)gridDim.x * 256ll) {
        const long long b = i / a.bcells, c = i % a.bcells;
        for (int q = 0; q < 5; ++q) a.U[amr_u(a, q, dst[b], c)] = tmp[(b * 5 + q) * a.bcells + c];
    }
}

// staged host data (per block AoS: bcells x 5) -> pool slots
template <class Real>
__global__ void k_amr_scatter(AmrDevT<Real> a, const double* __restrict__ stage, const int* __restrict__ slots, int n) {
    const long long tot = (long long)n * a.bcells;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < tot; i += (long long)gridDim.x * 256ll) {
        const long long b = i / a.bcells, c = i % a.bcells;
        for (int q = 0; q < 5; ++q) a.U[amr_u(a, q, slots[b], c)] = Real(stage[i * 5 + q]);
    }
}
// pool slots -> staging (per block AoS)
template <class Real>
__global__ void k_amr_gather(AmrDevT<Real> a, double* __restrict__ stage, const Real* __restrict__ src, long long n) {
    const long long tot = n * a.bcells;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < tot; i += (long long)gridDim.x * 256ll) {
        const long long b = i / a.bcells, c = i % a.bcells;
        for (int q = 0; q < 5; ++q) stage[i * 5 + q] = (double)src[q * a.vs + b * a.bcells + c];
    }
}
// block data move (coarsening's swap-remove)
template <class Real>
__global__ void k_amr_move(AmrDevT<Real> a, int from, int to) {
    for (long long c = blockIdx.x * 256ll + threadIdx.x; c < a.bcells; c += (long long)gridDim.x * 256ll)
        for (int q = 0; q < 5; ++q) a.U[amr_u(a, q, to, c)] = a.U[amr_u(a, q, from, c)];
}

template <class Real>
__global__ void k_amr_copy_u0(AmrDevT<Real> a, const int* __restrict__ leaves, int n) {
    const long long tot = (long long)n * a.bcells;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < tot; i += (long long)gridDim.x * 256ll) {
        const long long s = leaves[i / a.bcells], c = i % a.bcells;
        for (int q = 0; q < 5; ++q) a.U0[amr_u(a, q, s, c)] = a.U[amr_u(a, q, s, c)];
    }
}

// SSP-RK3 update of the leaves (amr.rs:1476-1498)
template <class Real>
__global__ void k_amr_rk(AmrDevT<Real> a, const int* __restrict__ leaves, int n, int stage) {
    const Real dt = Real(a.sc->dt);
    const long long tot = (long long)n * a.bcells;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < tot; i += (long long)gridDim.x * 256ll) {
        const long long s = leaves[i / a.bcells], c = i % a.bcells;
        for (int q = 0; q < 5; ++q) {
            const long long ix = amr_u(a, q, s, c);
            const Real u = a.U[ix], r = a.R[ix];
            Real o;
            if (stage == 1)
                o = u + dt * r;
            else if (stage == 2)
                o = Real(0.75) * a.U0[ix] + Real(0.25) * (u + dt * r);
            else
                o = a.U0[ix] / Real(3.0) + Real(2.0) / Real(3.0) * (u + dt * r);
            a.U[ix] = o;
        }
    }
}

// CFL rate over leaf cells: sum over active axes of max_speed / dx_level
// (amr.rs:1409-1430)
template <class Real>
__global__ __launch_bounds__(256) void k_amr_rate(AmrDevT<Real> a, const int* __restrict__ leaves,
                                                  const int* __restrict__ lvl, int n) {
    __shared__ double red[4];
    double r = 0.0;  // a max of Real values, exact in double
    const long long tot = (long long)n * a.bcells;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < tot; i += (long long)gridDim.x * 256ll) {
        const long long li = i / a.bcells, c = i % a.bcells;
        const long long s = leaves[li];
        Real u[5], w[5];
        for (int q = 0; q < 5; ++q) u[q] = a.U[amr_u(a, q, s, c)];
        to_prim(a.eos, u, w);
        const Real cs = sound_speed(a.eos, w[0], w[4]);
        Real ac = -Real(0.0);
        for (int d = 0; d < 3; ++d)
            if (a.act[d]) ac += (fabs(w[1 + d]) + cs) / a.dx[lvl[li]][d];
        r = fmax(r, (double)ac);
    }
    for (int off = 32; off > 0; off >>= 1) r = fmax(r, __shfl_xor(r, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        m = m > 0.0 ? m : 0.0;
        atomicMax(&a.sc->rate[blockIdx.x % kSlots][0], (unsigned long long)__double_as_longlong(m));
    }
}

// Threads per workgroup of the patch kernels (restriction, prediction,
// ghost slabs): their gathers are chains of dependent loads (block map, then
// data), latency-bound at a few hundred workgroups, so a workgroup keeps
// sixteen waves of them in flight.
constexpr int kPatchThreads = 1024;
constexpr int kGhostThreads = 256;  // the slab prediction needs more registers

// Ghost slabs of the leaves, one task per workgroup (persistent): same-level
// copy, physical boundary (bc_box on a face slab), or 6th-order prediction
// (fill_scratch, amr.rs:1113-1221).

template <class Real>
__global__ __launch_bounds__(kGhostThreads) void k_amr_ghost(AmrDevT<Real> a, AmrWeightsT<Real> W, const GhostTask* __restrict__ tasks, int n,
                                                  Real* scratch, long long cap) {
    Real* scr = scratch + (long long)blockIdx.x * 12 * cap;
    for (int ti = blockIdx.x; ti < n; ti += gridDim.x) {
        const GhostTask tk = tasks[ti];
        const int d = tk.axis, side = tk.side;
        const int E0 = d == 0 ? 1 : 0, E1 = d == 2 ? 1 : 2;
        const int B = a.bd[d];
        const int nt = a.bcells / B;
        const int ng = a.ng;
        const int s = 2 * d + side;
        if (tk.kind == 0 || tk.kind == 1) {
            for (int it = threadIdx.x; it < ng * nt; it += blockDim.x) {
                const int layer = it / nt, t = it % nt;
                int cc[3];
                cc[E0] = t / a.bd[E1];
                cc[E1] = t % a.bd[E1];
                const int x = side == 0 ? layer - ng : B + layer;
                Real u[5];
                if (tk.kind == 0) {
                    cc[d] = side == 0 ? B + x : x - B;
                    const long long cell = ((long long)cc[0] * a.bd[1] + cc[1]) * a.bd[2] + cc[2];
                    for (int q = 0; q < 5; ++q) u[q] = a.U[amr_u(a, q, tk.src, cell)];
                } else {
                    auto ld = [&](int xx, Real* o) {
                        int c3[3] = {cc[0], cc[1], cc[2]};
                        c3[d] = xx;
                        const long long cell = ((long long)c3[0] * a.bd[1] + c3[1]) * a.bd[2] + c3[2];
                        for (int q = 0; q < 5; ++q) o[q] = a.U[amr_u(a, q, tk.slot, cell)];
                    };
                    const int kind[2] = {a.bc_kind[d][0], a.bc_kind[d][1]};
                    if (d == 0)
                        bc_ghost<0>(a.eos, B, x, kind, a.bc_state[d], ld, u);
                    else if (d == 1)
                        bc_ghost<1>(a.eos, B, x, kind, a.bc_state[d], ld, u);
                    else
                        bc_ghost<2>(a.eos, B, x, kind, a.bc_state[d], ld, u);
                }
                for (int q = 0; q < 5; ++q) a.G[amr_g(a, q, tk.slot, s, layer, t)] = u[q];
            }
        } else {
            int fmin[3], fmax[3];
            for (int e = 0; e < 3; ++e) {
                if (!a.act[e]) {
                    fmin[e] = fmax[e] = 0;
                } else if (e == d) {
                    fmin[e] = side == 0 ? tk.idx[e] * B - ng : tk.idx[e] * B + B;
                    fmax[e] = fmin[e] + ng - 1;
                } else {
                    fmin[e] = tk.idx[e] * a.bd[e];
                    fmax[e] = fmin[e] + a.bd[e] - 1;
                }
            }
            const Real* res = predict_region(a, W, tk.level, fmin, fmax, scr, cap);
            // result box: extents ng along d, bd along the others (row-major)
            int rsz[3];
            for (int e = 0; e < 3; ++e) rsz[e] = (int)(fmax[e] - fmin[e] + 1);
            for (int it = threadIdx.x; it < ng * nt; it += blockDim.x) {
                const int layer = it / nt, t = it % nt;
                int cc[3];
                cc[E0] = t / a.bd[E1];
                cc[E1] = t % a.bd[E1];
                cc[d] = layer;
                const int li = (cc[0] * rsz[1] + cc[1]) * rsz[2] + cc[2];
                for (int q = 0; q < 5; ++q) a.G[amr_g(a, q, tk.slot, s, layer, t)] = res[q * cap + li];
            }
        }
        __syncthreads();
    }
}

// The same-level copies and boundary images of the ghost slabs (task kinds
// 0 and 1), one thread per slab cell: every task has ng * B^(D-1) cells.
// Light enough to run at full occupancy, while the coarse-fine predictions
// (kind 2) keep k_amr_ghost and its register budget.  Tasks write disjoint
// slabs and read only block data, so the split changes no value.
constexpr int kGhostCopyThreads = 256;
template <class Real>
__global__ __launch_bounds__(kGhostCopyThreads) void k_amr_ghost_copy(AmrDevT<Real> a, const GhostTask* __restrict__ tasks,
                                                                       long long ntot, int per) {
    for (long long gi = blockIdx.x * (long long)kGhostCopyThreads + threadIdx.x; gi < ntot;
         gi += (long long)gridDim.x * kGhostCopyThreads) {
        const GhostTask tk = tasks[gi / per];
        const int it = (int)(gi % per);
        const int d = tk.axis, side = tk.side;
        const int E0 = d == 0 ? 1 : 0, E1 = d == 2 ? 1 : 2;
        const int B = a.bd[d];
        const int nt = a.bcells / B;
        const int ng = a.ng;
        const int s = 2 * d + side;
        const int layer = it / nt, t = it % nt;
        int cc[3];
        cc[E0] = t / a.bd[E1];
        cc[E1] = t % a.bd[E1];
        const int x = side == 0 ? layer - ng : B + layer;
        Real u[5];
        if (tk.kind == 0) {
            cc[d] = side == 0 ? B + x : x - B;
            const long long cell = ((long long)cc[0] * a.bd[1] + cc[1]) * a.bd[2] + cc[2];
            for (int q = 0; q < 5; ++q) u[q] = a.U[amr_u(a, q, tk.src, cell)];
        } else {
            auto ld = [&](int xx, Real* o) {
                int c3[3] = {cc[0], cc[1], cc[2]};
                c3[d] = xx;
                const long long cell = ((long long)c3[0] * a.bd[1] + c3[1]) * a.bd[2] + c3[2];
                for (int q = 0; q < 5; ++q) o[q] = a.U[amr_u(a, q, tk.slot, cell)];
            };
            const int kind[2] = {a.bc_kind[d][0], a.bc_kind[d][1]};
            if (d == 0)
                bc_ghost<0>(a.eos, B, x, kind, a.bc_state[d], ld, u);
            else if (d == 1)
                bc_ghost<1>(a.eos, B, x, kind, a.bc_state[d], ld, u);
            else
                bc_ghost<2>(a.eos, B, x, kind, a.bc_state[d], ld, u);
        }
        for (int q = 0; q < 5; ++q) a.G[amr_g(a, q, tk.slot, s, layer, t)] = u[q];
    }
}

// Coarse-fine ghost slabs (task kind 2) with the prediction's gathered
// coarse patch (capA cells and their validity) and its intermediate passes
// (capB cells) in LDS, the last pass storing straight into the slab: the
// arithmetic of k_amr_ghost without its global scratch.  Launched when the
// largest slab's buffers fit in LDS.
template <class Real>
__global__ __launch_bounds__(kGhostThreads) void k_amr_ghost_lds(AmrDevT<Real> a, AmrWeightsT<Real> W,
                                                      const GhostTask* __restrict__ tasks, int n, int capA,
                                                      int capB) {
    extern __shared__ __align__(16) unsigned char lds_bytes[];  // one name for every Real
    Real* lds = reinterpret_cast<Real*>(lds_bytes);
    Real* A = lds;
    Real* ok = A + 5 * capA;
    Real* Bf = ok + capA;
    const XcdStride xs = xcd_stride();  // neighbouring slabs share one L2
    for (long long t = xs.first;; t += xs.step) {
        const long long ti = xs.item(t);
        if (ti >= n) break;
        const GhostTask tk = tasks[ti];
        const int d = tk.axis, side = tk.side;
        const int E0 = d == 0 ? 1 : 0, E1 = d == 2 ? 1 : 2;
        const int B = a.bd[d];
        const int ng = a.ng;
        const int s = 2 * d + side;
        int fmin[3], fmax[3];
        for (int e = 0; e < 3; ++e) {
            if (!a.act[e]) {
                fmin[e] = fmax[e] = 0;
            } else if (e == d) {
                fmin[e] = side == 0 ? tk.idx[e] * B - ng : tk.idx[e] * B + B;
                fmax[e] = fmin[e] + ng - 1;
            } else {
                fmin[e] = tk.idx[e] * a.bd[e];
                fmax[e] = fmin[e] + a.bd[e] - 1;
            }
        }
        predict_box(a, W, tk.level, fmin, fmax, A, ok, capA, Bf, capB, [&](int, const int c[3], const Real* v) {
            const int t = c[E0] * a.bd[E1] + c[E1];
#pragma unroll
            for (int q = 0; q < 5; ++q) a.G[amr_g(a, q, tk.slot, s, c[d], t)] = v[q];
        });
    }
}

// Restriction of the ancestors of one level cut into sub-boxes
// (restrict_subbox), one (ancestor, sub-box) per workgroup at a time, the
// passes after the first in LDS (cap1 + cap2 cells and their validity).
// Same entries and write masks as k_amr_restrict; no blocks are excluded.
constexpr int kRestrictSubThreads = 256;
template <class Real>
__global__ __launch_bounds__(kRestrictSubThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_amr_restrict_sub(AmrDevT<Real> a, AmrWeightsT<Real> W,
                                                               const int* __restrict__ ids,
                                                               const int* __restrict__ ent, int n, int level, int w,
                                                               RestrictSub rs) {
    extern __shared__ __align__(16) unsigned char lds_bytes[];  // one name for every Real
    Real* lds = reinterpret_cast<Real*>(lds_bytes);
    Real* T1 = lds;
    Real* ok1 = T1 + 5 * rs.cap1;
    Real* T2 = ok1 + rs.cap1;
    Real* ok2 = T2 + 5 * rs.cap2;
    __shared__ int bids[kBoxBlocks];
    __shared__ int pmap[3 * kPatchLen];
    const long long total = (long long)n * rs.nsubt;
    // the sub-boxes of an ancestor read overlapping fine data: keep them on
    // one XCD (its L2), in chunks of consecutive items
    const XcdStride xs = xcd_stride();
    for (long long t = xs.first;; t += xs.step) {
        const long long wi = xs.item(t);
        if (wi >= total) break;
        const int i = (int)(wi / rs.nsubt), sb = (int)(wi - (long long)i * rs.nsubt);
        const int bi[3] = {ent[4 * i], ent[4 * i + 1], ent[4 * i + 2]};
        restrict_subbox(a, W, level, ids[i], bi, ent[4 * i + 3], w, rs, sb, T1, ok1, T2, ok2, bids, pmap);
    }
}

// Restriction of a list of ancestors of one level.  Entry i holds the block
// coordinates and a write mask: bit (o0+1)*9 + (o1+1)*3 + (o2+1) selects the
// edge box towards neighbour offset o (cells within w of that side), bit 27
// the whole block (restrict_all's boxes, amr.rs:760-800).  Values do not
// depend on the box they are computed in, so the whole block is computed and
// only the selected cells are stored.  xoff/xs (optional, coarsening
// replay): ancestor i treats the blocks xs[xoff[i] .. xoff[i+1]) as absent.
template <class Real>
__global__ __launch_bounds__(kPatchThreads) void k_amr_restrict(AmrDevT<Real> a, AmrWeightsT<Real> W, const int* __restrict__ ids,
                                                     const int* __restrict__ ent, int n, int level, int w,
                                                     Real* scratch, long long cap, const int* __restrict__ xoff,
                                                     const int* __restrict__ xs) {
    Real* scr = scratch + (long long)blockIdx.x * 12 * cap;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int bi[3] = {ent[4 * i], ent[4 * i + 1], ent[4 * i + 2]};
        const int nx = xoff ? xoff[i + 1] - xoff[i] : 0;
        restrict_block(a, W, level, ids[i], bi, ent[4 * i + 3], w, scr, cap, nx ? xs + xoff[i] : nullptr, nx);
    }
}

// Prediction of whole refined children from their parent level
// (refine_block, amr.rs:989-996).
template <class Real>
__global__ __launch_bounds__(kPatchThreads) void k_amr_predict_blocks(AmrDevT<Real> a, AmrWeightsT<Real> W, const int* __restrict__ slots,
                                                           const int* __restrict__ idx, int n, int level,
                                                           Real* scratch, long long cap) {
    Real* scr = scratch + (long long)blockIdx.x * 12 * cap;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        int fmin[3], fmax[3];
        for (int e = 0; e < 3; ++e) {
            fmin[e] = idx[3 * i + e] * a.bd[e];
            fmax[e] = fmin[e] + a.bd[e] - 1;
            if (!a.act[e]) fmin[e] = fmax[e] = 0;
        }
        const Real* res = predict_region(a, W, level, fmin, fmax, scr, cap);
        for (int c = threadIdx.x; c < a.bcells; c += blockDim.x)
            for (int q = 0; q < 5; ++q) a.U[amr_u(a, q, slots[i], c)] = res[q * cap + c];
        __syncthreads();
    }
}

// Refluxing of one (axis, side): coarse rhs += +-(mean fine flux - coarse
// flux) / dx (amr.rs:1327-1406); one launch per (axis, side) in the
// reference's order, so no cell is corrected twice in one launch.
template <class Real>
__global__ void k_amr_reflux(AmrDevT<Real> a, const RefluxTask* __restrict__ t, int n, int d, int side) {
    for (int i = blockIdx.x * 128 + threadIdx.x; i < n; i += gridDim.x * 128) {
        const RefluxTask k = t[i];
        const Real dx = a.dx[k.level][d];
        const Real cnt = Real(k.nfine);
        for (int q = 0; q < 5; ++q) {
            Real sum = Real(0.0);
            for (int f = 0; f < k.nfine; ++f) sum += a.F[amr_f(a, q, k.fslot[f], 2 * d + (1 - side), k.fface[f])];
            const Real fc = a.F[amr_f(a, q, k.slot, 2 * d + side, k.face)];
            const Real df = sum / cnt - fc;
            const Real dr = side == 1 ? -df / dx : df / dx;
            const long long ix = amr_u(a, q, k.slot, k.cell);
            a.R[ix] += dr;
        }
    }
}

// Every reflux correction of the batched step in one launch: the tasks of
// all (axis, side) lists grouped by the coarse cell they correct, each group
// in the (axis, side) order of k_amr_reflux's launches, so a cell's residual
// takes its corrections in the same order and every cell is one thread's.
template <class Real>
__global__ void k_amr_reflux_cells(AmrDevT<Real> a, const RefluxTask* __restrict__ t, const int* __restrict__ grp, int ng) {
    for (int gi = blockIdx.x * 128 + threadIdx.x; gi < ng; gi += gridDim.x * 128) {
        for (int i = grp[gi]; i < grp[gi + 1]; ++i) {
            const RefluxTask k = t[i];
            const int d = k.axis, side = k.side;
            const Real dx = a.dx[k.level][d];
            const Real cnt = Real(k.nfine);
            for (int q = 0; q < 5; ++q) {
                Real sum = Real(0.0);
                for (int f = 0; f < k.nfine; ++f) sum += a.F[amr_f(a, q, k.fslot[f], 2 * d + (1 - side), k.fface[f])];
                const Real fc = a.F[amr_f(a, q, k.slot, 2 * d + side, k.face)];
                const Real df = sum / cnt - fc;
                const Real dr = side == 1 ? -df / dx : df / dx;
                const long long ix = amr_u(a, q, k.slot, k.cell);
                a.R[ix] += dr;
            }
        }
    }
}

// Largest relative wavelet detail of each leaf (block_detail, amr.rs:823-867):
// |s - sum_m wr[m] s(+-1, +-3, +-5)| of the sensor values ln rho,
// ln(p + p_inf) along every active axis; infinite if not finite.
template <class Real>
__global__ __launch_bounds__(256) void k_amr_details(AmrDevT<Real> a, AmrWeightsT<Real> W, const int* __restrict__ leaves, int n,
                                                    double* __restrict__ det) {
    __shared__ Real red[4];
    for (int li = blockIdx.x; li < n; li += gridDim.x) {
        const long long slot = leaves[li];
        Real mx = Real(0.0);
        for (int c = threadIdx.x; c < a.bcells; c += blockDim.x) {
            int ci[3];
            unflat(c, a.bd, ci);
            auto sens = [&](int d, int x, Real s[2]) {
                Real u[5], w[5];
                if (x >= 0 && x < a.bd[d]) {
                    int c3[3] = {ci[0], ci[1], ci[2]};
                    c3[d] = x;
                    const int cell = (c3[0] * a.bd[1] + c3[1]) * a.bd[2] + c3[2];
                    for (int q = 0; q < 5; ++q) u[q] = a.U[amr_u(a, q, slot, cell)];
                } else {
                    const int E0 = d == 0 ? 1 : 0, E1 = d == 2 ? 1 : 2;
                    const int t = ci[E0] * a.bd[E1] + ci[E1];
                    const int layer = x < 0 ? x + a.ng : x - a.bd[d];
                    for (int q = 0; q < 5; ++q) u[q] = a.G[amr_g(a, q, slot, 2 * d + (x < 0 ? 0 : 1), layer, t)];
                }
                to_prim(a.eos, u, w);
                s[0] = log(fmax(w[0], Real(1e-300)));
                s[1] = log(fmax(w[4] + a.eos.p_inf, Real(1e-300)));
            };
            for (int d = 0; d < 3; ++d) {
                if (!a.act[d]) continue;
                Real s0[2];
                sens(d, ci[d], s0);
                Real pred[2] = {Real(0.0), Real(0.0)};
                const int offs[6] = {-5, -3, -1, 1, 3, 5};
                for (int m = 0; m < 6; ++m) {
                    Real sv[2];
                    sens(d, ci[d] + offs[m], sv);
                    for (int q = 0; q < 2; ++q) pred[q] += W.wr[m] * sv[q];
                }
                for (int q = 0; q < 2; ++q) {
                    const Real dd = fabs(s0[q] - pred[q]);
                    mx = isfinite(dd) ? fmax(mx, dd) : INFINITY;
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0) det[li] = (double)fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        __syncthreads();
    }
}

// k_amr_details for blocks of at most 512 cells with face slabs of at most
// 64 cells: every sensor value a leaf's stencils read (its cells, and the
// five slab layers nearest each face) is computed once into LDS, instead of
// once per stencil that reads it; the values and the order of the sums are
// those of k_amr_details.
constexpr int kDetCells = 512, kDetFace = 64;
template <class Real>
__global__ __launch_bounds__(256) void k_amr_details_lds(AmrDevT<Real> a, AmrWeightsT<Real> W, const int* __restrict__ leaves, int n,
                                                        double* __restrict__ det) {
    __shared__ Real red[4];
    __shared__ Real si[2][kDetCells];
    __shared__ Real sg[2][6][5][kDetFace];
    auto sensor = [&](const Real u[5], Real s[2]) {
        Real w[5];
        to_prim(a.eos, u, w);
        s[0] = log(fmax(w[0], Real(1e-300)));
        s[1] = log(fmax(w[4] + a.eos.p_inf, Real(1e-300)));
    };
    for (int li = blockIdx.x; li < n; li += gridDim.x) {
        const long long slot = leaves[li];
        for (int c = threadIdx.x; c < a.bcells; c += blockDim.x) {
            Real u[5], sv[2];
            for (int q = 0; q < 5; ++q) u[q] = a.U[amr_u(a, q, slot, c)];
            sensor(u, sv);
            si[0][c] = sv[0];
            si[1][c] = sv[1];
        }
        // slab entry (face f = 2 d + side, layer k = 0..4 from the face outwards, t)
        for (int e = threadIdx.x; e < 6 * 5 * kDetFace; e += blockDim.x) {
            const int f = e / (5 * kDetFace), k = (e / kDetFace) % 5, t = e % kDetFace;
            const int d = f >> 1, side = f & 1;
            if (!a.act[d] || t >= a.bcells / a.bd[d]) continue;
            const int layer = side == 0 ? a.ng - 1 - k : k;
            Real u[5], sv[2];
            for (int q = 0; q < 5; ++q) u[q] = a.G[amr_g(a, q, slot, f, layer, t)];
            sensor(u, sv);
            sg[0][f][k][t] = sv[0];
            sg[1][f][k][t] = sv[1];
        }
        __syncthreads();
        Real mx = Real(0.0);
        for (int c = threadIdx.x; c < a.bcells; c += blockDim.x) {
            int ci[3];
            unflat(c, a.bd, ci);
            auto sens = [&](int d, int x, Real s[2]) {
                if (x >= 0 && x < a.bd[d]) {
                    int c3[3] = {ci[0], ci[1], ci[2]};
                    c3[d] = x;
                    const int cell = (c3[0] * a.bd[1] + c3[1]) * a.bd[2] + c3[2];
                    s[0] = si[0][cell];
                    s[1] = si[1][cell];
                } else {
                    const int E0 = d == 0 ? 1 : 0, E1 = d == 2 ? 1 : 2;
                    const int t = ci[E0] * a.bd[E1] + ci[E1];
                    const int f = 2 * d + (x < 0 ? 0 : 1);
                    const int k = x < 0 ? -1 - x : x - a.bd[d];
                    s[0] = sg[0][f][k][t];
                    s[1] = sg[1][f][k][t];
                }
            };
            for (int d = 0; d < 3; ++d) {
                if (!a.act[d]) continue;
                Real s0[2];
                sens(d, ci[d], s0);
                Real pred[2] = {Real(0.0), Real(0.0)};
                const int offs[6] = {-5, -3, -1, 1, 3, 5};
                for (int m = 0; m < 6; ++m) {
                    Real sv[2];
                    sens(d, ci[d] + offs[m], sv);
                    for (int q = 0; q < 2; ++q) pred[q] += W.wr[m] * sv[q];
                }
                for (int q = 0; q < 2; ++q) {
                    const Real dd = fabs(s0[q] - pred[q]);
                    mx = isfinite(dd) ? fmax(mx, dd) : INFINITY;
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0) det[li] = (double)fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        __syncthreads();
    }
}

// ------------------------------------------------------------------ host

static void lagrange6_h(double x0, double x, double out[6]) {  // amr.rs:41-47
    for (int k = 0; k < 6; ++k) {
        const double xk = x0 + (double)k;
        double p = 1.0;
        for (int j = 0; j < 6; ++j) {
            if (j == k) continue;
            p *= (x - (x0 + (double)j)) / (xk - (x0 + (double)j));
        }
        out[k] = p;
    }
}

static inline int64_t hdiv(int64_t a, int64_t b) {
    int64_t q = a / b;
    if (a % b < 0) q = b > 0 ? q - 1 : q + 1;
    return q;
}
static inline int64_t hrem(int64_t a, int64_t b) {
    int64_t r = a % b;
    return r < 0 ? r + (b < 0 ? -b : b) : r;
}

struct HBlock {
    int level;
    int64_t idx[3];
    bool leaf;
};

using InitFn = void (*)(const double x[3], const double* prm, double w[5]);

// The host hierarchy and the device pool of one block-AMR solver; Real is
// the type of the pool data and its arithmetic (double; float for the FP32
// variant, the reference's GpuAmr<f32>).  Times, dt, mesh spacing,
// wavelet details and the tree logic stay double on the host.
template <class Real>
struct AmrT {
    // configuration
    int B, levels, ng, regrid_every;
    double eps;
    int64_t base[3];
    double lo[3], hi[3];
    Eos eos;
    int bck[3][2];
    double bcs[3][2][5];
    int mode, positivity;
    double cfl, tol;
    int precise = 1;
    bool batched = true;
    bool act[3];
    int bd[3];
    int ndim, bcells, tmax;
    AmrWeightsT<Real> W;
    // hierarchy (ids = pool slots)
    std::vector<HBlock> blocks;
    std::vector<std::vector<int>> map;  // per level, dense
    std::vector<long long> map_off;
    double t = 0.0;
    int64_t steps = 0, regrid_interval, next_regrid;
    // device
    hipStream_t stream = nullptr;
    AmrDevT<Real> dev{};
    long long cap = 0;
    DevScalars* sc = nullptr;
    int* d_map = nullptr;
    long long map_total = 0;
    Real* big_scr = nullptr;
    Real* small_scr = nullptr;
    long long big_cap = 0, small_cap = 0;
    int big_wg = 64, small_wg = 512;
    // LDS launches of the batched stages (0 bytes: the global-scratch kernel)
    int ghost_capA = 0, ghost_capB = 0;
    size_t ghost_lds = 0;
    RestrictSub rsub{};
    size_t rsub_lds = 0;
    // tables
    std::vector<int> leaves, leaf_level;
    std::vector<std::vector<int>> edge, edge_ent;
    std::vector<GhostTask> ghost;
    std::vector<RefluxTask> reflux[3][2];
    int* d_leaves = nullptr;
    int* d_leaf_level = nullptr;
    GhostTask* d_ghost = nullptr;
    std::vector<GhostTask> ghost_copy, ghost_pred;  // the batched launch's split of `ghost` by kind
    int n_bnd_copy = 0;  // ghost_copy holds its boundary images (kind 1) first, then the same-level copies
    std::vector<int> gsrc;  // same-level source per (block, face) for the batched passes (AmrPassArgs::gsrc)
    int* d_gsrc = nullptr;
    GhostTask* d_ghost_copy = nullptr;
    GhostTask* d_ghost_pred = nullptr;
    std::vector<int*> d_edge_ids, d_edge_idx;
    RefluxTask* d_reflux[3][2] = {};
    std::vector<RefluxTask> reflux_cells;  // all reflux tasks grouped by coarse cell (k_amr_reflux_cells)
    std::vector<int> reflux_grp;           // group starts, and the end
    RefluxTask* d_reflux_cells = nullptr;
    int* d_reflux_grp = nullptr;
    int* d_tmp = nullptr;
    long long d_tmp_n = 0;
    double* d_det = nullptr;
    long long d_det_n = 0;
    bool tables_dirty = true;

    int64_t nbl(int l, int d) const { return act[d] ? base[d] << l : 1; }
    int64_t n_level(int l, int d) const { return act[d] ? base[d] * B * ((int64_t)1 << l) : 1; }
    double dxl(int l, int d) const { return (hi[d] - lo[d]) / (double)n_level(l, d); }
    std::vector<int> dims() const {
        std::vector<int> v;
        for (int d = 0; d < 3; ++d)
            if (act[d]) v.push_back(d);
        return v;
    }

    AmrT(const int64_t base_[3], int B_, int levels_, int ng_, double eps_, int regrid_every_, const double lo_[3],
        const double hi_[3], const double e[4], const int32_t bk[6], const double bs[30], int mode_, int pos_,
        double cfl_, double tol_) {
        B = B_;
        levels = levels_;
        ng = std::max(ng_, 5);
        eps = eps_;
        regrid_every = regrid_every_;
        if (levels < 1 || levels > kMaxLevels) throw AmrError(EM_ERR_ARG, "levels must be in 1..8");
        if (B != 8 && B != 16 && B != 32) throw AmrError(EM_ERR_ARG, "block size must be 8, 16 or 32");
        if (mode_ < 0 || mode_ > 2) throw AmrError(EM_ERR_ARG, "mode must be 0, 1 or 2");
        ndim = 0;
        for (int d = 0; d < 3; ++d) {
            base[d] = base_[d];
            lo[d] = lo_[d];
            hi[d] = hi_[d];
            act[d] = base[d] > 0 && (hi[d] - lo[d]) > 0.0 && B > 1;  // amr.rs:250-252
            bd[d] = base[d] == 0 ? 1 : B;
            ndim += act[d] ? 1 : 0;
            for (int s = 0; s < 2; ++s) {
                bck[d][s] = bk[2 * d + s];
                for (int q = 0; q < 5; ++q) bcs[d][s][q] = bs ? bs[(2 * d + s) * 5 + q] : 0.0;
            }
        }
        if (ndim == 0) throw AmrError(EM_ERR_ARG, "no active axis");
        if (B < ng) throw AmrError(EM_ERR_ARG, "block size must be at least the ghost depth");
        bcells = bd[0] * bd[1] * bd[2];
        int mn = B;
        tmax = bcells / mn;
        eos = Eos{e[0], e[1], e[2], e[3]};
        mode = mode_;
        positivity = pos_;
        cfl = cfl_;
        tol = tol_;
        regrid_interval = regrid_every;
        next_regrid = regrid_every;
        // the operator weights in double, then rounded to Real (the
        // reference's R::from_f64 of its f64 weight tables)
        AmrWeightsT<double> wd;
        lagrange6_h(-3.0, -0.25, wd.wp[0]);
        lagrange6_h(-2.0, 0.25, wd.wp[1]);
        const double wr[6] = {3.0 / 256.0, -25.0 / 256.0, 150.0 / 256.0, 150.0 / 256.0, -25.0 / 256.0, 3.0 / 256.0};
        for (int k = 0; k < 6; ++k) wd.wr[k] = wd.wts[0][k] = wr[k];
        const int shifts[5] = {0, -1, 1, -2, 2};
        for (int m = 1; m < 5; ++m) lagrange6_h(-2.5 + (double)shifts[m], 0.0, wd.wts[m]);
        for (int k = 0; k < 6; ++k) {
            W.wp[0][k] = Real(wd.wp[0][k]);
            W.wp[1][k] = Real(wd.wp[1][k]);
            W.wr[k] = Real(wd.wr[k]);
            for (int m = 0; m < 5; ++m) W.wts[m][k] = Real(wd.wts[m][k]);
        }
        // dense index
        map.resize(levels);
        map_off.resize(levels);
        map_total = 0;
        for (int l = 0; l < levels; ++l) {
            int64_t n = 1;
            for (int d = 0; d < 3; ++d) n *= nbl(l, d);
            map[l].assign((size_t)n, -1);
            map_off[l] = map_total;
            map_total += n;
        }
        HCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HCHK(hipMalloc(&sc, sizeof(DevScalars)));
        std::vector<DevScalars> hs(1);
        std::memset(hs.data(), 0, sizeof(DevScalars));
        hs[0].active = 1;
        HCHK(hipMemcpyAsync(sc, hs.data(), sizeof(DevScalars), hipMemcpyHostToDevice, stream));
        HCHK(hipMalloc(&d_map, sizeof(int) * map_total));
        std::vector<int> neg(map_total, -1);
        HCHK(hipMemcpyAsync(d_map, neg.data(), sizeof(int) * map_total, hipMemcpyHostToDevice, stream));
        HCHK(hipStreamSynchronize(stream));
        // scratch: restriction patch (2b+8)^D or prediction patch / block
        // Scratch per workgroup: 12 arrays of cap doubles (two 5-variable
        // buffers and two validity masks).  A separable pass turns one axis
        // from coarse to fine extent, so max(coarse, fine) per axis bounds
        // every intermediate.  Restriction patches are (2b+8)^D (and bound
        // whole-block predictions: b/2+7 coarse, b fine); ghost slabs have
        // at most ng/2+8 coarse and ng fine cells along their axis.
        long long rp = 1;
        for (int d = 0; d < 3; ++d) rp *= act[d] ? 2 * B + 8 : 1;
        long long sp = 1;
        for (int d = 0; d < 3; ++d) {
            if (!act[d]) continue;
            long long v = std::max(ng / 2 + 8, ng);
            for (int e = 0; e < 3; ++e)
                if (e != d && act[e]) v *= std::max(B / 2 + 7, B);
            sp = std::max(sp, v);
        }
        big_cap = std::max<long long>(rp, bcells);
        small_cap = sp;
        // workgroups of the restriction / block-prediction kernels: up to
        // 1024, with their scratch held under 2 GiB
        big_wg = (int)std::max<long long>(16, std::min<long long>(1024, (2ll << 30) / (96ll * big_cap)));
        HCHK(hipMalloc(&big_scr, sizeof(Real) * 12 * big_cap * big_wg));
        HCHK(hipMalloc(&small_scr, sizeof(Real) * 12 * small_cap * small_wg));
        lds_plans();
        dev.map = d_map;
        dev.sc = sc;
        for (int d = 0; d < 3; ++d) {
            dev.bd[d] = bd[d];
            dev.lb[d] = bd[d] == 32 ? 5 : bd[d] == 16 ? 4 : bd[d] == 8 ? 3 : 0;
            dev.act[d] = act[d] ? 1 : 0;
            dev.nb0[d] = (int)nbl(0, d);
            for (int s = 0; s < 2; ++s) {
                dev.bc_kind[d][s] = bck[d][s];
                for (int q = 0; q < 5; ++q) dev.bc_state[d][s][q] = Real(bcs[d][s][q]);
            }
        }
        dev.ndim = ndim;
        dev.bcells = bcells;
        dev.tmax = tmax;
        dev.ng = ng;
        dev.levels = levels;
        for (int l = 0; l < levels; ++l) {
            dev.map_off[l] = map_off[l];
            for (int d = 0; d < 3; ++d) dev.dx[l][d] = Real(dxl(l, d));
        }
        dev.eos = eos_as<Real>(eos);
        ensure_capacity(64);
    }

    ~AmrT() {
        auto f = [](void* p) {
            if (p) (void)hipFree(p);
        };
        f(dev.U);
        f(dev.U0);
        f(dev.R);
        f(dev.G);
        f(dev.F);
        f(sc);
        f(d_map);
        f(big_scr);
        f(small_scr);
        f(d_leaves);
        f(d_leaf_level);
        f(d_ghost);
        f(d_ghost_copy);
        f(d_ghost_pred);
        f(d_gsrc);
        for (auto p : d_edge_ids) f(p);
        for (auto p : d_edge_idx) f(p);
        f(d_reflux_cells);
        f(d_reflux_grp);
        for (auto& r : d_reflux)
            for (auto p : r) f(p);
        f(d_tmp);
        f(d_det);
        if (pin) (void)hipHostFree(pin);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (auto& e : prof_ev) {
            (void)hipEventDestroy(std::get<1>(e));
            (void)hipEventDestroy(std::get<2>(e));
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    void ensure_capacity(long long n) {
        if (n <= cap) return;
        long long nc = std::max<long long>(n, cap * 3 / 2 + 16);
        Real *U, *U0, *R, *G, *F;
        const long long vs = nc * bcells, gs = nc * 6 * ng * (long long)tmax, fs = nc * 6 * (long long)tmax;
        HCHK(hipMalloc(&U, sizeof(Real) * 5 * vs));
        HCHK(hipMalloc(&U0, sizeof(Real) * 5 * vs));
        HCHK(hipMalloc(&R, sizeof(Real) * 5 * vs));
        HCHK(hipMalloc(&G, sizeof(Real) * 5 * gs));
        HCHK(hipMalloc(&F, sizeof(Real) * 5 * fs));
        if (cap > 0) {
            for (int q = 0; q < 5; ++q)
                HCHK(hipMemcpyAsync(U + q * vs, dev.U + q * dev.vs, sizeof(Real) * dev.vs, hipMemcpyDeviceToDevice,
                                    stream));
            HCHK(hipStreamSynchronize(stream));
            (void)hipFree(dev.U);
            (void)hipFree(dev.U0);
            (void)hipFree(dev.R);
            (void)hipFree(dev.G);
            (void)hipFree(dev.F);
        }
        ++version;
        dev.U = U;
        dev.U0 = U0;
        dev.R = R;
        dev.G = G;
        dev.F = F;
        dev.cap = nc;
        dev.vs = vs;
        dev.gs = gs;
        dev.fs = fs;
        cap = nc;
    }

    template <class T>
    T* dev_vec(const std::vector<T>& v, T* old) {
        // device copy of a table, reusing the old buffer when it is large
        // enough (an empty table keeps the buffer; its count guards use)
        const size_t need = sizeof(T) * v.size();
        if (v.empty()) return old;
        if (old && dev_caps[old] >= need) {
            HCHK(hipMemcpyAsync(old, v.data(), need, hipMemcpyHostToDevice, stream));
            return old;
        }
        if (old) {
            HCHK(hipFree(old));
            dev_caps.erase(old);
        }
        const size_t cap = need + need / 2 + 256;
        T* p;
        HCHK(hipMalloc(&p, cap));
        dev_caps[p] = cap;
        HCHK(hipMemcpyAsync(p, v.data(), need, hipMemcpyHostToDevice, stream));
        return p;
    }
    std::unordered_map<const void*, size_t> dev_caps;
    // Pinned staging of the table uploads of build_tables: one host copy per
    // table into pinned memory, then an asynchronous copy from there; the
    // region is reused only after the stream has been synchronised.
    char* pin = nullptr;
    size_t pin_cap = 0, pin_off = 0;
    void pin_reserve(size_t bytes) {
        if (bytes <= pin_cap) return;
        if (pin) HCHK(hipHostFree(pin));
        pin_cap = bytes + bytes / 2 + 4096;
        HCHK(hipHostMalloc((void**)&pin, pin_cap, hipHostMallocDefault));
    }
    template <class T>
    static size_t staged_bytes(const std::vector<T>& v) {
        return (sizeof(T) * v.size() + 255) / 256 * 256;
    }
    template <class T>
    T* dev_vec_staged(const std::vector<T>& v, T* old) {
        const size_t need = sizeof(T) * v.size();
        if (v.empty()) return old;
        T* p = old;
        if (!(old && dev_caps[old] >= need)) {
            if (old) {
                HCHK(hipFree(old));
                dev_caps.erase(old);
            }
            const size_t cap = need + need / 2 + 256;
            HCHK(hipMalloc(&p, cap));
            dev_caps[p] = cap;
        }
        char* h = pin + pin_off;
        std::memcpy(h, v.data(), need);
        pin_off += staged_bytes(v);
        HCHK(hipMemcpyAsync(p, h, need, hipMemcpyHostToDevice, stream));
        return p;
    }

    // -------------------------------------------------------------- hierarchy
    long long slot(int l, const int64_t idx[3]) const {
        if (l < 0 || l >= levels) return -1;
        long long s = 0;
        for (int d = 0; d < 3; ++d) {
            const int64_t n = nbl(l, d);
            if (idx[d] < 0 || idx[d] >= n) return -1;
            s = s * n + idx[d];
        }
        return s;
    }
    int get(int l, const int64_t idx[3]) const {
        long long s = slot(l, idx);
        return s < 0 ? -1 : map[l][s];
    }
    void set_map(int l, const int64_t idx[3], int id, MapEdits& ed) {
        long long s = slot(l, idx);
        if (s < 0) return;
        map[l][s] = id;
        ed.off[ed.n] = map_off[l] + s;
        ed.val[ed.n] = id;
        ed.n++;
    }
    // Index edits are collected on the host and applied in one launch before
    // the next device use of the index (apply_edits).
    std::vector<long long> pend_off;
    std::vector<int> pend_val;
    void flush_edits(MapEdits& ed) {
        for (int i = 0; i < ed.n; ++i) {
            pend_off.push_back(ed.off[i]);
            pend_val.push_back(ed.val[i]);
        }
        ed.n = 0;
    }
    void map_host_set(int l, const int64_t idx[3], int id) {
        const long long s = slot(l, idx);
        if (s >= 0) map[l][s] = id;
    }
    long long* d_eoff = nullptr;
    int* d_eval = nullptr;
    // uploads and applies an edit list; the last edit of an entry wins
    void apply_edit_list(std::vector<long long>& off, std::vector<int>& val) {
        if (off.empty()) return;
        std::unordered_map<long long, size_t> last;
        for (size_t i = 0; i < off.size(); ++i) last[off[i]] = i;
        std::vector<long long> o;
        std::vector<int> v;
        for (size_t i = 0; i < off.size(); ++i)
            if (last[off[i]] == i) {
                o.push_back(off[i]);
                v.push_back(val[i]);
            }
        HCHK(hipStreamSynchronize(stream));  // the edit buffers may be reallocated
        d_eoff = dev_vec(o, d_eoff);
        d_eval = dev_vec(v, d_eval);
        const unsigned g = (unsigned)std::min<size_t>((o.size() + 255) / 256, 1024);
        hipLaunchKernelGGL(k_amr_map_set, dim3(g), dim3(256), 0, stream, d_map, d_eoff, d_eval, (int)o.size());
        HCHK(hipGetLastError());
        HCHK(hipStreamSynchronize(stream));  // o, v are copied from pageable memory
        off.clear();
        val.clear();
    }
    void apply_edits() { apply_edit_list(pend_off, pend_val); }

    std::vector<std::array<int64_t, 3>> neighbours(int l, const int64_t idx[3]) const {  // amr.rs:919-949
        std::vector<std::array<int64_t, 3>> out;
        int r0[3], r1[3];
        for (int d = 0; d < 3; ++d) {
            r0[d] = act[d] ? -1 : 0;
            r1[d] = act[d] ? 1 : 0;
        }
        for (int a = r0[0]; a <= r1[0]; ++a)
            for (int b = r0[1]; b <= r1[1]; ++b)
                for (int c = r0[2]; c <= r1[2]; ++c) {
                    if (a == 0 && b == 0 && c == 0) continue;
                    std::array<int64_t, 3> k = {idx[0] + a, idx[1] + b, idx[2] + c};
                    bool inside = true;
                    for (int d = 0; d < 3; ++d) {
                        const int64_t nb = nbl(l, d);
                        if (k[d] < 0 || k[d] >= nb) {
                            if (bck[d][k[d] >= nb ? 1 : 0] == BC_PERIODIC)
                                k[d] = hrem(k[d], nb);
                            else
                                inside = false;
                        }
                    }
                    if (inside) out.push_back(k);
                }
        return out;
    }

    std::vector<std::array<int64_t, 3>> children(const int64_t idx[3]) const {  // amr.rs:894-916
        std::vector<std::array<int64_t, 3>> out{{0, 0, 0}};
        for (int d = 0; d < 3; ++d) {
            std::vector<std::array<int64_t, 3>> next;
            for (auto c : out) {
                if (act[d]) {
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

    void x_of(int l, const int64_t g[3], double x[3]) const {  // amr.rs:283-285
        for (int d = 0; d < 3; ++d) x[d] = act[d] ? lo[d] + ((double)g[d] + 0.5) * dxl(l, d) : 0.5 * (lo[d] + hi[d]);
    }

    // block data from an initial condition (to_cons of init at cell centres)
    void init_data(int l, const int64_t idx[3], InitFn f, const double* prm, double* out) const {
        for (int i = 0; i < bd[0]; ++i)
            for (int j = 0; j < bd[1]; ++j)
                for (int k = 0; k < bd[2]; ++k) {
                    const int64_t g[3] = {idx[0] * bd[0] + i, idx[1] * bd[1] + j, idx[2] * bd[2] + k};
                    double x[3], w[5];
                    x_of(l, g, x);
                    f(x, prm, w);
                    const double rho = w[0];
                    const double ke = 0.5 * rho * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
                    const double v = 1.0 / rho;
                    const double rhoe = rho * ((w[4] + eos.gamma * eos.p_inf) * (v - eos.b) / (eos.gamma - 1.0) + eos.q);
                    double* o = out + ((size_t)((i * bd[1] + j) * bd[2] + k)) * 5;
                    o[0] = rho;
                    o[1] = rho * w[1];
                    o[2] = rho * w[2];
                    o[3] = rho * w[3];
                    o[4] = rhoe + ke;
                }
    }

    void upload_blocks(const std::vector<int>& slots, const double* data) {
        if (slots.empty()) return;
        const long long n = (long long)slots.size();
        double* stage;
        int* ds;
        HCHK(hipMalloc(&stage, sizeof(double) * 5 * bcells * n));
        HCHK(hipMalloc(&ds, sizeof(int) * n));
        HCHK(hipMemcpyAsync(stage, data, sizeof(double) * 5 * bcells * n, hipMemcpyHostToDevice, stream));
        HCHK(hipMemcpyAsync(ds, slots.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream));
        const unsigned g = (unsigned)std::min<long long>((n * bcells + 255) / 256, 4096);
        hipLaunchKernelGGL(k_amr_scatter<Real>, dim3(g), dim3(256), 0, stream, dev, stage, ds, (int)n);
        HCHK(hipGetLastError());
        HCHK(hipStreamSynchronize(stream));
        (void)hipFree(stage);
        (void)hipFree(ds);
    }

    int add_block(int l, const int64_t idx[3], MapEdits& ed) {
        HBlock b;
        b.level = l;
        for (int d = 0; d < 3; ++d) b.idx[d] = idx[d];
        b.leaf = true;
        const int id = (int)blocks.size();
        ensure_capacity(id + 1);
        set_map(l, idx, id, ed);
        blocks.push_back(b);
        if (ed.n >= 6) flush_edits(ed);
        return id;
    }

    // refine_block (amr.rs:954-1011): balance first, then 2^D children
    // predicted on the device (or evaluated from an initial condition).
    void refine_block(int id0, InitFn init, const double* prm) {
        const int l = blocks[id0].level;
        int64_t idx[3] = {blocks[id0].idx[0], blocks[id0].idx[1], blocks[id0].idx[2]};
        if (!blocks[id0].leaf) return;
        if (l > 0) {
            for (auto& nk : neighbours(l, idx)) {
                if (get(l, nk.data()) < 0) {
                    for (int m = l - 1; m >= 0; --m) {
                        int64_t pk[3];
                        for (int d = 0; d < 3; ++d) pk[d] = act[d] ? (nk[d] >> (l - m)) : 0;
                        const int pid = get(m, pk);
                        if (pid >= 0) {
                            refine_block(pid, init, prm);
                            break;
                        }
                    }
                }
            }
        }
        const int id = get(l, idx);
        if (id < 0 || !blocks[id].leaf) return;
        auto kids = children(idx);
        MapEdits ed{};
        std::vector<int> slots;
        std::vector<int> kidx;
        for (auto& c : kids) {
            // ids appended in children order, as the reference's push
            const int nid = (int)blocks.size() + (int)slots.size();
            slots.push_back(nid);
            for (int d = 0; d < 3; ++d) kidx.push_back((int)c[d]);
        }
        ensure_capacity((long long)blocks.size() + (long long)kids.size());
        if (init) {
            std::vector<double> data((size_t)5 * bcells * kids.size());
            for (size_t q = 0; q < kids.size(); ++q) init_data(l + 1, kids[q].data(), init, prm, data.data() + q * 5 * bcells);
            upload_blocks(slots, data.data());
        } else {
            // Predicted from level l, which a refinement never changes and
            // whose blocks around this one exist (balanced above): deferred
            // to one launch per level (flush_predictions).
            for (size_t q = 0; q < slots.size(); ++q)
                pending.push_back({l + 1, slots[q], kidx[3 * q], kidx[3 * q + 1], kidx[3 * q + 2]});
        }
        blocks[id].leaf = false;
        for (size_t q = 0; q < kids.size(); ++q) add_block(l + 1, kids[q].data(), ed);
        flush_edits(ed);
        tables_dirty = true;
    }

    // Children awaiting prediction: {level, slot, idx[3]}.  Levels run in
    // increasing order, so a child predicted from a block refined in the
    // same regrid reads that block's predicted data.
    std::vector<std::array<int, 5>> pending;
    std::vector<std::vector<int>> pending_host;

    void flush_predictions() {
        if (pending.empty()) return;
        apply_edits();  // the predictions read the index
        std::stable_sort(pending.begin(), pending.end(),
                         [](const std::array<int, 5>& a, const std::array<int, 5>& b) { return a[0] < b[0]; });
        ensure_tmp((long long)pending.size() * 4);
        size_t off = 0;
        pending_host.clear();
        for (size_t i = 0; i < pending.size();) {
            size_t j = i;
            while (j < pending.size() && pending[j][0] == pending[i][0]) ++j;
            const int n = (int)(j - i);
            std::vector<int> h(4 * n);
            for (int k = 0; k < n; ++k) {
                h[k] = pending[i + k][1];
                for (int d = 0; d < 3; ++d) h[n + 3 * k + d] = pending[i + k][2 + d];
            }
            HCHK(hipMemcpyAsync(d_tmp + off, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, stream));
            const unsigned g = (unsigned)std::min(n, big_wg);
            hipLaunchKernelGGL(k_amr_predict_blocks<Real>, dim3(g), dim3(kPatchThreads), 0, stream, dev, W, d_tmp + off,
                               d_tmp + off + n, n, pending[i][0], big_scr, big_cap);
            HCHK(hipGetLastError());
            pending_host.push_back(std::move(h));
            off += 4 * n;
            i = j;
        }
        HCHK(hipStreamSynchronize(stream));
        pending.clear();
        pending_host.clear();
    }

    void ensure_tmp(long long n) {
        if (n <= d_tmp_n) return;
        if (d_tmp) HCHK(hipFree(d_tmp));
        HCHK(hipMalloc(&d_tmp, sizeof(int) * n));
        d_tmp_n = n;
    }

    // coarsen (amr.rs:1014-1032), host part: the parent becomes a leaf and
    // its children are swap-removed (ids descending), exactly as the
    // reference orders them; the device work (restriction of the parent,
    // index edits, moves) is recorded and replayed in batches by
    // replay_coarsening.
    struct CoOp {
        int l;
        int64_t idx[3];
        int pid0;                       // the parent's pool slot when the coarsening phase began
        int batch;
        std::vector<long long> removed;  // index entries of its children
        std::vector<int> kids0;          // its children's pool slots when the phase began
    };
    std::vector<CoOp> co_ops;
    std::vector<int> orig_slot;  // phase-start pool slot of the data that belongs at slot s

    void coarsen(int l, const int64_t idx[3]) {
        auto kids = children(idx);
        const int pid = get(l, idx);
        CoOp op;
        op.l = l;
        for (int d = 0; d < 3; ++d) op.idx[d] = idx[d];
        op.pid0 = orig_slot[pid];
        op.batch = 0;
        blocks[pid].leaf = true;
        std::vector<int> ids;
        for (auto& c : kids) ids.push_back(get(l + 1, c.data()));
        std::sort(ids.begin(), ids.end(), [](int a, int b) { return a > b; });
        for (int i : ids) {
            op.kids0.push_back(orig_slot[i]);
            op.removed.push_back(map_off[blocks[i].level] + slot(blocks[i].level, blocks[i].idx));
            map_host_set(blocks[i].level, blocks[i].idx, -1);
            const int last = (int)blocks.size() - 1;
            if (i != last) {
                blocks[i] = blocks[last];
                orig_slot[i] = orig_slot[last];
            }
            blocks.pop_back();
            orig_slot.pop_back();
            if (i < (int)blocks.size()) map_host_set(blocks[i].level, blocks[i].idx, i);
        }
        co_ops.push_back(std::move(op));
        tables_dirty = true;
    }

    // index v of level L along d inside [lo, hi] (periodic axes wrap)
    bool in_span(int L, int d, int64_t v, int64_t lo, int64_t hi) const {
        if (!act[d]) return true;
        if (v >= lo && v <= hi) return true;
        if (bck[d][0] != BC_PERIODIC) return false;
        const int64_t nb = nbl(L, d);
        return (v - nb >= lo && v - nb <= hi) || (v + nb >= lo && v + nb <= hi);
    }
    // the restriction of k reads what coarsening j (earlier) changes: j's
    // removed children (same level, neighbours) or j's restricted parent
    // (one level finer, inside k's fine patch of blocks 2i-1 .. 2i+2)
    bool co_depends(const CoOp& k, const CoOp& j) const {
        if (j.l == k.l) {
            for (int d = 0; d < 3; ++d)
                if (!in_span(k.l, d, j.idx[d], k.idx[d] - 1, k.idx[d] + 1)) return false;
            return true;
        }
        if (j.l == k.l + 1) {
            for (int d = 0; d < 3; ++d)
                if (!in_span(j.l, d, j.idx[d], 2 * k.idx[d] - 1, 2 * k.idx[d] + 2)) return false;
            return true;
        }
        return false;
    }

    // Device side of the recorded coarsenings.  A restriction reads level
    // l+1 through the index and writes its parent; it sees the same data as
    // in the reference's sequential loop when every earlier coarsening it
    // depends on (co_depends) has been applied and no later one it would see
    // has.  A coarsening one level finer writes a parent the restriction
    // reads: levels run finest first, one launch each.  A coarsening on the
    // same level only removes children from the index: each restriction
    // masks the children of the earlier ones it depends on (their pool slots
    // when the phase began) out of its patch, so a level's restrictions run
    // together against the index and pool as they were when the phase began.
    // Then the removals are applied to the index, the swap-remove moves at
    // once as one block permutation, and the index entries of moved blocks
    // rewritten.
    Real* d_co_tmp = nullptr;
    long long d_co_tmp_n = 0;
    int* d_co_int = nullptr;

    void replay_coarsening() {
        const int n = (int)co_ops.size();
        if (n == 0) return;
        // int tables per level: parent slots, 4 ints each (index, whole-block
        // mask), exclusion offsets (n + 1) and the excluded slots
        struct Launch {
            int level, ids, ent, xoff, xs, n;
        };
        std::vector<int> ib;
        std::vector<Launch> rl;
        for (int l = levels - 1; l >= 0; --l) {
            std::vector<int> ids, ent, xoff, xs;
            for (int k = 0; k < n; ++k) {
                const CoOp& op = co_ops[k];
                if (op.l != l) continue;
                ids.push_back(op.pid0);
                for (int d = 0; d < 3; ++d) ent.push_back((int)op.idx[d]);
                ent.push_back(1 << 27);  // the whole block
                xoff.push_back((int)xs.size());
                for (int j = 0; j < k; ++j)
                    if (co_ops[j].l == l && co_depends(op, co_ops[j]))
                        xs.insert(xs.end(), co_ops[j].kids0.begin(), co_ops[j].kids0.end());
            }
            if (ids.empty()) continue;
            xoff.push_back((int)xs.size());
            Launch L{l, 0, 0, 0, 0, (int)ids.size()};
            L.ids = (int)ib.size();
            ib.insert(ib.end(), ids.begin(), ids.end());
            L.ent = (int)ib.size();
            ib.insert(ib.end(), ent.begin(), ent.end());
            L.xoff = (int)ib.size();
            ib.insert(ib.end(), xoff.begin(), xoff.end());
            L.xs = (int)ib.size();
            ib.insert(ib.end(), xs.begin(), xs.end());
            rl.push_back(L);
        }
        // block permutation of the moves
        std::vector<int> src, dst;
        for (size_t sl = 0; sl < orig_slot.size(); ++sl)
            if (orig_slot[sl] != (int)sl) {
                src.push_back(orig_slot[sl]);
                dst.push_back((int)sl);
            }
        const int psrc = (int)ib.size();
        ib.insert(ib.end(), src.begin(), src.end());
        const int pdst = (int)ib.size();
        ib.insert(ib.end(), dst.begin(), dst.end());
        // index edits: every removal, then the moved blocks' entries (all
        // entries distinct), uploaded once
        std::vector<long long> eo;
        std::vector<int> ev;
        for (const CoOp& op : co_ops)
            for (long long e : op.removed) {
                eo.push_back(e);
                ev.push_back(-1);
            }
        const int nrem = (int)eo.size();
        for (int sl : dst) {
            eo.push_back(map_off[blocks[sl].level] + slot(blocks[sl].level, blocks[sl].idx));
            ev.push_back(sl);
        }
        const int nall = (int)eo.size();
        apply_edits();  // earlier edits (refinement) first; synchronises
        d_co_int = dev_vec(ib, d_co_int);
        d_eoff = dev_vec(eo, d_eoff);
        d_eval = dev_vec(ev, d_eval);
        const long long need = (long long)src.size() * 5 * bcells;
        if (need > d_co_tmp_n) {
            if (d_co_tmp) HCHK(hipFree(d_co_tmp));
            HCHK(hipMalloc(&d_co_tmp, sizeof(Real) * (size_t)need));
            d_co_tmp_n = need;
        }
        auto edits = [&](int lo, int hi) {
            if (hi <= lo) return;
            const unsigned g = (unsigned)std::min((hi - lo + 255) / 256, 1024);
            hipLaunchKernelGGL(k_amr_map_set, dim3(g), dim3(256), 0, stream, d_map, d_eoff + lo, d_eval + lo, hi - lo);
            HCHK(hipGetLastError());
        };
        for (const Launch& L : rl) {
            const unsigned g = (unsigned)std::min(L.n, big_wg);
            hipLaunchKernelGGL(k_amr_restrict<Real>, dim3(g), dim3(kPatchThreads), 0, stream, dev, W, d_co_int + L.ids,
                               d_co_int + L.ent, L.n, L.level, box_w(), big_scr, big_cap, d_co_int + L.xoff,
                               d_co_int + L.xs);
            HCHK(hipGetLastError());
        }
        edits(0, nrem);
        if (!src.empty()) {
            const int np = (int)src.size();
            const unsigned g = (unsigned)std::min<long long>(((long long)np * bcells + 255) / 256, 4096);
            hipLaunchKernelGGL(k_amr_blocks_out<Real>, dim3(g), dim3(256), 0, stream, dev, d_co_int + psrc, np, d_co_tmp);
            HCHK(hipGetLastError());
            hipLaunchKernelGGL(k_amr_blocks_in<Real>, dim3(g), dim3(256), 0, stream, dev, d_co_int + pdst, np, d_co_tmp);
            HCHK(hipGetLastError());
        }
        edits(nrem, nall);
        HCHK(hipStreamSynchronize(stream));  // ib, eo, ev are copied from pageable memory
        co_ops.clear();
    }

    std::vector<int8_t> flags_from_details(const std::vector<double>& det) const {  // amr.rs:877-892
        std::vector<int8_t> f(blocks.size(), 0);
        for (size_t i = 0; i < blocks.size(); ++i) {
            if (!blocks[i].leaf)
                f[i] = 0;
            else if (det[i] > eps && blocks[i].level + 1 < levels)
                f[i] = 1;
            else if (det[i] < eps / 128.0)
                f[i] = -1;
            else
                f[i] = 0;
        }
        return f;
    }

    struct Key {
        int l;
        int64_t i[3];
        bool operator<(const Key& o) const {
            if (l != o.l) return l < o.l;
            for (int d = 0; d < 3; ++d)
                if (i[d] != o.i[d]) return i[d] < o.i[d];
            return false;
        }
        bool operator==(const Key& o) const { return l == o.l && i[0] == o.i[0] && i[1] == o.i[1] && i[2] == o.i[2]; }
    };

    std::set<Key> regrid_refine(const std::vector<int8_t>& fl) {  // amr.rs:1051-1077
        std::vector<Key> want;
        for (size_t i = 0; i < fl.size(); ++i) {
            if (fl[i] > 0) {
                const HBlock& b = blocks[i];
                want.push_back(Key{b.level, {b.idx[0], b.idx[1], b.idx[2]}});
                for (auto& nk : neighbours(b.level, b.idx)) {
                    const int j = get(b.level, nk.data());
                    if (j >= 0 && blocks[j].leaf) want.push_back(Key{b.level, {nk[0], nk[1], nk[2]}});
                }
            }
        }
        std::sort(want.begin(), want.end());
        want.erase(std::unique(want.begin(), want.end()), want.end());
        std::set<Key> prot(want.begin(), want.end());
        for (const Key& k : want) {
            const int id = get(k.l, k.i);
            if (id >= 0 && k.l + 1 < levels) refine_block(id, nullptr, nullptr);
        }
        flush_predictions();
        return prot;
    }

    void regrid_coarsen(const std::vector<int8_t>& fl, const std::set<Key>& prot) {  // amr.rs:1081-1106
        const size_t nkids = (size_t)1 << ndim;
        std::vector<Key> cand;
        for (size_t bi = 0; bi < blocks.size(); ++bi) {
            const HBlock& b = blocks[bi];
            if (b.leaf) continue;
            auto kids = children(b.idx);
            bool ok = kids.size() == nkids;
            for (auto& c : kids) {
                if (!ok) break;
                const int j = get(b.level + 1, c.data());
                ok = j >= 0 && blocks[j].leaf && fl[j] < 0 && prot.find(Key{b.level + 1, {c[0], c[1], c[2]}}) == prot.end();
            }
            if (ok) cand.push_back(Key{b.level, {b.idx[0], b.idx[1], b.idx[2]}});
        }
        std::stable_sort(cand.begin(), cand.end(), [](const Key& a, const Key& b) { return b.l < a.l; });
        orig_slot.resize(blocks.size());
        for (size_t i = 0; i < orig_slot.size(); ++i) orig_slot[i] = (int)i;
        co_ops.clear();
        for (const Key& k : cand) {
            bool ok = true;
            for (auto& nk : neighbours(k.l, k.i))
                for (auto& c : children(nk.data())) {
                    const int j = get(k.l + 1, c.data());
                    if (j >= 0 && !blocks[j].leaf) ok = false;
                }
            if (ok && get(k.l, k.i) >= 0) coarsen(k.l, k.i);
        }
        replay_coarsening();
    }

    // -------------------------------------------------------------- tables
    void build_tables() {
        apply_edits();
        HCHK(hipStreamSynchronize(stream));  // tables in use by queued work are freed below
        leaves.clear();
        leaf_level.clear();
        for (size_t i = 0; i < blocks.size(); ++i)
            if (blocks[i].leaf) {
                leaves.push_back((int)i);
                leaf_level.push_back(blocks[i].level);
            }
        // edge ancestors: an ancestor with a same-level neighbour that is a
        // leaf or absent (amr.rs:739-745)
        edge.assign(levels, {});
        for (size_t i = 0; i < blocks.size(); ++i) {
            const HBlock& b = blocks[i];
            if (b.leaf) continue;
            bool e = false;
            for (auto& nk : neighbours(b.level, b.idx)) {
                const int j = get(b.level, nk.data());
                if (j < 0 || blocks[j].leaf) e = true;
            }
            if (e) edge[b.level].push_back((int)i);
        }
        // ghost slabs of the leaves
        ghost.clear();
        for (size_t li = 0; li < leaves.size(); ++li) {
            const HBlock& b = blocks[leaves[li]];
            for (int d = 0; d < 3; ++d) {
                if (!act[d]) continue;
                for (int side = 0; side < 2; ++side) {
                    GhostTask t{};
                    t.slot = leaves[li];
                    t.axis = d;
                    t.side = side;
                    t.level = b.level;
                    for (int e = 0; e < 3; ++e) t.idx[e] = (int)b.idx[e];
                    int64_t nk[3] = {b.idx[0], b.idx[1], b.idx[2]};
                    nk[d] += side == 1 ? 1 : -1;
                    const int64_t nb = nbl(b.level, d);
                    bool inside = true;
                    if (nk[d] < 0 || nk[d] >= nb) {
                        if (bck[d][side] == BC_PERIODIC)
                            nk[d] = hrem(nk[d], nb);
                        else
                            inside = false;
                    }
                    if (!inside) {
                        t.kind = 1;
                    } else {
                        const int j = get(b.level, nk);
                        if (j >= 0) {
                            t.kind = 0;
                            t.src = j;
                        } else {
                            t.kind = 2;
                        }
                    }
                    ghost.push_back(t);
                }
            }
        }
        // reflux tasks per (axis, side), in the reference's block order
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s) reflux[d][s].clear();
        for (size_t id = 0; id < blocks.size(); ++id) {
            const HBlock& b = blocks[id];
            const int l = b.level;
            if (!b.leaf || l + 1 >= levels) continue;
            for (int d = 0; d < 3; ++d) {
                if (!act[d]) continue;
                const int e0 = d == 0 ? 1 : 0, e1 = d == 2 ? 1 : 2;
                const int64_t nbk = nbl(l, d);
                for (int side = 0; side < 2; ++side) {
                    int64_t nk[3] = {b.idx[0], b.idx[1], b.idx[2]};
                    nk[d] += side == 1 ? 1 : -1;
                    if (nk[d] < 0 || nk[d] >= nbk) {
                        if (bck[d][side] != BC_PERIODIC) continue;
                        nk[d] = hrem(nk[d], nbk);
                    }
                    const int nj = get(l, nk);
                    if (!(nj >= 0 && !blocks[nj].leaf)) continue;
                    // the fine leaves across the face: children (h0, h1) of the
                    // neighbour next to it, looked up once; fine face (fa, fb)
                    // of coarse face (a, bb) lies in child ((2a + fa) / B,
                    // (2bb + fb) / B) at ((2a + fa) % B, (2bb + fb) % B)
                    const int fa_n = act[e0] ? 2 : 1, fb_n = act[e1] ? 2 : 1;
                    int kid[2][2] = {{-1, -1}, {-1, -1}};
                    for (int h0 = 0; h0 < fa_n; ++h0)
                        for (int h1 = 0; h1 < fb_n; ++h1) {
                            int64_t ck[3];
                            ck[d] = 2 * nk[d] + (side == 1 ? 0 : 1);
                            ck[e0] = act[e0] ? 2 * b.idx[e0] + h0 : 0;
                            ck[e1] = act[e1] ? 2 * b.idx[e1] + h1 : 0;
                            const int cid = get(l + 1, ck);
                            kid[h0][h1] = (cid >= 0 && blocks[cid].leaf) ? cid : -1;
                        }
                    for (int a = 0; a < bd[e0]; ++a)
                        for (int bb = 0; bb < bd[e1]; ++bb) {
                            RefluxTask t{};
                            t.slot = (int)id;
                            t.level = l;
                            t.face = a * bd[e1] + bb;
                            int c[3];
                            c[d] = side == 1 ? bd[d] - 1 : 0;
                            c[e0] = a;
                            c[e1] = bb;
                            t.cell = (c[0] * bd[1] + c[1]) * bd[2] + c[2];
                            for (int fa = 0; fa < fa_n; ++fa)
                                for (int fb = 0; fb < fb_n; ++fb) {
                                    const int ga = 2 * a + fa, gb = 2 * bb + fb;
                                    const int cid = kid[act[e0] ? ga / bd[e0] : 0][act[e1] ? gb / bd[e1] : 0];
                                    if (cid < 0) continue;
                                    const int la = act[e0] ? ga % bd[e0] : 0;
                                    const int lb = act[e1] ? gb % bd[e1] : 0;
                                    t.fslot[t.nfine] = cid;
                                    t.fface[t.nfine] = la * bd[e1] + lb;
                                    t.nfine++;
                                }
                            if (t.nfine == 0) continue;
                            t.axis = d;
                            t.side = side;
                            reflux[d][side].push_back(t);
                        }
                }
            }
        }
        edge_ent.assign(levels, {});
        for (int l = 0; l < levels; ++l) {
            std::vector<int>& ix = edge_ent[l];
            for (int id : edge[l]) {
                for (int d = 0; d < 3; ++d) ix.push_back((int)blocks[id].idx[d]);
                ix.push_back(box_mask(id));
            }
        }
        // uploads through the pinned staging region (the stream was
        // synchronised above, so it is free)
        ghost_copy.clear();
        ghost_pred.clear();
        gsrc.assign(blocks.size() * 6, -1);
        for (const GhostTask& t : ghost) {
            if (t.kind == 1) ghost_copy.push_back(t);
            if (t.kind == 2) ghost_pred.push_back(t);
            if (t.kind == 0) gsrc[(size_t)t.slot * 6 + 2 * t.axis + t.side] = t.src;
        }
        n_bnd_copy = (int)ghost_copy.size();
        for (const GhostTask& t : ghost)
            if (t.kind == 0) ghost_copy.push_back(t);
        // reflux tasks by coarse cell, each cell's in (axis, side) order
        reflux_cells.clear();
        for (int d = 0; d < 3; ++d)
            for (int sd = 0; sd < 2; ++sd) reflux_cells.insert(reflux_cells.end(), reflux[d][sd].begin(), reflux[d][sd].end());
        std::stable_sort(reflux_cells.begin(), reflux_cells.end(), [](const RefluxTask& x, const RefluxTask& y) {
            return x.slot != y.slot ? x.slot < y.slot : x.cell < y.cell;
        });
        reflux_grp.clear();
        for (size_t i = 0; i < reflux_cells.size(); ++i)
            if (i == 0 || reflux_cells[i].slot != reflux_cells[i - 1].slot || reflux_cells[i].cell != reflux_cells[i - 1].cell)
                reflux_grp.push_back((int)i);
        reflux_grp.push_back((int)reflux_cells.size());
        size_t total = staged_bytes(leaves) + staged_bytes(leaf_level) + staged_bytes(ghost) +
                       staged_bytes(ghost_copy) + staged_bytes(ghost_pred) + staged_bytes(gsrc) +
                       staged_bytes(reflux_cells) + staged_bytes(reflux_grp);
        for (int l = 0; l < levels; ++l) total += staged_bytes(edge[l]) + staged_bytes(edge_ent[l]);
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s) total += staged_bytes(reflux[d][s]);
        pin_reserve(total);
        pin_off = 0;
        d_leaves = dev_vec_staged(leaves, d_leaves);
        d_leaf_level = dev_vec_staged(leaf_level, d_leaf_level);
        d_ghost = dev_vec_staged(ghost, d_ghost);
        d_ghost_copy = dev_vec_staged(ghost_copy, d_ghost_copy);
        d_ghost_pred = dev_vec_staged(ghost_pred, d_ghost_pred);
        d_gsrc = dev_vec_staged(gsrc, d_gsrc);
        d_edge_ids.resize(levels, nullptr);
        d_edge_idx.resize(levels, nullptr);
        for (int l = 0; l < levels; ++l) {
            d_edge_ids[l] = dev_vec_staged(edge[l], d_edge_ids[l]);
            d_edge_idx[l] = dev_vec_staged(edge_ent[l], d_edge_idx[l]);
        }
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s) d_reflux[d][s] = dev_vec_staged(reflux[d][s], d_reflux[d][s]);
        d_reflux_cells = dev_vec_staged(reflux_cells, d_reflux_cells);
        d_reflux_grp = dev_vec_staged(reflux_grp, d_reflux_grp);
        HCHK(hipStreamSynchronize(stream));
        tables_dirty = false;
        ++version;
    }

    void tables() {
        if (tables_dirty) build_tables();
        apply_edits();
    }

    // Buffers of the LDS kernels, in cells.  k_amr_ghost_lds: the coarse
    // patch of a slab and the outputs of the passes before the last, over
    // every slab axis, side and parity of the slab's first cell (as
    // predict_box sizes them).  k_amr_restrict_sub: sub-boxes shaped as
    // below.  Either runs when its buffers fit in 76 KiB (two workgroups per
    // CU) and, for the restriction, no axis has a characteristic boundary
    // (its patches take restrict_block's gathered path).
    void lds_plans() {
        constexpr size_t kLdsCap = 76 * 1024;
        const int dlast = act[2] ? 2 : act[1] ? 1 : 0;
        long long ca = 0, cb = 0;
        for (int ds = 0; ds < 3; ++ds) {
            if (!act[ds]) continue;
            for (int side = 0; side < 2; ++side)
                for (int k = 0; k < 4; ++k) {
                    long long fmin[3], fmax[3], sz[3];
                    for (int e = 0; e < 3; ++e) {
                        if (!act[e]) {
                            fmin[e] = fmax[e] = 0;
                        } else if (e == ds) {
                            fmin[e] = side == 0 ? (long long)k * bd[e] - ng : (long long)k * bd[e] + bd[e];
                            fmax[e] = fmin[e] + ng - 1;
                        } else {
                            fmin[e] = (long long)k * bd[e];
                            fmax[e] = fmin[e] + bd[e] - 1;
                        }
                        sz[e] = act[e] ? (hdiv(fmax[e], 2) + 3) - (hdiv(fmin[e], 2) - 3) + 1 : 1;
                    }
                    ca = std::max(ca, sz[0] * sz[1] * sz[2]);
                    int pass = 0;
                    for (int d = 0; d < 3; ++d) {
                        if (!act[d]) continue;
                        sz[d] = fmax[d] - fmin[d] + 1;
                        if (d == dlast) break;
                        long long& c = (pass & 1) ? ca : cb;  // Bf, then A, then Bf
                        c = std::max(c, sz[0] * sz[1] * sz[2]);
                        ++pass;
                    }
                }
        }
        const size_t gb = sizeof(Real) * (6 * ca + 5 * cb);
        if (gb <= kLdsCap) {
            ghost_capA = (int)ca;
            ghost_capB = (int)cb;
            ghost_lds = gb;
            HCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_amr_ghost_lds<Real>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)gb));
        }
        bool charac = false;
        int d0 = -1;
        for (int d = 0; d < 3; ++d)
            if (act[d]) {
                charac = charac || bck[d][0] == BC_CHARACTERISTIC || bck[d][1] == BC_CHARACTERISTIC;
                if (d0 < 0) d0 = d;
            }
        RestrictSub r{};
        auto F = [&](int e) { return act[e] ? 2 * r.S[e] + 8 : 1; };
        auto rbytes = [&] {
            r.cap1 = r.S[d0];
            for (int e = 0; e < 3; ++e)
                if (e != d0) r.cap1 *= F(e);
            r.cap2 = ndim == 3 ? r.S[0] * r.S[1] * F(2) : 0;
            return sizeof(Real) * 6 * ((size_t)r.cap1 + r.cap2);
        };
        for (int e = 0; e < 3; ++e) r.S[e] = act[e] ? bd[e] : 1;
        if (d0 >= 0) {
            // cuts across the first pass's axis repeat no work: one coarse
            // layer per box in 3-D, at most 16 in 2-D; then halve the widest
            // other axis while the buffers exceed the cap (each halving
            // recomputes part of the first pass's patch positions)
            r.S[d0] = ndim == 3 ? 1 : std::min(bd[d0], 16);
            while (rbytes() > kLdsCap) {
                int e = -1;
                for (int k = 0; k < 3; ++k)
                    if (act[k] && k != d0 && r.S[k] > 1 && (e < 0 || r.S[k] >= r.S[e])) e = k;
                if (e < 0 || r.S[e] <= 4) break;
                r.S[e] /= 2;
            }
        }
        r.nsubt = 1;
        for (int e = 0; e < 3; ++e) {
            r.nsub[e] = bd[e] / r.S[e];
            r.nsubt *= r.nsub[e];
        }
        const size_t rb = rbytes();
        if (!charac && d0 >= 0 && B >= 4 && 2 * B + 8 <= kPatchLen && rb <= kLdsCap) {
            rsub = r;
            rsub_lds = rb;
            HCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_amr_restrict_sub<Real>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)rb));
        }
    }

    // width of restrict_all's edge boxes (amr.rs:782)
    int box_w() const { return std::min(ng + 3, std::max(bd[0], std::max(bd[1], bd[2]))); }

    // write mask of an edge ancestor (amr.rs:766-800): a box towards every
    // offset whose neighbour is absent or a leaf (outside non-periodic
    // boundaries: none); the whole block when the boxes cover more than half
    int box_mask(int id) const {
        const HBlock& b = blocks[id];
        const int w = box_w();
        int mask = 0;
        long long vol = 0;
        int r0[3], r1[3];
        for (int d = 0; d < 3; ++d) {
            r0[d] = act[d] ? -1 : 0;
            r1[d] = act[d] ? 1 : 0;
        }
        for (int o0 = r0[0]; o0 <= r1[0]; ++o0)
            for (int o1 = r0[1]; o1 <= r1[1]; ++o1)
                for (int o2 = r0[2]; o2 <= r1[2]; ++o2) {
                    if (o0 == 0 && o1 == 0 && o2 == 0) continue;
                    const int o[3] = {o0, o1, o2};
                    int64_t nk[3];
                    bool skip = false;
                    for (int d = 0; d < 3; ++d) {
                        const int64_t nb = nbl(b.level, d);
                        nk[d] = b.idx[d] + o[d];
                        if (nk[d] < 0 || nk[d] >= nb) {
                            if (bck[d][nk[d] >= nb ? 1 : 0] == BC_PERIODIC)
                                nk[d] = hrem(nk[d], nb);
                            else
                                skip = true;
                        }
                    }
                    if (skip) continue;
                    const int j = get(b.level, nk);
                    if (j >= 0 && !blocks[j].leaf) continue;
                    mask |= 1 << ((o0 + 1) * 9 + (o1 + 1) * 3 + (o2 + 1));
                    long long v = 1;
                    for (int d = 0; d < 3; ++d) {
                        const int lo_ = o[d] == 1 ? (bd[d] >= w ? bd[d] - w : 0) : 0;
                        const int hi_ = o[d] == -1 ? std::min(w, bd[d]) : bd[d];
                        v *= hi_ - lo_;
                    }
                    vol += v;
                }
        if (2 * vol > (long long)bcells) mask = 1 << 27;
        return mask;
    }

    // -------------------------------------------------------------- stages
    void restrict_all() {  // amr.rs:737-816, finest level first
        for (int l = levels - 2; l >= 0; --l) {
            const int n = (int)edge[l].size();
            if (!n) continue;
            if (batched && rsub_lds) {
                const unsigned g = (unsigned)std::min<long long>(((long long)n * rsub.nsubt + 7) / 8 * 8, 16384);
                hipLaunchKernelGGL(k_amr_restrict_sub<Real>, dim3(g), dim3(kRestrictSubThreads), rsub_lds, stream, dev, W,
                                   d_edge_ids[l], d_edge_idx[l], n, l, box_w(), rsub);
            } else if (batched) {
                const unsigned g = (unsigned)std::min(n, big_wg);
                hipLaunchKernelGGL(k_amr_restrict<Real>, dim3(g), dim3(kPatchThreads), 0, stream, dev, W, d_edge_ids[l], d_edge_idx[l],
                                   n, l, box_w(), big_scr, big_cap, nullptr, nullptr);
            } else {
                for (int i = 0; i < n; ++i)
                    hipLaunchKernelGGL(k_amr_restrict<Real>, dim3(1), dim3(kPatchThreads), 0, stream, dev, W, d_edge_ids[l] + i,
                                       d_edge_idx[l] + 4 * i, 1, l, box_w(), big_scr, big_cap, nullptr, nullptr);
            }
            HCHK(hipGetLastError());
        }
    }

    // same_level false: the batched passes read same-level neighbours through
    // gsrc, so their slab copies are left out (the detail pass, which reads
    // every slab, fills them all)
    void fill_ghosts(bool same_level = true) {
        const int n = (int)ghost.size();
        if (!n) return;
        if (batched) {
            const int nc = same_level ? (int)ghost_copy.size() : n_bnd_copy, np = (int)ghost_pred.size();
            if (nc) {
                const int per = ng * (bcells / B);
                const long long ntot = (long long)nc * per;
                const unsigned g = (unsigned)std::min<long long>((ntot + kGhostCopyThreads - 1) / kGhostCopyThreads, 65536);
                hipLaunchKernelGGL(k_amr_ghost_copy<Real>, dim3(g), dim3(kGhostCopyThreads), 0, stream, dev, d_ghost_copy, ntot,
                                   per);
                HCHK(hipGetLastError());
            }
            if (np && ghost_lds) {
                hipLaunchKernelGGL(k_amr_ghost_lds<Real>, dim3((unsigned)std::min((np + 7) / 8 * 8, 16384)), dim3(kGhostThreads), ghost_lds,
                                   stream, dev, W, d_ghost_pred, np, ghost_capA, ghost_capB);
                HCHK(hipGetLastError());
            } else if (np) {
                const unsigned g = (unsigned)std::min(np, small_wg);
                hipLaunchKernelGGL(k_amr_ghost<Real>, dim3(g), dim3(kGhostThreads), 0, stream, dev, W, d_ghost_pred, np, small_scr,
                                   small_cap);
                HCHK(hipGetLastError());
            }
        } else {
            // one launch per leaf (its 2*D slabs)
            const int per = 2 * ndim;
            for (int i = 0; i < n; i += per) {
                hipLaunchKernelGGL(k_amr_ghost<Real>, dim3((unsigned)per), dim3(kGhostThreads), 0, stream, dev, W, d_ghost + i, per,
                                   small_scr, small_cap);
                HCHK(hipGetLastError());
            }
        }
    }

    int shape[3] = {0, 0, 0};  // pass shape per axis (amr_shape)

    hipError_t launch_pass(int d, const AmrPassArgsT<Real>& pa, unsigned grid) {
        return precise ? em::precise::launch_amr_pass(d, mode, B, shape[d], pa, grid, stream)
                       : em::fast::launch_amr_pass(d, mode, B, shape[d], pa, grid, stream);
    }

    void passes(int stage_dt_host, double stage_dt, int only_axis = -1) {
        int last = -1;
        for (int d = 0; d < 3; ++d)
            if (act[d]) last = d;
        bool first = true;
        for (int d = 0; d < 3; ++d) {
            if (!act[d]) continue;
            if (only_axis >= 0 && d != only_axis) continue;
            const int P = amr_shape(B, shape[d]).P;
            AmrPassArgsT<Real> pa{};
            pa.a = dev;
            pa.leaves = d_leaves;
            pa.leaf_level = d_leaf_level;
            pa.ngroup = (bcells / bd[d] + P - 1) / P;
            pa.first = first ? 1 : 0;
            pa.dfast = d == last ? 1 : 0;
            pa.ndim_pos = positivity ? (double)ndim : 0.0;
            pa.tol = Real(tol);
            pa.stage_dt_host = stage_dt_host;
            pa.stage_dt = stage_dt;
            pa.gsrc = batched ? d_gsrc : nullptr;
            const int nl = (int)leaves.size();
            if (batched) {
                pa.leaf0 = 0;
                pa.nleaves = nl;
                const long long nw = (long long)nl * pa.ngroup;
                const unsigned grid = (unsigned)std::min<long long>(nw, 256ll * 8);
                HCHK(launch_pass(d, pa, grid));
            } else {
                for (int li = 0; li < nl; ++li) {
                    pa.leaf0 = li;
                    pa.nleaves = 1;
                    HCHK(launch_pass(d, pa, (unsigned)pa.ngroup));
                }
            }
            first = false;
        }
    }

    void refluxes() {
        if (batched) {
            const int ng = (int)reflux_grp.size() - 1;
            if (ng > 0) {
                hipLaunchKernelGGL(k_amr_reflux_cells<Real>, dim3((unsigned)std::min((ng + 127) / 128, 4096)), dim3(128), 0,
                                   stream, dev, d_reflux_cells, d_reflux_grp, ng);
                HCHK(hipGetLastError());
            }
            return;
        }
        for (int d = 0; d < 3; ++d) {
            if (!act[d]) continue;
            for (int s = 0; s < 2; ++s) {
                const int n = (int)reflux[d][s].size();
                if (!n) continue;
                hipLaunchKernelGGL(k_amr_reflux<Real>, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, stream, dev,
                                   d_reflux[d][s], n, d, s);
                HCHK(hipGetLastError());
            }
        }
    }

    // compute_rhs (amr.rs:1302-1323): restriction, ghosts, fluxes, reflux
    void compute_rhs(int stage_dt_host, double stage_dt) {
        restrict_all();
        fill_ghosts(false);
        passes(stage_dt_host, stage_dt);
        refluxes();
    }

    // stage == 0: copy U to U0; 1..3: SSP-RK3 update
    void rk(int stage) {
        const int n = (int)leaves.size();
        if (batched) {
            const unsigned g = (unsigned)std::min<long long>(((long long)n * bcells + 255) / 256, 8192);
            if (stage == 0)
                hipLaunchKernelGGL(k_amr_copy_u0<Real>, dim3(g), dim3(256), 0, stream, dev, d_leaves, n);
            else
                hipLaunchKernelGGL(k_amr_rk<Real>, dim3(g), dim3(256), 0, stream, dev, d_leaves, n, stage);
        } else {
            const unsigned g = (unsigned)std::max(1, (bcells + 255) / 256);
            for (int li = 0; li < n; ++li) {
                if (stage == 0)
                    hipLaunchKernelGGL(k_amr_copy_u0<Real>, dim3(g), dim3(256), 0, stream, dev, d_leaves + li, 1);
                else
                    hipLaunchKernelGGL(k_amr_rk<Real>, dim3(g), dim3(256), 0, stream, dev, d_leaves + li, 1, stage);
            }
        }
        HCHK(hipGetLastError());
    }

    // Per-phase time (ms) as GpuAmr::prof (compressible-cuda/src/amr.rs:
    // 110-113): device time of restriction, ghost slabs, flux passes,
    // refluxing + RK update (events recorded only while profiling), and host
    // wall time of regrids.
    bool profile = false;
    double prof_ms[5] = {0, 0, 0, 0, 0};
    std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> prof_ev;
    size_t prof_used = 0;

    template <class F>
    void phase(int k, F&& f) {
        if (!profile) {
            f();
            return;
        }
        if (prof_used == prof_ev.size()) {
            hipEvent_t a, b;
            HCHK(hipEventCreate(&a));
            HCHK(hipEventCreate(&b));
            prof_ev.emplace_back(k, a, b);
        }
        auto& e = prof_ev[prof_used++];
        std::get<0>(e) = k;
        HCHK(hipEventRecord(std::get<1>(e), stream));
        f();
        HCHK(hipEventRecord(std::get<2>(e), stream));
    }

    void collect_profile() {
        if (!prof_used) return;
        HCHK(hipStreamSynchronize(stream));
        for (size_t i = 0; i < prof_used; ++i) {
            float ms = 0.f;
            HCHK(hipEventElapsedTime(&ms, std::get<1>(prof_ev[i]), std::get<2>(prof_ev[i])));
            prof_ms[std::get<0>(prof_ev[i])] += ms;
        }
        prof_used = 0;
    }

    void step_kernels() {
        phase(3, [&] { rk(0); });
        for (int s = 1; s <= 3; ++s) {
            phase(0, [&] { restrict_all(); });
            phase(1, [&] { fill_ghosts(false); });
            phase(2, [&] { passes(0, 0.0); });
            phase(3, [&] {
                refluxes();
                rk(s);
            });
        }
        hipLaunchKernelGGL(k_amr_advance, dim3(1), dim3(1), 0, stream, sc);
        HCHK(hipGetLastError());
    }

    // Amr::step (amr.rs:1470-1512): global SSP-RK3 step, regrid when due
    void step(double dt) {
        tables();
        hipLaunchKernelGGL(k_amr_set_run, dim3(1), dim3(1), 0, stream, sc);
        hipLaunchKernelGGL(k_amr_set_dt, dim3(1), dim3(1), 0, stream, sc, dt);
        step_kernels();
        collect_profile();
        t += dt;
        steps += 1;
        check_err();
        regrid_if_due();
    }

    void check_err() {
        std::vector<DevScalars> h(1);
        HCHK(hipMemcpyAsync(h.data(), sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
        HCHK(hipStreamSynchronize(stream));
        if (h[0].err & 6)
            throw AmrError(EM_ERR_STATE, "ghost cells could not be filled (hierarchy not 2:1 balanced?)");
        if (h[0].err & 8) throw AmrError(EM_ERR_STATE, "restriction patch outside the sub-box kernel's range");
        if (h[0].err & 1) throw AmrError(EM_ERR_COLLAPSE, "time step collapsed");
    }

    // One CFL step with dt computed on the device (rate, dt, RK stages).
    void device_step() {
        rate_kernels();
        hipLaunchKernelGGL(k_amr_dt_from_rate, dim3(1), dim3(1), 0, stream, sc, cfl);
        step_kernels();
    }

    // The launch sequence of a device step is fixed between hierarchy edits,
    // so batched stepping replays it as one HIP graph, captured again when
    // the tables or the pool change (kernel arguments are captured by value).
    hipGraphExec_t gexec = nullptr;
    long long graph_version = -1;
    long long version = 0;

    // Times every pass shape of every active axis on the residual of the
    // current hierarchy (ghosts filled, limiter on with the CFL dt) and keeps
    // the fastest; the state is untouched and the counters are restored.
    // Shapes give identical results: a cell's flux differences and
    // accumulation are the same in any pencil grouping.
    void autotune(int repeats, int32_t out[3]) {
        tables();
        const double dt = dt_cfl();
        std::vector<DevScalars> saved(1);
        HCHK(hipMemcpyAsync(saved.data(), sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
        hipLaunchKernelGGL(k_amr_set_run, dim3(1), dim3(1), 0, stream, sc);
        HCHK(hipGetLastError());
        restrict_all();
        fill_ghosts();
        hipEvent_t e0, e1;
        HCHK(hipEventCreate(&e0));
        HCHK(hipEventCreate(&e1));
        for (int d = 0; d < 3; ++d) {
            if (!act[d]) continue;
            double best = 1e300;
            int bs = shape[d];
            for (int c = 0; c < amr_num_shapes(B); ++c) {
                shape[d] = c;
                passes(1, std::isfinite(dt) ? dt : 0.0, d);  // warm-up
                HCHK(hipEventRecord(e0, stream));
                for (int r = 0; r < repeats; ++r) passes(1, std::isfinite(dt) ? dt : 0.0, d);
                HCHK(hipEventRecord(e1, stream));
                HCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HCHK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    bs = c;
                }
            }
            shape[d] = bs;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HCHK(hipMemcpyAsync(sc, saved.data(), sizeof(DevScalars), hipMemcpyHostToDevice, stream));
        HCHK(hipStreamSynchronize(stream));
        ++version;
        if (out)
            for (int d = 0; d < 3; ++d) out[d] = shape[d];
    }

    void graph_step() {
        if (!gexec || graph_version != version) {
            hipGraph_t g;
            HCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            device_step();
            HCHK(hipStreamEndCapture(stream, &g));
            // a new hierarchy usually records the same launches with other
            // arguments: update the executable graph in place (instantiating
            // costs milliseconds), else instantiate anew
            hipError_t e = hipSuccess;
            bool updated = false;
            if (gexec) {
                HCHK(hipStreamSynchronize(stream));  // no launch of it in flight
                hipGraphNode_t bad = nullptr;
                hipGraphExecUpdateResult res = hipGraphExecUpdateError;
                updated = hipGraphExecUpdate(gexec, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
                if (!updated) {
                    (void)hipGetLastError();
                    (void)hipGraphExecDestroy(gexec);
                    gexec = nullptr;
                }
            }
            if (!updated) e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HCHK(e);
            graph_version = version;
        }
        HCHK(hipGraphLaunch(gexec, stream));
    }

    bool use_graph = true;

    void run_steps(int64_t nsteps) {
        hipLaunchKernelGGL(k_amr_set_run, dim3(1), dim3(1), 0, stream, sc);
        for (int64_t q = 0; q < nsteps; ++q) {
            tables();
            if (batched && use_graph && !profile)
                graph_step();
            else
                device_step();
            collect_profile();
            steps += 1;
            if (steps >= next_regrid) {
                t = read_t();
                check_err();
                regrid_if_due();
            }
        }
        t = read_t();
        check_err();
    }

    // GpuAmr::run_to (compressible-cuda/src/amr.rs:807-819)
    int64_t run_to(double t_end) {
        int64_t n = 0;
        while (t < t_end * (1.0 - 1e-14)) {
            const double dt = std::fmin(dt_cfl(), t_end - t);
            if (!(std::isfinite(dt) && dt > 0.0)) throw AmrError(EM_ERR_COLLAPSE, "time step collapsed");
            step(dt);
            ++n;
        }
        return n;
    }

    double read_t() {
        std::vector<DevScalars> h(1);
        HCHK(hipMemcpyAsync(h.data(), sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
        HCHK(hipStreamSynchronize(stream));
        return h[0].t;
    }

    void compute_rate() {
        tables();
        rate_kernels();
    }

    void rate_kernels() {
        hipLaunchKernelGGL(k_amr_zero_rate, dim3(1), dim3(kSlots), 0, stream, sc);
        const int n = (int)leaves.size();
        const unsigned g = (unsigned)std::min<long long>(((long long)n * bcells + 255) / 256, 2048);
        hipLaunchKernelGGL(k_amr_rate<Real>, dim3(g), dim3(256), 0, stream, dev, d_leaves, d_leaf_level, n);
        HCHK(hipGetLastError());
    }

    double dt_cfl() {
        compute_rate();
        std::vector<DevScalars> h(1);
        HCHK(hipMemcpyAsync(h.data(), sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
        HCHK(hipStreamSynchronize(stream));
        unsigned long long m = 0;
        for (int i = 0; i < kSlots; ++i) m = std::max(m, h[0].rate[i][0]);
        return cfl / __builtin_bit_cast(double, m);
    }

    std::vector<double> details() {
        tables();
        restrict_all();
        fill_ghosts();
        const int n = (int)leaves.size();
        if (d_det_n < n) {
            if (d_det) HCHK(hipFree(d_det));
            HCHK(hipMalloc(&d_det, sizeof(double) * std::max(n, 1)));
            d_det_n = n;
        }
        std::vector<double> det(blocks.size(), 0.0);
        if (n) {
            if (bcells <= kDetCells && tmax <= kDetFace)
                hipLaunchKernelGGL(k_amr_details_lds<Real>, dim3((unsigned)std::min(n, 2048)), dim3(256), 0, stream, dev, W,
                                   d_leaves, n, d_det);
            else
                hipLaunchKernelGGL(k_amr_details<Real>, dim3((unsigned)std::min(n, 2048)), dim3(256), 0, stream, dev, W, d_leaves,
                                   n, d_det);
            HCHK(hipGetLastError());
            std::vector<double> h(n);
            HCHK(hipMemcpyAsync(h.data(), d_det, sizeof(double) * n, hipMemcpyDeviceToHost, stream));
            HCHK(hipStreamSynchronize(stream));
            for (int i = 0; i < n; ++i) det[leaves[i]] = h[i];
        }
        check_err();
        return det;
    }

    void regrid() {  // amr.rs:1039-1046
        std::vector<int8_t> f = flags_from_details(details());
        std::set<Key> prot = regrid_refine(f);
        // the coarsening reads the details of the refined hierarchy; with no
        // refinement (tables still clean) tree and data are unchanged, and so
        // are the details and flags just computed
        if (tables_dirty) {
            tables();
            f = flags_from_details(details());
        }
        regrid_coarsen(f, prot);
        tables();
    }

    bool regrid_if_due() {  // amr.rs:1525-1537
        if (steps < next_regrid) return false;
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<std::tuple<int, int64_t, int64_t, int64_t, bool>> before, after;
        for (auto& b : blocks) before.emplace_back(b.level, b.idx[0], b.idx[1], b.idx[2], b.leaf);
        regrid();
        for (auto& b : blocks) after.emplace_back(b.level, b.idx[0], b.idx[1], b.idx[2], b.leaf);
        const bool changed = before != after;
        regrid_interval = changed ? regrid_every : std::min<int64_t>(2 * regrid_interval, 2 * (int64_t)regrid_every);
        next_regrid = steps + regrid_interval;
        if (profile) {
            HCHK(hipStreamSynchronize(stream));
            prof_ms[4] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        return true;
    }

    // Amr::new (amr.rs:289-343): level-0 blocks from the initial condition,
    // refined where the details exceed eps, children re-evaluated from it
    void build_from_case(InitFn f, const double* prm) {
        MapEdits ed{};
        std::vector<int> slots;
        std::vector<double> data;
        for (int64_t i = 0; i < nbl(0, 0); ++i)
            for (int64_t j = 0; j < nbl(0, 1); ++j)
                for (int64_t k = 0; k < nbl(0, 2); ++k) {
                    const int64_t idx[3] = {i, j, k};
                    const int id = add_block(0, idx, ed);
                    slots.push_back(id);
                    const size_t o = data.size();
                    data.resize(o + (size_t)5 * bcells);
                    init_data(0, idx, f, prm, data.data() + o);
                }
        flush_edits(ed);
        upload_blocks(slots, data.data());
        tables_dirty = true;
        for (int it = 0; it < levels; ++it) {
            tables();
            std::vector<int8_t> fl = flags_from_details(details());
            std::vector<Key> keys;
            for (size_t i = 0; i < fl.size(); ++i)
                if (fl[i] > 0 && blocks[i].leaf) keys.push_back(Key{blocks[i].level, {blocks[i].idx[0], blocks[i].idx[1], blocks[i].idx[2]}});
            if (keys.empty()) break;
            for (const Key& k : keys) {
                const int id = get(k.l, k.i);
                if (id >= 0) refine_block(id, f, prm);
            }
        }
        tables();
    }

    void build_from_blocks(int64_t n, const int32_t* level, const int64_t* idx, const uint8_t* leaf, const double* u) {
        MapEdits ed{};
        std::vector<int> slots;
        for (int64_t i = 0; i < n; ++i) {
            if (level[i] < 0 || level[i] >= levels) throw AmrError(EM_ERR_ARG, "block level out of range");
            const int id = add_block(level[i], idx + 3 * i, ed);
            blocks[id].leaf = leaf[i] != 0;
            slots.push_back(id);
        }
        flush_edits(ed);
        upload_blocks(slots, u);
        tables_dirty = true;
        tables();
    }

    void download(int32_t* level, int64_t* idx, uint8_t* leaf, double* u) {
        const long long n = (long long)blocks.size();
        for (long long i = 0; i < n; ++i) {
            if (level) level[i] = blocks[i].level;
            if (idx)
                for (int d = 0; d < 3; ++d) idx[3 * i + d] = blocks[i].idx[d];
            if (leaf) leaf[i] = blocks[i].leaf ? 1 : 0;
        }
        if (u) gather(dev.U, u, n);
    }

    void gather(const Real* src, double* out, long long n) {
        if (n == 0) return;
        double* stage;
        HCHK(hipMalloc(&stage, sizeof(double) * 5 * bcells * n));
        const unsigned g = (unsigned)std::min<long long>((n * bcells + 255) / 256, 4096);
        hipLaunchKernelGGL(k_amr_gather<Real>, dim3(g), dim3(256), 0, stream, dev, stage, src, n);
        HCHK(hipGetLastError());
        HCHK(hipMemcpyAsync(out, stage, sizeof(double) * 5 * bcells * n, hipMemcpyDeviceToHost, stream));
        HCHK(hipStreamSynchronize(stream));
        (void)hipFree(stage);
    }

    void upload(const double* u) {
        std::vector<int> slots(blocks.size());
        for (size_t i = 0; i < blocks.size(); ++i) slots[i] = (int)i;
        upload_blocks(slots, u);
    }

    // residual of every leaf for a given stage dt (non-leaves zero)
    void rhs(double stage_dt, double* out, uint64_t stats[4]) {
        tables();
        std::vector<DevScalars> before(1), after(1);
        HCHK(hipMemcpyAsync(before.data(), sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
        hipLaunchKernelGGL(k_amr_set_run, dim3(1), dim3(1), 0, stream, sc);
        compute_rhs(1, stage_dt);
        HCHK(hipMemcpyAsync(after.data(), sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
        const long long n = (long long)blocks.size();
        gather(dev.R, out, n);
        for (long long i = 0; i < n; ++i)
            if (!blocks[i].leaf) std::fill(out + i * 5 * bcells, out + (i + 1) * 5 * bcells, 0.0);
        check_err();
        if (stats) {
            uint64_t faces = 0;
            for (int d = 0; d < 3; ++d)
                if (act[d]) faces += (uint64_t)(bd[d] + 1) * (bcells / bd[d]);
            stats[0] = faces * leaves.size();
            for (int q = 0; q < 3; ++q) {
                uint64_t s = 0;
                for (int i = 0; i < kSlots; ++i) s += after[0].counters[q][i][0] - before[0].counters[q][i][0];
                stats[1 + q] = s;
            }
        }
    }

    void counters(uint64_t out[3]) {
        std::vector<DevScalars> h(1);
        HCHK(hipMemcpyAsync(h.data(), sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
        HCHK(hipStreamSynchronize(stream));
        for (int q = 0; q < 3; ++q) {
            out[q] = 0;
            for (int i = 0; i < kSlots; ++i) out[q] += h[0].counters[q][i][0];
        }
    }
};
using Amr = AmrT<double>;

// ------------------------------------------------------------------ cases
// Primitive initial conditions of the AMR workloads (ideal gas, gamma 1.4):
//  0 four-quadrant 2-D Riemann problem, split (p[0], p[1])
//  1 smooth wave 1 + 0.1 sin(2 pi x) cos(2 pi y), velocity (0.3, 0.2, 0.1), p 1
//  2 shock-bubble: Mach p[0] shock at x = p[1] moving into (rho 1, p 1), sphere
//    centre (p[2], p[3], p[4]), radius p[5], density p[6], p 1
//  3 isentropic vortex (examples/amr_validate.rs:44-51)
//  4 state at rest rho = 1 + sum c_d x_d^5/5, p = 1 + sum c_d x_d^4/4, c = p[0..2]
static void amr_case(const double x[3], const double* p, double w[5]) {
    const int id = (int)p[15];
    w[1] = w[2] = w[3] = 0.0;
    if (id == 0) {
        const bool e = x[0] > p[0], n = x[1] > p[1];
        if (e && n) {
            w[0] = 1.5; w[4] = 1.5;
        } else if (!e && n) {
            w[0] = 0.5323; w[1] = 1.206; w[4] = 0.3;
        } else if (!e && !n) {
            w[0] = 0.138; w[1] = 1.206; w[2] = 1.206; w[4] = 0.029;
        } else {
            w[0] = 0.5323; w[2] = 1.206; w[4] = 0.3;
        }
    } else if (id == 1) {
        const double TAU = 6.283185307179586;
        w[0] = 1.0 + 0.1 * std::sin(TAU * x[0]) * std::cos(TAU * x[1]);
        w[1] = 0.3; w[2] = 0.2; w[3] = 0.1; w[4] = 1.0;
    } else if (id == 2) {
        const double g = 1.4, M = p[0];
        double r2 = 0.0;
        for (int d = 0; d < 3; ++d) r2 += (x[d] - p[2 + d]) * (x[d] - p[2 + d]);
        if (x[0] < p[1]) {
            const double rho2 = (g + 1.0) * M * M / ((g - 1.0) * M * M + 2.0);
            const double p2 = 1.0 + 2.0 * g / (g + 1.0) * (M * M - 1.0);
            w[0] = rho2;
            w[1] = M * std::sqrt(g) * (1.0 - 1.0 / rho2);
            w[4] = p2;
        } else if (r2 < p[5] * p[5]) {
            w[0] = p[6]; w[4] = 1.0;
        } else {
            w[0] = 1.0; w[4] = 1.0;
        }
    } else if (id == 3) {
        const double gam = 1.4, beta = 5.0, pi = 3.141592653589793;
        const double r2 = x[0] * x[0] + x[1] * x[1];
        const double f = beta / (2.0 * pi) * std::exp(0.5 * (1.0 - r2));
        const double tt = 1.0 - (gam - 1.0) * beta * beta / (8.0 * gam * pi * pi) * std::exp(1.0 - r2);
        const double rho = std::pow(tt, 1.0 / (gam - 1.0));
        w[0] = rho; w[1] = 1.0 - f * x[1]; w[2] = 1.0 + f * x[0]; w[4] = rho * tt;
    } else {
        double rho = 1.0, pr = 1.0;
        for (int d = 0; d < 3; ++d) {
            const double x2 = x[d] * x[d];
            rho += p[d] * x2 * x2 * x[d] / 5.0;
            pr += p[d] * x2 * x2 / 4.0;
        }
        w[0] = rho; w[4] = pr;
    }
}

// Several hierarchies stepped together, each with its own block size (8,
// 16 or 32 cells a side), dt, tables and regrid schedule: per stage, the
// restriction, ghost, reflux and RK launches of every member, and one
// persistent flux-pass launch per axis over the leaves of all members
// (amr_multi).  Members must share the flux mode, arithmetic variant and
// active axes, and run batched.  The group borrows the members (it does not
// own them) and runs them on its own stream; a member's results are those
// of the member stepped alone, bit for bit.
struct AmrGroup {
    std::vector<Amr*> m;
    hipStream_t stream = nullptr;
    AmrPassArgs* d_args[3] = {nullptr, nullptr, nullptr};
    long long* d_prefix[3] = {nullptr, nullptr, nullptr};
    long long total[3] = {0, 0, 0};
    int mode = 0, precise = 1;
    int ncu = 256;  // compute units of the device
    // host images of the uploaded arguments (kept alive for the async copies;
    // an unchanged step skips the upload)
    std::vector<AmrPassArgs> h_args[3];
    std::vector<long long> h_prefix[3];

    explicit AmrGroup(const std::vector<Amr*>& members) : m(members) {
        if (m.empty()) throw AmrError(EM_ERR_ARG, "an AMR group needs at least one member");
        mode = m[0]->mode;
        precise = m[0]->precise;
        for (Amr* a : m) {
            if (!a) throw AmrError(EM_ERR_ARG, "null AMR member");
            if (a->mode != mode || a->precise != precise)
                throw AmrError(EM_ERR_ARG, "AMR group members must share the flux mode and arithmetic variant");
            for (int d = 0; d < 3; ++d)
                if (a->act[d] != m[0]->act[d])
                    throw AmrError(EM_ERR_ARG, "AMR group members must share their active axes");
            if (!a->batched) throw AmrError(EM_ERR_ARG, "AMR group members must run batched");
        }
        hchk(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        int dev = 0;
        hchk(hipGetDevice(&dev), "hipGetDevice");
        hchk(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        const int n = (int)m.size();
        for (int d = 0; d < 3; ++d) {
            hchk(hipMalloc(&d_args[d], sizeof(AmrPassArgs) * n), "hipMalloc");
            hchk(hipMalloc(&d_prefix[d], sizeof(long long) * (n + 1)), "hipMalloc");
        }
    }
    ~AmrGroup() {
        for (int d = 0; d < 3; ++d) {
            if (d_args[d]) (void)hipFree(d_args[d]);
            if (d_prefix[d]) (void)hipFree(d_prefix[d]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    // the members' work goes to the group's stream while the group runs
    struct Borrow {
        AmrGroup& g;
        std::vector<hipStream_t> own;
        explicit Borrow(AmrGroup& g_) : g(g_) {
            for (Amr* a : g.m) {
                hchk(hipStreamSynchronize(a->stream), "hipStreamSynchronize");
                own.push_back(a->stream);
                a->stream = g.stream;
            }
        }
        ~Borrow() {
            (void)hipStreamSynchronize(g.stream);
            for (size_t i = 0; i < g.m.size(); ++i) g.m[i]->stream = own[i];
        }
    };

    // per-axis pass arguments of every member (their tables are fixed for a step)
    void build_args() {
        const int n = (int)m.size();
        int last = -1;
        for (int d = 0; d < 3; ++d)
            if (m[0]->act[d]) last = d;
        bool first = true;
        for (int d = 0; d < 3; ++d) {
            if (!m[0]->act[d]) continue;
            std::vector<AmrPassArgs> args(n);
            std::vector<long long> prefix(n + 1, 0);
            for (int i = 0; i < n; ++i) {
                Amr& a = *m[i];
                const int P = amr_multi_pencils(a.B);
                AmrPassArgs& pa = args[i];
                pa.a = a.dev;
                pa.leaves = a.d_leaves;
                pa.leaf_level = a.d_leaf_level;
                pa.leaf0 = 0;
                pa.nleaves = (int)a.leaves.size();
                pa.ngroup = (a.bcells / a.bd[d] + P - 1) / P;
                pa.first = first ? 1 : 0;
                pa.dfast = d == last ? 1 : 0;
                pa.ndim_pos = a.positivity ? (double)a.ndim : 0.0;
                pa.tol = a.tol;
                pa.stage_dt_host = 0;
                pa.stage_dt = 0.0;
                pa.gsrc = a.d_gsrc;
                prefix[i + 1] = prefix[i] + (long long)pa.nleaves * pa.ngroup;
            }
            total[d] = prefix[n];
            first = false;
            const bool same = h_args[d].size() == args.size() &&
                              std::memcmp(h_args[d].data(), args.data(), sizeof(AmrPassArgs) * n) == 0 &&
                              h_prefix[d] == prefix;
            if (same) continue;
            // the previous step's kernels may still read the old arguments
            hchk(hipStreamSynchronize(stream), "hipStreamSynchronize");
            h_args[d] = std::move(args);
            h_prefix[d] = std::move(prefix);
            hchk(hipMemcpyAsync(d_args[d], h_args[d].data(), sizeof(AmrPassArgs) * n, hipMemcpyHostToDevice, stream),
                 "hipMemcpyAsync");
            hchk(hipMemcpyAsync(d_prefix[d], h_prefix[d].data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice,
                                stream),
                 "hipMemcpyAsync");
        }
    }

    void passes() {
        for (int d = 0; d < 3; ++d) {
            if (!m[0]->act[d] || total[d] == 0) continue;
            const AmrMultiArgs ma{d_args[d], d_prefix[d], (int)m.size(), total[d]};
            // three full rounds of resident workgroups (three per CU): a
            // partial last round (2,048 workgroups were 2.67 rounds) left
            // CUs idle at the end of each pass, 2% of the group's step
            const unsigned grid = (unsigned)std::min<long long>(total[d], 9ll * ncu);
            hchk(precise ? em::precise::launch_amr_multi(d, mode, ma, grid, stream)
                         : em::fast::launch_amr_multi(d, mode, ma, grid, stream),
                 "amr_multi");
        }
    }

    // nsteps CFL steps of every member (dt of each member from its own
    // state, on the device), with each member's regrids when due.
    void run_steps(int64_t nsteps) {
        Borrow b(*this);
        for (Amr* a : m) hipLaunchKernelGGL(k_amr_set_run, dim3(1), dim3(1), 0, stream, a->sc);
        for (int64_t q = 0; q < nsteps; ++q) {
            for (Amr* a : m) a->tables();
            build_args();
            for (Amr* a : m) {
                a->rate_kernels();
                hipLaunchKernelGGL(k_amr_dt_from_rate, dim3(1), dim3(1), 0, stream, a->sc, a->cfl);
                a->rk(0);
            }
            for (int s = 1; s <= 3; ++s) {
                for (Amr* a : m) {
                    a->restrict_all();
                    a->fill_ghosts(false);
                }
                passes();
                for (Amr* a : m) {
                    a->refluxes();
                    a->rk(s);
                }
            }
            for (Amr* a : m) {
                hipLaunchKernelGGL(k_amr_advance, dim3(1), dim3(1), 0, stream, a->sc);
                a->steps += 1;
            }
            hchk(hipGetLastError(), "AMR group step");
            for (Amr* a : m)
                if (a->steps >= a->next_regrid) {
                    a->t = a->read_t();
                    a->check_err();
                    a->regrid_if_due();
                }
        }
        for (Amr* a : m) {
            a->t = a->read_t();
            a->check_err();
        }
    }
};

thread_local std::string g_last_error_amr;

}  // namespace em

// ------------------------------------------------------------------ C ABI

struct em_amr {
    em::AmrT<double>* a;   // FP64 hierarchy (null for an FP32 handle)
    em::AmrT<float>* a32;  // FP32 hierarchy (null for an FP64 handle)
};

struct em_amr_group {
    em::AmrGroup* g;
};

template <class F>
static int aguard(F&& f) {
    try {
        f();
        em::g_last_error_amr.clear();
        return EM_OK;
    } catch (const em::AmrError& e) {
        em::g_last_error_amr = e.what();
        return e.code;
    } catch (const std::exception& e) {
        em::g_last_error_amr = e.what();
        return EM_ERR_STATE;
    }
}

#define ANEED(h)                                           \
    if (!(h) || (!(h)->a && !(h)->a32)) {                  \
        em::g_last_error_amr = "null AMR handle";          \
        return EM_ERR_ARG;                                 \
    }

// f(hierarchy) on the handle's hierarchy, whichever real type it holds
template <class F>
static void aeach(em_amr* h, F&& f) {
    if (h->a)
        f(*h->a);
    else
        f(*h->a32);
}
template <class F>
static int aon(em_amr* h, F&& f) {
    return aguard([&] { aeach(h, f); });
}

template <class Real, class Fill>
static int amr_make(em_amr** out, const int64_t base[3], int32_t b, int32_t levels, int32_t ng, double eps,
                    int32_t regrid_every, const double lo[3], const double hi[3], const double eos[4],
                    const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity, double cfl,
                    double tol, Fill&& fill) {
    if (!out || !base || !lo || !hi || !eos || !bc_kind) {
        em::g_last_error_amr = "null argument";
        return EM_ERR_ARG;
    }
    return aguard([&] {
        auto* a = new em::AmrT<Real>(base, b, levels, ng, eps, regrid_every, lo, hi, eos, bc_kind, bc_state, mode,
                                     positivity, cfl, tol);
        try {
            fill(*a);
        } catch (...) {
            delete a;
            throw;
        }
        if constexpr (sizeof(Real) == 8)
            *out = new em_amr{a, nullptr};
        else
            *out = new em_amr{nullptr, a};
    });
}

template <class Real>
static int amr_create_t(em_amr** out, const int64_t base[3], int32_t b, int32_t levels, int32_t ng, double eps,
                        int32_t regrid_every, const double lo[3], const double hi[3], const double eos[4],
                        const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity,
                        double cfl, double tol, int64_t nblocks, const int32_t* level, const int64_t* idx,
                        const uint8_t* leaf, const double* u) {
    if (nblocks < 1 || !level || !idx || !leaf || !u) {
        em::g_last_error_amr = "block arrays missing";
        return EM_ERR_ARG;
    }
    return amr_make<Real>(out, base, b, levels, ng, eps, regrid_every, lo, hi, eos, bc_kind, bc_state, mode,
                          positivity, cfl, tol, [&](auto& a) { a.build_from_blocks(nblocks, level, idx, leaf, u); });
}

template <class Real>
static int amr_create_case_t(em_amr** out, const int64_t base[3], int32_t b, int32_t levels, int32_t ng, double eps,
                             int32_t regrid_every, const double lo[3], const double hi[3], const double eos[4],
                             const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity,
                             double cfl, double tol, int32_t case_id, const double* params) {
    if (case_id < 0 || case_id > 4) {
        em::g_last_error_amr = "unknown AMR case";
        return EM_ERR_ARG;
    }
    double prm[16] = {0};
    if (params)
        for (int i = 0; i < 15; ++i) prm[i] = params[i];
    prm[15] = (double)case_id;
    return amr_make<Real>(out, base, b, levels, ng, eps, regrid_every, lo, hi, eos, bc_kind, bc_state, mode,
                          positivity, cfl, tol, [&](auto& a) { a.build_from_case(em::amr_case, prm); });
}

extern "C" {

const char* em_amr_last_error(void) { return em::g_last_error_amr.c_str(); }

int em_amr_create(em_amr** out, const int64_t base[3], int32_t b, int32_t levels, int32_t ng, double eps,
                  int32_t regrid_every, const double lo[3], const double hi[3], const double eos[4],
                  const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity, double cfl,
                  double tol, int64_t nblocks, const int32_t* level, const int64_t* idx, const uint8_t* leaf,
                  const double* u) {
    return amr_create_t<double>(out, base, b, levels, ng, eps, regrid_every, lo, hi, eos, bc_kind, bc_state, mode,
                                positivity, cfl, tol, nblocks, level, idx, leaf, u);
}

int em_amr_create_f32(em_amr** out, const int64_t base[3], int32_t b, int32_t levels, int32_t ng, double eps,
                      int32_t regrid_every, const double lo[3], const double hi[3], const double eos[4],
                      const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity,
                      double cfl, double tol, int64_t nblocks, const int32_t* level, const int64_t* idx,
                      const uint8_t* leaf, const double* u) {
    return amr_create_t<float>(out, base, b, levels, ng, eps, regrid_every, lo, hi, eos, bc_kind, bc_state, mode,
                               positivity, cfl, tol, nblocks, level, idx, leaf, u);
}

int em_amr_create_case(em_amr** out, const int64_t base[3], int32_t b, int32_t levels, int32_t ng, double eps,
                       int32_t regrid_every, const double lo[3], const double hi[3], const double eos[4],
                       const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity,
                       double cfl, double tol, int32_t case_id, const double* params) {
    return amr_create_case_t<double>(out, base, b, levels, ng, eps, regrid_every, lo, hi, eos, bc_kind, bc_state,
                                     mode, positivity, cfl, tol, case_id, params);
}

int em_amr_create_case_f32(em_amr** out, const int64_t base[3], int32_t b, int32_t levels, int32_t ng, double eps,
                           int32_t regrid_every, const double lo[3], const double hi[3], const double eos[4],
                           const int32_t bc_kind[6], const double bc_state[30], int32_t mode, int32_t positivity,
                           double cfl, double tol, int32_t case_id, const double* params) {
    return amr_create_case_t<float>(out, base, b, levels, ng, eps, regrid_every, lo, hi, eos, bc_kind, bc_state,
                                    mode, positivity, cfl, tol, case_id, params);
}

int em_amr_destroy(em_amr* h) {
    if (!h) return EM_OK;
    delete h->a;
    delete h->a32;
    delete h;
    return EM_OK;
}

int em_amr_set_variant(em_amr* h, int32_t precise) {
    ANEED(h);
    aeach(h, [&](auto& a) {
        a.precise = precise ? 1 : 0;
        ++a.version;
    });
    return EM_OK;
}

int em_amr_set_batched(em_amr* h, int32_t batched) {
    ANEED(h);
    aeach(h, [&](auto& a) {
        a.batched = batched != 0;
        ++a.version;
    });
    return EM_OK;
}

int em_amr_set_shapes(em_amr* h, const int32_t shape[3]) {
    ANEED(h);
    if (!shape) {
        em::g_last_error_amr = "null argument";
        return EM_ERR_ARG;
    }
    const int bsz = h->a ? h->a->B : h->a32->B;
    for (int d = 0; d < 3; ++d)
        if (shape[d] < 0 || shape[d] >= em::amr_num_shapes(bsz)) {
            em::g_last_error_amr = "no such pass shape";
            return EM_ERR_ARG;
        }
    aeach(h, [&](auto& a) {
        for (int d = 0; d < 3; ++d) a.shape[d] = shape[d];
        ++a.version;
    });
    return EM_OK;
}

int em_amr_autotune(em_amr* h, int32_t repeats, int32_t shape_out[3]) {
    ANEED(h);
    return aon(h, [&](auto& a) { a.autotune(repeats > 0 ? repeats : 3, shape_out); });
}

int em_amr_set_graph(em_amr* h, int32_t on) {
    ANEED(h);
    aeach(h, [&](auto& a) { a.use_graph = on != 0; });
    return EM_OK;
}

int em_amr_set_profile(em_amr* h, int32_t on) {
    ANEED(h);
    aeach(h, [&](auto& a) { a.profile = on != 0; });
    return EM_OK;
}

int em_amr_profile(em_amr* h, double out[5]) {
    ANEED(h);
    aeach(h, [&](auto& a) {
        for (int i = 0; i < 5; ++i) out[i] = a.prof_ms[i];
    });
    return EM_OK;
}

int em_amr_set_next_regrid(em_amr* h, int64_t step) {
    ANEED(h);
    aeach(h, [&](auto& a) { a.next_regrid = step; });
    return EM_OK;
}

int em_amr_nblocks(em_amr* h, int64_t* n, int64_t* leaves) {
    ANEED(h);
    aeach(h, [&](auto& a) {
        if (n) *n = (int64_t)a.blocks.size();
        if (leaves) {
            int64_t k = 0;
            for (auto& b : a.blocks) k += b.leaf ? 1 : 0;
            *leaves = k;
        }
    });
    return EM_OK;
}

int em_amr_download(em_amr* h, int32_t* level, int64_t* idx, uint8_t* leaf, double* u) {
    ANEED(h);
    return aon(h, [&](auto& a) { a.download(level, idx, leaf, u); });
}

int em_amr_upload(em_amr* h, const double* u) {
    ANEED(h);
    return aon(h, [&](auto& a) { a.upload(u); });
}

int em_amr_rhs(em_amr* h, double stage_dt, double* rhs, uint64_t stats[4]) {
    ANEED(h);
    return aon(h, [&](auto& a) { a.rhs(stage_dt, rhs, stats); });
}

int em_amr_step(em_amr* h, double dt) {
    ANEED(h);
    return aon(h, [&](auto& a) { a.step(dt); });
}

int em_amr_run_steps(em_amr* h, int64_t nsteps) {
    ANEED(h);
    return aon(h, [&](auto& a) { a.run_steps(nsteps); });
}

int em_amr_run_to(em_amr* h, double t_end, int64_t* steps) {
    ANEED(h);
    return aon(h, [&](auto& a) {
        const int64_t n = a.run_to(t_end);
        if (steps) *steps = n;
    });
}

int em_amr_dt_cfl(em_amr* h, double* dt) {
    ANEED(h);
    return aon(h, [&](auto& a) { *dt = a.dt_cfl(); });
}

int em_amr_details(em_amr* h, double* det) {
    ANEED(h);
    return aon(h, [&](auto& a) {
        std::vector<double> d = a.details();
        std::copy(d.begin(), d.end(), det);
    });
}

int em_amr_regrid(em_amr* h) {
    ANEED(h);
    return aon(h, [&](auto& a) { a.regrid(); });
}

int em_amr_counters(em_amr* h, uint64_t out[3]) {
    ANEED(h);
    return aon(h, [&](auto& a) { a.counters(out); });
}

int em_amr_time(em_amr* h, double* t) {
    ANEED(h);
    aeach(h, [&](auto& a) { *t = a.t; });
    return EM_OK;
}

int em_amr_synchronize(em_amr* h) {
    ANEED(h);
    return aon(h, [&](auto& a) { em::hchk(hipStreamSynchronize(a.stream), "hipStreamSynchronize"); });
}

int em_amr_group_create(em_amr_group** out, em_amr* const* members, int32_t n) {
    if (!out || !members || n < 1) {
        em::g_last_error_amr = "null argument or empty group";
        return EM_ERR_ARG;
    }
    std::vector<em::Amr*> m;
    for (int i = 0; i < n; ++i) {
        if (!members[i] || (!members[i]->a && !members[i]->a32)) {
            em::g_last_error_amr = "null AMR handle";
            return EM_ERR_ARG;
        }
        if (!members[i]->a) {
            em::g_last_error_amr = "AMR group members must be FP64 hierarchies";
            return EM_ERR_ARG;
        }
        m.push_back(members[i]->a);
    }
    return aguard([&] { *out = new em_amr_group{new em::AmrGroup(m)}; });
}

int em_amr_group_destroy(em_amr_group* g) {
    if (!g) return EM_OK;
    delete g->g;
    delete g;
    return EM_OK;
}

int em_amr_group_run_steps(em_amr_group* g, int64_t nsteps) {
    if (!g || !g->g) {
        em::g_last_error_amr = "null AMR group handle";
        return EM_ERR_ARG;
    }
    return aguard([&] { g->g->run_steps(nsteps); });
}

}  // extern "C"
