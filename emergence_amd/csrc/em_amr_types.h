// Device descriptors of the block-AMR pool (host/device shared).
//
// All blocks (leaves and ancestors) live in one pool indexed by block id
// (the host hierarchy keeps ids in the reference's order).  Per variable q:
//   U[q][slot][cell]          conserved state, cell = (i*b1 + j)*b2 + k
//   U0, R                     step-start state and residual (leaves)
//   G[q][slot][s][layer][t]   ghost slab s = 2*axis + side; on the low
//                             side layer l is the cell at axis coordinate
//                             l - ng, on the high side at b + l
//   F[q][slot][s][t]          boundary face fluxes (for refluxing)
// where t is the transverse cell index of a face, (t0*bd[e1] + t1) over the
// two other axes e0 < e1.
#pragma once
#include <cstdint>

#include "em_device.h"
#include "em_types.h"

namespace em {

constexpr int kMaxLevels = 8;

// Real: the type of the pool data and of the arithmetic on it (double, or
// float for the FP32 variant); the device scalars (dt, time, CFL rate) stay
// double.
template <class Real>
struct AmrDevT {
    Real* U;
    Real* U0;
    Real* R;
    Real* G;
    Real* F;
    long long cap;        // pool capacity (blocks)
    long long vs;         // variable stride of U/U0/R = cap * bcells
    long long gs;         // variable stride of G = cap * 6 * ng * tmax
    long long fs;         // variable stride of F = cap * 6 * tmax
    int bd[3];            // block extent per axis (1 on inactive axes)
    int lb[3];            // log2 of bd
    int act[3];
    int ndim;
    int bcells;
    int tmax;             // largest face cell count, bcells / min active bd
    int ng;               // ghost layers held in the slabs (>= 5)
    int levels;
    int nb0[3];           // level-0 blocks per axis (1 on inactive axes)
    long long map_off[kMaxLevels];  // dense per-level block index offsets
    const int* map;       // block id at (level, block coords) or -1
    Real dx[kMaxLevels][3];
    EosT<Real> eos;
    int bc_kind[3][2];
    Real bc_state[3][2][5];
    DevScalars* sc;
};
using AmrDev = AmrDevT<double>;

// One ghost slab of one leaf: kind 0 same-level copy from block `src`,
// 1 physical boundary condition, 2 6th-order prediction from level-1.
struct GhostTask {
    int slot, axis, side, kind, src, level;
    int idx[3];
};

// One coarse boundary face corrected by refluxing (amr.rs:1327-1406):
// coarse leaf slot, cell and face index, the covering fine faces.
struct RefluxTask {
    int slot, cell, face, nfine, level;
    int fslot[4], fface[4];
    int axis, side;
};

template <class Real>
struct AmrPassArgsT {
    AmrDevT<Real> a;
    const int* leaves;      // leaf slots (block ids)
    const int* leaf_level;  // level of each leaf
    int leaf0, nleaves;     // leaves [leaf0, leaf0 + nleaves) of the list
    int ngroup;             // pencil groups per leaf along the axis
    int first;              // residual input is zero (first active axis)
    int dfast;              // the pass axis is the contiguous one in a block
    double ndim_pos;        // active dimensions if the limiter is on, else 0
    Real tol;
    int stage_dt_host;
    double stage_dt;
    // same-level source block of each (block, face) ghost slab, indexed
    // block * 6 + 2 * axis + side, -1 where the slab itself is filled; null:
    // every ghost is read from the slabs
    const int* gsrc;
};
using AmrPassArgs = AmrPassArgsT<double>;

// Several hierarchies (of any block sizes 8, 16, 32) in one batched pass:
// per-hierarchy pass arguments (leaf0 = 0, nleaves = all its leaves, ngroup
// for the group kernel's pencils per work item amr_multi_pencils(b)) and the
// prefix sums of their work items (n + 1 entries, total = prefix[n]).
struct AmrMultiArgs {
    const AmrPassArgs* args;
    const long long* prefix;
    int n;
    long long total;
};
constexpr int kAmrMultiThreads = 256;
// pencils per work item of the group kernel: (b + 5) * P midpoints fill
// about the 256 threads (208, 252, 222)
inline int amr_multi_pencils(int b) { return b == 8 ? 16 : b == 16 ? 12 : 6; }

// Shapes of the batched pass with a compiled kernel (launch_amr_b): P pencils
// per work item, NT threads; shape 0 is the default, em_amr_autotune times
// the others.
struct AmrShape {
    int P, NT;
};
inline int amr_num_shapes(int b) { return b == 8 ? 4 : 2; }
inline AmrShape amr_shape(int b, int shape) {
    static const AmrShape s8[4] = {{8, 128}, {16, 256}, {32, 512}, {64, 512}};
    static const AmrShape s16[2] = {{16, 384}, {32, 512}};
    static const AmrShape s32[2] = {{8, 320}, {16, 512}};
    return b == 8 ? s8[shape] : b == 16 ? s16[shape] : s32[shape];
}

template <class Real>
__host__ __device__ inline long long amr_u(const AmrDevT<Real>& a, int q, long long slot, long long cell) {
    return q * a.vs + slot * a.bcells + cell;
}
template <class Real>
__host__ __device__ inline long long amr_g(const AmrDevT<Real>& a, int q, long long slot, int s, int layer, int t) {
    return q * a.gs + ((slot * 6 + s) * a.ng + layer) * (long long)a.tmax + t;
}
template <class Real>
__host__ __device__ inline long long amr_f(const AmrDevT<Real>& a, int q, long long slot, int s, int t) {
    return q * a.fs + (slot * 6 + s) * (long long)a.tmax + t;
}

}  // namespace em
