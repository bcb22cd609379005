// Host/device shared descriptors of the uniform-grid axis pass.
#pragma once
#include <cstdint>

#include "em_device.h"

namespace em {

enum BcKind : int { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECT = 2, BC_DIRICHLET = 3, BC_CHARACTERISTIC = 4 };
enum Mode : int { MODE_HYBRID = 0, MODE_WENO = 1, MODE_CENTRAL = 2 };

// Device scalars of one solver.  Workgroup-level accumulators (counters and
// the CFL rate) are spread over kSlots slots on separate 128-byte lines so
// that tens of thousands of workgroups per launch do not serialise on one
// address; readers reduce over the slots.
constexpr int kSlots = 64;
struct alignas(256) DevScalars {
    double dt;                    // step size used by every kernel of a step
    double t;                     // solver time
    double t_end;                 // run_to target (run mode 1)
    int active;                   // 0 once run_to reached t_end or failed
    int err;                      // bit 0: dt collapsed
    long long steps;              // steps taken by device-driven runs
    // max CFL rate of the state as IEEE bits of a double >= 0, per slot
    unsigned long long rate[kSlots][16];
    // [shock-path faces, fallbacks, limited faces] per slot
    unsigned long long counters[3][kSlots][16];
};

// One axis pass over interior-only SoA state arrays (variable plane stride
// `ncell`, cell index (i*n1 + j)*n2 + k), in real type Real (double, or
// float for the FP32 variant).  The limiter's lambda is formed in double
// from dt (the device scalars hold dt in double) and then rounded, as the
// reference's R::from_f64(ndim * dt / dx).
template <class Real>
struct PassArgsT {
    const Real* U;         // stage input
    Real* R;               // residual accumulator
    const Real* U0;        // U^n (RK combination)
    Real* Uout;            // RK output (may equal U0 at stage 3)
    DevScalars* sc;
    int n[3];
    long long stride[3];
    long long ncell;
    EosT<Real> eos;
    int bc_kind[2];
    Real bc_state[2][5];
    Real inv_dx;           // 1 / dx_d (divergence)
    double ndim_pos;       // D if the positivity limiter is on, else 0
    double dx_d;           // dx_d (limiter lambda = ndim*dt/dx_d)
    Real dx_all[3];        // CFL rate: sum over active axes of max_speed / dx
    int active[3];
    Real tol;              // hybrid sensor tolerance
    int first;             // residual input is zero (first active axis)
    int rk_stage;          // 0: store residual only; 1..3: fused SSP-RK3 stage
    int want_rate;         // accumulate the CFL rate of the RK output
    int fast_axis, slow_axis;
    int dfast;             // stride along the pass axis is 1
    int ntile, ngroup;     // tiles along the axis, pencil groups along fast_axis
    int stage_dt_host;     // 0: lambda from sc->dt
    double stage_dt;       // used when stage_dt_host != 0
};
using PassArgs = PassArgsT<double>;

// Ensemble pass: per-instance pass arguments and the prefix sums of the
// instances' tile counts (ninst + 1 entries; total = prefix[ninst]).
struct EnsPassArgs {
    const PassArgs* args;
    const long long* prefix;
    int ninst;
    long long total;
};

struct TileCfg {
    int T, P, NT;
};

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup w runs on
// XCD w % 8, and each XCD has its own L2).  Tiles that share cache lines are
// neighbours in tile order: adjacent tiles of a pencil share 10 halo cells,
// and the pencil groups 2k and 2k+1 of a strided pass read the two halves of
// the same 128-byte lines.  xcd_major hands out the tile order in chunks of
// kXcdChunk consecutive tiles, chunk g to XCD g % 8, so that neighbours run
// on one XCD at about the same time and a line fetched for one tile is still
// in that L2 for the next; the chunks of one XCD are spread over the whole
// tile range, so work that varies in space stays balanced over the XCDs.
// A bijection of [0, nblk) for any nblk (the last nblk mod 256 tiles keep
// their own index).
#ifndef EM_XCD_ORDER
#define EM_XCD_ORDER 1
#endif
constexpr int kXcds = 8;
constexpr int kXcdChunk = 32;
__device__ __forceinline__ int xcd_major(int w, int nblk) {
    constexpr int span = kXcds * kXcdChunk;
    if (!EM_XCD_ORDER || w >= nblk / span * span) return w;
    const int x = w % kXcds, m = w / kXcds;
    return ((m / kXcdChunk) * kXcds + x) * kXcdChunk + m % kXcdChunk;
}

// The same order for a persistent grid of workgroups striding over work
// items: the grid/8 workgroups of XCD x take turns over the items of the
// chunks g = x, x + 8, x + 16, ... in order.  Loop:
//   for (long long t = s.first;; t += s.step) {
//       const long long w = s.item(t);
//       if (w >= nwork) break;  ...
// item() increases with t, so every item is visited once.  A grid that is
// not a multiple of 8 keeps the plain stride.
struct XcdStride {
    long long first, step;
    int x, chunked;
    __device__ __forceinline__ long long item(long long t) const {
        if (!chunked) return t;
        return ((t / kXcdChunk) * kXcds + x) * kXcdChunk + t % kXcdChunk;
    }
};
__device__ __forceinline__ XcdStride xcd_stride() {
    const int grid = (int)gridDim.x, b = (int)blockIdx.x;
    if (!EM_XCD_ORDER || grid % kXcds != 0) return XcdStride{b, grid, 0, 0};
    return XcdStride{b / kXcds, grid / kXcds, b % kXcds, 1};
}

}  // namespace em
