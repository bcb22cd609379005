// Discrete adjoint of the Weno-mode residual (host-side declarations).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

#include "em_device.h"

namespace em {

// One axis of MU (+)= (d R_axis / d U)^T LAM at the state U (SoA interior
// arrays of 5 * ncell doubles).
struct AdjArgs {
    const double* U;
    const double* LAM;
    double* MU;
    int n[3];
    long long stride[3];
    long long ncell;
    Eos eos;
    int bc_kind[2];          // boundary kinds of the pass axis (low, high)
    double bc_state[2][5];
    double inv_dx;
    double lam_lim;          // limiter lambda D * dt / dx (0: limiter off)
    int first;               // MU = contribution (else MU += contribution)
    int ntile;               // tiles per pencil (adj_tiles)
    int ne;                  // output positions of the longest tile (adj_tile_extent)
    // SSP-RK3 adjoint combination fused into the store of T = MU + this
    // axis's contribution (the last axis of a stage; k_adj_combine's
    // updates, T itself not stored): 0 none (MU = T); 3: cBout = 2/3 cA +
    // 2/3 cdt T, cZ = cA / 3; 2: cZ += 3/4 cB, cBout = 1/4 cB + 1/4 cdt T
    // (cBout must not be this pass's LAM); 1: cZ = cZ + cB + cdt T
    int combine;
    double cdt;
    const double* cA;
    const double* cB;
    double* cZ;
    double* cBout;
};

// Ensemble VJP: per-instance arguments and prefix sums of their work items
// (pencils x tiles; ninst + 1 entries; total = prefix[ninst]).
struct EnsAdjArgs {
    const AdjArgs* args;
    const long long* prefix;
    int ninst;
    long long total;
};

// One instance of the batched RK adjoint combination (dts[dt_index] is the
// instance's dt of the step being reversed).
struct AdjCombineTask {
    double* A;
    const double* T;
    double* B;
    double* Z;
    long long nv;
    long long dt_index;
};

constexpr int kAdjThreads = 256;     // threads per adjoint workgroup (4 waves; three workgroups per CU)
constexpr int kAdjTileCells = 128;   // interior cells per tile at most (LDS of a tile <= 45 KB)
int adj_tiles(int n, int kind_lo, int kind_hi);
int adj_tile_extent(int n, int ntile, int kind_lo, int kind_hi);
hipError_t launch_ens_adj_pass(int d, const EnsAdjArgs& e, size_t lds, unsigned grid, hipStream_t s);
hipError_t launch_ens_adj_combine(int stage, const AdjCombineTask* tasks, int ninst, const double* dts,
                                  long long maxnv, hipStream_t s);
size_t adj_lds_bytes(int ne);
hipError_t launch_adj_pass(int d, const AdjArgs& a, hipStream_t s);
hipError_t launch_adj_combine(int stage, long long nv, double dt, const double* A, const double* T, double* B,
                              double* Z, hipStream_t s);

}  // namespace em
