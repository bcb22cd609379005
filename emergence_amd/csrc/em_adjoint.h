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
    // SSP-RK3 adjoint update fused into the store of T = MU + this axis's
    // contribution on the last axis of a stage (T itself is not stored).
    // For the forward u1 = u0 + dt L(u0); u2 = 3/4 u0 + 1/4 (u1 + dt L(u1));
    // u3 = u0 / 3 + 2/3 (u2 + dt L(u2)) and T = L'(u_s)^T (cotangent):
    //   0: none (MU = T);
    //   3: cBout = 2/3 cA + 2/3 cdt T, cZ = cA / 3;
    //   2: cZ += 3/4 cB, cBout = 1/4 cB + 1/4 cdt T (cBout is not this
    //      pass's LAM: other tiles still read it);
    //   1: cZ = cZ + cB + cdt T (the cotangent of the step's input).
    int combine;
    double cdt;
    const double* cA;
    const double* cB;
    double* cZ;
    double* cBout;
};

constexpr int kAdjThreads = 256;     // threads per adjoint workgroup (4 waves; three workgroups per CU)
constexpr int kAdjTileCells = 128;   // interior cells per tile at most (LDS of a tile <= 45 KB)
int adj_tiles(int n, int kind_lo, int kind_hi, int cap = kAdjTileCells);
int adj_tile_extent(int n, int ntile, int kind_lo, int kind_hi);
size_t adj_lds_bytes(int ne);
hipError_t launch_adj_pass(int d, const AdjArgs& a, hipStream_t s);

}  // namespace em
