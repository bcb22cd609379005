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
// `ncell`, cell index (i*n1 + j)*n2 + k).
struct PassArgs {
    const double* U;       // stage input
    double* R;             // residual accumulator
    const double* U0;      // U^n (RK combination)
    double* Uout;          // RK output (may equal U0 at stage 3)
    DevScalars* sc;
    int n[3];
    long long stride[3];
    long long ncell;
    Eos eos;
    int bc_kind[2];
    double bc_state[2][5];
    double inv_dx;         // 1 / dx_d (divergence)
    double ndim_pos;       // D if the positivity limiter is on, else 0
    double dx_d;           // dx_d (limiter lambda = ndim*dt/dx_d)
    double dx_all[3];      // CFL rate: sum over active axes of max_speed / dx
    int active[3];
    double tol;            // hybrid sensor tolerance
    int first;             // residual input is zero (first active axis)
    int rk_stage;          // 0: store residual only; 1..3: fused SSP-RK3 stage
    int want_rate;         // accumulate the CFL rate of the RK output
    int fast_axis, slow_axis;
    int dfast;             // stride along the pass axis is 1
    int ntile, ngroup;     // tiles along the axis, pencil groups along fast_axis
    int stage_dt_host;     // 0: lambda from sc->dt
    double stage_dt;       // used when stage_dt_host != 0
};

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

}  // namespace em
