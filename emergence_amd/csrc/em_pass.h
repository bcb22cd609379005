// Launchers of the fused axis pass, one per arithmetic variant.
#pragma once
#include <hip/hip_runtime.h>

#include "em_amr_types.h"
#include "em_types.h"

namespace em {

// Tile shapes (cells along the axis x pencils, threads); index = cfg id.
// Shapes 4..9 give exactly one midpoint per thread ((T+5)*P == NT); shape 9
// puts 16 pencils in a workgroup, so that a strided pass reads whole 128-byte
// lines (16 neighbouring cells of the innermost axis).
constexpr TileCfg kTileCfgs[] = {{32, 16, 640}, {64, 8, 576}, {128, 4, 576}, {256, 2, 576}, {27, 8, 256},
                                 {59, 4, 256},  {59, 8, 512}, {27, 16, 512}, {123, 4, 512}, {59, 16, 1024}};
constexpr int kNumTileCfgs = 10;

#define EM_DECLARE_LAUNCHERS                                                               \
    hipError_t launch_pass_hybrid(int d, int cfg, const PassArgs& a, hipStream_t s);      \
    hipError_t launch_pass_weno(int d, int cfg, const PassArgs& a, hipStream_t s);        \
    hipError_t launch_pass_central(int d, int cfg, const PassArgs& a, hipStream_t s);     \
    hipError_t launch_amr_hybrid(int d, int b, int shape, const AmrPassArgs& pa, unsigned grid, hipStream_t s); \
    hipError_t launch_amr_weno(int d, int b, int shape, const AmrPassArgs& pa, unsigned grid, hipStream_t s);   \
    hipError_t launch_amr_central(int d, int b, int shape, const AmrPassArgs& pa, unsigned grid, hipStream_t s); \
    inline hipError_t launch_amr_pass(int d, int mode, int b, int shape, const AmrPassArgs& pa,       \
                                      unsigned grid, hipStream_t s) {                                \
        switch (mode) {                                                                    \
            case MODE_HYBRID: return launch_amr_hybrid(d, b, shape, pa, grid, s);          \
            case MODE_WENO: return launch_amr_weno(d, b, shape, pa, grid, s);              \
            case MODE_CENTRAL: return launch_amr_central(d, b, shape, pa, grid, s);        \
            default: return hipErrorInvalidValue;                                          \
        }                                                                                  \
    }                                                                                      \
    hipError_t launch_amr32_hybrid(int d, int b, int shape, const AmrPassArgsT<float>& pa, unsigned grid,  \
                                   hipStream_t s);                                                          \
    hipError_t launch_amr32_weno(int d, int b, int shape, const AmrPassArgsT<float>& pa, unsigned grid,    \
                                 hipStream_t s);                                                            \
    hipError_t launch_amr32_central(int d, int b, int shape, const AmrPassArgsT<float>& pa, unsigned grid, \
                                    hipStream_t s);                                                         \
    inline hipError_t launch_amr_pass(int d, int mode, int b, int shape, const AmrPassArgsT<float>& pa,    \
                                      unsigned grid, hipStream_t s) {                                      \
        switch (mode) {                                                                    \
            case MODE_HYBRID: return launch_amr32_hybrid(d, b, shape, pa, grid, s);        \
            case MODE_WENO: return launch_amr32_weno(d, b, shape, pa, grid, s);            \
            case MODE_CENTRAL: return launch_amr32_central(d, b, shape, pa, grid, s);      \
            default: return hipErrorInvalidValue;                                          \
        }                                                                                  \
    }                                                                                      \
    hipError_t launch_amr_multi_hybrid(int d, const AmrMultiArgs& m, unsigned grid, hipStream_t s);   \
    hipError_t launch_amr_multi_weno(int d, const AmrMultiArgs& m, unsigned grid, hipStream_t s);     \
    hipError_t launch_amr_multi_central(int d, const AmrMultiArgs& m, unsigned grid, hipStream_t s);  \
    inline hipError_t launch_amr_multi(int d, int mode, const AmrMultiArgs& m, unsigned grid, hipStream_t s) { \
        switch (mode) {                                                                    \
            case MODE_HYBRID: return launch_amr_multi_hybrid(d, m, grid, s);               \
            case MODE_WENO: return launch_amr_multi_weno(d, m, grid, s);                   \
            case MODE_CENTRAL: return launch_amr_multi_central(d, m, grid, s);             \
            default: return hipErrorInvalidValue;                                          \
        }                                                                                  \
    }                                                                                      \
    hipError_t launch_ens_hybrid(int d, int cfg, const EnsPassArgs& e, unsigned grid, hipStream_t s); \
    hipError_t launch_ens_weno(int d, int cfg, const EnsPassArgs& e, unsigned grid, hipStream_t s);   \
    hipError_t launch_ens_central(int d, int cfg, const EnsPassArgs& e, unsigned grid, hipStream_t s); \
    inline hipError_t launch_ens_pass(int d, int mode, int cfg, const EnsPassArgs& e, unsigned grid,  \
                                      hipStream_t s) {                                               \
        switch (mode) {                                                                    \
            case MODE_HYBRID: return launch_ens_hybrid(d, cfg, e, grid, s);                \
            case MODE_WENO: return launch_ens_weno(d, cfg, e, grid, s);                    \
            case MODE_CENTRAL: return launch_ens_central(d, cfg, e, grid, s);              \
            default: return hipErrorInvalidValue;                                          \
        }                                                                                  \
    }                                                                                      \
    inline hipError_t launch_pass(int d, int mode, int cfg, const PassArgs& a, hipStream_t s) { \
        switch (mode) {                                                                    \
            case MODE_HYBRID: return launch_pass_hybrid(d, cfg, a, s);                     \
            case MODE_WENO: return launch_pass_weno(d, cfg, a, s);                         \
            case MODE_CENTRAL: return launch_pass_central(d, cfg, a, s);                   \
            default: return hipErrorInvalidValue;                                          \
        }                                                                                  \
    }                                                                                      \
    hipError_t launch_pass32_hybrid(int d, int cfg, const PassArgsT<float>& a, hipStream_t s); \
    hipError_t launch_pass32_weno(int d, int cfg, const PassArgsT<float>& a, hipStream_t s);   \
    hipError_t launch_pass32_central(int d, int cfg, const PassArgsT<float>& a, hipStream_t s); \
    inline hipError_t launch_pass(int d, int mode, int cfg, const PassArgsT<float>& a, hipStream_t s) { \
        switch (mode) {                                                                    \
            case MODE_HYBRID: return launch_pass32_hybrid(d, cfg, a, s);                   \
            case MODE_WENO: return launch_pass32_weno(d, cfg, a, s);                       \
            case MODE_CENTRAL: return launch_pass32_central(d, cfg, a, s);                 \
            default: return hipErrorInvalidValue;                                          \
        }                                                                                  \
    }

namespace precise {
EM_DECLARE_LAUNCHERS
}
namespace fast {
EM_DECLARE_LAUNCHERS
}

}  // namespace em
