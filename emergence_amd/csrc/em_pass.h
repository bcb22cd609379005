// Launchers of the fused axis pass, one per arithmetic variant.
#pragma once
#include <hip/hip_runtime.h>

#include "em_types.h"

namespace em {

// Tile shapes (cells along the axis x pencils, threads); index = cfg id.
constexpr TileCfg kTileCfgs[] = {{32, 16, 640}, {64, 8, 576}, {128, 4, 576}, {256, 2, 576}};
constexpr int kNumTileCfgs = 4;

namespace precise {
hipError_t launch_pass(int d, int mode, int cfg, const PassArgs& a, hipStream_t s);
}
namespace fast {
hipError_t launch_pass(int d, int mode, int cfg, const PassArgs& a, hipStream_t s);
}

}  // namespace em
