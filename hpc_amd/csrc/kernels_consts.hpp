// Tile-plan limits shared by host code (operator.cpp) and the device code
// (kernels.hpp, kTileNnz / kTileRows); kept in one place so they cannot drift.
#pragma once

namespace xj {
constexpr int kTileNnzHost = 2048;
constexpr int kTileRowsHost = 256;
}  // namespace xj
