// Tile-plan limits shared by host code (operator.cpp) and the device code
// (kernels.hpp); kept in one place so they cannot drift.
#pragma once

namespace xj {
constexpr int kTileNnzHost = 2048;   // LDS slots for products: entries one tile may stream
constexpr int kTileRowsHost = 256;   // rows per tile (one lane per row)
// Entries the planner puts in a short-row tile: the streaming loop widens a
// tile's range down to a 16-byte-aligned start (up to 3 extra entries for
// fp32 values), so a tile holds at most kTileNnzHost - 4 entries.  A row with
// more entries is a tile of its own, summed by the whole workgroup.
constexpr int kTilePlanNnz = kTileNnzHost - 4;
}  // namespace xj
