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
// Row bins of the tile SpMV: inside a short-row tile, a row with at most
// kWaveRowMinHost entries is summed by its own lane left to right; a longer
// one by its wavefront (lane l adds entries l, l+64, ... in order, then a
// fixed xor butterfly).  Rows above kTilePlanNnz are tiles of their own,
// summed by the whole workgroup.
constexpr int kWaveRowMinHost = 64;
}  // namespace xj

namespace xj {
// Panel-swept layout (spmv_panels / cg_spmv_panels in jpcg.hip): column
// panels of kPanelColsHost columns, so an entry's column is stored as a
// uint16 offset in its panel; row blocks of at most kPanelRowsMaxHost rows,
// whose running sums stay in LDS (uint16 row offsets); workgroups of
// kPanelBlockHost threads stage kPanelChunkHost entries at a time.
constexpr int kPanelColsHost = 65536;
constexpr int kPanelRowsMaxHost = 4096;
constexpr int kPanelBlockHost = 1024;
constexpr int kPanelChunkHost = 2048;
// An entry's uint16 row word: the row offset in its block in the low
// kPanelRowBits bits; above them, for the first entry of a row's run inside a
// staged chunk, the run's length (kPanelRunScan: that many or more, summed by
// a wavefront that scans for the run's end), 0 for the run's other entries.
constexpr int kPanelRowBits = 12;
constexpr int kPanelRunScan = 15;
static_assert(kPanelRowsMaxHost <= (1 << kPanelRowBits) && kPanelRunScan < (1 << (16 - kPanelRowBits)),
              "row offset and run code share 16 bits");
// LDS bytes of one panel workgroup: row sums + staged products + staged row words
constexpr int panel_lds_bytes(int block_rows) { return block_rows * 8 + kPanelChunkHost * 10; }
// column-offset pattern tiles run two tiles per workgroup; the CG loop keeps
// one p.Ap partial per workgroup
constexpr int kPatternTilesHost = 2;
constexpr int pattern_groups(int ntiles) { return (ntiles + kPatternTilesHost - 1) / kPatternTilesHost; }
// tiles per workgroup of the row-tile SpMV kernels (measured on the 160^3
// stencil: 1 beats 2, which costs registers — 238 vs 260 us for fp64 values)
constexpr int kRowTilesHost = 1;
constexpr int tile_groups(int ntiles) { return (ntiles + kRowTilesHost - 1) / kRowTilesHost; }
}  // namespace xj
