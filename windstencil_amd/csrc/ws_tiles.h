// ws_tiles.h -- tile shapes of the tiled RHS kernels, shared by the kernels, the
// batched launcher's tile lists and the autotuner.
#pragma once

#include "windstencil_amd.h"

namespace ws {

// x tiles: L0 cells along x by W0 rows; y/z tiles: L1 cells along the line by W1
// columns of x; K faces per thread.  Indexed by WS_TILE_*.
struct TileShapeDef {
    int L0, W0, L1, W1, K;
};

constexpr TileShapeDef kTileShapes[WS_TILE_COUNT] = {
    {15, 16, 16, 15, 1},   // WS_TILE_15x16_K1
    {30, 16, 32, 15, 2},   // WS_TILE_30x16_K2
    {30, 16, 32, 16, 2},   // WS_TILE_30x16_A16_K2
};

// y/z tiles of 16 columns start on a multiple of 16 ghosted columns
constexpr int tile_xalign(int dir, int w) { return (dir != 0 && w == 16) ? 16 : 1; }

}  // namespace ws
