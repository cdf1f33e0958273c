// ws_tiles.h -- tile shapes of the tiled RHS kernels, shared by the kernels, the
// batched launcher's tile lists and the autotuner.
#pragma once

#include "windstencil_amd.h"

namespace ws {

// x tiles: L0 cells along x by W0 rows; y tiles: L1 cells along y by W1 columns of x;
// z tiles: L2 cells along z by W2 columns of x; K faces per thread.  Indexed by WS_TILE_*.
// A work-group of W * ceil((L + 1) / K) lanes holds 10 * W * (L + 8) doubles of LDS: within
// four wavefronts and about 50 KiB, three work-groups share a compute unit (the
// five-wavefront 32x16 y/z tiles fit only two, which is why they are slow); the
// two-wavefront 126x2 x tiles (21 KiB) fit seven, fourteen wavefronts per compute unit.
struct TileShapeDef {
    int L0, W0, L1, W1, L2, W2, K;
};

constexpr TileShapeDef kTileShapes[WS_TILE_COUNT] = {
    {15, 16, 16, 15, 16, 15, 1},   // WS_TILE_15x16_K1
    {30, 16, 32, 15, 32, 15, 2},   // WS_TILE_30x16_K2
    {30, 16, 32, 16, 32, 16, 2},   // WS_TILE_30x16_A16_K2 (five wavefronts; a reference point)
    {62, 8, 32, 15, 32, 15, 2},    // WS_TILE_62x8_K2
    {126, 2, 32, 15, 32, 15, 2},   // WS_TILE_126x2_K2
};

// y/z tiles of 16 columns start on a multiple of 16 ghosted columns
constexpr int tile_xalign(int dir, int w) { return (dir != 0 && w == 16) ? 16 : 1; }

}  // namespace ws
