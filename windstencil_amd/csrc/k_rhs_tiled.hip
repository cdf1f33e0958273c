// k_rhs_tiled.hip -- LDS-tiled inviscid RHS, FP64 (the fast path); the kernel
// template is in rhs_tile.h.
#include "rhs_tile.h"

namespace ws {

int launch_rhs_tiled(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                     const ws_metric_fields *m, ws_soa du, ws_job job, int accumulate, hipStream_t s)
{
    return launch_rhs_tile<false>(c, d, u, v, w, cc, m, du, job, accumulate, s);
}

int launch_rhs_tiled_stage(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc, ws_soa du,
                           ws_job job, const StageArgs &sa, hipStream_t s)
{
    return launch_rhs_tile<false>(c, d, u, v, w, cc, nullptr, du, job, 0, s, &sa);
}

long stage_launch_tiles(const Ctx &c)
{
    Box box;
    for (int d = 0; d < 3; ++d) box.s[d] = LAP;
    box.e[0] = c.kp.nx + LAP; box.e[1] = c.kp.ny + LAP; box.e[2] = c.kp.nz + LAP;
    switch (c.stage_shape >= 0 ? c.stage_shape : c.tile_shape) {
    case 0: return stage_tiles_n<0>(box, c.stage_order);
    case 1: return stage_tiles_n<1>(box, c.stage_order);
    case 2: return stage_tiles_n<2>(box, c.stage_order);
    case 3: return stage_tiles_n<3>(box, c.stage_order);
    default: return stage_tiles_n<4>(box, c.stage_order);
    }
}

int launch_rhs_batch_tiled(const KParams &k, const BatchTiles &bt, hipStream_t s)
{
    return launch_rhs_batch_tile<false>(k, bt, s);
}

}  // namespace ws
