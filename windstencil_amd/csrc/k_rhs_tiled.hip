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

int launch_rhs_batch_tiled(const KParams &k, const BatchTiles &bt, hipStream_t s)
{
    return launch_rhs_batch_tile<false>(k, bt, s);
}

}  // namespace ws
