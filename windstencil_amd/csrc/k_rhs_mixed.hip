// k_rhs_mixed.hip -- mixed-precision inviscid RHS: FP64 splitting, WENO7 weights and
// candidates in FP32 on first differences taken relative to FP64 base points, FP64
// merge and flux difference.  The kernel template is in rhs_tile.h.
#include "rhs_tile.h"

namespace ws {

int launch_rhs_mixed(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                     const ws_metric_fields *m, ws_soa du, ws_job job, int accumulate, hipStream_t s)
{
    return launch_rhs_tile<true>(c, d, u, v, w, cc, m, du, job, accumulate, s);
}

int launch_rhs_mixed_stage(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc, ws_soa du,
                           ws_job job, const StageArgs &sa, hipStream_t s)
{
    return launch_rhs_tile<true>(c, d, u, v, w, cc, nullptr, du, job, 0, s, &sa);
}

int launch_rhs_batch_mixed(const KParams &k, const BatchTiles &bt, hipStream_t s)
{
    return launch_rhs_batch_tile<true>(k, bt, s);
}

}  // namespace ws
