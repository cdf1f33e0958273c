// k_rhs_mixed.hip -- mixed-precision inviscid RHS (FP32 splitting and candidate
// stencils, FP64 weights near discontinuities, FP64 flux difference).
#include <hip/hip_runtime.h>

#include "ws_device.h"

namespace ws {

int launch_rhs_mixed(const Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                     const ws_metric_fields *m, ws_soa du, ws_job job, int accumulate, hipStream_t s)
{
    (void)c; (void)d; (void)u; (void)v; (void)w; (void)cc; (void)m; (void)du; (void)job; (void)accumulate; (void)s;
    return set_error(WS_ERR_UNSUPPORTED, "mixed-precision RHS is not built yet");
}

}  // namespace ws
