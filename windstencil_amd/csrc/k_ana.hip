// k_ana.hip -- failure detection on the device: the reference's ana_NAN_and_NT (NaN and
// negative temperature over the ghosted box, src/runtime/OCFD_ana.cpp:115-225) and
// ana_Jac (negative Jacobian, :35-56), plus negative density and the CFL number of a
// given time step.
//
// One pass over the ghosted box: each work-group reduces its counts, first offending
// cells (smallest linear ghosted index) and the largest max(|u|,|v|,|w|) + c of its
// interior cells in LDS and writes one partial record; one work-group reduces the
// records.  No atomics: the report is reproducible.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ws_device.h"

namespace ws {

namespace {

constexpr int AT = 256;
constexpr int ABLOCKS = 512;
constexpr long long NONE = 0x7fffffffffffffffLL;

// per-quantity count and first offending ghosted index
struct AnaRec {
    long long count[4];      // NaN T, T < 0, d < 0, Ajac < 0
    long long first[4];
    double smax;             // largest max(|u|,|v|,|w|) + c over interior cells
};

__device__ __forceinline__ void rec_merge(AnaRec &a, const AnaRec &b)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        a.count[q] += b.count[q];
        a.first[q] = min(a.first[q], b.first[q]);
    }
    a.smax = fmax(a.smax, b.smax);
}

__device__ __forceinline__ void rec_reduce(AnaRec &r, AnaRec *red)
{
    red[threadIdx.x] = r;
    __syncthreads();
    for (int s = AT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) rec_merge(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    r = red[0];
}

__global__ void __launch_bounds__(AT) k_ana_partial(KParams k, ws_field d, ws_field u, ws_field v, ws_field w,
                                                    ws_field T, ws_field ajac, AnaRec *partial)
{
    __shared__ AnaRec red[AT];
    AnaRec r;
#pragma unroll
    for (int q = 0; q < 4; ++q) { r.count[q] = 0; r.first[q] = NONE; }
    r.smax = 0.0;
    const int nxg = k.nx + 2 * LAP;
    const long long total = (long long)nxg * k.nyg * k.nzg;
    for (long long i = (long long)blockIdx.x * AT + threadIdx.x; i < total; i += (long long)gridDim.x * AT) {
        const int x = (int)(i % nxg), y = (int)((i / nxg) % k.nyg), z = (int)(i / ((long long)nxg * k.nyg));
        const double tv = T.ptr[fidx(T.pitch, k.nyg, x, y, z)];
        const double dv = d.ptr[fidx(d.pitch, k.nyg, x, y, z)];
        const double jv = ajac.ptr ? ajac.ptr[fidx(ajac.pitch, k.nyg, x, y, z)] : k.ajac;
        const bool bad[4] = {isnan(tv), tv < 0.0, dv < 0.0, jv < 0.0};
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (bad[q]) {
                ++r.count[q];
                r.first[q] = min(r.first[q], i);
            }
        const bool interior = x >= LAP && x < LAP + k.nx && y >= LAP && y < LAP + k.ny && z >= LAP && z < LAP + k.nz;
        if (interior && tv >= 0.0) {
            const double c = sqrt(tv) / k.ama;
            const double a = fmax(fabs(u.ptr[fidx(u.pitch, k.nyg, x, y, z)]),
                                  fmax(fabs(v.ptr[fidx(v.pitch, k.nyg, x, y, z)]),
                                       fabs(w.ptr[fidx(w.pitch, k.nyg, x, y, z)])));
            r.smax = fmax(r.smax, a + c);
        }
    }
    rec_reduce(r, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ void __launch_bounds__(AT) k_ana_final(const AnaRec *partial, int n, AnaRec *out)
{
    __shared__ AnaRec red[AT];
    AnaRec r;
#pragma unroll
    for (int q = 0; q < 4; ++q) { r.count[q] = 0; r.first[q] = NONE; }
    r.smax = 0.0;
    for (int i = threadIdx.x; i < n; i += AT) rec_merge(r, partial[i]);
    rec_reduce(r, red);
    if (threadIdx.x == 0) *out = r;
}

// Device scratch of ws_ana_state_noalloc: the partial records of k_ana_partial and the
// final record.  Fixed size (ABLOCKS + 1 records of 72 bytes, about 37 KB) whatever the
// box, since the pass is grid-strided over ABLOCKS work-groups; each device holds its own
// copy until the process ends.  In host code `scratch` names the host-side shadow, not the
// device array: kernels get the address hipGetSymbolAddress returns for the current
// device (ana_scratch_ptr), never `scratch` itself.
constexpr int ANA_SCRATCH_RECS = ABLOCKS + 1;
static __device__ __align__(256) AnaRec scratch[ANA_SCRATCH_RECS];

constexpr int ANA_MAX_DEVICES = 64;
std::mutex g_ana_mutex;
AnaRec *g_ana_scratch[ANA_MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per device.
int ana_scratch_ptr(AnaRec **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_error(e, "hipGetDevice");
    if (dev < 0 || dev >= ANA_MAX_DEVICES) return set_error(WS_ERR_UNSUPPORTED, "device index beyond 64");
    std::lock_guard<std::mutex> lock(g_ana_mutex);
    if (!g_ana_scratch[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return hip_error(e, "analysis scratch: hipGetSymbolAddress");
        g_ana_scratch[dev] = static_cast<AnaRec *>(ptr);
    }
    *out = g_ana_scratch[dev];
    return WS_OK;
}

// The pass and its report; buf holds ABLOCKS + 1 records of device memory.  Every record
// is written before it is read, so buf needs no clearing between calls.
int ana_run(ws_ctx *c, const ws_state *st, const ws_metric_fields *metrics, double dt, ws_ana_report *rep,
            hipStream_t s, AnaRec *buf)
{
    const KParams &k = c->kp;
    const ws_field aj = metrics ? metrics->ajac : ws_field{};
    hipLaunchKernelGGL(k_ana_partial, dim3(ABLOCKS), dim3(AT), 0, s, k, st->d, st->u, st->v, st->w, st->T, aj, buf);
    hipLaunchKernelGGL(k_ana_final, dim3(1), dim3(AT), 0, s, (const AnaRec *)buf, ABLOCKS, buf + ABLOCKS);
    AnaRec r{};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&r, buf + ABLOCKS, sizeof(AnaRec), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(e, "state analysis");
    const long long nxg = k.nx + 2 * LAP, nyg = k.nyg;
    int64_t *cnt[4] = {&rep->nan_T, &rep->neg_T, &rep->neg_d, &rep->neg_jac};
    int32_t *first[4] = {rep->first_nan_T, rep->first_neg_T, rep->first_neg_d, rep->first_neg_jac};
    for (int q = 0; q < 4; ++q) {
        *cnt[q] = r.count[q];
        if (r.first[q] == NONE) {
            first[q][0] = first[q][1] = first[q][2] = WS_ANA_NONE;
        } else {
            const long long i = r.first[q];
            first[q][0] = (int32_t)(i % nxg) - LAP;
            first[q][1] = (int32_t)((i / nxg) % nyg) - LAP;
            first[q][2] = (int32_t)(i / (nxg * nyg)) - LAP;
        }
    }
    const double hmin = fmin(k.h[0], fmin(k.h[1], k.h[2]));
    rep->cfl = dt > 0.0 ? dt * r.smax / hmin : 0.0;
    rep->fatal = (rep->nan_T > 0 || rep->neg_T > 10) ? 1 : 0;
    return WS_OK;
}

int ana_check(ws_ctx *c, const ws_state *st, const ws_metric_fields *metrics, ws_ana_report *rep)
{
    if (!c || !st || !rep) return set_error(WS_ERR_INVALID, "null argument");
    if (!st->d.ptr || !st->u.ptr || !st->v.ptr || !st->w.ptr || !st->T.ptr)
        return set_error(WS_ERR_INVALID, "state primitives not allocated");
    if (metrics && !metrics->ajac.ptr) return set_error(WS_ERR_INVALID, "metric fields without Ajac");
    // the ghosted linear index of the pass and of the first offending cells is 64-bit
    const long long cells = (long long)(c->kp.nx + 2 * LAP) * c->kp.nyg * c->kp.nzg;
    if (cells <= 0 || cells > (1LL << 40)) return set_error(WS_ERR_INVALID, "ghosted box too large for the analysis pass");
    return WS_OK;
}

}  // namespace

}  // namespace ws

using namespace ws;

extern "C" int ws_ana_state(ws_ctx *c, const ws_state *st, const ws_metric_fields *metrics, double dt,
                            ws_ana_report *rep, hipStream_t s)
{
    int r = ana_check(c, st, metrics, rep);
    if (r) return r;
    AnaRec *buf = nullptr;
    hipError_t e = hipMalloc(&buf, sizeof(AnaRec) * ANA_SCRATCH_RECS);
    if (e != hipSuccess) return hip_error(e, "analysis scratch");
    r = ana_run(c, st, metrics, dt, rep, s, buf);
    (void)hipFree(buf);
    return r;
}

extern "C" int ws_ana_state_noalloc(ws_ctx *c, const ws_state *st, const ws_metric_fields *metrics, double dt,
                                    ws_ana_report *rep, hipStream_t s)
{
    int r = ana_check(c, st, metrics, rep);
    if (r) return r;
    AnaRec *buf = nullptr;
    if ((r = ana_scratch_ptr(&buf))) return r;
    return ana_run(c, st, metrics, dt, rep, s, buf);
}
