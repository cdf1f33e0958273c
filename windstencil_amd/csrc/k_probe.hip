// k_probe.hip -- vector ALU throughput of the device: FMA, multiply and add in FP64 and
// scalar FP32, the attainable VALU roof that the RHS kernels are compared against beside
// the published peak.
//
// Every lane runs eight independent chains (FMA: x = x * a + b with |a| < 1; multiply:
// x = x * a; add: x = x + y with y fixed per chain; the values stay bounded), unrolled so
// that the loop overhead is scalar work; the grid holds 16 work-groups of 256 lanes per
// compute unit.  FP32 is the scalar (non-packed) form the kernels use.
#include <hip/hip_runtime.h>

#include "ws_internal.h"

namespace ws {

namespace {

constexpr int PT = 256;
constexpr int CHAINS = 8;
constexpr int UNROLL = 16;

template <typename T, int OP>
__global__ void __launch_bounds__(PT) k_valu_probe(T a, T b, int iters, T *out)
{
    T x[CHAINS], y[CHAINS];
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) {
        x[c] = (T)(threadIdx.x + c) * (T)1e-3;
        y[c] = (T)(c + 1) * b;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
#pragma unroll
            for (int c = 0; c < CHAINS; ++c) {
                if (OP == WS_PROBE_FMA) x[c] = fma(x[c], a, b);
                else if (OP == WS_PROBE_MUL) x[c] = x[c] * a;
                else x[c] = x[c] + y[c];
            }
    }
    T s = 0;
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) s += x[c];
    out[(size_t)blockIdx.x * PT + threadIdx.x] = s;
}

template <typename T, int OP>
int probe(double *rate, hipStream_t s)
{
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_error(e, "hipGetDevice");
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return hip_error(e, "compute unit count");
    const int blocks = 16 * cus;
    const int iters = 256;
    T *out = nullptr;
    e = hipMalloc(&out, sizeof(T) * (size_t)blocks * PT);
    if (e != hipSuccess) return hip_error(e, "probe output");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int r = WS_OK;
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
        r = hip_error(e, "probe event");
    } else {
        const T a = (T)0.999999, b = (T)1e-6;
        float best = 1e30f;
        // WARM untimed launches (about 40 ms: the clock settles at its loaded value), then
        // the best of five
        constexpr int WARM = 40;
        for (int rep = 0; rep < WARM + 5 && !r; ++rep) {
            if ((e = hipEventRecord(e0, s)) != hipSuccess) { r = hip_error(e, "probe event"); break; }
            hipLaunchKernelGGL((k_valu_probe<T, OP>), dim3(blocks), dim3(PT), 0, s, a, b, iters, out);
            if ((e = hipGetLastError()) != hipSuccess) { r = hip_error(e, "probe launch"); break; }
            if ((e = hipEventRecord(e1, s)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess) {
                r = hip_error(e, "probe sync");
                break;
            }
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            if (rep >= WARM && ms < best) best = ms;
        }
        if (!r) {
            const double ops = (double)CHAINS * UNROLL * (double)iters * (double)blocks * PT;
            *rate = ops / (best * 1e-3) / 1e12;
        }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(out);
    return r;
}

template <typename T>
int probe_op(int32_t op, double *rate, hipStream_t s)
{
    if (op == WS_PROBE_FMA) return probe<T, WS_PROBE_FMA>(rate, s);
    if (op == WS_PROBE_MUL) return probe<T, WS_PROBE_MUL>(rate, s);
    if (op == WS_PROBE_ADD) return probe<T, WS_PROBE_ADD>(rate, s);
    return set_error(WS_ERR_INVALID, "unknown probe operation");
}

}  // namespace

}  // namespace ws

extern "C" int ws_probe_valu(int32_t precision, int32_t op, double *tera_ops, hipStream_t stream)
{
    if (!tera_ops) return ws::set_error(WS_ERR_INVALID, "null output pointer");
    if (precision == WS_PREC_FP64) return ws::probe_op<double>(op, tera_ops, stream);
    if (precision == WS_PREC_MIXED) return ws::probe_op<float>(op, tera_ops, stream);
    return ws::set_error(WS_ERR_INVALID, "unknown precision");
}
