// k_rhs_anchor.hip -- reference-order inviscid RHS (FP64), the GPU correctness anchor.
//
// Per direction: one kernel splits every point of every line (ghosts included) into a
// scratch buffer, a second kernel reconstructs both faces of every cell from it and
// accumulates the positive then the negative flux difference into du, in the order of
// the reference's decomposed kernels (OCFD_Schemes.cpp:461-484, 639-662 with
// put_du_p/m_kernel, :221-296).  No floating-point contraction, so the result equals
// the CPU restatement bit for bit.  Simple and memory-hungry by design; the tiled
// kernels are the fast path.
#include <hip/hip_runtime.h>

#include "ws_device.h"

namespace ws {

namespace {

struct Box {
    int s[3];      // start, ghosted coordinates
    int n[3];      // extents
};

__device__ __forceinline__ void metric_at(const KParams &k, const ws_metric_fields *mf, bool fields, int dir,
                                          uint32_t off_pitchless_x, int x, int y, int z, double &A1,
                                          double &A2, double &A3)
{
    (void)off_pitchless_x;
    if (fields) {
        const ws_field &f1 = mf->A[3 * dir + 0];
        const ws_field &f2 = mf->A[3 * dir + 1];
        const ws_field &f3 = mf->A[3 * dir + 2];
        A1 = f1.ptr[fidx(f1.pitch, k.nyg, x, y, z)];
        A2 = f2.ptr[fidx(f2.pitch, k.nyg, x, y, z)];
        A3 = f3.ptr[fidx(f3.pitch, k.nyg, x, y, z)];
    } else {
        A1 = k.A[3 * dir + 0];
        A2 = k.A[3 * dir + 1];
        A3 = k.A[3 * dir + 2];
    }
}

// scratch layout: [q (0..9)][b][a][p], p along the direction (n_dir + 2 LAP points),
// (a, b) the other two axes in increasing order.
template <int DIR>
__global__ void __launch_bounds__(256) k_anchor_split(KParams k, ws_field d, ws_field u, ws_field v, ws_field w,
                                                      ws_field cc, ws_metric_fields mf, int fields, Box box,
                                                      double *scratch)
{
    const int np = box.n[DIR] + 2 * LAP;
    const int A_AX = DIR == 0 ? 1 : 0;
    const int B_AX = DIR == 2 ? 1 : 2;
    const int na = box.n[A_AX], nb = box.n[B_AX];
    const long total = (long)np * na * nb;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int p = (int)(i % np);
    const int a = (int)((i / np) % na);
    const int b = (int)(i / ((long)np * na));
    int c[3];
    c[DIR] = box.s[DIR] - LAP + p;
    c[A_AX] = box.s[A_AX] + a;
    c[B_AX] = box.s[B_AX] + b;
    const int x = c[0], y = c[1], z = c[2];
    double A1, A2, A3;
    metric_at(k, &mf, fields != 0, DIR, 0, x, y, z, A1, A2, A3);
    double fp[5], fm[5];
    split_ref(k, d.ptr[fidx(d.pitch, k.nyg, x, y, z)], u.ptr[fidx(u.pitch, k.nyg, x, y, z)],
              v.ptr[fidx(v.pitch, k.nyg, x, y, z)], w.ptr[fidx(w.pitch, k.nyg, x, y, z)],
              cc.ptr[fidx(cc.pitch, k.nyg, x, y, z)], A1, A2, A3, fp, fm);
    const size_t plane = (size_t)total;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        scratch[(size_t)m * plane + i] = fp[m];
        scratch[(size_t)(5 + m) * plane + i] = fm[m];
    }
}

template <int DIR>
__global__ void __launch_bounds__(256) k_anchor_weno(KParams k, ws_metric_fields mf, int fields, Box box,
                                                     const double *scratch, ws_soa du)
{
#pragma clang fp contract(off)
    const int nd = box.n[DIR];
    const int np = nd + 2 * LAP;
    const int A_AX = DIR == 0 ? 1 : 0;
    const int B_AX = DIR == 2 ? 1 : 2;
    const int na = box.n[A_AX], nb = box.n[B_AX];
    const long total = (long)nd * na * nb;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int cpos = (int)(i % nd);
    const int a = (int)((i / nd) % na);
    const int b = (int)(i / ((long)nd * na));
    int c[3];
    c[DIR] = box.s[DIR] + cpos;
    c[A_AX] = box.s[A_AX] + a;
    c[B_AX] = box.s[B_AX] + b;
    const int x = c[0], y = c[1], z = c[2];
    const double aj = fields ? mf.ajac.ptr[fidx(mf.ajac.pitch, k.nyg, x, y, z)] : k.ajac;
    const double h = k.h[DIR];
    const size_t plane = (size_t)np * na * nb;
    const size_t line = ((size_t)b * na + a) * np;
    // face c-1/2 uses points cpos-4 .. cpos+3, i.e. line indices cpos .. cpos+7
    const size_t o = (size_t)x - LAP + (size_t)du.pitch * ((size_t)(y - LAP) + (size_t)k.ny * (size_t)(z - LAP));
    const size_t vstride = (size_t)du.pitch * k.ny * k.nz;
    // one-sided closures at closed domain faces: global face indices of this cell's faces
    const int n_dom = DIR == 0 ? k.nx : DIR == 1 ? k.ny : k.nz;
    const int g0 = box.s[DIR] - LAP;                 // global index of the box's first cell
    const int fl = g0 + cpos, fr = fl + 1;
    const bool lo = k.bnd[2 * DIR] != 0;
    const int hi = k.bnd[2 * DIR + 1];   // the high face's closure form (0: none)
    const bool skipP = lo && k.nonref[2 * DIR] && fl == 0;
    const bool skipM = hi && k.nonref[2 * DIR + 1] && fl == n_dom - 1;
#pragma unroll 1
    for (int m = 0; m < 5; ++m) {
        const double *q = scratch + (size_t)m * plane + line;
        auto cell = [&](int cg) { return q[cg - g0 + LAP]; };
        double hr, hl;
        if (!(k.bnd_any[DIR] && closure_p(cell, n_dom, fr, lo, hi, hr))) hr = weno_ref<false>(q + cpos + 1);
        if (!(k.bnd_any[DIR] && closure_p(cell, n_dom, fl, lo, hi, hl))) hl = weno_ref<false>(q + cpos);
        double *dp = du.ptr + o + (size_t)m * vstride;
        if (!skipP) *dp = *dp + -aj * (hr - hl) / h;
    }
#pragma unroll 1
    for (int m = 0; m < 5; ++m) {
        const double *q = scratch + (size_t)(5 + m) * plane + line;
        auto cell = [&](int cg) { return q[cg - g0 + LAP]; };
        double hr, hl;
        if (!(k.bnd_any[DIR] && closure_m(cell, n_dom, fr, lo, hi, hr))) hr = weno_ref<true>(q + cpos + 1);
        if (!(k.bnd_any[DIR] && closure_m(cell, n_dom, fl, lo, hi, hl))) hl = weno_ref<true>(q + cpos);
        double *dp = du.ptr + o + (size_t)m * vstride;
        if (!skipM) *dp = *dp + -aj * (hr - hl) / h;
    }
}

__global__ void k_zero_box(ws_soa du, int ny, int nz, Box box)
{
    const long total = (long)box.n[0] * box.n[1] * box.n[2];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int x = box.s[0] - LAP + (int)(i % box.n[0]);
    const int y = box.s[1] - LAP + (int)((i / box.n[0]) % box.n[1]);
    const int z = box.s[2] - LAP + (int)(i / ((long)box.n[0] * box.n[1]));
    for (int m = 0; m < 5; ++m) du.ptr[sidx(du.pitch, ny, nz, x, y, z, m)] = 0.0;
}

}  // namespace

// The context's split-flux scratch of the anchor kernels: 10 doubles per point of the
// largest directional sweep of the whole interior, so that every job box of the context
// fits and the buffer never grows (a graph captured around an anchor RHS keeps its
// pointer valid until ws_ctx_destroy).  Allocated by the first anchor RHS of the context,
// which must not run under stream capture.
static int anchor_scratch(Ctx &c, hipStream_t s, double **out)
{
    if (!c.anchor_scratch) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
            return set_error(WS_ERR_UNSUPPORTED,
                             "anchor RHS: its scratch is allocated by the context's first anchor call, which "
                             "must run outside a stream capture");
        const size_t n[3] = {(size_t)c.kp.nx, (size_t)c.kp.ny, (size_t)c.kp.nz};
        size_t need = 0;
        for (int dir = 0; dir < 3; ++dir) {
            size_t pts = n[dir] + 2 * LAP;
            for (int ax = 0; ax < 3; ++ax)
                if (ax != dir) pts *= n[ax];
            need = need > pts ? need : pts;
        }
        need *= 10 * sizeof(double);
        hipError_t e = hipMalloc(&c.anchor_scratch, need);
        if (e != hipSuccess) { c.anchor_scratch = nullptr; return hip_error(e, "anchor: scratch allocation"); }
        c.anchor_scratch_bytes = need;
    }
    *out = c.anchor_scratch;
    return WS_OK;
}

int launch_rhs_anchor(Ctx &c, ws_field d, ws_field u, ws_field v, ws_field w, ws_field cc,
                      const ws_metric_fields *m, ws_soa du, ws_job job, int accumulate, hipStream_t s)
{
    Box box;
    box.s[0] = (int)job.start.x; box.s[1] = (int)job.start.y; box.s[2] = (int)job.start.z;
    box.n[0] = (int)(job.end.x - job.start.x);
    box.n[1] = (int)(job.end.y - job.start.y);
    box.n[2] = (int)(job.end.z - job.start.z);
    double *scratch = nullptr;
    int r = anchor_scratch(c, s, &scratch);
    if (r) return r;
    ws_metric_fields mf{};
    const int fields = m != nullptr;
    if (m) mf = *m;
    const int T = 256;
    if (!accumulate) {
        long total = (long)box.n[0] * box.n[1] * box.n[2];
        hipLaunchKernelGGL(k_zero_box, dim3((unsigned)((total + T - 1) / T)), dim3(T), 0, s, du, c.kp.ny, c.kp.nz, box);
    }
#define WS_ANCHOR_DIR(DIR)                                                                                  \
    {                                                                                                       \
        long pts = (long)(box.n[DIR] + 2 * LAP) * box.n[DIR == 0 ? 1 : 0] * box.n[DIR == 2 ? 1 : 2];        \
        long cells = (long)box.n[0] * box.n[1] * box.n[2];                                                  \
        hipLaunchKernelGGL(k_anchor_split<DIR>, dim3((unsigned)((pts + T - 1) / T)), dim3(T), 0, s, c.kp, d, \
                           u, v, w, cc, mf, fields, box, scratch);                                        \
        hipLaunchKernelGGL(k_anchor_weno<DIR>, dim3((unsigned)((cells + T - 1) / T)), dim3(T), 0, s, c.kp,  \
                           mf, fields, box, (const double *)scratch, du);                                 \
    }
    WS_ANCHOR_DIR(0)
    WS_ANCHOR_DIR(1)
    WS_ANCHOR_DIR(2)
#undef WS_ANCHOR_DIR
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "anchor RHS launch");
    return WS_OK;
}

}  // namespace ws
