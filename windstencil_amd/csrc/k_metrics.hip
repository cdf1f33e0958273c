// k_metrics.hip -- metric terms from mesh coordinates (the reference's Comput_Jacobian3d /
// comput_Jac3d, src/solver/OCFD_Comput_Jacobian3d.cpp:215-429), run once at set-up.
//
// One kernel, one thread per interior cell: the nine derivatives of the coordinates
// x, y, z along the three index directions (sixth-order central differences over the
// caller's ghost layers; with jacbound[d] the one-sided and low-order forms on the four
// cells next to each face of direction d), the cofactors Akx..Asz of d(x,y,z)/d(i,j,k) and
// Ajac = 1/det; the cell also writes its value into the LAP ghost cells of a face it
// touches.  A periodic ghost fill then replaces the ghosts of periodic directions.
// Expressions in the reference's order with contraction off: bit-identical to the CPU
// restatement.
#include <hip/hip_runtime.h>

#include "ws_device.h"

namespace ws {

namespace {

constexpr int MT = 128;

struct MetricParams {
    int n[3];
    int nyg;
    double h[3];
    int jacbound[3];
};

struct Coords {
    ws_field f[3];
};

struct MetricOut {
    ws_field f[10];          // Akx, Aky, Akz, Aix, Aiy, Aiz, Asx, Asy, Asz, Ajac
};

// derivative of ghosted field f along direction d at interior cell (i, j, k) (OCFD_dx0_CD6_kernel,
// src/kernels/OCFD_Schemes.cpp:1229-1297; boundary forms OCFD_Dx0_bound_kernel_m/_p and the
// y/z copies, src/boundary/OCFD_bound_Scheme.cpp:268-390)
__device__ __forceinline__ double metric_deriv(const MetricParams &p, const ws_field &f, int d, int i, int j, int k)
{
#pragma clang fp contract(off)
    const int n = p.n[d];
    const int c = d == 0 ? i : d == 1 ? j : k;
    const double h = p.h[d];
    auto fv = [&](int o) {
        const int x = (d == 0 ? o : i) + LAP, y = (d == 1 ? o : j) + LAP, z = (d == 2 ? o : k) + LAP;
        return f.ptr[fidx(f.pitch, p.nyg, x, y, z)];
    };
    if (p.jacbound[d]) {
        const int t = n - 1;
        if (c == 0) return (fv(1) - fv(0)) / h;
        if (c == 1) return (fv(2) - fv(0)) * 0.5 / h;
        if (c == 2) return (fv(0) - 8.0 * fv(1) + 8.0 * fv(3) - fv(4)) / (12.0 * h);
        if (c == 3) return (fv(6) - fv(0) - 9.0 * (fv(5) - fv(1)) + 45.0 * (fv(4) - fv(2))) / (60.0 * h);
        if (c == t) return (fv(t) - fv(t - 1)) / h;
        if (c == t - 1) return (fv(t) - fv(t - 2)) * 0.5 / h;
        if (c == t - 2) return (fv(t - 4) - 8.0 * fv(t - 3) + 8.0 * fv(t - 1) - fv(t)) / (12.0 * h);
        if (c == t - 3)
            return (fv(t) - fv(t - 6) - 9.0 * (fv(t - 1) - fv(t - 5)) + 45.0 * (fv(t - 2) - fv(t - 4))) / (60.0 * h);
    }
    return (45.0 * (fv(c + 1) - fv(c - 1)) - 9.0 * (fv(c + 2) - fv(c - 2)) + (fv(c + 3) - fv(c - 3))) / (60.0 * h);
}

__global__ void __launch_bounds__(MT) k_comput_jacobian(MetricParams p, Coords xyz, MetricOut out)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * MT + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= p.n[0]) return;
    double dv[3][3];          // dv[c][d]: derivative of coordinate c along direction d
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) dv[c][d] = metric_deriv(p, xyz.f[c], d, i, j, k);
    const double xi1 = dv[0][0], xj1 = dv[0][1], xk1 = dv[0][2];
    const double yi1 = dv[1][0], yj1 = dv[1][1], yk1 = dv[1][2];
    const double zi1 = dv[2][0], zj1 = dv[2][1], zk1 = dv[2][2];
    double v[10];
    v[9] = 1.0 / (xi1 * yj1 * zk1 + yi1 * zj1 * xk1 + zi1 * xj1 * yk1 - zi1 * yj1 * xk1 - yi1 * xj1 * zk1 -
                  xi1 * zj1 * yk1);
    v[0] = (yj1 * zk1 - zj1 * yk1);
    v[1] = (zj1 * xk1 - xj1 * zk1);
    v[2] = (xj1 * yk1 - yj1 * xk1);
    v[3] = (yk1 * zi1 - zk1 * yi1);
    v[4] = (zk1 * xi1 - xk1 * zi1);
    v[5] = (xk1 * yi1 - yk1 * xi1);
    v[6] = (yi1 * zj1 - zi1 * yj1);
    v[7] = (zi1 * xj1 - xi1 * zj1);
    v[8] = (xi1 * yj1 - yi1 * xj1);
    const int x = i + LAP, y = j + LAP, z = k + LAP;
#pragma unroll
    for (int m = 0; m < 10; ++m) {
        const ws_field &o = out.f[m];
        auto put = [&](int gx, int gy, int gz) { o.ptr[fidx(o.pitch, p.nyg, gx, gy, gz)] = v[m]; };
        put(x, y, z);
        // the face ghosts repeat the adjacent interior cell
        for (int g = 1; g <= LAP; ++g) {
            if (i == 0) put(x - g, y, z);
            if (i == p.n[0] - 1) put(x + g, y, z);
            if (j == 0) put(x, y - g, z);
            if (j == p.n[1] - 1) put(x, y + g, z);
            if (k == 0) put(x, y, z - g);
            if (k == p.n[2] - 1) put(x, y, z + g);
        }
    }
}

}  // namespace

int launch_comput_jacobian(const Ctx &c, const ws_field *coords, const ws_metric_fields &mf, const int32_t *jacbound,
                           hipStream_t s)
{
    const ws_config &cfg = c.cfg;
    MetricParams p{};
    p.n[0] = cfg.nx; p.n[1] = cfg.ny; p.n[2] = cfg.nz;
    p.nyg = cfg.ny + 2 * LAP;
    p.h[0] = cfg.hx; p.h[1] = cfg.hy; p.h[2] = cfg.hz;
    for (int d = 0; d < 3; ++d) p.jacbound[d] = jacbound ? jacbound[d] : 0;
    Coords xyz;
    for (int q = 0; q < 3; ++q) xyz.f[q] = coords[q];
    MetricOut out;
    for (int q = 0; q < 9; ++q) out.f[q] = mf.A[q];
    out.f[9] = mf.ajac;
    const dim3 g((unsigned)((cfg.nx + MT - 1) / MT), (unsigned)cfg.ny, (unsigned)cfg.nz);
    hipLaunchKernelGGL(k_comput_jacobian, g, dim3(MT), 0, s, p, xyz, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "metric launch");
    // periodic directions: the exchange of the metric terms (Comput_Jacobian3d)
    return launch_ghost_fill(c, out.f, 10, s);
}

}  // namespace ws
