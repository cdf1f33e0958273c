// restart.cpp -- restart files in the reference's format (OCFD%08d.dat) and its
// metric file (OCFD3d-Jacobi.dat).
//
// Layout (src/io/OCFD_IO.c:153-226 writes it, :24-90 reads it; the per-plane records
// are written by src/io/OCFD_IO_mpi.c:605-701 and read by :563-604), native byte order:
//   header record:  int32 12 | int32 Istep | float64 tt | int32 12
//   then rho, u, v, w, T, each as NZ records, one per z plane (z ascending):
//                   int32 NX*NY*8 | NX*NY float64 (x fastest, then y) | int32 NX*NY*8
// i.e. Fortran sequential unformatted records.  Only the interior is stored.
// OCFD3d-Jacobi.dat (read by Init_Jacobian3d, src/solver/OCFD_Comput_Jacobian3d.cpp:138-205)
// has no header: x, y, z, Akx, Aky, Akz, Aix, Aiy, Aiz, Asx, Asy, Asz, Ajac as NZ plane
// records each; OCFD3d-Mesh.dat (read there when no Jacobi file exists, :120-135) holds only
// x, y, z, the same way.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ws_internal.h"

using namespace ws;

namespace {

struct File {
    FILE *f = nullptr;
    ~File() { if (f) std::fclose(f); }
};

bool put_i32(FILE *f, int32_t v) { return std::fwrite(&v, sizeof(v), 1, f) == 1; }

bool get_i32(FILE *f, int32_t *v) { return std::fread(v, sizeof(*v), 1, f) == 1; }

// nf fields of nz plane records each
bool write_planes(FILE *f, int32_t nx, int32_t ny, int32_t nz, int nf, const double *const *fields)
{
    const size_t plane = (size_t)nx * ny;
    const int32_t rec = (int32_t)(plane * sizeof(double));
    bool ok = true;
    for (int m = 0; m < nf && ok; ++m)
        for (int z = 0; z < nz && ok; ++z)
            ok = put_i32(f, rec) && std::fwrite(fields[m] + (size_t)z * plane, sizeof(double), plane, f) == plane &&
                 put_i32(f, rec);
    return ok;
}

int read_planes(FILE *f, int32_t nx, int32_t ny, int32_t nz, int nf, double *const *fields)
{
    const size_t plane = (size_t)nx * ny;
    const int32_t rec = (int32_t)(plane * sizeof(double));
    int32_t a = 0, b = 0;
    for (int m = 0; m < nf; ++m)
        for (int z = 0; z < nz; ++z) {
            if (!get_i32(f, &a) || a != rec)
                return set_error(WS_ERR_INVALID, "plane record does not match nx*ny of this grid");
            if (std::fread(fields[m] + (size_t)z * plane, sizeof(double), plane, f) != plane || !get_i32(f, &b) ||
                b != rec)
                return set_error(WS_ERR_INVALID, "truncated plane record");
        }
    // the file must end here: a longer file belongs to a larger grid
    if (std::fgetc(f) != EOF) return set_error(WS_ERR_INVALID, "file longer than this grid's data");
    return WS_OK;
}

int check_extents(int32_t nx, int32_t ny, int32_t nz)
{
    if (nx <= 0 || ny <= 0 || nz <= 0) return set_error(WS_ERR_INVALID, "restart: bad extents");
    if ((long long)nx * ny * 8 > 0x7fffffffLL) return set_error(WS_ERR_INVALID, "restart: plane too large for an int32 record marker");
    return WS_OK;
}

// interior of a ghosted device field <-> contiguous host array (x fastest)
int field_to_host(const ws_field &fld, int nx, int ny, int nz, double *host)
{
    const size_t nyg = (size_t)ny + 2 * LAP;
    for (int z = 0; z < nz; ++z) {
        const double *src = fld.ptr + LAP + (size_t)fld.pitch * (LAP + nyg * (size_t)(z + LAP));
        hipError_t e = hipMemcpy2D(host + (size_t)z * nx * ny, (size_t)nx * sizeof(double), src,
                                   (size_t)fld.pitch * sizeof(double), (size_t)nx * sizeof(double), ny,
                                   hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_error(e, "restart: device to host copy");
    }
    return WS_OK;
}

int host_to_field(const double *host, int nx, int ny, int nz, const ws_field &fld)
{
    const size_t nyg = (size_t)ny + 2 * LAP;
    for (int z = 0; z < nz; ++z) {
        double *dst = fld.ptr + LAP + (size_t)fld.pitch * (LAP + nyg * (size_t)(z + LAP));
        hipError_t e = hipMemcpy2D(dst, (size_t)fld.pitch * sizeof(double), host + (size_t)z * nx * ny,
                                   (size_t)nx * sizeof(double), (size_t)nx * sizeof(double), ny,
                                   hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_error(e, "restart: host to device copy");
    }
    return WS_OK;
}

}  // namespace

extern "C" {

int ws_restart_name(int32_t step, char *buf, int32_t buflen)
{
    if (!buf || buflen <= 0 || step < 0) return set_error(WS_ERR_INVALID, "restart name: bad arguments");
    const int n = std::snprintf(buf, (size_t)buflen, "OCFD%08d.dat", step);
    if (n < 0 || n >= buflen) return set_error(WS_ERR_INVALID, "restart name: buffer too small");
    return WS_OK;
}

int ws_restart_write_host(const char *path, int32_t nx, int32_t ny, int32_t nz, int32_t istep, double tt,
                          const double *const *fields)
{
    if (!path || !fields) return set_error(WS_ERR_INVALID, "restart write: null argument");
    int r = check_extents(nx, ny, nz);
    if (r) return r;
    for (int m = 0; m < 5; ++m)
        if (!fields[m]) return set_error(WS_ERR_INVALID, "restart write: null field");
    File fh;
    fh.f = std::fopen(path, "wb");
    if (!fh.f) return set_error(WS_ERR_INVALID, std::string("restart write: cannot open ") + path);
    const int32_t hdr = (int32_t)(sizeof(int32_t) + sizeof(double));
    bool ok = put_i32(fh.f, hdr) && put_i32(fh.f, istep) && std::fwrite(&tt, sizeof(tt), 1, fh.f) == 1 &&
              put_i32(fh.f, hdr);
    ok = ok && write_planes(fh.f, nx, ny, nz, 5, fields);
    if (!ok) return set_error(WS_ERR_INVALID, std::string("restart write: short write to ") + path);
    if (std::fclose(fh.f) != 0) { fh.f = nullptr; return set_error(WS_ERR_INVALID, "restart write: close failed"); }
    fh.f = nullptr;
    return WS_OK;
}

int ws_restart_read_host(const char *path, int32_t nx, int32_t ny, int32_t nz, int32_t *istep, double *tt,
                         double *const *fields)
{
    if (!path || !fields || !istep || !tt) return set_error(WS_ERR_INVALID, "restart read: null argument");
    int r = check_extents(nx, ny, nz);
    if (r) return r;
    for (int m = 0; m < 5; ++m)
        if (!fields[m]) return set_error(WS_ERR_INVALID, "restart read: null field");
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return set_error(WS_ERR_INVALID, std::string("restart read: cannot open ") + path);
    const int32_t hdr = (int32_t)(sizeof(int32_t) + sizeof(double));
    int32_t a = 0, b = 0;
    if (!get_i32(fh.f, &a) || !get_i32(fh.f, istep) || std::fread(tt, sizeof(*tt), 1, fh.f) != 1 ||
        !get_i32(fh.f, &b) || a != hdr || b != hdr)
        return set_error(WS_ERR_INVALID, "restart read: bad header record");
    if ((r = read_planes(fh.f, nx, ny, nz, 5, fields)))
        return set_error(r, std::string("restart read: ") + ws_last_error());
    return WS_OK;
}

int ws_save_restart(ws_ctx *c, const ws_state *st, int32_t istep, double tt, const char *path, hipStream_t s)
{
    if (!c || !st || !path) return set_error(WS_ERR_INVALID, "restart save: null argument");
    const int nx = c->cfg.nx, ny = c->cfg.ny, nz = c->cfg.nz;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(e, "restart save: stream sync");
    const size_t n = (size_t)nx * ny * nz;
    std::vector<double> host(5 * n);
    const ws_field fl[5] = {st->d, st->u, st->v, st->w, st->T};
    for (int m = 0; m < 5; ++m) {
        if (!fl[m].ptr || fl[m].pitch < (uint32_t)(nx + 2 * LAP)) return set_error(WS_ERR_INVALID, "restart save: bad field");
        int r = field_to_host(fl[m], nx, ny, nz, host.data() + m * n);
        if (r) return r;
    }
    const double *p[5] = {host.data(), host.data() + n, host.data() + 2 * n, host.data() + 3 * n, host.data() + 4 * n};
    return ws_restart_write_host(path, nx, ny, nz, istep, tt, p);
}

int ws_load_restart(ws_ctx *c, ws_state *st, const char *path, int32_t *istep, double *tt, hipStream_t s)
{
    if (!c || !st || !path || !istep || !tt) return set_error(WS_ERR_INVALID, "restart load: null argument");
    const int nx = c->cfg.nx, ny = c->cfg.ny, nz = c->cfg.nz;
    const size_t n = (size_t)nx * ny * nz;
    std::vector<double> host(5 * n);
    double *p[5] = {host.data(), host.data() + n, host.data() + 2 * n, host.data() + 3 * n, host.data() + 4 * n};
    int r = ws_restart_read_host(path, nx, ny, nz, istep, tt, p);
    if (r) return r;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(e, "restart load: stream sync");
    const ws_field fl[5] = {st->d, st->u, st->v, st->w, st->T};
    for (int m = 0; m < 5; ++m) {
        if (!fl[m].ptr || fl[m].pitch < (uint32_t)(nx + 2 * LAP)) return set_error(WS_ERR_INVALID, "restart load: bad field");
        if ((r = host_to_field(p[m], nx, ny, nz, fl[m]))) return r;
    }
    return WS_OK;
}

int ws_jacobi_write_host(const char *path, int32_t nx, int32_t ny, int32_t nz, const double *const *fields)
{
    if (!path || !fields) return set_error(WS_ERR_INVALID, "Jacobi write: null argument");
    int r = check_extents(nx, ny, nz);
    if (r) return r;
    for (int m = 0; m < 13; ++m)
        if (!fields[m]) return set_error(WS_ERR_INVALID, "Jacobi write: null field");
    File fh;
    fh.f = std::fopen(path, "wb");
    if (!fh.f) return set_error(WS_ERR_INVALID, std::string("Jacobi write: cannot open ") + path);
    if (!write_planes(fh.f, nx, ny, nz, 13, fields))
        return set_error(WS_ERR_INVALID, std::string("Jacobi write: short write to ") + path);
    if (std::fclose(fh.f) != 0) { fh.f = nullptr; return set_error(WS_ERR_INVALID, "Jacobi write: close failed"); }
    fh.f = nullptr;
    return WS_OK;
}

int ws_jacobi_read_host(const char *path, int32_t nx, int32_t ny, int32_t nz, double *const *fields)
{
    if (!path || !fields) return set_error(WS_ERR_INVALID, "Jacobi read: null argument");
    int r = check_extents(nx, ny, nz);
    if (r) return r;
    for (int m = 0; m < 13; ++m)
        if (!fields[m]) return set_error(WS_ERR_INVALID, "Jacobi read: null field");
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return set_error(WS_ERR_INVALID, std::string("Jacobi read: cannot open ") + path);
    if ((r = read_planes(fh.f, nx, ny, nz, 13, fields)))
        return set_error(r, std::string("Jacobi read: ") + ws_last_error());
    return WS_OK;
}

int ws_load_jacobi(ws_ctx *c, const char *path, const ws_metric_fields *mf, const ws_field *coords,
                   hipStream_t s)
{
    if (!c || !path || !mf) return set_error(WS_ERR_INVALID, "Jacobi load: null argument");
    const int nx = c->cfg.nx, ny = c->cfg.ny, nz = c->cfg.nz;
    ws_field dst[13];
    for (int m = 0; m < 3; ++m) dst[m] = coords ? coords[m] : ws_field{};
    for (int m = 0; m < 9; ++m) dst[3 + m] = mf->A[m];
    dst[12] = mf->ajac;
    for (int m = 0; m < 13; ++m)
        if ((m >= 3 || coords) && (!dst[m].ptr || dst[m].pitch < (uint32_t)(nx + 2 * LAP)))
            return set_error(WS_ERR_INVALID, "Jacobi load: bad destination field");
    const size_t n = (size_t)nx * ny * nz;
    std::vector<double> host(13 * n);
    double *p[13];
    for (int m = 0; m < 13; ++m) p[m] = host.data() + m * n;
    int r = ws_jacobi_read_host(path, nx, ny, nz, p);
    if (r) return r;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(e, "Jacobi load: stream sync");
    for (int m = coords ? 0 : 3; m < 13; ++m)
        if ((r = host_to_field(p[m], nx, ny, nz, dst[m]))) return r;
    // periodic face ghosts of the metric terms (the reference's exchange_boundary_xyz)
    if ((r = launch_ghost_fill(*c, dst + 3, 10, s))) return r;
    if (coords && (r = launch_ghost_fill(*c, dst, 3, s))) return r;
    return WS_OK;
}

int ws_mesh_write_host(const char *path, int32_t nx, int32_t ny, int32_t nz, const double *const *xyz)
{
    if (!path || !xyz || !xyz[0] || !xyz[1] || !xyz[2]) return set_error(WS_ERR_INVALID, "mesh write: null argument");
    int r = check_extents(nx, ny, nz);
    if (r) return r;
    File fh;
    fh.f = std::fopen(path, "wb");
    if (!fh.f) return set_error(WS_ERR_INVALID, std::string("mesh write: cannot open ") + path);
    if (!write_planes(fh.f, nx, ny, nz, 3, xyz))
        return set_error(WS_ERR_INVALID, std::string("mesh write: short write to ") + path);
    if (std::fclose(fh.f) != 0) { fh.f = nullptr; return set_error(WS_ERR_INVALID, "mesh write: close failed"); }
    fh.f = nullptr;
    return WS_OK;
}

int ws_mesh_read_host(const char *path, int32_t nx, int32_t ny, int32_t nz, double *const *xyz)
{
    if (!path || !xyz || !xyz[0] || !xyz[1] || !xyz[2]) return set_error(WS_ERR_INVALID, "mesh read: null argument");
    int r = check_extents(nx, ny, nz);
    if (r) return r;
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return set_error(WS_ERR_INVALID, std::string("mesh read: cannot open ") + path);
    if ((r = read_planes(fh.f, nx, ny, nz, 3, xyz))) return set_error(r, std::string("mesh read: ") + ws_last_error());
    return WS_OK;
}

static int check_metric_args(const ws_ctx *c, const ws_field *coords, const ws_metric_fields *mf,
                             const int32_t *jacbound)
{
    if (!c || !coords || !mf) return set_error(WS_ERR_INVALID, "metrics: null argument");
    const uint32_t need = (uint32_t)(c->cfg.nx + 2 * LAP);
    for (int q = 0; q < 3; ++q)
        if (!coords[q].ptr || coords[q].pitch < need) return set_error(WS_ERR_INVALID, "metrics: bad coordinate field");
    for (int q = 0; q < 10; ++q) {
        const ws_field &f = q < 9 ? mf->A[q] : mf->ajac;
        if (!f.ptr || f.pitch < need) return set_error(WS_ERR_INVALID, "metrics: bad metric field");
    }
    const int n[3] = {c->cfg.nx, c->cfg.ny, c->cfg.nz};
    for (int d = 0; d < 3 && jacbound; ++d) {
        if (jacbound[d] != 0 && jacbound[d] != 1) return set_error(WS_ERR_INVALID, "metrics: jacbound entries must be 0 or 1");
        if (jacbound[d] && n[d] < 8)
            return set_error(WS_ERR_INVALID, "metrics: one-sided derivative forms need at least 8 cells");
    }
    return WS_OK;
}

int ws_comput_jacobian(ws_ctx *c, const ws_field *coords, const ws_metric_fields *mf, const int32_t *jacbound,
                       hipStream_t s)
{
    int r = check_metric_args(c, coords, mf, jacbound);
    return r ? r : launch_comput_jacobian(*c, coords, *mf, jacbound, s);
}

int ws_load_mesh(ws_ctx *c, const char *path, const ws_field *coords, const ws_metric_fields *mf,
                 const int32_t *jacbound, hipStream_t s)
{
    int r = check_metric_args(c, coords, mf, jacbound);
    if (r) return r;
    if (!path) return set_error(WS_ERR_INVALID, "mesh load: null path");
    const int nx = c->cfg.nx, ny = c->cfg.ny, nz = c->cfg.nz;
    const size_t n = (size_t)nx * ny * nz;
    std::vector<double> host(3 * n);
    double *p[3] = {host.data(), host.data() + n, host.data() + 2 * n};
    if ((r = ws_mesh_read_host(path, nx, ny, nz, p))) return r;
    // Init_Jacobian3d: every coordinate and metric value starts at 1.0 (ghosts included)
    ws_field all[13];
    for (int q = 0; q < 3; ++q) all[q] = coords[q];
    for (int q = 0; q < 9; ++q) all[3 + q] = mf->A[q];
    all[12] = mf->ajac;
    for (const ws_field &f : all)
        if ((r = launch_fill(1.0, f.ptr, f.pitch, (uint32_t)(nx + 2 * LAP), (uint32_t)(ny + 2 * LAP),
                             (uint32_t)(nz + 2 * LAP), s)))
            return r;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(e, "mesh load: stream sync");
    for (int q = 0; q < 3; ++q)
        if ((r = host_to_field(p[q], nx, ny, nz, coords[q]))) return r;
    // the coordinates' periodic face halo (exchange_boundary_xyz after the read), then the metrics
    if ((r = launch_ghost_fill(*c, coords, 3, s))) return r;
    return launch_comput_jacobian(*c, coords, *mf, jacbound, s);
}

}  // extern "C"
