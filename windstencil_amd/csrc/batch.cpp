// batch.cpp -- many periodic blocks of uneven size advanced together.
//
// Each block keeps its own context (extents, spacing 1/(n-1)) and its caller-owned
// state.  The inviscid RHS of all blocks runs as one launch per direction over a device
// tile list (block, tile coordinates) built here for every tile shape: one work-group
// per tile, or a persistent grid walking contiguous runs of the list (the batch's
// schedule); the ghost fill, sound speed, fused RK3 stage and CFL step run as one launch
// each over all blocks (a table of per-block work-group ranges), per block where that
// table is not available (a first call under a stream capture).  A
// batch step uses one dt for every block: the minimum of the blocks' CFL steps.  A
// whole SSP-RK3 step (CFL reduction included) can be captured and replayed as one
// hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "ws_internal.h"
#include "ws_tiles.h"

using namespace ws;

struct ws_batch {
    ws_config base;
    int precision = WS_PREC_FP64;
    int kernel = WS_KERNEL_TILED;
    int tile_shape = WS_TILE_30x16_K2;
    int schedule = WS_SCHED_PREFETCH;
    // which launch of a stage-applying RHS applies the stage (WS_ORDER_*: the last
    // direction; ws_batch_autotune picks)
    int stage_order = WS_ORDER_XYZ;
    // and its tile shape (-1: tile_shape)
    int stage_shape = -1;
    // KParams::off32 for the RHS inputs of d_blocks (every block's own fields) and of
    // d_blocks2 (the second primitive sets): every block's five fields pass prim_off32
    int off32_own = 0, off32_second = 0;
    std::vector<ws_ctx *> ctx;
    std::vector<ws_state> st;
    BatchBlock *d_blocks = nullptr;
    // the same table with the primitive inputs taken from each block's second primitive set
    // and the outputs (od..occ) the block's own fields; d_blocks' outputs are the second set
    BatchBlock *d_blocks2 = nullptr;
    std::vector<BatchBlock> host_blocks;
    std::vector<double *> fas, fbs;   // each block's f and fn as created (the table's fa, fb)
    double *fb0 = nullptr;
    int4 *d_tiles[WS_TILE_COUNT][3] = {};
    int ntiles[WS_TILE_COUNT][3] = {};
    double *d_dts = nullptr;    // per-block CFL steps
    double *d_dt = nullptr;     // the batch step (their minimum)
    // the table of the batched helper kernels (built at the first use outside a capture;
    // aux_second: its second field sets are the blocks' second primitive sets)
    AuxBlock *d_aux = nullptr;
    bool aux_second = false;
    long ghost_wgs = 0, cell_wgs = 0;
    double *d_cfl_part = nullptr;   // the CFL maxima of the batched stage kernel (cell_wgs)
    bool cfl_in_batch = false;      // the last step's stage kernel wrote them there
    // per tile shape and stage order: each block's range of the tile list of the
    // stage-applying launch (z's for WS_ORDER_XYZ, x's for WS_ORDER_ZYX)
    int2 *d_tile_ranges[WS_TILE_COUNT][WS_ORDER_COUNT] = {};
    // the CFL maxima per tile of stage fusion 3's stage-3 launch (the longest x or z list)
    double *d_cfl_tiles = nullptr;
    hipStream_t capture_stream = nullptr;
    long long cells = 0;
};

namespace {

int fail(ws_batch *b, int r)
{
    ws_batch_destroy(b);
    return r;
}

// tiles of one block for one direction and shape, in launch order (bx fastest)
void block_tiles(int b, const int n[3], int shape, int dir, std::vector<int4> &out)
{
    const TileShapeDef &D = kTileShapes[shape];
    const int L = dir == 0 ? D.L0 : dir == 1 ? D.L1 : D.L2, W = dir == 0 ? D.W0 : dir == 1 ? D.W1 : D.W2;
    const int xa = tile_xalign(dir, W);
    // the box starts at LAP; aligned y/z tiles start at the aligned column below it
    const int nxa = n[0] + LAP - LAP / xa * xa;
    int gx, gy, gz;
    if (dir == 0) { gx = (n[0] + L - 1) / L; gy = (n[1] + W - 1) / W; gz = n[2]; }
    else if (dir == 1) { gx = (nxa + W - 1) / W; gy = (n[1] + L - 1) / L; gz = n[2]; }
    else { gx = (nxa + W - 1) / W; gy = (n[2] + L - 1) / L; gz = n[1]; }
    for (int z = 0; z < gz; ++z)
        for (int y = 0; y < gy; ++y)
            for (int x = 0; x < gx; ++x) out.push_back(make_int4(b, x, y, z));
}

// the tile shape of the batch's stage-applying RHS
int stage_shape_of(const ws_batch *b) { return b->stage_shape >= 0 ? b->stage_shape : b->tile_shape; }

// the launch tables of the batch's RHS, or with `stage` of its stage-applying RHS
BatchTiles tiles_of(const ws_batch *b, bool stage = false)
{
    BatchTiles t;
    t.shape = stage ? stage_shape_of(b) : b->tile_shape;
    t.schedule = b->schedule;
    t.blocks = b->d_blocks;
    t.stage = nullptr;
    t.stage_order = b->stage_order;
    for (int d = 0; d < 3; ++d) {
        t.tiles[d] = b->d_tiles[t.shape][d];
        t.ntiles[d] = b->ntiles[t.shape][d];
    }
    return t;
}

int batch_rhs(ws_batch *b, hipStream_t s, const BatchStage *stage = nullptr, const BatchBlock *table = nullptr)
{
    if (b->kernel == WS_KERNEL_ANCHOR) {
        // the reference-order kernels have no batched form: one RHS per block
        for (size_t i = 0; i < b->ctx.size(); ++i) {
            int r = rhs_for_ctx(b->ctx[i], &b->st[i], s);
            if (r) return r;
        }
        return WS_OK;
    }
    KParams k = b->ctx[0]->kp;   // physics constants and metrics are the batch's
    k.off32 = table && table == b->d_blocks2 ? b->off32_second : b->off32_own;
    BatchTiles t = tiles_of(b, stage != nullptr);
    t.stage = stage;
    if (table) t.blocks = table;
    return b->precision == WS_PREC_MIXED ? launch_rhs_batch_mixed(k, t, s) : launch_rhs_batch_tiled(k, t, s);
}

// Both block tables for RK3 steps whose stage-applying launch also derives the primitives:
// the blocks' second primitive sets (their contexts allocate them) become the outputs of
// d_blocks and the inputs of d_blocks2.  Built once, outside a stream capture.
bool prim_tables(ws_batch *b, hipStream_t s)
{
    if (b->d_blocks2) return true;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
    const size_t n = b->host_blocks.size();
    std::vector<BatchBlock> a = b->host_blocks, c = b->host_blocks;
    int second32 = 1;
    for (size_t i = 0; i < n; ++i) {
        ws_field sec[6];
        if (!prim_set(b->ctx[i], s, sec)) return false;
        if (!prim_off32(b->ctx[i]->kp, sec[0], sec[1], sec[2], sec[3], sec[5])) second32 = 0;
        const ws_state &st = b->st[i];
        a[i].od = sec[0]; a[i].ou = sec[1]; a[i].ov = sec[2]; a[i].ow = sec[3]; a[i].oT = sec[4]; a[i].occ = sec[5];
        c[i].d = sec[0]; c[i].u = sec[1]; c[i].v = sec[2]; c[i].w = sec[3]; c[i].cc = sec[5];
        c[i].od = st.d; c[i].ou = st.u; c[i].ov = st.v; c[i].ow = st.w; c[i].oT = st.T; c[i].occ = st.cc;
    }
    BatchBlock *d2 = nullptr;
    if (hipMalloc(&d2, sizeof(BatchBlock) * n) != hipSuccess) return false;
    if (hipStreamSynchronize(s) != hipSuccess ||
        hipMemcpy(d2, c.data(), sizeof(BatchBlock) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_blocks, a.data(), sizeof(BatchBlock) * n, hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(d2);
        return false;
    }
    b->host_blocks = a;
    b->d_blocks2 = d2;
    b->off32_second = second32;
    return true;
}

// The table of the batched helper kernels (ghost fill, sound speed, stage, CFL step: one
// launch each for every block), with the blocks' second primitive sets once prim_tables
// has made them.  Built or refreshed outside a stream capture; false if it is not there.
bool aux_table(ws_batch *b, hipStream_t s)
{
    const bool second = b->d_blocks2 != nullptr;
    if (b->d_aux && b->aux_second == second) return true;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return b->d_aux != nullptr;
    const size_t n = b->ctx.size();
    std::vector<AuxBlock> t(n);
    long gw = 0, cw = 0;
    for (size_t i = 0; i < n; ++i) {
        const ws_state &st = b->st[i];
        const BatchBlock &h = b->host_blocks[i];
        const KParams &k = b->ctx[i]->kp;
        AuxBlock &a = t[i];
        std::memset(&a, 0, sizeof(a));
        const ws_field s0[6] = {st.d, st.u, st.v, st.w, st.T, st.cc};
        const ws_field s1[6] = {h.od, h.ou, h.ov, h.ow, h.oT, h.occ};
        for (int q = 0; q < 6; ++q) {
            a.g[0][q] = s0[q];
            a.g[1][q] = second ? s1[q] : s0[q];
        }
        a.P = st.P;
        a.fa = h.fa;
        a.fb = h.fb;
        a.n[0] = k.nx; a.n[1] = k.ny; a.n[2] = k.nz;
        a.nyg = k.nyg;
        a.hmin = std::fmin(k.h[0], std::fmin(k.h[1], k.h[2]));
        const long ghosts = 2L * LAP * ((long)k.ny * k.nz + (long)k.nx * k.nz + (long)k.nx * k.ny);
        const long cells = (long)k.nx * k.ny * k.nz;
        a.ghost_wg0 = gw; gw += (ghosts + 255) / 256; a.ghost_wg1 = gw;
        a.cell_wg0 = cw; cw += (cells + 255) / 256; a.cell_wg1 = cw;
    }
    if (gw > 0x7fffffffL || cw > 0x7fffffffL) return false;
    if (!b->d_aux) {
        if (hipMalloc(&b->d_aux, sizeof(AuxBlock) * n) != hipSuccess) { b->d_aux = nullptr; return false; }
        if (hipMalloc(&b->d_cfl_part, sizeof(double) * cw) != hipSuccess) {
            hipFree(b->d_aux);
            b->d_aux = nullptr;
            b->d_cfl_part = nullptr;
            return false;
        }
    }
    // a refresh overwrites a table earlier launches may still read
    if (hipStreamSynchronize(s) != hipSuccess ||
        hipMemcpy(b->d_aux, t.data(), sizeof(AuxBlock) * n, hipMemcpyHostToDevice) != hipSuccess)
        return false;
    b->aux_second = second;
    b->ghost_wgs = gw;
    b->cell_wgs = cw;
    return true;
}

// ghost fill and sound speed of every block (the batched form when the table is there)
int batch_prepare(ws_batch *b, hipStream_t s)
{
    if (aux_table(b, s))
        return launch_prepare_batch(b->ctx[0]->kp, b->d_aux, (int)b->ctx.size(), b->ghost_wgs, b->cell_wgs, s);
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        int r = prepare_rhs_inputs(b->ctx[i], &b->st[i], s);
        if (r) return r;
    }
    return WS_OK;
}

}  // namespace

extern "C" {

int ws_batch_create(const ws_config *base, int32_t nblocks, const int32_t *extents, const ws_state *states,
                    ws_batch **out)
{
    if (!base || !extents || !states || !out || nblocks <= 0)
        return set_error(WS_ERR_INVALID, "bad batch request");
    for (int i = 0; i < 3; ++i)
        if (base->periodic[i] != 1 || base->d0_bound[2 * i] || base->d0_bound[2 * i + 1])
            return set_error(WS_ERR_UNSUPPORTED, "batches hold periodic blocks only");
    *out = nullptr;
    ws_batch *b = new (std::nothrow) ws_batch;
    if (!b) return set_error(WS_ERR_NOMEM, "batch allocation failed");
    b->base = *base;
    b->precision = base->precision;
    b->kernel = base->kernel;
    std::vector<BatchBlock> host(nblocks);
    for (int i = 0; i < nblocks; ++i) {
        ws_config cfg = *base;
        cfg.nx = extents[3 * i]; cfg.ny = extents[3 * i + 1]; cfg.nz = extents[3 * i + 2];
        if (cfg.nx < 2 || cfg.ny < 2 || cfg.nz < 2) return fail(b, set_error(WS_ERR_INVALID, "block extent below 2"));
        cfg.hx = 1.0 / (cfg.nx - 1); cfg.hy = 1.0 / (cfg.ny - 1); cfg.hz = 1.0 / (cfg.nz - 1);
        ws_ctx *c = nullptr;
        int r = ws_ctx_create(&cfg, &c);
        if (r) return fail(b, r);
        b->ctx.push_back(c);
        if ((r = check_state(c, &states[i]))) return fail(b, r);
        b->st.push_back(states[i]);
        BatchBlock &h = host[i];
        std::memset(&h, 0, sizeof(h));
        h.d = states[i].d; h.u = states[i].u; h.v = states[i].v; h.w = states[i].w; h.cc = states[i].cc;
        h.P = states[i].P;
        h.du = states[i].du;
        h.fa = states[i].f;
        h.fb = states[i].fn;
        b->fas.push_back(states[i].f.ptr);
        b->fbs.push_back(states[i].fn.ptr);
        if (i == 0) b->fb0 = states[i].fn.ptr;
        h.n[0] = cfg.nx; h.n[1] = cfg.ny; h.n[2] = cfg.nz;
        h.rh[0] = 1.0 / cfg.hx; h.rh[1] = 1.0 / cfg.hy; h.rh[2] = 1.0 / cfg.hz;
        b->cells += (long long)cfg.nx * cfg.ny * cfg.nz;
    }
    b->off32_own = 1;
    for (int i = 0; i < nblocks; ++i) {
        const ws_state &st = states[i];
        if (!prim_off32(b->ctx[i]->kp, st.d, st.u, st.v, st.w, st.cc)) b->off32_own = 0;
    }
    hipError_t e = hipMalloc(&b->d_blocks, sizeof(BatchBlock) * nblocks);
    if (e != hipSuccess) return fail(b, set_error(WS_ERR_NOMEM, "batch block table"));
    e = hipMemcpy(b->d_blocks, host.data(), sizeof(BatchBlock) * nblocks, hipMemcpyHostToDevice);
    b->host_blocks = host;
    if (e != hipSuccess) return fail(b, hip_error(e, "batch block table upload"));
    for (int shape = 0; shape < WS_TILE_COUNT; ++shape)
        for (int dir = 0; dir < 3; ++dir) {
            std::vector<int4> list;
            std::vector<int2> ranges(nblocks);
            for (int i = 0; i < nblocks; ++i) {
                ranges[i].x = (int)list.size();
                block_tiles(i, host[i].n, shape, dir, list);
                ranges[i].y = (int)list.size();
            }
            if (dir != 1) {
                int2 *&dr = b->d_tile_ranges[shape][dir == 2 ? WS_ORDER_XYZ : WS_ORDER_ZYX];
                e = hipMalloc(&dr, sizeof(int2) * nblocks);
                if (e != hipSuccess) return fail(b, set_error(WS_ERR_NOMEM, "batch tile ranges"));
                e = hipMemcpy(dr, ranges.data(), sizeof(int2) * nblocks, hipMemcpyHostToDevice);
                if (e != hipSuccess) return fail(b, hip_error(e, "batch tile ranges upload"));
            }
            if (list.size() > (size_t)0x7fffffff) return fail(b, set_error(WS_ERR_INVALID, "too many tiles"));
            b->ntiles[shape][dir] = (int)list.size();
            e = hipMalloc(&b->d_tiles[shape][dir], sizeof(int4) * list.size());
            if (e != hipSuccess) return fail(b, set_error(WS_ERR_NOMEM, "batch tile list"));
            e = hipMemcpy(b->d_tiles[shape][dir], list.data(), sizeof(int4) * list.size(), hipMemcpyHostToDevice);
            if (e != hipSuccess) return fail(b, hip_error(e, "batch tile list upload"));
        }
    e = hipMalloc(&b->d_dts, sizeof(double) * nblocks);
    if (e == hipSuccess) e = hipMalloc(&b->d_dt, sizeof(double));
    if (e != hipSuccess) return fail(b, set_error(WS_ERR_NOMEM, "batch dt buffers"));
    e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(b, hip_error(e, "batch capture stream"));
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(b, hip_error(e, "batch setup"));
    *out = b;
    return WS_OK;
}

int ws_batch_destroy(ws_batch *b)
{
    if (!b) return WS_OK;
    hipDeviceSynchronize();
    for (ws_ctx *c : b->ctx) ws_ctx_destroy(c);
    if (b->d_blocks) hipFree(b->d_blocks);
    if (b->d_blocks2) hipFree(b->d_blocks2);
    for (int shape = 0; shape < WS_TILE_COUNT; ++shape)
        for (int dir = 0; dir < 3; ++dir)
            if (b->d_tiles[shape][dir]) hipFree(b->d_tiles[shape][dir]);
    if (b->d_dts) hipFree(b->d_dts);
    if (b->d_dt) hipFree(b->d_dt);
    if (b->d_aux) hipFree(b->d_aux);
    if (b->d_cfl_part) hipFree(b->d_cfl_part);
    for (int shape = 0; shape < WS_TILE_COUNT; ++shape)
        for (int o = 0; o < WS_ORDER_COUNT; ++o)
            if (b->d_tile_ranges[shape][o]) hipFree(b->d_tile_ranges[shape][o]);
    if (b->d_cfl_tiles) hipFree(b->d_cfl_tiles);
    if (b->capture_stream) hipStreamDestroy(b->capture_stream);
    delete b;
    return WS_OK;
}

int ws_batch_block(const ws_batch *b, int32_t i, ws_state *st)
{
    if (!b || !st || i < 0 || i >= (int32_t)b->st.size()) return set_error(WS_ERR_INVALID, "bad block index");
    *st = b->st[i];
    return WS_OK;
}

int64_t ws_batch_cells(const ws_batch *b) { return b ? (int64_t)b->cells : 0; }

int ws_batch_set_precision(ws_batch *b, int32_t precision)
{
    if (!b) return set_error(WS_ERR_INVALID, "null batch");
    if (precision != WS_PREC_FP64 && precision != WS_PREC_MIXED) return set_error(WS_ERR_INVALID, "unknown precision");
    b->precision = precision;
    for (ws_ctx *c : b->ctx) ws_ctx_set_precision(c, precision);
    return WS_OK;
}

int ws_batch_set_kernel(ws_batch *b, int32_t kernel)
{
    if (!b) return set_error(WS_ERR_INVALID, "null batch");
    if (kernel != WS_KERNEL_TILED && kernel != WS_KERNEL_ANCHOR) return set_error(WS_ERR_INVALID, "unknown kernel");
    b->kernel = kernel;
    for (ws_ctx *c : b->ctx) ws_ctx_set_kernel(c, kernel);
    return WS_OK;
}

int ws_batch_set_tile_shape(ws_batch *b, int32_t shape)
{
    if (!b) return set_error(WS_ERR_INVALID, "null batch");
    if (shape < 0 || shape >= WS_TILE_COUNT) return set_error(WS_ERR_INVALID, "unknown tile shape");
    b->tile_shape = shape;
    return WS_OK;
}

int ws_batch_set_stage_fusion(ws_batch *b, int32_t on)
{
    if (!b) return set_error(WS_ERR_INVALID, "null batch");
    // 3 (the contexts' default: alternating primitive sets within ws_rk3_run) acts as 2 here
    if (on < 0 || on > 3) return set_error(WS_ERR_INVALID, "stage fusion must be 0, 1, 2 or 3");
    for (ws_ctx *c : b->ctx) c->fuse_stage = on;
    return WS_OK;
}

int ws_batch_set_stage_order(ws_batch *b, int32_t order)
{
    if (!b) return set_error(WS_ERR_INVALID, "null batch");
    if (order < 0 || order >= WS_ORDER_COUNT) return set_error(WS_ERR_INVALID, "unknown direction order");
    b->stage_order = order;
    return WS_OK;
}

int ws_batch_set_stage_shape(ws_batch *b, int32_t shape)
{
    if (!b) return set_error(WS_ERR_INVALID, "null batch");
    if (shape < -1 || shape >= WS_TILE_COUNT) return set_error(WS_ERR_INVALID, "unknown tile shape");
    b->stage_shape = shape;
    return WS_OK;
}

int ws_batch_get_stage_shape(const ws_batch *b, int32_t *shape)
{
    if (!b || !shape) return set_error(WS_ERR_INVALID, "null argument");
    *shape = b->stage_shape;
    return WS_OK;
}

int ws_batch_get_stage_order(const ws_batch *b, int32_t *order)
{
    if (!b || !order) return set_error(WS_ERR_INVALID, "null argument");
    *order = b->stage_order;
    return WS_OK;
}

int ws_batch_get_tile_shape(const ws_batch *b, int32_t *shape)
{
    if (!b || !shape) return set_error(WS_ERR_INVALID, "null argument");
    *shape = b->tile_shape;
    return WS_OK;
}

int ws_batch_set_schedule(ws_batch *b, int32_t schedule)
{
    if (!b) return set_error(WS_ERR_INVALID, "null batch");
    if (schedule < 0 || schedule >= WS_SCHED_COUNT) return set_error(WS_ERR_INVALID, "unknown schedule");
    b->schedule = schedule;
    return WS_OK;
}

int ws_batch_get_schedule(const ws_batch *b, int32_t *schedule)
{
    if (!b || !schedule) return set_error(WS_ERR_INVALID, "null argument");
    *schedule = b->schedule;
    return WS_OK;
}

int ws_batch_inviscid_rhs(ws_batch *b, hipStream_t s)
{
    if (!b) return set_error(WS_ERR_INVALID, "null batch");
    return batch_rhs(b, s);
}

int ws_batch_du_comput(ws_batch *b, hipStream_t s)
{
    if (!b) return set_error(WS_ERR_INVALID, "null batch");
    int r = batch_prepare(b, s);
    if (!r) r = batch_rhs(b, s);
    if (r || !b->base.viscous) return r;
    // viscous terms, block by block (memory-bound kernels; each block owns its scratch)
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        const ws_state &st = b->st[i];
        if ((r = launch_viscous(*b->ctx[i], st.u, st.v, st.w, st.T, nullptr, st.du, s))) return r;
    }
    return WS_OK;
}

int ws_batch_cfl_dt(ws_batch *b, double cfl, double *dt_dev, hipStream_t s)
{
    if (!b || !dt_dev || !(cfl > 0.0)) return set_error(WS_ERR_INVALID, "bad CFL request");
    for (size_t i = 0; i < b->ctx.size(); ++i) {
        int r = launch_cfl(*b->ctx[i], b->st[i], cfl, b->d_dts + i, s);
        if (r) return r;
    }
    return launch_min_reduce(b->d_dts, (int)b->ctx.size(), dt_dev, s);
}

// One batched RK3 step; cfl_next: the step leaves the blocks' CFL maxima for the next
// step's dt (STEP_PLAIN: its stage kernels, in the batch's buffer or each block's
// cfl_stage_buffer; the alternating modes: its stage-3 launch, per tile in d_cfl_tiles).
// mode and prepare as for a single box's rk3_step_fused (stage fusion 3): an alternating
// step from set A runs stage 1 into B, stage 2 into A and stage 3 into B with the batched
// stage-applying launch writing every block's primitives, sound speed and their periodic
// face halo; the last step of a run from the second sets copies them to the blocks'
// fields and lets the stage-3 launch write the primitives and P over their interior.
static int batch_step(ws_batch *b, const double *dt_dev, hipStream_t s, bool cfl_next, StepMode mode = STEP_PLAIN,
                      bool prepare = true)
{
    if (!b) return set_error(WS_ERR_INVALID, "null batch");
    for (ws_state &st : b->st) std::swap(st.f, st.fn);
    const ws_soa nolap{};
    int r = WS_OK;
    // With the tiled kernels, the grid schedule and no viscous terms the last batched launch
    // applies each stage (as a single box's RK3 step does); the block table holds every
    // block's f and fn as created, and whether they are exchanged now is the same for all
    // blocks (they swap together).
    bool fuse = b->kernel == WS_KERNEL_TILED && !b->base.viscous && b->schedule == WS_SCHED_GRID &&
                b->ctx[0]->fuse_stage;
    const int swapped = b->st[0].f.ptr == b->fb0 ? 1 : 0;
    for (size_t i = 0; i < b->st.size() && fuse; ++i)
        fuse = b->st[i].f.ptr == (swapped ? b->fbs[i] : b->fas[i]);
    // Fusion level 2 (as for a single box): the last launch of stages 1 and 2 also writes the
    // primitives, stage 1 from the blocks' own fields into their second sets (d_blocks),
    // stage 2 back (d_blocks2), with the periodic face halo of the six written fields.
    const bool prims = fuse && b->ctx[0]->fuse_stage > 1 && prim_tables(b, s);
    // the helper kernels of all blocks in one launch each (the table holds the second
    // primitive sets when prims is on: prim_tables runs first and aux_table follows it)
    const bool aux = aux_table(b, s) && (!prims || b->aux_second);
    const int nb = (int)b->ctx.size();
    const bool planes = b->base.bc_planes != 0;
    b->cfl_in_batch = false;
    if (mode != STEP_PLAIN) {
        if (!(fuse && prims && aux) || b->ctx[0]->fuse_stage < 3 || b->kernel != WS_KERNEL_TILED)
            return set_error(WS_ERR_INVALID, "alternating batched RK3 step without stage fusion 3");
        const bool from_second = mode != STEP_FROM_STATE, last = mode == STEP_LAST_FROM_SECOND;
        // d_blocks reads the blocks' fields and writes their second sets, d_blocks2 the reverse
        const BatchBlock *toB = from_second ? b->d_blocks2 : b->d_blocks;
        const BatchBlock *toA = from_second ? b->d_blocks : b->d_blocks2;
        if (prepare && !from_second)
            r = launch_prepare_batch(b->ctx[0]->kp, b->d_aux, nb, b->ghost_wgs, b->cell_wgs, s);
        for (int krk = 1; krk <= 3 && !r; ++krk) {
            BatchStage bs{dt_dev, b->base.dt, krk, swapped};
            bs.prims = 1;
            bs.planes = planes;
            if (krk == 3 && last) {
                bs.cc = 0;
                bs.pressure = 1;
                for (size_t i = 0; i < b->ctx.size() && !r; ++i) {
                    const BatchBlock &h = b->host_blocks[i];
                    const ws_state &st = b->st[i];
                    const ws_field sec[6] = {h.od, h.ou, h.ov, h.ow, h.oT, h.occ};
                    const ws_field own[6] = {st.d, st.u, st.v, st.w, st.T, st.cc};
                    r = launch_copy_fields(*b->ctx[i], sec, own, 6, s);
                }
            } else {
                bs.halo = 1;   // the launch fills the output sets' periodic halos itself
                if (krk == 3 && cfl_next) bs.cfl_part = b->d_cfl_tiles;
            }
            if (!r) r = batch_rhs(b, s, &bs, krk == 2 ? toA : toB);
        }
        return r;
    }
    // the blocks are periodic (ws_batch_create); stages 1 and 2 leave the next stage's
    // inputs prepared (the last launch or the stage kernel writing the sound speed, the halo
    // by that launch itself or a ghost fill after the stage kernel)
    for (int krk = 1; krk <= 3 && !r; ++krk) {
        if (krk == 1) {
            if (aux) r = launch_prepare_batch(b->ctx[0]->kp, b->d_aux, nb, b->ghost_wgs, b->cell_wgs, s);
            for (size_t i = 0; i < b->ctx.size() && !r && !aux; ++i) r = prepare_rhs_inputs(b->ctx[i], &b->st[i], s);
        }
        BatchStage bs{dt_dev, b->base.dt, krk, swapped};
        if (prims && krk < 3) {
            bs.prims = 1;
            bs.planes = planes;
            bs.halo = 1;   // the launch fills the written sets' periodic halos itself
            // stage 1 writes the second sets, stage 2 the state's fields
            if (!r) r = batch_rhs(b, s, &bs, krk == 2 ? b->d_blocks2 : b->d_blocks);
            continue;
        }
        if (!r) r = batch_rhs(b, s, fuse ? &bs : nullptr);
        for (size_t i = 0; i < b->ctx.size() && !r && b->base.viscous; ++i) {
            const ws_state &st = b->st[i];
            r = launch_viscous(*b->ctx[i], st.u, st.v, st.w, st.T, nullptr, st.du, s);
        }
        if (fuse && aux) {
            // the last RHS launch applied the stage: the primitives of every block's f
            const bool last = cfl_next && krk == 3;
            if (!r) r = launch_stage_batch(b->ctx[0]->kp, b->d_aux, nb, b->cell_wgs, swapped, planes, krk < 3,
                                           last ? b->d_cfl_part : nullptr, s);
            if (!r && krk < 3) r = launch_ghost_fill_batch(b->d_aux, nb, b->ghost_wgs, 0, s);
            if (!r && last) b->cfl_in_batch = true;
            continue;
        }
        for (size_t i = 0; i < b->ctx.size() && !r; ++i) {
            ws_ctx *c = b->ctx[i];
            const ws_state &st = b->st[i];
            r = launch_stage(*c, fuse ? 0 : krk, st, nolap, c->cfg.dt, dt_dev, c->cfg.bc_planes != 0, krk < 3, s,
                             cfl_next && krk == 3 ? c->cfl_stage : nullptr);
            if (!r && krk < 3) {
                const ws_field g[6] = {st.d, st.u, st.v, st.w, st.T, st.cc};
                r = launch_ghost_fill(*c, g, 6, s);
            }
        }
    }
    return r;
}

int ws_batch_rk3_step(ws_batch *b, const double *dt_dev, hipStream_t s)
{
    return batch_step(b, dt_dev, s, false);
}

// Whether ws_batch_rk3_run can alternate the primitive sets (stage fusion 3 and every
// condition of the batched primitives-deriving fusion), allocating what that needs.
static bool batch_alternating(ws_batch *b, hipStream_t s, bool cfl)
{
    if (b->kernel != WS_KERNEL_TILED || b->base.viscous || b->schedule != WS_SCHED_GRID ||
        b->ctx[0]->fuse_stage < 3)
        return false;
    if (!prim_tables(b, s) || !aux_table(b, s) || !b->aux_second) return false;
    if (!cfl || b->d_cfl_tiles) return true;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
    long n = 0;   // room for the x and z tile lists of every shape
    for (int shape = 0; shape < WS_TILE_COUNT; ++shape)
        n = std::max(n, (long)std::max(b->ntiles[shape][0], b->ntiles[shape][2]));
    if (hipMalloc(&b->d_cfl_tiles, sizeof(double) * n) != hipSuccess) {
        b->d_cfl_tiles = nullptr;
        (void)hipGetLastError();
        return false;
    }
    return true;
}

int ws_batch_rk3_run(ws_batch *b, int32_t nsteps, double cfl, int32_t use_graph, hipStream_t s)
{
    if (!b || nsteps < 0) return set_error(WS_ERR_INVALID, "bad RK3 request");
    double *dt_dev = cfl > 0.0 ? b->d_dt : nullptr;
    // After the first step each block's CFL maxima come from its stage kernel that ended
    // the previous step (as in ws_rk3_run): per block k_cfl_final over those, then the
    // minimum over the blocks; the same dt bits as ws_batch_cfl_dt.
    bool staged = dt_dev != nullptr;
    for (ws_ctx *c : b->ctx) staged = staged && cfl_stage_buffer(c);
    // Stage fusion 3: the steps alternate the primitive sets (batch_step), the stage-3
    // launch of each step but the last leaves the maxima per tile, and the last step ends
    // in the blocks' fields.
    const bool alt = nsteps > 1 && batch_alternating(b, s, dt_dev != nullptr);
    auto mode_of = [&](int n) {
        if (!alt) return STEP_PLAIN;
        if (n == nsteps - 1) return (n & 1) ? STEP_LAST_FROM_SECOND : STEP_PLAIN;
        return (n & 1) ? STEP_FROM_SECOND : STEP_FROM_STATE;
    };
    const int nb = (int)b->ctx.size();
    auto cfl_dt = [&](int n, hipStream_t ss) {
        if (n == 0 || !staged) return ws_batch_cfl_dt(b, cfl, dt_dev, ss);
        // the previous step left its maxima per tile (alternating) or in the batch's buffer
        if (alt)
            return launch_cfl_batch(b->d_aux, nb, b->d_cfl_tiles, cfl, b->d_dts, dt_dev, ss,
                                    b->d_tile_ranges[stage_shape_of(b)][b->stage_order]);
        if (b->cfl_in_batch) return launch_cfl_batch(b->d_aux, nb, b->d_cfl_part, cfl, b->d_dts, dt_dev, ss);
        for (size_t i = 0; i < b->ctx.size(); ++i) {
            ws_ctx *c = b->ctx[i];
            const int q = launch_cfl_final(*c, c->cfl_stage, (int)c->cfl_stage_n, cfl, b->d_dts + i, ss);
            if (q) return q;
        }
        return launch_min_reduce(b->d_dts, nb, dt_dev, ss);
    };
    // the last step's maxima would never be read
    auto step = [&](int n, hipStream_t ss) {
        int q = WS_OK;
        if (dt_dev) q = cfl_dt(n, ss);
        if (!q) q = batch_step(b, dt_dev, ss, staged && n < nsteps - 1, mode_of(n), n == 0);
        return q;
    };
    int r = WS_OK;
    if (!use_graph) {
        for (int n = 0; n < nsteps && !r; ++n) r = step(n, s);
    } else {
        // one captured graph per parity of the f/fn swap (a step ends with them exchanged,
        // and, alternating, with the primitives in the other set); step 0 runs eagerly, and
        // so does an alternating run's last step, which ends differently
        hipGraphExec_t exec[2] = {nullptr, nullptr};
        const int replayed = alt ? nsteps - 1 : nsteps;
        for (int n = 0; n < replayed && !r; ++n) {
            if (n == 0) {
                r = step(0, s);
                continue;
            }
            const int p = n & 1;
            if (!exec[p]) {
                const std::vector<ws_state> saved = b->st;
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) { r = hip_error(e, "begin capture"); break; }
                r = step(n, b->capture_stream);
                e = hipStreamEndCapture(b->capture_stream, &g);
                b->st = saved;   // capture recorded the step; the swap happens at replay
                if (r) { if (g) hipGraphDestroy(g); break; }
                if (e != hipSuccess) { r = hip_error(e, "end capture"); break; }
                e = hipGraphInstantiate(&exec[p], g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                if (e != hipSuccess) { r = hip_error(e, "graph instantiate"); break; }
            }
            hipError_t e = hipGraphLaunch(exec[p], s);
            if (e != hipSuccess) { r = hip_error(e, "graph launch"); break; }
            for (ws_state &st : b->st) std::swap(st.f, st.fn);
        }
        if (!r && alt) r = step(nsteps - 1, s);
        for (int p = 0; p < 2; ++p)
            if (exec[p]) {
                hipStreamSynchronize(s);
                hipGraphExecDestroy(exec[p]);
            }
    }
    hipError_t e = hipStreamSynchronize(s);
    if (!r && e != hipSuccess) r = hip_error(e, "batch RK3 run");
    return r;
}

// The batch's stage shape and order, timed as ws_autotune times a context's: the batched RHS of an
// RK3 stage 2 whose last launch writes every block's new f into scratch allocated here (a
// copy of the block's f, which is the stage's fn and is only read) and, where the blocks
// keep second primitive sets, the primitives, sound speed and their halo into those (the
// sets a run's first stage overwrites anyway), in every tile shape and both orders.  Sets
// *shape and *order to the fastest; false when the scratch cannot be had (both then stay
// as they were) or on an error (r set).
static bool batch_tune_stage(ws_batch *b, int reps, hipEvent_t e0, hipEvent_t e1, hipStream_t s, int &r,
                             int *shape_out, int *order_out)
{
    const bool prims = b->ctx[0]->fuse_stage > 1 && prim_tables(b, s);
    const size_t nb = b->st.size();
    std::vector<size_t> off(nb), len(nb);   // each block's f in the scratch (doubles)
    size_t total = 0;
    for (size_t i = 0; i < nb; ++i) {
        off[i] = total;
        len[i] = 5 * (size_t)b->st[i].f.pitch * b->host_blocks[i].n[1] * b->host_blocks[i].n[2];
        total += len[i];
    }
    double *scratch = nullptr;
    BatchBlock *d_tab = nullptr;
    if (hipMalloc(&scratch, total * sizeof(double)) != hipSuccess ||
        hipMalloc(&d_tab, sizeof(BatchBlock) * nb) != hipSuccess) {
        (void)hipGetLastError();
        if (scratch) hipFree(scratch);
        return -1;
    }
    // the block table with fa the scratch copy (the stage's f: BatchStage::swapped = 0) and
    // fb the block's f; its outputs are the second sets once prim_tables has made them
    std::vector<BatchBlock> tab = b->host_blocks;
    for (size_t i = 0; i < nb && !r; ++i) {
        const ws_soa f = b->st[i].f;
        hipError_t e = hipMemcpyAsync(scratch + off[i], f.ptr, len[i] * sizeof(double), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) r = hip_error(e, "autotune stage scratch copy");
        tab[i].fa = f;
        tab[i].fa.ptr = scratch + off[i];
        tab[i].fb = f;
    }
    if (!r) {
        hipError_t e = hipMemcpy(d_tab, tab.data(), sizeof(BatchBlock) * nb, hipMemcpyHostToDevice);
        if (e != hipSuccess) r = hip_error(e, "autotune stage table upload");
    }
    BatchStage bs{nullptr, b->base.dt, 2, 0};
    bs.prims = prims ? 1 : 0;
    bs.planes = b->base.bc_planes != 0;
    bs.halo = bs.prims;
    const int saved = b->stage_order, saved_shape = b->stage_shape;
    float best_ms = 1e30f;
    for (int sh = 0; sh < WS_TILE_COUNT && !r; ++sh)
        for (int o = 0; o < WS_ORDER_COUNT && !r; ++o) {
            b->stage_shape = sh;
            b->stage_order = o;
            for (int i = 0; i < 2 && !r; ++i) r = batch_rhs(b, s, &bs, d_tab);
            if (r) break;
            hipEventRecord(e0, s);
            for (int i = 0; i < reps && !r; ++i) r = batch_rhs(b, s, &bs, d_tab);
            hipEventRecord(e1, s);
            if (r) break;
            hipError_t e = hipEventSynchronize(e1);
            if (e != hipSuccess) { r = hip_error(e, "autotune sync"); break; }
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            if (ms < best_ms) {
                best_ms = ms;
                *shape_out = sh;
                *order_out = o;
            }
        }
    b->stage_order = saved;
    b->stage_shape = saved_shape;
    hipStreamSynchronize(s);
    hipFree(d_tab);
    hipFree(scratch);
    return !r;
}

int ws_batch_autotune(ws_batch *b, int32_t reps, float *times_ms, hipStream_t s)
{
    if (!b || reps <= 0) return set_error(WS_ERR_INVALID, "bad autotune request");
    if (b->kernel == WS_KERNEL_ANCHOR) return set_error(WS_ERR_INVALID, "autotune applies to the tiled kernels");
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return hip_error(e, "autotune event");
    e = hipEventCreate(&e1);
    if (e != hipSuccess) { hipEventDestroy(e0); return hip_error(e, "autotune event"); }
    const int saved_shape = b->tile_shape, saved_sched = b->schedule;
    int best_shape = saved_shape, best_sched = saved_sched, r = WS_OK;
    float best_ms = 1e30f;
    for (int shape = 0; shape < WS_TILE_COUNT && !r; ++shape)
        for (int sched = 0; sched < WS_SCHED_COUNT && !r; ++sched) {
            b->tile_shape = shape;
            b->schedule = sched;
            for (int i = 0; i < 2 && !r; ++i) r = batch_rhs(b, s);
            if (r) break;
            hipEventRecord(e0, s);
            for (int i = 0; i < reps && !r; ++i) r = batch_rhs(b, s);
            hipEventRecord(e1, s);
            if (r) break;
            e = hipEventSynchronize(e1);
            if (e != hipSuccess) { r = hip_error(e, "autotune sync"); break; }
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            ms /= (float)reps;
            if (times_ms) times_ms[shape * WS_SCHED_COUNT + sched] = ms;
            if (ms < best_ms) { best_ms = ms; best_shape = shape; best_sched = sched; }
        }
    if (!r && best_sched == WS_SCHED_GRID && !b->base.viscous && b->ctx[0]->fuse_stage) {
        b->tile_shape = best_shape;
        b->schedule = best_sched;
        // the stage candidates differ by 1-3 %: three times the repetitions
        int sh = b->stage_shape, o = b->stage_order;
        if (batch_tune_stage(b, 3 * reps, e0, e1, s, r, &sh, &o)) {
            b->stage_shape = sh;
            b->stage_order = o;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    b->tile_shape = r ? saved_shape : best_shape;
    b->schedule = r ? saved_sched : best_sched;
    return r;
}

}  // extern "C"
