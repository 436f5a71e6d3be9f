// dvc_api.cpp -- C ABI of libdvc_amd.so (declared in include/dvc_amd.h).
//
// Host responsibilities: the device context, HBM-resident volumes, the FP64 geometry
// that Search::Search and Interpolate::Interpolate fix once per run
// (Core/Search.cpp:24-82, Core/Interpolate.cpp:34-58), the sphere sample template
// (Core/FloatingCloud.cpp:40-82), launching the kernels of dvc_kernels.hip, and the
// host-side DataCloud ordering (Core/DataCloud.cpp:69-144).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "dvc_geometry.h"

using namespace dvc;

namespace {

struct Volume {
    uint8_t *d = nullptr;
    bool owned = false;
    int bpv = 0;
    int64_t W = 0, H = 0, T = 0;
};

// The largest cloud a run allows: a cube of 37^3 samples for subvol_npts 50,000
// (cube counts round up to a cube, Search.cpp:30-31).
constexpr int kNoAllocSamples = 50653;
// Device scratch of one dvc_obj_val_at_noalloc call: stable and moved clouds, sample
// values, residual and objective (9 n + 1 doubles), rounded up to 256 bytes.
constexpr size_t kNoAllocScratchBytes = (sizeof(double) * (9 * (size_t)kNoAllocSamples + 1) + 255) / 256 * 256;
constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;                              // one no-alloc call at a time in this process
void *g_scratch[kMaxDevices] = {};                    // the scratch array's address on each device
double g_host_stage[9 * (size_t)kNoAllocSamples];     // sphere offsets, stable and moved clouds

}  // namespace

// dvc_obj_val_at_noalloc's device scratch: one copy on every GPU the library runs on,
// held until the process ends.  Host code never names it: it uses the address that
// hipGetSymbolAddress returns for each device (g_scratch).
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Set-up of the no-alloc entry point on `device`, which must be current: the scratch
// array's address there, looked up once per device.
static int scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch[device]) return 0;
    void *ptr = nullptr;
    if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return -1;
    g_scratch[device] = ptr;
    return 0;
}

struct dvc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    Volume vol[2];
    int cus = 0;
    size_t lds_max = 0;
    int block_threads = 0;
    float last_ms = 0.0f;
    bool timing_pending = false;     // ev0/ev1 of an execute not yet read
    // search plan
    bool plan_ready = false;
    dvc_run_params rc{};
    Geometry geo{};
    int npts = 0, nproc = 0, k = 0;
    double *d_pts = nullptr, *d_sph = nullptr;
    int *d_order = nullptr, *d_pos = nullptr, *d_neigh = nullptr, *d_levels = nullptr;
    dvc_point_result *d_res = nullptr;
    double off_min[3]{}, off_max[3]{};
    int chosen_threads = 256;
    int ref_threads = 256;           // threads per point of reference_kernel (one launch over all points)
    int round_cap = 0;               // points per chain launch: 0 automatic, -1 levels only
    std::vector<char> searched;
    std::vector<int> level_off;      // dependency levels (dvc_geometry.h)
    float *d_rvals = nullptr;        // reference samples of every searched point (or null)
    // per-point geometry for the next prepare (empty: the run parameters' for all points)
    std::vector<int32_t> pg_geom, pg_size, pg_npts;
    PtGeo *d_pgeo = nullptr;
    int *d_ref_order = nullptr;      // reference_kernel's workgroup order (spatial, per XCD)
    RefStats *d_rstat = nullptr;
    // start modes (DVC_START_*): the mode set for the next prepare and the prepared one
    int start_mode = DVC_START_CHAIN;
    int plan_mode = DVC_START_CHAIN;
    bool starts_set = false;
    double *d_start = nullptr;                 // [npts*12] given start warps
    dvc_point_result *d_res_prop = nullptr;    // proposal pass (DVC_START_JUDGED)
    dvc_point_result *d_res_judge = nullptr;   // judge pass (DVC_START_JUDGED)
    int *d_kept = nullptr;                     // 1 proposal kept, 0 judge kept, 2 repaired
    int *d_repair = nullptr;                   // 1 where a judged point is searched again
    int *d_rep_levels = nullptr;               // repair launches (the serial chain's launches)
    std::vector<int> rep_off, rep_pts_h;       // their offsets, and their points on the host
    int *h_repair = nullptr;                   // pinned host copy of d_repair (one int per point)
    // slice launches of the chain (SliceArgs): largest workgroup count of a sliced launch
    // (0 automatic, -1 off), slice launches per chain launch before the one-workgroup
    // search resumes the points still searching (0 automatic), samples per workgroup
    int slice_wgs = 0, slice_evals = 0, slice_per = 512;
    std::vector<char> sliced;                  // chain launches run as slice launches
    std::vector<int> slice_nsl;                // workgroups per point of each chain launch (when sliced)
    std::vector<int> slice_cap;                // per-point geometry: cap on a point's workgroups (0: none)
    int slice_cnt8 = 0, slice_n = 0;           // buffer shape of the current plan
    std::vector<int> level_pts_h;              // points of the chain launches (host copy of d_levels)
    std::vector<int> pt_nsamp;                 // samples of each point (per-point geometry only)
    unsigned char *d_state = nullptr;          // [2][cnt8][kStateStride]
    double *d_part = nullptr;                  // [2][cnt8][nslice][kPartStride]
    // the strain fit of the search's results after every execute (ndp 0: none)
    int strain_ndp = 0, strain_refill = 0, strain_planar = 0;
    double strain_objt = 1.0;
    double pts_mn[3]{}, pts_mx[3]{};           // bounding box of the prepared points (plane detection)
    double *d_sdisp = nullptr, *d_sobj = nullptr, *d_sout = nullptr;   // inputs; outputs [3n disp, n swr, 27n E L DG]
    int *d_sstat = nullptr, *d_snpts = nullptr;
    // the level launches of a prepared search, captured once and replayed
    int use_graph = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

namespace {

int fail(dvc_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(dvc_ctx *c, hipError_t e, const char *what) {
    return fail(c, DVC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(call)                                        \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

template <typename T>
int dev_alloc(dvc_ctx *ctx, T **p, size_t count) {
    if (*p) { hipFree(*p); *p = nullptr; }
    if (count == 0) count = 1;
    HIPCHK(hipMalloc((void **)p, count * sizeof(T)));
    return 0;
}

void drop_graph(dvc_ctx *c) {
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

// HBM for the reference sub-volumes computed ahead of the level chain
constexpr size_t kRefBudgetBytes = (size_t)48 << 30;
// accelerator complex dies of an MI355X (gfx950), each with its own L2
constexpr int kXcds = 8;

void free_plan(dvc_ctx *c) {
    drop_graph(c);
    hipFree(c->d_pts); hipFree(c->d_sph); hipFree(c->d_order); hipFree(c->d_pos);
    hipFree(c->d_neigh); hipFree(c->d_levels); hipFree(c->d_res);
    hipFree(c->d_rvals); hipFree(c->d_rstat); hipFree(c->d_pgeo); hipFree(c->d_ref_order);
    c->d_pgeo = nullptr;
    c->d_ref_order = nullptr;
    hipFree(c->d_state); hipFree(c->d_part);
    c->d_state = nullptr; c->d_part = nullptr;
    c->slice_cnt8 = c->slice_n = 0;
    c->sliced.clear();
    c->slice_nsl.clear();
    c->slice_cap.clear();
    hipFree(c->d_start); hipFree(c->d_res_prop); hipFree(c->d_res_judge); hipFree(c->d_kept);
    hipFree(c->d_repair); hipFree(c->d_rep_levels);
    if (c->h_repair) hipHostFree(c->h_repair);
    hipFree(c->d_sdisp); hipFree(c->d_sobj); hipFree(c->d_sout); hipFree(c->d_sstat); hipFree(c->d_snpts);
    c->d_sdisp = c->d_sobj = c->d_sout = nullptr;
    c->d_sstat = c->d_snpts = nullptr;
    c->d_start = nullptr; c->d_res_prop = c->d_res_judge = nullptr; c->d_kept = nullptr;
    c->d_repair = nullptr; c->d_rep_levels = nullptr;
    c->h_repair = nullptr;
    c->rep_off.clear();
    c->rep_pts_h.clear();
    c->starts_set = false;
    c->d_rvals = nullptr;
    c->d_rstat = nullptr;
    c->d_pts = c->d_sph = nullptr;
    c->d_order = c->d_pos = c->d_neigh = c->d_levels = nullptr;
    c->d_res = nullptr;
    c->plan_ready = false;
}

int check_params(dvc_ctx *ctx, const dvc_run_params *rc) {
    if (!rc) return fail(ctx, DVC_ERR_ARG, "null run parameters");
    if (rc->ndof != 3 && rc->ndof != 6 && rc->ndof != 12) return fail(ctx, DVC_ERR_ARG, "ndof must be 3, 6 or 12");
    if (rc->obj < 0 || rc->obj > 4) return fail(ctx, DVC_ERR_ARG, "unknown objective function");
    if (rc->interp < 0 || rc->interp > 2) return fail(ctx, DVC_ERR_ARG, "unknown interpolation type");
    if (rc->geom != DVC_GEOM_CUBE && rc->geom != DVC_GEOM_SPHERE) return fail(ctx, DVC_ERR_ARG, "unknown subvolume geometry");
    if (rc->subvol_size < 1) return fail(ctx, DVC_ERR_ARG, "subvol_size must be >= 1");
    if (rc->subvol_npts < 1 || rc->subvol_npts > 50000) return fail(ctx, DVC_ERR_ARG, "subvol_npts must be in 1..50000");
    if (rc->disp_max < 0) return fail(ctx, DVC_ERR_ARG, "disp_max must be >= 0");
    if (!(rc->basin_radius >= 0.0) || rc->basin_radius > rc->disp_max) return fail(ctx, DVC_ERR_ARG, "basin_radius must be in 0..disp_max");
    for (int q = 0; q < 3; q++)
        if (!(rc->aspect[q] >= 0.1 && rc->aspect[q] <= 10.0)) return fail(ctx, DVC_ERR_ARG, "subvol_aspect must be in 0.1..10");
    if (rc->geom == DVC_GEOM_CUBE) {
        int g = (int)ceil(pow((double)rc->subvol_npts, 1.0 / 3.0));
        if (g < 2) return fail(ctx, DVC_ERR_ARG, "a cube needs at least 2 samples per side");
    }
    return 0;
}

int check_volumes(dvc_ctx *ctx) {
    if (!ctx->vol[0].d || !ctx->vol[1].d) return fail(ctx, DVC_ERR_STATE, "both volumes must be uploaded first");
    const Volume &a = ctx->vol[0], &b = ctx->vol[1];
    if (a.bpv != b.bpv || a.W != b.W || a.H != b.H || a.T != b.T)
        return fail(ctx, DVC_ERR_ARG, "reference and correlated volumes differ in size or depth");
    return 0;
}

}  // namespace

namespace dvc {

int ctx_device(dvc_ctx *ctx) { return ctx->device; }
void *ctx_stream(dvc_ctx *ctx) { return (void *)ctx->stream; }

int ctx_volumes(dvc_ctx *ctx, const uint8_t **v0, const uint8_t **v1, int *bpv, int *W, int *H, int *T) {
    if (check_volumes(ctx) != 0) return -1;
    if (ctx->vol[0].W > INT32_MAX || ctx->vol[0].H > INT32_MAX || ctx->vol[0].T > INT32_MAX) return -1;
    *v0 = ctx->vol[0].d; *v1 = ctx->vol[1].d; *bpv = ctx->vol[0].bpv;
    *W = (int)ctx->vol[0].W; *H = (int)ctx->vol[0].H; *T = (int)ctx->vol[0].T;
    return 0;
}

int ctx_volume_info(dvc_ctx *ctx, int slot, const uint8_t **p, int *bpv, int *W, int *H, int *T) {
    if (slot != 0 && slot != 1) return -1;
    const Volume &v = ctx->vol[slot];
    if (!v.d || v.W > INT32_MAX || v.H > INT32_MAX || v.T > INT32_MAX) return -1;
    *p = v.d; *bpv = v.bpv; *W = (int)v.W; *H = (int)v.H; *T = (int)v.T;
    return 0;
}

uint8_t *ctx_volume_alloc(dvc_ctx *ctx, int slot, int bpv, int64_t w, int64_t h, int64_t t) {
    if (slot != 0 && slot != 1) return nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return nullptr;
    drop_graph(ctx);
    Volume &v = ctx->vol[slot];
    if (v.owned) hipFree(v.d);
    v = Volume();
    const size_t bytes = (size_t)w * h * t * bpv;
    if (hipMalloc((void **)&v.d, bytes + 64) != hipSuccess) { v.d = nullptr; return nullptr; }
    v.owned = true;
    v.bpv = bpv; v.W = w; v.H = h; v.T = t;
    if (hipMemsetAsync(v.d + bytes, 0, 64, ctx->stream) != hipSuccess) return nullptr;
    return v.d;
}

int ctx_prepared(dvc_ctx *ctx, const double **pts, const int **order, int *npts, int *nproc, double **start) {
    if (!ctx->plan_ready || ctx->plan_mode == DVC_START_CHAIN) return -1;
    *pts = ctx->d_pts; *order = ctx->d_order; *npts = ctx->npts; *nproc = ctx->nproc; *start = ctx->d_start;
    return 0;
}

void ctx_mark_starts(dvc_ctx *ctx) { ctx->starts_set = true; }

}  // namespace dvc

extern "C" {

int dvc_create(dvc_ctx **out, int device) {
    if (!out) return DVC_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return DVC_ERR_NODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return DVC_ERR_NODEV;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return DVC_ERR_NODEV;
    dvc_ctx *ctx = new dvc_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete ctx; return DVC_ERR_NODEV; }
    ctx->cus = prop.multiProcessorCount;
    ctx->lds_max = prop.sharedMemPerBlock;
    if (search_set_lds_limit(ctx->lds_max) != 0) { delete ctx; return DVC_ERR_HIP; }
    if (scratch_setup(device) != 0) { delete ctx; return DVC_ERR_HIP; }
    if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        delete ctx;
        return DVC_ERR_HIP;
    }
    ctx->stream = ctx->own_stream;
    *out = ctx;
    return DVC_OK;
}

void dvc_destroy(dvc_ctx *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    free_plan(ctx);
    for (int s = 0; s < 2; s++)
        if (ctx->vol[s].owned) hipFree(ctx->vol[s].d);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

const char *dvc_last_error(const dvc_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int dvc_set_stream(dvc_ctx *ctx, void *s) {
    if (!ctx) return DVC_ERR_ARG;
    ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
    return DVC_OK;
}

int dvc_last_kernel_ms(dvc_ctx *ctx, float *ms) {
    if (!ctx || !ms) return DVC_ERR_ARG;
    if (ctx->timing_pending) {
        HIPCHK(hipEventSynchronize(ctx->ev1));
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
        ctx->last_ms = t;
        ctx->timing_pending = false;
    }
    *ms = ctx->last_ms;
    return DVC_OK;
}

int dvc_set_round_capacity(dvc_ctx *ctx, int32_t points) {
    if (!ctx || points < -1) return DVC_ERR_ARG;
    ctx->round_cap = points;
    return DVC_OK;
}

int dvc_set_block_threads(dvc_ctx *ctx, int threads) {
    if (!ctx) return DVC_ERR_ARG;
    if (threads != 0 && threads != 64 && threads != 128 && threads != 256 && threads != 512 && threads != 1024)
        return fail(ctx, DVC_ERR_ARG, "threads per point must be 0, 64, 128, 256, 512 or 1024");
    ctx->block_threads = threads;
    return DVC_OK;
}

int dvc_volume_upload(dvc_ctx *ctx, int slot, const void *host, int64_t header, int bpv, int big_endian,
                      int64_t w, int64_t h, int64_t t) {
    if (!ctx) return DVC_ERR_ARG;
    if (slot != 0 && slot != 1) return fail(ctx, DVC_ERR_ARG, "slot must be DVC_VOL_REF or DVC_VOL_COR");
    if (!host || (bpv != 1 && bpv != 2) || w < 4 || h < 4 || t < 4 || header < 0)
        return fail(ctx, DVC_ERR_ARG, "invalid volume description");
    if (w > 2147483647LL / 2 || h > 2147483647LL || t > 2147483647LL) return fail(ctx, DVC_ERR_ARG, "volume too large");
    HIPCHK(hipSetDevice(ctx->device));
    drop_graph(ctx);
    Volume &v = ctx->vol[slot];
    if (v.owned) hipFree(v.d);
    v = Volume();
    size_t bytes = (size_t)w * h * t * bpv;
    HIPCHK(hipMalloc((void **)&v.d, bytes + 64));
    v.owned = true;
    v.bpv = bpv; v.W = w; v.H = h; v.T = t;
    HIPCHK(hipMemsetAsync(v.d + bytes, 0, 64, ctx->stream));
    HIPCHK(hipMemcpyAsync(v.d, (const uint8_t *)host + header, bytes, hipMemcpyHostToDevice, ctx->stream));
    if (bpv == 2 && big_endian)
        if (launch_swap16((uint16_t *)v.d, (long long)w * h * t, ctx->stream) != 0) return fail(ctx, DVC_ERR_HIP, "byte swap launch failed");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DVC_OK;
}

int dvc_volume_attach(dvc_ctx *ctx, int slot, const void *dptr, int bpv, int64_t w, int64_t h, int64_t t) {
    if (!ctx) return DVC_ERR_ARG;
    if (slot != 0 && slot != 1) return fail(ctx, DVC_ERR_ARG, "slot must be DVC_VOL_REF or DVC_VOL_COR");
    if (!dptr || (bpv != 1 && bpv != 2) || w < 4 || h < 4 || t < 4) return fail(ctx, DVC_ERR_ARG, "invalid volume description");
    if (((uintptr_t)dptr & 3) != 0) return fail(ctx, DVC_ERR_ARG, "device volume must be 4-byte aligned");
    drop_graph(ctx);
    Volume &v = ctx->vol[slot];
    if (v.owned) hipFree(v.d);
    v = Volume();
    v.d = (uint8_t *)dptr; v.owned = false; v.bpv = bpv; v.W = w; v.H = h; v.T = t;
    return DVC_OK;
}

// Interpolate::center_on + kernels + tri_cub_Lek / tri_lin / nearest on the boxes of an
// Interpolate object (est box, act box = est box less one voxel, kern_4d extent box[3]).
// Interpolate::center_on (Interpolate.cpp:569-574) and the range rule of tri_cub_Lek /
// tri_lin / nearest (Interpolate.cpp:404-567): the box's first voxel b0 after centring
// on `center`, DVC_ERR_RANGE when the points' box leaves the active box, a point's cell
// leaves the coefficient box or the box leaves the volume.
static int interp_place(dvc_ctx *ctx, const Volume &v, HBox est, HBox act, const int box[3], const double center[3],
                        const double *xyz, int32_t n, const double bbox[6], int method, int b0[3]) {
    est.center_on(center);
    act.center_on(center);
    bool ok = true;
    int64_t dims[3] = {v.W, v.H, v.T};
    for (int q = 0; q < 3; q++) {
        b0[q] = trunc_i(est.mn[q]);
        if (!(est.mn[q] == est.mn[q]) || b0[q] < 0 || (int64_t)b0[q] + box[q] > dims[q]) ok = false;
        if (!(bbox[q] >= act.mn[q]) || !(bbox[3 + q] <= act.mx[q])) ok = false;   // BoundBox::contains
    }
    int lim = method == DVC_INTERP_NEAREST ? 1 : 2;
    if (ok)
        for (int i = 0; i < n && ok; i++)
            for (int q = 0; q < 3; q++) {
                double c = xyz[3 * i + q];
                if (!(c == c) || std::fabs(c) > 1e15) { ok = false; break; }
                int cell = trunc_i(c) - b0[q];
                if (cell < 0 || cell > box[q] - lim) { ok = false; break; }
            }
    if (!ok) return fail(ctx, DVC_ERR_RANGE, "points outside the active interpolation box");
    return DVC_OK;
}

static int interp_boxes(dvc_ctx *ctx, int slot, HBox est, HBox act, const int box[3], const double center[3],
                        const double *xyz, int32_t n, const double bbox[6], int method, double *out, double *grad) {
    if (slot != 0 && slot != 1) return fail(ctx, DVC_ERR_ARG, "bad slot");
    const Volume &v = ctx->vol[slot];
    if (!v.d) return fail(ctx, DVC_ERR_STATE, "volume not uploaded");
    if (n < 0 || (n > 0 && (!xyz || !out || !bbox || !center))) return fail(ctx, DVC_ERR_ARG, "bad arrays");
    if (method < 0 || method > 2) return fail(ctx, DVC_ERR_ARG, "bad interpolation method");
    if (grad && method != DVC_INTERP_TRICUBIC) return fail(ctx, DVC_ERR_ARG, "gradients are tricubic only");
    for (int q = 0; q < 3; q++)
        if (box[q] < 4) return fail(ctx, DVC_ERR_ARG, "interpolation box smaller than 4 voxels");
    int b0[3];
    if (int e = interp_place(ctx, v, est, act, box, center, xyz, n, bbox, method, b0)) return e;
    if (n == 0) return DVC_OK;
    HIPCHK(hipSetDevice(ctx->device));
    double *d_xyz = nullptr, *d_out = nullptr, *d_grad = nullptr;
    HIPCHK(hipMalloc((void **)&d_xyz, sizeof(double) * 3 * n));
    HIPCHK(hipMalloc((void **)&d_out, sizeof(double) * n));
    if (grad) HIPCHK(hipMalloc((void **)&d_grad, sizeof(double) * 3 * n));
    HIPCHK(hipMemcpyAsync(d_xyz, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
    int r = launch_interp(v.d, v.bpv, (int)v.W, (int)v.H, (int)v.T, b0, box, d_xyz, n, method, d_out, d_grad,
                          ctx->stream);
    if (r == 0) {
        hipMemcpyAsync(out, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream);
        if (grad) hipMemcpyAsync(grad, d_grad, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream);
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    hipFree(d_xyz); hipFree(d_out); hipFree(d_grad);
    if (r != 0) return fail(ctx, DVC_ERR_HIP, "interpolation launch failed");
    if (e != hipSuccess) return hip_fail(ctx, e, "interpolation");
    return DVC_OK;
}

int dvc_interpolate(dvc_ctx *ctx, int slot, const dvc_run_params *rc, const double center[3], const double *xyz,
                    int32_t n, const double bbox[6], int method, double *out, double *grad) {
    if (!ctx) return DVC_ERR_ARG;
    if (int e = check_params(ctx, rc)) return e;
    Geometry G = make_geometry(*rc);
    HBox est(G.est_mn, G.est_mx);
    HBox act(G.act_mn, G.act_mx);
    for (int q = 0; q < 3; q++) { act.cn[q] = G.act_cn[q]; est.cn[q] = G.est_cn[q]; }
    return interp_boxes(ctx, slot, est, act, G.box, center, xyz, n, bbox, method, out, grad);
}

int dvc_interpolate_box(dvc_ctx *ctx, int slot, const double region_min[3], const double region_max[3],
                        const double center[3], const double *xyz, int32_t n, const double bbox[6], int method,
                        double *out, double *grad) {
    if (!ctx) return DVC_ERR_ARG;
    if (!region_min || !region_max) return fail(ctx, DVC_ERR_ARG, "bad region");
    for (int q = 0; q < 3; q++)
        if (!std::isfinite(region_min[q]) || !std::isfinite(region_max[q]) || region_max[q] < region_min[q] ||
            std::fabs(region_min[q]) > 1e9 || std::fabs(region_max[q]) > 1e9)
            return fail(ctx, DVC_ERR_ARG, "bad region");
    // Interpolate::Interpolate (Interpolate.cpp:34-58)
    HBox est(region_min, region_max);
    HBox act(region_min, region_max);
    act.grow_by(-1.0);
    int box[3];
    for (int q = 0; q < 3; q++) box[q] = trunc_i(est.mx[q]) - trunc_i(est.mn[q]);
    return interp_boxes(ctx, slot, est, act, box, center, xyz, n, bbox, method, out, grad);
}

int dvc_objective(dvc_ctx *ctx, int kind, const double *ref, const double *tar, int32_t n, double *obj, double *res) {
    if (!ctx) return DVC_ERR_ARG;
    if (kind < 0 || kind > 4) return fail(ctx, DVC_ERR_ARG, "unknown objective");
    if (n < 1 || !ref || !tar || !obj) return fail(ctx, DVC_ERR_ARG, "bad arrays");
    HIPCHK(hipSetDevice(ctx->device));
    double *d = nullptr;
    HIPCHK(hipMalloc((void **)&d, sizeof(double) * (3 * (size_t)n + 1)));
    double *dr = d, *dt = d + n, *dres = d + 2 * (size_t)n, *dobj = d + 3 * (size_t)n;
    hipMemcpyAsync(dr, ref, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream);
    hipMemcpyAsync(dt, tar, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream);
    int r = launch_objective(kind, dr, dt, n, dobj, res ? dres : nullptr, ctx->stream);
    if (r == 0) {
        hipMemcpyAsync(obj, dobj, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (res) hipMemcpyAsync(res, dres, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream);
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    hipFree(d);
    if (r != 0) return fail(ctx, DVC_ERR_HIP, "objective launch failed");
    if (e != hipSuccess) return hip_fail(ctx, e, "objective");
    return DVC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ obj_val_at

namespace {

// Device buffers of one point evaluation of n samples: the stable and moved clouds,
// their sample values (ref_subvol, tar_subvol of Search), the residual and the objective.
struct PointWork {
    double *xyz_s, *xyz_m, *vref, *vtar, *res, *obj;
};

size_t point_work_bytes(int n) { return sizeof(double) * (9 * (size_t)n + 1); }

PointWork carve_point_work(void *base, int n) {
    double *d = (double *)base;
    const size_t m = (size_t)n;
    PointWork w;
    w.xyz_s = d; w.xyz_m = d + 3 * m; w.vref = d + 6 * m; w.vtar = d + 7 * m; w.res = d + 8 * m; w.obj = d + 9 * m;
    return w;
}

// Host side of one point evaluation: room for the stable and moved clouds ([3n] each)
// and the sphere offsets ([3n], spheres only).
struct PointClouds {
    double *sph, *stable, *moving;
};

int obj_val_args(dvc_ctx *ctx, const dvc_run_params *rc, const double point[3], const double *x, int32_t nx,
                 const double *obj, const double *residual, int32_t residual_len, Geometry &G) {
    if (int e = check_params(ctx, rc)) return e;
    if (int e = check_volumes(ctx)) return e;
    if (!point || !x || !obj) return fail(ctx, DVC_ERR_ARG, "null point, parameters or objective");
    if (nx < 1 || nx > 12) return fail(ctx, DVC_ERR_ARG, "nx must be in 1..12");
    for (int q = 0; q < 3; q++)
        if (!std::isfinite(point[q])) return fail(ctx, DVC_ERR_ARG, "non-finite point");
    G = make_geometry(*rc);
    if (residual && residual_len < G.subv_num) return fail(ctx, DVC_ERR_ARG, "residual shorter than the cloud");
    return DVC_OK;
}

// Search::obj_val_at (Search.cpp:448-488) in the state Search::search_pt_setup
// (Search.cpp:248-275) leaves for one point: the clouds are built on the host in C, the
// sample values, residual and objective live in w on the device.  The reference cloud
// is sampled with the interpolation box centred on the point, the moved cloud with the
// box centred on the point moved by the start translation.  Returns after the stream
// has drained.
int obj_val_run(dvc_ctx *ctx, const dvc_run_params *rc, const Geometry &G, const double point[3],
                const double start[3], const double *x, int32_t nx, const PointClouds &C, const PointWork &w,
                double *obj, double *residual) {
    const int n = G.subv_num;
    double sbox[6], sref[3], mbox[6];
    if (rc->geom == DVC_GEOM_SPHERE) sphere_offsets(*rc, G, C.sph);
    stable_cloud(*rc, G, C.sph, point, C.stable, sbox, sref);
    HBox est(G.est_mn, G.est_mx), act(G.act_mn, G.act_mx);
    for (int q = 0; q < 3; q++) { act.cn[q] = G.act_cn[q]; est.cn[q] = G.est_cn[q]; }
    int b0r[3], b0t[3];
    if (interp_place(ctx, ctx->vol[0], est, act, G.box, point, C.stable, n, sbox, rc->interp, b0r))
        return fail(ctx, DVC_ERR_RANGE, "the reference cloud leaves the interpolation box (Intrp_Fail)");
    for (int q = 0; q < nx; q++)
        if (!std::isfinite(x[q])) return fail(ctx, DVC_ERR_RANGE, "non-finite parameters (Range_Fail)");
    affine_cloud(C.stable, n, sbox, sref, x, nx, C.moving, mbox);
    double cen[3];
    for (int q = 0; q < 3; q++) cen[q] = point[q] + start[q];
    if (interp_place(ctx, ctx->vol[1], est, act, G.box, cen, C.moving, n, mbox, rc->interp, b0t))
        return fail(ctx, DVC_ERR_RANGE, "the moved cloud leaves the interpolation box (Range_Fail)");
    HIPCHK(hipSetDevice(ctx->device));
    const Volume &vr = ctx->vol[0], &vt = ctx->vol[1];
    hipStream_t st = ctx->stream;
    hipError_t he = hipMemcpyAsync(w.xyz_s, C.stable, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st);
    if (he == hipSuccess) he = hipMemcpyAsync(w.xyz_m, C.moving, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st);
    int e = he == hipSuccess ? 0 : -1;
    if (e == 0)
        e = launch_interp(vr.d, vr.bpv, (int)vr.W, (int)vr.H, (int)vr.T, b0r, G.box, w.xyz_s, n, rc->interp, w.vref,
                          nullptr, st);
    if (e == 0)
        e = launch_interp(vt.d, vt.bpv, (int)vt.W, (int)vt.H, (int)vt.T, b0t, G.box, w.xyz_m, n, rc->interp, w.vtar,
                          nullptr, st);
    if (e == 0) e = launch_objective(rc->obj, w.vref, w.vtar, n, w.obj, residual ? w.res : nullptr, st);
    if (e == 0) he = hipMemcpyAsync(obj, w.obj, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == 0 && he == hipSuccess && residual)
        he = hipMemcpyAsync(residual, w.res, sizeof(double) * n, hipMemcpyDeviceToHost, st);
    hipError_t hs = hipStreamSynchronize(st);
    if (he != hipSuccess) return hip_fail(ctx, he, "obj_val_at copy");
    if (e != 0) return fail(ctx, DVC_ERR_HIP, "obj_val_at launch failed");
    if (hs != hipSuccess) return hip_fail(ctx, hs, "obj_val_at");
    return DVC_OK;
}

}  // namespace

extern "C" {

int dvc_obj_val_at(dvc_ctx *ctx, const dvc_run_params *rc, const double point[3], const double start[3],
                   const double *x, int32_t nx, double *obj, double *residual, int32_t residual_len, int32_t *nsamp) {
    if (!ctx) return DVC_ERR_ARG;
    Geometry G;
    if (int e = obj_val_args(ctx, rc, point, x, nx, obj, residual, residual_len, G)) return e;
    const int n = G.subv_num;
    if (nsamp) *nsamp = n;
    std::vector<double> host((size_t)9 * n);
    PointClouds C{host.data(), host.data() + 3 * (size_t)n, host.data() + 6 * (size_t)n};
    HIPCHK(hipSetDevice(ctx->device));
    void *buf = nullptr;
    HIPCHK(hipMalloc(&buf, point_work_bytes(n)));
    int e = obj_val_run(ctx, rc, G, point, start ? start : rc->rigid_trans, x, nx, C, carve_point_work(buf, n), obj,
                        residual);
    hipFree(buf);
    return e;
}

int dvc_obj_val_at_noalloc(dvc_ctx *ctx, const dvc_run_params *rc, const double point[3], const double start[3],
                           const double *x, int32_t nx, double *obj, double *residual, int32_t residual_len,
                           int32_t *nsamp) {
    if (!ctx) return DVC_ERR_ARG;
    Geometry G;
    if (int e = obj_val_args(ctx, rc, point, x, nx, obj, residual, residual_len, G)) return e;
    const int n = G.subv_num;
    if (n > kNoAllocSamples) return fail(ctx, DVC_ERR_ARG, "cloud larger than the no-alloc scratch (50,653 samples)");
    if (nsamp) *nsamp = n;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    void *dev = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
    if (!dev) return fail(ctx, DVC_ERR_STATE, "no scratch address on this device");
    const size_t m = (size_t)n;
    PointClouds C{g_host_stage, g_host_stage + 3 * m, g_host_stage + 6 * m};
    return obj_val_run(ctx, rc, G, point, start ? start : rc->rigid_trans, x, nx, C, carve_point_work(dev, n), obj,
                       residual);
}

int dvc_search_prepare(dvc_ctx *ctx, const dvc_run_params *rc, const double *points, int32_t npts,
                       const int32_t *order, int32_t nproc, const int32_t *neigh, int32_t k) {
    if (!ctx) return DVC_ERR_ARG;
    ctx->plan_ready = false;
    drop_graph(ctx);
    if (int e = check_params(ctx, rc)) return e;
    if (int e = check_volumes(ctx)) return e;
    if (npts < 1 || !points || !order || !neigh) return fail(ctx, DVC_ERR_ARG, "bad point arrays");
    if (k < 1 || k > kMaxNeigh || k > npts) return fail(ctx, DVC_ERR_ARG, "k must be in 1..min(npts, 128)");
    // dvc.cpp:229-232
    int N = npts;
    if (nproc > 1 && nproc < npts) N = nproc;
    std::vector<int> pos;
    if (!order_positions(order, npts, N, pos)) return fail(ctx, DVC_ERR_ARG, "order must be a permutation of 0..npts-1");
    for (int64_t i = 0; i < (int64_t)npts * k; i++)
        if (neigh[i] < 0 || neigh[i] >= npts) return fail(ctx, DVC_ERR_ARG, "neighbour index out of range");
    for (int i = 0; i < npts * 3; i++)
        if (!std::isfinite(points[i])) return fail(ctx, DVC_ERR_ARG, "non-finite point coordinate");
    Geometry G = make_geometry(*rc);
    HIPCHK(hipSetDevice(ctx->device));
    // reference samples of all searched points go to an HBM table ahead of the level
    // chain when it fits the budget; otherwise each point computes its own into LDS
    const int bpv = ctx->vol[0].bpv;
    int nsamp_pad = (G.subv_num + 3) & ~3;
    const bool table = (size_t)N * (size_t)nsamp_pad * sizeof(float) <= kRefBudgetBytes;
    // threads per point: the requested count, else the largest of 512/256/128/64 whose
    // LDS (MFMA staging, point state, and the samples without the table) fits
    const int auto_cands[4] = {512, 256, 128, 64};
    const int *cands = ctx->block_threads ? &ctx->block_threads : auto_cands;
    const int ncand = ctx->block_threads ? 1 : 4;
    int bt = 0;
    for (int i = 0; i < ncand && bt == 0; i++) {
        size_t st = 0;
        if (search_static_lds(bpv, cands[i], &st) != 0) return fail(ctx, DVC_ERR_HIP, "kernel attributes");
        if (search_lds_bytes(nsamp_pad, cands[i], !table) + st <= ctx->lds_max) bt = cands[i];
    }
    if (bt == 0) return fail(ctx, DVC_ERR_ARG, "too many samples per point for one workgroup's LDS");
    // sphere template: identical offsets for every point (one seed, Search.cpp:102-105)
    std::vector<double> sph;
    double omn[3] = {0, 0, 0}, omx[3] = {0, 0, 0};
    const bool uneven = !ctx->pg_geom.empty();
    std::vector<PtGeo> pgeo;
    size_t table_floats = (size_t)N * (size_t)nsamp_pad;
    if (uneven) {
        // one geometry class per distinct (shape, size, samples): Search::Search of a
        // RunControl with those fields (Search.cpp:24-82), the sphere template of its
        // seed; the reference samples of each point at its own offset of one table
        if ((int)ctx->pg_geom.size() != npts) return fail(ctx, DVC_ERR_ARG, "point geometry has another point count");
        if (!table) return fail(ctx, DVC_ERR_ARG, "per-point geometry needs the HBM sample table");
        std::map<std::tuple<int, int, int>, int> index;      // class of each distinct triple
        std::vector<std::tuple<int, int, int>> keys;
        std::vector<int> cls(npts);
        for (int i = 0; i < npts; i++) {
            const auto key = std::make_tuple(ctx->pg_geom[i], ctx->pg_size[i], ctx->pg_npts[i]);
            auto it = index.find(key);
            if (it == index.end()) {
                dvc_run_params ri = *rc;
                ri.geom = ctx->pg_geom[i]; ri.subvol_size = ctx->pg_size[i]; ri.subvol_npts = ctx->pg_npts[i];
                if (int e = check_params(ctx, &ri)) return e;
                it = index.emplace(key, (int)keys.size()).first;
                keys.push_back(key);
            }
            cls[i] = it->second;
        }
        // every class: its geometry and (spheres) its sample template, on host threads
        const long long wb = (long long)ctx->vol[0].W * bpv, whb = wb * (long long)ctx->vol[0].H;
        const int nc = (int)keys.size();
        std::vector<PtGeo> cg(nc, PtGeo{});
        std::vector<std::vector<double>> tmpl(nc);
        auto build = [&](int c0, int c1) {
            for (int c = c0; c < c1; c++) {
                dvc_run_params ri = *rc;
                ri.geom = std::get<0>(keys[c]); ri.subvol_size = std::get<1>(keys[c]); ri.subvol_npts = std::get<2>(keys[c]);
                const Geometry Gi = make_geometry(ri);
                PtGeo &g = cg[c];
                g.geom = ri.geom; g.nsamp = Gi.subv_num; g.g = Gi.g; g.subv_rad = Gi.subv_rad;
                g.gather32 = (wb % 4 == 0) && ((long long)(Gi.box[2] + 1) * whb + 64 < (1ll << 31)) ? 1 : 0;
                for (int q = 0; q < 3; q++) {
                    g.box[q] = Gi.box[q];
                    g.est_mn[q] = Gi.est_mn[q]; g.est_mx[q] = Gi.est_mx[q]; g.est_cn[q] = Gi.est_cn[q];
                    g.act_mn[q] = Gi.act_mn[q]; g.act_mx[q] = Gi.act_mx[q]; g.act_cn[q] = Gi.act_cn[q];
                }
                if (ri.geom == DVC_GEOM_SPHERE) {
                    double mn[3], mx[3];
                    sphere_template(ri, Gi, tmpl[c], mn, mx);
                    locality_order(tmpl[c]);
                    for (int q = 0; q < 3; q++) { g.off_min[q] = mn[q]; g.off_max[q] = mx[q]; }
                }
            }
        };
        const int nth = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (int t = 0; t < nth; t++) pool.emplace_back(build, (int)((long long)nc * t / nth), (int)((long long)nc * (t + 1) / nth));
        for (auto &th : pool) th.join();
        for (int c = 0; c < nc; c++)
            if (cg[c].geom == DVC_GEOM_SPHERE) {
                cg[c].sph_off = (long long)(sph.size() / 3);
                sph.insert(sph.end(), tmpl[c].begin(), tmpl[c].end());
            }
        pgeo.assign(npts, PtGeo{});
        for (int i = 0; i < npts; i++) pgeo[i] = cg[cls[i]];
        // reference-sample rows in processing order
        size_t off = 0;
        for (int q = 0; q < N; q++) {
            PtGeo &g = pgeo[order[q]];
            g.rval_off = (long long)off;
            off += (size_t)((g.nsamp + 3) & ~3);
        }
        table_floats = off;
        if (table_floats * sizeof(float) > kRefBudgetBytes)
            return fail(ctx, DVC_ERR_ARG, "reference samples of all points exceed the HBM table budget");
    } else if (rc->geom == DVC_GEOM_SPHERE) {
        sphere_template(*rc, G, sph, omn, omx);
        locality_order(sph);
    }
    if (int e = dev_alloc(ctx, &ctx->d_pts, (size_t)3 * npts)) return e;
    if (int e = dev_alloc(ctx, &ctx->d_sph, sph.size())) return e;
    if (int e = dev_alloc(ctx, &ctx->d_order, (size_t)npts)) return e;
    if (int e = dev_alloc(ctx, &ctx->d_pos, (size_t)npts)) return e;
    if (int e = dev_alloc(ctx, &ctx->d_neigh, (size_t)npts * k)) return e;
    std::vector<int> level_pts, level_off, rep_pts, rep_off;
    const int mode = ctx->start_mode;
    // the serial chain's launches: dependency levels, or, where a level is wider than the
    // chip holds at once, launches of at most one round filled with later points whose
    // neighbours are done
    auto chain_launches = [&](std::vector<int> &lp, std::vector<int> &lo) {
        dependency_levels(order, pos, neigh, npts, k, N, lp, lo);
        int bpc = 0;          // workgroups of the search kernel one CU holds
        if (search_max_blocks(bpv, bt, search_lds_bytes(nsamp_pad, bt, !table), &bpc) != 0 || bpc < 1)
            bpc = std::max(1, 512 / bt);
        const int cap = ctx->round_cap > 0 ? ctx->round_cap : ctx->cus * bpc;
        int widest = 0;
        for (size_t L = 0; L + 1 < lo.size(); L++) widest = std::max(widest, lo[L + 1] - lo[L]);
        if (ctx->round_cap >= 0 && widest > cap) capacity_schedule(order, pos, neigh, k, N, cap, lp, lo);
    };
    if (mode == DVC_START_CHAIN) {
        chain_launches(level_pts, level_off);
    } else {
        // no point reads a result of the same pass: one launch over the whole order
        level_pts.assign(order, order + N);
        level_off = {0, N};
        // the judged mode's repair launches run the serial chain's launches
        if (mode == DVC_START_JUDGED) chain_launches(rep_pts, rep_off);
    }
    // within each launch, workgroups in the XCD-aware spatial order (the points of one
    // launch do not read each other's results, so any order gives the same results)
    auto spatial = [&](std::vector<int> &lp, const std::vector<int> &lo) {
        std::vector<int> seg, idx;
        for (size_t L = 0; L + 1 < lo.size(); L++) {
            const int l0 = lo[L], cnt = lo[L + 1] - l0;
            if (cnt < 2) continue;
            spatial_block_order(points, lp.data() + l0, cnt, kXcds, idx);
            seg.assign(lp.begin() + l0, lp.begin() + l0 + cnt);
            for (int b = 0; b < cnt; b++) lp[l0 + b] = seg[idx[b]];
        }
    };
    spatial(level_pts, level_off);
    spatial(rep_pts, rep_off);
    if (int e = dev_alloc(ctx, &ctx->d_levels, level_pts.size())) return e;
    if (int e = dev_alloc(ctx, &ctx->d_res, (size_t)npts)) return e;
    ctx->starts_set = false;
    if (mode != DVC_START_CHAIN) {
        if (int e = dev_alloc(ctx, &ctx->d_start, (size_t)npts * 12)) return e;
        HIPCHK(hipMemsetAsync(ctx->d_start, 0, sizeof(double) * 12 * (size_t)npts, ctx->stream));
    }
    if (mode == DVC_START_JUDGED) {
        if (int e = dev_alloc(ctx, &ctx->d_res_prop, (size_t)npts)) return e;
        if (int e = dev_alloc(ctx, &ctx->d_res_judge, (size_t)npts)) return e;
        if (int e = dev_alloc(ctx, &ctx->d_kept, (size_t)npts)) return e;
        if (int e = dev_alloc(ctx, &ctx->d_repair, (size_t)npts)) return e;
        if (int e = dev_alloc(ctx, &ctx->d_rep_levels, rep_pts.size())) return e;
        HIPCHK(hipMemsetAsync(ctx->d_kept, 0, sizeof(int) * (size_t)npts, ctx->stream));
        HIPCHK(hipMemsetAsync(ctx->d_repair, 0, sizeof(int) * (size_t)npts, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->d_rep_levels, rep_pts.data(), sizeof(int) * rep_pts.size(), hipMemcpyHostToDevice,
                              ctx->stream));
        HIPCHK(hipHostMalloc((void **)&ctx->h_repair, sizeof(int) * (size_t)npts, hipHostMallocDefault));
    }
    std::vector<int> ref_order;
    if (table) {
        if (int e = dev_alloc(ctx, &ctx->d_rvals, table_floats)) return e;
        if (int e = dev_alloc(ctx, &ctx->d_rstat, (size_t)N)) return e;
        spatial_block_order(points, order, N, kXcds, ref_order);
        if (int e = dev_alloc(ctx, &ctx->d_ref_order, (size_t)N)) return e;
        HIPCHK(hipMemcpyAsync(ctx->d_ref_order, ref_order.data(), sizeof(int) * (size_t)N, hipMemcpyHostToDevice,
                              ctx->stream));
    } else if (ctx->d_ref_order) {
        hipFree(ctx->d_ref_order);
        ctx->d_ref_order = nullptr;
    }
    if (uneven) {
        if (int e = dev_alloc(ctx, &ctx->d_pgeo, (size_t)npts)) return e;
        HIPCHK(hipMemcpyAsync(ctx->d_pgeo, pgeo.data(), sizeof(PtGeo) * npts, hipMemcpyHostToDevice, ctx->stream));
    } else if (ctx->d_pgeo) {
        hipFree(ctx->d_pgeo);
        ctx->d_pgeo = nullptr;
    }
    HIPCHK(hipMemcpyAsync(ctx->d_pts, points, sizeof(double) * 3 * npts, hipMemcpyHostToDevice, ctx->stream));
    if (!sph.empty()) HIPCHK(hipMemcpyAsync(ctx->d_sph, sph.data(), sizeof(double) * sph.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_order, order, sizeof(int) * npts, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_pos, pos.data(), sizeof(int) * npts, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_neigh, neigh, sizeof(int) * (size_t)npts * k, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_levels, level_pts.data(), sizeof(int) * level_pts.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->rc = *rc;
    ctx->geo = G;
    ctx->npts = npts; ctx->nproc = N; ctx->k = k;
    for (int q = 0; q < 3; q++) { ctx->off_min[q] = omn[q]; ctx->off_max[q] = omx[q]; }
    for (int q = 0; q < 3; q++) { ctx->pts_mn[q] = 1e300; ctx->pts_mx[q] = -1e300; }
    for (int i = 0; i < npts; i++)
        for (int q = 0; q < 3; q++) {
            ctx->pts_mn[q] = std::min(ctx->pts_mn[q], points[3 * i + q]);
            ctx->pts_mx[q] = std::max(ctx->pts_mx[q], points[3 * i + q]);
        }
    ctx->chosen_threads = bt;
    ctx->ref_threads = 256;
    ctx->level_off = level_off;
    ctx->level_pts_h = level_pts;
    ctx->rep_off = rep_off;
    ctx->rep_pts_h = rep_pts;
    ctx->pt_nsamp.clear();
    if (uneven) {
        ctx->pt_nsamp.resize(npts);
        for (int i = 0; i < npts; i++) ctx->pt_nsamp[i] = pgeo[i].nsamp;
    }
    ctx->plan_mode = mode;
    ctx->searched.assign(npts, 0);
    for (int i = 0; i < npts; i++) ctx->searched[i] = pos[i] < N ? 1 : 0;
    ctx->plan_ready = true;
    return DVC_OK;
}

// Slice launches per chain launch before the one-workgroup search resumes.
static int slice_evals_of(const dvc_ctx *ctx) { return ctx->slice_evals > 0 ? ctx->slice_evals : 4; }

// Samples per point split into workgroups of ctx->slice_per samples (with per-point
// geometry: of the point with the most samples).
static int slices_of(const dvc_ctx *ctx) {
    int ns = ctx->geo.subv_num;
    for (int v : ctx->pt_nsamp) ns = std::max(ns, v);
    return (ns + ctx->slice_per - 1) / ctx->slice_per;
}

// Threads per slice workgroup: one sample per thread up to 512, more above.
static int slice_threads_of(const dvc_ctx *ctx) { return ctx->slice_per >= kSliceThreads[1] ? kSliceThreads[1] : kSliceThreads[0]; }

// Workgroups per point of a chain launch run as slice launches (points pts[0..cnt), in
// workgroup-slot order): one per slice_per samples of its largest point.  A point's
// slices share one XCD (its slot mod 8): with more of them than the XCD has CUs, some
// CUs would run two slices one after the other, so wide launches take fewer, larger
// slices (the kernels deal a point's 64-sample chunks to however many slices it has).
// With per-point geometry *cap is the largest cap c >= 2 on every point's slices that
// leaves no XCD more slices than CUs (0: none needed, or even 2 does not fit).
static int launch_slices(const dvc_ctx *ctx, const int *pts, int cnt, int *cap) {
    const bool uneven = !ctx->pt_nsamp.empty();
    auto slices = [&](int pt) {
        return ((uneven ? ctx->pt_nsamp[pt] : ctx->geo.subv_num) + ctx->slice_per - 1) / ctx->slice_per;
    };
    int nsl = 0;
    for (int q = 0; q < cnt; q++) nsl = std::max(nsl, slices(pts[q]));
    *cap = 0;
    const int cus_xcd = std::max(1, ctx->cus / kXcds);
    if (!uneven) {
        const int c = cus_xcd / std::max(1, (cnt + kXcds - 1) / kXcds);
        return c >= 2 && nsl > c ? c : nsl;
    }
    auto fits = [&](int c) {
        int load[kXcds] = {0};
        for (int q = 0; q < cnt; q++) load[q % kXcds] += std::min(slices(pts[q]), c);
        for (int x = 0; x < kXcds; x++)
            if (load[x] > cus_xcd) return false;
        return true;
    };
    if (fits(nsl)) return nsl;
    int c = nsl - 1;
    while (c >= 2 && !fits(c)) c--;
    if (c < 2) return nsl;
    *cap = c;
    return c;
}

// Which chain launches of the prepared search run as slice launches, their workgroups per
// point, and their buffers.  Slice launches serve the serial chain with no translation
// grid and the HBM table of reference samples.  With per-point geometry a chain launch
// takes the slice count of its point with the most samples.  Called outside stream
// capture.
static int plan_slices(dvc_ctx *ctx) {
    const int nl = ctx->level_off.empty() ? 0 : (int)ctx->level_off.size() - 1;
    ctx->sliced.assign(nl, 0);
    ctx->slice_nsl.assign(nl, 0);
    ctx->slice_cap.assign(nl, 0);
    const bool eligible = ctx->plan_mode == DVC_START_CHAIN && ctx->d_rvals && ctx->rc.basin_radius == 0.0 &&
                          ctx->slice_wgs >= 0;
    if (!eligible) return DVC_OK;
    const int max_wgs = ctx->slice_wgs > 0 ? ctx->slice_wgs : 4 * ctx->cus;
    for (int L = 0; L < nl; L++)
        ctx->slice_nsl[L] = launch_slices(ctx, ctx->level_pts_h.data() + ctx->level_off[L],
                                          ctx->level_off[L + 1] - ctx->level_off[L], &ctx->slice_cap[L]);
    const int cnt8 = slice_plan(ctx->level_off, ctx->slice_nsl, max_wgs, kXcds, ctx->sliced);
    if (cnt8 == 0) return DVC_OK;
    const int nslice = slices_of(ctx);
    if (cnt8 > ctx->slice_cnt8 || nslice > ctx->slice_n || !ctx->d_state) {
        const int c8 = std::max(cnt8, ctx->slice_cnt8), ns = std::max(nslice, ctx->slice_n);
        if (int e = dev_alloc(ctx, &ctx->d_state, (size_t)2 * c8 * kStateStride)) return e;
        if (int e = dev_alloc(ctx, &ctx->d_part, (size_t)2 * c8 * ns * kPartStride)) return e;
        ctx->slice_cnt8 = c8;
        ctx->slice_n = ns;
    }
    return DVC_OK;
}

static int enqueue_search_only(dvc_ctx *ctx, hipStream_t st);

// The strain fit of the search's results (strain.cpp:402-591), enqueued after the search.
static int enqueue_strain(dvc_ctx *ctx, hipStream_t st) {
    const int n = ctx->npts;
    if (launch_fields(ctx->d_res, ctx->d_pos, n, ctx->nproc, ctx->d_sdisp, ctx->d_sstat, ctx->d_sobj, st) != 0)
        return fail(ctx, DVC_ERR_HIP, "strain input launch failed");
    StrainArgs a;
    a.pts = ctx->d_pts; a.disp = ctx->d_sdisp; a.objmin = ctx->d_sobj; a.status = ctx->d_sstat; a.neigh = ctx->d_neigh;
    a.k = ctx->k; a.n = n; a.ndp = ctx->strain_ndp; a.refill = ctx->strain_refill;
    // plane detection (strain.cpp:340-365)
    a.yz = (ctx->strain_planar & 1) || std::fabs(ctx->pts_mx[0] - ctx->pts_mn[0]) < 0.001;
    a.zx = (ctx->strain_planar & 2) || std::fabs(ctx->pts_mx[1] - ctx->pts_mn[1]) < 0.001;
    a.xy = (ctx->strain_planar & 4) || std::fabs(ctx->pts_mx[2] - ctx->pts_mn[2]) < 0.001;
    a.objt = ctx->strain_objt;
    double *o = ctx->d_sout;
    a.out_disp = o; a.out_swr = o + 3 * (size_t)n; a.out_E = o + 4 * (size_t)n; a.out_L = o + 13 * (size_t)n;
    a.out_DG = o + 22 * (size_t)n;
    a.out_npts = ctx->d_snpts;
    if (launch_strain(a, st) != 0) return fail(ctx, DVC_ERR_HIP, "strain launch failed");
    return DVC_OK;
}

// Everything a prepared search enqueues up to its results; in the judged mode up to the
// selection (the repair launches and the strain fit follow in dvc_search_execute, once
// the points to repair are known on the host).
static int enqueue_search(dvc_ctx *ctx, hipStream_t st) {
    if (int e = enqueue_search_only(ctx, st)) return e;
    if (ctx->strain_ndp > 0 && ctx->plan_mode != DVC_START_JUDGED) return enqueue_strain(ctx, st);
    return DVC_OK;
}

// Enqueue the whole search on `st`: clear the results, then one launch per dependency
// level (a level's points only read results of earlier levels).  A level run as slice
// launches takes slice_evals_of(ctx) launches of the slice kernel, then the
// one-workgroup search resumes the points that are still searching.
// The search kernels' arguments of a prepared search (serial chain: results and start
// guesses in d_res).
static void prepared_args(dvc_ctx *ctx, SearchArgs &a) {
    fill_search_args(a, ctx->rc, ctx->geo, ctx->vol[0].bpv, (int)ctx->vol[0].W, (int)ctx->vol[0].H,
                     (int)ctx->vol[0].T, ctx->npts, ctx->nproc, ctx->k, ctx->off_min, ctx->off_max);
    a.vol[0] = ctx->vol[0].d; a.vol[1] = ctx->vol[1].d;
    if (((uintptr_t)a.vol[0] | (uintptr_t)a.vol[1]) & 3) a.gather32 = 0;   // attached, unaligned
    a.sph = ctx->d_sph; a.pts = ctx->d_pts; a.order = ctx->d_order; a.pos = ctx->d_pos; a.neigh = ctx->d_neigh;
    a.res = ctx->d_res;
    a.nbres = ctx->d_res;
    a.start = nullptr;
    a.rvals = ctx->d_rvals;
    a.rstat = ctx->d_rstat;
    a.pgeo = ctx->d_pgeo;
}

// The judged mode's repair: the points where the passes do not both converge and agree
// (h_repair, copied from the selection), searched again in the serial chain's launches
// from the reference's start guess over the final results.  Only the launches that hold
// such a point are enqueued; in them every other point leaves its workgroup at once.
// Where the serial chain would slice a launch (same rules as plan_slices), the repair
// runs it as slice launches too: a repaired point's search then spreads over several
// CUs instead of one workgroup.
static int enqueue_repairs(dvc_ctx *ctx, hipStream_t st) {
    SearchArgs a;
    prepared_args(ctx, a);
    a.repair = ctx->d_repair;
    const int bt = ctx->chosen_threads;
    const size_t lds = search_lds_bytes(a.nsamp_pad, bt, a.rvals == nullptr);
    const int depth = (int)ctx->rep_off.size() - 1;
    const bool sliceable = ctx->d_rvals && ctx->rc.basin_radius == 0.0 && ctx->slice_wgs >= 0;
    const int nsl0 = slices_of(ctx);
    const int max_wgs = ctx->slice_wgs > 0 ? ctx->slice_wgs : 4 * ctx->cus;
    for (int L = 0; L < depth; L++) {
        const int q0 = ctx->rep_off[L], q1 = ctx->rep_off[L + 1];
        bool any = false;
        for (int q = q0; q < q1 && !any; q++) any = ctx->h_repair[ctx->rep_pts_h[q]] != 0;
        if (!any) continue;
        const int cnt = q1 - q0, cnt8 = (cnt + kXcds - 1) / kXcds * kXcds;
        int cap = 0;
        const int nsl = launch_slices(ctx, ctx->rep_pts_h.data() + q0, cnt, &cap);
        if (sliceable && nsl >= 2 && (long long)cnt8 * nsl <= max_wgs) {
            if (cnt8 > ctx->slice_cnt8 || nsl0 > ctx->slice_n || !ctx->d_state) {
                const int c8 = std::max(cnt8, ctx->slice_cnt8), ns = std::max(nsl0, ctx->slice_n);
                if (int e = dev_alloc(ctx, &ctx->d_state, (size_t)2 * c8 * kStateStride)) return e;
                if (int e = dev_alloc(ctx, &ctx->d_part, (size_t)2 * c8 * ns * kPartStride)) return e;
                ctx->slice_cnt8 = c8;
                ctx->slice_n = ns;
            }
            SliceArgs sa{};
            sa.per = ctx->slice_per;
            for (int q = 0; q < 2; q++) {
                sa.state[q] = ctx->d_state + (size_t)q * ctx->slice_cnt8 * kStateStride;
                sa.part[q] = ctx->d_part + (size_t)q * ctx->slice_cnt8 * ctx->slice_n * kPartStride;
            }
            sa.level_pts = ctx->d_rep_levels + q0;
            sa.cnt = cnt;
            sa.cnt8 = cnt8;
            sa.nslice = nsl;
            sa.maxsl = cap;
            for (int t = 0; t < slice_evals_of(ctx); t++) {
                sa.t = t;
                if (launch_slice(a, sa, slice_threads_of(ctx), st) != 0)
                    return fail(ctx, DVC_ERR_HIP, "repair slice launch failed");
            }
            sa.t = slice_evals_of(ctx);
            if (launch_resume(a, sa, st) != 0) return fail(ctx, DVC_ERR_HIP, "repair resume launch failed");
        } else if (launch_search(a, ctx->d_rep_levels + q0, bt, cnt, lds, st) != 0) {
            return fail(ctx, DVC_ERR_HIP, "repair launch failed");
        }
    }
    return DVC_OK;
}

static int enqueue_search_only(dvc_ctx *ctx, hipStream_t st) {
    SearchArgs a;
    prepared_args(ctx, a);
    int bt = ctx->chosen_threads;
    size_t lds = search_lds_bytes(a.nsamp_pad, bt, a.rvals == nullptr);
    const long long res_words = (long long)ctx->npts * (long long)(sizeof(dvc_point_result) / 4);
    if (launch_zero_int((int *)ctx->d_res, res_words, st) != 0)
        return fail(ctx, DVC_ERR_HIP, "zero launch");
    a.ref_order = ctx->d_ref_order;
    if (a.rvals && launch_reference(a, ctx->ref_threads, st) != 0)
        return fail(ctx, DVC_ERR_HIP, "reference kernel launch failed");
    a.ref_order = nullptr;
    if (ctx->plan_mode == DVC_START_CHAIN) {
        const int depth = (int)ctx->level_off.size() - 1;
        SliceArgs sa{};
        sa.per = ctx->slice_per;
        for (int q = 0; q < 2; q++) {
            sa.state[q] = ctx->d_state ? ctx->d_state + (size_t)q * ctx->slice_cnt8 * kStateStride : nullptr;
            sa.part[q] = ctx->d_part ? ctx->d_part + (size_t)q * ctx->slice_cnt8 * ctx->slice_n * kPartStride : nullptr;
        }
        const int sthreads = slice_threads_of(ctx);
        for (int L = 0; L < depth; L++) {
            const int cnt = ctx->level_off[L + 1] - ctx->level_off[L];
            const int *lp = ctx->d_levels + ctx->level_off[L];
            if (L < (int)ctx->sliced.size() && ctx->sliced[L]) {
                sa.level_pts = lp;
                sa.cnt = cnt;
                sa.nslice = ctx->slice_nsl[L];
                sa.maxsl = ctx->slice_cap[L];
                sa.cnt8 = (cnt + kXcds - 1) / kXcds * kXcds;
                for (int t = 0; t < slice_evals_of(ctx); t++) {
                    sa.t = t;
                    if (launch_slice(a, sa, sthreads, st) != 0) return fail(ctx, DVC_ERR_HIP, "slice kernel launch failed");
                }
                sa.t = slice_evals_of(ctx);
                if (launch_resume(a, sa, st) != 0) return fail(ctx, DVC_ERR_HIP, "resume kernel launch failed");
            } else if (launch_search(a, lp, bt, cnt, lds, st) != 0) {
                return fail(ctx, DVC_ERR_HIP, "search kernel launch failed");
            }
        }
        return DVC_OK;
    }
    // proposal pass: every point from its given start warp, one launch
    a.start = ctx->d_start;
    if (ctx->plan_mode == DVC_START_JUDGED) {
        a.res = ctx->d_res_prop;
        if (launch_zero_int((int *)ctx->d_res_prop, res_words, st) != 0 ||
            launch_zero_int((int *)ctx->d_res_judge, res_words, st) != 0)
            return fail(ctx, DVC_ERR_HIP, "zero launch");
    }
    if (launch_search(a, ctx->d_levels, bt, ctx->nproc, lds, st) != 0)
        return fail(ctx, DVC_ERR_HIP, "search kernel launch failed");
    if (ctx->plan_mode == DVC_START_JUDGED) {
        // judge pass: the reference's start guess on the proposal pass's results
        a.start = nullptr;
        a.nbres = ctx->d_res_prop;
        a.res = ctx->d_res_judge;
        if (launch_search(a, ctx->d_levels, bt, ctx->nproc, lds, st) != 0)
            return fail(ctx, DVC_ERR_HIP, "search kernel launch failed");
        if (launch_judge(ctx->d_order, ctx->nproc, ctx->d_res_prop, ctx->d_res_judge, ctx->d_res, ctx->d_kept,
                         ctx->d_repair, st) != 0)
            return fail(ctx, DVC_ERR_HIP, "judge launch failed");
    }
    return DVC_OK;
}

int dvc_volume_download(dvc_ctx *ctx, int slot, void *host) {
    if (!ctx || !host || slot < 0 || slot > 1) return DVC_ERR_ARG;
    const Volume &v = ctx->vol[slot];
    if (!v.d) return fail(ctx, DVC_ERR_STATE, "no volume in that slot");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(host, v.d, (size_t)(v.W * v.H * v.T * v.bpv), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DVC_OK;
}

int dvc_device_alloc(dvc_ctx *ctx, void **ptr, int64_t bytes) {
    if (!ctx || !ptr || bytes < 0) return DVC_ERR_ARG;
    *ptr = nullptr;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMalloc(ptr, (size_t)std::max<int64_t>(bytes, 1)));
    return DVC_OK;
}

int dvc_device_free(dvc_ctx *ctx, void *ptr) {
    if (!ctx) return DVC_ERR_ARG;
    if (!ptr) return DVC_OK;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipFree(ptr));
    return DVC_OK;
}

int dvc_device_copy(dvc_ctx *ctx, void *dst, const void *src, int64_t bytes, int kind) {
    if (!ctx || bytes < 0 || (bytes > 0 && (!dst || !src)) || kind < 0 || kind > 2) return DVC_ERR_ARG;
    if (bytes == 0) return DVC_OK;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(dst, src, (size_t)bytes, k, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DVC_OK;
}

int dvc_synchronize(dvc_ctx *ctx) {
    if (!ctx) return DVC_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DVC_OK;
}

int dvc_set_point_geometry(dvc_ctx *ctx, const int32_t *geom, const int32_t *size, const int32_t *nsamp, int32_t n) {
    if (!ctx || n < 0) return DVC_ERR_ARG;
    if (n == 0 || !geom || !size || !nsamp) {
        ctx->pg_geom.clear(); ctx->pg_size.clear(); ctx->pg_npts.clear();
        return DVC_OK;
    }
    ctx->pg_geom.assign(geom, geom + n);
    ctx->pg_size.assign(size, size + n);
    ctx->pg_npts.assign(nsamp, nsamp + n);
    return DVC_OK;
}

int dvc_set_start_mode(dvc_ctx *ctx, int mode) {
    if (!ctx) return DVC_ERR_ARG;
    if (mode != DVC_START_CHAIN && mode != DVC_START_GIVEN && mode != DVC_START_JUDGED)
        return fail(ctx, DVC_ERR_ARG, "unknown start mode");
    ctx->start_mode = mode;
    return DVC_OK;
}

int dvc_search_set_starts(dvc_ctx *ctx, const double *start, int on_device) {
    if (!ctx || !start) return DVC_ERR_ARG;
    if (!ctx->plan_ready) return fail(ctx, DVC_ERR_STATE, "dvc_search_prepare first");
    if (ctx->plan_mode == DVC_START_CHAIN)
        return fail(ctx, DVC_ERR_STATE, "the prepared search takes its starts from the serial chain");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(ctx->d_start, start, sizeof(double) * 12 * (size_t)ctx->npts,
                          on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    ctx->starts_set = true;
    return DVC_OK;
}

int dvc_search_fetch_passes(dvc_ctx *ctx, dvc_point_result *proposal, dvc_point_result *judge, int32_t *kept) {
    if (!ctx) return DVC_ERR_ARG;
    if (!ctx->plan_ready) return fail(ctx, DVC_ERR_STATE, "dvc_search_prepare first");
    if (ctx->plan_mode != DVC_START_JUDGED) return fail(ctx, DVC_ERR_STATE, "the prepared search has no judge pass");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)ctx->npts;
    if (proposal) HIPCHK(hipMemcpyAsync(proposal, ctx->d_res_prop, sizeof(dvc_point_result) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (judge) HIPCHK(hipMemcpyAsync(judge, ctx->d_res_judge, sizeof(dvc_point_result) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (kept) HIPCHK(hipMemcpyAsync(kept, ctx->d_kept, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < n; i++)
        if (!ctx->searched[i]) {
            dvc_point_result *rs[2] = {proposal, judge};
            for (dvc_point_result *r : rs)
                if (r) {
                    r[i].status = DVC_NOT_SEARCH; r[i].iters = 0; r[i].obj = 0.0;
                    for (int q = 0; q < 12; q++) r[i].par[q] = 0.0;
                }
            if (kept) kept[i] = -1;
        }
    return DVC_OK;
}

int dvc_search_execute(dvc_ctx *ctx) {
    if (!ctx) return DVC_ERR_ARG;
    if (!ctx->plan_ready) return fail(ctx, DVC_ERR_STATE, "dvc_search_prepare first");
    if (ctx->plan_mode != DVC_START_CHAIN && !ctx->starts_set)
        return fail(ctx, DVC_ERR_STATE, "dvc_search_set_starts first");
    if (int e = check_volumes(ctx)) return e;
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->gexec) {
        if (int e = plan_slices(ctx)) return e;
        if (ctx->strain_ndp > 0 && !ctx->d_sout) {
            const size_t n = (size_t)ctx->npts;
            if (int e = dev_alloc(ctx, &ctx->d_sdisp, 3 * n)) return e;
            if (int e = dev_alloc(ctx, &ctx->d_sobj, n)) return e;
            if (int e = dev_alloc(ctx, &ctx->d_sstat, n)) return e;
            if (int e = dev_alloc(ctx, &ctx->d_snpts, n)) return e;
            if (int e = dev_alloc(ctx, &ctx->d_sout, 31 * n)) return e;
        }
    }
    if (ctx->use_graph && !ctx->gexec) {
        if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return fail(ctx, DVC_ERR_HIP, "stream capture could not start");
        int e = enqueue_search(ctx, ctx->stream);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        if (e != 0) { if (g) hipGraphDestroy(g); return e; }
        if (ce != hipSuccess) return hip_fail(ctx, ce, "hipStreamEndCapture");
        ctx->graph = g;
        hipError_t ie = hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0);
        if (ie != hipSuccess) { drop_graph(ctx); return hip_fail(ctx, ie, "hipGraphInstantiate"); }
    }
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    if (ctx->use_graph) {
        HIPCHK(hipGraphLaunch(ctx->gexec, ctx->stream));
    } else if (int e = enqueue_search(ctx, ctx->stream)) {
        return e;
    }
    if (ctx->plan_mode == DVC_START_JUDGED) {
        // the points to repair, then the repair launches that hold one, and the strain fit
        HIPCHK(hipMemcpyAsync(ctx->h_repair, ctx->d_repair, sizeof(int) * (size_t)ctx->npts, hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (int e = enqueue_repairs(ctx, ctx->stream)) return e;
        if (ctx->strain_ndp > 0)
            if (int e = enqueue_strain(ctx, ctx->stream)) return e;
    }
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    ctx->timing_pending = true;
    return DVC_OK;
}

int dvc_search_autotune(dvc_ctx *ctx, int reps, int32_t *best_threads, float *best_ms) {
    if (!ctx || reps < 1 || reps > 100) return DVC_ERR_ARG;
    if (!ctx->plan_ready) return fail(ctx, DVC_ERR_STATE, "dvc_search_prepare first");
    if (ctx->plan_mode != DVC_START_CHAIN && !ctx->starts_set)
        return fail(ctx, DVC_ERR_STATE, "dvc_search_set_starts first");
    const int bpv = ctx->vol[0].bpv;
    const int nsamp_pad = (ctx->geo.subv_num + 3) & ~3;
    const int cands[5] = {64, 128, 256, 512, 1024};
    int best = 0;
    float bms = 0.0f;
    for (int c : cands) {
        size_t st = 0;
        if (search_static_lds(bpv, c, &st) != 0) return fail(ctx, DVC_ERR_HIP, "kernel attributes");
        if (search_lds_bytes(nsamp_pad, c, ctx->d_rvals == nullptr) + st > ctx->lds_max) continue;
        ctx->chosen_threads = c;
        drop_graph(ctx);
        if (int e = dvc_search_execute(ctx)) return e;          // warm-up, captures the graph
        float tot = 0.0f;
        for (int r = 0; r < reps; r++) {
            if (int e = dvc_search_execute(ctx)) return e;
            float ms = 0.0f;
            if (int e = dvc_last_kernel_ms(ctx, &ms)) return e;
            tot += ms;
        }
        const float avg = tot / (float)reps;
        if (best == 0 || avg < bms) { best = c; bms = avg; }
    }
    if (best == 0) return fail(ctx, DVC_ERR_STATE, "no workgroup size fits");
    ctx->chosen_threads = best;
    // the reference kernel is one launch over every point, throughput-bound: its own
    // workgroup size, timed with the chosen search size
    if (ctx->d_rvals) {
        int best_ref = ctx->ref_threads;
        for (int c : {128, 512, 256}) {
            if (c == best_ref) continue;
            const int keep = ctx->ref_threads;
            ctx->ref_threads = c;
            drop_graph(ctx);
            if (int e = dvc_search_execute(ctx)) return e;
            float tot = 0.0f;
            for (int r = 0; r < reps; r++) {
                if (int e = dvc_search_execute(ctx)) return e;
                float ms = 0.0f;
                if (int e = dvc_last_kernel_ms(ctx, &ms)) return e;
                tot += ms;
            }
            const float avg = tot / (float)reps;
            if (avg < bms) { bms = avg; best_ref = c; }
            else ctx->ref_threads = keep;
        }
        ctx->ref_threads = best_ref;
    }
    // slice launches of the chain (unless switched off): off, and a few widths, launch
    // counts and samples per workgroup (more launches and larger slices for the long
    // searches and large clouds of per-point geometry), the fastest kept
    if (ctx->plan_mode == DVC_START_CHAIN && ctx->slice_wgs >= 0 && ctx->d_rvals && ctx->rc.basin_radius == 0.0) {
        const int C = ctx->cus;
        const int cands[][3] = {{-1, 0, 512},      {2 * C, 4, 512},  {4 * C, 3, 512},  {4 * C, 4, 512},
                                {4 * C, 6, 512},   {8 * C, 4, 512},  {4 * C, 7, 1024}, {8 * C, 7, 1024},
                                {8 * C, 14, 1024}, {8 * C, 14, 2048}};
        int bw = ctx->slice_wgs, be = ctx->slice_evals, bp = ctx->slice_per;
        float bsl = -1.0f;
        for (const auto &c : cands) {
            ctx->slice_wgs = c[0];
            ctx->slice_evals = c[1];
            ctx->slice_per = c[2];
            drop_graph(ctx);
            if (int e = dvc_search_execute(ctx)) return e;
            float tot = 0.0f;
            for (int r = 0; r < reps; r++) {
                if (int e = dvc_search_execute(ctx)) return e;
                float ms = 0.0f;
                if (int e = dvc_last_kernel_ms(ctx, &ms)) return e;
                tot += ms;
            }
            const float avg = tot / (float)reps;
            if (bsl < 0.0f || avg < bsl) { bsl = avg; bw = c[0]; be = c[1]; bp = c[2]; }
        }
        ctx->slice_wgs = bw;
        ctx->slice_evals = be;
        ctx->slice_per = bp;
        bms = bsl;
    }
    drop_graph(ctx);
    if (best_threads) *best_threads = best;
    if (best_ms) *best_ms = bms;
    return DVC_OK;
}

int dvc_set_slice_launches(dvc_ctx *ctx, int32_t max_workgroups, int32_t evaluations, int32_t samples_per_workgroup) {
    if (!ctx) return DVC_ERR_ARG;
    if (max_workgroups < -1) return fail(ctx, DVC_ERR_ARG, "max_workgroups must be -1, 0 or positive");
    if (evaluations < 0 || evaluations > 20) return fail(ctx, DVC_ERR_ARG, "evaluations must be in 0..20");
    if (samples_per_workgroup != 0 && samples_per_workgroup != 256 &&
        (samples_per_workgroup < 512 || samples_per_workgroup > 8192 || samples_per_workgroup % 512 != 0))
        return fail(ctx, DVC_ERR_ARG, "samples_per_workgroup must be 0, 256 or a multiple of 512 up to 8192");
    drop_graph(ctx);
    ctx->slice_wgs = max_workgroups;
    ctx->slice_evals = evaluations;
    ctx->slice_per = samples_per_workgroup ? samples_per_workgroup : 512;
    return DVC_OK;
}

int dvc_search_slice_info(dvc_ctx *ctx, int32_t *sliced_launches, int32_t *slices_per_point, int32_t *evaluations) {
    if (!ctx) return DVC_ERR_ARG;
    if (!ctx->plan_ready) return fail(ctx, DVC_ERR_STATE, "dvc_search_prepare first");
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->gexec)
        if (int e = plan_slices(ctx)) return e;
    int n = 0;
    for (char c : ctx->sliced) n += c ? 1 : 0;
    if (sliced_launches) *sliced_launches = n;
    if (slices_per_point) *slices_per_point = n ? slices_of(ctx) : 0;
    if (evaluations) *evaluations = n ? slice_evals_of(ctx) : 0;
    return DVC_OK;
}

int dvc_search_set_strain(dvc_ctx *ctx, int32_t ndp, double objt, int32_t refill, int32_t planar) {
    if (!ctx) return DVC_ERR_ARG;
    if (ndp < 0 || ndp > 75) return fail(ctx, DVC_ERR_ARG, "ndp must be 0 (off) or in 1..75");
    if (planar < 0 || planar > 7) return fail(ctx, DVC_ERR_ARG, "planar must be a 3-bit mask");
    drop_graph(ctx);
    ctx->strain_ndp = ndp;
    ctx->strain_objt = objt;
    ctx->strain_refill = refill ? 1 : 0;
    ctx->strain_planar = planar;
    return DVC_OK;
}

int dvc_search_fetch_strain(dvc_ctx *ctx, double *out_disp, int32_t *out_npts, double *out_swr, double *out_E,
                            double *out_L, double *out_DG) {
    if (!ctx) return DVC_ERR_ARG;
    if (!ctx->plan_ready || ctx->strain_ndp <= 0 || !ctx->d_sout)
        return fail(ctx, DVC_ERR_STATE, "no strain fit: dvc_search_set_strain, prepare and execute first");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)ctx->npts;
    const double *o = ctx->d_sout;
    struct { void *dst; const void *src; size_t bytes; } cp[6] = {
        {out_disp, o, 3 * n * sizeof(double)}, {out_swr, o + 3 * n, n * sizeof(double)},
        {out_E, o + 4 * n, 9 * n * sizeof(double)}, {out_L, o + 13 * n, 9 * n * sizeof(double)},
        {out_DG, o + 22 * n, 9 * n * sizeof(double)}, {out_npts, ctx->d_snpts, n * sizeof(int32_t)}};
    for (auto &c : cp)
        if (c.dst) HIPCHK(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DVC_OK;
}

int dvc_search_info(dvc_ctx *ctx, int32_t *levels, int32_t *threads, int32_t *searched) {
    if (!ctx) return DVC_ERR_ARG;
    if (!ctx->plan_ready) return fail(ctx, DVC_ERR_STATE, "dvc_search_prepare first");
    if (levels) *levels = (int32_t)ctx->level_off.size() - 1;
    if (threads) *threads = ctx->chosen_threads;
    if (searched) *searched = ctx->nproc;
    return DVC_OK;
}

int dvc_set_launch_mode(dvc_ctx *ctx, int graph) {
    if (!ctx) return DVC_ERR_ARG;
    if ((graph != 0) != (ctx->use_graph != 0)) drop_graph(ctx);
    ctx->use_graph = graph ? 1 : 0;
    return DVC_OK;
}

int dvc_search_fetch(dvc_ctx *ctx, dvc_point_result *results) {
    if (!ctx || !results) return DVC_ERR_ARG;
    if (!ctx->plan_ready) return fail(ctx, DVC_ERR_STATE, "dvc_search_prepare first");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(results, ctx->d_res, sizeof(dvc_point_result) * ctx->npts, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    if (ctx->timing_pending && hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_ms = ms;
    ctx->timing_pending = false;
    // points never searched keep the not_search record of DataCloud.cpp:168-170
    for (int i = 0; i < ctx->npts; i++)
        if (!ctx->searched[i]) {
            results[i].status = DVC_NOT_SEARCH;
            results[i].iters = 0;
            results[i].obj = 0.0;
            for (int q = 0; q < 12; q++) results[i].par[q] = 0.0;
        }
    return DVC_OK;
}

int dvc_search_points(dvc_ctx *ctx, const dvc_run_params *rc, const double *points, int32_t npts,
                      const int32_t *order, int32_t nproc, const int32_t *neigh, int32_t k,
                      dvc_point_result *results) {
    if (!ctx) return DVC_ERR_ARG;
    // always the serial loop's start guess, whatever mode later prepares use
    const int mode = ctx->start_mode;
    ctx->start_mode = DVC_START_CHAIN;
    int e = dvc_search_prepare(ctx, rc, points, npts, order, nproc, neigh, k);
    ctx->start_mode = mode;
    if (e) return e;
    if (int e2 = dvc_search_execute(ctx)) return e2;
    return dvc_search_fetch(ctx, results);
}

int dvc_strain_fit(dvc_ctx *ctx, const double *pts, const double *disp, const int32_t *status, const double *objmin,
                   const int32_t *neigh, int32_t k, int32_t n, int32_t ndp, double objt, int32_t refill, int32_t planar,
                   double *out_disp, int32_t *out_npts, double *out_swr, double *out_E, double *out_L, double *out_DG) {
    if (!ctx) return DVC_ERR_ARG;
    if (n < 1 || !pts || !disp || !status || !objmin || !neigh) return fail(ctx, DVC_ERR_ARG, "bad input arrays");
    if (!out_disp || !out_npts || !out_swr || !out_E || !out_L || !out_DG) return fail(ctx, DVC_ERR_ARG, "bad output arrays");
    if (k < 1 || k > kMaxNeigh) return fail(ctx, DVC_ERR_ARG, "k must be in 1..128");
    if (ndp < 1 || ndp > 75) return fail(ctx, DVC_ERR_ARG, "ndp must be in 1..75");
    for (int64_t i = 0; i < (int64_t)n * k; i++)
        if (neigh[i] < 0 || neigh[i] >= n) return fail(ctx, DVC_ERR_ARG, "neighbour index out of range");
    // plane detection (strain.cpp:340-365)
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n; i++)
        for (int q = 0; q < 3; q++) { mn[q] = std::min(mn[q], pts[3 * i + q]); mx[q] = std::max(mx[q], pts[3 * i + q]); }
    int yz = (planar & 1) || std::fabs(mx[0] - mn[0]) < 0.001;
    int zx = (planar & 2) || std::fabs(mx[1] - mn[1]) < 0.001;
    int xy = (planar & 4) || std::fabs(mx[2] - mn[2]) < 0.001;
    HIPCHK(hipSetDevice(ctx->device));
    size_t nd = (size_t)n;
    size_t bytes = sizeof(double) * (3 * nd + 3 * nd + nd + 3 * nd + nd + 9 * nd * 3) + sizeof(int) * (nd + nd * k + nd);
    uint8_t *buf = nullptr;
    HIPCHK(hipMalloc((void **)&buf, bytes + 256));
    double *d_pts = (double *)buf, *d_disp = d_pts + 3 * nd, *d_obj = d_disp + 3 * nd;
    double *d_odisp = d_obj + nd, *d_swr = d_odisp + 3 * nd, *d_E = d_swr + nd, *d_L = d_E + 9 * nd, *d_DG = d_L + 9 * nd;
    int *d_st = (int *)(d_DG + 9 * nd), *d_nb = d_st + nd, *d_np = d_nb + nd * k;
    hipMemcpyAsync(d_pts, pts, sizeof(double) * 3 * nd, hipMemcpyHostToDevice, ctx->stream);
    hipMemcpyAsync(d_disp, disp, sizeof(double) * 3 * nd, hipMemcpyHostToDevice, ctx->stream);
    hipMemcpyAsync(d_obj, objmin, sizeof(double) * nd, hipMemcpyHostToDevice, ctx->stream);
    hipMemcpyAsync(d_st, status, sizeof(int) * nd, hipMemcpyHostToDevice, ctx->stream);
    hipMemcpyAsync(d_nb, neigh, sizeof(int) * nd * k, hipMemcpyHostToDevice, ctx->stream);
    StrainArgs a;
    a.pts = d_pts; a.disp = d_disp; a.objmin = d_obj; a.status = d_st; a.neigh = d_nb;
    a.k = k; a.n = n; a.ndp = ndp; a.refill = refill ? 1 : 0;
    a.yz = yz; a.zx = zx; a.xy = xy; a.objt = objt;
    a.out_disp = d_odisp; a.out_swr = d_swr; a.out_E = d_E; a.out_L = d_L; a.out_DG = d_DG; a.out_npts = d_np;
    hipEventRecord(ctx->ev0, ctx->stream);
    int r = launch_strain(a, ctx->stream);
    hipEventRecord(ctx->ev1, ctx->stream);
    if (r == 0) {
        hipMemcpyAsync(out_disp, d_odisp, sizeof(double) * 3 * nd, hipMemcpyDeviceToHost, ctx->stream);
        hipMemcpyAsync(out_swr, d_swr, sizeof(double) * nd, hipMemcpyDeviceToHost, ctx->stream);
        hipMemcpyAsync(out_E, d_E, sizeof(double) * 9 * nd, hipMemcpyDeviceToHost, ctx->stream);
        hipMemcpyAsync(out_L, d_L, sizeof(double) * 9 * nd, hipMemcpyDeviceToHost, ctx->stream);
        hipMemcpyAsync(out_DG, d_DG, sizeof(double) * 9 * nd, hipMemcpyDeviceToHost, ctx->stream);
        hipMemcpyAsync(out_npts, d_np, sizeof(int) * nd, hipMemcpyDeviceToHost, ctx->stream);
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    float ms = 0.0f;
    if (e == hipSuccess && hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_ms = ms;
    ctx->timing_pending = false;
    hipFree(buf);
    if (r != 0) return fail(ctx, DVC_ERR_HIP, "strain launch failed");
    if (e != hipSuccess) return hip_fail(ctx, e, "strain");
    return DVC_OK;
}

// DataCloud::sort_order_neighbors (DataCloud.cpp:69-144) with the same element type,
// comparator and std::sort, so ties resolve exactly as in the reference.
struct DualSort { int index; double value; };
static bool sortByValue(const DualSort &lhs, const DualSort &rhs) { return lhs.value < rhs.value; }
static inline double pdist(const double *a, const double *b) {
    return std::sqrt(std::pow(a[0] - b[0], 2) + std::pow(a[1] - b[1], 2) + std::pow(a[2] - b[2], 2));
}

int dvc_sort_order_neighbors(const double *pts, int32_t n, const double start[3], int32_t *order, int32_t *neigh,
                             int32_t nthreads) {
    if (n < 1 || !pts || !start || !order || !neigh) return DVC_ERR_ARG;
    int k = 75 < n ? 75 : n;
    std::vector<DualSort> idx(n);
    for (int i = 0; i < n; i++) { idx[i].index = i; idx[i].value = pdist(start, &pts[3 * i]); }
    std::sort(idx.begin(), idx.end(), sortByValue);
    for (int i = 0; i < n; i++) order[i] = idx[i].index;
    int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > n) nt = n;
    auto work = [&](int t) {
        std::vector<DualSort> cp(n);
        for (int i = t; i < n; i += nt) {
            for (int j = 0; j < n; j++) { cp[j].index = j; cp[j].value = pdist(&pts[3 * i], &pts[3 * j]); }
            std::sort(cp.begin(), cp.end(), sortByValue);
            for (int j = 0; j < k; j++) neigh[(size_t)i * k + j] = cp[j].index;
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    return k;
}

// k nearest neighbours through a uniform grid of cells: exact k-NN with ties broken by
// point index (the order std::stable_sort gives), O(n k) instead of the full O(n^2 log n)
// sort per point.  Identical to dvc_sort_order_neighbors whenever no two candidate
// distances of a point are equal; on lattices, equal distances appear in index order.
// The processing order is the same full sort of distances from `start`.
int dvc_knn_order_neighbors(const double *pts, int32_t n, const double start[3], int32_t k, int32_t *order,
                            int32_t *neigh, int32_t nthreads) {
    if (n < 1 || !pts || !start || !order || !neigh || k < 1) return DVC_ERR_ARG;
    if (k > n) k = n;
    for (int i = 0; i < 3 * n; i++)
        if (!std::isfinite(pts[i])) return DVC_ERR_ARG;
    {
        std::vector<DualSort> idx(n);
        for (int i = 0; i < n; i++) { idx[i].index = i; idx[i].value = pdist(start, &pts[3 * i]); }
        std::sort(idx.begin(), idx.end(), sortByValue);
        for (int i = 0; i < n; i++) order[i] = idx[i].index;
    }
    // grid: about 8 points per cell, at most 2^21 cells
    double lo[3], hi[3];
    for (int q = 0; q < 3; q++) { lo[q] = pts[q]; hi[q] = pts[q]; }
    for (int i = 1; i < n; i++)
        for (int q = 0; q < 3; q++) { lo[q] = std::min(lo[q], pts[3 * i + q]); hi[q] = std::max(hi[q], pts[3 * i + q]); }
    double ext[3], vol = 1.0;
    int flat_axes = 0;
    for (int q = 0; q < 3; q++) {
        ext[q] = hi[q] - lo[q];
        if (ext[q] > 0.0) vol *= ext[q]; else flat_axes += 1;
    }
    double h = std::pow(vol * 8.0 / (double)n, 1.0 / (double)(3 - std::min(flat_axes, 2)));
    if (!(h > 0.0) || !std::isfinite(h)) h = 1.0;
    long long dims[3];
    for (;;) {
        long long cells = 1;
        for (int q = 0; q < 3; q++) { dims[q] = (long long)(ext[q] / h) + 1; cells *= dims[q]; }
        if (cells <= (1ll << 21)) break;
        h *= 1.25;
    }
    auto cell_of = [&](const double *p, long long c[3]) {
        for (int q = 0; q < 3; q++) {
            long long v = (long long)((p[q] - lo[q]) / h);
            c[q] = v < 0 ? 0 : (v >= dims[q] ? dims[q] - 1 : v);
        }
    };
    const long long ncell = dims[0] * dims[1] * dims[2];
    std::vector<int32_t> cstart(ncell + 1, 0), cpts(n);
    std::vector<long long> pc(n);
    for (int i = 0; i < n; i++) {
        long long c[3];
        cell_of(&pts[3 * i], c);
        pc[i] = (c[2] * dims[1] + c[1]) * dims[0] + c[0];
        cstart[pc[i] + 1] += 1;
    }
    for (long long c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
    {
        std::vector<int32_t> fill(cstart.begin(), cstart.end() - 1);
        for (int i = 0; i < n; i++) cpts[fill[pc[i]]++] = i;
    }
    int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > n) nt = n;
    const long long maxdim = std::max(dims[0], std::max(dims[1], dims[2]));
    auto work = [&](int t) {
        std::vector<DualSort> cand;
        auto better = [](const DualSort &a, const DualSort &b) {
            return a.value < b.value || (a.value == b.value && a.index < b.index);
        };
        for (int i = t; i < n; i += nt) {
            const double *pi = &pts[3 * i];
            long long c[3];
            cell_of(pi, c);
            cand.clear();
            for (long long r = 0;; r++) {
                // shell of cells at Chebyshev distance r from the query's cell
                for (long long z = c[2] - r; z <= c[2] + r; z++) {
                    if (z < 0 || z >= dims[2]) continue;
                    for (long long y = c[1] - r; y <= c[1] + r; y++) {
                        if (y < 0 || y >= dims[1]) continue;
                        const bool yz_edge = (z == c[2] - r || z == c[2] + r || y == c[1] - r || y == c[1] + r);
                        for (long long x = c[0] - r; x <= c[0] + r; x += (yz_edge || r == 0) ? 1 : 2 * r) {
                            if (x < 0 || x >= dims[0]) continue;
                            const long long id = (z * dims[1] + y) * dims[0] + x;
                            for (int32_t s = cstart[id]; s < cstart[id + 1]; s++) {
                                const int j = cpts[s];
                                cand.push_back({j, pdist(pi, &pts[3 * j])});
                            }
                        }
                    }
                }
                if ((int)cand.size() >= k) {
                    std::nth_element(cand.begin(), cand.begin() + (k - 1), cand.end(), better);
                    // every point outside the searched cells is at least r*h away
                    if (cand[k - 1].value < (double)r * h || r > maxdim) break;
                } else if (r > maxdim) {
                    break;
                }
            }
            std::partial_sort(cand.begin(), cand.begin() + k, cand.end(), better);
            for (int j = 0; j < k; j++) neigh[(size_t)i * k + j] = cand[j].index;
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    return k;
}

}  // extern "C"
