// dvc_internal.h -- structures shared by the host code (dvc_api.cpp) and the device
// code (dvc_kernels.hip).  Not part of the public ABI.
#pragma once

#include <stdint.h>
#include "../../include/dvc_amd.h"

namespace dvc {

constexpr int kMaxNeigh = 128;       // neighbour slots read per point (reference: 75)

// Sums of one point's reference sub-volume (reference_kernel -> search_kernel).
struct RefStats {
    double rmean, rbmean, R2, Src, Rn2;
};

// Sub-volume geometry of one point when points differ in shape, size and sample count
// (one RunControl per point; the reference has one for all).  The search kernels apply
// it over their arguments for that point.
struct PtGeo {
    int geom, nsamp, g, gather32;
    int box[3], pad_;
    double subv_rad;
    double est_mn[3], est_mx[3], est_cn[3];
    double act_mn[3], act_mx[3], act_cn[3];
    double off_min[3], off_max[3];
    long long sph_off;               // first sample of its sphere template in SearchArgs::sph (samples)
    long long rval_off;              // first entry of its reference samples in SearchArgs::rvals
};

// Arguments of the search kernel.  Geometry mirrors Search::Search
// (Search.cpp:24-82) and Interpolate::Interpolate (Interpolate.cpp:34-56).
struct SearchArgs {
    const uint8_t *vol[2];           // reference and correlated volume bytes
    int W, H, T, bpv;
    int ndof, obj, interp, geom;
    int nsamp, g;                    // samples per point; cube grid side
    int npts, nproc, k, nsamp_pad;
    int box[3];                      // interpolation box size in voxels (kern_4d extent)
    int gather32;                    // box rows addressable by 32-bit offsets (see Gather)
    double est_mn[3], est_mx[3], est_cn[3];   // est_box_nom after the two grow_by calls
    double act_mn[3], act_mx[3], act_cn[3];   // act_box = est_box shrunk by one voxel
    double subv_rad, aspect[3], rigid[3], basin, disp_max;
    double off_min[3], off_max[3];   // sphere offsets' bounding box
    const double *sph;               // sphere offsets aspect*rel [nsamp*3]
    const double *pts;               // [npts*3]
    const int *order;                // [npts]
    const int *pos;                  // position of each point in order[] (npts if unsearched)
    const int *neigh;                // [npts*k]
    dvc_point_result *res;           // [npts] results written by the launch
    const dvc_point_result *nbres;   // [npts] results the start guess averages (res in the serial chain)
    const double *start;             // [npts*12] given start warps (a proposal pass), or null
    float *rvals;                    // [nproc * nsamp_pad] reference samples by position, or null
    RefStats *rstat;                 // [nproc]
    const PtGeo *pgeo;               // [npts] per-point geometry, or null (one geometry for all)
    const int *ref_order;            // [nproc] processing position of each reference_kernel workgroup, or null
    const int *repair;               // [npts] search only points whose entry is nonzero (repair launches), or null
};

// Slice launches of the serial chain: the points of one chain launch run one launch per
// Gauss-Newton evaluation, each point's samples split over nslice workgroups (with
// per-point geometry, ceil(samples / per) of them; nslice is the launch's largest).
// Block b works on point slot b % cnt8 (level_pts[slot]) and slice b / cnt8, which takes
// the point's chunks of 64 samples slice, slice + m, slice + 2 m, ... (m: the point's
// slice count); cnt8 is cnt rounded up to a multiple of the XCD count, so the slices of
// one point share an XCD (and its L2).  No workgroup communicates with another
// inside a launch: launch t writes every slice's sums and, from slice 0, the point's state
// into buffers t & 1; launch t + 1 reads them (every workgroup of a point adds the
// slices up, in slice order, and takes the same step) and writes buffers (t + 1) & 1.
constexpr int kStateStride = 2048;   // bytes of one point's carried state
constexpr int kPartStride = 272;     // doubles of one slice's partial sums: 3 sums, 6 bbox, 7 pad, 256 V
struct SliceArgs {
    const int *level_pts;            // points of the chain launch
    int cnt, cnt8;                   // points; cnt rounded up to a multiple of 8
    int nslice, per;                 // workgroups per point (the launch's largest); samples per workgroup
    int maxsl;                       // per-point geometry: at most this many workgroups per point (0: no cap)
    int t;                           // launch index within the chain launch (0: set-up and start guess)
    unsigned char *state[2];         // [cnt8][kStateStride] carried point state, by launch parity
    double *part[2];                 // [cnt8][nslice][kPartStride] slice sums, by launch parity
};

struct StrainArgs {
    const double *pts, *disp, *objmin;
    const int *status, *neigh;
    int k, n, ndp, refill;
    int yz, zx, xy;
    double objt;
    double *out_disp, *out_swr, *out_E, *out_L, *out_DG;
    int *out_npts;
};

// launchers implemented in dvc_kernels.hip
int launch_search(const SearchArgs &a, const int *level_pts, int block_threads, int grid_blocks, size_t lds_bytes,
                  void *stream);
// one slice launch with `threads` (kSliceThreads[]) per workgroup
constexpr int kSliceThreads[2] = {256, 512};
int launch_slice(const SearchArgs &a, const SliceArgs &sa, int threads, void *stream);
// the points of rs.level_pts after rs.t slice launches, one 512-thread workgroup each
int launch_resume(const SearchArgs &a, const SliceArgs &rs, void *stream);
// the geometry a point's search uses (pgeo[n] applied, or a itself)
int search_set_lds_limit(size_t lds_total);
int search_static_lds(int bpv, int block_threads, size_t *bytes);
int search_max_blocks(int bpv, int block_threads, size_t lds_bytes, int *blocks_per_cu);
size_t search_lds_bytes(int nsamp_pad, int block_threads, bool samples_in_lds);
int launch_interp(const uint8_t *vol, int bpv, int W, int H, int T, const int b0[3], const int box[3],
                  const double *xyz, int n, int method, double *out, double *grad, void *stream);
int launch_objective(int kind, const double *ref, const double *tar, int n, double *obj_out,
                     double *residual, void *stream);
int launch_strain(const StrainArgs &a, void *stream);
int launch_fields(const dvc_point_result *res, const int *pos, int npts, int nproc, double *disp, int *status,
                  double *objmin, void *stream);
int launch_reference(const SearchArgs &a, int block_threads, void *stream);
int launch_zero_int(int *p, long long n, void *stream);
int launch_judge(const int *order, int n, const dvc_point_result *prop, const dvc_point_result *judge,
                 dvc_point_result *out, int *kept, int *repair, void *stream);
int launch_swap16(uint16_t *p, long long n, void *stream);

// context internals for the proposal network (dvc_net.hip), implemented in dvc_api.cpp
int ctx_device(dvc_ctx *ctx);
void *ctx_stream(dvc_ctx *ctx);
// both volume slots filled with the same shape and type: 0, else -1
int ctx_volumes(dvc_ctx *ctx, const uint8_t **v0, const uint8_t **v1, int *bpv, int *W, int *H, int *T);
// a prepared given-start or judged search: its device points, order, sizes and start table
int ctx_prepared(dvc_ctx *ctx, const double **pts, const int **order, int *npts, int *nproc, double **start);
void ctx_mark_starts(dvc_ctx *ctx);
// one volume slot: 0, else -1 (empty)
int ctx_volume_info(dvc_ctx *ctx, int slot, const uint8_t **p, int *bpv, int *W, int *H, int *T);
// (re)allocate a volume slot on the device (64 zero bytes of padding); null on failure
uint8_t *ctx_volume_alloc(dvc_ctx *ctx, int slot, int bpv, int64_t w, int64_t h, int64_t t);

}  // namespace dvc
