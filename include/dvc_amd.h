/*
 * dvc_amd.h -- C ABI of the MI355X (gfx950) implementation of the CCPi DVC
 * per-point correlation path.
 *
 * Plain pointers and sizes only.  All entry points return 0 on success and a
 * negative DVC_ERR_* code on failure; dvc_last_error() describes the last failure.
 * No C++ exception crosses this boundary.  One context per host thread.
 *
 * Each declaration names the reference interface it replaces (file:line in the
 * TomographicImaging/DigitalVolumeCorrelation sources, Core/ directory).
 */
#ifndef DVC_AMD_H
#define DVC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Point status codes, identical to Core/Utility.h:38-41. */
#define DVC_POINT_GOOD   0
#define DVC_RANGE_FAIL  -1
#define DVC_CONVG_FAIL  -2
#define DVC_NOT_SEARCH  -3

/* Return codes. */
#define DVC_OK            0
#define DVC_ERR_ARG      -1   /* invalid argument or size */
#define DVC_ERR_HIP      -2   /* HIP runtime failure */
#define DVC_ERR_STATE    -3   /* call out of order (e.g. no volume uploaded) */
#define DVC_ERR_NODEV    -5   /* no usable gfx950 device */
#define DVC_ERR_RANGE    -6   /* dvc_interpolate: points outside the active box (Intrp_Fail) */

/* Objective functions, in the order of Core/Utility.h:65-68. */
#define DVC_OBJ_SAD   0
#define DVC_OBJ_SSD   1
#define DVC_OBJ_ZSSD  2
#define DVC_OBJ_NSSD  3
#define DVC_OBJ_ZNSSD 4

/* Interpolation methods, in the order of Core/Utility.h:70-73. */
#define DVC_INTERP_NEAREST   0
#define DVC_INTERP_TRILINEAR 1
#define DVC_INTERP_TRICUBIC  2

/* Sub-volume geometry, Core/Utility.h:55-58. */
#define DVC_GEOM_CUBE   0
#define DVC_GEOM_SPHERE 1

/* Volume slots: the reference (RunControl::ref_fname) and the correlated
 * (RunControl::cor_fname) image volumes. */
#define DVC_VOL_REF 0
#define DVC_VOL_COR 1

typedef struct dvc_ctx dvc_ctx;

/* The RunControl fields read by Search/Interpolate (Core/Utility.h:90-134). */
typedef struct {
    int32_t ndof;            /* num_srch_dof: 3, 6 or 12 */
    int32_t obj;             /* DVC_OBJ_* */
    int32_t interp;          /* DVC_INTERP_* */
    int32_t geom;            /* DVC_GEOM_* */
    int32_t subvol_size;     /* side or diameter in voxels */
    int32_t subvol_npts;     /* requested samples (cube: rounded up to a cube, Search.cpp:30) */
    int32_t disp_max;        /* voxels */
    uint32_t sphere_seed;    /* mt19937 seed of every sphere cloud (Search.cpp:102-105 uses 1) */
    double basin_radius;     /* 0 disables the translation grid (Search.cpp:941) */
    double rigid_trans[3];   /* start guess without converged neighbours (Search.cpp:351-356) */
    double aspect[3];        /* subvol_aspect */
} dvc_run_params;

/* One search result: ResultRecord (Core/DataCloud.h:45-56) plus the iteration count. */
typedef struct {
    int32_t status;          /* DVC_POINT_GOOD / DVC_RANGE_FAIL / DVC_NOT_SEARCH */
    int32_t iters;           /* Gauss-Newton evaluations (LM_prep_at calls) */
    double  obj;             /* obj_min, 0 unless good */
    double  par[12];         /* par_min: u v w phi the psi exx eyy ezz exy eyz exz */
} dvc_point_result;

/* ---------------------------------------------------------------- context */

/* Create a context on HIP device `device` (replaces nothing: the reference is a
 * single-process CPU program; this owns the device buffers, stream and events). */
int  dvc_create(dvc_ctx **ctx, int device);
void dvc_destroy(dvc_ctx *ctx);
const char *dvc_last_error(const dvc_ctx *ctx);
/* Run subsequent work on an existing hipStream_t (NULL: the context's own stream). */
int  dvc_set_stream(dvc_ctx *ctx, void *hip_stream);
/* Milliseconds of the last search execute / strain fit on the device, timed with HIP
 * events on the stream the kernels ran on.  Waits for that work to finish. */
int  dvc_last_kernel_ms(dvc_ctx *ctx, float *ms);

/* Device memory on the context's device, for callers that keep their own inputs
 * resident (no reference counterpart).  dvc_device_copy: kind 0 host to device,
 * 1 device to host, 2 device to device; it waits for the copy and for all earlier work
 * on the context stream.  dvc_device_free waits for that stream first.
 * dvc_synchronize waits for all work queued on the context stream. */
int dvc_device_alloc(dvc_ctx *ctx, void **ptr, int64_t bytes);
int dvc_device_free(dvc_ctx *ctx, void *ptr);
int dvc_device_copy(dvc_ctx *ctx, void *dst, const void *src, int64_t bytes, int kind);
int dvc_synchronize(dvc_ctx *ctx);

/* ---------------------------------------------------------------- volumes */

/* Make an image volume resident in HBM in its native uint8 / uint16 form.
 * Replaces the per-point disk reads of Interpolate::kernels (Interpolate.cpp:144-193):
 * `host` holds the raw file contents after `header_bytes`; big_endian swaps uint16
 * as Interpolate.cpp:233-242 does.  Layout: x fastest, then y, then z. */
int dvc_volume_upload(dvc_ctx *ctx, int slot, const void *host, int64_t header_bytes,
                      int bytes_per, int big_endian, int64_t wide, int64_t high, int64_t tall);
/* Borrow a volume that is already in device memory (x fastest, native endianness,
 * readable for 16 bytes past its end).  The caller keeps ownership. */
int dvc_volume_attach(dvc_ctx *ctx, int slot, const void *device_ptr, int bytes_per,
                      int64_t wide, int64_t high, int64_t tall);

/* Copy the volume in `slot` back to host memory (wide*high*tall*bytes_per bytes, native
 * endianness); no reference counterpart. */
int dvc_volume_download(dvc_ctx *ctx, int slot, void *host);

/* Synthetic volumes made on the device, for workloads too large to generate on the host
 * (no reference counterpart; the reference's test data are files).
 * dvc_synth_noise: fill `slot` with a wide x high x tall volume of bytes_per-byte voxels:
 * uniform noise from `seed` (splitmix64 of the voxel index), separable Gaussian
 * smoothing with `sigma` (truncated at 4 sigma, mirror-reflected at the faces), rescaled
 * so that its minimum and maximum map to 0 and the type's maximum, rounded.
 * dvc_synth_warp_field: fill the correlated slot (1) with the reference (slot 0) under
 * the displacement field u(X) = sum_b a_b exp(-|X - c_b|^2 / (2 w_b^2))
 * + sum_s a_s sin(k_s . X + phi_s) + strain (X - center): target(x) = ref(X) where
 * X + u(X) = x (`iterations` fixed-point steps), separable Catmull-Rom, rounded and
 * clamped.  bumps[7*b..] = c_b (xyz), w_b, a_b (xyz); waves[7*s..] = k_s (xyz), phi_s,
 * a_s (xyz). */
int dvc_synth_noise(dvc_ctx *ctx, int slot, int bytes_per, int64_t wide, int64_t high, int64_t tall, double sigma,
                    uint64_t seed);
int dvc_synth_warp_field(dvc_ctx *ctx, const double *bumps, int32_t nbumps, const double *waves, int32_t nwaves,
                         double strain, const double center[3], int32_t iterations);

/* ---------------------------------------------------------- interpolation */

/* Interpolate::center_on + Interpolate::kernels + tri_cub_Lek / tri_lin / nearest
 * (Interpolate.cpp:569-574, 144-258, 472-567, 429-470, 404-427): place the
 * interpolation box of `rc` (Search.cpp:64-80) on `center`, then evaluate the volume
 * in `slot` at n points xyz[3n] whose bounding box is bbox[6] (min xyz, max xyz).
 * out[n] receives values; grad[3n] (may be NULL, tricubic only) the x,y,z derivatives.
 * Returns DVC_ERR_RANGE where tri_cub_Lek throws Intrp_Fail. */
int dvc_interpolate(dvc_ctx *ctx, int slot, const dvc_run_params *rc, const double center[3],
                    const double *xyz, int32_t n, const double bbox[6], int method,
                    double *out, double *grad);

/* The same on the boxes of Interpolate(const BoundBox *region) (Interpolate.cpp:34-58):
 * est box = region [min, max], active box = region less one voxel on every side,
 * coefficient extent = trunc(max) - trunc(min) voxels (at least 4 per axis). */
int dvc_interpolate_box(dvc_ctx *ctx, int slot, const double region_min[3], const double region_max[3],
                        const double center[3], const double *xyz, int32_t n, const double bbox[6], int method,
                        double *out, double *grad);

/* obj_SAD / obj_SSD / obj_ZSSD / obj_NSSD / obj_ZNSSD with residual
 * (ObjectiveFunctions.h:31-45, ObjectiveFunctions.cpp:24-263), on the device in FP64.
 * residual[n] may be NULL. */
int dvc_objective(dvc_ctx *ctx, int kind, const double *ref, const double *tar, int32_t n,
                  double *obj, double *residual);

/* Search::obj_val_at(x) and obj_val_at(x, residual) (Search.h:88-89,
 * Search.cpp:448-488) for one point, in the state Search::search_pt_setup
 * (Search.cpp:248-275) leaves: the point's FloatingCloud (rc's geometry, aspect and
 * sphere seed, Search.cpp:99-118) is sampled in the reference volume with the
 * interpolation box centred on point[3] (ref_subvol); the box is then centred on
 * point + start[0..2], the start guess's translation (par_min after starting_param;
 * NULL: rc->rigid_trans, the guess of a point without good neighbours); the cloud
 * under FloatingCloud::affine_to(x, nx) (1 <= nx <= 12) is sampled in the correlated
 * volume (tar_subvol), and *obj = the objective of rc->obj over the two.  residual
 * (may be NULL, else residual_len >= the cloud's size) receives obj_fcn_res's
 * residual vector; *nsamp (may be NULL) the cloud's size.  DVC_ERR_RANGE where the
 * reference throws Range_Fail (or, sampling the reference cloud, Intrp_Fail). */
int dvc_obj_val_at(dvc_ctx *ctx, const dvc_run_params *rc, const double point[3], const double start[3],
                   const double *x, int32_t nx, double *obj, double *residual, int32_t residual_len,
                   int32_t *nsamp);

/* dvc_obj_val_at without any allocation: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops): the same arguments, results and errors, with the clouds, sample values,
 * residual and objective in a fixed device array of 3.5 MiB on each GPU (set up by
 * dvc_create, held until the process ends) and the host-side clouds in a fixed host
 * array.  Clouds of at most 50,653 samples (every geometry a run allows); larger ones
 * are DVC_ERR_ARG.  All calls on one GPU share its array, which keeps the last call's
 * contents; the library serialises the calls of one process, and calls of several
 * processes each use their own copy. */
int dvc_obj_val_at_noalloc(dvc_ctx *ctx, const dvc_run_params *rc, const double point[3], const double start[3],
                           const double *x, int32_t nx, double *obj, double *residual, int32_t residual_len,
                           int32_t *nsamp);

/* ----------------------------------------------------------------- search */

/* The dvc point loop (dvc.cpp:229-303) around Search::process_point
 * (Search.cpp:95-192): points[3*npts]; order[npts] = DataCloud::order; neigh[npts*k]
 * = DataCloud::neigh (k = min(75, npts), self first); the first nproc entries of order
 * are searched (nproc <= 1 or >= npts: all, dvc.cpp:229-232).  Every point's start
 * guess averages the converged translations of its neighbours searched before it
 * (Search.cpp:277-377); the points run level by level of that dependency, so results
 * match the serial loop.  results[npts]. */
int dvc_search_points(dvc_ctx *ctx, const dvc_run_params *rc, const double *points,
                      int32_t npts, const int32_t *order, int32_t nproc,
                      const int32_t *neigh, int32_t k, dvc_point_result *results);

/* The same loop split for callers that keep inputs resident: prepare uploads the
 * point cloud and its order/neighbour tables and derives the chain's launches (the
 * dependency levels, or launches of the round capacity, dvc_set_round_capacity);
 * execute enqueues them on the context stream without synchronising (in the judged mode,
 * DVC_START_JUDGED, it waits once, for the selection, and then enqueues only the repair
 * launches that hold a point to repair); fetch waits and copies results back. */
int dvc_search_prepare(dvc_ctx *ctx, const dvc_run_params *rc, const double *points,
                       int32_t npts, const int32_t *order, int32_t nproc,
                       const int32_t *neigh, int32_t k);
int dvc_search_execute(dvc_ctx *ctx);
int dvc_search_fetch(dvc_ctx *ctx, dvc_point_result *results);

/* Per-point sub-volume geometry for the next dvc_search_prepare (no reference
 * counterpart: the reference has one RunControl, so one shape, size and sample count for
 * all points).  Point i is searched as Search::process_point would with a RunControl
 * whose subvol_geom, subvol_size and subvol_npts are geom[i], size[i], nsamp[i] and whose
 * other fields are the prepare's run parameters; start guesses still come from the
 * neighbours' results.  n = npts of the prepare; n = 0 (or NULL arrays) returns to one
 * geometry for all.  Needs the HBM table of reference samples (their total under 48 GiB). */
int dvc_set_point_geometry(dvc_ctx *ctx, const int32_t *geom, const int32_t *size, const int32_t *nsamp, int32_t n);

/* Start modes of a prepared search.  The reference has one: the serial loop's start
 * guess (Search.cpp:277-377), DVC_START_CHAIN.  The other two have no counterpart in
 * the reference; they serve starting warps proposed for every point at once (e.g. by a
 * learned proposal) and drop the dependency between points, so each pass is one launch:
 *   DVC_START_GIVEN  : every point searched from its given start warp start[12n..12n+ndof).
 *   DVC_START_JUDGED : the given-start pass, then a judge pass in which every point starts
 *                      from the reference's own guess (mean translation of good neighbours
 *                      earlier in order) computed on the first pass's results.  Where both
 *                      passes converged and agree (translations within 1e-3 voxel,
 *                      objectives within 1e-5 relative) the judge's result is the point's.
 *                      Every other point is searched again (repair) in the serial chain's
 *                      launches, from the reference's start guess over the final results;
 *                      of those launches only the ones that hold such a point run (as
 *                      slice launches where the chain would slice them).
 *                      Every result is thus the plain search from the reference's start
 *                      guess; the given starts only make the neighbours' results available
 *                      at once.
 * dvc_set_start_mode applies to the next dvc_search_prepare (default DVC_START_CHAIN;
 * dvc_search_points always runs the chain). */
#define DVC_START_CHAIN  0
#define DVC_START_GIVEN  1
#define DVC_START_JUDGED 2
int dvc_set_start_mode(dvc_ctx *ctx, int mode);
/* Start warps of a prepared DVC_START_GIVEN / DVC_START_JUDGED search: start[npts*12]
 * (u v w phi the psi exx eyy ezz exy eyz exz per point; entries past ndof ignored), in
 * host memory, or in device memory when on_device is 1 (copied on the context stream, so
 * a producer on another stream must be finished).  Required before dvc_search_execute;
 * later calls replace them for the next execute.  A non-finite start fails its point
 * with DVC_RANGE_FAIL in the given-start pass. */
int dvc_search_set_starts(dvc_ctx *ctx, const double *start, int on_device);
/* After a DVC_START_JUDGED execute: the given-start pass's results, the judge pass's
 * results (each [npts], either may be NULL) and kept[npts] (0: the judge's result,
 * 2: repaired, -1: not searched; may be NULL).  dvc_search_fetch returns the
 * selection. */
int dvc_search_fetch_passes(dvc_ctx *ctx, dvc_point_result *proposal, dvc_point_result *judge, int32_t *kept);

/* Launch tuning: threads per point (64, 128, 256, 512 or 1024; 0 = automatic), applied
 * by the next dvc_search_prepare. */
int dvc_set_block_threads(dvc_ctx *ctx, int threads);
/* Autotuner of a prepared search (no reference counterpart): times reps executes at
 * each threads-per-point count whose LDS fits (64, 128, 256, 512, 1024; one point per
 * workgroup) and keeps the fastest for later executes (until the next prepare); then
 * times 128, 256 and 512 threads per point for the launch that computes every point's
 * reference samples (256 by default) and keeps the fastest of those too; then, for the
 * serial chain, slice launches off and nine shapes of them (dvc_set_slice_launches:
 * widths of 2-8 workgroups per CU, 3-14 slice launches, 512, 1024 or 2048 samples per
 * workgroup), keeping the fastest.  The given starts must be set first in the
 * given-start and judged modes.  Results change only in the summation order of the
 * per-point sums.  best_threads / best_ms (device milliseconds per execute) may be
 * NULL. */
int dvc_search_autotune(dvc_ctx *ctx, int reps, int32_t *best_threads, float *best_ms);
/* 1 (default): the level launches of a prepared search are captured into a HIP graph
 * once and replayed by dvc_search_execute; 0: launched one by one. */
int dvc_set_launch_mode(dvc_ctx *ctx, int graph);
/* Launches of the serial chain (no reference counterpart), applied by the next
 * dvc_search_prepare.  The chain runs as one launch per dependency level: a level's
 * points read only results of earlier levels.  When a level is wider than the chip runs
 * at once (`points` per launch; 0 = automatic: CUs x workgroups of the chosen size per
 * CU), the chain instead runs in launches of at most that many points: in processing
 * order, each point joins the first launch after those holding its earlier neighbours
 * that has room.  Each point still reads only finished results, so the results are
 * identical; partly empty rounds at the ends of wide levels fill with later points.
 * -1: dependency levels only. */
int dvc_set_round_capacity(dvc_ctx *ctx, int32_t points);
/* Slice launches of the serial chain (no reference counterpart).  A chain launch of few
 * points leaves most of the chip idle with one workgroup per point.  Such a launch can
 * instead run as slice launches: one launch per Gauss-Newton evaluation, each point's
 * samples split over workgroups of samples_per_workgroup samples (0: 512; 256 or a
 * multiple of 512 up to 8192), every workgroup of a point adding the previous launch's
 * slice sums up in slice order and taking the step of Search.cpp:194-246.  After
 * `evaluations` such launches (0: 4) the one-workgroup search resumes the points still
 * searching.  A launch is sliced when its points times their slices are at most
 * max_workgroups (0: 4 x the chip's CUs; -1: never).  A point's slices run on one XCD:
 * where a launch's points would need more of them than an XCD has CUs, each point takes
 * fewer, larger slices (at least 2).  With per-point geometry (dvc_set_point_geometry)
 * each point uses as many slices as its own samples need, up to the launch's cap, and a
 * launch's slice count is that of its point with the most.  Applies to the serial
 * chain without a translation grid and with the HBM table of reference samples; results
 * differ from the one-workgroup search only in the summation order of the per-point sums.
 * Takes effect at the next execute. */
int dvc_set_slice_launches(dvc_ctx *ctx, int32_t max_workgroups, int32_t evaluations, int32_t samples_per_workgroup);
/* Slice launches of the prepared search: chain launches run that way, workgroups per
 * point (the most any launch uses) and slice launches per chain launch (all 0 when
 * none).  Any pointer may be NULL. */
int dvc_search_slice_info(dvc_ctx *ctx, int32_t *sliced_launches, int32_t *slices_per_point, int32_t *evaluations);
/* Shape of the prepared search: chain launches (dependency levels, or launches of the
 * round capacity), threads per point, points searched.  Any pointer may be NULL. */
int dvc_search_info(dvc_ctx *ctx, int32_t *levels, int32_t *threads_per_point, int32_t *points_searched);

/* -------------------------------------------------------- learned proposal */

/* A learned start-warp proposal.  The reference has no counterpart (its only start
 * guess is the neighbours' mean translation, Search.cpp:277-377); this network proposes
 * the 12 warp parameters of a point from the 32^3 reference/target crops around it, for
 * the given-start and judged search modes above.  Network: four 3x3x3 stride-2
 * convolutions with ReLU (2->16->32->64->128 channels, BF16 weights and activations, FP32
 * accumulation on the matrix cores), the mean over the last 2^3 voxels and a 128->12
 * linear head.  Parameters are one FP32 vector of dvc_net_num_params() values: for each
 * convolution its weights [cout][27*cin] (tap-major, tap = (kz*3+ky)*3+kx, then input
 * channel) and biases [cout], then the head's weights [12][128] and biases [12].  The
 * head's output j is scaled by 8 (u v w, voxels), 0.05 (angles, radians) or 0.02
 * (strains).  All device buffers are on the context's device and stream. */
typedef struct dvc_net dvc_net;
/* max_batch crops per forward (1..4096); training != 0 also allocates gradient and
 * optimiser buffers (then max_batch must be a multiple of 4). */
int  dvc_net_create(dvc_ctx *ctx, int max_batch, int training, dvc_net **net);
/* Destroy every network of a context before the context itself. */
void dvc_net_destroy(dvc_net *net);
const char *dvc_net_last_error(const dvc_net *net);
int64_t dvc_net_num_params(void);
/* set != 0: upload count FP32 parameters from host (and reset the optimiser);
 * set == 0: download them. */
int  dvc_net_params(dvc_net *net, float *host, int64_t count, int set);
/* Crops of the context's volume pair around points_dev[3*idx_dev[i]] (idx_dev NULL:
 * point i), i < count, into slot slot_dev[i] (NULL: slot i) of out_dev (BF16
 * [slots][32][32][32][2], channel 0 reference, 1 target, each normalised to zero mean
 * and unit variance; out_dev NULL: the network's own input buffer, count <= max_batch).
 * flip_dev[i] (NULL: 0) mirrors crop i along x (bit 0), y (bit 1), z (bit 2), for
 * training-set augmentation.  All pointers are device pointers; the caller sizes
 * out_dev for every slot named. */
int  dvc_net_crop(dvc_net *net, const double *points_dev, const int32_t *idx_dev, int32_t count, void *out_dev,
                  const int32_t *slot_dev, const int32_t *flip_dev);
/* Forward pass of B crops (in_dev: device BF16 crops, NULL: the network's input buffer)
 * into par_dev[B*12] (FP64 warp parameters, may be NULL) and y_dev[B*12] (FP32 unscaled
 * head output, may be NULL). */
int  dvc_net_forward(dvc_net *net, const void *in_dev, int32_t B, double *par_dev, float *y_dev);
/* Loss (mean over the batch of the squared error of the scaled outputs against
 * label_dev[B*12] / scale) and its gradient with respect to every parameter, into
 * grad_host[dvc_net_num_params()] and loss_host (either may be NULL). */
int  dvc_net_gradients(dvc_net *net, const void *in_dev, const float *label_dev, int32_t B, float *grad_host,
                       float *loss_host);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on that gradient. */
int  dvc_net_train_step(dvc_net *net, const void *in_dev, const float *label_dev, int32_t B, float lr,
                        float *loss_host);
/* Training pairs for the network: fill the context's correlated volume (slot 1, made
 * the shape and type of slot 0) with the reference volume (slot 0) resampled tile by
 * tile, each tile x tile x tile
 * block t (numbered x fastest) under its own affine map out(x) = ref(M_t x + o_t),
 * maps[12*t..] = M row-major then o, nmaps = the number of tiles; separable Catmull-Rom
 * with clamped indices, rounded and clamped to the voxel range. */
int  dvc_synth_warp_tiles(dvc_ctx *ctx, int tile, const double *maps, int64_t nmaps);
/* Proposals for every point of a prepared DVC_START_GIVEN / DVC_START_JUDGED search,
 * written straight into its start table (replaces dvc_search_set_starts). */
int  dvc_search_propose(dvc_ctx *ctx, dvc_net *net);

/* The strain fit of a prepared search's results, on the device after every
 * dvc_search_execute (the reference's strain program on the dvc program's output,
 * strain.cpp:402-591 over dvc.cpp's .disp and .sort.csv): the search's neighbour table,
 * window ndp (1..75; 0: no fit), objective threshold objt, refill and planar as
 * dvc_strain_fit.  Takes effect at the next execute; dvc_search_fetch_strain copies the
 * last fit's outputs back (layouts of dvc_strain_fit; any pointer may be NULL). */
int dvc_search_set_strain(dvc_ctx *ctx, int32_t ndp, double objt, int32_t refill, int32_t planar);
int dvc_search_fetch_strain(dvc_ctx *ctx, double *out_disp, int32_t *out_npts, double *out_swr, double *out_E,
                            double *out_L, double *out_DG);

/* ----------------------------------------------------------------- strain */

/* The strain-window quadratic fit of strain.cpp:402-591 for n points: pts[3n],
 * disp[3n], status[n], objmin[n], neigh[n*k]; ndp window size (10..75,
 * strain.cpp:204-207), objt threshold, refill (-r), planar bit mask (1 yz, 2 zx,
 * 4 xy; auto-detected planes are added as strain.cpp:340-365).  Outputs:
 * fit displacement [3n], points in window [n], window radius [n], engineering and
 * Lagrangian strain [9n] (exx eyy ezz exy eyz exz p1 p2 p3), gradient [9n]. */
int dvc_strain_fit(dvc_ctx *ctx, const double *pts, const double *disp, const int32_t *status,
                   const double *objmin, const int32_t *neigh, int32_t k, int32_t n,
                   int32_t ndp, double objt, int32_t refill, int32_t planar,
                   double *out_disp, int32_t *out_npts, double *out_swr,
                   double *out_E, double *out_L, double *out_DG);

/* --------------------------------------------------------- point ordering */

/* DataCloud::sort_order_neighbors (DataCloud.cpp:69-144) on the host: the processing
 * order by distance from start[3] and the k = min(75, n) nearest neighbours of every
 * point, with std::sort tie order.  order[n], neigh[n*k]; returns k. */
int dvc_sort_order_neighbors(const double *pts, int32_t n, const double start[3],
                             int32_t *order, int32_t *neigh, int32_t nthreads);

/* The same tables for large clouds through a uniform grid: the order as above, and the
 * k (<= n) nearest neighbours of every point with equal distances in point-index order
 * (std::stable_sort order).  Identical to dvc_sort_order_neighbors when a point's
 * candidate distances are distinct; O(n k) rather than the reference's O(n^2 log n)
 * (DataCloud.cpp:95-140).  Returns k. */
int dvc_knn_order_neighbors(const double *pts, int32_t n, const double start[3], int32_t k,
                            int32_t *order, int32_t *neigh, int32_t nthreads);

#ifdef __cplusplus
}
#endif
#endif
