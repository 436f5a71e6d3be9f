/*
 * dvc_oracle.h -- C interface of the FP64 CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in double precision and in
 * the reference's own operation order, the CCPi DVC per-point correlation path
 * (Core/Search.cpp, Core/Interpolate.cpp, Core/Matrix_4d.*, Core/FloatingCloud.cpp,
 * Core/SearchParams.cpp, Core/ObjectiveFunctions.cpp, Core/DataCloud.cpp,
 * Core/strain.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, for the results to compare against and as the timed CPU baseline.  The
 * product (libdvc_amd.so) never links or calls it.
 */
#ifndef DVC_ORACLE_H
#define DVC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run description: the RunControl fields the path reads (Core/Utility.h:90-134). */
typedef struct {
    int ndof;            /* num_srch_dof: 3, 6 or 12 */
    int obj;             /* obj_function: 0 sad, 1 ssd, 2 zssd, 3 nssd, 4 znssd */
    int interp;          /* interp_type: 0 nearest, 1 trilinear, 2 tricubic */
    int geom;            /* subvol_geom: 0 cube, 1 sphere */
    int subvol_size;
    int subvol_npts;
    int disp_max;
    int bytes_per;       /* vol_bit_depth / 8 */
    int vol_w, vol_h, vol_t;
    int jacobian;        /* 0: forward difference h=1e-10 (Search.cpp:501); 1: analytic */
    unsigned int sphere_seed;   /* seed of every sphere cloud (Search.cpp:104 uses 1) */
    int pad_;
    double basin_radius;
    double rigid_trans[3];
    double aspect[3];
} orc_run;

typedef struct {
    int status;          /* 0 point_good, -1 range_fail, -2 convg_fail, -3 not_search */
    int iters;           /* Gauss-Newton LM_prep_at calls */
    double obj;          /* obj_min */
    double par[12];      /* par_min (ndof used) */
} orc_result;

int    orc_version(void);
void   orc_bi_matrix(double *out64x64);
void   orc_constraint_matrix(double *out64x64);
int    orc_subv_num(int geom, int npts);

/* Stable sample cloud of one search point: xyz[n*3], bbox[6] = min xyz, max xyz,
 * refpt[3].  Returns the number of samples. */
int    orc_cloud(const orc_run *rc, const double pt[3], double *xyz, double *bbox, double *refpt);
/* Sphere offsets aspect*rel for a seed (FloatingCloud.cpp:51-78). */
int    orc_sphere_offsets(unsigned int seed, double rad, int num, const double aspect[3], double *off);

/* Moving cloud under params (FloatingCloud::affine_to, FloatingCloud.cpp:236-286). */
int    orc_affine(const orc_run *rc, const double pt[3], const double *par, int ndof,
                  double *xyz, double *bbox);

/* One Interpolate placement: build est_box_nom from rc (Search.cpp:66-80), center it on
 * center[3], load the box from vol (Interpolate.cpp:144-258) and evaluate points with
 * method 0 nearest / 1 trilinear / 2 tricubic.  bbox[6] is the points' bounding box
 * checked against act_box.  grad (may be NULL, tricubic only) receives d/dx,d/dy,d/dz.
 * Returns 0, or -1 on Intrp_Fail. */
int    orc_interp(const orc_run *rc, const void *vol, const double center[3],
                  const double *xyz, int n, const double *bbox, int method,
                  double *out, double *grad);

double orc_objective(int kind, const double *ref, const double *tar, int n, double *residual);
/* Search::obj_val_at after search_pt_setup for one point (box of the correlated volume on
 * pt + start): returns the cloud's size, or -1 on Range_Fail / Intrp_Fail. */
int orc_obj_val_at(const orc_run *rc, const void *refv, const void *tarv, const double pt[3], const double start[3],
                   const double *x, int nx, double *obj, double *res);

/* DataCloud::sort_order_neighbors (DataCloud.cpp:69-144). neigh is n x k, k=min(75,n). */
int    orc_sort_order_neighbors(const double *pts, int n, const double start[3],
                                int *order, int *neigh, int nthreads);

/* The dvc point loop (dvc.cpp:229-303): nproc points of order[] are searched serially.
 * results[npts] are initialised to not_search.  seconds receives the loop wall time
 * measured like dvc.cpp:227,307-310.  nthreads is the OpenMP team of the interpolator. */
int    orc_dvc_run(const orc_run *rc, const void *ref, const void *tar,
                   const double *pts, int npts, const int *order, const int *neigh, int k,
                   int nproc, int nthreads, orc_result *results, double *seconds);

/* The serial loop with per-point subvolume geometry (no reference counterpart): point n
 * uses rc with subvol_geom = geom[n], subvol_size = size[n], subvol_npts = npp[n]. */
int    orc_dvc_run_geo(const orc_run *rc, const void *ref, const void *tar,
                       const double *pts, int npts, const int *order, const int *neigh, int k,
                       int nproc, int nthreads, const int *geom, const int *size, const int *npp,
                       orc_result *results, double *seconds);

/* Given-start passes (no reference counterpart): mode 1 searches every point of the
 * first nproc of order[] from start[12n..]; mode 2 adds the judge pass (the reference's
 * start guess over the first pass's results at earlier positions).  Where both passes
 * converged and agree (translations within 1e-3 voxel, objectives within 1e-5
 * relative) the judge's result is the point's; the other points are searched again in
 * processing order from the reference's start guess over the final results (repair).
 * prop/judge/kept are required for mode 2 (kept: 0 judge, 2 repaired, -1 unsearched). */
int    orc_dvc_run_start(const orc_run *rc, const void *ref, const void *tar,
                         const double *pts, int npts, const int *order, const int *neigh, int k,
                         int nproc, int nthreads, int mode, const double *start, orc_result *prop,
                         orc_result *judge, orc_result *results, int *kept, double *seconds);

/* strain.cpp:402-591.  out_dis[n*3], out_npts[n], out_swr[n], out_E[n*9], out_L[n*9],
 * out_DG[n*9].  planar: bit0 yz_only, bit1 zx_only, bit2 xy_only forced; auto-detected
 * planes are added as strain.cpp:340-365 does. */
int    orc_strain(const double *pts, const double *dis, const int *status, const double *objmin,
                  const int *neigh, int k, int n, int ndp, double objt, int refill, int planar,
                  double *out_dis, int *out_npts, double *out_swr, double *out_E, double *out_L,
                  double *out_DG);

/* Dense solvers restated from Eigen's ColPivHouseholderQR / FullPivHouseholderQR. */
int    orc_colpiv_solve(const double *A, const double *b, int n, double *x);
int    orc_fullpiv_solve(const double *A, const double *b, int n, double *x);
void   orc_sym3_eigen(const double *S, double *lam_desc);

#ifdef __cplusplus
}
#endif
#endif
