// dvc_search.cpp -- see dvc_search.h.
#include "dvc_search.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>

namespace dvc {

namespace {

struct Shared {
    dvc_ctx *ctx = nullptr;
    std::string slot_file[2];        // file resident in each volume slot
    int last_slot = 1;
    ~Shared() {
        if (ctx) dvc_destroy(ctx);
    }
};

Shared &shared() {
    static Shared s;
    return s;
}

void check(int rc, dvc_ctx *ctx, const char *what) {
    if (rc != DVC_OK) throw Device_Fail{rc, std::string(what) + ": " + (ctx ? dvc_last_error(ctx) : "")};
}

double objective(int kind, const std::vector<double> &r, const std::vector<double> &t, std::vector<double> *res) {
    if (r.size() != t.size() || r.empty()) throw Device_Fail{DVC_ERR_ARG, "objective: sizes differ or are zero"};
    dvc_ctx *ctx = shared_context();
    double v = 0.0;
    if (res) res->assign(r.size(), 0.0);
    check(dvc_objective(ctx, kind, r.data(), t.data(), (int32_t)r.size(), &v, res ? res->data() : nullptr), ctx,
          "dvc_objective");
    return v;
}

}  // namespace

dvc_ctx *shared_context() {
    Shared &s = shared();
    if (!s.ctx) {
        int dev = 0;
        if (const char *e = std::getenv("DVC_DEVICE")) dev = std::atoi(e);
        dvc_ctx *c = nullptr;
        int rc = dvc_create(&c, dev);
        if (rc != DVC_OK) throw Device_Fail{rc, "no usable gfx950 device"};
        s.ctx = c;
    }
    return s.ctx;
}

double obj_SAD(const std::vector<double> &r, const std::vector<double> &t) { return objective(DVC_OBJ_SAD, r, t, nullptr); }
double obj_SSD(const std::vector<double> &r, const std::vector<double> &t) { return objective(DVC_OBJ_SSD, r, t, nullptr); }
double obj_ZSSD(const std::vector<double> &r, const std::vector<double> &t) { return objective(DVC_OBJ_ZSSD, r, t, nullptr); }
double obj_NSSD(const std::vector<double> &r, const std::vector<double> &t) { return objective(DVC_OBJ_NSSD, r, t, nullptr); }
double obj_ZNSSD(const std::vector<double> &r, const std::vector<double> &t) { return objective(DVC_OBJ_ZNSSD, r, t, nullptr); }
double obj_SAD(const std::vector<double> &r, const std::vector<double> &t, std::vector<double> &res) {
    return objective(DVC_OBJ_SAD, r, t, &res);
}
double obj_SSD(const std::vector<double> &r, const std::vector<double> &t, std::vector<double> &res) {
    return objective(DVC_OBJ_SSD, r, t, &res);
}
double obj_ZSSD(const std::vector<double> &r, const std::vector<double> &t, std::vector<double> &res) {
    return objective(DVC_OBJ_ZSSD, r, t, &res);
}
double obj_NSSD(const std::vector<double> &r, const std::vector<double> &t, std::vector<double> &res) {
    return objective(DVC_OBJ_NSSD, r, t, &res);
}
double obj_ZNSSD(const std::vector<double> &r, const std::vector<double> &t, std::vector<double> &res) {
    return objective(DVC_OBJ_ZNSSD, r, t, &res);
}

void load_volume_file(dvc_ctx *ctx, int slot, const std::string &fname, int bytes_per, const std::string &endian,
                      unsigned long long header, long long wide, long long high, long long tall) {
    Shared &s = shared();
    if (ctx == s.ctx && s.slot_file[slot] == fname) return;
    int fd = ::open(fname.c_str(), O_RDONLY);
    if (fd < 0) throw Input_Fail();
    struct stat st;
    if (::fstat(fd, &st) != 0) { ::close(fd); throw Input_Fail(); }
    const unsigned long long need = header + (unsigned long long)wide * high * tall * bytes_per;
    if ((unsigned long long)st.st_size < need) { ::close(fd); throw Input_Fail(); }
    void *map = ::mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    ::close(fd);
    if (map == MAP_FAILED) throw Input_Fail();
    const int big = (bytes_per == 2 && endian == "big") ? 1 : 0;
    int rc = dvc_volume_upload(ctx, slot, map, (int64_t)header, bytes_per, big, wide, high, tall);
    ::munmap(map, (size_t)st.st_size);
    check(rc, ctx, "dvc_volume_upload");
    if (ctx == s.ctx) s.slot_file[slot] = fname;
}

dvc_run_params run_params(const RunControl *run) {
    dvc_run_params p;
    std::memset(&p, 0, sizeof(p));
    p.ndof = run->num_srch_dof;
    p.obj = (int32_t)run->obj_fcn;
    p.interp = (int32_t)run->int_typ;
    p.geom = run->sub_geo == dvcio::sphere ? DVC_GEOM_SPHERE : DVC_GEOM_CUBE;
    p.subvol_size = run->subvol_size;
    p.subvol_npts = run->subvol_npts;
    p.disp_max = run->disp_max;
    p.sphere_seed = 1;                       // the reference's fixed seed (Search.cpp:102-105)
    p.basin_radius = run->basin_radius;
    for (int q = 0; q < 3; q++) {
        p.rigid_trans[q] = run->rigid_trans.size() > (size_t)q ? run->rigid_trans[q] : 0.0;
        p.aspect[q] = run->subvol_aspect.size() > (size_t)q ? run->subvol_aspect[q] : 1.0;
    }
    return p;
}

// ------------------------------------------------------------------ DataCloud

bool grid_neighbours(long long n) {
    const char *e = std::getenv("DVC_NEIGHBOURS");
    if (e && std::string(e) == "exact") return false;
    if (e && std::string(e) == "grid") return true;
    return n > kExactNeighbourLimit;
}

std::string neighbour_method(long long n) {
    if (!grid_neighbours(n)) return "neighbour tables: full sort per point (DataCloud.cpp:95-140 tie order)";
    return "neighbour tables: grid k-NN, equal distances in point-index order (the reference's std::sort tie order "
           "is not reproduced above " + std::to_string(kExactNeighbourLimit) +
           " points; DVC_NEIGHBOURS=exact forces the full sort)";
}

void DataCloud::sort_order_neighbors(Point start) {
    const int n = (int)points.size();
    std::vector<double> p((size_t)3 * n);
    for (int i = 0; i < n; i++) { p[3 * i] = points[i].x(); p[3 * i + 1] = points[i].y(); p[3 * i + 2] = points[i].z(); }
    const int k = n < 75 ? n : 75;
    std::vector<int32_t> ord(n), nb((size_t)n * k);
    const double s[3] = {start.x(), start.y(), start.z()};
    int rc = grid_neighbours(n) ? dvc_knn_order_neighbors(p.data(), n, s, k, ord.data(), nb.data(), 0)
                                : dvc_sort_order_neighbors(p.data(), n, s, ord.data(), nb.data(), 0);
    if (rc != k) throw Device_Fail{rc, "neighbour tables"};
    order.assign(ord.begin(), ord.end());
    neigh.assign(n, std::vector<int>(k));
    for (int i = 0; i < n; i++)
        for (int j = 0; j < k; j++) neigh[i][j] = nb[(size_t)i * k + j];
}

void DataCloud::organize_cloud(RunControl *run) {
    Point start = points[0];
    // DataCloud.cpp:149-153: a vector of NaNs never compares equal, so the run's
    // starting_point is always taken, NaN when the keyword was absent
    if (run->starting_point.size() == 3)
        start = Point(run->starting_point[0], run->starting_point[1], run->starting_point[2]);
    sort_order_neighbors(start);
    const int n = (int)points.size(), k = n ? (int)neigh[0].size() : 0;
    std::vector<int32_t> flat((size_t)n * k);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < k; j++) flat[(size_t)i * k + j] = neigh[i][j];
    dvcio::write_sort_file(run->pts_fname, flat.data(), n, k);
    results.assign(1, std::vector<ResultRecord>(points.size()));
    for (auto &r : results[0]) {
        r.status = dvcio::not_search;
        r.obj_min = 0.0;
        r.par_min.assign(run->num_srch_dof, 0.0);
    }
}

// ---------------------------------------------------------------- Interpolate

Interpolate::Interpolate(const BoundBox *region) {
    rmin_[0] = region->min().x(); rmin_[1] = region->min().y(); rmin_[2] = region->min().z();
    rmax_[0] = region->max().x(); rmax_[1] = region->max().y(); rmax_[2] = region->max().z();
    for (int q = 0; q < 3; q++) center_[q] = (rmin_[q] + rmax_[q]) / 2.0;
}

void Interpolate::kernels(const std::string &voxfname, const BoundBox *vox_box, int bytes_per,
                          const std::string &endian, unsigned int offset) {
    dvc_ctx *ctx = shared_context();
    Shared &s = shared();
    int slot = -1;
    for (int q = 0; q < 2; q++)
        if (s.slot_file[q] == voxfname) slot = q;
    if (slot < 0) slot = 1 - s.last_slot;
    load_volume_file(ctx, slot, voxfname, bytes_per, endian, offset, (long long)vox_box->max().x(),
                     (long long)vox_box->max().y(), (long long)vox_box->max().z());
    s.last_slot = slot;
    slot_ = slot;
}

void Interpolate::center_on(Point pt) {
    center_[0] = pt.x(); center_[1] = pt.y(); center_[2] = pt.z();
}

void Interpolate::run(int method, const std::vector<Point> &pts, const BoundBox *bbox, std::vector<double> &ivals) {
    if (slot_ < 0) throw Intrp_Fail();
    dvc_ctx *ctx = shared_context();
    const int n = (int)pts.size();
    std::vector<double> xyz((size_t)3 * n);
    for (int i = 0; i < n; i++) { xyz[3 * i] = pts[i].x(); xyz[3 * i + 1] = pts[i].y(); xyz[3 * i + 2] = pts[i].z(); }
    const double bb[6] = {bbox->min().x(), bbox->min().y(), bbox->min().z(),
                          bbox->max().x(), bbox->max().y(), bbox->max().z()};
    ivals.assign(n, 0.0);
    int rc = dvc_interpolate_box(ctx, slot_, rmin_, rmax_, center_, xyz.data(), n, bb, method, ivals.data(), nullptr);
    if (rc == DVC_ERR_RANGE) throw Intrp_Fail();
    check(rc, ctx, "dvc_interpolate_box");
}

void Interpolate::nearest(const std::vector<Point> &pts, const BoundBox *bbox, std::vector<double> &ivals) {
    run(DVC_INTERP_NEAREST, pts, bbox, ivals);
}
void Interpolate::tri_lin(const std::vector<Point> &pts, const BoundBox *bbox, std::vector<double> &ivals) {
    run(DVC_INTERP_TRILINEAR, pts, bbox, ivals);
}
void Interpolate::tri_cub_Lek(const std::vector<Point> &pts, const BoundBox *bbox, std::vector<double> &ivals) {
    run(DVC_INTERP_TRICUBIC, pts, bbox, ivals);
}

// --------------------------------------------------------------------- Search

Search::Search(RunControl *run) : rc(run) {
    params_ = run_params(run);
    ctx_ = shared_context();
    bytes_per = run->vol_bit_depth / 8;
    subv_rad = run->subvol_size / 2.0;
    if (run->sub_geo == dvcio::cube) subv_num = (int)std::pow(std::ceil(std::pow((double)run->subvol_npts, 1.0 / 3.0)), 3.0);
    else subv_num = run->subvol_npts;
    par_min.assign(run->num_srch_dof, 0.0);
    obj_min = 0.0;
    load_volume_file(ctx_, DVC_VOL_REF, run->ref_fname, bytes_per, run->vol_endian, (unsigned long long)run->vol_hdr_lngth,
                     run->vol_wide, run->vol_high, run->vol_tall);
    load_volume_file(ctx_, DVC_VOL_COR, run->cor_fname, bytes_per, run->vol_endian, (unsigned long long)run->vol_hdr_lngth,
                     run->vol_wide, run->vol_high, run->vol_tall);
}

Search::~Search() {}

void Search::starting_param(Point /*srch_pt*/, std::vector<ResultRecord> &neigh_res) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    int good = 0;
    for (const ResultRecord &r : neigh_res)
        if (r.status == dvcio::point_good) {
            sx += r.par_min[0]; sy += r.par_min[1]; sz += r.par_min[2];
            good += 1;
        }
    for (size_t q = 0; q < par_min.size(); q++) par_min[q] = 0.0;
    if (good != 0) {
        sx /= good; sy /= good; sz /= good;
        par_min[0] = sx; par_min[1] = sy; par_min[2] = sz;
    } else {
        par_min[0] = params_.rigid_trans[0]; par_min[1] = params_.rigid_trans[1]; par_min[2] = params_.rigid_trans[2];
    }
}

void Search::search_pt_setup(Point srch_pt, std::vector<ResultRecord> &neigh_res) {
    starting_param(srch_pt, neigh_res);
    pt_[0] = srch_pt.x(); pt_[1] = srch_pt.y(); pt_[2] = srch_pt.z();
    for (int q = 0; q < 3; q++) start_[q] = par_min[q];
    setup_ = true;
}

double Search::evaluate(const std::vector<double> &x, std::vector<double> *residual) {
    if (!setup_) throw Device_Fail{DVC_ERR_STATE, "obj_val_at before search_pt_setup"};
    int32_t n = 0;
    double v = 0.0;
    int rcode = dvc_obj_val_at(ctx_, &params_, pt_, start_, x.data(), (int32_t)x.size(), &v, nullptr, 0, &n);
    if (rcode == DVC_OK && residual) {
        residual->assign(n, 0.0);
        rcode = dvc_obj_val_at(ctx_, &params_, pt_, start_, x.data(), (int32_t)x.size(), &v, residual->data(), n, &n);
    }
    if (rcode == DVC_ERR_RANGE) throw Range_Fail();
    check(rcode, ctx_, "dvc_obj_val_at");
    return v;
}

double Search::obj_val_at(const std::vector<double> x) { return evaluate(x, nullptr); }

double Search::obj_val_at(const std::vector<double> x, std::vector<double> &residual) {
    return evaluate(x, &residual);
}

void Search::process_point(int t, int n, DataCloud *data, int /*test*/) {
    // Search::search_pt_setup / starting_param (Search.cpp:248-377) from the neighbours' records
    std::vector<ResultRecord> neigh_res;
    for (int m : data->neigh[n]) neigh_res.push_back(data->results[t][m]);
    search_pt_setup(data->points[n], neigh_res);
    dvc_run_params p = params_;
    for (int q = 0; q < 3; q++) p.rigid_trans[q] = start_[q];
    // one point whose only neighbour is itself: its start guess is p.rigid_trans
    const double pt[3] = {data->points[n].x(), data->points[n].y(), data->points[n].z()};
    const int32_t ord = 0, nb = 0;
    dvc_point_result res;
    check(dvc_search_points(ctx_, &p, pt, 1, &ord, 0, &nb, 1, &res), ctx_, "dvc_search_points");
    if (res.status == DVC_POINT_GOOD) {
        for (size_t q = 0; q < par_min.size(); q++) par_min[q] = res.par[q];
        obj_min = res.obj;
        throw Point_Good();
    }
    if (res.status == DVC_CONVG_FAIL) throw Convg_Fail();
    throw Range_Fail();
}

void Search::process_all(int t, DataCloud *data, unsigned int N, std::vector<dvc_point_result> &out) {
    const int n = (int)data->points.size();
    const int k = n ? (int)data->neigh[0].size() : 0;
    std::vector<double> p((size_t)3 * n);
    for (int i = 0; i < n; i++) {
        p[3 * i] = data->points[i].x(); p[3 * i + 1] = data->points[i].y(); p[3 * i + 2] = data->points[i].z();
    }
    std::vector<int32_t> ord(data->order.begin(), data->order.end()), nb((size_t)n * k);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < k; j++) nb[(size_t)i * k + j] = data->neigh[i][j];
    out.assign(n, dvc_point_result());
    check(dvc_search_points(ctx_, &params_, p.data(), n, ord.data(), (int32_t)N, nb.data(), k, out.data()), ctx_,
          "dvc_search_points");
    // dvc.cpp:263-283: records keep par_min / obj_min of good points only
    const unsigned int M = (N > 1 && N < (unsigned)n) ? N : (unsigned)n;
    for (unsigned int i = 0; i < M; i++) {
        const int m = data->order[i];
        ResultRecord &r = data->results[t][m];
        r.status = out[m].status;
        if (out[m].status == DVC_POINT_GOOD) {
            r.obj_min = out[m].obj;
            for (size_t q = 0; q < r.par_min.size(); q++) r.par_min[q] = out[m].par[q];
        }
    }
}

}  // namespace dvc
