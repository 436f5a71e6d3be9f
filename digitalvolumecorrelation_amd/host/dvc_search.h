// dvc_search.h -- the reference's C++ call surface for the correlation path, implemented
// over the C ABI (include/dvc_amd.h), so C++ callers written against Search,
// Interpolate, DataCloud and the obj_* functions keep working:
//
//   obj_SAD .. obj_ZNSSD       Core/ObjectiveFunctions.h:31-45
//   Point, BoundBox            Core/Point.h, Core/BoundBox.h (the parts the path uses)
//   DataCloud                  Core/DataCloud.h:59-110 (points, labels, order, neigh, results)
//   Interpolate                Core/Interpolate.h:48-80
//   Search                     Core/Search.h:51-110 (process_point throws Point_Good,
//                              Range_Fail or Convg_Fail as Core/Utility.h:77-79)
//
// Search::process_all runs the whole dvc point loop (Core/dvc.cpp:229-303) on the device
// in one call; process_point runs one point at a time with its start guess taken from
// the DataCloud results, as the reference loop does.
#pragma once

#include <cmath>
#include <string>
#include <vector>

#include "../../include/dvc_amd.h"
#include "dvc_io.h"

namespace dvc {

using dvcio::RunControl;

// Core/Utility.h:77-83
class Point_Good {};
class Range_Fail {};
class Convg_Fail {};
class Bound_Fail {};
class Intrp_Fail {};
class Input_Fail {};

// Device failure (no GPU, HIP error): not part of the reference's set.
struct Device_Fail {
    int code;
    std::string what;
};

class Point {
public:
    Point() : x_(0), y_(0), z_(0) {}
    Point(double x, double y, double z) : x_(x), y_(y), z_(z) {}
    double x() const { return x_; }
    double y() const { return y_; }
    double z() const { return z_; }
    void move_to(double x, double y, double z) { x_ = x; y_ = y; z_ = z; }
    void move_by(double u, double v, double w) { x_ += u; y_ += v; z_ += w; }
    // Point::pt_dist (Core/Point.h:47-48)
    double pt_dist(const Point &p) const {
        return std::sqrt(std::pow(x_ - p.x_, 2) + std::pow(y_ - p.y_, 2) + std::pow(z_ - p.z_, 2));
    }

private:
    double x_, y_, z_;
};

class BoundBox {
public:
    BoundBox(const Point &mn, const Point &mx) : mn_(mn), mx_(mx) {}
    const Point &min() const { return mn_; }
    const Point &max() const { return mx_; }
    void move_to(const Point &mn, const Point &mx) { mn_ = mn; mx_ = mx; }

private:
    Point mn_, mx_;
};

// Core/DataCloud.h:45-56
struct ResultRecord {
    int status;
    double obj_min;
    std::vector<double> par_min;
};

class DataCloud {
public:
    // DataCloud::organize_cloud (DataCloud.cpp:146-175): order and neighbours from the
    // starting point (run->starting_point, which the reference takes even when it was
    // not given and is NaN), the <pts_fname>.sort.csv file, and result records set to
    // not_search.
    void organize_cloud(RunControl *run);
    // DataCloud::sort_order_neighbors (DataCloud.cpp:69-144)
    void sort_order_neighbors(Point starting_point);
    int nbr_num_save() const { return 75; }

    std::vector<Point> points;
    std::vector<int> labels;
    std::vector<int> order;
    std::vector<std::vector<int>> neigh;
    std::vector<std::vector<ResultRecord>> results;
};

// Neighbour tables: the reference's full sort per point (exact tie order, O(n^2 log n))
// up to kExactNeighbourLimit points, the grid k-NN (ties in index order, O(n k)) above;
// DVC_NEIGHBOURS=exact or =grid chooses explicitly.
constexpr long long kExactNeighbourLimit = 20000;
bool grid_neighbours(long long n);
// One line naming the neighbour method used for n points, for the .stat file.
std::string neighbour_method(long long n);

// A process-wide device context for the objects below (device from DVC_DEVICE, else 0).
dvc_ctx *shared_context();

double obj_SAD(const std::vector<double> &ref_subvol, const std::vector<double> &tar_subvol);
double obj_SSD(const std::vector<double> &ref_subvol, const std::vector<double> &tar_subvol);
double obj_ZSSD(const std::vector<double> &ref_subvol, const std::vector<double> &tar_subvol);
double obj_NSSD(const std::vector<double> &ref_subvol, const std::vector<double> &tar_subvol);
double obj_ZNSSD(const std::vector<double> &ref_subvol, const std::vector<double> &tar_subvol);
double obj_SAD(const std::vector<double> &ref_subvol, const std::vector<double> &tar_subvol,
               std::vector<double> &residual);
double obj_SSD(const std::vector<double> &ref_subvol, const std::vector<double> &tar_subvol,
               std::vector<double> &residual);
double obj_ZSSD(const std::vector<double> &ref_subvol, const std::vector<double> &tar_subvol,
                std::vector<double> &residual);
double obj_NSSD(const std::vector<double> &ref_subvol, const std::vector<double> &tar_subvol,
                std::vector<double> &residual);
double obj_ZNSSD(const std::vector<double> &ref_subvol, const std::vector<double> &tar_subvol,
                 std::vector<double> &residual);

// Volume file contents after the header, uploaded once per (file, slot) and kept
// resident.  Interpolate::kernels and Search use it in place of per-point file reads.
void load_volume_file(dvc_ctx *ctx, int slot, const std::string &fname, int bytes_per, const std::string &endian,
                      unsigned long long header, long long wide, long long high, long long tall);

class Interpolate {
public:
    explicit Interpolate(const BoundBox *region);
    // Interpolate::kernels: make the volume in `voxfname` (x fastest, `bytes_per`
    // bytes per voxel, `endian`, `offset` header bytes, size vox_box->max()) the source
    void kernels(const std::string &voxfname, const BoundBox *vox_box, int bytes_per, const std::string &endian,
                 unsigned int offset);
    void kernels(const std::string &voxfname, const BoundBox *vox_box, int bytes_per, const std::string &endian) {
        kernels(voxfname, vox_box, bytes_per, endian, 0);
    }
    // interpolation at pts whose bounding box is bbox; throw Intrp_Fail outside the
    // active box (Interpolate.cpp:404-567)
    void nearest(const std::vector<Point> &pts, const BoundBox *bbox, std::vector<double> &ivals);
    void tri_lin(const std::vector<Point> &pts, const BoundBox *bbox, std::vector<double> &ivals);
    void tri_cub_Lek(const std::vector<Point> &pts, const BoundBox *bbox, std::vector<double> &ivals);
    void center_on(Point pt);

private:
    void run(int method, const std::vector<Point> &pts, const BoundBox *bbox, std::vector<double> &ivals);
    double rmin_[3], rmax_[3], center_[3];
    int slot_ = -1;
};

class Search {
public:
    explicit Search(RunControl *run);
    ~Search();
    // one point, start guess from srch_data->results[t] of its neighbours; throws
    void process_point(int t, int n, DataCloud *srch_data, int test = 0);
    // the whole loop over the first N points of srch_data->order, on the device;
    // fills srch_data->results[t] as the reference loop does
    void process_all(int t, DataCloud *srch_data, unsigned int N, std::vector<dvc_point_result> &out);

    // Search::search_pt_setup (Search.cpp:248-275): make srch_pt the point obj_val_at
    // evaluates, with par_min set by starting_param from the neighbours' records; the
    // correlated volume's interpolation box then sits on srch_pt moved by that start
    void search_pt_setup(Point srch_pt, std::vector<ResultRecord> &neigh_res);
    // Search::starting_param (Search.cpp:277-377): par_min = the mean translation of the
    // good records, else rc->rigid_trans; rotation and strain 0
    void starting_param(Point srch_pt, std::vector<ResultRecord> &neigh_res);
    // Search::obj_val_at (Search.h:88-89, Search.cpp:448-488) at the point of the last
    // search_pt_setup or process_point; throws Range_Fail
    double obj_val_at(const std::vector<double> x);
    double obj_val_at(const std::vector<double> x, std::vector<double> &residual);

    int bytes_per;
    double subv_rad;
    int subv_num;
    std::vector<double> par_min;
    double obj_min;

private:
    RunControl *rc;
    dvc_run_params params_;
    dvc_ctx *ctx_;
    bool setup_ = false;             // a point is set up for obj_val_at
    double pt_[3] = {0.0, 0.0, 0.0};
    double start_[3] = {0.0, 0.0, 0.0};
    double evaluate(const std::vector<double> &x, std::vector<double> *residual);
};

dvc_run_params run_params(const RunControl *run);

}  // namespace dvc
