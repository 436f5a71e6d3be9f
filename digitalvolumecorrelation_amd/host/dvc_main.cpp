// dvc -- the dvc program of the reference (Core/dvc.cpp:35-349) on the MI355X path:
// the same keyword input file, ROI file, <roi>.sort.csv, <output>.disp and
// <output>.stat, with the point loop run on the device by Search::process_all.
//
//   dvc <input file>     run
//   dvc help             how to run
//   dvc example          write dvc_in_example (every keyword with an example value)
//   dvc version          print the version
#include <chrono>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "dvc_io.h"
#include "dvc_search.h"

namespace {

const char *kVersion = "dvc-mi355x 0.1 (gfx950)";

void usage() {
    std::cout << "\nOptions:\n"
                 "dvc dvc_in\t\t// execute dvc code with dvc_in controlling the run\n"
                 "dvc help\t\t// provide additional detail about running the dvc code\n"
                 "dvc example\t\t// print dvc_in_example with brief keyword descriptions\n"
                 "dvc version\t\t// print the code version\n\n";
}

void help() {
    std::cout << "\nProgram execution is controlled by a key_word based input file.\n"
                 "To run the code type dvc followed by the name of the input file.\n"
                 "Results are written to output_filename.disp, run information to output_filename.stat,\n"
                 "and point order / neighbours to point_cloud_filename.sort.csv.\n"
                 "All points are correlated on the GPU; the output matches the serial point loop.\n\n";
}

void example() {
    std::ofstream f("dvc_in_example");
    f << "### dvc input file example: keyword, tab, value; lines starting with # are ignored\n\n"
         "reference_filename\treference_volume.raw\t### reference image volume\n"
         "correlate_filename\tcorrelate_volume.raw\t### image volume to correlate\n"
         "point_cloud_filename\tROI_points_nxyz.txt\t### lines of: label x y z\n"
         "output_filename\tmyproject_results\t### base name of the .disp and .stat files\n\n"
         "vol_bit_depth\t8\t\t### 8 or 16\n"
         "vol_endian\tlittle\t\t### little or big (16 bit only)\n"
         "vol_hdr_lngth\t0\t\t### header bytes before the voxels\n"
         "vol_wide\t4008\t\t### voxels along x\n"
         "vol_high\t2670\t\t### voxels along y\n"
         "vol_tall\t4016\t\t### voxels along z (slices)\n\n"
         "subvol_geom\tcube\t\t### cube or sphere\n"
         "subvol_size\t30\t\t### side or diameter in voxels\n"
         "subvol_npts\t5000\t\t### samples per subvolume (cube: rounded up to a cube)\n"
         "subvol_thresh\toff\t\t### on or off\n\n"
         "disp_max\t15\t\t### voxels, search range\n"
         "num_srch_dof\t6\t\t### 3, 6 or 12\n"
         "obj_function\tznssd\t\t### sad ssd zssd nssd znssd\n"
         "interp_type\ttricubic\t### nearest trilinear tricubic\n\n"
         "rigid_trans\t0.0 0.0 0.0\t### start translation without converged neighbours\n"
         "basin_radius\t0.0\t\t### translation grid step, 0.0 = none\n"
         "subvol_aspect\t1.0 1.0 1.0\t### subvolume aspect ratio\n"
         "starting_point\t0.0 0.0 0.0\t### processing starts at the point nearest this location\n"
         "num_points_to_process\t0\t### 0 or 1: all points\n";
    std::cout << "\ndvc_in_example printed in the current working directory\n\n";
}

}  // namespace

int main(int argc, char *argv[]) {
    using namespace dvcio;
    if (argc == 1) { usage(); return 0; }
    const std::string arg1(argv[1]);
    if (arg1 == "help") { help(); return 0; }
    if (arg1 == "example") { example(); return 0; }
    if (arg1 == "version") { std::cout << kVersion << std::endl; return 0; }

    InputRead in;
    if (!in.input_file_accessible(arg1)) return 0;
    RunControl run;
    if (!in.input_file_read(&run)) return 0;

    dvc::DataCloud data;
    std::vector<double> pts;
    if (!in.read_point_cloud(&run, pts, data.labels)) return 0;
    for (size_t i = 0; i < data.labels.size(); i++) data.points.emplace_back(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    data.organize_cloud(&run);

    const long long t_start = (long long)std::time(nullptr);
    in.echo_input(&run, kVersion);
    in.append_time_date(run.sta_fname, "Run start: ", t_start);

    const unsigned int npts = (unsigned int)data.points.size();
    unsigned int N = npts;
    if (run.num_points_to_process > 1 && run.num_points_to_process < N) N = run.num_points_to_process;

    std::vector<dvc_point_result> out;
    double elapsed = 0.0;
    try {
        dvc::Search optimize(&run);
        auto t0 = std::chrono::steady_clock::now();
        optimize.process_all(0, &data, run.num_points_to_process, out);
        auto t1 = std::chrono::steady_clock::now();
        elapsed = std::chrono::duration<double>(t1 - t0).count();
    } catch (const dvc::Device_Fail &e) {
        std::cout << "\n-> device failure: " << e.what << " (code " << e.code << ")\n\n";
        return 1;
    } catch (const dvc::Input_Fail &) {
        std::cout << "\n-> cannot read the image volumes\n\n";
        return 1;
    }

    // console lines of the point loop (dvc.cpp:248-294), in processing order
    int count_good = 0, count_range = 0, count_convg = 0;
    for (unsigned int i = 0; i < N; i++) {
        const int n = data.order[i];
        const dvc::ResultRecord &r = data.results[0][n];
        std::cout << i << "/" << N << " " << data.labels[n] << "\t" << std::setprecision(3) << std::fixed
                  << data.points[n].x() << " " << data.points[n].y() << " " << data.points[n].z() << " \t";
        if (r.status == point_good) {
            count_good += 1;
            std::cout << "Point_Good\t" << std::setprecision(6) << "obj= " << r.obj_min << std::setw(12) << "dx= "
                      << r.par_min[0] << std::setw(12) << "dy= " << r.par_min[1] << std::setw(12) << "dz= "
                      << r.par_min[2];
        } else if (r.status == range_fail) {
            count_range += 1;
            std::cout << "Range_Fail";
        } else if (r.status == convg_fail) {
            count_convg += 1;
            std::cout << "Convg_Fail";
        }
        std::cout << "\n";
    }
    std::cout.unsetf(std::ios::floatfield);
    const double count = (double)N;
    std::cout << count / elapsed << " pt/sec average" << std::endl;

    // .disp in processing order from the stored records (dvc.cpp:312-320)
    std::string disp = disp_header();
    disp.reserve((size_t)N * 80);
    for (unsigned int i = 0; i < N; i++) {
        const int n = data.order[i];
        const dvc::ResultRecord &r = data.results[0][n];
        const double p[3] = {data.points[n].x(), data.points[n].y(), data.points[n].z()};
        const double u[3] = {r.par_min.size() > 0 ? r.par_min[0] : 0.0, r.par_min.size() > 1 ? r.par_min[1] : 0.0,
                             r.par_min.size() > 2 ? r.par_min[2] : 0.0};
        disp_row(disp, data.labels[n], p, r.status, r.obj_min, u);
    }
    if (!write_text(run.res_fname, disp)) {
        std::cout << "\n-> cannot write " << run.res_fname << "\n\n";
        return 1;
    }

    // .stat (dvc.cpp:322-338)
    const long long t_end = (long long)std::time(nullptr);
    in.echo_input(&run, kVersion);
    in.append_time_date(run.sta_fname, "Run start:\t", t_start);
    in.append_time_date(run.sta_fname, "Run finish:\t", t_end);
    std::ofstream sta(run.sta_fname, std::ios::out | std::ios::app);
    sta << dvc::neighbour_method(npts) << std::endl << std::endl;
    sta << std::setprecision(0) << std::fixed;
    sta << N << " points processed in " << elapsed << " seconds" << std::endl;
    sta << std::setprecision(3) << std::fixed;
    sta << elapsed / count << " sec/pt" << std::endl;
    sta << count / elapsed << " pt/sec" << std::endl;
    sta << std::endl;
    sta << "number successful = " << count_good << "\t(" << 100 * ((double)count_good / count) << "%)" << std::endl;
    sta << "number range fail = " << count_range << "\t(" << 100 * ((double)count_range / count) << "%)" << std::endl;
    sta << "number convg fail = " << count_convg << "\t(" << 100 * ((double)count_convg / count) << "%)" << std::endl;
    sta << std::endl;
    return 0;
}
