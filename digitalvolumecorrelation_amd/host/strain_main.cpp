// strain -- the strain program of the reference (Core/strain.cpp:136-660) on the MI355X
// path: reads a .disp file (and optionally a .sort.csv file), fits the quadratic
// displacement model over each point's strain window on the device (dvc_strain_fit) and
// writes <base>-sw<ndp>.Lstr.csv (and .Estr.csv / .dgrd.csv with -E / -D / -A).
//
//   strain [flags] file.disp [file.sort.csv]
//   -sw N  strain window points (10..75, default 25)    -t X  objmin threshold (1e-4..1)
//   -r     refill the window past failing points         -E / -D / -A  more outputs
//   -xy / -yz / -zx  planar analysis (also detected from the point positions)
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/dvc_amd.h"
#include "dvc_io.h"
#include "dvc_search.h"

namespace {

constexpr int kNdpMin = 10, kNdpDef = 25, kNdpMax = 75;
constexpr double kObjtMin = 0.0001, kObjtMax = 1.0, kObjtDef = 1.0;

// StrainCalc::find_flag (strain.cpp:39-135): remove the flag (and its value) from argv.
bool take_flag(const std::string &flag, int &argc, char *argv[]) {
    for (int i = 0; i < argc; i++)
        if (flag == argv[i]) {
            for (int j = i; j < argc - 1; j++) argv[j] = argv[j + 1];
            argc -= 1;
            return true;
        }
    return false;
}

template <typename T>
bool take_flag(const std::string &flag, int &argc, char *argv[], T &val) {
    for (int i = 0; i < argc; i++) {
        if (flag != argv[i]) continue;
        if (argc > i + 1) {
            try {
                if constexpr (std::is_same<T, int>::value) val = std::stoi(argv[i + 1]);
                else val = std::stod(argv[i + 1]);
                for (int j = i; j < argc - 2; j++) argv[j] = argv[j + 2];
                argc -= 2;
                return true;
            } catch (const std::exception &) {
                for (int j = i; j < argc - 1; j++) argv[j] = argv[j + 1];
                argc -= 1;
                std::cout << "argument after " << flag << " not valid" << std::endl;
                return false;
            }
        }
        std::cout << "no argument after " << flag << " flag" << std::endl;
        argc -= 1;
        return false;
    }
    return false;
}

}  // namespace

int main(int argc, char *argv[]) {
    using namespace dvcio;
    std::cout << std::endl;
    int ndp = kNdpDef;
    double objt = kObjtDef;
    bool refill = false, xy = false, yz = false, zx = false, out_E = false, out_L = true, out_D = false;
    if (argc == 1) {
        std::cout << "Strain command needs more input.\n\nCommand line:\nstrain dvc.disp\nstrain dvc.disp dvc.sort.csv\n\n"
                     "Flags:\n-sw 30\t\tstrain window points (" << kNdpMin << ".." << kNdpMax << ", default " << kNdpDef
                  << ")\n-t 0.01\t\tobjmin threshold\n-r\t\trefill strain windows\n-E\t\tadd engineering strain output\n"
                     "-D\t\tadd displacement gradient output\n-A\t\twrite all outputs\n\n";
        return 0;
    }
    if (take_flag("-sw", argc, argv, ndp)) ndp = std::min(std::max(ndp, kNdpMin), kNdpMax);
    if (take_flag("-t", argc, argv, objt)) objt = std::min(std::max(objt, kObjtMin), kObjtMax);
    if (take_flag("-r", argc, argv)) refill = true;
    if (take_flag("-xy", argc, argv)) xy = true;
    if (take_flag("-yz", argc, argv)) yz = true;
    if (take_flag("-zx", argc, argv)) zx = true;
    if (take_flag("-E", argc, argv)) out_E = true;
    if (take_flag("-D", argc, argv)) out_D = true;
    if (take_flag("-A", argc, argv)) { out_E = true; out_D = true; }
    for (int i = 1; i < argc;) {
        if (argv[i][0] == '-') {
            std::cout << "-> unused flag " << argv[i] << " found, ignored" << std::endl;
            for (int j = i; j < argc - 1; j++) argv[j] = argv[j + 1];
            argc -= 1;
        } else {
            i++;
        }
    }
    if (argc < 2) return 0;
    const std::string fdisp = argv[1];
    {
        std::ifstream t(fdisp);
        if (!t.good()) { std::cout << "-> cannot find .disp file '" << fdisp << std::endl; return 0; }
    }
    std::cout << "using disp file: " << fdisp << std::endl;
    const size_t dot = fdisp.find_first_of('.');
    const std::string base = dot == std::string::npos ? fdisp : fdisp.substr(0, dot);

    std::vector<int> labels, status;
    std::vector<double> pts, objmin, disp;
    if (!read_disp_file(fdisp, labels, pts, status, objmin, disp)) return 0;
    const int n = (int)labels.size();
    if (n == 0) { std::cout << "-> no points in " << fdisp << std::endl; return 0; }

    std::vector<std::vector<int>> neigh;
    bool sorted = false;
    if (argc > 2) {
        if (read_sort_file(argv[2], neigh)) {
            std::cout << "using sort file: " << argv[2] << std::endl;
            sorted = true;
        } else {
            std::cout << "-> cannot find .sort file " << argv[2] << std::endl;
        }
    }
    if (sorted && (int)neigh.size() != n) {
        sorted = false;
        std::cout << "-> different number of points in .disp and .sort" << std::endl;
    }
    int k = 0;
    std::vector<int32_t> nb;
    if (sorted) {
        k = (int)neigh[0].size();
        for (const auto &row : neigh)
            if ((int)row.size() != k) { std::cout << "-> ragged .sort file" << std::endl; return 0; }
        nb.resize((size_t)n * k);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < k; j++) nb[(size_t)i * k + j] = neigh[i][j];
    } else {
        std::cout << "-> creating a new sort file: " << base + ".sort.csv" << std::endl;
        k = n < kNdpMax ? n : kNdpMax;
        std::vector<int32_t> order(n);
        nb.resize((size_t)n * k);
        const int rc = dvc::grid_neighbours(n) ? dvc_knn_order_neighbors(pts.data(), n, pts.data(), k, order.data(),
                                                                         nb.data(), 0)
                                               : dvc_sort_order_neighbors(pts.data(), n, pts.data(), order.data(),
                                                                          nb.data(), 0);
        if (rc != k) return 1;
        write_sort_file(base, nb.data(), n, k);
    }
    if (yz) std::cout << "using 2d yz coordinate plane analsis" << std::endl;
    if (zx) std::cout << "using 2d zx coordinate plane analsis" << std::endl;
    if (xy) std::cout << "using 2d xy coordinate plane analsis" << std::endl;
    std::cout << "using strain window: " << ndp << std::endl;
    std::cout << "using objmin thresh: " << objt << std::endl;
    if (refill) std::cout << "using strain window refill " << std::endl;

    dvc_ctx *ctx = nullptr;
    if (dvc_create(&ctx, 0) != DVC_OK) { std::cout << "\n-> no usable gfx950 device\n\n"; return 1; }
    std::vector<double> fdisp3((size_t)3 * n), swr(n), E((size_t)9 * n), L((size_t)9 * n), DG((size_t)9 * n);
    std::vector<int32_t> npin(n);
    const int planar = (yz ? 1 : 0) | (zx ? 2 : 0) | (xy ? 4 : 0);
    int rc = dvc_strain_fit(ctx, pts.data(), disp.data(), status.data(), objmin.data(), nb.data(), k, n, ndp, objt,
                            refill ? 1 : 0, planar, fdisp3.data(), npin.data(), swr.data(), E.data(), L.data(),
                            DG.data());
    if (rc != DVC_OK) {
        std::cout << "\n-> strain fit failed: " << dvc_last_error(ctx) << "\n\n";
        dvc_destroy(ctx);
        return 1;
    }
    dvc_destroy(ctx);

    const std::string hdr = "n,x,y,z,u_fit,v_fit,w_fit,pts_in_sw,sw_radius,exx,eyy,ezz,exy,eyz,exz,ep1,ep2,ep3\n";
    auto write_csv = [&](const std::string &suffix, const std::string &head, const std::vector<double> &vals) {
        const std::string name = base + "-sw" + std::to_string(ndp) + suffix;
        std::cout << "output file: " << name << std::endl;
        std::string text = head;
        text.reserve((size_t)n * 200);
        for (int i = 0; i < n; i++)
            csv_row(text, labels[i], &pts[3 * i], &fdisp3[3 * i], 3, npin[i], swr[i], &vals[9 * (size_t)i], 9);
        write_text(name, text);
    };
    if (out_E) write_csv(".Estr.csv", hdr, E);
    if (out_L) write_csv(".Lstr.csv", hdr, L);
    if (out_D)
        write_csv(".dgrd.csv", "n,x,y,z,u_fit,v_fit,w_fit,pts_in_sw,sw_radius,ux,uy,uz,vx,vy,vz,wx,wy,wz\n", DG);
    std::cout << std::endl;
    return 0;
}
