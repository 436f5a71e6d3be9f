// dvc_io.h -- host-side run control, input-file parsing, point-cloud and result files
// of the dvc and strain programs.  Pure C++ (no device code): the formats and the
// parsing rules follow the reference's InputRead / DispRead / DataCloud file code so
// that existing input files, ROI files, .disp, .stat, .sort.csv and strain CSV files
// are read and written the same way.
//
//   RunControl            Core/Utility.h:90-134
//   InputRead             Core/InputRead.cpp:477-716 (keyword lines, limits, defaults)
//   echo / result files   Core/InputRead.cpp:1033-1196
//   DispRead              Core/InputRead.cpp:718-860
//   sort file             Core/DataCloud.cpp:43-67
#pragma once

#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

namespace dvcio {

// Core/Utility.h:31-41
constexpr int input_line_ok = 1;
constexpr int keywd_missing = -1;
constexpr int param_invalid = -2;

constexpr int point_good = 0;
constexpr int range_fail = -1;
constexpr int convg_fail = -2;
constexpr int not_search = -3;

enum Subvol_Type { cube, sphere };
enum Objfcn_Type { SAD, SSD, ZSSD, NSSD, ZNSSD };
enum Interp_Type { nearest, trilinear, tricubic };

// The values read and checked from the input file (Core/Utility.h:90-134).
struct RunControl {
    std::string ref_fname, cor_fname, pts_fname, out_fname, res_fname, sta_fname;
    int vol_bit_depth = 8;
    std::string vol_endian = "little";
    int vol_hdr_lngth = 0, vol_wide = 0, vol_high = 0, vol_tall = 0;
    std::string subvol_geom;
    int subvol_size = 0, subvol_npts = 0;
    std::string subvol_thresh;
    double gray_thresh_min = 0.0, gray_thresh_max = 0.0, min_vol_fract = 0.0;
    int disp_max = 0, num_srch_dof = 0;
    unsigned int num_points_to_process = 0;
    Objfcn_Type obj_fcn = ZNSSD;
    Interp_Type int_typ = tricubic;
    Subvol_Type sub_geo = cube;
    std::string obj_function, interp_type;
    std::vector<double> rigid_trans, subvol_aspect, starting_point;
    double basin_radius = 0.0;
};

// Keyword-driven input file (Core/InputRead.cpp:477-608).  Messages go to stdout as in
// the reference; every read returns 1 on success and 0 on failure.
class InputRead {
public:
    int input_file_accessible(const std::string &fname);
    int input_file_read(RunControl *run);
    // ROI file: "label x y z" lines, anything else skipped (InputRead.cpp:632-716)
    int read_point_cloud(const RunControl *run, std::vector<double> &points, std::vector<int> &labels);
    // .stat echo (InputRead.cpp:1033-1106); truncates the file
    int echo_input(const RunControl *run, const std::string &version) const;
    int append_time_date(const std::string &fname, const std::string &label, long long t) const;

    double box_min[3] = {0, 0, 0}, box_max[3] = {0, 0, 0};
    long long num_pts = 0;

private:
    int get_line_with_keyword(const std::string &keyword, std::string &keyline, bool req);
    int parse_old_file(const char *word, std::string &arg, unsigned long long &bytes, bool req);
    int parse_new_file(const char *word, std::string &arg, bool req);
    int parse_int_in(const char *word, const std::vector<int> &vals, int &arg, bool req);
    int parse_str_in(const char *word, const std::vector<std::string> &vals, std::string &arg, bool req);
    int parse_int(const char *word, int lo, int hi, int &arg, bool req);
    int parse_uint(const char *word, unsigned lo, unsigned hi, unsigned &arg, bool req);
    int parse_double(const char *word, double lo, double hi, double &arg, bool req);
    int parse_dvect(const char *word, const std::vector<double> &lim, std::vector<double> &vals, bool req);
    int parse_dvect(const char *word, double lo, double hi, std::vector<double> &vals, bool req);
    void invalid(const char *word) const;

    std::string text_;           // whole input file
    char eol_ = '\n';
    unsigned long long ref_bytes_ = 0, cor_bytes_ = 0, pts_bytes_ = 0;
};

// Line terminator of a text file: '\n' unless '\r' splits it into more lines
// (InputRead.cpp:490-516).
char detect_eol(const std::string &text);
bool read_text(const std::string &fname, std::string &out);

// .disp file (InputRead.cpp:1131-1196): header, then one row per point.
std::string disp_header();
void disp_row(std::string &out, int label, const double p[3], int status, double objmin, const double uvw[3]);
bool write_text(const std::string &fname, const std::string &text);

// DispRead::read_disp_file_cst_sv (InputRead.cpp:799-860): returns false without an
// 'n' header.  DispRead::read_sort_file_cst_sv (InputRead.cpp:718-751).
bool read_disp_file(const std::string &fname, std::vector<int> &labels, std::vector<double> &points,
                    std::vector<int> &status, std::vector<double> &objmin, std::vector<double> &disp);
bool read_sort_file(const std::string &fname, std::vector<std::vector<int>> &neigh);
// DataCloud::write_sort_file: base + ".sort.csv", comma separated rows
bool write_sort_file(const std::string &base, const int32_t *neigh, long long n, int k);

// Strain CSV row (strain.cpp:595-656): label, x, y, z then the values, comma separated.
void csv_row(std::string &out, int label, const double p[3], const double *vals, int nvals, int npts_in_sw,
             double sw_radius, const double *tail, int ntail);

}  // namespace dvcio
