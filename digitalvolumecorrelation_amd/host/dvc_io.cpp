// dvc_io.cpp -- see dvc_io.h.
#include "dvc_io.h"

#include <cmath>
#include <cstdio>
#include <ctime>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>

namespace dvcio {

namespace {

// Keyword help lines of the parameters (InputRead.cpp:123-455): only the words and the
// limits the parser checks are needed here.
const std::vector<int> kBitDepth = {8, 16};
const std::vector<int> kSrchDof = {3, 6, 12};
const std::vector<std::string> kEndian = {"little", "big"};
const std::vector<std::string> kGeom = {"cube", "sphere"};
const std::vector<std::string> kOnOff = {"on", "off"};
const std::vector<std::string> kObj = {"sad", "ssd", "zssd", "nssd", "znssd"};
const std::vector<std::string> kInterp = {"nearest", "trilinear", "tricubic"};
constexpr int kHdrMax = 4096, kDimMax = 8000, kNptsMax = 50000;
constexpr double kFractMax = 1.0, kAspectMin = 0.1, kAspectMax = 10.0;

// Lines of `text` split at `eol`, each with "\n" appended (the reference appends its
// line terminator to every line it reads).
std::vector<std::string> split_lines(const std::string &text, char eol) {
    std::vector<std::string> out;
    std::string cur;
    std::istringstream in(text);
    while (std::getline(in, cur, eol)) out.push_back(cur + "\n");
    return out;
}

// i-th tab-separated non-empty field (the file-name keywords, InputRead.cpp:1295-1299);
// unlike the reference, a name that ends the line is accepted.
bool tab_field(const std::string &line, int want, std::string &out) {
    std::istringstream in(line);
    std::string f;
    int got = 0;
    while (std::getline(in, f, '\t')) {
        // the line terminator (and stray spaces) end the last field
        while (!f.empty() && (f.back() == '\n' || f.back() == '\r' || f.back() == ' ')) f.pop_back();
        if (f.empty()) continue;
        if (got == want) { out = f; return true; }
        got += 1;
    }
    return false;
}

}  // namespace

char detect_eol(const std::string &text) {
    long long n = 0, r = 0;
    {
        std::istringstream in(text);
        std::string l;
        while (std::getline(in, l, '\n')) n += 1;
    }
    {
        std::istringstream in(text);
        std::string l;
        while (std::getline(in, l, '\r')) r += 1;
    }
    return (r > n) ? '\r' : '\n';
}

bool read_text(const std::string &fname, std::string &out) {
    std::ifstream f(fname, std::ios::binary);
    if (!f.good()) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

bool write_text(const std::string &fname, const std::string &text) {
    std::ofstream f(fname, std::ios::binary | std::ios::out | std::ios::trunc);
    if (!f.is_open()) return false;
    f.write(text.data(), (std::streamsize)text.size());
    return f.good();
}

// ------------------------------------------------------------------ InputRead

int InputRead::input_file_accessible(const std::string &fname) {
    if (!read_text(fname, text_)) {
        std::cout << "\nCannot find input file '" << fname << "'\n\n";
        return 0;
    }
    eol_ = detect_eol(text_);
    return 1;
}

void InputRead::invalid(const char *word) const {
    std::cout << "\nInput Error: " << word << " contains an invalid parameter.\n\n";
}

// First line not starting with '#' that contains the keyword anywhere
// (InputRead.cpp:1250-1277).
int InputRead::get_line_with_keyword(const std::string &keyword, std::string &keyline, bool req) {
    for (const std::string &l : split_lines(text_, eol_)) {
        if (l[0] != '#' && l.find(keyword) != std::string::npos) {
            keyline = l;
            return 1;
        }
    }
    if (req) std::cout << "\n\nInput Error: keyword '" << keyword << "' not found in input file.\n\n";
    return 0;
}

int InputRead::parse_old_file(const char *word, std::string &arg, unsigned long long &bytes, bool req) {
    std::string line, name;
    if (!get_line_with_keyword(word, line, req)) return keywd_missing;
    if (!tab_field(line, 1, name)) { invalid(word); return param_invalid; }
    std::ifstream f(name, std::ios::binary);
    if (f.good()) {
        arg = name;
        f.seekg(0, std::ios::end);
        bytes = (unsigned long long)f.tellg();
        return input_line_ok;
    }
    std::cout << "\nInput Error: " << word << " cannot find file: " << name << "\n\n";
    return param_invalid;
}

int InputRead::parse_new_file(const char *word, std::string &arg, bool req) {
    std::string line, name;
    if (!get_line_with_keyword(word, line, req)) return keywd_missing;
    if (!tab_field(line, 1, name)) { invalid(word); return param_invalid; }
    // the name must be creatable (the reference opens and removes it)
    std::ifstream existing(name);
    bool existed = existing.good();
    existing.close();
    std::ofstream f(name, std::ios::app);
    if (f.is_open()) {
        f.close();
        if (!existed) std::remove(name.c_str());
        arg = name;
        return input_line_ok;
    }
    invalid(word);
    return param_invalid;
}

int InputRead::parse_int_in(const char *word, const std::vector<int> &vals, int &arg, bool req) {
    std::string line, w;
    if (!get_line_with_keyword(word, line, req)) return keywd_missing;
    std::istringstream in(line);
    double v = 0.0;
    in >> w >> v;
    for (int x : vals)
        if (v == x) { arg = x; return input_line_ok; }
    invalid(word);
    return param_invalid;
}

int InputRead::parse_str_in(const char *word, const std::vector<std::string> &vals, std::string &arg, bool req) {
    std::string line, w, v;
    if (!get_line_with_keyword(word, line, req)) return keywd_missing;
    std::istringstream in(line);
    in >> w >> v;
    for (const std::string &x : vals)
        if (v == x) { arg = v; return input_line_ok; }
    invalid(word);
    return param_invalid;
}

int InputRead::parse_int(const char *word, int lo, int hi, int &arg, bool req) {
    std::string line, w;
    if (!get_line_with_keyword(word, line, req)) return keywd_missing;
    std::istringstream in(line);
    int v = 0;
    in >> w >> v;
    if (v >= lo && v <= hi) { arg = v; return input_line_ok; }
    invalid(word);
    return param_invalid;
}

int InputRead::parse_uint(const char *word, unsigned lo, unsigned hi, unsigned &arg, bool req) {
    std::string line, w;
    if (!get_line_with_keyword(word, line, req)) return keywd_missing;
    std::istringstream in(line);
    unsigned v = 0;
    in >> w >> v;
    if (v >= lo && v <= hi) { arg = v; return input_line_ok; }
    invalid(word);
    return param_invalid;
}

int InputRead::parse_double(const char *word, double lo, double hi, double &arg, bool req) {
    std::string line, w;
    if (!get_line_with_keyword(word, line, req)) return keywd_missing;
    std::istringstream in(line);
    double v = 0.0;
    in >> w >> v;
    if (v >= lo && v <= hi) { arg = v; return input_line_ok; }
    invalid(word);
    return param_invalid;
}

// Vector keyword: each value within +-lim[i]; values before the first bad one are kept
// (InputRead.cpp:1530-1556).
int InputRead::parse_dvect(const char *word, const std::vector<double> &lim, std::vector<double> &vals, bool req) {
    std::string line, w;
    if (!get_line_with_keyword(word, line, req)) return keywd_missing;
    std::istringstream in(line);
    in >> w;
    size_t ok = 0;
    for (size_t i = 0; i < vals.size(); i++) {
        double v = 0.0;
        in >> v;
        if (std::fabs(v) > std::fabs(lim[i])) break;
        vals[i] = v;
        ok += 1;
    }
    if (ok == vals.size()) return input_line_ok;
    invalid(word);
    return param_invalid;
}

int InputRead::parse_dvect(const char *word, double lo, double hi, std::vector<double> &vals, bool req) {
    std::string line, w;
    if (!get_line_with_keyword(word, line, req)) return keywd_missing;
    std::istringstream in(line);
    in >> w;
    size_t ok = 0;
    for (size_t i = 0; i < vals.size(); i++) {
        double v = 0.0;
        in >> v;
        if (v < lo || v > hi) break;
        vals[i] = v;
        ok += 1;
    }
    if (ok == vals.size()) return input_line_ok;
    invalid(word);
    return param_invalid;
}

// InputRead::input_file_read (InputRead.cpp:518-630): required keywords, limits,
// defaults of the optional ones, and the volume file sizes.
int InputRead::input_file_read(RunControl *run) {
    if (parse_old_file("reference_filename", run->ref_fname, ref_bytes_, true) != input_line_ok) return 0;
    if (parse_old_file("correlate_filename", run->cor_fname, cor_bytes_, true) != input_line_ok) return 0;
    if (parse_old_file("point_cloud_filename", run->pts_fname, pts_bytes_, true) != input_line_ok) return 0;
    if (parse_new_file("output_filename", run->out_fname, true) != input_line_ok) return 0;
    run->res_fname = run->out_fname + ".disp";
    run->sta_fname = run->out_fname + ".stat";

    if (parse_int_in("vol_bit_depth", kBitDepth, run->vol_bit_depth, true) != input_line_ok) return 0;
    if (run->vol_bit_depth != 8)
        if (parse_str_in("vol_endian", kEndian, run->vol_endian, true) != input_line_ok) return 0;
    if (parse_int("vol_hdr_lngth", 0, kHdrMax, run->vol_hdr_lngth, true) != input_line_ok) return 0;
    if (parse_int("vol_wide", 0, kDimMax, run->vol_wide, true) != input_line_ok) return 0;
    if (parse_int("vol_high", 0, kDimMax, run->vol_high, true) != input_line_ok) return 0;
    if (parse_int("vol_tall", 0, kDimMax, run->vol_tall, true) != input_line_ok) return 0;

    int size_max = run->vol_wide;
    if (run->vol_high < size_max) size_max = run->vol_high;
    if (run->vol_tall < size_max) size_max = run->vol_tall;

    if (parse_str_in("subvol_geom", kGeom, run->subvol_geom, true) != input_line_ok) return 0;
    run->sub_geo = (run->subvol_geom == "sphere") ? sphere : cube;
    if (parse_int("subvol_size", 0, size_max, run->subvol_size, true) != input_line_ok) return 0;
    if (parse_int("subvol_npts", 0, kNptsMax, run->subvol_npts, true) != input_line_ok) return 0;
    if (parse_str_in("subvol_thresh", kOnOff, run->subvol_thresh, true) != input_line_ok) return 0;
    if (run->subvol_thresh == "on") {
        const double top = std::pow(2.0, (double)run->vol_bit_depth);
        if (parse_double("gray_thresh_min", 0.0, top, run->gray_thresh_min, true) != input_line_ok) return 0;
        if (parse_double("gray_thresh_max", run->gray_thresh_min, top, run->gray_thresh_max, true) != input_line_ok)
            return 0;
        if (parse_double("min_vol_fract", 0.0, kFractMax, run->min_vol_fract, true) != input_line_ok) return 0;
    }

    if (parse_int("disp_max", 0, size_max, run->disp_max, true) != input_line_ok) return 0;
    if (parse_int_in("num_srch_dof", kSrchDof, run->num_srch_dof, true) != input_line_ok) return 0;
    if (parse_str_in("obj_function", kObj, run->obj_function, true) != input_line_ok) return 0;
    for (size_t i = 0; i < kObj.size(); i++)
        if (run->obj_function == kObj[i]) run->obj_fcn = (Objfcn_Type)i;
    if (parse_str_in("interp_type", kInterp, run->interp_type, true) != input_line_ok) return 0;
    for (size_t i = 0; i < kInterp.size(); i++)
        if (run->interp_type == kInterp[i]) run->int_typ = (Interp_Type)i;

    run->basin_radius = 0.0;
    if (parse_double("basin_radius", 0.0, (double)run->disp_max, run->basin_radius, false) == param_invalid) return 0;
    run->rigid_trans.assign(3, 0.0);
    const std::vector<double> dims = {(double)run->vol_wide, (double)run->vol_high, (double)run->vol_tall};
    if (parse_dvect("rigid_trans", dims, run->rigid_trans, false) == param_invalid) return 0;
    run->subvol_aspect.assign(3, 1.0);
    if (parse_dvect("subvol_aspect", kAspectMin, kAspectMax, run->subvol_aspect, false) == param_invalid) return 0;
    run->starting_point.assign(3, std::nan(""));
    if (parse_dvect("starting_point", dims, run->starting_point, false) == param_invalid) return 0;
    if (parse_uint("num_points_to_process", 0u, std::numeric_limits<unsigned>::max(), run->num_points_to_process,
                   true) != input_line_ok)
        return 0;
    std::cout << "Max points to process is " << run->num_points_to_process << std::endl;

    const unsigned long long expect = (unsigned long long)run->vol_hdr_lngth +
                                      (unsigned long long)run->vol_wide * (unsigned long long)run->vol_high *
                                          (unsigned long long)run->vol_tall *
                                          (unsigned long long)(run->vol_bit_depth / 8);
    if (ref_bytes_ != expect) {
        std::cout << "\nReference volume file size does not match the volume description.\n"
                     "Check vol_bit_depth, vol_hdr_lngth, vol_high, vol_wide, and vol_tall.\n\n";
        return 0;
    }
    if (cor_bytes_ != expect) {
        std::cout << "\nCorrelate volume file size does not match the volume description.\n"
                     "Check vol_bit_depth, vol_hdr_lngth, vol_high, vol_wide, and vol_tall.\n\n";
        return 0;
    }
    return 1;
}

// InputRead::read_point_cloud (InputRead.cpp:632-716): an integer label followed by
// three numbers; lines that do not parse (headers) are skipped, and a label that reads
// as the start of a real number is rejected.
int InputRead::read_point_cloud(const RunControl *run, std::vector<double> &points, std::vector<int> &labels) {
    std::string text;
    if (!read_text(run->pts_fname, text)) {
        std::cout << "\nCannot read point cloud file " << run->pts_fname << "\n";
        return 0;
    }
    const char eol = detect_eol(text);
    double mn[3] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                    std::numeric_limits<double>::max()};
    double mx[3] = {std::numeric_limits<double>::min(), std::numeric_limits<double>::min(),
                    std::numeric_limits<double>::min()};
    for (const std::string &line : split_lines(text, eol)) {
        std::istringstream ss(line);
        int n = 0;
        double p[3];
        ss >> n;
        if (ss.fail()) continue;
        int c = ss.peek();
        if (c == '.' || c == 'E') continue;
        if (!(ss >> p[0]) || !(ss >> p[1]) || !(ss >> p[2])) continue;
        labels.push_back(n);
        for (int q = 0; q < 3; q++) {
            points.push_back(p[q]);
            if (p[q] < mn[q]) mn[q] = p[q];
            if (p[q] > mx[q]) mx[q] = p[q];
        }
    }
    if (labels.empty()) {
        std::cout << "\nNo points were read, the Point Cloud file may be improperly formatted.\n"
                     "The expected format is plain text, tab (or other 'white space') delimited.\n"
                     "Header info is OK as long as it does not match the format of a point description.\n"
                     "Each line of the file should contain an integer point label and x,y,z coordinates, e.g.:\n\n"
                     "5\t10.72\t15.87\t23.45\n\n\n";
        return 0;
    }
    for (int q = 0; q < 3; q++) { box_min[q] = mn[q]; box_max[q] = mx[q]; }
    num_pts = (long long)labels.size();
    return 1;
}

int InputRead::echo_input(const RunControl *run, const std::string &version) const {
    std::ofstream f(run->sta_fname, std::ios::out);
    f << "\n### echo of the input file for this run\n\n";
    f << "reference_filename\t" << run->ref_fname << "\n";
    f << "correlate_filename\t" << run->cor_fname << "\n";
    f << "point_cloud_filename\t" << run->pts_fname << "\n";
    f << "output_filename\t" << run->out_fname << "\n\n";
    f << "vol_bit_depth\t" << run->vol_bit_depth << "\n";
    if (run->vol_bit_depth != 8) f << "vol_endian\t" << run->vol_endian << "\n";
    f << "vol_hdr_lngth\t" << run->vol_hdr_lngth << "\n";
    f << "vol_wide\t" << run->vol_wide << "\n";
    f << "vol_high\t" << run->vol_high << "\n";
    f << "vol_tall\t" << run->vol_tall << "\n\n";
    f << "subvol_geom\t" << run->subvol_geom << "\n";
    f << "subvol_size\t" << run->subvol_size << "\n";
    f << "subvol_npts\t" << run->subvol_npts << "\n\n";
    f << "subvol_thresh\t" << run->subvol_thresh << "\n";
    if (run->subvol_thresh == "on") {
        f << "gray_thresh_min\t" << run->gray_thresh_min << "\n";
        f << "gray_thresh_max\t" << run->gray_thresh_max << "\n";
        f << "min_vol_fract\t" << run->min_vol_fract << "\n";
    }
    f << "\n";
    f << "disp_max\t" << run->disp_max << "\n";
    f << "num_srch_dof\t" << run->num_srch_dof << "\n";
    f << "obj_function\t" << run->obj_function << "\n";
    f << "interp_type\t" << run->interp_type << "\n\n";
    f << "rigid_trans\t" << run->rigid_trans[0] << "\t" << run->rigid_trans[1] << "\t" << run->rigid_trans[2] << "\n";
    f << "basin_radius\t" << run->basin_radius << "\n";
    f << "subvol_aspect\t" << run->subvol_aspect[0] << "\t" << run->subvol_aspect[1] << "\t" << run->subvol_aspect[2]
      << "\n";
    f << "\n### end of input file echo\n";
    f << "\nPoint Cloud contains " << num_pts << " points\n\n";
    f << "\tbounding box min = [" << box_min[0] << " " << box_min[1] << " " << box_min[2] << "]\n";
    f << "\tbounding box max = [" << box_max[0] << " " << box_max[1] << " " << box_max[2] << "]\n";
    f << "\nrunning under dvc code version: " << version << "\n\n";
    return 1;
}

int InputRead::append_time_date(const std::string &fname, const std::string &label, long long t) const {
    std::ofstream f(fname, std::ios::out | std::ios::app);
    std::time_t tt = (std::time_t)t;
    char ss[30];
    std::strftime(ss, sizeof(ss), "%Y-%m-%d %H:%M:%S", std::localtime(&tt));
    f << label << ss << "\n";
    return 1;
}

// ------------------------------------------------------------- result files

std::string disp_header() { return "n\tx\ty\tz\tstatus\tobjmin\tu\tv\tw\n"; }

// InputRead::append_result (InputRead.cpp:1165-1196): label, position and objmin in the
// stream's default format, displacements fixed with 6 decimals.
void disp_row(std::string &out, int label, const double p[3], int status, double objmin, const double uvw[3]) {
    std::ostringstream f;
    f << label << "\t" << p[0] << "\t" << p[1] << "\t" << p[2] << "\t" << status << "\t" << objmin;
    f << std::fixed << std::setprecision(6);
    for (int i = 0; i < 3; i++) f << "\t" << uvw[i];
    f << "\n";
    out += f.str();
}

namespace {
bool get_num(std::stringstream &ss, double &v) {
    if (ss >> v) {
        if (ss.peek() == ',') ss.ignore();
        if (ss.peek() == ' ') ss.ignore();
        if (ss.peek() == '\t') ss.ignore();
        return true;
    }
    return false;
}
bool get_num(std::stringstream &ss, int &v) {
    if (ss >> v) {
        if (ss.peek() == ',') ss.ignore();
        if (ss.peek() == ' ') ss.ignore();
        if (ss.peek() == '\t') ss.ignore();
        return true;
    }
    return false;
}
}  // namespace

bool read_disp_file(const std::string &fname, std::vector<int> &labels, std::vector<double> &points,
                    std::vector<int> &status, std::vector<double> &objmin, std::vector<double> &disp) {
    std::ifstream in(fname);
    std::string line;
    long long count = 0;
    while (std::getline(in, line)) {
        std::stringstream ss(line);
        if (count == 0 && ss.peek() != 'n') {
            std::cout << "-> .disp file lacks proper header, check command line arguments" << std::endl;
            return false;
        }
        count += 1;
        int n = 0, st = 0, cols = 0;
        double p[3], om = 0.0, d[3];
        if (get_num(ss, n)) cols += 1;
        if (get_num(ss, p[0]) && get_num(ss, p[1]) && get_num(ss, p[2])) cols += 3;
        if (get_num(ss, st)) cols += 1;
        if (get_num(ss, om)) cols += 1;
        if (get_num(ss, d[0]) && get_num(ss, d[1]) && get_num(ss, d[2])) cols += 3;
        if (cols == 9) {
            labels.push_back(n);
            status.push_back(st);
            objmin.push_back(om);
            for (int q = 0; q < 3; q++) { points.push_back(p[q]); disp.push_back(d[q]); }
        }
    }
    return true;
}

bool read_sort_file(const std::string &fname, std::vector<std::vector<int>> &neigh) {
    std::ifstream in(fname);
    if (!in.good()) return false;
    std::string line;
    while (std::getline(in, line)) {
        std::stringstream ss(line);
        std::vector<int> row;
        int v;
        while (get_num(ss, v)) row.push_back(v);
        if (!row.empty()) neigh.push_back(row);
    }
    return true;
}

bool write_sort_file(const std::string &base, const int32_t *neigh, long long n, int k) {
    std::string out;
    out.reserve((size_t)n * (size_t)k * 6);
    char buf[16];
    for (long long i = 0; i < n; i++) {
        for (int j = 0; j < k; j++) {
            if (j > 0) out += ',';
            int len = std::snprintf(buf, sizeof(buf), "%d", neigh[i * k + j]);
            out.append(buf, (size_t)len);
        }
        out += '\n';
    }
    return write_text(base + ".sort.csv", out);
}

void csv_row(std::string &out, int label, const double p[3], const double *vals, int nvals, int npts_in_sw,
             double sw_radius, const double *tail, int ntail) {
    std::ostringstream f;
    f << label << "," << p[0] << "," << p[1] << "," << p[2];
    for (int i = 0; i < nvals; i++) f << "," << vals[i];
    f << "," << npts_in_sw << "," << sw_radius;
    for (int i = 0; i < ntail; i++) f << "," << tail[i];
    f << "\n";
    out += f.str();
}

}  // namespace dvcio
