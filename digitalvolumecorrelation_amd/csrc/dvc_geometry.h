// dvc_geometry.h -- host-side FP64 run geometry shared by the C ABI (dvc_api.cpp) and
// the CPU build of the kernels used by the tests.  No HIP dependency.
//
// Mirrors Search::Search (Core/Search.cpp:24-82), Interpolate::Interpolate
// (Core/Interpolate.cpp:34-58) and the sphere cloud of FloatingCloud
// (Core/FloatingCloud.cpp:40-82).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <utility>
#include <vector>

#include "dvc_internal.h"

namespace dvc {

// Point truncation of Core/Point.cpp:59-61.
inline int trunc_i(double v) { return (int)(v / 1); }

// BoundBox arithmetic (Core/BoundBox.cpp:33-85).
struct HBox {
    double mn[3], mx[3], cn[3];
    HBox(const double a[3], const double b[3]) {
        for (int q = 0; q < 3; q++) { mn[q] = a[q]; mx[q] = b[q]; cn[q] = (a[q] + b[q]) / 2.0; }
    }
    void grow_by(double f) {
        for (int q = 0; q < 3; q++) { mn[q] = mn[q] + -f; mx[q] = mx[q] + f; }
        for (int q = 0; q < 3; q++) cn[q] = (mn[q] + mx[q]) / 2.0;
    }
    void center_on(const double c[3]) {
        for (int q = 0; q < 3; q++) {
            double d = c[q] - cn[q];
            mn[q] = mn[q] + d; mx[q] = mx[q] + d; cn[q] = cn[q] + d;
        }
    }
};

struct Geometry {
    double subv_rad;
    int subv_num, g;
    double est_mn[3], est_mx[3], est_cn[3];
    double act_mn[3], act_mx[3], act_cn[3];
    int box[3];
};

inline Geometry make_geometry(const dvc_run_params &rc) {
    Geometry G;
    G.subv_rad = rc.subvol_size / 2.0;                                        // Search.cpp:29
    if (rc.geom == DVC_GEOM_CUBE) G.subv_num = (int)pow(ceil(pow((double)rc.subvol_npts, 1.0 / 3.0)), 3.0);
    else G.subv_num = rc.subvol_npts;                                         // Search.cpp:30-31
    G.g = (int)ceil(pow((double)G.subv_num, 1.0 / 3.0));                      // Search.cpp:117
    double z[3] = {0.0, 0.0, 0.0};
    double m[3] = {2 * G.subv_rad * rc.aspect[0], 2 * G.subv_rad * rc.aspect[1], 2 * G.subv_rad * rc.aspect[2]};
    HBox est(z, m);                                                           // Search.cpp:66-76
    est.grow_by(rc.disp_max);
    est.grow_by(2.0);
    HBox e2(est.mn, est.mx);                                                  // Interpolate.cpp:51
    HBox act(est.mn, est.mx);                                                 // Interpolate.cpp:52-55
    act.grow_by(-1.0);
    for (int q = 0; q < 3; q++) {
        G.est_mn[q] = e2.mn[q]; G.est_mx[q] = e2.mx[q]; G.est_cn[q] = e2.cn[q];
        G.act_mn[q] = act.mn[q]; G.act_mx[q] = act.mx[q]; G.act_cn[q] = act.cn[q];
        G.box[q] = trunc_i(e2.mx[q]) - trunc_i(e2.mn[q]);                     // kern_4d extent, Interpolate.cpp:58
    }
    return G;
}

// Sphere sample offsets aspect*rel of one seed (FloatingCloud.cpp:51-78): the same
// libstdc++ mt19937 and uniform_real_distribution<double> stream, rejection in the
// unit-aspect ball, then the aspect scaling.
inline void sphere_offsets(const dvc_run_params &rc, const Geometry &G, double *sph) {
    std::mt19937 mt(rc.sphere_seed);
    std::uniform_real_distribution<double> dist(-G.subv_rad, G.subv_rad);
    int count = 0;
    while (count < G.subv_num) {
        double rx = dist(mt), ry = dist(mt), rz = dist(mt);
        double radius = std::sqrt(rx * rx + ry * ry + rz * rz);
        if (radius <= G.subv_rad) {
            sph[3 * count + 0] = rc.aspect[0] * (rx);
            sph[3 * count + 1] = rc.aspect[1] * (ry);
            sph[3 * count + 2] = rc.aspect[2] * (rz);
            count += 1;
        }
    }
}

inline void sphere_template(const dvc_run_params &rc, const Geometry &G, std::vector<double> &sph,
                            double omn[3], double omx[3]) {
    sph.assign((size_t)3 * G.subv_num, 0.0);
    sphere_offsets(rc, G, sph.data());
    for (int q = 0; q < 3; q++) { omn[q] = 1e300; omx[q] = -1e300; }
    for (int i = 0; i < G.subv_num; i++)
        for (int q = 0; q < 3; q++) {
            omn[q] = std::min(omn[q], sph[3 * i + q]);
            omx[q] = std::max(omx[q], sph[3 * i + q]);
        }
}

// The stable cloud of one point (FloatingCloud.cpp:40-128, as Search.cpp:99-118 builds
// it): xyz[3 * subv_num], its bounding box box[6] (min xyz, max xyz) grown point by point
// from the empty box of Cloud's constructor, and its reference point.  Spheres take the
// offsets sph of sphere_offsets; cubes the g^3 grid over the aspect-scaled cube.
inline void stable_cloud(const dvc_run_params &rc, const Geometry &G, const double *sph, const double pt[3],
                         double *xyz, double box[6], double ref[3]) {
#pragma clang fp contract(off)
    const int n = G.subv_num;
    if (rc.geom == DVC_GEOM_CUBE) {
        double mn[3], mx[3], inc[3];
        for (int q = 0; q < 3; q++) {
            mn[q] = pt[q] + -G.subv_rad * rc.aspect[q];
            mx[q] = pt[q] + G.subv_rad * rc.aspect[q];
            ref[q] = (mx[q] + mn[q]) / 2.0;
            inc[q] = (mx[q] - mn[q]) / (G.g - 1);
        }
        int j = 0;
        for (int i = 0; i < G.g; i++)
            for (int k = 0; k < G.g; k++)
                for (int l = 0; l < G.g; l++, j++) {
                    xyz[3 * j + 0] = mn[0] + l * inc[0];
                    xyz[3 * j + 1] = mn[1] + k * inc[1];
                    xyz[3 * j + 2] = mn[2] + i * inc[2];
                }
    } else {
        for (int q = 0; q < 3; q++) ref[q] = pt[q];
        for (int j = 0; j < n; j++)
            for (int q = 0; q < 3; q++) xyz[3 * j + q] = pt[q] + sph[3 * j + q];
    }
    for (int q = 0; q < 3; q++) { box[q] = 9999.99; box[3 + q] = -9999.99; }
    for (int j = 0; j < n; j++) {
        double lo[3], hi[3];
        bool reset = false;
        for (int q = 0; q < 3; q++) {
            lo[q] = box[q]; hi[q] = box[3 + q];
            if (xyz[3 * j + q] < lo[q]) { lo[q] = xyz[3 * j + q]; reset = true; }
        }
        for (int q = 0; q < 3; q++)
            if (xyz[3 * j + q] > hi[q]) { hi[q] = xyz[3 * j + q]; reset = true; }
        if (reset)
            for (int q = 0; q < 3; q++) { box[q] = lo[q]; box[3 + q] = hi[q]; }
    }
}

// FloatingCloud::affine_to(x, nx) (FloatingCloud.cpp:236-286) with
// SearchParams::set_param_vect (SearchParams.cpp:53-107) on a fresh parameter vector:
// translate the stable cloud, then (nx > 3) rotate about the moved reference point,
// (nx > 6) strain about it, and take the moved cloud's box.
inline void affine_cloud(const double *stable, int n, const double sbox[6], const double sref[3], const double *x,
                         int nx, double *moving, double mbox[6]) {
#pragma clang fp contract(off)
    double v[12] = {0.0};
    for (int q = 0; q < nx && q < 12; q++) v[q] = x[q];
    double A[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (nx > 3) {
        const double phi = v[3], the = v[4], psi = v[5];
        A[0] = cos(the) * cos(phi);
        A[1] = cos(the) * sin(phi);
        A[2] = -sin(the);
        A[3] = sin(psi) * sin(the) * cos(phi) - cos(psi) * sin(phi);
        A[4] = sin(psi) * sin(the) * sin(phi) + cos(psi) * cos(phi);
        A[5] = cos(the) * sin(psi);
        A[6] = cos(psi) * sin(the) * cos(phi) + sin(psi) * sin(phi);
        A[7] = cos(psi) * sin(the) * sin(phi) - sin(psi) * cos(phi);
        A[8] = cos(the) * cos(psi);
    }
    if (nx > 6) {
        S[0] = v[6];  S[1] = v[9];  S[2] = v[11];
        S[3] = v[9];  S[4] = v[7];  S[5] = v[10];
        S[6] = v[11]; S[7] = v[10]; S[8] = v[8];
    }
    double mref[3];
    for (int q = 0; q < 3; q++) {
        mref[q] = sref[q] + v[q];
        mbox[q] = sbox[q] + v[q];
        mbox[3 + q] = sbox[3 + q] + v[q];
    }
    for (int j = 0; j < n; j++) {
        double p[3];
        for (int q = 0; q < 3; q++) p[q] = stable[3 * j + q] + v[q];
        if (nx > 3) {
            double r[3], o[3];
            for (int q = 0; q < 3; q++) r[q] = p[q] - mref[q];
            for (int q = 0; q < 3; q++) o[q] = A[3 * q] * r[0] + A[3 * q + 1] * r[1] + A[3 * q + 2] * r[2];
            for (int q = 0; q < 3; q++) p[q] = p[q] + (o[q] - r[q]);
            if (nx > 6) {
                for (int q = 0; q < 3; q++) r[q] = p[q] - mref[q];
                for (int q = 0; q < 3; q++) o[q] = S[3 * q] * r[0] + S[3 * q + 1] * r[1] + S[3 * q + 2] * r[2];
                for (int q = 0; q < 3; q++) p[q] = p[q] + o[q];
            }
        }
        for (int q = 0; q < 3; q++) moving[3 * j + q] = p[q];
    }
    if (nx > 3) {
        double lo[3] = {9999.9, 9999.9, 9999.9}, hi[3] = {-9999.9, -9999.9, -9999.9};
        bool reset = false;
        for (int j = 0; j < n; j++)
            for (int q = 0; q < 3; q++) {
                if (moving[3 * j + q] < lo[q]) { lo[q] = moving[3 * j + q]; reset = true; }
                if (moving[3 * j + q] > hi[q]) { hi[q] = moving[3 * j + q]; reset = true; }
            }
        if (reset)
            for (int q = 0; q < 3; q++) { mbox[q] = lo[q]; mbox[3 + q] = hi[q]; }
    }
}

// Reorder sample offsets by their unit cells, z slowest and x fastest.  A sample's
// value does not depend on its position in the cloud, only the order of the objective's
// sums does.  In cell order the 64 lanes of a wavefront fall on a few rows (y, z) of
// the volume, so each of a sample's 16 row loads touches a few cache lines for the
// whole wavefront instead of one line per lane; the L1 tag lookups of these gathers are
// what bounds the sample pass.  Stable: samples of one cell keep their relative order.
inline void locality_order(std::vector<double> &sph) {
    const size_t n = sph.size() / 3;
    if (n == 0) return;
    double lo[3] = {1e300, 1e300, 1e300};
    for (size_t i = 0; i < n; i++)
        for (int q = 0; q < 3; q++) lo[q] = std::min(lo[q], sph[3 * i + q]);
    std::vector<std::pair<uint64_t, size_t>> key(n);
    for (size_t i = 0; i < n; i++) {
        uint64_t c[3];
        for (int q = 0; q < 3; q++) {
            double f = std::floor(sph[3 * i + q] - lo[q]);
            c[q] = (uint64_t)std::min(std::max(f, 0.0), 2097151.0);
        }
        key[i] = {(c[2] << 42) | (c[1] << 21) | c[0], i};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint64_t, size_t> &a, const std::pair<uint64_t, size_t> &b) {
                         return a.first < b.first;
                     });
    std::vector<double> out(sph.size());
    for (size_t i = 0; i < n; i++)
        for (int q = 0; q < 3; q++) out[3 * i + q] = sph[3 * key[i].second + q];
    sph.swap(out);
}

// Kernel arguments of one search (pointers filled by the caller).
inline void fill_search_args(SearchArgs &a, const dvc_run_params &rc, const Geometry &G, int bpv, int W, int H,
                             int T, int npts, int nproc, int k, const double omn[3], const double omx[3]) {
    std::memset(&a, 0, sizeof(a));
    a.W = W; a.H = H; a.T = T; a.bpv = bpv;
    a.ndof = rc.ndof; a.obj = rc.obj; a.interp = rc.interp; a.geom = rc.geom;
    a.nsamp = G.subv_num; a.g = G.g;
    a.npts = npts; a.nproc = nproc; a.k = k;
    a.nsamp_pad = (G.subv_num + 3) & ~3;
    for (int q = 0; q < 3; q++) {
        a.box[q] = G.box[q];
        a.est_mn[q] = G.est_mn[q]; a.est_mx[q] = G.est_mx[q]; a.est_cn[q] = G.est_cn[q];
        a.act_mn[q] = G.act_mn[q]; a.act_mx[q] = G.act_mx[q]; a.act_cn[q] = G.act_cn[q];
        a.aspect[q] = rc.aspect[q]; a.rigid[q] = rc.rigid_trans[q];
        a.off_min[q] = omn[q]; a.off_max[q] = omx[q];
    }
    a.subv_rad = G.subv_rad; a.basin = rc.basin_radius; a.disp_max = (double)rc.disp_max;
    // 32-bit row offsets: whole-dword row strides and the box's slices within 2 GiB
    const long long wb = (long long)W * bpv, whb = wb * (long long)H;
    a.gather32 = (wb % 4 == 0) && ((long long)(G.box[2] + 1) * whb + 64 < (1ll << 31)) ? 1 : 0;
}

// Slice launches of the serial chain (SliceArgs): which chain launches split their points'
// samples over several workgroups.  A launch L of cnt points is sliced when its cnt8 *
// nslice[L] workgroups (cnt8: cnt rounded up to a multiple of xcds) are at most max_wgs
// and there are at least two slices; narrow launches leave most of the chip idle with one
// workgroup per point.  sliced[L] marks launch L; returns the largest cnt8 among them
// (0: none).
inline int slice_plan(const std::vector<int> &level_off, const std::vector<int> &nslice, int max_wgs, int xcds,
                      std::vector<char> &sliced) {
    const int nl = (int)level_off.size() - 1;
    sliced.assign(nl > 0 ? nl : 0, 0);
    int widest = 0;
    if (max_wgs <= 0) return 0;
    for (int L = 0; L < nl; L++) {
        const int cnt = level_off[L + 1] - level_off[L];
        const int cnt8 = (cnt + xcds - 1) / xcds * xcds;
        if (cnt > 0 && nslice[L] >= 2 && (long long)cnt8 * nslice[L] <= max_wgs) {
            sliced[L] = 1;
            widest = std::max(widest, cnt8);
        }
    }
    return widest;
}

// Position of every point in order[] (npts when it is never searched, dvc.cpp:229-232).
inline bool order_positions(const int *order, int npts, int N, std::vector<int> &pos) {
    pos.assign(npts, npts);
    for (int i = 0; i < npts; i++) {
        int o = order[i];
        if (o < 0 || o >= npts || pos[o] != npts) return false;
        pos[o] = i;
    }
    for (int i = 0; i < npts; i++)
        if (pos[i] >= N) pos[i] = npts;
    return true;
}

// Dependency levels of the serial point loop (dvc.cpp:233-303 with the start guess of
// Search.cpp:277-377): the point at position i of order[] reads the results of its
// neighbours at earlier positions, so its level is one more than the highest level
// among them.  Points of one level are independent; running the levels in order
// reproduces the serial loop exactly.  level_pts lists the searched points level by
// level (ascending position inside a level), level_off[L]..level_off[L+1] indexes it.
inline int dependency_levels(const int *order, const std::vector<int> &pos, const int *neigh, int npts, int k, int N,
                             std::vector<int> &level_pts, std::vector<int> &level_off) {
    std::vector<int> lv(N, 0);
    int depth = 0;
    for (int i = 0; i < N; i++) {
        int n = order[i], l = 0;
        for (int q = 0; q < k; q++) {
            int m = neigh[(size_t)n * k + q];
            int pm = pos[m];
            if (pm < i && lv[pm] + 1 > l) l = lv[pm] + 1;
        }
        lv[i] = l;
        if (l + 1 > depth) depth = l + 1;
    }
    level_off.assign(depth + 1, 0);
    for (int i = 0; i < N; i++) level_off[lv[i] + 1] += 1;
    for (int L = 0; L < depth; L++) level_off[L + 1] += level_off[L];
    std::vector<int> fill(level_off.begin(), level_off.end() - 1);
    level_pts.assign(N, 0);
    for (int i = 0; i < N; i++) level_pts[fill[lv[i]]++] = order[i];
    return depth;
}

// Launch schedule of the serial chain for a chip that runs `cap` points at once.  In
// processing order, each point goes into the first launch after every launch that holds
// one of its earlier neighbours and that still has room (first fit).  As with
// dependency levels, a point's start guess reads only results of earlier launches, so
// the results are the same.  A level wider than `cap` then needs no partly empty launch
// of its own: later points whose neighbours are done fill it.  Same output format as
// dependency_levels; returns the number of launches.
inline int capacity_schedule(const int *order, const std::vector<int> &pos, const int *neigh, int k, int N, int cap,
                             std::vector<int> &level_pts, std::vector<int> &level_off) {
    std::vector<int> at(N, 0);          // launch of the point at position i
    std::vector<int> cnt;               // points per launch
    std::vector<int> nxt;               // union-find: first launch >= t with room
    auto find = [&](int t) {
        int r = t;
        while (r < (int)nxt.size() && nxt[r] != r) r = nxt[r];
        while (t < (int)nxt.size() && nxt[t] != t) { const int u = nxt[t]; nxt[t] = r; t = u; }
        return r;
    };
    for (int i = 0; i < N; i++) {
        const int n = order[i];
        int e = 0;
        for (int q = 0; q < k; q++) {
            const int m = neigh[(size_t)n * k + q];
            const int pm = pos[m];
            if (pm < i && at[pm] + 1 > e) e = at[pm] + 1;
        }
        while ((int)cnt.size() <= e) { nxt.push_back((int)cnt.size()); cnt.push_back(0); }
        const int t = find(e);
        if (t == (int)cnt.size()) { nxt.push_back(t); cnt.push_back(0); }
        at[i] = t;
        if (++cnt[t] == cap) nxt[t] = t + 1;
    }
    const int depth = (int)cnt.size();
    level_off.assign(depth + 1, 0);
    for (int t = 0; t < depth; t++) level_off[t + 1] = level_off[t] + cnt[t];
    std::vector<int> fill(level_off.begin(), level_off.end() - 1);
    level_pts.assign(N, 0);
    for (int i = 0; i < N; i++) level_pts[fill[at[i]]++] = order[i];
    return depth;
}

// Workgroup order of the reference kernel: block b computes processing position
// out[b].  The positions are sorted by the Morton code of their point's 16-voxel cell,
// and the chips' L2 caches are per XCD, with workgroups dealt to the XCDs in turn
// (block b to XCD b % xcds); so XCD x gets one contiguous run of the spatial order, and
// points whose sub-volumes overlap share an L2.  Any order gives the same results: each
// workgroup fills only its own point's row of the table.
inline void spatial_block_order(const double *points, const int *order, int N, int xcds, std::vector<int> &out) {
    auto spread = [](uint64_t v) {                 // 21 bits -> every third bit
        v &= 0x1fffff;
        v = (v | v << 32) & 0x1f00000000ffffull;
        v = (v | v << 16) & 0x1f0000ff0000ffull;
        v = (v | v << 8) & 0x100f00f00f00f00full;
        v = (v | v << 4) & 0x10c30c30c30c30c3ull;
        v = (v | v << 2) & 0x1249249249249249ull;
        return v;
    };
    std::vector<std::pair<uint64_t, int>> key(N);
    for (int i = 0; i < N; i++) {
        const double *q = points + 3 * (size_t)order[i];
        uint64_t c[3];
        for (int d = 0; d < 3; d++) {
            const double v = std::isfinite(q[d]) ? q[d] / 16.0 : 0.0;
            c[d] = v <= 0.0 ? 0 : (v >= 2097151.0 ? 2097151 : (uint64_t)v);
        }
        key[i] = {spread(c[0]) | spread(c[1]) << 1 | spread(c[2]) << 2, i};
    }
    std::sort(key.begin(), key.end());
    out.assign(N, 0);
    size_t start = 0;
    for (int x = 0; x < xcds; x++) {
        const int cnt = x < N ? (N - x + xcds - 1) / xcds : 0;   // blocks x, x + xcds, ...
        for (int j = 0; j < cnt; j++) out[(size_t)x + (size_t)j * xcds] = key[start + j].second;
        start += cnt;
    }
}

}  // namespace dvc
