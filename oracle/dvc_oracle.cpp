/*
 * dvc_oracle.cpp -- FP64 CPU restatement of the CCPi DVC per-point correlation path.
 *
 * TEST INFRASTRUCTURE ONLY (see dvc_oracle.h).  Every function cites the reference
 * file:line it follows.  Floating-point expressions keep the reference's operation
 * order and are compiled with -ffp-contract=off so that no multiply-add is fused.
 *
 * Pinning (what checks each part; the tests live in tests/test_oracle_*.py):
 *   BI matrix            closed form: BI = B^-1 with B built from monomial derivatives
 *   tricubic             Catmull-Rom identity in the box interior, exact reproduction
 *                        of tri-quadratic fields, node values, clamped-face rule
 *   objectives           identities (ZNSSD offset/scale invariance, bounds) + brute sums
 *   clouds               cube grid by formula; sphere radius/count/seed determinism
 *   affine warp          closed-form composite x' = c+u+(I+S)A(x-c), rotation orthogonality
 *   neighbour sort       the known 10x10 answer of Core/tests/test_DataCloud.cpp:129-139
 *   Gauss-Newton         synthetic ground truth; analytic vs forward-difference Jacobian
 *   QR solvers           residual of random well-conditioned systems vs numpy.linalg
 *   strain               exact recovery of linear / quadratic displacement fields
 *   Eigen rounding       parity unpinned: Eigen is not present here, the solvers follow
 *                        Eigen 3.4's published algorithms, not its binary.
 */
#include "dvc_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <map>
#include <tuple>
#include <vector>
#include <omp.h>

namespace {

/* ------------------------------------------------------------------ geometry */

/* Point: value plus truncated integer part and remainder (Core/Point.cpp:57-66). */
struct Pt {
    double x, y, z;
    int ix, iy, iz;
    double rx, ry, rz;
    Pt() : x(0), y(0), z(0), ix(0), iy(0), iz(0), rx(0), ry(0), rz(0) {}
    Pt(double a, double b, double c) { move_to(a, b, c); }
    void move_to(double a, double b, double c) {
        x = a; y = b; z = c;
        ix = (int)(x / 1); iy = (int)(y / 1); iz = (int)(z / 1);
        rx = x - ix; ry = y - iy; rz = z - iz;
    }
    void move_by(double u, double v, double w) { move_to(x + u, y + v, z + w); }
};

/* pt_dist (Core/Point.h:47-48). */
static inline double pt_dist(const Pt &a, const Pt &b) {
    return std::sqrt(std::pow(a.x - b.x, 2) + std::pow(a.y - b.y, 2) + std::pow(a.z - b.z, 2));
}

/* BoundBox (Core/BoundBox.cpp). */
struct Box {
    Pt mn, mx, cn;
    Box() {}
    Box(const Pt &a, const Pt &b) {
        mn = Pt(a.x, a.y, a.z);
        mx = Pt(b.x, b.y, b.z);
        cn = Pt((a.x + b.x) / 2.0, (a.y + b.y) / 2.0, (a.z + b.z) / 2.0);   /* :37-43 */
    }
    void set_center() { cn.move_to((mn.x + mx.x) / 2.0, (mn.y + mx.y) / 2.0, (mn.z + mx.z) / 2.0); }
    void move_to(const Pt &a, const Pt &b) { mn.move_to(a.x, a.y, a.z); mx.move_to(b.x, b.y, b.z); set_center(); }
    void grow_by(double f) { mn.move_by(-f, -f, -f); mx.move_by(f, f, f); set_center(); }  /* :65-73 */
    void center_on(const Pt &c) {                                                         /* :75-85 */
        double dx = c.x - cn.x, dy = c.y - cn.y, dz = c.z - cn.z;
        mn.move_by(dx, dy, dz); mx.move_by(dx, dy, dz); cn.move_by(dx, dy, dz);
    }
    /* contains (:87-99): true when b lies inside, failing on strict < / >. */
    bool contains(const Box &b) const {
        return !(b.mn.x < mn.x || b.mn.y < mn.y || b.mn.z < mn.z ||
                 b.mx.x > mx.x || b.mx.y > mx.y || b.mx.z > mx.z);
    }
    int iwide() const { return mx.ix - mn.ix; }
    int ihigh() const { return mx.iy - mn.iy; }
    int itall() const { return mx.iz - mn.iz; }
};

/* Cloud with AddPoint box tracking (Core/Cloud.cpp:28-74). */
struct Cloud {
    std::vector<Pt> pts;
    Box box;
    Cloud() : box(Pt(9999.99, 9999.99, 9999.99), Pt(-9999.99, -9999.99, -9999.99)) {}
    void add(const Pt &p) {
        pts.push_back(p);
        double mnx = box.mn.x, mny = box.mn.y, mnz = box.mn.z;
        double mxx = box.mx.x, mxy = box.mx.y, mxz = box.mx.z;
        bool reset = false;
        if (p.x < mnx) { mnx = p.x; reset = true; }
        if (p.y < mny) { mny = p.y; reset = true; }
        if (p.z < mnz) { mnz = p.z; reset = true; }
        if (p.x > mxx) { mxx = p.x; reset = true; }
        if (p.y > mxy) { mxy = p.y; reset = true; }
        if (p.z > mxz) { mxz = p.z; reset = true; }
        if (reset) box.move_to(Pt(mnx, mny, mnz), Pt(mxx, mxy, mxz));
    }
};

/* --------------------------------------------------------- search parameters */

/* SearchParams::set_param_vect (Core/SearchParams.cpp:53-107). */
struct Params {
    double v[12];
    double A[3][3];
    double S[3][3];
    Params() {
        for (int i = 0; i < 12; i++) v[i] = 0.0;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) { A[i][j] = (i == j) ? 1 : 0; S[i][j] = 0.0; }
    }
    void set(const double *t, int tsize, int ndof) {
        for (int i = 0; i < tsize; i++) v[i] = t[i];
        if (ndof > 3) {
            double phi = v[3], the = v[4], psi = v[5];
            A[0][0] = cos(the) * cos(phi);
            A[0][1] = cos(the) * sin(phi);
            A[0][2] = -sin(the);
            A[1][0] = sin(psi) * sin(the) * cos(phi) - cos(psi) * sin(phi);
            A[1][1] = sin(psi) * sin(the) * sin(phi) + cos(psi) * cos(phi);
            A[1][2] = cos(the) * sin(psi);
            A[2][0] = cos(psi) * sin(the) * cos(phi) + sin(psi) * sin(phi);
            A[2][1] = cos(psi) * sin(the) * sin(phi) - sin(psi) * cos(phi);
            A[2][2] = cos(the) * cos(psi);
        }
        if (ndof > 6) {
            S[0][0] = v[6];  S[0][1] = v[9];  S[0][2] = v[11];
            S[1][0] = v[9];  S[1][1] = v[7];  S[1][2] = v[10];
            S[2][0] = v[11]; S[2][1] = v[10]; S[2][2] = v[8];
        }
    }
};

/* Derivatives of the rotation matrix with respect to phi, the, psi (for the analytic
 * Jacobian; differentiates SearchParams.cpp:77-87). */
static void rot_derivs(double phi, double the, double psi, double dA[3][3][3]) {
    double cf = cos(phi), sf = sin(phi), ct = cos(the), st = sin(the), cp = cos(psi), sp = sin(psi);
    double (*P)[3] = dA[0];
    P[0][0] = -ct * sf; P[0][1] = ct * cf; P[0][2] = 0;
    P[1][0] = -sp * st * sf - cp * cf; P[1][1] = sp * st * cf - cp * sf; P[1][2] = 0;
    P[2][0] = -cp * st * sf + sp * cf; P[2][1] = cp * st * cf + sp * sf; P[2][2] = 0;
    double (*T)[3] = dA[1];
    T[0][0] = -st * cf; T[0][1] = -st * sf; T[0][2] = -ct;
    T[1][0] = sp * ct * cf; T[1][1] = sp * ct * sf; T[1][2] = -st * sp;
    T[2][0] = cp * ct * cf; T[2][1] = cp * ct * sf; T[2][2] = -st * cp;
    double (*Q)[3] = dA[2];
    Q[0][0] = 0; Q[0][1] = 0; Q[0][2] = 0;
    Q[1][0] = cp * st * cf + sp * sf; Q[1][1] = cp * st * sf - sp * cf; Q[1][2] = ct * cp;
    Q[2][0] = -sp * st * cf + cp * sf; Q[2][1] = -sp * st * sf - cp * cf; Q[2][2] = -ct * sp;
}

/* ------------------------------------------------------------- FloatingCloud */

struct FCloud {
    Cloud stable, moving;
    Params params;
    Pt sref, mref;

    /* sphere constructor (Core/FloatingCloud.cpp:40-82) */
    void make_sphere(const Pt &cen, double rad_max, int num, double ax, double ay, double az, unsigned int seed) {
        sref = Pt(cen.x, cen.y, cen.z); mref = sref;
        std::mt19937 mt(seed);
        std::uniform_real_distribution<double> dist(-rad_max, rad_max);
        int count = 0;
        while (count < num) {
            double rx = dist(mt);
            double ry = dist(mt);
            double rz = dist(mt);
            double radius = std::sqrt(rx * rx + ry * ry + rz * rz);
            if (radius <= rad_max) {
                count += 1;
                Pt p(cen.x + ax * (rx), cen.y + ay * (ry), cen.z + az * (rz));
                stable.add(p); moving.add(p);
            }
        }
    }
    /* cube constructor (Core/FloatingCloud.cpp:84-128) */
    void make_cube(const Pt &bmin, const Pt &bmax, int nx, int ny, int nz) {
        double avx = (bmax.x + bmin.x) / 2.0, avy = (bmax.y + bmin.y) / 2.0, avz = (bmax.z + bmin.z) / 2.0;
        sref = Pt(avx, avy, avz); mref = sref;
        double incx = (bmax.x - bmin.x) / (nx - 1);
        double incy = (bmax.y - bmin.y) / (ny - 1);
        double incz = (bmax.z - bmin.z) / (nz - 1);
        for (int i = 0; i < nz; i++) {
            double pz = bmin.z + i * incz;
            for (int j = 0; j < ny; j++) {
                double py = bmin.y + j * incy;
                for (int k = 0; k < nx; k++) {
                    double px = bmin.x + k * incx;
                    Pt p(px, py, pz);
                    stable.add(p); moving.add(p);
                }
            }
        }
    }
    /* rotate_ptvect (:174-204) */
    void rotate() {
        const double (*A)[3] = params.A;
        for (size_t i = 0; i < moving.pts.size(); i++) {
            Pt &p = moving.pts[i];
            double rx = p.x - mref.x, ry = p.y - mref.y, rz = p.z - mref.z;
            double ox = A[0][0] * rx + A[0][1] * ry + A[0][2] * rz;
            double oy = A[1][0] * rx + A[1][1] * ry + A[1][2] * rz;
            double oz = A[2][0] * rx + A[2][1] * ry + A[2][2] * rz;
            p.move_by(ox - rx, oy - ry, oz - rz);
        }
    }
    /* strain_ptvect (:148-172) */
    void strain() {
        const double (*S)[3] = params.S;
        for (size_t i = 0; i < moving.pts.size(); i++) {
            Pt &p = moving.pts[i];
            double rx = p.x - mref.x, ry = p.y - mref.y, rz = p.z - mref.z;
            double sx = S[0][0] * rx + S[0][1] * ry + S[0][2] * rz;
            double sy = S[1][0] * rx + S[1][1] * ry + S[1][2] * rz;
            double sz = S[2][0] * rx + S[2][1] * ry + S[2][2] * rz;
            p.move_by(sx, sy, sz);
        }
    }
    /* resize_bbox (:206-233) */
    void resize_bbox() {
        double mnx = 9999.9, mny = 9999.9, mnz = 9999.9, mxx = -9999.9, mxy = -9999.9, mxz = -9999.9;
        bool reset = false;
        for (size_t i = 0; i < moving.pts.size(); i++) {
            const Pt &p = moving.pts[i];
            if (p.x < mnx) { mnx = p.x; reset = true; }
            if (p.y < mny) { mny = p.y; reset = true; }
            if (p.z < mnz) { mnz = p.z; reset = true; }
            if (p.x > mxx) { mxx = p.x; reset = true; }
            if (p.y > mxy) { mxy = p.y; reset = true; }
            if (p.z > mxz) { mxz = p.z; reset = true; }
        }
        if (reset) moving.box.move_to(Pt(mnx, mny, mnz), Pt(mxx, mxy, mxz));
    }
    /* affine_to (:236-286) */
    void affine_to(const double *t, int tsize, int ndof) {
        params.set(t, tsize, ndof);
        double dx = params.v[0], dy = params.v[1], dz = params.v[2];
        mref.move_to(sref.x + dx, sref.y + dy, sref.z + dz);
        moving.box.move_to(Pt(stable.box.mn.x + dx, stable.box.mn.y + dy, stable.box.mn.z + dz),
                           Pt(stable.box.mx.x + dx, stable.box.mx.y + dy, stable.box.mx.z + dz));
        for (size_t i = 0; i < moving.pts.size(); i++)
            moving.pts[i].move_to(stable.pts[i].x + dx, stable.pts[i].y + dy, stable.pts[i].z + dz);
        if (ndof > 3) {
            rotate();
            if (ndof > 6) strain();
            resize_bbox();
        }
    }
};

/* ------------------------------------------------------------ BI (Lekien) */

/* Inverse of the 1D cubic Hermite constraint matrix: rows are the monomial
 * coefficients a0..a3 of p(t), columns the data (f0, f1, f'0, f'1). */
static const int H1[4][4] = {
    { 1,  0,  0,  0},
    { 0,  0,  1,  0},
    {-3,  3, -2, -1},
    { 2, -2,  1,  1}};

/* Quantity order of Matrix_4d (Core/Matrix_4d.h:105-106): f, fx, fy, fz, fxy, fxz, fyz, fxyz
 * as (dx,dy,dz) derivative orders; corner order (Core/Matrix_4d.cpp:68-76). */
static const int QD[8][3] = {{0,0,0},{1,0,0},{0,1,0},{0,0,1},{1,1,0},{1,0,1},{0,1,1},{1,1,1}};
static const int CO[8][3] = {{0,0,0},{1,0,0},{0,1,0},{1,1,0},{0,0,1},{1,0,1},{0,1,1},{1,1,1}};

static double g_BI[64][64];
static bool g_BI_ready = false;

/* BI (Interpolate.cpp:61-127) is the tricubic Lekien-Marsden inverse: a Kronecker
 * product of H1 over the three axes, rows = coefficient i+4j+16k, columns = data
 * index quantity*8 + corner. */
static void build_BI() {
    if (g_BI_ready) return;
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) for (int k = 0; k < 4; k++) {
        int row = i + 4 * j + 16 * k;
        for (int q = 0; q < 8; q++) for (int c = 0; c < 8; c++) {
            /* 1D data index: value at 0 -> 0, value at 1 -> 1, slope at 0 -> 2, slope at 1 -> 3 */
            int dx = QD[q][0] * 2 + CO[c][0];
            int dy = QD[q][1] * 2 + CO[c][1];
            int dz = QD[q][2] * 2 + CO[c][2];
            g_BI[row][q * 8 + c] = (double)(H1[i][dx] * H1[j][dy] * H1[k][dz]);
        }
    }
    g_BI_ready = true;
}

/* sparse column storage of BI (Interpolate.cpp:129-135) */
struct SparseBI {
    std::vector<int> colptr, row;
    std::vector<double> val;
    SparseBI() {
        build_BI();
        colptr.push_back(0);
        for (int j = 0; j < 64; j++) {
            for (int i = 0; i < 64; i++) if (g_BI[i][j] != 0) { row.push_back(i); val.push_back(g_BI[i][j]); }
            colptr.push_back((int)row.size());
        }
    }
};

/* ---------------------------------------------------------------- volume */

struct Vol {
    const void *data;
    int bpv;
    long W, H, T;
    double at(long x, long y, long z) const {
        size_t idx = (size_t)((z * H + y) * W + x);
        if (bpv == 1) return (double)((const uint8_t *)data)[idx];
        return (double)((const uint16_t *)data)[idx];
    }
};

/* ------------------------------------------------------------- Interpolate */

/* Interpolate with the Matrix_4d storage (72 doubles per voxel, Matrix_4d.cpp:25-57).
 * Box extent is the allocated size; the reference loops over the current box size,
 * which equals the allocation for every box that fits (the two differ only in the
 * undefined cases, which are reported as range failures here). */
struct Interp {
    Box est, act;
    int nx, ny, nz, nk;
    int xh, yh, zh;
    std::vector<double> m;
    std::vector<unsigned char> avail;
    const SparseBI *sbi;
    bool box_ok;
    int nthreads;

    Interp(const Box &region, const SparseBI *s, int nthr) : sbi(s), box_ok(false), nthreads(nthr) {
        est = Box(region.mn, region.mx);                    /* Interpolate.cpp:51-55 */
        act = Box(region.mn, region.mx);
        act.grow_by(-1.0);
        nx = est.iwide(); ny = est.ihigh(); nz = est.itall(); nk = 72;
        xh = ny * nz * nk; yh = nz * nk; zh = nk;
        m.assign((size_t)nx * ny * nz * nk, 0.0);
        avail.assign((size_t)nx * ny * nz, 0);
    }
    inline size_t c0(int x, int y, int z) const { return (size_t)x * xh + (size_t)y * yh + (size_t)z * zh; }
    void center_on(const Pt &p) { est.center_on(p); act.center_on(p); }   /* :569-574 */

    /* kernels (Interpolate.cpp:144-258): copy the box from the volume, then derivatives.
     * A box reaching outside the volume is flagged (reference: undefined reads). */
    void load(const Vol &v) {
        long bx = est.mn.ix, by = est.mn.iy, bz = est.mn.iz;
        box_ok = !(bx < 0 || by < 0 || bz < 0 || bx + nx > v.W || by + ny > v.H || bz + nz > v.T);
        if (!box_ok) return;
#pragma omp parallel for num_threads(nthreads)
        for (int is = 0; is < nz; is++)
            for (int ir = 0; ir < ny; ir++)
                for (int ic = 0; ic < nx; ic++) {
                    m[c0(ic, ir, is)] = v.at(bx + ic, by + ir, bz + is);
                    avail[((size_t)ic * ny + ir) * nz + is] = 0;
                }
        derivs();
    }

    /* kernels_derivs (Interpolate.cpp:260-337) with Matrix_4d::set_c0_fsp/fdp/ftp
     * (Matrix_4d.h:108-133), then the face copies. */
    void derivs() {
        double *M = m.data();
#pragma omp parallel num_threads(nthreads)
        {
#pragma omp for
            for (int ic = 1; ic < nx - 1; ic++)
                for (int ir = 1; ir < ny - 1; ir++)
                    for (int is = 1; is < nz - 1; is++) {
                        size_t c = c0(ic, ir, is);
                        M[c + 1] = (M[c + xh] - M[c - xh]) / 2.0;
                        M[c + 2] = (M[c + yh] - M[c - yh]) / 2.0;
                        M[c + 3] = (M[c + zh] - M[c - zh]) / 2.0;
                        M[c + 4] = (M[c + xh + yh] - M[c - xh + yh] + M[c - xh - yh] - M[c + xh - yh]) / 4.0;
                        M[c + 5] = (M[c + xh + zh] - M[c - xh + zh] + M[c - xh - zh] - M[c + xh - zh]) / 4.0;
                        M[c + 6] = (M[c + yh + zh] - M[c - yh + zh] + M[c - yh - zh] - M[c + yh - zh]) / 4.0;
                        double t1 = (M[c + xh + yh - zh] - M[c - xh + yh - zh] + M[c - xh - yh - zh] - M[c + xh - yh - zh]) / 4.0;
                        double t2 = (M[c + xh + yh + zh] - M[c - xh + yh + zh] + M[c - xh - yh + zh] - M[c + xh - yh + zh]) / 4.0;
                        M[c + 7] = (t2 - t1) / 2.0;
                    }
#pragma omp for
            for (int ir = 0; ir < ny; ir++)
                for (int is = 0; is < nz; is++)
                    for (int n = 1; n < 8; n++) {
                        M[c0(0, ir, is) + n] = M[c0(1, ir, is) + n];
                        M[c0(nx - 1, ir, is) + n] = M[c0(nx - 2, ir, is) + n];
                    }
#pragma omp for
            for (int ic = 0; ic < nx; ic++)
                for (int is = 0; is < nz; is++)
                    for (int n = 1; n < 8; n++) {
                        M[c0(ic, 0, is) + n] = M[c0(ic, 1, is) + n];
                        M[c0(ic, ny - 1, is) + n] = M[c0(ic, ny - 2, is) + n];
                    }
#pragma omp for
            for (int ic = 0; ic < nx; ic++)
                for (int ir = 0; ir < ny; ir++)
                    for (int n = 1; n < 8; n++) {
                        M[c0(ic, ir, 0) + n] = M[c0(ic, ir, 1) + n];
                        M[c0(ic, ir, nz - 1) + n] = M[c0(ic, ir, nz - 2) + n];
                    }
        }
    }

    /* kernels_Lekien_one (Interpolate.cpp:377-402) with Matrix_4d::load_vb_k (Matrix_4d.cpp:65-77). */
    void lekien_one(int ic, int ir, int is) {
        double vb[64], va[64];
        size_t c = c0(ic, ir, is);
        const size_t sh[8] = {0, (size_t)xh, (size_t)yh, (size_t)(xh + yh),
                              (size_t)zh, (size_t)(xh + zh), (size_t)(yh + zh), (size_t)(xh + yh + zh)};
        for (int k = 0; k < 8; k++)
            for (int cc = 0; cc < 8; cc++) vb[k * 8 + cc] = m[c + sh[cc] + k];
        for (int i = 0; i < 64; i++) va[i] = 0.0;
        for (int j = 0; j < 64; j++)
            for (int p = sbi->colptr[j]; p < sbi->colptr[j + 1]; p++) va[sbi->row[p]] += sbi->val[p] * vb[j];
        for (int n = 0; n < 64; n++) m[c + 8 + n] = va[n];
    }

    /* Range rule shared by all methods: act_box.contains(bbox) (Interpolate.cpp:474-481),
     * box inside the volume, and every sample cell with its upper corner inside the box
     * (the reference indexes one cell past kern_4d otherwise, Interpolate.cpp:494-500). */
    bool range_ok(const Pt *pts, int n, const Box &bbox, bool need_upper) const {
        // A non-finite box (parameters gone NaN after a singular solve) would pass the
        // comparisons of BoundBox::contains and index kern_4d out of range in the
        // reference; it is a range failure here.
        const double b6[6] = {bbox.mn.x, bbox.mn.y, bbox.mn.z, bbox.mx.x, bbox.mx.y, bbox.mx.z};
        for (int q = 0; q < 6; q++)
            if (!(b6[q] == b6[q]) || std::fabs(b6[q]) > 1.0e15) return false;
        if (!act.contains(bbox)) return false;
        if (!box_ok) return false;
        int bx = est.mn.ix, by = est.mn.iy, bz = est.mn.iz;
        for (int i = 0; i < n; i++) {
            int cx = pts[i].ix - bx, cy = pts[i].iy - by, cz = pts[i].iz - bz;
            int lim = need_upper ? 2 : 1;
            if (cx < 0 || cy < 0 || cz < 0 || cx > nx - lim || cy > ny - lim || cz > nz - lim) return false;
        }
        return true;
    }

    /* nearest (Interpolate.cpp:404-427) */
    int nearest(const Pt *pts, int n, const Box &bbox, double *out) {
        if (!range_ok(pts, n, bbox, false)) return -1;
        for (int i = 0; i < n; i++)
            out[i] = m[c0(pts[i].ix - est.mn.ix, pts[i].iy - est.mn.iy, pts[i].iz - est.mn.iz)];
        return 0;
    }

    /* tri_lin (Interpolate.cpp:429-470) */
    int tri_lin(const Pt *pts, int n, const Box &bbox, double *out) {
        if (!range_ok(pts, n, bbox, true)) return -1;
        for (int i = 0; i < n; i++) {
            int cx = pts[i].ix - est.mn.ix, cy = pts[i].iy - est.mn.iy, cz = pts[i].iz - est.mn.iz;
            double rx = pts[i].rx, ry = pts[i].ry, rz = pts[i].rz;
            out[i] = m[c0(cx + 0, cy + 0, cz + 0)] * (1.0 - rx) * (1.0 - ry) * (1.0 - rz) +
                     m[c0(cx + 1, cy + 0, cz + 0)] * (rx) * (1.0 - ry) * (1.0 - rz) +
                     m[c0(cx + 0, cy + 1, cz + 0)] * (1.0 - rx) * (ry) * (1.0 - rz) +
                     m[c0(cx + 0, cy + 0, cz + 1)] * (1.0 - rx) * (1.0 - ry) * (rz) +
                     m[c0(cx + 1, cy + 0, cz + 1)] * (rx) * (1.0 - ry) * (rz) +
                     m[c0(cx + 0, cy + 1, cz + 1)] * (1.0 - rx) * (ry) * (rz) +
                     m[c0(cx + 1, cy + 1, cz + 0)] * (rx) * (ry) * (1.0 - rz) +
                     m[c0(cx + 1, cy + 1, cz + 1)] * (rx) * (ry) * (rz);
        }
        return 0;
    }

    /* tri_cub_Lek (Interpolate.cpp:472-567); grad (optional) adds the analytic
     * derivative of the same polynomial with respect to x, y, z. */
    int tri_cub(const Pt *pts, int n, const Box &bbox, double *out, double *grad) {
        if (!range_ok(pts, n, bbox, true)) return -1;
        int bx = est.mn.ix, by = est.mn.iy, bz = est.mn.iz;
#pragma omp parallel num_threads(nthreads)
        {
#pragma omp for
            for (int i = 0; i < n; i++) {
                int cx = pts[i].ix - bx, cy = pts[i].iy - by, cz = pts[i].iz - bz;
                unsigned char *f = &avail[((size_t)cx * ny + cy) * nz + cz];
                if (!__atomic_test_and_set(f, __ATOMIC_ACQ_REL)) lekien_one(cx, cy, cz);
            }
#pragma omp for
            for (int i = 0; i < n; i++) {
                int cx = pts[i].ix - bx, cy = pts[i].iy - by, cz = pts[i].iz - bz;
                const double *a = &m[c0(cx, cy, cz) + 8];
                double px[4], py[4], pz[4];
                px[0] = 1.0; py[0] = 1.0; pz[0] = 1.0;
                px[1] = pts[i].rx; py[1] = pts[i].ry; pz[1] = pts[i].rz;
                px[2] = px[1] * px[1]; py[2] = py[1] * py[1]; pz[2] = pz[1] * pz[1];
                px[3] = px[2] * px[1]; py[3] = py[2] * py[1]; pz[3] = pz[2] * pz[1];
                double val = 0;
                for (int ii = 0; ii < 4; ii++) {
                    for (int jj = 0; jj < 4; jj++) {
                        double pij = px[ii] * py[jj];
                        int idx = ii + 4 * jj;
                        val += a[idx] * pij;
                        idx += 16; val += a[idx] * pij * pz[1];
                        idx += 16; val += a[idx] * pij * pz[2];
                        idx += 16; val += a[idx] * pij * pz[3];
                    }
                }
                out[i] = val;
                if (grad) {
                    double dpx[4] = {0.0, 1.0, 2.0 * px[1], 3.0 * px[2]};
                    double dpy[4] = {0.0, 1.0, 2.0 * py[1], 3.0 * py[2]};
                    double dpz[4] = {0.0, 1.0, 2.0 * pz[1], 3.0 * pz[2]};
                    double gx = 0, gy = 0, gz = 0;
                    for (int kk = 0; kk < 4; kk++)
                        for (int jj = 0; jj < 4; jj++)
                            for (int ii = 0; ii < 4; ii++) {
                                double al = a[ii + 4 * jj + 16 * kk];
                                gx += al * dpx[ii] * py[jj] * pz[kk];
                                gy += al * px[ii] * dpy[jj] * pz[kk];
                                gz += al * px[ii] * py[jj] * dpz[kk];
                            }
                    grad[3 * i + 0] = gx; grad[3 * i + 1] = gy; grad[3 * i + 2] = gz;
                }
            }
        }
        return 0;
    }
};

/* --------------------------------------------------------- objective functions */

/* ObjectiveFunctions.cpp:24-263 (with-residual overloads; residual may be NULL). */
static double objective(int kind, const double *r, const double *t, int n, double *res) {
    double obj = 0.0;
    if (kind == 0 || kind == 1) {                       /* SAD :35-45, SSD :59-70 */
        for (int i = 0; i < n; i++) {
            double d = t[i] - r[i];
            if (res) res[i] = d;
            obj += (kind == 0) ? fabs(d) : d * d;
        }
        return obj;
    }
    if (kind == 2) {                                    /* ZSSD :98-123 */
        double ar = 0.0, at = 0.0;
        for (int i = 0; i < n; i++) { ar += r[i]; at += t[i]; }
        ar = ar / n; at = at / n;
        for (int i = 0; i < n; i++) {
            double d = (t[i] - at) - (r[i] - ar);
            if (res) res[i] = d;
            obj += d * d;
        }
        return obj;
    }
    if (kind == 3) {                                    /* NSSD :155-184 */
        double sr = 0.0, st = 0.0;
        for (int i = 0; i < n; i++) { sr += r[i] * r[i]; st += t[i] * t[i]; }
        sr = std::sqrt(sr); st = std::sqrt(st);
        for (int i = 0; i < n; i++) {
            double d = t[i] / st - r[i] / sr;
            if (res) res[i] = d;
            obj += d * d;
        }
        return obj / 2.0;
    }
    /* ZNSSD :225-263 */
    double ar = 0.0, at = 0.0;
    for (int i = 0; i < n; i++) { ar += r[i]; at += t[i]; }
    ar = ar / n; at = at / n;
    double sr = 0.0, st = 0.0;
    for (int i = 0; i < n; i++) { sr += (r[i] - ar) * (r[i] - ar); st += (t[i] - at) * (t[i] - at); }
    sr = std::sqrt(sr); st = std::sqrt(st);
    for (int i = 0; i < n; i++) {
        double d = ((t[i] - at) / st) - ((r[i] - ar) / sr);
        if (res) res[i] = d;
        obj += d * d;
    }
    return obj / 4.0;
}

/* ------------------------------------------------------------- dense solvers */

/* Householder vector of x[0..m): Eigen's makeHouseholder (Householder.h). */
static void make_householder(double *x, int m, double &tau, double &beta) {
    double tail = 0.0;
    for (int i = 1; i < m; i++) tail += x[i] * x[i];
    double c0 = x[0];
    const double tol = std::numeric_limits<double>::min();
    if (tail <= tol) {
        tau = 0.0; beta = c0;
        for (int i = 1; i < m; i++) x[i] = 0.0;
    } else {
        beta = std::sqrt(c0 * c0 + tail);
        if (c0 >= 0.0) beta = -beta;
        for (int i = 1; i < m; i++) x[i] = x[i] / (c0 - beta);
        tau = (beta - c0) / beta;
    }
}

/* Apply H = I - tau [1;v][1;v]^T from the left to the m x nc block starting at (r0, c0)
 * of a column-major matrix with leading dimension ld (Eigen applyHouseholderOnTheLeft). */
static void apply_householder_left(double *M, int ld, int r0, int c0, int m, int nc,
                                   const double *v, double tau) {
    if (m == 1) {
        for (int j = 0; j < nc; j++) M[(size_t)(c0 + j) * ld + r0] *= (1.0 - tau);
        return;
    }
    if (tau == 0.0) return;
    for (int j = 0; j < nc; j++) {
        double *col = &M[(size_t)(c0 + j) * ld + r0];
        double tmp = 0.0;
        for (int i = 1; i < m; i++) tmp += v[i - 1] * col[i];
        tmp += col[0];
        col[0] -= tau * tmp;
        for (int i = 1; i < m; i++) col[i] -= tau * v[i - 1] * tmp;
    }
}

/* ColPivHouseholderQR(A).solve(b) for square n x n A, restating Eigen 3.4's
 * ColPivHouseholderQR::computeInPlace and _solve_impl (used at Search.cpp:234). */
static void colpiv_solve(const double *Arow, const double *b, int n, double *x) {
    std::vector<double> Q((size_t)n * n);                   /* column major */
    for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) Q[(size_t)j * n + i] = Arow[i * n + j];
    std::vector<double> hc(n), nu(n), nd(n);
    std::vector<int> tr(n);
    for (int k = 0; k < n; k++) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += Q[(size_t)k * n + i] * Q[(size_t)k * n + i];
        nd[k] = std::sqrt(s); nu[k] = nd[k];
    }
    double maxn = 0.0;
    for (int k = 0; k < n; k++) maxn = std::max(maxn, nu[k]);
    double eps = std::numeric_limits<double>::epsilon();
    double thr_helper = (maxn * eps) * (maxn * eps) / (double)n;
    double down_thr = std::sqrt(eps);
    int nonzero = n;
    for (int k = 0; k < n; k++) {
        int big = k; double bn = nu[k];
        for (int j = k + 1; j < n; j++) if (nu[j] > bn) { bn = nu[j]; big = j; }
        double bsq = bn * bn;
        if (nonzero == n && bsq < thr_helper * (double)(n - k)) nonzero = k;
        tr[k] = big;
        if (k != big) {
            for (int i = 0; i < n; i++) std::swap(Q[(size_t)k * n + i], Q[(size_t)big * n + i]);
            std::swap(nu[k], nu[big]); std::swap(nd[k], nd[big]);
        }
        double tau, beta;
        make_householder(&Q[(size_t)k * n + k], n - k, tau, beta);
        hc[k] = tau;
        Q[(size_t)k * n + k] = beta;
        apply_householder_left(Q.data(), n, k, k + 1, n - k, n - k - 1, &Q[(size_t)k * n + k + 1], tau);
        for (int j = k + 1; j < n; j++) {
            if (nu[j] != 0.0) {
                double t = fabs(Q[(size_t)j * n + k]) / nu[j];
                t = (1.0 + t) * (1.0 - t);
                t = t < 0.0 ? 0.0 : t;
                double r = nu[j] / nd[j];
                double t2 = t * (r * r);
                if (t2 <= down_thr) {
                    double s = 0.0;
                    for (int i = k + 1; i < n; i++) s += Q[(size_t)j * n + i] * Q[(size_t)j * n + i];
                    nd[j] = std::sqrt(s); nu[j] = nd[j];
                } else {
                    nu[j] *= std::sqrt(t);
                }
            }
        }
    }
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int k = 0; k < n; k++) std::swap(perm[k], perm[tr[k]]);
    if (nonzero == 0) { for (int i = 0; i < n; i++) x[i] = 0.0; return; }
    std::vector<double> c(b, b + n);
    for (int k = 0; k < nonzero; k++)
        apply_householder_left(c.data(), n, k, 0, n - k, 1, &Q[(size_t)k * n + k + 1], hc[k]);
    for (int i = nonzero - 1; i >= 0; i--) {
        double s = c[i];
        for (int j = i + 1; j < nonzero; j++) s -= Q[(size_t)j * n + i] * c[j];
        c[i] = s / Q[(size_t)i * n + i];
    }
    for (int i = 0; i < n; i++) x[i] = 0.0;
    for (int i = 0; i < nonzero; i++) x[perm[i]] = c[i];
}

/* FullPivHouseholderQR(A).solve(b), restating Eigen 3.4's FullPivHouseholderQR
 * (used at strain.cpp:533). */
static void fullpiv_solve(const double *Arow, const double *b, int n, double *x) {
    std::vector<double> Q((size_t)n * n);
    for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) Q[(size_t)j * n + i] = Arow[i * n + j];
    std::vector<double> hc(n, 0.0);
    std::vector<int> rt(n), ct(n);
    double prec = std::numeric_limits<double>::epsilon() * (double)n;
    double biggest = 0.0, maxpivot = 0.0;
    int nonzero = n;
    for (int k = 0; k < n; k++) {
        int rb = k, cb = k; double sc = -1.0;
        for (int j = k; j < n; j++)
            for (int i = k; i < n; i++) {
                double a = fabs(Q[(size_t)j * n + i]);
                if (a > sc) { sc = a; rb = i; cb = j; }
            }
        double bic = sc;
        if (k == 0) biggest = bic;
        if (bic <= biggest * prec) {                      /* isMuchSmallerThan */
            nonzero = k;
            for (int i = k; i < n; i++) { rt[i] = i; ct[i] = i; hc[i] = 0.0; }
            break;
        }
        rt[k] = rb; ct[k] = cb;
        if (k != rb) for (int j = k; j < n; j++) std::swap(Q[(size_t)j * n + k], Q[(size_t)j * n + rb]);
        if (k != cb) for (int i = 0; i < n; i++) std::swap(Q[(size_t)k * n + i], Q[(size_t)cb * n + i]);
        double tau, beta;
        make_householder(&Q[(size_t)k * n + k], n - k, tau, beta);
        hc[k] = tau;
        Q[(size_t)k * n + k] = beta;
        if (fabs(beta) > maxpivot) maxpivot = fabs(beta);
        apply_householder_left(Q.data(), n, k, k + 1, n - k, n - k - 1, &Q[(size_t)k * n + k + 1], tau);
    }
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int k = 0; k < n; k++) std::swap(perm[k], perm[ct[k]]);
    double pthr = fabs(maxpivot) * (std::numeric_limits<double>::epsilon() * (double)n);
    int rank = 0;
    for (int i = 0; i < nonzero; i++) rank += (fabs(Q[(size_t)i * n + i]) > pthr);
    for (int i = 0; i < n; i++) x[i] = 0.0;
    if (rank == 0) return;
    std::vector<double> c(b, b + n);
    for (int k = 0; k < rank; k++) {
        std::swap(c[k], c[rt[k]]);
        apply_householder_left(c.data(), n, k, 0, n - k, 1, &Q[(size_t)k * n + k + 1], hc[k]);
    }
    for (int i = rank - 1; i >= 0; i--) {
        double s = c[i];
        for (int j = i + 1; j < rank; j++) s -= Q[(size_t)j * n + i] * c[j];
        c[i] = s / Q[(size_t)i * n + i];
    }
    for (int i = 0; i < rank; i++) x[perm[i]] = c[i];
}

/* Eigenvalues of a symmetric 3x3 by cyclic Jacobi rotations, sorted descending.
 * (strain.cpp:553-556 takes EigenSolver real parts of a symmetric matrix and sorts.)
 * From the fifth sweep on, an off-diagonal entry too small to change either diagonal
 * entry it couples (|a_pp| + 100 |a_pq| == |a_pp|, same for a_qq) is set to zero: the
 * classical cyclic-Jacobi stopping rule, since rotations alone stall at the rounding
 * level of the diagonal. */
static void sym3_eigen(const double *S, double *lam) {
    double a[3][3];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) a[i][j] = 0.5 * (S[i * 3 + j] + S[j * 3 + i]);
    for (int sweep = 0; sweep < 50; sweep++) {
        double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        if (off == 0.0) break;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                if (a[p][q] == 0.0) continue;
                const double g = 100.0 * fabs(a[p][q]);
                if (sweep > 3 && fabs(a[p][p]) + g == fabs(a[p][p]) && fabs(a[q][q]) + g == fabs(a[q][q])) {
                    a[p][q] = 0.0;
                    a[q][p] = 0.0;
                    continue;
                }
                double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; k++) {
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; k++) {
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
            }
    }
    double l[3] = {a[0][0], a[1][1], a[2][2]};
    std::sort(l, l + 3);
    lam[0] = l[2]; lam[1] = l[1]; lam[2] = l[0];
}

/* ------------------------------------------------------------------- Search */

struct Result {
    int status; int iters; double obj; double par[12];
};

struct Searcher {
    orc_run rc;
    Vol ref, tar;
    double subv_rad; int subv_num;
    Box est_nom;
    Interp *interp;
    std::vector<double> ref_sv, tar_sv;
    double par_min[12]; double obj_min;
    FCloud *fc;
    int nthreads;
    int iters;

    Searcher(const orc_run &r, const Vol &rv, const Vol &tv, const SparseBI *sbi, int nthr)
        : rc(r), ref(rv), tar(tv), fc(nullptr), nthreads(nthr), iters(0) {
        subv_rad = rc.subvol_size / 2.0;                                     /* Search.cpp:29 */
        if (rc.geom == 0) subv_num = (int)pow(ceil(pow((double)rc.subvol_npts, 1.0 / 3.0)), 3.0);
        else subv_num = rc.subvol_npts;
        ref_sv.assign(subv_num, 0.0); tar_sv.assign(subv_num, 0.0);
        for (int i = 0; i < 12; i++) par_min[i] = 0.0;
        obj_min = 0.0;
        est_nom = Box(Pt(0.0, 0.0, 0.0),                                     /* :66-76 */
                      Pt(2 * subv_rad * rc.aspect[0], 2 * subv_rad * rc.aspect[1], 2 * subv_rad * rc.aspect[2]));
        est_nom.grow_by(rc.disp_max);
        est_nom.grow_by(2.0);
        interp = new Interp(est_nom, sbi, nthr);
    }
    ~Searcher() { delete interp; delete fc; }

    /* interpolate cloud c into out (tricubic / trilinear / nearest per rc.interp). */
    int sample(const Cloud &c, double *out, double *grad) {
        if (rc.interp == 0) return interp->nearest(c.pts.data(), (int)c.pts.size(), c.box, out);
        if (rc.interp == 1) return interp->tri_lin(c.pts.data(), (int)c.pts.size(), c.box, out);
        return interp->tri_cub(c.pts.data(), (int)c.pts.size(), c.box, out, grad);
    }

    /* obj_val_at (Search.cpp:448-488); returns 0 or -1 (Range_Fail). */
    int obj_val_at(const double *x, int xs, double &obj, double *res, double *grad) {
        // Non-finite parameters (a singular normal matrix) make every sample NaN; the
        // reference then indexes kern_4d with undefined cells.  Range failure here.
        for (int q = 0; q < xs; q++)
            if (!std::isfinite(x[q])) return -1;
        fc->affine_to(x, xs, xs);
        if (sample(fc->moving, tar_sv.data(), grad) != 0) return -1;
        obj = objective(rc.obj, ref_sv.data(), tar_sv.data(), subv_num, res);
        return 0;
    }

    /* starting_param (Search.cpp:277-377) */
    void starting_param(const std::vector<const Result *> &nres) {
        double sx = 0.0, sy = 0.0, sz = 0.0; int good = 0;
        for (size_t i = 0; i < nres.size(); i++)
            if (nres[i]->status == 0) { sx += nres[i]->par[0]; sy += nres[i]->par[1]; sz += nres[i]->par[2]; good += 1; }
        if (good != 0) { sx /= good; sy /= good; sz /= good; }
        if (good != 0) { par_min[0] = sx; par_min[1] = sy; par_min[2] = sz; }
        else { par_min[0] = rc.rigid_trans[0]; par_min[1] = rc.rigid_trans[1]; par_min[2] = rc.rigid_trans[2]; }
        for (int i = 3; i < rc.ndof; i++) par_min[i] = 0.0;
    }

    /* trgrid_global (Search.cpp:934-1043) */
    int trgrid_global(double dmax, double basin) {
        if (basin == 0.0) return 0;
        int num_inc = (int)(dmax / basin);
        double inc = dmax / num_inc;
        double omin = std::numeric_limits<double>::max();
        int mx = 0, my = 0, mz = 0;
        double pc[3];
        for (int z = -num_inc; z <= num_inc; z++)
            for (int y = -num_inc; y <= num_inc; y++)
                for (int x = -num_inc; x <= num_inc; x++) {
                    pc[0] = par_min[0] + (double)x * inc;
                    pc[1] = par_min[1] + (double)y * inc;
                    pc[2] = par_min[2] + (double)z * inc;
                    double ov;
                    if (obj_val_at(pc, 3, ov, nullptr, nullptr) != 0) return -1;
                    if (ov < omin) { omin = ov; mx = x; my = y; mz = z; }
                }
        par_min[0] += (double)mx * inc;
        par_min[1] += (double)my * inc;
        par_min[2] += (double)mz * inc;
        return 0;
    }

    /* Residual Jacobian (npts x ndof, row major) at a.  Forward differences as
     * LM_prep_at (Search.cpp:490-525) or the analytic derivative of the same residual. */
    int lm_prep(const double *a, int nd, std::vector<double> &e, std::vector<double> &J, double &obj) {
        int np = subv_num;
        std::vector<double> base(np), step(np);
        if (rc.jacobian == 0) {
            if (obj_val_at(a, nd, obj, base.data(), nullptr) != 0) return -1;
            double h = 1E-10;
            double sx[12];
            for (int i = 0; i < nd; i++) {
                for (int q = 0; q < nd; q++) sx[q] = a[q];
                sx[i] += h;
                double dummy;
                if (obj_val_at(sx, nd, dummy, step.data(), nullptr) != 0) return -1;
                for (int j = 0; j < np; j++) J[(size_t)j * nd + i] = (step[j] - base[j]) / h;
            }
            for (int j = 0; j < np; j++) e[j] = base[j];
            return 0;
        }
        /* analytic: residual derivative = d(residual)/d(t) * grad(t) . dx/dp */
        std::vector<double> grad((size_t)np * 3);
        if (obj_val_at(a, nd, obj, base.data(), grad.data()) != 0) return -1;
        double dA[3][3][3];
        rot_derivs(nd > 3 ? a[3] : 0.0, nd > 3 ? a[4] : 0.0, nd > 3 ? a[5] : 0.0, dA);
        const Params &P = fc->params;
        std::vector<double> g((size_t)np * nd);
        for (int j = 0; j < np; j++) {
            const Pt &s = fc->stable.pts[j];
            double rel[3] = {s.x - fc->sref.x, s.y - fc->sref.y, s.z - fc->sref.z};
            double gr[3] = {grad[3 * j], grad[3 * j + 1], grad[3 * j + 2]};
            double *gj = &g[(size_t)j * nd];
            gj[0] = gr[0]; gj[1] = gr[1]; gj[2] = gr[2];
            if (nd > 3) {
                double w[3];
                for (int r = 0; r < 3; r++) w[r] = P.A[r][0] * rel[0] + P.A[r][1] * rel[1] + P.A[r][2] * rel[2];
                /* (I+S) dA rel for each angle */
                for (int q = 0; q < 3; q++) {
                    double d[3], dd[3];
                    for (int r = 0; r < 3; r++) d[r] = dA[q][r][0] * rel[0] + dA[q][r][1] * rel[1] + dA[q][r][2] * rel[2];
                    for (int r = 0; r < 3; r++) {
                        dd[r] = d[r];
                        if (nd > 6) dd[r] += P.S[r][0] * d[0] + P.S[r][1] * d[1] + P.S[r][2] * d[2];
                    }
                    gj[3 + q] = gr[0] * dd[0] + gr[1] * dd[1] + gr[2] * dd[2];
                }
                if (nd > 6) {
                    gj[6] = gr[0] * w[0];
                    gj[7] = gr[1] * w[1];
                    gj[8] = gr[2] * w[2];
                    gj[9] = gr[0] * w[1] + gr[1] * w[0];
                    gj[10] = gr[1] * w[2] + gr[2] * w[1];
                    gj[11] = gr[0] * w[2] + gr[2] * w[0];
                }
            }
        }
        const double *r = ref_sv.data(), *t = tar_sv.data();
        int kind = rc.obj;
        if (kind == 0 || kind == 1) {
            for (size_t q = 0; q < (size_t)np * nd; q++) J[q] = g[q];
        } else if (kind == 2) {
            std::vector<double> gm(nd, 0.0);
            for (int j = 0; j < np; j++) for (int p = 0; p < nd; p++) gm[p] += g[(size_t)j * nd + p];
            for (int p = 0; p < nd; p++) gm[p] /= np;
            for (int j = 0; j < np; j++) for (int p = 0; p < nd; p++) J[(size_t)j * nd + p] = g[(size_t)j * nd + p] - gm[p];
        } else if (kind == 3) {
            double T2 = 0.0;
            for (int j = 0; j < np; j++) T2 += t[j] * t[j];
            double T = std::sqrt(T2);
            std::vector<double> c(nd, 0.0);
            for (int j = 0; j < np; j++) for (int p = 0; p < nd; p++) c[p] += t[j] * g[(size_t)j * nd + p];
            for (int j = 0; j < np; j++)
                for (int p = 0; p < nd; p++)
                    J[(size_t)j * nd + p] = g[(size_t)j * nd + p] / T - t[j] * c[p] / (T * T * T);
        } else {
            double tm = 0.0;
            for (int j = 0; j < np; j++) tm += t[j];
            tm /= np;
            double N2 = 0.0;
            for (int j = 0; j < np; j++) N2 += (t[j] - tm) * (t[j] - tm);
            double N = std::sqrt(N2);
            std::vector<double> gm(nd, 0.0), c(nd, 0.0);
            for (int j = 0; j < np; j++)
                for (int p = 0; p < nd; p++) { gm[p] += g[(size_t)j * nd + p]; c[p] += (t[j] - tm) * g[(size_t)j * nd + p]; }
            for (int p = 0; p < nd; p++) gm[p] /= np;
            for (int j = 0; j < np; j++)
                for (int p = 0; p < nd; p++)
#ifdef ORC_PLANT_ZNSSD_DROP_TERM
                    J[(size_t)j * nd + p] = (g[(size_t)j * nd + p] - gm[p]) / N;   /* planted mistake (tests only) */
#else
                    J[(size_t)j * nd + p] = (g[(size_t)j * nd + p] - gm[p]) / N - (t[j] - tm) * c[p] / (N * N * N);
#endif
        }
        (void)r;
        for (int j = 0; j < np; j++) e[j] = base[j];
        return 0;
    }

    /* min_Lev_Mar (Search.cpp:194-246) */
    int min_lev_mar(const double *start, int nd, double obj_tol, double mag_tol, double *jump) {
        int np = subv_num;
        std::vector<double> e(np), J((size_t)np * nd);
        double JTJ[144], JTe[12], upd[12], rhs[12];
        for (int i = 0; i < nd; i++) { jump[i] = start[i]; upd[i] = 0.0; }
        double obj_old = 0.0;
        iters = 0;
        for (int i = 0; i < 20; i++) {
            double obj;
            iters += 1;
            if (lm_prep(jump, nd, e, J, obj) != 0) return -1;
#ifdef ORC_PLANT_LATE_STOP
            if (i > 1) {            /* planted mistake (tests only): the test skips iteration 1 */
#else
            if (i > 0) {
#endif
                double del_obj = fabs(obj - obj_old);
                double del_mag = std::sqrt(upd[0] * upd[0] + upd[1] * upd[1] + upd[2] * upd[2]);
#ifdef ORC_PLANT_DELMAG_ALL
                /* planted mistake (tests only): every parameter of the step in the step size */
                del_mag = 0.0;
                for (int q = 0; q < nd; q++) del_mag += upd[q] * upd[q];
                del_mag = std::sqrt(del_mag);
#endif
#ifdef ORC_PLANT_MAGTOL
                mag_tol = 0.5 * mag_tol;   /* planted mistake (tests only): half the step tolerance */
#endif
                if ((del_obj < obj_tol) || (del_mag < mag_tol)) break;
            }
            for (int a = 0; a < nd; a++) {
                for (int b = 0; b < nd; b++) {
                    double s = 0.0;
                    for (int j = 0; j < np; j++) s += J[(size_t)j * nd + a] * J[(size_t)j * nd + b];
                    JTJ[a * nd + b] = s;
                }
                double s = 0.0;
                for (int j = 0; j < np; j++) s += J[(size_t)j * nd + a] * e[j];
                JTe[a] = s;
                rhs[a] = -JTe[a];
            }
            colpiv_solve(JTJ, rhs, nd, upd);
            for (int q = 0; q < nd; q++) jump[q] += upd[q];
            obj_old = obj;
        }
        return 0;
    }

    /* process_point (Search.cpp:95-192) returning the status code (Utility.h:38-41). */
    int process_point(const Pt &sp, const std::vector<const Result *> &nres, const double *start = nullptr) {
        delete fc;
        fc = new FCloud();
        if (rc.geom == 1)
            fc->make_sphere(sp, subv_rad, subv_num, rc.aspect[0], rc.aspect[1], rc.aspect[2], rc.sphere_seed);
        if (rc.geom == 0) {
            Pt smin = sp; smin.move_by(-subv_rad * rc.aspect[0], -subv_rad * rc.aspect[1], -subv_rad * rc.aspect[2]);
            Pt smax = sp; smax.move_by(subv_rad * rc.aspect[0], subv_rad * rc.aspect[1], subv_rad * rc.aspect[2]);
            int g = (int)ceil(pow((double)subv_num, 1.0 / 3.0));
            fc->make_cube(smin, smax, g, g, g);
        }
        iters = 0;
        /* search_pt_setup (Search.cpp:248-275) */
        interp->center_on(sp);
        interp->load(ref);
        if (sample(fc->stable, ref_sv.data(), nullptr) != 0) return -1;
        if (start) {
            /* a given start warp (no reference counterpart): all ndof parameters */
            for (int i = 0; i < rc.ndof; i++) {
                if (!std::isfinite(start[i])) return -1;
                par_min[i] = start[i];
            }
        } else {
            starting_param(nres);
        }
        Pt off = sp; off.move_by(par_min[0], par_min[1], par_min[2]);
        interp->center_on(off);
        interp->load(tar);
        if (trgrid_global(rc.disp_max, rc.basin_radius) != 0) return -1;
        double jump[12];
        if (min_lev_mar(par_min, rc.ndof, 0.000001, 0.01, jump) != 0) return -1;
        for (int i = 0; i < rc.ndof; i++) par_min[i] = jump[i];
        double ov;
        if (obj_val_at(par_min, rc.ndof, ov, nullptr, nullptr) != 0) return -1;
        obj_min = ov;
        return 0;
    }
};

static SparseBI *g_sbi() {
    static SparseBI s;
    return &s;
}

static Vol make_vol(const orc_run *rc, const void *p) {
    Vol v; v.data = p; v.bpv = rc->bytes_per; v.W = rc->vol_w; v.H = rc->vol_h; v.T = rc->vol_t;
    return v;
}

}  // namespace

/* ======================================================================= C API */

extern "C" {

int orc_version(void) { return 1; }

void orc_bi_matrix(double *out) {
    build_BI();
    for (int i = 0; i < 64; i++) for (int j = 0; j < 64; j++) out[i * 64 + j] = g_BI[i][j];
}

void orc_constraint_matrix(double *out) {
    /* B[q*8+c][i+4j+16k] = d^q (x^i y^j z^k) at corner c. */
    for (int q = 0; q < 8; q++) for (int c = 0; c < 8; c++)
        for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) for (int k = 0; k < 4; k++) {
            int e[3] = {i, j, k}; double v = 1.0;
            for (int a = 0; a < 3; a++) {
                int d = QD[q][a], x = CO[c][a];
                double coef = 1.0; int pw = e[a];
                if (d == 1) { if (pw == 0) { coef = 0.0; } else { coef = pw; pw -= 1; } }
                double pv = (pw == 0) ? 1.0 : (x == 0 ? 0.0 : 1.0);
                v *= coef * pv;
            }
            out[(q * 8 + c) * 64 + i + 4 * j + 16 * k] = v;
        }
}

int orc_subv_num(int geom, int npts) {
    if (geom == 0) return (int)pow(ceil(pow((double)npts, 1.0 / 3.0)), 3.0);
    return npts;
}

int orc_sphere_offsets(unsigned int seed, double rad, int num, const double aspect[3], double *off) {
    std::mt19937 mt(seed);
    std::uniform_real_distribution<double> dist(-rad, rad);
    int count = 0;
    while (count < num) {
        double rx = dist(mt), ry = dist(mt), rz = dist(mt);
        double radius = std::sqrt(rx * rx + ry * ry + rz * rz);
        if (radius <= rad) {
            off[3 * count + 0] = aspect[0] * rx;
            off[3 * count + 1] = aspect[1] * ry;
            off[3 * count + 2] = aspect[2] * rz;
            count += 1;
        }
    }
    return count;
}

static FCloud *cloud_for(const orc_run *rc, const double pt[3]) {
    double subv_rad = rc->subvol_size / 2.0;
    int subv_num = orc_subv_num(rc->geom, rc->subvol_npts);
    FCloud *fc = new FCloud();
    Pt sp(pt[0], pt[1], pt[2]);
    if (rc->geom == 1) fc->make_sphere(sp, subv_rad, subv_num, rc->aspect[0], rc->aspect[1], rc->aspect[2], rc->sphere_seed);
    else {
        Pt smin = sp; smin.move_by(-subv_rad * rc->aspect[0], -subv_rad * rc->aspect[1], -subv_rad * rc->aspect[2]);
        Pt smax = sp; smax.move_by(subv_rad * rc->aspect[0], subv_rad * rc->aspect[1], subv_rad * rc->aspect[2]);
        int g = (int)ceil(pow((double)subv_num, 1.0 / 3.0));
        fc->make_cube(smin, smax, g, g, g);
    }
    return fc;
}

int orc_cloud(const orc_run *rc, const double pt[3], double *xyz, double *bbox, double *refpt) {
    FCloud *fc = cloud_for(rc, pt);
    int n = (int)fc->stable.pts.size();
    for (int i = 0; i < n; i++) { xyz[3 * i] = fc->stable.pts[i].x; xyz[3 * i + 1] = fc->stable.pts[i].y; xyz[3 * i + 2] = fc->stable.pts[i].z; }
    if (bbox) {
        const Box &b = fc->stable.box;
        bbox[0] = b.mn.x; bbox[1] = b.mn.y; bbox[2] = b.mn.z; bbox[3] = b.mx.x; bbox[4] = b.mx.y; bbox[5] = b.mx.z;
    }
    if (refpt) { refpt[0] = fc->sref.x; refpt[1] = fc->sref.y; refpt[2] = fc->sref.z; }
    delete fc;
    return n;
}

int orc_affine(const orc_run *rc, const double pt[3], const double *par, int ndof, double *xyz, double *bbox) {
    FCloud *fc = cloud_for(rc, pt);
    fc->affine_to(par, ndof, ndof);
    int n = (int)fc->moving.pts.size();
    for (int i = 0; i < n; i++) { xyz[3 * i] = fc->moving.pts[i].x; xyz[3 * i + 1] = fc->moving.pts[i].y; xyz[3 * i + 2] = fc->moving.pts[i].z; }
    const Box &b = fc->moving.box;
    bbox[0] = b.mn.x; bbox[1] = b.mn.y; bbox[2] = b.mn.z; bbox[3] = b.mx.x; bbox[4] = b.mx.y; bbox[5] = b.mx.z;
    delete fc;
    return n;
}

int orc_interp(const orc_run *rc, const void *vol, const double center[3], const double *xyz, int n,
               const double *bbox, int method, double *out, double *grad) {
    double subv_rad = rc->subvol_size / 2.0;
    Box nom(Pt(0.0, 0.0, 0.0), Pt(2 * subv_rad * rc->aspect[0], 2 * subv_rad * rc->aspect[1], 2 * subv_rad * rc->aspect[2]));
    nom.grow_by(rc->disp_max);
    nom.grow_by(2.0);
    Interp it(nom, g_sbi(), 1);
    it.center_on(Pt(center[0], center[1], center[2]));
    it.load(make_vol(rc, vol));
    std::vector<Pt> pts(n);
    for (int i = 0; i < n; i++) pts[i] = Pt(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    Box bb(Pt(bbox[0], bbox[1], bbox[2]), Pt(bbox[3], bbox[4], bbox[5]));
    if (method == 0) return it.nearest(pts.data(), n, bb, out);
    if (method == 1) return it.tri_lin(pts.data(), n, bb, out);
    return it.tri_cub(pts.data(), n, bb, out, grad);
}

/* Search::obj_val_at (Search.cpp:448-488) after search_pt_setup (Search.cpp:248-275)
 * for one point: the reference cloud sampled with the box on pt, the box moved to pt +
 * start (the start guess's translation), the cloud under affine_to(x, nx) sampled in the
 * correlated volume, the objective (and residual, may be NULL).  Returns the cloud's size,
 * or -1 where the reference throws (Intrp_Fail sampling the reference, Range_Fail). */
int orc_obj_val_at(const orc_run *rc, const void *refv, const void *tarv, const double pt[3], const double start[3],
                   const double *x, int nx, double *obj, double *res) {
    Searcher s(*rc, make_vol(rc, refv), make_vol(rc, tarv), g_sbi(), 1);
    s.fc = cloud_for(rc, pt);
    Pt sp(pt[0], pt[1], pt[2]);
    s.interp->center_on(sp);
    s.interp->load(s.ref);
    if (s.sample(s.fc->stable, s.ref_sv.data(), nullptr) != 0) return -1;
    Pt off = sp;
    off.move_by(start[0], start[1], start[2]);
    s.interp->center_on(off);
    s.interp->load(s.tar);
    double o = 0.0;
    if (s.obj_val_at(x, nx, o, res, nullptr) != 0) return -1;
    *obj = o;
    return s.subv_num;
}

double orc_objective(int kind, const double *ref, const double *tar, int n, double *residual) {
    return objective(kind, ref, tar, n, residual);
}

struct DualSort { int index; double value; };
static bool sortByValue(const DualSort &l, const DualSort &r) { return l.value < r.value; }

int orc_sort_order_neighbors(const double *p, int n, const double start[3], int *order, int *neigh, int nthreads) {
    int k = 75 < n ? 75 : n;                                      /* DataCloud.cpp:71 */
    std::vector<Pt> pts(n);
    for (int i = 0; i < n; i++) pts[i] = Pt(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    Pt sp(start[0], start[1], start[2]);
    std::vector<DualSort> idx(n);
    for (int i = 0; i < n; i++) { idx[i].index = i; idx[i].value = pt_dist(sp, pts[i]); }
    std::sort(idx.begin(), idx.end(), sortByValue);
    for (int i = 0; i < n; i++) order[i] = idx[i].index;
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<DualSort> cp(idx);
#pragma omp for
        for (int i = 0; i < n; i++) {
            for (int j = 0; j < n; j++) { cp[j].index = j; cp[j].value = pt_dist(pts[i], pts[j]); }
            std::sort(cp.begin(), cp.end(), sortByValue);
            for (int j = 0; j < k; j++) neigh[(size_t)i * k + j] = cp[j].index;
        }
    }
    return k;
}

int orc_dvc_run(const orc_run *rc, const void *refv, const void *tarv, const double *p, int npts,
                const int *order, const int *neigh, int k, int nproc, int nthreads,
                orc_result *results, double *seconds) {
    for (int i = 0; i < npts; i++) {                              /* DataCloud.cpp:163-172 */
        results[i].status = -3; results[i].iters = 0; results[i].obj = 0.0;
        for (int q = 0; q < 12; q++) results[i].par[q] = 0.0;
    }
    Searcher S(*rc, make_vol(rc, refv), make_vol(rc, tarv), g_sbi(), nthreads);
    unsigned int N = (unsigned int)npts;                          /* dvc.cpp:229-232 */
    if ((unsigned int)nproc > 1 && (unsigned int)nproc < N) N = (unsigned int)nproc;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<const Result *> nres;
    for (unsigned int i = 0; i < N; i++) {
        int n = order[i];
        nres.clear();
        for (int j = 0; j < k; j++) nres.push_back((const Result *)&results[neigh[(size_t)n * k + j]]);
        Pt sp(p[3 * n], p[3 * n + 1], p[3 * n + 2]);
        int st = S.process_point(sp, nres);
        results[n].iters = S.iters;
        if (st == 0) {
            results[n].status = 0;
            results[n].obj = S.obj_min;
            for (int q = 0; q < rc->ndof; q++) results[n].par[q] = S.par_min[q];
        } else {
            results[n].status = st;
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return (int)N;
}

/* The serial loop with one RunControl per point (no reference counterpart): point n is
 * searched by a Searcher of rc with subvol_geom/size/npts = geom[n], size[n], npp[n]
 * (one Searcher per distinct triple); start guesses read the shared results as in
 * orc_dvc_run. */
int orc_dvc_run_geo(const orc_run *rc, const void *refv, const void *tarv, const double *p, int npts,
                    const int *order, const int *neigh, int k, int nproc, int nthreads, const int *geom,
                    const int *size, const int *npp, orc_result *results, double *seconds) {
    for (int i = 0; i < npts; i++) {
        results[i].status = -3; results[i].iters = 0; results[i].obj = 0.0;
        for (int q = 0; q < 12; q++) results[i].par[q] = 0.0;
    }
    std::map<std::tuple<int, int, int>, Searcher *> searchers;
    unsigned int N = (unsigned int)npts;
    if ((unsigned int)nproc > 1 && (unsigned int)nproc < N) N = (unsigned int)nproc;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<const Result *> nres;
    for (unsigned int i = 0; i < N; i++) {
        int n = order[i];
        const auto key = std::make_tuple(geom[n], size[n], npp[n]);
        auto it = searchers.find(key);
        if (it == searchers.end()) {
            orc_run r = *rc;
            r.geom = geom[n]; r.subvol_size = size[n]; r.subvol_npts = npp[n];
            it = searchers.emplace(key, new Searcher(r, make_vol(&r, refv), make_vol(&r, tarv), g_sbi(), nthreads)).first;
        }
        Searcher &S = *it->second;
        nres.clear();
        for (int j = 0; j < k; j++) nres.push_back((const Result *)&results[neigh[(size_t)n * k + j]]);
        Pt sp(p[3 * n], p[3 * n + 1], p[3 * n + 2]);
        int st = S.process_point(sp, nres);
        results[n].iters = S.iters;
        if (st == 0) {
            results[n].status = 0;
            results[n].obj = S.obj_min;
            for (int q = 0; q < rc->ndof; q++) results[n].par[q] = S.par_min[q];
        } else {
            results[n].status = st;
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    for (auto &kv : searchers) delete kv.second;
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return (int)N;
}

/* Given-start passes (no reference counterpart).  mode 1: every searched point from
 * start[12n..]; mode 2: that pass into prop[], then the judge pass into judge[] with
 * the reference's start guess (starting_param, Search.cpp:277-377) over the good
 * neighbours at earlier positions of order[] taken from prop[], and results[] the
 * proposal's result where it is good and its objective <= the judge's + 1e-6, else the
 * judge's (kept[] = 1 / 0, -1 unsearched).  Points are independent within a pass. */
int orc_dvc_run_start(const orc_run *rc, const void *refv, const void *tarv, const double *p, int npts,
                      const int *order, const int *neigh, int k, int nproc, int nthreads, int mode,
                      const double *start, orc_result *prop, orc_result *judge, orc_result *results,
                      int *kept, double *seconds) {
    auto clear = [&](orc_result *r) {
        if (!r) return;
        for (int i = 0; i < npts; i++) {
            r[i].status = -3; r[i].iters = 0; r[i].obj = 0.0;
            for (int q = 0; q < 12; q++) r[i].par[q] = 0.0;
        }
    };
    clear(results); clear(prop); clear(judge);
    if (mode == 2 && (!prop || !judge || !kept)) return -1;
    Searcher S(*rc, make_vol(rc, refv), make_vol(rc, tarv), g_sbi(), nthreads);
    unsigned int N = (unsigned int)npts;
    if ((unsigned int)nproc > 1 && (unsigned int)nproc < N) N = (unsigned int)nproc;
    std::vector<int> pos(npts, npts);
    for (unsigned int i = 0; i < N; i++) pos[order[i]] = (int)i;
    auto store = [&](orc_result *r, int n, int st) {
        r[n].iters = S.iters;
        if (st == 0) {
            r[n].status = 0;
            r[n].obj = S.obj_min;
            for (int q = 0; q < rc->ndof; q++) r[n].par[q] = S.par_min[q];
        } else {
            r[n].status = st;
        }
    };
    auto t0 = std::chrono::steady_clock::now();
    orc_result *first = (mode == 2) ? prop : results;
    std::vector<const Result *> nres;
    for (unsigned int i = 0; i < N; i++) {
        int n = order[i];
        Pt sp(p[3 * n], p[3 * n + 1], p[3 * n + 2]);
        store(first, n, S.process_point(sp, nres, start + (size_t)12 * n));
    }
    if (mode == 2) {
        static const orc_result none = {-3, 0, 0.0, {0}};
        for (unsigned int i = 0; i < N; i++) {
            int n = order[i];
            nres.clear();
            for (int j = 0; j < k; j++) {
                int m = neigh[(size_t)n * k + j];
                nres.push_back(pos[m] < (int)i ? (const Result *)&prop[m] : (const Result *)&none);
            }
            Pt sp(p[3 * n], p[3 * n + 1], p[3 * n + 2]);
            store(judge, n, S.process_point(sp, nres));
        }
        /* selection: where both passes converged and agree (translations within 1e-3
           voxel, objectives within 1e-5 relative) the judge's result (the reference's
           start guess over the given-start results) is the point's; every other point is
           searched again (repair) */
        for (int n = 0; n < npts; n++) {
            if (pos[n] >= (int)N) { kept[n] = -1; continue; }
            const orc_result &a = prop[n], &b = judge[n];
            bool agree = a.status == 0 && b.status == 0;
            if (agree) {
                double du = 0.0;
                for (int q = 0; q < 3; q++) du = std::max(du, fabs(a.par[q] - b.par[q]));
                agree = du <= 1e-3 && fabs(a.obj - b.obj) <= 1e-5 * fabs(b.obj);
            }
            if (agree) {
                results[n] = b;
                kept[n] = 0;
            } else {
                kept[n] = 2;
            }
        }
        /* repair: the disagreeing points in processing order, each from the reference's
           start guess (Search.cpp:277-377) over the final results of its earlier
           neighbours, as the serial loop (dvc.cpp:233-303) runs them */
        for (unsigned int i = 0; i < N; i++) {
            int n = order[i];
            if (kept[n] != 2) continue;
            nres.clear();
            for (int j = 0; j < k; j++) {
                int m = neigh[(size_t)n * k + j];
                nres.push_back(pos[m] < (int)i ? (const Result *)&results[m] : (const Result *)&none);
            }
            results[n].status = -3; results[n].iters = 0; results[n].obj = 0.0;
            for (int q = 0; q < 12; q++) results[n].par[q] = 0.0;
            Pt sp(p[3 * n], p[3 * n + 1], p[3 * n + 2]);
            store(results, n, S.process_point(sp, nres));
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return (int)N;
}

int orc_strain(const double *pts, const double *dis, const int *status, const double *objmin,
               const int *neigh, int k, int n, int ndp, double objt, int refill, int planar,
               double *out_dis, int *out_npts, double *out_swr, double *out_E, double *out_L, double *out_DG) {
    bool yz = planar & 1, zx = planar & 2, xy = planar & 4;
    /* plane auto-detection (strain.cpp:340-365) */
    double mnx = std::numeric_limits<double>::max(), mxx = -std::numeric_limits<double>::max();
    double mny = mnx, mxy = mxx, mnz = mnx, mxz = mxx;
    for (int i = 0; i < n; i++) {
        mnx = std::min(mnx, pts[3 * i]); mxx = std::max(mxx, pts[3 * i]);
        mny = std::min(mny, pts[3 * i + 1]); mxy = std::max(mxy, pts[3 * i + 1]);
        mnz = std::min(mnz, pts[3 * i + 2]); mxz = std::max(mxz, pts[3 * i + 2]);
    }
    if (fabs(mxx - mnx) < 0.001) yz = true;
    if (fabs(mxy - mny) < 0.001) zx = true;
    if (fabs(mxz - mnz) < 0.001) xy = true;
    const int ndp_max = 75;
    for (int i = 0; i < n; i++) {
        const int *nb = &neigh[(size_t)i * k];
        Pt P(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        double distance = pt_dist(P, Pt(0.0, 0.0, 0.0));
        double scale = 1.0 / distance;
        Pt base(scale * P.x, scale * P.y, scale * P.z);
        std::vector<Pt> npos, ndis;
        auto take = [&](int idx) {
            int q = nb[idx];
            npos.push_back(Pt(scale * pts[3 * q], scale * pts[3 * q + 1], scale * pts[3 * q + 2]));
            ndis.push_back(Pt(scale * dis[3 * q], scale * dis[3 * q + 1], scale * dis[3 * q + 2]));
        };
        if (!refill) {                                               /* :424-434 */
            for (int j = 0; j < ndp && j < k; j++)
                if (status[nb[j]] == 0 && objmin[nb[j]] <= objt) take(j);
        } else {                                                     /* :435-448 */
            int indx = 0, np = 0;
            while (np < ndp && indx < ndp_max && indx < k) {
                if (status[nb[indx]] == 0 && objmin[nb[indx]] <= objt) { take(indx); np += 1; }
                indx += 1;
            }
        }
        int ngp = (int)npos.size();
        out_npts[i] = ngp;
        out_swr[i] = (ngp == 0) ? 0.0 : distance * pt_dist(base, npos.back());
        double Xm[75 * 10], dvs[75 * 3];
        for (int j = 0; j < ngp; j++) {                               /* :478-527 */
            double *X = &Xm[j * 10];
            X[0] = 1; X[1] = npos[j].x; X[2] = npos[j].y; X[3] = npos[j].z;
            X[4] = npos[j].x * npos[j].x; X[5] = npos[j].y * npos[j].y; X[6] = npos[j].z * npos[j].z;
            X[7] = npos[j].x * npos[j].y; X[8] = npos[j].y * npos[j].z; X[9] = npos[j].x * npos[j].z;
            dvs[j * 3] = ndis[j].x; dvs[j * 3 + 1] = ndis[j].y; dvs[j * 3 + 2] = ndis[j].z;
            if (yz) { X[1] = 0.0; X[4] = 0.0; X[7] = 0.0; X[9] = 0.0; dvs[j * 3] = 0.0; }
            if (zx) { X[2] = 0.0; X[5] = 0.0; X[7] = 0.0; X[8] = 0.0; dvs[j * 3 + 1] = 0.0; }
            if (xy) { X[3] = 0.0; X[6] = 0.0; X[8] = 0.0; X[9] = 0.0; dvs[j * 3 + 2] = 0.0; }
        }
        double XtX[100];
        for (int a = 0; a < 10; a++) for (int b = 0; b < 10; b++) {
            double s = 0.0;
            for (int j = 0; j < ngp; j++) s += Xm[j * 10 + a] * Xm[j * 10 + b];
            XtX[a * 10 + b] = s;
        }
        double DG[9], displ[3];
        for (int c = 0; c < 3; c++) {                                  /* :531-545 */
            double rhs[10], par[10];
            for (int a = 0; a < 10; a++) {
                double s = 0.0;
                for (int j = 0; j < ngp; j++) s += Xm[j * 10 + a] * dvs[j * 3 + c];
                rhs[a] = s;
            }
            fullpiv_solve(XtX, rhs, 10, par);
            DG[c * 3 + 0] = par[1] + 2 * par[4] * base.x + par[7] * base.y + par[9] * base.z;
            DG[c * 3 + 1] = par[2] + 2 * par[5] * base.y + par[7] * base.x + par[8] * base.z;
            DG[c * 3 + 2] = par[3] + 2 * par[6] * base.z + par[8] * base.y + par[9] * base.x;
            displ[c] = distance * (par[0] + par[1] * base.x + par[2] * base.y + par[3] * base.z
                                   + par[4] * base.x * base.x + par[5] * base.y * base.y + par[6] * base.z * base.z
                                   + par[7] * base.x * base.y + par[8] * base.y * base.z + par[9] * base.x * base.z);
        }
        for (int c = 0; c < 3; c++) out_dis[3 * i + c] = displ[c];
        for (int q = 0; q < 9; q++) out_DG[9 * i + q] = DG[q];
        for (int pass = 0; pass < 2; pass++) {                        /* :552-590 */
            double ST[9];
            for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) {
                double v = DG[a * 3 + b] + DG[b * 3 + a];
                if (pass == 1) {
                    double s = 0.0;
                    for (int c = 0; c < 3; c++) s += DG[c * 3 + a] * DG[c * 3 + b];
                    v += s;
                }
                ST[a * 3 + b] = 0.5 * v;
            }
            double lam[3];
            sym3_eigen(ST, lam);
            if (xy || yz || zx) {
                if (lam[0] == 0.0) { lam[0] = lam[1]; lam[1] = lam[2]; }
                if (lam[1] == 0.0) { lam[1] = lam[2]; }
                lam[2] = 0.0;
            }
            double *o = (pass == 0 ? out_E : out_L) + 9 * i;
            o[0] = ST[0]; o[1] = ST[4]; o[2] = ST[8]; o[3] = ST[1]; o[4] = ST[5]; o[5] = ST[2];
            o[6] = lam[0]; o[7] = lam[1]; o[8] = lam[2];
        }
    }
    return 0;
}

int orc_colpiv_solve(const double *A, const double *b, int n, double *x) { colpiv_solve(A, b, n, x); return 0; }
int orc_fullpiv_solve(const double *A, const double *b, int n, double *x) { fullpiv_solve(A, b, n, x); return 0; }
void orc_sym3_eigen(const double *S, double *lam) { sym3_eigen(S, lam); }

}  /* extern "C" */
