// dvc_kernels.hip -- gfx950 kernels of the DVC per-point correlation path.
//
// Search kernels: one workgroup per ROI point and one launch per dependency level of the
// serial point order (search_kernel), or, for narrow levels, one launch per Gauss-Newton
// evaluation with every point's samples split over several workgroups (slice_kernel,
// then resume_kernel):
//   Search::process_point            Core/Search.cpp:95-192
//   Search::search_pt_setup          Core/Search.cpp:248-275
//   Search::starting_param           Core/Search.cpp:277-377
//   Search::trgrid_global            Core/Search.cpp:934-1043
//   Search::min_Lev_Mar / LM_prep_at Core/Search.cpp:194-246, 490-525
//   FloatingCloud ctors, affine_to   Core/FloatingCloud.cpp:40-128, 148-286
//   Interpolate::tri_cub_Lek etc.    Core/Interpolate.cpp:144-567, Core/Matrix_4d.h:108-133
//   obj_* with residual              Core/ObjectiveFunctions.cpp:24-263
//
// Design (gfx950, wave64):
//   * Volumes stay resident in HBM as uint8/uint16, x fastest.  A sample reads its
//     4x4x4 support as 16 row loads (one dwordx2 for uint8, dwordx3 for uint16).
//   * The tricubic Lekien-Marsden interpolant of the reference (64x64 BI times the 64
//     finite-difference data of a cell) is the tensor-product cubic Hermite
//     interpolant with central-difference slopes.  Inside the interpolation box it is
//     therefore separable Catmull-Rom: 3 x 4-tap filters, value and gradient in ~190
//     FP32 FMAs.  Cells touching the box faces, where the reference copies derivatives
//     from the inner neighbour (Interpolate.cpp:293-335), use the general Hermite form
//     with those clamped derivatives.
//   * Gauss-Newton uses the analytic Jacobian of the residual.  The normal-equation
//     moments of all objectives are entries of the 16x16 matrix V = sum_j v_j v_j^T with
//     v_j = [dt_j/dp (12), t_j - rbar, r_j - rbar, 1, 0]; each wave forms it on the
//     matrix cores with v_mfma_f32_16x16x4_f32 (A = B = the same staged register),
//     flushing the FP32 accumulator into FP64 every 64 samples.
//   * Objective values are FP64 sums over the FP32 sample values; the 12x12 step solve
//     (Cholesky or Gauss-Jordan of the normal equations, with column-pivoted Householder QR as
//     Eigen's colPivHouseholderQr for (near-)singular systems) and all point geometry
//     (cloud, warp, bounding boxes, range checks) are FP64 with the reference's
//     operation order and no fused multiply-add, so statuses follow the reference's
//     comparisons.
#ifdef DVC_EMU
#include "dvc_emu.h"
#else
#include <hip/hip_runtime.h>
#define DVC_DYN_LDS(type, name) extern __shared__ __align__(16) type name[]
#define DVC_SHARED(type, name) \
    using name##_t = type;     \
    __shared__ name##_t name
#endif
#include <math.h>
#include <stddef.h>
#include "dvc_internal.h"

namespace dvc {

#ifdef DVC_EMU
typedef emu::f4 floatx4;
#define DVC_GLOBAL
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) { return v; }
#else
typedef float floatx4 __attribute__((ext_vector_type(4)));
// global (not flat) address space: loads become global_load with an SGPR base
#define DVC_GLOBAL __attribute__((address_space(1)))
// a wave-uniform 64-bit value moved to scalar registers
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
#endif

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ double wave_sum(double x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}
__device__ __forceinline__ double wave_min(double x) {
    for (int o = 32; o > 0; o >>= 1) x = fmin(x, __shfl_xor(x, o, 64));
    return x;
}
__device__ __forceinline__ double wave_max(double x) {
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
    return x;
}

// Totals over the 64 lanes of a fully active wavefront by data-parallel primitives (DPP)
// instead of LDS permutes: pairs, quads, half-rows and rows of 16 by quad_perm and row
// mirrors, then row 15 into rows 1 and 3 (row_bcast:15) and row 31 into rows 2 and 3
// (row_bcast:31), the total landing in lane 63, which is read back as a uniform value.
// Each step moves the two 32-bit halves of the double.  Rows a broadcast step leaves out
// take `old` (0 for sums, the value itself for min / max).
#ifdef DVC_EMU
__device__ __forceinline__ double wave_total_sum(double x) { return wave_sum(x); }
__device__ __forceinline__ double wave_total_min(double x) { return wave_min(x); }
__device__ __forceinline__ double wave_total_max(double x) { return wave_max(x); }
#else
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double src, double old) {
    const long long s = __double_as_longlong(src), o = __double_as_longlong(old);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)s, CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(s >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <class Op>
__device__ __forceinline__ double wave_total(double x, Op op, bool sum) {
    constexpr int kXor1 = 0xB1, kXor2 = 0x4E;                  // quad_perm [1,0,3,2], [2,3,0,1]
    constexpr int kHalfMirror = 0x141, kMirror = 0x140, kBcast15 = 0x142, kBcast31 = 0x143;
    x = op(x, dpp_f64<kXor1, 0xf>(x, x));
    x = op(x, dpp_f64<kXor2, 0xf>(x, x));
    x = op(x, dpp_f64<kHalfMirror, 0xf>(x, x));
    x = op(x, dpp_f64<kMirror, 0xf>(x, x));
    x = op(x, dpp_f64<kBcast15, 0xa>(x, sum ? 0.0 : x));
    x = op(x, dpp_f64<kBcast31, 0xc>(x, sum ? 0.0 : x));
    long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double wave_total_sum(double x) {
    return wave_total(x, [](double u, double v) { return u + v; }, true);
}
__device__ __forceinline__ double wave_total_min(double x) {
    return wave_total(x, [](double u, double v) { return fmin(u, v); }, false);
}
__device__ __forceinline__ double wave_total_max(double x) {
    return wave_total(x, [](double u, double v) { return fmax(u, v); }, false);
}
#endif

// Deterministic block sum of K doubles per thread; results land in out[0..K).
template <int B, int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double *red, double *out) {
    constexpr int NW = B / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < K; q++) v[q] = wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < K; q++) red[w * K + q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0.0;
        for (int ww = 0; ww < NW; ww++) s += red[ww * K + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// Catmull-Rom weights (Keys a = -1/2) and their derivatives for t in [0,1].
__device__ __forceinline__ void cr_weights(float t, float w[4], float d[4]) {
    float t2 = t * t, t3 = t2 * t;
    w[0] = 0.5f * (-t3 + 2.0f * t2 - t);
    w[1] = 0.5f * (3.0f * t3 - 5.0f * t2 + 2.0f);
    w[2] = 0.5f * (-3.0f * t3 + 4.0f * t2 + t);
    w[3] = 0.5f * (t3 - t2);
    d[0] = 0.5f * (-3.0f * t2 + 4.0f * t - 1.0f);
    d[1] = 0.5f * (9.0f * t2 - 10.0f * t);
    d[2] = 0.5f * (-9.0f * t2 + 8.0f * t + 1.0f);
    d[3] = 0.5f * (3.0f * t2 - 2.0f * t);
}

struct VolView {
    const uint8_t *p;
    int W;
    long long WH;
};

template <int BPV>
__device__ __forceinline__ float voxel(const VolView &v, int x, int y, int z) {
    long long i = (long long)z * v.WH + (long long)y * v.W + x;
    if (BPV == 1) return (float)v.p[i];
    return (float)((const uint16_t *)v.p)[i];
}

// Four consecutive voxels x0..x0+3 of row (y, z).  The volume buffer is dword aligned
// and padded by 16 bytes, so the aligned window read stays inside the allocation.
template <int BPV>
__device__ __forceinline__ void load_row(const VolView &v, int x0, int y, int z, float bias, float r[4]) {
    long long i = (long long)z * v.WH + (long long)y * v.W + x0;
    if (BPV == 1) {
        uintptr_t ad = (uintptr_t)(v.p + i);
        const uint32_t *a4 = (const uint32_t *)(ad & ~(uintptr_t)3);
        uint32_t s = (uint32_t)(ad & 3);
        uint32_t lo = a4[0], hi = a4[1];
        uint32_t w = __builtin_amdgcn_alignbyte(hi, lo, s);
        r[0] = (float)(w & 0xffu) - bias;
        r[1] = (float)((w >> 8) & 0xffu) - bias;
        r[2] = (float)((w >> 16) & 0xffu) - bias;
        r[3] = (float)(w >> 24) - bias;
    } else {
        uintptr_t ad = (uintptr_t)(v.p + 2 * i);
        const uint32_t *a4 = (const uint32_t *)(ad & ~(uintptr_t)3);
        uint32_t s = (uint32_t)(ad & 3);
        uint32_t d0 = a4[0], d1 = a4[1], d2 = a4[2];
        uint32_t w0 = __builtin_amdgcn_alignbyte(d1, d0, s);
        uint32_t w1 = __builtin_amdgcn_alignbyte(d2, d1, s);
        r[0] = (float)(w0 & 0xffffu) - bias;
        r[1] = (float)(w0 >> 16) - bias;
        r[2] = (float)(w1 & 0xffffu) - bias;
        r[3] = (float)(w1 >> 16) - bias;
    }
}

// Interpolation box: base voxel (truncated est_box.min) and size.
struct BoxView {
    int b[3];
    int n[3];
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Row gathers of one point's box.  fast: the rows of the box lie within 2 GiB of `base`
// (the first voxel of the box's first row and slice, dword aligned) and row strides are
// whole dwords, so a sample's 16 rows share one 32-bit offset and byte shift and differ
// only by wave-uniform row bases.  Otherwise rows are addressed with 64-bit arithmetic.
struct Gather {
    const DVC_GLOBAL uint8_t *base;
    uint32_t Wb, WHb;                // row and slice strides in bytes
    bool fast;
};

template <int BPV>
__device__ __forceinline__ Gather make_gather(const VolView &v, const BoxView &bx, bool fast) {
    Gather g;
    g.fast = fast;
    g.Wb = (uint32_t)(v.W * BPV);
    g.WHb = (uint32_t)(v.WH * BPV);
    uint64_t b = (uint64_t)(uintptr_t)(v.p + BPV * ((long long)bx.b[2] * v.WH + (long long)bx.b[1] * v.W));
    g.base = (const DVC_GLOBAL uint8_t *)(uintptr_t)uniform_u64(b);
    return g;
}

// Cubic Hermite basis on [0,1]: h[a][q] is the weight of the value (q=0) or slope (q=1)
// at node a; dh its derivative.
__device__ __forceinline__ void hermite(float t, float h[2][2], float dh[2][2]) {
    float t2 = t * t, t3 = t2 * t;
    h[0][0] = 2.0f * t3 - 3.0f * t2 + 1.0f;
    h[1][0] = -2.0f * t3 + 3.0f * t2;
    h[0][1] = t3 - 2.0f * t2 + t;
    h[1][1] = t3 - t2;
    dh[0][0] = 6.0f * t2 - 6.0f * t;
    dh[1][0] = -6.0f * t2 + 6.0f * t;
    dh[0][1] = 3.0f * t2 - 4.0f * t + 1.0f;
    dh[1][1] = 3.0f * t2 - 2.0f * t;
}

// Tricubic value (and gradient) of a cell with a corner on the box faces: the
// 8 Matrix_4d quantities (f, fx, fy, fz, fxy, fxz, fyz, fxyz; Matrix_4d.h:105-133) at
// each corner, derivatives taken at the corner clamped into the box interior as the
// face copies of Interpolate.cpp:293-335 leave them, combined with the Hermite basis.
template <int BPV>
__device__ float tricubic_faces(const VolView &v, const BoxView &bx, int ix, int iy, int iz,
                                float tx, float ty, float tz, float bias, bool want_grad,
                                float &gx, float &gy, float &gz) {
    float hx[2][2], hy[2][2], hz[2][2], dhx[2][2], dhy[2][2], dhz[2][2];
    hermite(tx, hx, dhx);
    hermite(ty, hy, dhy);
    hermite(tz, hz, dhz);
    float val = 0.0f;
    gx = gy = gz = 0.0f;
    for (int c = 0; c < 8; c++) {
        int a0 = c & 1, a1 = (c >> 1) & 1, a2 = (c >> 2) & 1;
        int vx = ix + a0, vy = iy + a1, vz = iz + a2;
        int qx = clampi(vx, bx.b[0] + 1, bx.b[0] + bx.n[0] - 2);
        int qy = clampi(vy, bx.b[1] + 1, bx.b[1] + bx.n[1] - 2);
        int qz = clampi(vz, bx.b[2] + 1, bx.b[2] + bx.n[2] - 2);
#define F(dx, dy, dz) (voxel<BPV>(v, qx + (dx), qy + (dy), qz + (dz)) - bias)
        float D[8];
        D[0] = voxel<BPV>(v, vx, vy, vz) - bias;
        D[1] = (F(1, 0, 0) - F(-1, 0, 0)) / 2.0f;
        D[2] = (F(0, 1, 0) - F(0, -1, 0)) / 2.0f;
        D[3] = (F(0, 0, 1) - F(0, 0, -1)) / 2.0f;
        D[4] = (F(1, 1, 0) - F(-1, 1, 0) + F(-1, -1, 0) - F(1, -1, 0)) / 4.0f;
        D[5] = (F(1, 0, 1) - F(-1, 0, 1) + F(-1, 0, -1) - F(1, 0, -1)) / 4.0f;
        D[6] = (F(0, 1, 1) - F(0, -1, 1) + F(0, -1, -1) - F(0, 1, -1)) / 4.0f;
        float t1 = (F(1, 1, -1) - F(-1, 1, -1) + F(-1, -1, -1) - F(1, -1, -1)) / 4.0f;
        float t2 = (F(1, 1, 1) - F(-1, 1, 1) + F(-1, -1, 1) - F(1, -1, 1)) / 4.0f;
        D[7] = (t2 - t1) / 2.0f;
#undef F
        const int qd[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
        for (int q = 0; q < 8; q++) {
            float wx = hx[a0][qd[q][0]], wy = hy[a1][qd[q][1]], wz = hz[a2][qd[q][2]];
            val += wx * wy * wz * D[q];
            if (want_grad) {
                gx += dhx[a0][qd[q][0]] * wy * wz * D[q];
                gy += wx * dhy[a1][qd[q][1]] * wz * D[q];
                gz += wx * wy * dhz[a2][qd[q][2]] * D[q];
            }
        }
    }
    return val;
}

// Interpolated value minus `bias` at (x, y, z) with the reference's cell split
// (Point.cpp:57-66: truncated integer part, remainder) and the box semantics above.
// method: 0 nearest, 1 trilinear, 2 tricubic.  An integer bias near the local mean is
// subtracted from the voxels exactly before the weighted sums, which keeps the FP32
// rounding proportional to the local contrast instead of the full intensity range.
// Callers guarantee every sample lies in a valid cell; the clamps below only keep
// stray (NaN) coordinates inside the box.
template <int BPV>
__device__ __forceinline__ float interp_at(const VolView &v, const BoxView &bx, double x, double y,
                                           double z, int method, float bias, bool want_grad,
                                           float &gx, float &gy, float &gz) {
    int ix = (int)x, iy = (int)y, iz = (int)z;
    float tx = (float)(x - (double)ix), ty = (float)(y - (double)iy), tz = (float)(z - (double)iz);
    int cx = ix - bx.b[0], cy = iy - bx.b[1], cz = iz - bx.b[2];
    gx = gy = gz = 0.0f;
    if (method == 0) {
        ix = bx.b[0] + clampi(cx, 0, bx.n[0] - 1);
        iy = bx.b[1] + clampi(cy, 0, bx.n[1] - 1);
        iz = bx.b[2] + clampi(cz, 0, bx.n[2] - 1);
        return voxel<BPV>(v, ix, iy, iz) - bias;
    }
    if (cx < 0 || cx > bx.n[0] - 2 || cy < 0 || cy > bx.n[1] - 2 || cz < 0 || cz > bx.n[2] - 2 ||
        !(tx >= 0.0f && tx <= 1.0f && ty >= 0.0f && ty <= 1.0f && tz >= 0.0f && tz <= 1.0f)) {
        cx = clampi(cx, 0, bx.n[0] - 2);
        cy = clampi(cy, 0, bx.n[1] - 2);
        cz = clampi(cz, 0, bx.n[2] - 2);
        tx = fminf(fmaxf(tx, 0.0f), 1.0f);
        ty = fminf(fmaxf(ty, 0.0f), 1.0f);
        tz = fminf(fmaxf(tz, 0.0f), 1.0f);
        if (!(tx == tx)) tx = 0.0f;
        if (!(ty == ty)) ty = 0.0f;
        if (!(tz == tz)) tz = 0.0f;
        ix = bx.b[0] + cx; iy = bx.b[1] + cy; iz = bx.b[2] + cz;
    }
    if (method == 1) {
        // tri_lin (Interpolate.cpp:429-470)
        float c000 = voxel<BPV>(v, ix, iy, iz) - bias, c100 = voxel<BPV>(v, ix + 1, iy, iz) - bias;
        float c010 = voxel<BPV>(v, ix, iy + 1, iz) - bias, c110 = voxel<BPV>(v, ix + 1, iy + 1, iz) - bias;
        float c001 = voxel<BPV>(v, ix, iy, iz + 1) - bias, c101 = voxel<BPV>(v, ix + 1, iy, iz + 1) - bias;
        float c011 = voxel<BPV>(v, ix, iy + 1, iz + 1) - bias, c111 = voxel<BPV>(v, ix + 1, iy + 1, iz + 1) - bias;
        float x0 = 1.0f - tx, y0 = 1.0f - ty, z0 = 1.0f - tz;
        float a00 = c000 * x0 + c100 * tx, a10 = c010 * x0 + c110 * tx;
        float a01 = c001 * x0 + c101 * tx, a11 = c011 * x0 + c111 * tx;
        float b0 = a00 * y0 + a10 * ty, b1 = a01 * y0 + a11 * ty;
        if (want_grad) {
            float e00 = c100 - c000, e10 = c110 - c010, e01 = c101 - c001, e11 = c111 - c011;
            gx = (e00 * y0 + e10 * ty) * z0 + (e01 * y0 + e11 * ty) * tz;
            gy = (a10 - a00) * z0 + (a11 - a01) * tz;
            gz = b1 - b0;
        }
        return b0 * z0 + b1 * tz;
    }
    cx = ix - bx.b[0]; cy = iy - bx.b[1]; cz = iz - bx.b[2];
    if (cx < 1 || cx > bx.n[0] - 3 || cy < 1 || cy > bx.n[1] - 3 || cz < 1 || cz > bx.n[2] - 3)
        return tricubic_faces<BPV>(v, bx, ix, iy, iz, tx, ty, tz, bias, want_grad, gx, gy, gz);
    float wx[4], dx[4], wy[4], dy[4], wz[4], dz[4];
    cr_weights(tx, wx, dx);
    cr_weights(ty, wy, dy);
    cr_weights(tz, wz, dz);
    float val = 0.0f;
#pragma unroll
    for (int kz = 0; kz < 4; kz++) {
        float pv = 0.0f, px = 0.0f, py = 0.0f;
#pragma unroll
        for (int ky = 0; ky < 4; ky++) {
            float r[4];
            load_row<BPV>(v, ix - 1, iy - 1 + ky, iz - 1 + kz, bias, r);
            float rv = r[0] * wx[0] + r[1] * wx[1] + r[2] * wx[2] + r[3] * wx[3];
            pv += wy[ky] * rv;
            if (want_grad) {
                float rd = r[0] * dx[0] + r[1] * dx[1] + r[2] * dx[2] + r[3] * dx[3];
                px += wy[ky] * rd;
                py += dy[ky] * rv;
            }
        }
        val += wz[kz] * pv;
        if (want_grad) {
            gx += wz[kz] * px;
            gy += wz[kz] * py;
            gz += dz[kz] * pv;
        }
    }
    return val;
}

// A tricubic sample split into address, load and arithmetic stages, so that a thread
// can have the 16 row loads of several samples in flight before it needs any of them.
// The arithmetic is the interior path of interp_at; cells on the box faces take
// tricubic_faces at evaluation time (their rows are loaded at the nearest interior
// cell, always inside the box, and not used).
// Samples per thread whose volume loads are in flight together in the sample passes.
#ifndef DVC_INFLIGHT
#define DVC_INFLIGHT 1
#endif
constexpr int kInflight = DVC_INFLIGHT;

template <int BPV>
struct Tap {
    static constexpr int K = BPV == 1 ? 2 : 3;   // dwords per row window
    int ix, iy, iz;
    float tx, ty, tz;
    bool face;
    uint32_t shift;                              // 2-bit byte shift of each row window
    uint32_t raw[16][K];
};

template <int BPV>
__device__ __forceinline__ void tap_locate(const BoxView &bx, double x, double y, double z, Tap<BPV> &t) {
    int ix = (int)x, iy = (int)y, iz = (int)z;
    float tx = (float)(x - (double)ix), ty = (float)(y - (double)iy), tz = (float)(z - (double)iz);
    int cx = ix - bx.b[0], cy = iy - bx.b[1], cz = iz - bx.b[2];
    if (cx < 0 || cx > bx.n[0] - 2 || cy < 0 || cy > bx.n[1] - 2 || cz < 0 || cz > bx.n[2] - 2 ||
        !(tx >= 0.0f && tx <= 1.0f && ty >= 0.0f && ty <= 1.0f && tz >= 0.0f && tz <= 1.0f)) {
        cx = clampi(cx, 0, bx.n[0] - 2);
        cy = clampi(cy, 0, bx.n[1] - 2);
        cz = clampi(cz, 0, bx.n[2] - 2);
        tx = fminf(fmaxf(tx, 0.0f), 1.0f);
        ty = fminf(fmaxf(ty, 0.0f), 1.0f);
        tz = fminf(fmaxf(tz, 0.0f), 1.0f);
        if (!(tx == tx)) tx = 0.0f;
        if (!(ty == ty)) ty = 0.0f;
        if (!(tz == tz)) tz = 0.0f;
    }
    t.ix = bx.b[0] + cx; t.iy = bx.b[1] + cy; t.iz = bx.b[2] + cz;
    t.tx = tx; t.ty = ty; t.tz = tz;
    t.face = cx < 1 || cx > bx.n[0] - 3 || cy < 1 || cy > bx.n[1] - 3 || cz < 1 || cz > bx.n[2] - 3;
}

template <int BPV>
__device__ __forceinline__ void tap_load(const VolView &v, const Gather &g, const BoxView &bx, Tap<BPV> &t) {
    // rows of the nearest interior cell: inside the box for every cell of the box
    const int lx = clampi(t.ix, bx.b[0] + 1, bx.b[0] + bx.n[0] - 3);
    const int ly = clampi(t.iy, bx.b[1] + 1, bx.b[1] + bx.n[1] - 3);
    const int lz = clampi(t.iz, bx.b[2] + 1, bx.b[2] + bx.n[2] - 3);
    if (g.fast) {
        const uint32_t o = (uint32_t)(lz - 1 - bx.b[2]) * g.WHb + (uint32_t)(ly - 1 - bx.b[1]) * g.Wb +
                           (uint32_t)(lx - 1) * (uint32_t)BPV;
        const uint32_t sh = o & 3u, vo = o & ~3u;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int ky = r & 3, kz = r >> 2;
            const DVC_GLOBAL uint8_t *row = g.base + (kz * g.WHb + ky * g.Wb);
            const DVC_GLOBAL uint32_t *a4 = (const DVC_GLOBAL uint32_t *)(row + vo);
#pragma unroll
            for (int q = 0; q < Tap<BPV>::K; q++) t.raw[r][q] = a4[q];
        }
        t.shift = sh * 0x55555555u;
        return;
    }
    uint32_t sh = 0;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int ky = r & 3, kz = r >> 2;
        long long i = (long long)(lz - 1 + kz) * v.WH + (long long)(ly - 1 + ky) * v.W + (lx - 1);
        uintptr_t ad = (uintptr_t)(v.p + BPV * i);
        const uint32_t *a4 = (const uint32_t *)(ad & ~(uintptr_t)3);
        sh |= (uint32_t)(ad & 3) << (2 * r);
#pragma unroll
        for (int q = 0; q < Tap<BPV>::K; q++) t.raw[r][q] = a4[q];
    }
    t.shift = sh;
}

template <int BPV>
__device__ __forceinline__ float tap_eval(const VolView &v, const BoxView &bx, const Tap<BPV> &t, float bias,
                                          bool want_grad, float &gx, float &gy, float &gz) {
    gx = gy = gz = 0.0f;
    if (t.face)
        return tricubic_faces<BPV>(v, bx, t.ix, t.iy, t.iz, t.tx, t.ty, t.tz, bias, want_grad, gx, gy, gz);
    float wx[4], dx[4], wy[4], dy[4], wz[4], dz[4];
    cr_weights(t.tx, wx, dx);
    cr_weights(t.ty, wy, dy);
    cr_weights(t.tz, wz, dz);
    float val = 0.0f;
#pragma unroll
    for (int kz = 0; kz < 4; kz++) {
        float pv = 0.0f, px = 0.0f, py = 0.0f;
#pragma unroll
        for (int ky = 0; ky < 4; ky++) {
            const int rr = kz * 4 + ky;
            const uint32_t s = (t.shift >> (2 * rr)) & 3u;
            float r[4];
            if (BPV == 1) {
                uint32_t w = __builtin_amdgcn_alignbyte(t.raw[rr][1], t.raw[rr][0], s);
                r[0] = (float)(w & 0xffu) - bias;
                r[1] = (float)((w >> 8) & 0xffu) - bias;
                r[2] = (float)((w >> 16) & 0xffu) - bias;
                r[3] = (float)(w >> 24) - bias;
            } else {
                uint32_t w0 = __builtin_amdgcn_alignbyte(t.raw[rr][1], t.raw[rr][0], s);
                uint32_t w1 = __builtin_amdgcn_alignbyte(t.raw[rr][Tap<BPV>::K - 1], t.raw[rr][1], s);
                r[0] = (float)(w0 & 0xffffu) - bias;
                r[1] = (float)(w0 >> 16) - bias;
                r[2] = (float)(w1 & 0xffffu) - bias;
                r[3] = (float)(w1 >> 16) - bias;
            }
            float rv = r[0] * wx[0] + r[1] * wx[1] + r[2] * wx[2] + r[3] * wx[3];
            pv += wy[ky] * rv;
            if (want_grad) {
                float rd = r[0] * dx[0] + r[1] * dx[1] + r[2] * dx[2] + r[3] * dx[3];
                px += wy[ky] * rd;
                py += dy[ky] * rv;
            }
        }
        val += wz[kz] * pv;
        if (want_grad) {
            gx += wz[kz] * px;
            gy += wz[kz] * py;
            gz += dz[kz] * pv;
        }
    }
    return val;
}

// ------------------------------------------------------------- point state

// State of one point's search, in LDS.  Everything before nb_sum is the point's state
// between evaluations: slice launches keep it in HBM from one launch to the next.
struct PtShared {
    int status, iters, fail, stop, pfail;
    int b0[3], box_ok;
    int it, done, last, pad_;        // next Gauss-Newton iteration; search finished; last slice of a launch
    double obj_old;                  // objective of the previous iteration (Search.cpp:217-230)
    double sp[3], sref[3], smin[3], inc[3], sbb_mn[3], sbb_mx[3];
    double act_mn[3], act_mx[3];
    double par[12], upd[12], evalpar[12];
    double u[3], mref[3], A[9], Sm[9];
    float M[3][9], Af[9];
    double bb[6];
    double rmean, R2, Rn2, Src;
    float fbias;                     // integer intensity subtracted before interpolation sums
    double rbmean;                   // mean of the biased reference samples
    double tmean, N2, Str, T2, Strr;
    double obj, omin;
    int gmin[3];
    double sums[8];
    double bbr[6];                   // moving-cloud bounding box of a sample pass (min xyz, max xyz)
    // start guess inputs, not part of the carried state
    double nb_sum[3];                // translations of the good earlier neighbours, summed
    int nb_good;                     // their number
    double dA[3][9];                 // d A / d angle of the current warp (set_warp to set_warp_rows)
};
constexpr int kStateBytes = (int)offsetof(PtShared, nb_sum);
static_assert(kStateBytes % 8 == 0, "carried point state is copied in 8-byte words");

// Stable cloud sample j (FloatingCloud.cpp:84-128 cube grid, z outer / x inner;
// FloatingCloud.cpp:71-76 sphere = centre + aspect*rel).
__device__ __forceinline__ void stable_pt(const SearchArgs &a, const PtShared &S, int j,
                                          double &x, double &y, double &z) {
#pragma clang fp contract(off)
    if (a.geom == DVC_GEOM_CUBE) {
        int g = a.g;
        int kx = j % g, ky = (j / g) % g, kz = j / (g * g);
        x = S.smin[0] + (double)kx * S.inc[0];
        y = S.smin[1] + (double)ky * S.inc[1];
        z = S.smin[2] + (double)kz * S.inc[2];
    } else {
        x = S.sp[0] + a.sph[3 * j + 0];
        y = S.sp[1] + a.sph[3 * j + 1];
        z = S.sp[2] + a.sph[3 * j + 2];
    }
}

// Sphere offsets of sample j, loaded one loop iteration ahead of their use in the sample
// passes (zero for cubes, whose samples are computed from j).
struct SphOff {
    double x, y, z;
};

__device__ __forceinline__ SphOff sph_off(const SearchArgs &a, int j, bool valid) {
    SphOff o = {0.0, 0.0, 0.0};
    if (a.geom == DVC_GEOM_SPHERE && valid) {
        o.x = a.sph[3 * j + 0];
        o.y = a.sph[3 * j + 1];
        o.z = a.sph[3 * j + 2];
    }
    return o;
}

// stable_pt with the sphere offsets already loaded (same arithmetic)
__device__ __forceinline__ void stable_pt_off(const SearchArgs &a, const PtShared &S, int j, const SphOff &o,
                                              double &x, double &y, double &z) {
#pragma clang fp contract(off)
    if (a.geom == DVC_GEOM_CUBE) {
        int g = a.g;
        int kx = j % g, ky = (j / g) % g, kz = j / (g * g);
        x = S.smin[0] + (double)kx * S.inc[0];
        y = S.smin[1] + (double)ky * S.inc[1];
        z = S.smin[2] + (double)kz * S.inc[2];
    } else {
        x = S.sp[0] + o.x;
        y = S.sp[1] + o.y;
        z = S.sp[2] + o.z;
    }
}

// FloatingCloud::affine_to for one point (FloatingCloud.cpp:236-286 with rotate_ptvect
// :174-204 and strain_ptvect :148-172), same operation order, no contraction.
__device__ __forceinline__ void warp_pt(const PtShared &S, int ndof, double xs, double ys, double zs,
                                        double &xo, double &yo, double &zo) {
#pragma clang fp contract(off)
    double x = xs + S.u[0], y = ys + S.u[1], z = zs + S.u[2];
    if (ndof > 3) {
        double rx = x - S.mref[0], ry = y - S.mref[1], rz = z - S.mref[2];
        double ox = S.A[0] * rx + S.A[1] * ry + S.A[2] * rz;
        double oy = S.A[3] * rx + S.A[4] * ry + S.A[5] * rz;
        double oz = S.A[6] * rx + S.A[7] * ry + S.A[8] * rz;
        double dx = ox - rx, dy = oy - ry, dz = oz - rz;
        x = x + dx; y = y + dy; z = z + dz;
        if (ndof > 6) {
            rx = x - S.mref[0]; ry = y - S.mref[1]; rz = z - S.mref[2];
            double sx = S.Sm[0] * rx + S.Sm[1] * ry + S.Sm[2] * rz;
            double sy = S.Sm[3] * rx + S.Sm[4] * ry + S.Sm[5] * rz;
            double sz = S.Sm[6] * rx + S.Sm[7] * ry + S.Sm[8] * rz;
            x = x + sx; y = y + sy; z = z + sz;
        }
    }
    xo = x; yo = y; zo = z;
}

// Interpolate::center_on (Interpolate.cpp:569-574, BoundBox.cpp:75-85) applied to the
// nominal boxes, then the volume extent check.  Thread 0 only.
__device__ __forceinline__ void place_box(const SearchArgs &a, PtShared &S, double cx, double cy, double cz) {
#pragma clang fp contract(off)
    double c[3] = {cx, cy, cz};
    S.box_ok = 1;
    int dims[3] = {a.W, a.H, a.T};
    for (int q = 0; q < 3; q++) {
        double d = c[q] - a.est_cn[q];
        double mn = a.est_mn[q] + d;
        double da = c[q] - a.act_cn[q];
        S.act_mn[q] = a.act_mn[q] + da;
        S.act_mx[q] = a.act_mx[q] + da;
        int b = (int)(mn / 1);
        S.b0[q] = b;
        if (!(mn == mn) || b < 0 || (long long)b + a.box[q] > (long long)dims[q]) S.box_ok = 0;
    }
}

// Range rule of tri_cub_Lek / tri_lin / nearest: act_box.contains(bbox)
// (Interpolate.cpp:474-481, BoundBox.cpp:87-99), box inside the volume, and the cells
// of the extreme samples inside the box (the reference indexes past kern_4d there).
__device__ __forceinline__ bool range_ok(const SearchArgs &a, const PtShared &S, const double bb[6]) {
    for (int q = 0; q < 6; q++)
        if (!(bb[q] == bb[q]) || fabs(bb[q]) > 1.0e15) return false;
    if (bb[0] < S.act_mn[0] || bb[1] < S.act_mn[1] || bb[2] < S.act_mn[2] ||
        bb[3] > S.act_mx[0] || bb[4] > S.act_mx[1] || bb[5] > S.act_mx[2])
        return false;
    if (!S.box_ok) return false;
    int lim = (a.interp == DVC_INTERP_NEAREST) ? 1 : 2;
    for (int q = 0; q < 3; q++) {
        int lo = (int)(bb[q] / 1) - S.b0[q];
        int hi = (int)(bb[3 + q] / 1) - S.b0[q];
        if (lo < 0 || hi > a.box[q] - lim) return false;
    }
    return true;
}

// Rotation matrix of SearchParams::set_param_vect (SearchParams.cpp:77-87) and its
// derivatives with respect to (phi, the, psi).
// Cross-lane moves of one wavefront.  rdlane: value of lane l (l uniform);
// shfl_from: value of lane src (src per lane).
#ifdef DVC_EMU
__device__ __forceinline__ double rdlane(double v, int l) { return emu::shfl_idx(v, l); }
__device__ __forceinline__ double shfl_from(double v, int src) { return emu::shfl_idx(v, src); }
__device__ __forceinline__ int shfl_from(int v, int src) { return (int)emu::shfl_idx((double)v, src); }
__device__ __forceinline__ int rdlane_i(int v, int l) { return (int)emu::shfl_idx((double)v, l); }
#else
__device__ __forceinline__ int rdlane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ double rdlane(double v, int l) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double shfl_from(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ int shfl_from(int v, int src) { return __shfl(v, src, 64); }
#endif

// trig: cos/sin of phi, the, psi (cf sf ct st cp sp)
__device__ void rotation(const double trig[6], double A[9], double dA[3][9]) {
    const double cf = trig[0], sf = trig[1], ct = trig[2], st = trig[3], cp = trig[4], sp = trig[5];
    A[0] = ct * cf; A[1] = ct * sf; A[2] = -st;
    A[3] = sp * st * cf - cp * sf; A[4] = sp * st * sf + cp * cf; A[5] = ct * sp;
    A[6] = cp * st * cf + sp * sf; A[7] = cp * st * sf - sp * cf; A[8] = ct * cp;
    double *P = dA[0];
    P[0] = -ct * sf; P[1] = ct * cf; P[2] = 0;
    P[3] = -sp * st * sf - cp * cf; P[4] = sp * st * cf - cp * sf; P[5] = 0;
    P[6] = -cp * st * sf + sp * cf; P[7] = cp * st * cf + sp * sf; P[8] = 0;
    double *T = dA[1];
    T[0] = -st * cf; T[1] = -st * sf; T[2] = -ct;
    T[3] = sp * ct * cf; T[4] = sp * ct * sf; T[5] = -st * sp;
    T[6] = cp * ct * cf; T[7] = cp * ct * sf; T[8] = -st * cp;
    double *Q = dA[2];
    Q[0] = 0; Q[1] = 0; Q[2] = 0;
    Q[3] = cp * st * cf + sp * sf; Q[4] = cp * st * sf - sp * cf; Q[5] = ct * cp;
    Q[6] = -sp * st * cf + cp * sf; Q[7] = -sp * st * sf - cp * cf; Q[8] = -ct * sp;
}

// Warp state for parameter vector evalpar (thread 0), given the cosines and sines of
// its three angles (trig): translation, rotation A, strain Sm and, in S.dA, the rotation's
// derivatives for set_warp_rows.  SearchParams.cpp:53-107.
__device__ __forceinline__ void set_warp(PtShared &S, int ndof, const double trig[6]) {
#pragma clang fp contract(off)
    for (int q = 0; q < 3; q++) {
        S.u[q] = S.evalpar[q];
        S.mref[q] = S.sref[q] + S.u[q];
    }
    if (ndof > 3) {
        rotation(trig, S.A, S.dA);
    } else {
        for (int q = 0; q < 9; q++) S.A[q] = (q % 4 == 0) ? 1.0 : 0.0;
        for (int r = 0; r < 3; r++) for (int q = 0; q < 9; q++) S.dA[r][q] = 0.0;
    }
    for (int q = 0; q < 9; q++) S.Sm[q] = 0.0;
    if (ndof > 6) {
        const double *v = S.evalpar;
        S.Sm[0] = v[6];  S.Sm[1] = v[9];  S.Sm[2] = v[11];
        S.Sm[3] = v[9];  S.Sm[4] = v[7];  S.Sm[5] = v[10];
        S.Sm[6] = v[11]; S.Sm[7] = v[10]; S.Sm[8] = v[8];
    }
}

// The FP32 factors of the Jacobian rows after set_warp, one entry per lane (lanes 0-26
// of wavefront 0): d x'/d angle = (I+S) dA rel, formed in FP32, and A in FP32.
__device__ __forceinline__ void set_warp_rows(PtShared &S, int ndof, int lane) {
    if (lane < 27) {
        const int r = lane / 9, q = lane % 9, i = q / 3, j = q % 3;
        float s = (float)S.dA[r][q];
        if (ndof > 6)
            for (int l = 0; l < 3; l++) s = __builtin_fmaf((float)S.Sm[i * 3 + l], (float)S.dA[r][l * 3 + j], s);
        S.M[r][q] = s;
    }
    if (lane < 9) S.Af[lane] = (float)S.A[lane];
}

// ------------------------------------------------------------- evaluation

// One objective evaluation of the moving cloud at S.evalpar (first ndof entries),
// Search::obj_val_at (Search.cpp:448-488), in three parts: eval_begin (warp and the
// checks that need no samples), eval_samples (the sums over a range of samples, block
// totals in S.sums[0..2], S.bbr and, with need_j, the Jacobian moments V) and eval_close
// (the range rule of the moving bounding box, then the objective).  A workgroup runs all
// three over every sample; slice launches split eval_samples over several workgroups.

// Warp of S.evalpar and the checks before any sample: non-finite parameters and, with
// 3 DOF, the range rule of the shifted stable box.  False (uniformly) on Range_Fail.
template <int B>
__device__ __forceinline__ bool eval_begin(const SearchArgs &a, PtShared &S, int ndof) {
    const int tid = threadIdx.x, lane = tid & 63;
    if (tid < 64) {
        // cosine and sine of the three angles, one angle per lane of wavefront 0
        double sv = 0.0, cv = 1.0;
        if (ndof > 3 && lane < 3) sincos(S.evalpar[3 + lane], &sv, &cv);
        const double trig[6] = {rdlane(cv, 0), rdlane(sv, 0), rdlane(cv, 1), rdlane(sv, 1), rdlane(cv, 2),
                                rdlane(sv, 2)};
        if (tid == 0) {
#pragma clang fp contract(off)
            set_warp(S, ndof, trig);
            // non-finite parameters (singular normal matrix) are a range failure
            int pfail = 0;
            for (int q = 0; q < ndof; q++)
                if (!isfinite(S.evalpar[q])) pfail = 1;
            // Range rule.  With 3 DOF the moving bounding box is the stable one shifted;
            // with more it is accumulated in the sample pass, which then runs
            // speculatively: every sample is clamped into the interpolation box (inside
            // the volume), and a failed range check discards the sums.
            int fail = pfail;
            if (!pfail) {
                if (ndof <= 3) {
                    for (int q = 0; q < 3; q++) { S.bb[q] = S.sbb_mn[q] + S.u[q]; S.bb[3 + q] = S.sbb_mx[q] + S.u[q]; }
                    fail = range_ok(a, S, S.bb) ? 0 : 1;
                } else {
                    fail = S.box_ok ? 0 : 1;
                }
            }
            S.pfail = pfail;
            S.fail = fail;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        set_warp_rows(S, ndof, lane);
    }
    __syncthreads();
    return S.fail == 0;
}

// Sums over a workgroup's samples of the moving cloud at the warp of S: block totals of
// sum t, and of t^2 and t r (centred objectives) or (t - r)^2 and |t - r| (SSD, SAD) in
// S.sums[0..2]; with more than 3 DOF the samples' bounding box in S.bbr; with need_j the
// moments V = sum v v^T (16x16) of the block's samples.  Samples go to wavefronts in
// chunks of 64 consecutive samples (one per lane): the workgroup takes chunks c0,
// c0 + cstride, c0 + 2 cstride, ... (c0 = 0, cstride = 1: every sample, j = tid + it*B).
template <int BPV, int B, int U = kInflight>
__device__ __forceinline__ void eval_samples(const SearchArgs &a, PtShared &S, int ndof, bool need_j,
                                             const float *rsrc, float *stage, double *red, double *V, int c0,
                                             int cstride) {
    constexpr int NW = B / 64;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const bool merged = ndof > 3;
    const VolView vv = {a.vol[1], a.W, (long long)a.W * a.H};
    const BoxView bx = {{S.b0[0], S.b0[1], S.b0[2]}, {a.box[0], a.box[1], a.box[2]}};
    const Gather gv = make_gather<BPV>(vv, bx, a.gather32 != 0);
    const double rbmean = S.rbmean;
    const float fbias = S.fbias;
    const int method = a.interp;
    const bool jrow = need_j;
    double acc64[4] = {0.0, 0.0, 0.0, 0.0};
    float *wst = stage + wid * 64 * 17;
    // one pass of sums: t, then t^2 and t r (zero-mean and normalised objectives) or
    // (t - r)^2 and |t - r| (SSD, SAD); products of FP32 values are exact in FP64
    const bool centred = a.obj == DVC_OBJ_ZSSD || a.obj == DVC_OBJ_NSSD || a.obj == DVC_OBJ_ZNSSD;
    double s_tc = 0.0, s_x1 = 0.0, s_x2 = 0.0;
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};

    // moving-cloud position of sample j (and its stable position)
    auto position = [&](int j, const SphOff &o, double &xs, double &ys, double &zs, double &x, double &y,
                        double &z) {
        stable_pt_off(a, S, j, o, xs, ys, zs);
        warp_pt(S, ndof, xs, ys, zs, x, y, z);
        if (merged) {
            mn[0] = fmin(mn[0], x); mn[1] = fmin(mn[1], y); mn[2] = fmin(mn[2], z);
            mx[0] = fmax(mx[0], x); mx[1] = fmax(mx[1], y); mx[2] = fmax(mx[2], z);
        }
    };
    // accumulate one sample (every lane of the wave calls it; ok = false adds nothing)
    auto accumulate = [&](bool ok, float rc, float t, float gx, float gy, float gz, double xs, double ys,
                          double zs) {
        float v[16];
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = 0.0f;
        if (ok) {
            float tc = (float)((double)t - rbmean);
            const double tcd = (double)tc, rcd = (double)rc;
            s_tc += tcd;
            if (centred) {
                s_x1 += tcd * tcd;
                s_x2 += tcd * rcd;
            } else {
                const double d = tcd - rcd;
                s_x1 += d * d;
                s_x2 += fabs(d);
            }
            if (jrow) {
                v[0] = gx; v[1] = gy; v[2] = gz;
                if (ndof > 3) {
                    float rx = (float)(xs - S.sref[0]), ry = (float)(ys - S.sref[1]), rz = (float)(zs - S.sref[2]);
#pragma unroll
                    for (int r = 0; r < 3; r++) {
                        const float *M = S.M[r];
                        float dx = M[0] * rx + M[1] * ry + M[2] * rz;
                        float dy = M[3] * rx + M[4] * ry + M[5] * rz;
                        float dz = M[6] * rx + M[7] * ry + M[8] * rz;
                        v[3 + r] = gx * dx + gy * dy + gz * dz;
                    }
                    if (ndof > 6) {
                        float w0 = S.Af[0] * rx + S.Af[1] * ry + S.Af[2] * rz;
                        float w1 = S.Af[3] * rx + S.Af[4] * ry + S.Af[5] * rz;
                        float w2 = S.Af[6] * rx + S.Af[7] * ry + S.Af[8] * rz;
                        v[6] = gx * w0;
                        v[7] = gy * w1;
                        v[8] = gz * w2;
                        v[9] = gx * w1 + gy * w0;
                        v[10] = gy * w2 + gz * w1;
                        v[11] = gx * w2 + gz * w0;
                    }
                }
                v[12] = tc;
                v[13] = rc;
                v[14] = 1.0f;
            }
        }
        if (jrow) {
            // stage the 64 rows of this wave, then V += rows^T rows on the matrix cores
#pragma unroll
            for (int q = 0; q < 16; q++) wst[lane * 17 + q] = v[q];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // two accumulators: independent MFMA chains of 8
            floatx4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < 16; s += 2) {
                float xv0 = wst[(4 * s + (lane >> 4)) * 17 + (lane & 15)];
                float xv1 = wst[(4 * (s + 1) + (lane >> 4)) * 17 + (lane & 15)];
                acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(xv0, xv0, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(xv1, xv1, acc1, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) acc64[r] += (double)acc0[r] + (double)acc1[r];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    };

    constexpr int NWv = B / 64;
    const int n = a.nsamp;
    const int nchunk = (n + 63) / 64;
    const int mych = c0 < nchunk ? (nchunk - c0 + cstride - 1) / cstride : 0;   // this workgroup's chunks
    const int steps = (mych + NWv - 1) / NWv;
    // sample of this thread at step it (n or beyond: none)
    auto sample_of = [&](int it) {
        const int k = it * NWv + wid;
        return k < mych ? (c0 + k * cstride) * 64 + lane : n;
    };
    const int j1 = n;
    if (method == DVC_INTERP_TRICUBIC) {
        // U samples per thread: all their row loads are issued before any is used
        SphOff pre[U];
#pragma unroll
        for (int u = 0; u < U; u++) pre[u] = sph_off(a, sample_of(u), sample_of(u) < j1);
        for (int it = 0; it < steps; it += U) {
            double xs[U], ys[U], zs[U];
            Tap<BPV> t[U];
            SphOff cur[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                cur[u] = pre[u];
                const int jn = sample_of(it + U + u);
                pre[u] = sph_off(a, jn, jn < j1);
            }
            float rc[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int j = sample_of(it + u);
                double x = 0.0, y = 0.0, z = 0.0;
                xs[u] = ys[u] = zs[u] = 0.0;
                rc[u] = 0.0f;
                if (j < j1) {
                    rc[u] = rsrc[j];
                    position(j, cur[u], xs[u], ys[u], zs[u], x, y, z);
                }
                tap_locate<BPV>(bx, x, y, z, t[u]);
                // a lane without a sample reads interior rows and skips the face-cell
                // path (its value is not used)
                if (!(j < j1)) t[u].face = false;
            }
#pragma unroll
            for (int u = 0; u < U; u++) tap_load<BPV>(vv, gv, bx, t[u]);
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (it + u < steps) {
                    const int j = sample_of(it + u);
                    float gx, gy, gz;
                    float tv = tap_eval<BPV>(vv, bx, t[u], fbias, jrow, gx, gy, gz);
                    accumulate(j < j1, rc[u], tv, gx, gy, gz, xs[u], ys[u], zs[u]);
                }
            }
        }
    } else {
        for (int it = 0; it < steps; it++) {
            const int j = sample_of(it);
            const bool ok = j < j1;
            double xs = 0.0, ys = 0.0, zs = 0.0, x = 0.0, y = 0.0, z = 0.0;
            float t = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f, rc = 0.0f;
            if (ok) {
                rc = rsrc[j];
                position(j, sph_off(a, j, true), xs, ys, zs, x, y, z);
                t = interp_at<BPV>(vv, bx, x, y, z, method, fbias, jrow, gx, gy, gz);
            }
            accumulate(ok, rc, t, gx, gy, gz, xs, ys, zs);
        }
    }
    // block totals of the sums and of the moving bounding box
    {
        double v3[3] = {wave_total_sum(s_tc), wave_total_sum(s_x1), wave_total_sum(s_x2)};
        if (merged)
            for (int q = 0; q < 3; q++) { mn[q] = wave_total_min(mn[q]); mx[q] = wave_total_max(mx[q]); }
        if (lane == 0) {
            for (int q = 0; q < 3; q++) red[wid * 9 + q] = v3[q];
            for (int q = 0; q < 3; q++) { red[wid * 9 + 3 + q] = mn[q]; red[wid * 9 + 6 + q] = mx[q]; }
        }
        __syncthreads();
        if (tid < 3) {
            double sacc = 0.0;
            for (int w = 0; w < NW; w++) sacc += red[w * 9 + tid];
            S.sums[tid] = sacc;
        } else if (tid < 9) {
            const int q = tid - 3;
            double v = red[3 + q];
            for (int w = 1; w < NW; w++) v = q < 3 ? fmin(v, red[w * 9 + 3 + q]) : fmax(v, red[w * 9 + 3 + q]);
            S.bbr[q] = v;
        }
    }
    if (jrow) {
        double *st64 = (double *)stage;
        __syncthreads();            // every wave is done with its staging tile
        // D row = 4*(lane/16) + r, column = lane%16 (symmetric, so either reading works)
#pragma unroll
        for (int r = 0; r < 4; r++) st64[wid * 256 + (4 * (lane >> 4) + r) * 16 + (lane & 15)] = acc64[r];
        __syncthreads();
        for (int e = tid; e < 256; e += B) {
            double s = 0.0;
            for (int w = 0; w < NW; w++) s += st64[w * 256 + e];
            V[e] = s;
        }
    }
    __syncthreads();
}

// The range rule on the moving bounding box of the whole cloud (S.bbr; with more than
// 3 DOF, Interpolate.cpp:474-481 on the box of FloatingCloud::resize_bbox), then the
// objective value and the moments the step needs (ObjectiveFunctions.cpp:24-263) from
// the sums in S.sums.  The centred sums of the reference's second and third passes
// follow from the one pass, e.g. sum (t - tm)^2 = sum t^2 - tm sum t, and the normalised
// objectives are (2 - 2 rho)/4 and (2 - 2 rho)/2 with rho the normalised correlation.
// False (uniformly) on Range_Fail.
__device__ __forceinline__ bool eval_close(const SearchArgs &a, PtShared &S, int ndof) {
    const int tid = threadIdx.x;
    if (tid == 0 && ndof > 3) {
        for (int q = 0; q < 3; q++) {
            // resize_bbox sentinels (FloatingCloud.cpp:208-214)
            S.bb[q] = fmin(S.bbr[q], 9999.9);
            S.bb[3 + q] = fmax(S.bbr[3 + q], -9999.9);
        }
        S.fail = range_ok(a, S, S.bb) ? 0 : 1;
    }
    if (tid == 0 && !S.fail) {
        const int kind = a.obj;
        const double nn = (double)a.nsamp;
        const double st = S.sums[0], sx1 = S.sums[1], sx2 = S.sums[2];
        const double tm = st / nn;
        S.tmean = tm;
        if (kind == DVC_OBJ_SAD) {
            S.obj = sx2;
        } else if (kind == DVC_OBJ_SSD) {
            S.obj = sx1;
        } else {
            const double rm = S.rmean;
            const double N2 = sx1 - tm * st;
            const double Str = sx2 - tm * S.Src;
            const double T2 = sx1 + 2.0 * rm * st + nn * rm * rm;
            const double Strr = sx2 + rm * st + rm * S.Src + nn * rm * rm;
            S.N2 = N2; S.Str = Str; S.T2 = T2; S.Strr = Strr;
            if (kind == DVC_OBJ_ZSSD) {
                S.obj = N2 - 2.0 * Str + S.R2;
            } else if (kind == DVC_OBJ_ZNSSD) {
                S.obj = (2.0 - 2.0 * Str / (sqrt(N2) * sqrt(S.R2))) / 4.0;
            } else {
                S.obj = (2.0 - 2.0 * Strr / (sqrt(T2) * sqrt(S.Rn2))) / 2.0;
            }
        }
    }
    __syncthreads();
    return S.fail == 0;
}

// The whole evaluation on one workgroup.  With need_j the analytic-Jacobian moments are
// left in V (16x16, FP64).  Returns false on Range_Fail.
template <int BPV, int B>
__device__ __forceinline__ bool evaluate(const SearchArgs &a, PtShared &S, int ndof, bool need_j,
                                         const float *rsrc, float *stage, double *red, double *V) {
    if (!eval_begin<B>(a, S, ndof)) return false;
    eval_samples<BPV, B>(a, S, ndof, need_j, rsrc, stage, red, V, 0, 1);
    return eval_close(a, S, ndof);
}

// ------------------------------------------------------------- step solve

// Cholesky solve of the symmetric normal equations held one column per lane (lane j < N:
// column j of A, lane N: right-hand side b), on wavefront 0.  Step k takes the pivot
// A_kk and the column L_ik (i > k) of the updated trailing matrix from lane k, both
// uniform; lane j > k applies the rank-1 update A_ij -= L_ik L_jk to its column, and
// lane N runs the forward substitution L y = b alongside.  Only the lower triangle is
// read.  The back substitution L^T x = y is uniform on every lane.  Returns false,
// having written nothing, when a pivot is not above 1e-9 of the largest diagonal entry
// (or is not finite): the caller then solves by column-pivoted QR, whose rank rule
// covers the (near-)singular systems.
template <int N>
__device__ __forceinline__ bool chol_solve(double (&col)[N], int lane, double (&x)[N]) {
    double dmax = 0.0;
#pragma unroll
    for (int k = 0; k < N; k++) dmax = fmax(dmax, rdlane(col[k], k));
    const double tol = 1e-9 * dmax;
    double lkk[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
        const double d = rdlane(col[k], k);
        if (!(d > tol) || !(d < 1.0e300)) return false;
        const double l = sqrt(d), r = 1.0 / l;
        lkk[k] = l;
        double lik[N];
        double ljk = (lane == N) ? col[k] * r : 0.0;   // lane j > k: L_jk; lane N: y_k
#pragma unroll
        for (int i = k + 1; i < N; i++) {
            lik[i] = rdlane(col[i], k) * r;
            if (lane == i) ljk = lik[i];
        }
        if (lane > k && lane <= N)
#pragma unroll
            for (int i = k + 1; i < N; i++) col[i] -= lik[i] * ljk;
        if (lane == N) col[k] = ljk;                    // y_k
        if (lane == k) {                                // column k of L
            col[k] = l;
#pragma unroll
            for (int i = k + 1; i < N; i++) col[i] = lik[i];
        }
    }
#pragma unroll
    for (int i = N - 1; i >= 0; i--) {
        double sacc = rdlane(col[i], N);
#pragma unroll
        for (int q = i + 1; q < N; q++) sacc -= rdlane(col[q], i) * x[q];
        x[i] = sacc / lkk[i];
    }
    return true;
}

// Reciprocals of the scale factors of the normal equations (normal_cols), once per step
// for the slice launches' solve: its entries are then products, not quotients.
struct NormalScale {
    double inv_n, inv_a, inv_b, inv_c, inv_d;
};

__device__ __forceinline__ NormalScale normal_scale(const SearchArgs &a, const PtShared &S) {
#pragma clang fp contract(off)
    NormalScale f = {1.0 / (double)a.nsamp, 0.0, 0.0, 0.0, 0.0};
    if (a.obj == DVC_OBJ_NSSD) {
        const double T2 = S.T2, T = sqrt(T2), Rn = sqrt(S.Rn2);
        f.inv_a = 1.0 / T2; f.inv_b = 1.0 / (T2 * T2); f.inv_c = 1.0 / (T * Rn); f.inv_d = 1.0 / (T2 * T * Rn);
    } else if (a.obj == DVC_OBJ_ZNSSD) {
        const double N2 = S.N2, Nn = sqrt(N2), R = sqrt(S.R2);
        f.inv_a = 1.0 / N2; f.inv_b = 1.0 / (N2 * N2); f.inv_c = 1.0 / (Nn * R); f.inv_d = 1.0 / (N2 * Nn * R);
    }
    return f;
}

// 1 / d for a positive finite d: the hardware reciprocal estimate and two Newton steps
// r' = r + r (1 - d r), which carry its error below double rounding (within an ulp of
// the quotient).  Shorter on the dependency chain than the division.
__device__ __forceinline__ double fast_rcp(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
    return r;
}

// Row i of the objective's Gauss-Newton normal equations (Search.cpp:234's system, as
// LM_prep_at assembles it) from the block's moments V, for gj_rows: r[j] = (J^T J)_ij
// from the lower triangle's entry (max(i, j), min(i, j)), as the other solves read it,
// and r[N] = (-J^T e)_i.  The arithmetic of normal_cols, with the quotients replaced by
// products with the reciprocals in f.
template <int N>
__device__ __forceinline__ void normal_row(const SearchArgs &a, const PtShared &S, const double *V, int i,
                                           const NormalScale &f, double (&r)[N + 1]) {
#pragma clang fp contract(off)
    const int kind = a.obj;
    const double Sg = V[i * 16 + 14];
    auto vl = [&](int j) { return i > j ? V[i * 16 + j] : V[j * 16 + i]; };
    if (kind == DVC_OBJ_SAD || kind == DVC_OBJ_SSD) {
#pragma unroll
        for (int j = 0; j < N; j++) r[j] = vl(j);
        r[N] = -(V[i * 16 + 12] - V[i * 16 + 13]);
    } else if (kind == DVC_OBJ_ZSSD) {
        const double tm = S.tmean, Src = V[13 * 16 + 14];
#pragma unroll
        for (int j = 0; j < N; j++) r[j] = vl(j) - Sg * V[j * 16 + 14] * f.inv_n;
        r[N] = -((V[i * 16 + 12] - tm * Sg) - V[i * 16 + 13] + (Sg * f.inv_n) * Src);
    } else if (kind == DVC_OBJ_NSSD) {
        const double rm = S.rmean;
        const double ci = V[i * 16 + 12] + rm * Sg;
#pragma unroll
        for (int j = 0; j < N; j++) r[j] = vl(j) * f.inv_a - ci * (V[j * 16 + 12] + rm * V[j * 16 + 14]) * f.inv_b;
        r[N] = -(-(V[i * 16 + 13] + rm * Sg) * f.inv_c + ci * S.Strr * f.inv_d);
    } else {
        const double tm = S.tmean;
        const double ci = V[i * 16 + 12] - tm * Sg;
#pragma unroll
        for (int j = 0; j < N; j++) {
            const double Sgj = V[j * 16 + 14];
            r[j] = (vl(j) - Sg * Sgj * f.inv_n) * f.inv_a - ci * (V[j * 16 + 12] - tm * Sgj) * f.inv_b;
        }
        const double gr = V[i * 16 + 13] - Sg * S.Src * f.inv_n;
        r[N] = -(-gr * f.inv_c + ci * S.Str * f.inv_d);
    }
}

// Solve of the symmetric normal equations A x = b by Gauss-Jordan elimination with one
// row per lane: lane i < N holds row i of A and b_i in r[0..N].  Step k reads the pivot
// row k from lane k into uniform registers, and every other row subtracts its multiple
// A_ik / A_kk of it, rows above the pivot as well as below; A ends diagonal, and lane i
// returns x_i = b_i / A_ii.  The pivots A_kk are those of LDL^T (the diagonals of the
// successive Schur complements), tested as there: returns false, having computed nothing
// usable, when a pivot is not above 1e-9 of the largest diagonal entry (or is not
// finite), and the caller solves by column-pivoted QR.  The test is taken once, after
// the last step: a failed pivot only spoils values that are then discarded.  Lanes at
// or above N compute nothing used.  Compared with an LDL^T computed whole on every lane,
// a step needs ~N + 1 dependent fused multiply-adds instead of the whole trailing
// update, and a lane holds N + 1 FP64 values instead of the whole system.
template <int N>
__device__ __forceinline__ bool gj_rows(double (&r)[N + 1], int lane, double &x) {
    double dmax = 0.0;
#pragma unroll
    for (int k = 0; k < N; k++) dmax = fmax(dmax, rdlane(r[k], k));
    const double tol = 1e-9 * dmax;
    double rinv_own = 0.0;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < N; k++) {
        double pr[N + 1];
#pragma unroll
        for (int j = k; j <= N; j++) pr[j] = rdlane(r[j], k);
        const double d = pr[k];
        ok = ok && (d > tol) && (d < 1.0e300);
        const double rinv = fast_rcp(d);
        if (lane == k) {
            rinv_own = rinv;
        } else {
            const double f = r[k] * rinv;
            r[k] = 0.0;
#pragma unroll
            for (int j = k + 1; j <= N; j++) r[j] = __builtin_fma(-f, pr[j], r[j]);
        }
    }
    x = r[N] * rinv_own;
    return ok;
}

// Column `lane` (lane < N) of the objective's Gauss-Newton normal matrix J^T J, or for
// lane N the right-hand side -J^T e, from the block's moments V (Search.cpp:234's
// system, as LM_prep_at assembles it); other lanes get column 0.
template <int N>
__device__ __forceinline__ void normal_cols(const SearchArgs &a, const PtShared &S, const double *V, int lane,
                                            double (&col)[N]) {
#pragma clang fp contract(off)
    const double n = (double)a.nsamp;
    const int kind = a.obj;
    const int j = lane < N ? lane : 0;      // column of this lane (lane N: right-hand side)
    {
        const double Sgj = V[j * 16 + 14];
        if (kind == DVC_OBJ_SAD || kind == DVC_OBJ_SSD) {
#pragma unroll
            for (int i = 0; i < N; i++) col[i] = V[i * 16 + j];
            if (lane == N)
#pragma unroll
                for (int i = 0; i < N; i++) col[i] = -(V[i * 16 + 12] - V[i * 16 + 13]);
        } else if (kind == DVC_OBJ_ZSSD) {
            const double tm = S.tmean, Src = V[13 * 16 + 14];
#pragma unroll
            for (int i = 0; i < N; i++) col[i] = V[i * 16 + j] - V[i * 16 + 14] * Sgj / n;
            if (lane == N)
#pragma unroll
                for (int i = 0; i < N; i++) {
                    const double Sg = V[i * 16 + 14];
                    col[i] = -((V[i * 16 + 12] - tm * Sg) - V[i * 16 + 13] + (Sg / n) * Src);
                }
        } else if (kind == DVC_OBJ_NSSD) {
            const double T2 = S.T2, T = sqrt(T2), Rn = sqrt(S.Rn2), rm = S.rmean;
            const double cj = V[j * 16 + 12] + rm * Sgj;
#pragma unroll
            for (int i = 0; i < N; i++) {
                const double ci = V[i * 16 + 12] + rm * V[i * 16 + 14];
                col[i] = V[i * 16 + j] / T2 - ci * cj / (T2 * T2);
            }
            if (lane == N)
#pragma unroll
                for (int i = 0; i < N; i++) {
                    const double Sg = V[i * 16 + 14];
                    const double ci = V[i * 16 + 12] + rm * Sg;
                    const double Sgr = V[i * 16 + 13] + rm * Sg;
                    col[i] = -(-Sgr / (T * Rn) + ci * S.Strr / (T2 * T * Rn));
                }
        } else {
            const double tm = S.tmean, N2 = S.N2, Nn = sqrt(N2), R = sqrt(S.R2);
            const double cj = V[j * 16 + 12] - tm * Sgj;
#pragma unroll
            for (int i = 0; i < N; i++) {
                const double ci = V[i * 16 + 12] - tm * V[i * 16 + 14];
                col[i] = (V[i * 16 + j] - V[i * 16 + 14] * Sgj / n) / N2 - ci * cj / (N2 * N2);
            }
            if (lane == N)
#pragma unroll
                for (int i = 0; i < N; i++) {
                    // sum_j (g_ij - mean g_i) (r_j - mean r): the FP32-centred reference
                    // samples sum to Src, not exactly zero
                    const double Sg = V[i * 16 + 14];
                    const double ci = V[i * 16 + 12] - tm * Sg;
                    const double gr = V[i * 16 + 13] - Sg * S.Src / n;
                    col[i] = -(-gr / (Nn * R) + ci * S.Str / (N2 * Nn * R));
                }
        }
    }
}

// One Gauss-Newton step on wavefront 0 (all 64 lanes call it): the normal equations
// J^T J dx = -J^T e of the objective's residual, assembled from the moments in V, then
// solved by Cholesky with one column per lane (chol_solve), or with UNIFORM by
// Gauss-Jordan with one row per lane (gj_rows; the slice launches' solve, with the
// shorter dependency chain).  `ne` is LDS scratch of N (N + 1) / 2 + N doubles, unused
// by these two.  When the factorisation rejects a pivot the step is
// solved by column-pivoted Householder QR
// (Eigen 3.4 ColPivHouseholderQR::computeInPlace and _solve_impl, as used at
// Search.cpp:234): lane j keeps column j of J^T J in registers (lane N the right-hand
// side); the pivot scan, the reflector and the back substitution are uniform across
// lanes, and every column is updated by its own lane with the same operation order as a
// serial loop.  Adds the step to S.par and stores it in S.upd (Search.cpp:236-240).
template <int N, bool CHOL = true, bool UNIFORM = false>
__device__ __forceinline__ void gn_step(const SearchArgs &a, PtShared &S, const double *V, double *ne) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    if (CHOL && UNIFORM) {
        // row `lane` of the system (the lower triangle's entries, as the other solves
        // read them), then Gauss-Jordan with one row per lane
        const NormalScale f = normal_scale(a, S);
        const int i = lane < N ? lane : 0;
        double r[N + 1];
        normal_row<N>(a, S, V, i, f, r);
        double x = 0.0;
        if (gj_rows<N>(r, lane, x)) {
            if (lane < N) {
                S.upd[lane] = x;
                S.par[lane] += x;
            }
            return;
        }
    }
    double col[N];
    normal_cols<N>(a, S, V, lane, col);
    if (CHOL && !UNIFORM) {
        double x[N];
        if (chol_solve<N>(col, lane, x)) {
#pragma unroll
            for (int i = 0; i < N; i++)
                if (lane == i) {
                    S.upd[i] = x[i];
                    S.par[i] += x[i];
                }
            return;
        }
        normal_cols<N>(a, S, V, lane, col);       // the factorisation overwrote them
    }
    // column norms (exact and updated), the permutation carried with the columns
    double cn = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) cn += col[i] * col[i];
    double nrm_d = sqrt(cn), nrm_u = nrm_d;
    int orig = lane;
    double maxn = 0.0;
#pragma unroll
    for (int k = 0; k < N; k++) maxn = fmax(maxn, rdlane(nrm_u, k));
    const double eps = 2.220446049250313e-16;
    const double thr = (maxn * eps) * (maxn * eps) / (double)N;
    const double dthr = sqrt(eps);
    int nonzero = N;
#pragma unroll
    for (int k = 0; k < N; k++) {
        int big = k;
        double bn = rdlane(nrm_u, k);
#pragma unroll
        for (int q = k + 1; q < N; q++) {
            double v = rdlane(nrm_u, q);
            if (v > bn) { bn = v; big = q; }
        }
        if (nonzero == N && bn * bn < thr * (double)(N - k)) nonzero = k;
        if (big != k) {
            const int src = (lane == k) ? big : ((lane == big) ? k : lane);
#pragma unroll
            for (int i = 0; i < N; i++) col[i] = shfl_from(col[i], src);
            nrm_u = shfl_from(nrm_u, src);
            nrm_d = shfl_from(nrm_d, src);
            orig = shfl_from(orig, src);
        }
        // makeHouseholder of pivot column k, rows k..N-1 (uniform on every lane)
        const int m = N - k;
        double ess[N];
        double tail = 0.0;
#pragma unroll
        for (int i = 1; i < m; i++) { ess[i] = rdlane(col[k + i], k); tail += ess[i] * ess[i]; }
        const double c0 = rdlane(col[k], k);
        double tau, beta;
        if (tail <= 2.2250738585072014e-308) {
            tau = 0.0; beta = c0;
#pragma unroll
            for (int i = 1; i < m; i++) ess[i] = 0.0;
        } else {
            beta = sqrt(c0 * c0 + tail);
            if (c0 >= 0.0) beta = -beta;
#pragma unroll
            for (int i = 1; i < m; i++) ess[i] = ess[i] / (c0 - beta);
            tau = (beta - c0) / beta;
        }
        if (lane == k) {
            col[k] = beta;
#pragma unroll
            for (int i = 1; i < m; i++) col[k + i] = ess[i];
        }
        // apply the reflector to the later columns and, while k < nonzero, to the
        // right-hand side
        const bool later = lane > k && lane < N;
        const bool rhs = lane == N && k < nonzero;
        if (later || rhs) {
            if (m > 1 && tau != 0.0) {
                double t = 0.0;
#pragma unroll
                for (int i = 1; i < m; i++) t += ess[i] * col[k + i];
                t += col[k];
                col[k] -= tau * t;
#pragma unroll
                for (int i = 1; i < m; i++) col[k + i] -= tau * ess[i] * t;
            } else if (m == 1) {
                col[k] *= (1.0 - tau);
            }
        }
        // norm downdate of the later columns
        if (later && nrm_u != 0.0) {
            double t = fabs(col[k]) / nrm_u;
            t = (1.0 + t) * (1.0 - t);
            t = t < 0.0 ? 0.0 : t;
            const double r = nrm_u / nrm_d;
            const double t2 = t * (r * r);
            if (t2 <= dthr) {
                double s2 = 0.0;
#pragma unroll
                for (int i = k + 1; i < N; i++) s2 += col[i] * col[i];
                nrm_d = sqrt(s2);
                nrm_u = nrm_d;
            } else {
                nrm_u *= sqrt(t);
            }
        }
    }
    // back substitution with R (upper triangle of the columns), uniform on every lane
    double c[N];
#pragma unroll
    for (int i = 0; i < N; i++) c[i] = rdlane(col[i], N);
#pragma unroll
    for (int i = N - 1; i >= 0; i--) {
        if (i < nonzero) {
            double sacc = c[i];
#pragma unroll
            for (int q = i + 1; q < N; q++)
                if (q < nonzero) sacc -= rdlane(col[i], q) * c[q];
            c[i] = sacc / rdlane(col[i], i);
        }
    }
    // x[perm[i]] = c[i] for i < nonzero, 0 otherwise; the step is added to the parameters
#pragma unroll
    for (int i = 0; i < N; i++)
        if (lane == i) {
            const double x = (i < nonzero) ? c[i] : 0.0;
            S.upd[orig] = x;
            S.par[orig] += x;
        }
}

// ------------------------------------------------------------- the search

// Point set-up (Search.cpp:98-119, 252-258): cloud geometry, the reference box on the
// point, its range rule and the integer bias.  Returns false (uniformly) on Range_Fail.
template <int BPV>
__device__ __forceinline__ void point_setup_t0(const SearchArgs &a, int n, PtShared &S) {
#pragma clang fp contract(off)
    // ---- cloud geometry and reference box (Search.cpp:98-119, 252-258)
    {
#pragma clang fp contract(off)
        S.status = DVC_POINT_GOOD; S.iters = 0; S.fail = 0; S.stop = 0;
        for (int q = 0; q < 12; q++) { S.par[q] = 0.0; S.upd[q] = 0.0; }
        for (int q = 0; q < 3; q++) S.sp[q] = a.pts[3 * n + q];
        if (a.geom == DVC_GEOM_CUBE) {
            double smn[3], smx[3];
            for (int q = 0; q < 3; q++) {
                smn[q] = S.sp[q] + -a.subv_rad * a.aspect[q];
                smx[q] = S.sp[q] + a.subv_rad * a.aspect[q];
            }
            for (int q = 0; q < 3; q++) {
                S.sref[q] = (smx[q] + smn[q]) / 2.0;
                S.inc[q] = (smx[q] - smn[q]) / (a.g - 1);
                S.smin[q] = smn[q];
                S.sbb_mn[q] = smn[q] + 0.0 * S.inc[q];
                S.sbb_mx[q] = smn[q] + (double)(a.g - 1) * S.inc[q];
            }
        } else {
            for (int q = 0; q < 3; q++) {
                S.sref[q] = S.sp[q];
                S.sbb_mn[q] = S.sp[q] + a.off_min[q];
                S.sbb_mx[q] = S.sp[q] + a.off_max[q];
            }
        }
        place_box(a, S, S.sp[0], S.sp[1], S.sp[2]);
        double bb[6] = {S.sbb_mn[0], S.sbb_mn[1], S.sbb_mn[2], S.sbb_mx[0], S.sbb_mx[1], S.sbb_mx[2]};
        S.fail = range_ok(a, S, bb) ? 0 : 1;
        S.fbias = 0.0f;
        if (!S.fail) {
            // the search point lies inside the (valid) reference box
            const VolView vr = {a.vol[0], a.W, (long long)a.W * a.H};
            S.fbias = voxel<BPV>(vr, (int)(S.sp[0] / 1), (int)(S.sp[1] / 1), (int)(S.sp[2] / 1));
        }
    }
}

template <int BPV, int B>
__device__ __forceinline__ bool point_setup(const SearchArgs &a, int n, PtShared &S) {
    if (threadIdx.x == 0) point_setup_t0<BPV>(a, n, S);
    __syncthreads();
    return S.fail == 0;
}

// The start guess's inputs (Search.cpp:277-377), by one wavefront: the translations of
// the neighbours of point n (at position p) that come earlier in the order and are
// good, summed (wavefront sums) into S.nb_sum, and their number into S.nb_good.
__device__ __forceinline__ void gather_neighbours(const SearchArgs &a, int p, int n, PtShared &S) {
    double sx = 0.0, sy = 0.0, sz = 0.0, good = 0.0;
    for (int q = threadIdx.x & 63; q < a.k; q += 64) {
        const int m = a.neigh[(long long)n * a.k + q];
        if (m >= 0 && m < a.npts && a.pos[m] < p && a.nbres[m].status == DVC_POINT_GOOD) {
            sx += a.nbres[m].par[0]; sy += a.nbres[m].par[1]; sz += a.nbres[m].par[2];
            good += 1.0;
        }
    }
    sx = wave_sum(sx); sy = wave_sum(sy); sz = wave_sum(sz); good = wave_sum(good);
    if ((threadIdx.x & 63) == 0) {
        S.nb_sum[0] = sx; S.nb_sum[1] = sy; S.nb_sum[2] = sz;
        S.nb_good = (int)good;
    }
}

// The reference sub-volume (Search.cpp:252-258): sample values at the stable cloud,
// centred, left in rcb[], and their sums in S.
template <int BPV, int B>
__device__ __forceinline__ void reference_pass(const SearchArgs &a, PtShared &S, float *rcb, double *red) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x;
    const int ns = a.nsamp;
    {
        const VolView vr = {a.vol[0], a.W, (long long)a.W * a.H};
        const BoxView bx = {{S.b0[0], S.b0[1], S.b0[2]}, {a.box[0], a.box[1], a.box[2]}};
        const Gather gr = make_gather<BPV>(vr, bx, a.gather32 != 0);
        const float fbias = S.fbias;
        double s[1] = {0.0};
        if (a.interp == DVC_INTERP_TRICUBIC) {
            // kInflight samples per thread in flight (see evaluate)
            constexpr int U = kInflight;
            SphOff pre[U];
#pragma unroll
            for (int u = 0; u < U; u++) pre[u] = sph_off(a, tid + u * B, tid + u * B < ns);
            for (int j0 = tid; j0 < ns; j0 += U * B) {
                Tap<BPV> t[U];
                SphOff cur[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    cur[u] = pre[u];
                    const int jn = j0 + (U + u) * B;
                    pre[u] = sph_off(a, jn, jn < ns);
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int j = j0 + u * B;
                    double x = 0.0, y = 0.0, z = 0.0;
                    if (j < ns) stable_pt_off(a, S, j, cur[u], x, y, z);
                    tap_locate<BPV>(bx, x, y, z, t[u]);
                }
#pragma unroll
                for (int u = 0; u < U; u++) tap_load<BPV>(vr, gr, bx, t[u]);
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int j = j0 + u * B;
                    if (j < ns) {
                        float gx, gy, gz;
                        float r = tap_eval<BPV>(vr, bx, t[u], fbias, false, gx, gy, gz);
                        rcb[j] = r;
                        s[0] += (double)r;
                    }
                }
            }
        } else {
            for (int j = tid; j < ns; j += B) {
                double x, y, z;
                stable_pt(a, S, j, x, y, z);
                float gx, gy, gz;
                float r = interp_at<BPV>(vr, bx, x, y, z, a.interp, fbias, false, gx, gy, gz);
                rcb[j] = r;
                s[0] += (double)r;
            }
        }
        block_sum<B, 1>(s, red, S.sums);
        const double rbmean = S.sums[0] / (double)ns;
        const double rmean = rbmean + (double)fbias;
        double q3[3] = {0.0, 0.0, 0.0};
        for (int j = tid; j < ns; j += B) {
            float rc = (float)((double)rcb[j] - rbmean);
            rcb[j] = rc;
            q3[0] += (double)rc * (double)rc;
            q3[1] += (double)rc;
            double r = (double)rc + rmean;
            q3[2] += r * r;
        }
        block_sum<B, 3>(q3, red, &S.sums[1]);
        if (tid == 0) {
            S.rmean = rmean; S.rbmean = rbmean;
            S.R2 = S.sums[1]; S.Src = S.sums[2]; S.Rn2 = S.sums[3];
        }
    }
}

// Search::process_point up to its Gauss-Newton loop (Search.cpp:98-119, 248-377,
// 934-1043): set-up, reference samples, start guess and (GRID) the translation grid.
// rsrc receives the point's reference samples.  False (uniformly) when the point
// failed; S.status then holds its status.
template <int BPV, int B, bool GRID>
__device__ __forceinline__ bool point_start(const SearchArgs &a, int p, int n, PtShared &S, float *rcb, float *stage,
                                            double *red, double *V, const float *&rsrc) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x, wid = tid >> 6;
    // independent loads side by side: the point's set-up (thread 0), the neighbours'
    // results for the start guess (one wavefront) and the reference-sample sums
    constexpr int kGatherWave = B >= 128 ? 1 : 0;
    constexpr int kStatThread = B >= 192 ? 128 : 0;
    if (tid == 0) point_setup_t0<BPV>(a, n, S);
    if (!a.start && wid == kGatherWave) gather_neighbours(a, p, n, S);
    if (a.rvals && tid == kStatThread) {
        const RefStats &st = a.rstat[p];
        S.rmean = st.rmean; S.rbmean = st.rbmean;
        S.R2 = st.R2; S.Src = st.Src; S.Rn2 = st.Rn2;
    }
    __syncthreads();
    if (S.fail) {
        if (tid == 0) S.status = DVC_RANGE_FAIL;
        return false;
    }
    // reference samples: computed for every point before the levels run
    // (reference_kernel) and read from HBM, or computed here into LDS
    rsrc = rcb;
    if (a.rvals) rsrc = a.rvals + (size_t)p * a.nsamp_pad;
    else reference_pass<BPV, B>(a, S, rcb, red);
    if (a.start) {
        // ---- a given start warp (proposal pass): all ndof parameters, no neighbours
        if (tid == 0) {
            int bad = 0;
            for (int q = 0; q < 12; q++) {
                const double v = (q < a.ndof) ? a.start[(long long)n * 12 + q] : 0.0;
                if (!isfinite(v)) bad = 1;
                S.par[q] = v;
            }
            S.fail = bad;
            if (!bad) place_box(a, S, S.sp[0] + S.par[0], S.sp[1] + S.par[1], S.sp[2] + S.par[2]);
        }
        __syncthreads();
        if (S.fail) {
            if (tid == 0) S.status = DVC_RANGE_FAIL;
            return false;
        }
    } else {
        // ---- starting_param: good neighbours at earlier positions of the order
        // (gathered above).  In the serial chain (nbres == res) launches run the
        // dependency levels in order, so every such neighbour's result is already in
        // memory; a judge pass reads them from the finished table of a proposal pass.
        if (tid == 0) {
#pragma clang fp contract(off)
            double sx = S.nb_sum[0], sy = S.nb_sum[1], sz = S.nb_sum[2];
            const int good = S.nb_good;
            if (good != 0) { sx /= good; sy /= good; sz /= good; }
            if (good != 0) { S.par[0] = sx; S.par[1] = sy; S.par[2] = sz; }
            else { S.par[0] = a.rigid[0]; S.par[1] = a.rigid[1]; S.par[2] = a.rigid[2]; }
            // target box on the offset point (Search.cpp:267-273)
            double ox = S.sp[0] + S.par[0], oy = S.sp[1] + S.par[1], oz = S.sp[2] + S.par[2];
            place_box(a, S, ox, oy, oz);
        }
        __syncthreads();
    }
    // ---- translation grid (Search.cpp:934-1043)
    if (GRID && a.basin != 0.0) {
        const int ni = (int)(a.disp_max / a.basin);
        const double inc = a.disp_max / ni;
        if (tid == 0) { S.omin = 1.7976931348623157e308; S.gmin[0] = S.gmin[1] = S.gmin[2] = 0; }
        for (int z = -ni; z <= ni; z++)
            for (int y = -ni; y <= ni; y++)
                for (int x = -ni; x <= ni; x++) {
                    if (tid == 0) {
#pragma clang fp contract(off)
                        S.evalpar[0] = S.par[0] + (double)x * inc;
                        S.evalpar[1] = S.par[1] + (double)y * inc;
                        S.evalpar[2] = S.par[2] + (double)z * inc;
                    }
                    __syncthreads();
                    if (!evaluate<BPV, B>(a, S, 3, false, rsrc, stage, red, V)) {
                        if (tid == 0) S.status = DVC_RANGE_FAIL;
                        return false;
                    }
                    if (tid == 0 && S.obj < S.omin) { S.omin = S.obj; S.gmin[0] = x; S.gmin[1] = y; S.gmin[2] = z; }
                    __syncthreads();
                }
        if (tid == 0) {
#pragma clang fp contract(off)
            S.par[0] += (double)S.gmin[0] * inc;
            S.par[1] += (double)S.gmin[1] * inc;
            S.par[2] += (double)S.gmin[2] * inc;
        }
        __syncthreads();
    }
    return true;
}

// Gauss-Newton from iteration i0 (Search.cpp:194-246; obj_old: the objective of
// iteration i0 - 1), then the final objective when the iteration limit ends the loop
// (Search.cpp:187).  S.status is set on a range failure.
template <int BPV, int B>
__device__ __forceinline__ void gn_iterations(const SearchArgs &a, PtShared &S, const float *rsrc, float *stage,
                                              double *red, double *V, int i0, double obj_old) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x;
    const int nd = a.ndof;
    bool converged = false;
    for (int i = i0; i < 20; i++) {
        if (tid == 0) {
            for (int q = 0; q < 12; q++) S.evalpar[q] = S.par[q];
            S.iters += 1;                 // LM_prep_at calls, counted when attempted
        }
        __syncthreads();
        // the step-size test needs only the previous step: when it already stops the
        // loop, this evaluation only supplies the final objective and needs no Jacobian
        const double del_mag = sqrt(S.upd[0] * S.upd[0] + S.upd[1] * S.upd[1] + S.upd[2] * S.upd[2]);
        const bool small_step = i > 0 && del_mag < 0.01;
        if (!evaluate<BPV, B>(a, S, nd, !small_step, rsrc, stage, red, V)) {
            if (tid == 0) S.status = DVC_RANGE_FAIL;
            return;
        }
        if (tid < 64) {
            // stopping test on the previous step (Search.cpp:217-230), uniform on wave 0
            bool stop = false;
            if (i > 0) {
                double del_obj = fabs(S.obj - obj_old);
                if ((del_obj < 0.000001) || small_step) stop = true;
            }
            if (!stop) {
                double *ne = (double *)stage;
                if (nd == 3) gn_step<3>(a, S, V, ne);
                else if (nd == 6) gn_step<6>(a, S, V, ne);
                else gn_step<12>(a, S, V, ne);
            }
            if (tid == 0) S.stop = stop ? 1 : 0;
        }
        __syncthreads();
        if (S.stop) { converged = true; break; }
        obj_old = S.obj;
    }
    // ---- final objective (Search.cpp:187): needed when the iteration limit ended the loop
    if (!converged) {
        if (tid == 0)
            for (int q = 0; q < 12; q++) S.evalpar[q] = S.par[q];
        __syncthreads();
        if (!evaluate<BPV, B>(a, S, nd, false, rsrc, stage, red, V)) {
            if (tid == 0) S.status = DVC_RANGE_FAIL;
            return;
        }
    }
}

template <int BPV, int B>
__device__ __forceinline__ void point_search(const SearchArgs &a, int p, int n, PtShared &S, float *rcb, float *stage,
                                             double *red, double *V) {
    const float *rsrc = nullptr;
    if (point_start<BPV, B, true>(a, p, n, S, rcb, stage, red, V, rsrc))
        gn_iterations<BPV, B>(a, S, rsrc, stage, red, V, 0, 0.0);
}

// Per-point geometry (PtGeo of point n) applied over the launch's arguments, by thread 0
// into the workgroup's shared copy.  The point's reference samples start at
// rvals + rval_off, so its copy carries that base and a row stride of 0: the search
// functions address them as rvals + p * nsamp_pad.
__device__ __forceinline__ void apply_geo(SearchArgs &la, const SearchArgs &a, int n) {
    la = a;
    const PtGeo &g = a.pgeo[n];
    la.geom = g.geom; la.nsamp = g.nsamp; la.g = g.g; la.gather32 = g.gather32;
    la.subv_rad = g.subv_rad;
    for (int q = 0; q < 3; q++) {
        la.box[q] = g.box[q];
        la.est_mn[q] = g.est_mn[q]; la.est_mx[q] = g.est_mx[q]; la.est_cn[q] = g.est_cn[q];
        la.act_mn[q] = g.act_mn[q]; la.act_mx[q] = g.act_mx[q]; la.act_cn[q] = g.act_cn[q];
        la.off_min[q] = g.off_min[q]; la.off_max[q] = g.off_max[q];
    }
    la.sph = a.sph + 3 * g.sph_off;
    la.rvals = a.rvals + g.rval_off;
    la.nsamp_pad = 0;
}

// The reference sub-volumes of all searched points at once: they depend only on the
// point, not on its neighbours, so they are taken out of the level chain.  Workgroup p
// handles the point at position p of the order.
template <int BPV, int B>
__device__ __forceinline__ void reference_body(const SearchArgs &a, int p, int n, PtShared &S, double *red) {
    const int tid = threadIdx.x;
    if (!point_setup<BPV, B>(a, n, S)) return;
    // each thread writes and re-reads only its own samples
    reference_pass<BPV, B>(a, S, a.rvals + (size_t)p * a.nsamp_pad, red);
    __syncthreads();
    if (tid == 0) {
        RefStats &st = a.rstat[p];
        st.rmean = S.rmean; st.rbmean = S.rbmean;
        st.R2 = S.R2; st.Src = S.Src; st.Rn2 = S.Rn2;
    }
}

template <int BPV, int B, bool UNEVEN>
__global__ void __launch_bounds__(B) reference_kernel(SearchArgs a) {
    DVC_SHARED(PtShared, S);
    DVC_SHARED(double[(B / 64) * 9], red);
    const int p = a.ref_order ? a.ref_order[blockIdx.x] : (int)blockIdx.x;
    const int n = a.order[p];
    if constexpr (UNEVEN) {
        DVC_SHARED(SearchArgs, la);
        if (threadIdx.x == 0) apply_geo(la, a, n);
        __syncthreads();
        reference_body<BPV, B>(la, p, n, S, red);
    } else {
        reference_body<BPV, B>(a, p, n, S, red);
    }
}

// Ends evaluation S.it of a point whose samples ran as slices: adds the slices' sums up
// (sp: the point's nslice partial records in HBM, in slice order), applies the range rule
// and the objective (eval_close), the stopping test of Search.cpp:217-230 and, when it
// does not stop, the step.  Every workgroup of the point runs it on the same inputs, so
// all agree.  Returns 0 after a step (S.it and S.obj_old advanced), 1 when the search
// stopped (S holds its result), -1 on Range_Fail (S.status set).
// Slice sums prefetched per thread: entry tid of the first kPrefetch slices' records.
constexpr int kPrefetch = 16;
template <int B>
__device__ __forceinline__ void prefetch_slices(const double *sp, int nslice, double (&pre)[kPrefetch]) {
    const int e = threadIdx.x;
#pragma unroll
    for (int q = 0; q < kPrefetch; q++) pre[q] = (e < kPartStride && q < nslice) ? sp[(size_t)q * kPartStride + e] : 0.0;
}

template <int B, bool UNIFORM>
__device__ __forceinline__ int finish_sliced(const SearchArgs &a, PtShared &S, const double *sp, int nslice,
                                             double *V, float *stage, const double (&pre)[kPrefetch]) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x;
    const int i = S.it, nd = a.ndof;
    const double del_mag = sqrt(S.upd[0] * S.upd[0] + S.upd[1] * S.upd[1] + S.upd[2] * S.upd[2]);
    const bool small_step = i > 0 && del_mag < 0.01;
    // entry e of every slice record, added up in slice order (min / max for the bounding
    // box); thread e's own entry of the first kPrefetch slices is already in `pre`
    for (int e = tid; e < kPartStride; e += B) {
        if ((e >= 9 && e < 16) || (e >= 16 && small_step)) continue;
        const int op = e < 3 || e >= 16 ? 0 : (e < 6 ? -1 : 1);
        auto fold = [&](double acc, double w, int q) {
            if (q == 0) return w;
            return op == 0 ? acc + w : (op < 0 ? fmin(acc, w) : fmax(acc, w));
        };
        double acc = 0.0;
        int q0 = 0;
        if (e == tid) {
#pragma unroll
            for (int q = 0; q < kPrefetch; q++)
                if (q < nslice) acc = fold(acc, pre[q], q);
            q0 = kPrefetch;
        }
        for (; q0 < nslice; q0 += 8) {
            double w[8];
#pragma unroll
            for (int u = 0; u < 8; u++) w[u] = q0 + u < nslice ? sp[(size_t)(q0 + u) * kPartStride + e] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (q0 + u < nslice) acc = fold(acc, w[u], q0 + u);
        }
        if (e < 3) S.sums[e] = acc;
        else if (e < 9) S.bbr[e - 3] = acc;
        else V[e - 16] = acc;
    }
    __syncthreads();
    if (!eval_close(a, S, nd)) {
        if (tid == 0) S.status = DVC_RANGE_FAIL;
        return -1;
    }
    if (tid < 64) {
        bool stop = false;
        if (i > 0) {
            const double del_obj = fabs(S.obj - S.obj_old);
            if ((del_obj < 0.000001) || small_step) stop = true;
        }
        if (!stop) {
            double *ne = (double *)stage;
            if (nd == 3) gn_step<3, true, UNIFORM>(a, S, V, ne);
            else if (nd == 6) gn_step<6, true, UNIFORM>(a, S, V, ne);
            else gn_step<12, true, UNIFORM>(a, S, V, ne);
        }
        if (tid == 0) {
            S.stop = stop ? 1 : 0;
            if (!stop) { S.obj_old = S.obj; S.it = i + 1; }
        }
    }
    __syncthreads();
    return S.stop ? 1 : 0;
}

// Result record of a finished point (dvc.cpp:263-283 records only the status of a failed
// point).  Thread 0.
__device__ __forceinline__ void write_result(const SearchArgs &a, const PtShared &S, int n) {
    dvc_point_result *r = &a.res[n];
    r->status = S.status;
    r->iters = S.iters;
    if (S.status == DVC_POINT_GOOD) {
        r->obj = S.obj;
        for (int q = 0; q < 12; q++) r->par[q] = (q < a.ndof) ? S.par[q] : 0.0;
    } else {
        r->obj = 0.0;
        for (int q = 0; q < 12; q++) r->par[q] = 0.0;
    }
}

// Carried point state between launches: the head of PtShared (kStateBytes), copied by
// every thread of the workgroup in 8-byte words.  Callers synchronise around the copies.
template <int B>
__device__ __forceinline__ void state_load(PtShared &S, const unsigned char *g) {
    const unsigned long long *src = (const unsigned long long *)g;
    unsigned long long *dst = (unsigned long long *)&S;
    for (int w = threadIdx.x; w < kStateBytes / 8; w += B) dst[w] = src[w];
}
template <int B>
__device__ __forceinline__ void state_store(const PtShared &S, unsigned char *g) {
    const unsigned long long *src = (const unsigned long long *)&S;
    unsigned long long *dst = (unsigned long long *)g;
    for (int w = threadIdx.x; w < kStateBytes / 8; w += B) dst[w] = src[w];
}

// Wavefronts per SIMD the search kernel is compiled for: three for workgroups of up to
// 256 threads (at most 168 registers, so three such workgroups share a CU and hide each
// other's gather latency), else what the workgroup size allows.
template <int B>
constexpr int search_waves_per_simd() { return B <= 256 ? 3 : 1; }

// One dependency level of the point loop: workgroup b searches point level_pts[b].
// All neighbours a point depends on belong to earlier levels (earlier launches).
template <int BPV, int B, bool UNEVEN>
__global__ void __launch_bounds__(B, search_waves_per_simd<B>()) search_kernel(SearchArgs a, const int *level_pts) {
    DVC_DYN_LDS(float, dyn);
    // [MFMA staging tiles][reference samples, when not read from HBM]
    float *stage = dyn;
    float *rcb = dyn + (B / 64) * 64 * 17;
    DVC_SHARED(PtShared, S);
    DVC_SHARED(double[(B / 64) * 9], red);
    DVC_SHARED(double[256], V);
    const int tid = threadIdx.x;
    const int n = level_pts[blockIdx.x];
    if (a.repair && !a.repair[n]) return;         // a repair launch: this point's result stands
    const int p = a.pos[n];
    if constexpr (UNEVEN) {
        DVC_SHARED(SearchArgs, la);
        if (tid == 0) apply_geo(la, a, n);
        __syncthreads();
        point_search<BPV, B>(la, p, n, S, rcb, stage, red, V);
    } else {
        point_search<BPV, B>(a, p, n, S, rcb, stage, red, V);
    }
    __syncthreads();
    if (tid == 0) write_result(a, S, n);
}

// Workgroups of a point with per-point geometry in a slice launch: one per sa.per of its
// samples, at most sa.maxsl when that is set.
__device__ __forceinline__ int point_slices(int nsamp, const SliceArgs &sa) {
    const int n = (nsamp + sa.per - 1) / sa.per;
    return sa.maxsl > 0 && n > sa.maxsl ? sa.maxsl : n;
}

// The points of a chain launch after its slice launches (rs.t of them): workgroup b
// continues point level_pts[b] from the state and slice sums that slice launch rs.t - 1
// left.  Finished points are skipped; the others end their sliced evaluation and run
// their remaining Gauss-Newton iterations here, one workgroup per point.  nsl: the point's
// slices (rs.nslice, or fewer for a point with fewer samples than the launch's largest).
template <int BPV, int B>
__device__ __forceinline__ void resume_body(const SearchArgs &a, const int *level_pts, const SliceArgs &rs, int nsl,
                                            float *stage, PtShared &S, double *red, double *V) {
    const int n = level_pts[blockIdx.x];
    const int p = a.pos[n];
    const int par = (rs.t - 1) & 1;
    const double *sp = rs.part[par] + (size_t)blockIdx.x * rs.nslice * kPartStride;
    double pre[kPrefetch];
    prefetch_slices<B>(sp, nsl, pre);
    state_load<B>(S, rs.state[par] + (size_t)blockIdx.x * kStateStride);
    __syncthreads();
    if (S.done) return;
    const int r = finish_sliced<B, false>(a, S, sp, nsl, V, stage, pre);
    if (r == 0) gn_iterations<BPV, B>(a, S, a.rvals + (size_t)p * a.nsamp_pad, stage, red, V, S.it, S.obj_old);
    __syncthreads();
    if (threadIdx.x == 0) write_result(a, S, n);
}

// UNEVEN: every point its own geometry (SearchArgs::pgeo), copied into an LDS SearchArgs
// first, as search_kernel does.
template <int BPV, int B, bool UNEVEN>
__global__ void __launch_bounds__(B) resume_kernel(SearchArgs a, const int *level_pts, SliceArgs rs) {
    DVC_DYN_LDS(float, dyn);
    DVC_SHARED(PtShared, S);
    DVC_SHARED(double[(B / 64) * 9], red);
    DVC_SHARED(double[256], V);
    if constexpr (UNEVEN) {
        DVC_SHARED(SearchArgs, la);
        const int n = level_pts[blockIdx.x];
        if (threadIdx.x == 0) apply_geo(la, a, n);
        __syncthreads();
        resume_body<BPV, B>(la, level_pts, rs, point_slices(la.nsamp, rs), dyn, S, red, V);
    } else {
        resume_body<BPV, B>(a, level_pts, rs, rs.nslice, dyn, S, red, V);
    }
}

// One slice launch (SliceArgs) of the points of a chain launch.  Launch t ends evaluation
// t - 1 of every point still searching (finish_sliced, on the slice sums and state that
// launch t - 1 left; at t = 0 instead the point's set-up and start guess), then runs the
// set-up of evaluation t and its own slice of the samples, and writes the slice's sums;
// slice 0 also writes the point's state.  A point whose search ended writes its result
// (slice 0) and carries its finished state forward.  Every workgroup of a point computes
// the same set-up, sums and step from the same inputs, so all agree.  nsl: the point's
// slices; sa.nslice, the launch's largest, is the stride of the slice records.
template <int BPV, int BS>
__device__ __forceinline__ void slice_body(const SearchArgs &a, const SliceArgs &sa, int slot, int slice, int nsl,
                                           float *stage, PtShared &S, double *red, double *V) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int n = sa.level_pts[slot];
    const int p = a.pos[n];
    const int cur = sa.t & 1, prev = cur ^ 1;
    unsigned char *gst = sa.state[cur] + (size_t)slot * kStateStride;
    const float *rsrc = a.rvals + (size_t)p * a.nsamp_pad;
    // the point's search has ended: its result, and its state for the later launches
    auto finished = [&]() {
        if (slice != 0) return;
        if (tid == 0) { S.done = 1; write_result(a, S, n); }
        __syncthreads();
        state_store<BS>(S, gst);
    };
    if (sa.t == 0) {
        if (a.repair && !a.repair[n]) {
            // a judged search's repair launch: this point's result stands; its state
            // only carries "finished" to the later launches
            if (slice == 0) {
                if (tid == 0) S.done = 1;
                __syncthreads();
                state_store<BS>(S, gst);
            }
            return;
        }
        const float *rs = nullptr;
        if (!point_start<BPV, BS, false>(a, p, n, S, nullptr, stage, red, V, rs)) {
            finished();
            return;
        }
        if (tid == 0) { S.it = 0; S.obj_old = 0.0; S.done = 0; }
    } else {
        // the slice sums are loaded while the state is: both come from launch t - 1
        const double *sp = sa.part[prev] + (size_t)slot * sa.nslice * kPartStride;
        double pre[kPrefetch];
        prefetch_slices<BS>(sp, nsl, pre);
        state_load<BS>(S, sa.state[prev] + (size_t)slot * kStateStride);
        __syncthreads();
        if (S.done) {
            if (slice == 0) state_store<BS>(S, gst);
            return;
        }
        const int r = finish_sliced<BS, true>(a, S, sp, nsl, V, stage, pre);
        if (r != 0) {
            finished();
            return;
        }
    }
    // ---- evaluation S.it (Search.cpp:194-246): set-up, then this workgroup's samples
    if (tid == 0) {
        for (int q = 0; q < 12; q++) S.evalpar[q] = S.par[q];
        S.iters += 1;
    }
    __syncthreads();
    const int i = S.it, nd = a.ndof;
    const double del_mag = sqrt(S.upd[0] * S.upd[0] + S.upd[1] * S.upd[1] + S.upd[2] * S.upd[2]);
    const bool need_j = !(i > 0 && del_mag < 0.01);
    if (!eval_begin<BS>(a, S, nd)) {
        if (tid == 0) S.status = DVC_RANGE_FAIL;
        finished();
        return;
    }
    // chunks of 64 samples dealt to the slices in turn: every slice gets parts of the
    // whole cloud, so slices cost the same where rows of the cloud are short (its caps)
    eval_samples<BPV, BS, 1>(a, S, nd, need_j, rsrc, stage, red, V, slice, nsl);
    if (tid < 64) {
        double *gp = sa.part[cur] + ((size_t)slot * sa.nslice + slice) * kPartStride;
        const int ne = need_j ? kPartStride : 16;
        for (int e = lane; e < ne; e += 64)
            gp[e] = e < 3 ? S.sums[e] : (e < 9 ? S.bbr[e - 3] : (e < 16 ? 0.0 : V[e - 16]));
    }
    if (slice == 0) state_store<BS>(S, gst);
}

// UNEVEN: every point its own geometry (SearchArgs::pgeo) and its own slice count; the
// workgroups past a point's slices leave at once.
template <int BPV, int BS, bool UNEVEN>
__global__ void __launch_bounds__(BS) slice_kernel(SearchArgs a, SliceArgs sa) {
    DVC_DYN_LDS(float, dyn);
    DVC_SHARED(PtShared, S);
    DVC_SHARED(double[(BS / 64) * 9], red);
    DVC_SHARED(double[256], V);
    const int slot = (int)blockIdx.x % sa.cnt8, slice = (int)blockIdx.x / sa.cnt8;
    if (slot >= sa.cnt) return;
    if constexpr (UNEVEN) {
        const int n = sa.level_pts[slot];
        const int nsl = point_slices(a.pgeo[n].nsamp, sa);
        if (slice >= nsl) return;
        DVC_SHARED(SearchArgs, la);
        if (threadIdx.x == 0) apply_geo(la, a, n);
        __syncthreads();
        slice_body<BPV, BS>(la, sa, slot, slice, nsl, dyn, S, red, V);
    } else {
        slice_body<BPV, BS>(a, sa, slot, slice, sa.nslice, dyn, S, red, V);
    }
}

#ifndef DVC_EMU
// ---------------------------------------------------- launchers per instantiation

// The search and reference kernels of one (bytes per voxel, workgroup size) pair.  The
// library compiles this file in several translation units at once: with DVC_PART = k > 0
// a unit instantiates only the pairs of part k (below); DVC_PART = 0 holds everything
// else and refers to those instantiations.
template <int BPV, int B>
struct Inst {
    static int launch_search(const SearchArgs &a, const int *level_pts, int grid, size_t lds, hipStream_t st);
    static int lds_limit(size_t lds_total);
    static int launch_reference(const SearchArgs &a, hipStream_t st);
    static int occupancy(size_t lds, int *bpc);
    static int static_lds(size_t *bytes);
};

template <int BPV, int BS>
struct SliceInst {
    static int launch(const SearchArgs &a, const SliceArgs &sa, hipStream_t st);
    // resume_kernel of BS threads per point
    static int resume(const SearchArgs &a, const SliceArgs &rs, hipStream_t st);
};

template <int BPV, int B>
int Inst<BPV, B>::launch_search(const SearchArgs &a, const int *level_pts, int grid, size_t lds, hipStream_t st) {
    if (a.pgeo) hipLaunchKernelGGL((search_kernel<BPV, B, true>), dim3(grid), dim3(B), lds, st, a, level_pts);
    else hipLaunchKernelGGL((search_kernel<BPV, B, false>), dim3(grid), dim3(B), lds, st, a, level_pts);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Raise the dynamic-LDS limit of one instantiation to everything its static LDS leaves
// of `lds_total`.  Done once per context, outside any stream capture, so that launches
// (direct or replayed from a graph) never change function attributes.
template <int BPV, int B>
int Inst<BPV, B>::lds_limit(size_t lds_total) {
    const void *ks[4] = {(const void *)search_kernel<BPV, B, false>, (const void *)reference_kernel<BPV, B, false>,
                         (const void *)search_kernel<BPV, B, true>, (const void *)reference_kernel<BPV, B, true>};
    for (const void *k : ks) {
        hipFuncAttributes fa;
        if (hipFuncGetAttributes(&fa, k) != hipSuccess) return -1;
        if (fa.sharedSizeBytes >= lds_total) return -1;
        int dyn = (int)(lds_total - fa.sharedSizeBytes);
        if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, dyn) != hipSuccess) return -1;
    }
    return 0;
}

// reference_kernel uses static LDS only
template <int BPV, int B>
int Inst<BPV, B>::launch_reference(const SearchArgs &a, hipStream_t st) {
    if (a.pgeo) hipLaunchKernelGGL((reference_kernel<BPV, B, true>), dim3(a.nproc), dim3(B), 0, st, a);
    else hipLaunchKernelGGL((reference_kernel<BPV, B, false>), dim3(a.nproc), dim3(B), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int BPV, int B>
int Inst<BPV, B>::occupancy(size_t lds, int *bpc) {
    auto k = search_kernel<BPV, B, false>;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, (const void *)k, B, lds) == hipSuccess ? 0 : -1;
}

template <int BPV, int B>
int Inst<BPV, B>::static_lds(size_t *bytes) {
    hipFuncAttributes fa;
    // the per-point-geometry instantiation holds one more shared SearchArgs: the larger
    if (hipFuncGetAttributes(&fa, (const void *)search_kernel<BPV, B, true>) != hipSuccess) return -1;
    *bytes = fa.sharedSizeBytes;
    return 0;
}

template <int BPV, int BS>
int SliceInst<BPV, BS>::launch(const SearchArgs &a, const SliceArgs &sa, hipStream_t st) {
    const size_t lds = (size_t)(BS / 64) * 64 * 17 * sizeof(float);
    const unsigned grid = (unsigned)sa.cnt8 * (unsigned)sa.nslice;
    if (a.pgeo) hipLaunchKernelGGL((slice_kernel<BPV, BS, true>), dim3(grid), dim3(BS), lds, st, a, sa);
    else hipLaunchKernelGGL((slice_kernel<BPV, BS, false>), dim3(grid), dim3(BS), lds, st, a, sa);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int BPV, int BS>
int SliceInst<BPV, BS>::resume(const SearchArgs &a, const SliceArgs &rs, hipStream_t st) {
    const size_t lds = (size_t)(BS / 64) * 64 * 17 * sizeof(float);
    if (a.pgeo) hipLaunchKernelGGL((resume_kernel<BPV, BS, true>), dim3(rs.cnt), dim3(BS), lds, st, a, rs.level_pts, rs);
    else hipLaunchKernelGGL((resume_kernel<BPV, BS, false>), dim3(rs.cnt), dim3(BS), lds, st, a, rs.level_pts, rs);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

#define DVC_INST_LIST(X) X(1, 64) X(1, 128) X(1, 256) X(1, 512) X(1, 1024) X(2, 64) X(2, 128) X(2, 256) X(2, 512) X(2, 1024)

#if defined(DVC_PART) && DVC_PART > 0
#if DVC_PART == 1
template struct Inst<1, 64>;
template struct Inst<1, 128>;
template struct Inst<1, 256>;
#elif DVC_PART == 2
template struct Inst<1, 512>;
template struct Inst<1, 1024>;
#elif DVC_PART == 3
template struct Inst<2, 64>;
template struct Inst<2, 128>;
template struct Inst<2, 256>;
#elif DVC_PART == 4
template struct Inst<2, 512>;
template struct Inst<2, 1024>;
#elif DVC_PART == 5
template struct SliceInst<1, 256>;
template struct SliceInst<2, 256>;
template struct SliceInst<1, 512>;
template struct SliceInst<2, 512>;
#endif
#else  // the main unit
#ifdef DVC_PART
#define DVC_EXTERN(BPV_, B_) extern template struct Inst<BPV_, B_>;
DVC_INST_LIST(DVC_EXTERN)
extern template struct SliceInst<1, 256>;
extern template struct SliceInst<2, 256>;
extern template struct SliceInst<1, 512>;
extern template struct SliceInst<2, 512>;
#undef DVC_EXTERN
#endif

size_t search_lds_bytes(int nsamp_pad, int block_threads, bool samples_in_lds) {
    return (samples_in_lds ? (size_t)nsamp_pad * sizeof(float) : 0) +
           (size_t)(block_threads / 64) * 64 * 17 * sizeof(float);
}

#define DVC_DISPATCH(BPV_, B_, FN, ...) \
    if (bpv == BPV_ && block_threads == B_) return Inst<BPV_, B_>::FN(__VA_ARGS__);

int launch_search(const SearchArgs &a, const int *level_pts, int block_threads, int grid, size_t lds, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    const int bpv = a.bpv;
#define X(BPV_, B_) DVC_DISPATCH(BPV_, B_, launch_search, a, level_pts, grid, lds, st)
    DVC_INST_LIST(X)
#undef X
    return -1;
}

int launch_slice(const SearchArgs &a, const SliceArgs &sa, int threads, void *stream) {
    if (sa.cnt <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (a.bpv == 1 && threads == 256) return SliceInst<1, 256>::launch(a, sa, st);
    if (a.bpv == 2 && threads == 256) return SliceInst<2, 256>::launch(a, sa, st);
    if (a.bpv == 1 && threads == 512) return SliceInst<1, 512>::launch(a, sa, st);
    if (a.bpv == 2 && threads == 512) return SliceInst<2, 512>::launch(a, sa, st);
    return -1;
}

int launch_resume(const SearchArgs &a, const SliceArgs &rs, void *stream) {
    if (rs.cnt <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (a.bpv == 1) return SliceInst<1, 512>::resume(a, rs, st);
    if (a.bpv == 2) return SliceInst<2, 512>::resume(a, rs, st);
    return -1;
}

int search_static_lds(int bpv, int block_threads, size_t *bytes) {
#define X(BPV_, B_) DVC_DISPATCH(BPV_, B_, static_lds, bytes)
    DVC_INST_LIST(X)
#undef X
    return -1;
}

int launch_reference(const SearchArgs &a, int block_threads, void *stream) {
    if (a.nproc <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const int bpv = a.bpv;
#define X(BPV_, B_) DVC_DISPATCH(BPV_, B_, launch_reference, a, st)
    DVC_INST_LIST(X)
#undef X
    return -1;
}

int search_set_lds_limit(size_t lds_total) {
    for (int bpv = 1; bpv <= 2; bpv++)
        for (int block_threads = 64; block_threads <= 1024; block_threads *= 2) {
            int r = -1;
            [&] {
#define X(BPV_, B_) \
    if (bpv == BPV_ && block_threads == B_) { r = Inst<BPV_, B_>::lds_limit(lds_total); return; }
                DVC_INST_LIST(X)
#undef X
            }();
            if (r != 0) return -1;
        }
    return 0;
}

int search_max_blocks(int bpv, int block_threads, size_t lds, int *bpc) {
#define X(BPV_, B_) DVC_DISPATCH(BPV_, B_, occupancy, lds, bpc)
    DVC_INST_LIST(X)
#undef X
    return -1;
}
#endif  // the main unit
#endif  // DVC_EMU

// The remaining kernels and their launchers belong to the main unit.
#if !defined(DVC_PART) || DVC_PART == 0

// ------------------------------------------------------ interpolation kernel

template <int BPV>
__global__ void interp_kernel(VolView v, BoxView bx, const double *xyz, int n, int method,
                              double *out, double *grad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gx, gy, gz;
    double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    int cx = clampi((int)x - bx.b[0], 0, bx.n[0] - 1), cy = clampi((int)y - bx.b[1], 0, bx.n[1] - 1);
    int cz = clampi((int)z - bx.b[2], 0, bx.n[2] - 1);
    float bias = voxel<BPV>(v, bx.b[0] + cx, bx.b[1] + cy, bx.b[2] + cz);
    float val = interp_at<BPV>(v, bx, x, y, z, method, bias, grad != nullptr, gx, gy, gz);
    out[i] = (double)val + (double)bias;
    if (grad) { grad[3 * i] = gx; grad[3 * i + 1] = gy; grad[3 * i + 2] = gz; }
}

#ifndef DVC_EMU
int launch_interp(const uint8_t *vol, int bpv, int W, int H, int T, const int b0[3], const int box[3],
                  const double *xyz, int n, int method, double *out, double *grad, void *stream) {
    (void)T;
    VolView v = {vol, W, (long long)W * H};
    BoxView bx = {{b0[0], b0[1], b0[2]}, {box[0], box[1], box[2]}};
    int blocks = (n + 255) / 256;
    if (blocks == 0) return 0;
    if (bpv == 1) hipLaunchKernelGGL(interp_kernel<1>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, v, bx, xyz, n, method, out, grad);
    else hipLaunchKernelGGL(interp_kernel<2>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, v, bx, xyz, n, method, out, grad);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

#endif  // DVC_EMU

// ------------------------------------------------------------ objective kernel

// obj_* of ObjectiveFunctions.cpp:24-263 for one pair of sample vectors (FP64), one
// workgroup of 1024 threads.
__global__ void __launch_bounds__(1024) objective_kernel(int kind, const double *r, const double *t, int n,
                                                         double *obj, double *res) {
    DVC_SHARED(double[16 * 8], red);
    DVC_SHARED(double[8], sums);
    const int tid = threadIdx.x;
    double s[2] = {0.0, 0.0};
    for (int i = tid; i < n; i += 1024) { s[0] += r[i]; s[1] += t[i]; }
    block_sum<1024, 2>(s, red, sums);
    double ar = sums[0] / n, at = sums[1] / n;
    double q[2] = {0.0, 0.0};
    for (int i = tid; i < n; i += 1024) {
        if (kind == DVC_OBJ_NSSD) { q[0] += r[i] * r[i]; q[1] += t[i] * t[i]; }
        else { q[0] += (r[i] - ar) * (r[i] - ar); q[1] += (t[i] - at) * (t[i] - at); }
    }
    block_sum<1024, 2>(q, red, &sums[2]);
    double sr = sqrt(sums[2]), stt = sqrt(sums[3]);
    double o[1] = {0.0};
    for (int i = tid; i < n; i += 1024) {
        double d;
        if (kind == DVC_OBJ_SAD || kind == DVC_OBJ_SSD) d = t[i] - r[i];
        else if (kind == DVC_OBJ_ZSSD) d = (t[i] - at) - (r[i] - ar);
        else if (kind == DVC_OBJ_NSSD) d = t[i] / stt - r[i] / sr;
        else d = ((t[i] - at) / stt) - ((r[i] - ar) / sr);
        if (res) res[i] = d;
        o[0] += (kind == DVC_OBJ_SAD) ? fabs(d) : d * d;
    }
    block_sum<1024, 1>(o, red, &sums[4]);
    if (tid == 0) {
        double v = sums[4];
        if (kind == DVC_OBJ_NSSD) v /= 2.0;
        if (kind == DVC_OBJ_ZNSSD) v /= 4.0;
        *obj = v;
    }
}

#ifndef DVC_EMU
int launch_objective(int kind, const double *ref, const double *tar, int n, double *obj, double *res, void *stream) {
    hipLaunchKernelGGL(objective_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, kind, ref, tar, n, obj, res);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

#endif  // DVC_EMU

// ------------------------------------------------------------- small kernels

// Selection between a proposal pass and its judge pass (the reference's start guess
// applied to the proposal pass's results).  Where both converged and agree
// (translations within 1e-3 voxel, objectives within 1e-5 relative) the judge's result
// is the point's (kept 0).  Every other point is marked for repair (kept 2): repair
// launches search it again in the serial chain's order from the reference's start guess
// over the final results.
__global__ void judge_kernel(const int *order, int n, const dvc_point_result *prop,
                             const dvc_point_result *judge, dvc_point_result *out, int *kept, int *repair) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = order[i];
    const dvc_point_result &a = prop[m];
    const dvc_point_result &b = judge[m];
    bool agree = a.status == DVC_POINT_GOOD && b.status == DVC_POINT_GOOD;
    if (agree) {
        double du = 0.0;
        for (int q = 0; q < 3; q++) du = fmax(du, fabs(a.par[q] - b.par[q]));
        agree = du <= 1e-3 && fabs(a.obj - b.obj) <= 1e-5 * fabs(b.obj);
    }
    if (agree) {
        out[m] = b;
        kept[m] = 0;
        repair[m] = 0;
    } else {
        out[m] = b;
        kept[m] = 2;
        repair[m] = 1;
    }
}

#ifndef DVC_EMU
int launch_judge(const int *order, int n, const dvc_point_result *prop, const dvc_point_result *judge,
                 dvc_point_result *out, int *kept, int *repair, void *stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(judge_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, order, n, prop,
                       judge, out, kept, repair);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
#endif

__global__ void zero_int_kernel(int *p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long st = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = 0;
}

#ifndef DVC_EMU
int launch_zero_int(int *p, long long n, void *stream) {
    if (n <= 0) return 0;
    long long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(zero_int_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, p, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

#endif  // DVC_EMU

__global__ void swap16_kernel(uint16_t *p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long st = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += st) { uint16_t v = p[i]; p[i] = (uint16_t)((v >> 8) | (v << 8)); }
}

#ifndef DVC_EMU
int launch_swap16(uint16_t *p, long long n, void *stream) {
    if (n <= 0) return 0;
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(swap16_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, p, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

#endif  // DVC_EMU

// ---------------------------------------------------------------- strain fit

// FullPivHouseholderQR(A).solve(b) for the 10x10 normal matrix (Eigen 3.4
// FullPivHouseholderQR, used at strain.cpp:533).  Q is column major, overwritten.
__device__ void fullpiv_factor(double *Q, int n, double *hc, int *rt, int *ct, int *rank_out) {
#pragma clang fp contract(off)
    const double prec = 2.220446049250313e-16 * (double)n;
    double biggest = 0.0, maxpivot = 0.0;
    int nonzero = n;
    for (int k = 0; k < n; k++) {
        int rb = k, cb = k;
        double sc = -1.0;
        for (int j = k; j < n; j++)
            for (int i = k; i < n; i++) {
                double a = fabs(Q[j * n + i]);
                if (a > sc) { sc = a; rb = i; cb = j; }
            }
        if (k == 0) biggest = sc;
        if (sc <= biggest * prec) {
            nonzero = k;
            for (int i = k; i < n; i++) { rt[i] = i; ct[i] = i; hc[i] = 0.0; }
            break;
        }
        rt[k] = rb; ct[k] = cb;
        if (k != rb)
            for (int j = k; j < n; j++) { double t = Q[j * n + k]; Q[j * n + k] = Q[j * n + rb]; Q[j * n + rb] = t; }
        if (k != cb)
            for (int i = 0; i < n; i++) { double t = Q[k * n + i]; Q[k * n + i] = Q[cb * n + i]; Q[cb * n + i] = t; }
        double *col = &Q[k * n + k];
        int m = n - k;
        double tail = 0.0;
        for (int i = 1; i < m; i++) tail += col[i] * col[i];
        double c0 = col[0], tau, beta;
        if (tail <= 2.2250738585072014e-308) {
            tau = 0.0; beta = c0;
            for (int i = 1; i < m; i++) col[i] = 0.0;
        } else {
            beta = sqrt(c0 * c0 + tail);
            if (c0 >= 0.0) beta = -beta;
            for (int i = 1; i < m; i++) col[i] = col[i] / (c0 - beta);
            tau = (beta - c0) / beta;
        }
        hc[k] = tau;
        col[0] = beta;
        if (fabs(beta) > maxpivot) maxpivot = fabs(beta);
        if (m == 1) {
            for (int j = k + 1; j < n; j++) Q[j * n + k] *= (1.0 - tau);
        } else if (tau != 0.0) {
            for (int j = k + 1; j < n; j++) {
                double *cj = &Q[j * n + k];
                double t = 0.0;
                for (int i = 1; i < m; i++) t += col[i] * cj[i];
                t += cj[0];
                cj[0] -= tau * t;
                for (int i = 1; i < m; i++) cj[i] -= tau * col[i] * t;
            }
        }
    }
    double pthr = fabs(maxpivot) * (2.220446049250313e-16 * (double)n);
    int rank = 0;
    for (int i = 0; i < nonzero; i++) rank += (fabs(Q[i * n + i]) > pthr);
    *rank_out = rank;
}

__device__ void fullpiv_apply(const double *Q, int n, const double *hc, const int *rt, const int *ct, int rank,
                              const double *b, double *x) {
#pragma clang fp contract(off)
    double c[10];
    int perm[10];
    for (int i = 0; i < n; i++) { perm[i] = i; x[i] = 0.0; c[i] = b[i]; }
    for (int k = 0; k < n; k++) { int t = perm[k]; perm[k] = perm[ct[k]]; perm[ct[k]] = t; }
    if (rank == 0) return;
    for (int k = 0; k < rank; k++) {
        double t0 = c[k]; c[k] = c[rt[k]]; c[rt[k]] = t0;
        int m = n - k;
        double tau = hc[k];
        if (m == 1) { c[k] *= (1.0 - tau); continue; }
        if (tau == 0.0) continue;
        double t = 0.0;
        for (int i = 1; i < m; i++) t += Q[k * n + k + i] * c[k + i];
        t += c[k];
        c[k] -= tau * t;
        for (int i = 1; i < m; i++) c[k + i] -= tau * Q[k * n + k + i] * t;
    }
    for (int i = rank - 1; i >= 0; i--) {
        double s = c[i];
        for (int j = i + 1; j < rank; j++) s -= Q[j * n + i] * c[j];
        c[i] = s / Q[i * n + i];
    }
    for (int i = 0; i < rank; i++) x[perm[i]] = c[i];
}

// The same factorisation and solve on one wavefront, for the 10x10 strain fit with its
// three right-hand sides: lane j < 10 holds column j of the normal matrix, lanes 10..12
// the right-hand sides X^T d.  The pivot search takes each column's first largest
// entry and then the first column holding the largest (the serial column-major scan);
// row and column swaps, reflectors and the rank rule follow fullpiv_factor and
// fullpiv_apply operation by operation.  The reflectors are applied to the right-hand
// sides as they are formed: those past the rank only change entries the back
// substitution does not read.  Solutions land in par[3][10].
__device__ void fullpiv_wave10(const double *M, double (*par)[10]) {
#pragma clang fp contract(off)
    constexpr int N = 10;
    const int lane = threadIdx.x & 63;
    const bool is_col = lane < N, is_rhs = lane >= N && lane < N + 3;
    double col[N];
#pragma unroll
    for (int i = 0; i < N; i++)
        col[i] = is_col ? M[lane * N + i] : (is_rhs ? M[100 + (lane - N) * N + i] : 0.0);
    const double prec = 2.220446049250313e-16 * (double)N;
    double biggest = 0.0, maxpivot = 0.0;
    int nonzero = N;
    int rt[N], ct[N];
    double hc[N];
#pragma unroll
    for (int k = 0; k < N; k++) { rt[k] = k; ct[k] = k; hc[k] = 0.0; }
#pragma unroll
    for (int k = 0; k < N; k++) {
        if (nonzero != N) continue;      // (no break: keeps the loop fully unrolled)
        // first largest |entry| of this lane's column among rows k..N-1
        double cm = -1.0;
        int ci = k;
#pragma unroll
        for (int i = k; i < N; i++) {
            double av = fabs(col[i]);
            if (av > cm) { cm = av; ci = i; }
        }
        double sc = -1.0;
        int rb = k, cb = k;
#pragma unroll
        for (int j = k; j < N; j++) {
            double v = rdlane(cm, j);
            int r = rdlane_i(ci, j);
            if (v > sc) { sc = v; rb = r; cb = j; }
        }
        if (k == 0) biggest = sc;
        if (sc <= biggest * prec) {
            nonzero = k;
            continue;
        }
        rt[k] = rb; ct[k] = cb;
        // row swap k <-> rb in columns k.. (and in the right-hand sides)
        if (k != rb && ((lane >= k && lane < N) || is_rhs)) {
            double tk = col[k], tr = 0.0;
#pragma unroll
            for (int i = k + 1; i < N; i++)
                if (i == rb) tr = col[i];
            col[k] = tr;
#pragma unroll
            for (int i = k + 1; i < N; i++)
                if (i == rb) col[i] = tk;
        }
        // column swap k <-> cb
        if (k != cb) {
            const int src = (lane == k) ? cb : ((lane == cb) ? k : lane);
#pragma unroll
            for (int i = 0; i < N; i++) col[i] = shfl_from(col[i], src);
        }
        // makeHouseholder of column k, rows k..N-1 (uniform)
        const int m = N - k;
        double ess[N];
        double tail = 0.0;
#pragma unroll
        for (int i = 1; i < m; i++) { ess[i] = rdlane(col[k + i], k); tail += ess[i] * ess[i]; }
        const double c0 = rdlane(col[k], k);
        double tau, beta;
        if (tail <= 2.2250738585072014e-308) {
            tau = 0.0; beta = c0;
#pragma unroll
            for (int i = 1; i < m; i++) ess[i] = 0.0;
        } else {
            beta = sqrt(c0 * c0 + tail);
            if (c0 >= 0.0) beta = -beta;
#pragma unroll
            for (int i = 1; i < m; i++) ess[i] = ess[i] / (c0 - beta);
            tau = (beta - c0) / beta;
        }
        hc[k] = tau;
        if (lane == k) {
            col[k] = beta;
#pragma unroll
            for (int i = 1; i < m; i++) col[k + i] = ess[i];
        }
        if (fabs(beta) > maxpivot) maxpivot = fabs(beta);
        if ((lane > k && lane < N) || is_rhs) {
            if (m == 1) {
                col[k] *= (1.0 - tau);
            } else if (tau != 0.0) {
                double t = 0.0;
#pragma unroll
                for (int i = 1; i < m; i++) t += ess[i] * col[k + i];
                t += col[k];
                col[k] -= tau * t;
#pragma unroll
                for (int i = 1; i < m; i++) col[k + i] -= tau * ess[i] * t;
            }
        }
    }
    const double pthr = fabs(maxpivot) * (2.220446049250313e-16 * (double)N);
    int rank = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
        if (i < nonzero) rank += (fabs(rdlane(col[i], i)) > pthr);
    // back substitution in each right-hand-side lane
    double c[N];
#pragma unroll
    for (int i = 0; i < N; i++) c[i] = col[i];
#pragma unroll
    for (int i = N - 1; i >= 0; i--) {
        if (i < rank) {
            double sacc = c[i];
#pragma unroll
            for (int j = i + 1; j < N; j++)
                if (j < rank) sacc -= rdlane(col[i], j) * c[j];
            c[i] = sacc / rdlane(col[i], i);
        }
    }
    // column permutation from the transpositions; x[perm[i]] = c[i]
    int perm[N];
#pragma unroll
    for (int i = 0; i < N; i++) perm[i] = i;
#pragma unroll
    for (int k = 0; k < N; k++) {
        int t = perm[k], pk = 0;
#pragma unroll
        for (int q = 0; q < N; q++)
            if (q == ct[k]) pk = perm[q];
        perm[k] = pk;
#pragma unroll
        for (int q = 0; q < N; q++)
            if (q == ct[k]) perm[q] = t;
    }
    if (is_rhs) {
        double *x = par[lane - N];
#pragma unroll
        for (int i = 0; i < N; i++) x[i] = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++)
            if (i < rank) x[perm[i]] = c[i];
    }
}

// Eigenvalues of a symmetric 3x3 (cyclic Jacobi), sorted descending
// (strain.cpp:553-556 takes EigenSolver real parts and sorts them).  From the fifth
// sweep on, an off-diagonal entry that no longer changes either diagonal entry it
// couples (|a_pp| + 100 |a_pq| == |a_pp|, and the same for a_qq) is set to zero instead
// of rotated, the classical cyclic-Jacobi stopping rule: rotations cannot drive the
// entries below the rounding of the diagonal, so without it every call would run all 50 sweeps.
__device__ void sym3_eigen(const double *S, double *lam) {
#pragma clang fp contract(off)
    double a[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = 0.5 * (S[i * 3 + j] + S[j * 3 + i]);
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
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
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
    double l0 = a[0][0], l1 = a[1][1], l2 = a[2][2], t;
    if (l0 < l1) { t = l0; l0 = l1; l1 = t; }
    if (l1 < l2) { t = l1; l1 = l2; l2 = t; }
    if (l0 < l1) { t = l0; l0 = l1; l1 = t; }
    lam[0] = l0; lam[1] = l1; lam[2] = l2;
}

// One wavefront per point: window selection (strain.cpp:424-448), design rows
// (:478-527), normal equations by lanes, FP64 solve and strain tensors (:529-590).
// All FP64 arithmetic keeps the reference's operation order without fused
// multiply-add, so windows too small to determine the 10-term fit (rank detection of
// the full-pivot QR) resolve exactly as the serial restatement does.
constexpr int kStrainWaves = 4;
__global__ void __launch_bounds__(64 * kStrainWaves) strain_kernel(StrainArgs a) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * kStrainWaves + w;
    DVC_SHARED(double[kStrainWaves][75][10], Xs);
    DVC_SHARED(double[kStrainWaves][75][3], Ds);
    DVC_SHARED(double[kStrainWaves][100 + 30], Ms);
    DVC_SHARED(int[kStrainWaves][76], sel);
    DVC_SHARED(double[kStrainWaves][3][10], Ps);
    if (i >= a.n) return;      // whole wave leaves together; no block barrier below
    const int *nb = a.neigh + (long long)i * a.k;
    double px = a.pts[3 * i], py = a.pts[3 * i + 1], pz = a.pts[3 * i + 2];
    double distance = sqrt(px * px + py * py + pz * pz);
    double scale = 1.0 / distance;
    double bx = scale * px, by = scale * py, bz = scale * pz;
    {
        // candidates in neighbour order; the first ndp that pass form the window.  Every
        // lane tests its own candidates, a ballot keeps the order.
        const int lim = a.refill ? (a.k < 75 ? a.k : 75) : (a.ndp < a.k ? a.ndp : a.k);
        int np = 0;
        for (int q0 = 0; q0 < lim && np < a.ndp; q0 += 64) {
            const int q = q0 + lane;
            bool ok = false;
            int m = -1;
            if (q < lim) {
                m = nb[q];
                ok = m >= 0 && m < a.n && a.status[m] == 0 && a.objmin[m] <= a.objt;
            }
            const unsigned long long mask = __ballot(ok);
            const int before = __popcll(mask & ((1ull << lane) - 1ull));
            if (ok && np + before < a.ndp) sel[w][np + before] = m;
            const int cnt = __popcll(mask);
            np = (np + cnt < a.ndp) ? np + cnt : a.ndp;
        }
        if (lane == 0) sel[w][75] = np;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int ngp = sel[w][75];
    if (lane < ngp) {
        int m = sel[w][lane];
        double x = scale * a.pts[3 * m], y = scale * a.pts[3 * m + 1], z = scale * a.pts[3 * m + 2];
        double *X = Xs[w][lane];
        X[0] = 1; X[1] = x; X[2] = y; X[3] = z;
        X[4] = x * x; X[5] = y * y; X[6] = z * z;
        X[7] = x * y; X[8] = y * z; X[9] = x * z;
        double *D = Ds[w][lane];
        D[0] = scale * a.disp[3 * m]; D[1] = scale * a.disp[3 * m + 1]; D[2] = scale * a.disp[3 * m + 2];
        if (a.yz) { X[1] = 0.0; X[4] = 0.0; X[7] = 0.0; X[9] = 0.0; D[0] = 0.0; }
        if (a.zx) { X[2] = 0.0; X[5] = 0.0; X[7] = 0.0; X[8] = 0.0; D[1] = 0.0; }
        if (a.xy) { X[3] = 0.0; X[6] = 0.0; X[8] = 0.0; X[9] = 0.0; D[2] = 0.0; }
    }
    if (lane + 64 < ngp) {
        // windows are at most 75 points: lanes 0..10 also build rows 64..74
        int r = lane + 64;
        int m = sel[w][r];
        double x = scale * a.pts[3 * m], y = scale * a.pts[3 * m + 1], z = scale * a.pts[3 * m + 2];
        double *X = Xs[w][r];
        X[0] = 1; X[1] = x; X[2] = y; X[3] = z;
        X[4] = x * x; X[5] = y * y; X[6] = z * z;
        X[7] = x * y; X[8] = y * z; X[9] = x * z;
        double *D = Ds[w][r];
        D[0] = scale * a.disp[3 * m]; D[1] = scale * a.disp[3 * m + 1]; D[2] = scale * a.disp[3 * m + 2];
        if (a.yz) { X[1] = 0.0; X[4] = 0.0; X[7] = 0.0; X[9] = 0.0; D[0] = 0.0; }
        if (a.zx) { X[2] = 0.0; X[5] = 0.0; X[7] = 0.0; X[8] = 0.0; D[1] = 0.0; }
        if (a.xy) { X[3] = 0.0; X[6] = 0.0; X[8] = 0.0; X[9] = 0.0; D[2] = 0.0; }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // X^T X (100 entries, column-major into Ms[0..99]) and X^T d (30 entries)
    for (int e = lane; e < 130; e += 64) {
        double s = 0.0;
        if (e < 100) {
            int r = e % 10, c = e / 10;
            for (int j = 0; j < ngp; j++) s += Xs[w][j][r] * Xs[w][j][c];
        } else {
            int r = (e - 100) % 10, c = (e - 100) / 10;
            for (int j = 0; j < ngp; j++) s += Xs[w][j][r] * Ds[w][j][c];
        }
        Ms[w][e] = s;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    fullpiv_wave10(Ms[w], Ps[w]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane != 0) return;
    double DG[9], displ[3];
    for (int c = 0; c < 3; c++) {
        const double *par = Ps[w][c];
        DG[c * 3 + 0] = par[1] + 2 * par[4] * bx + par[7] * by + par[9] * bz;
        DG[c * 3 + 1] = par[2] + 2 * par[5] * by + par[7] * bx + par[8] * bz;
        DG[c * 3 + 2] = par[3] + 2 * par[6] * bz + par[8] * by + par[9] * bx;
        displ[c] = distance * (par[0] + par[1] * bx + par[2] * by + par[3] * bz + par[4] * bx * bx +
                               par[5] * by * by + par[6] * bz * bz + par[7] * bx * by + par[8] * by * bz +
                               par[9] * bx * bz);
    }
    a.out_npts[i] = ngp;
    if (ngp == 0) {
        a.out_swr[i] = 0.0;
    } else {
        int m = sel[w][ngp - 1];
        double dx = bx - scale * a.pts[3 * m], dy = by - scale * a.pts[3 * m + 1], dz = bz - scale * a.pts[3 * m + 2];
        a.out_swr[i] = distance * sqrt(dx * dx + dy * dy + dz * dz);
    }
    for (int c = 0; c < 3; c++) a.out_disp[3 * i + c] = displ[c];
    for (int q = 0; q < 9; q++) a.out_DG[9 * i + q] = DG[q];
    const bool planar = a.xy || a.yz || a.zx;
    for (int pass = 0; pass < 2; pass++) {
        double ST[9];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                double v = DG[r * 3 + c] + DG[c * 3 + r];
                if (pass == 1) {
                    double s = 0.0;
                    for (int l = 0; l < 3; l++) s += DG[l * 3 + r] * DG[l * 3 + c];
                    v += s;
                }
                ST[r * 3 + c] = 0.5 * v;
            }
        double lam[3];
        sym3_eigen(ST, lam);
        if (planar) {
            if (lam[0] == 0.0) { lam[0] = lam[1]; lam[1] = lam[2]; }
            if (lam[1] == 0.0) { lam[1] = lam[2]; }
            lam[2] = 0.0;
        }
        double *o = (pass == 0 ? a.out_E : a.out_L) + 9 * i;
        o[0] = ST[0]; o[1] = ST[4]; o[2] = ST[8]; o[3] = ST[1]; o[4] = ST[5]; o[5] = ST[2];
        o[6] = lam[0]; o[7] = lam[1]; o[8] = lam[2];
    }
}

#ifndef DVC_EMU
// The strain fit's inputs from a search's result records: displacement, status and
// objective per point; points outside the searched prefix of the order are not_search
// (DataCloud.cpp:168-170).
__global__ void fields_kernel(const dvc_point_result *res, const int *pos, int npts, int nproc, double *disp,
                              int *status, double *objmin) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= npts) return;
    const dvc_point_result &r = res[n];
    const bool searched = pos[n] < nproc;
    for (int q = 0; q < 3; q++) disp[3 * n + q] = searched ? r.par[q] : 0.0;
    status[n] = searched ? r.status : DVC_NOT_SEARCH;
    objmin[n] = searched ? r.obj : 0.0;
}

int launch_fields(const dvc_point_result *res, const int *pos, int npts, int nproc, double *disp, int *status,
                  double *objmin, void *stream) {
    if (npts <= 0) return 0;
    hipLaunchKernelGGL(fields_kernel, dim3((npts + 255) / 256), dim3(256), 0, (hipStream_t)stream, res, pos, npts,
                       nproc, disp, status, objmin);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_strain(const StrainArgs &a, void *stream) {
    if (a.n <= 0) return 0;
    int blocks = (a.n + kStrainWaves - 1) / kStrainWaves;
    hipLaunchKernelGGL(strain_kernel, dim3(blocks), dim3(64 * kStrainWaves), 0, (hipStream_t)stream, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

#endif  // DVC_EMU

#endif  // the main unit

}  // namespace dvc
