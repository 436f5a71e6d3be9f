// dvc_net.hip -- the learned start-warp proposal: a small 3D convolutional network that
// maps a 32^3 reference/target crop pair around a search point to a 12-parameter warp
// (u v w phi the psi exx eyy ezz exy eyz exz, the parameter vector of SearchParams,
// Core/SearchParams.cpp:53-107).  The reference has no counterpart: it starts every
// point from its neighbours' mean translation (Search.cpp:277-377).  The proposal only
// feeds the start of the same Gauss-Newton search, and a judge pass from the
// reference's own start guess decides which result is kept (dvc_amd.h, start modes).
//
// Network (channels-last activations, BF16 storage, FP32 accumulation):
//   crop 2 x 32^3 (each channel normalised to zero mean, unit variance)
//   4 x [conv 3x3x3, stride 2, pad 1, + bias, ReLU]: 2->16 (16^3), 16->32 (8^3),
//        32->64 (4^3), 64->128 (2^3)
//   mean over the 2^3 voxels -> 128 features -> linear 128 -> 12 (FP32), times the
//   per-parameter scale (voxels for u v w, radians for the angles, strain units).
// Every convolution is an im2col matrix times the weights on the matrix cores
// (v_mfma_f32_16x16x32_bf16); its weight gradient is the transposed pair of the same
// GEMM split over rows, and its input gradient the GEMM with the transposed weights
// followed by a col2im gather.  Training uses Adam on FP32 master weights.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <algorithm>
#include <string>
#include <vector>
#include "dvc_internal.h"

namespace dvc {
namespace net {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
    return __uint_as_float((uint32_t)h << 16);
}
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);           // round to nearest even
    return (uint16_t)(u >> 16);
}

// ------------------------------------------------------------------ crops

// One workgroup per point: the 32^3 box whose voxel 15 (per axis) holds the point's
// truncated position, from both volumes, each normalised to zero mean and unit
// variance, stored channels-last [32][32][32][2] in BF16.  Voxels outside the volume
// repeat the nearest face voxel.
constexpr int kCropThreads = 1024;      // threads of the crop kernels (16 waves)
constexpr int kCropBatch = 8;           // voxels per thread whose loads are issued together

// Sums of both channels over the box (every thread its voxels i = tid + k*1024 in order,
// then the wave sums, then the 16 wave results in order) -> mean and 1/std per channel.
// Shared by crop_kernel and crop_conv1_kernel so that both normalise identically.
__device__ __forceinline__ void crop_stats(double s0, double q0, double s1, double q1, double (*red)[16], float *stat) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    double r4[4] = {s0, q0, s1, q1};
    for (int j = 0; j < 4; j++) {
        double x = r4[j];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) red[j][wid] = x;
    }
    __syncthreads();
    if (tid == 0) {
        double t[4];
        for (int j = 0; j < 4; j++) {
            t[j] = 0.0;
            for (int w = 0; w < kCropThreads / 64; w++) t[j] += red[j][w];
        }
        for (int c = 0; c < 2; c++) {
            double mean = t[2 * c] / 32768.0;
            double var = t[2 * c + 1] / 32768.0 - mean * mean;
            stat[2 * c] = (float)mean;
            stat[2 * c + 1] = var > 1e-12 ? (float)(1.0 / sqrt(var)) : 0.0f;
        }
    }
    __syncthreads();
}

// Raw voxel of crop position i (x fastest) of the box starting at (ox, oy, oz), mirrored
// along the axes of fl, coordinates clamped to the volume.
__device__ __forceinline__ uint32_t crop_voxel(const uint8_t *v, int bpv, int W, int H, int T, int ox, int oy, int oz,
                                               int fl, int i) {
    const int cx = i & 31, cy = (i >> 5) & 31, cz = i >> 10;
    int x = ox + ((fl & 1) ? 31 - cx : cx), y = oy + ((fl & 2) ? 31 - cy : cy), z = oz + ((fl & 4) ? 31 - cz : cz);
    x = x < 0 ? 0 : (x >= W ? W - 1 : x);
    y = y < 0 ? 0 : (y >= H ? H - 1 : y);
    z = z < 0 ? 0 : (z >= T ? T - 1 : z);
    const long long o = ((long long)z * H + y) * W + x;
    return bpv == 1 ? (uint32_t)v[o] : (uint32_t)((const uint16_t *)v)[o];
}

// One workgroup per point: the 32^3 box whose voxel 15 (per axis) holds the point's
// truncated position, from both volumes, each normalised to zero mean and unit
// variance, stored channels-last [32][32][32][2] in BF16.  Voxels outside the volume
// repeat the nearest face voxel.  The raw box is kept in registers between the two
// passes (32 voxels per thread and channel).
__global__ void __launch_bounds__(kCropThreads) crop_kernel(const uint8_t *v0, const uint8_t *v1, int bpv, int W,
                                                            int H, int T, const double *pts, const int *idx, int n,
                                                            uint16_t *out, const int *slot, const int *flip) {
    __shared__ double red[4][16];
    __shared__ float stat[4];
    const int p = blockIdx.x;
    if (p >= n) return;
    const int pt = idx ? idx[p] : p;
    const int ox = (int)floor(pts[3 * (size_t)pt + 0]) - 15;
    const int oy = (int)floor(pts[3 * (size_t)pt + 1]) - 15;
    const int oz = (int)floor(pts[3 * (size_t)pt + 2]) - 15;
    const int tid = threadIdx.x;
    const int fl = flip ? flip[p] : 0;          // bit 0/1/2: mirror the crop along x/y/z
    constexpr int PER = 32768 / kCropThreads;
    double s0 = 0.0, q0 = 0.0, s1 = 0.0, q1 = 0.0;
    for (int b0 = 0; b0 < PER; b0 += kCropBatch) {
        uint32_t ra[kCropBatch], rb[kCropBatch];
#pragma unroll
        for (int k = 0; k < kCropBatch; k++) {
            ra[k] = crop_voxel(v0, bpv, W, H, T, ox, oy, oz, fl, tid + (b0 + k) * kCropThreads);
            rb[k] = crop_voxel(v1, bpv, W, H, T, ox, oy, oz, fl, tid + (b0 + k) * kCropThreads);
        }
#pragma unroll
        for (int k = 0; k < kCropBatch; k++) {
            const float a = (float)ra[k], b = (float)rb[k];
            s0 += a; q0 += (double)a * a;
            s1 += b; q1 += (double)b * b;
        }
    }
    crop_stats(s0, q0, s1, q1, red, stat);
    const float m0 = stat[0], i0 = stat[1], m1 = stat[2], i1 = stat[3];
    uint32_t *o32 = (uint32_t *)(out + (size_t)(slot ? slot[p] : p) * 65536);
    for (int b0 = 0; b0 < PER; b0 += kCropBatch) {
        uint32_t ra[kCropBatch], rb[kCropBatch];
#pragma unroll
        for (int k = 0; k < kCropBatch; k++) {
            ra[k] = crop_voxel(v0, bpv, W, H, T, ox, oy, oz, fl, tid + (b0 + k) * kCropThreads);
            rb[k] = crop_voxel(v1, bpv, W, H, T, ox, oy, oz, fl, tid + (b0 + k) * kCropThreads);
        }
#pragma unroll
        for (int k = 0; k < kCropBatch; k++) {
            const uint32_t a = f2bf(((float)ra[k] - m0) * i0), b = f2bf(((float)rb[k] - m1) * i1);
            o32[tid + (b0 + k) * kCropThreads] = a | (b << 16);
        }
    }
}

// Crop, normalisation and the first convolution in one workgroup per point (inference):
// the point's two 32^3 boxes are read once into LDS (raw voxels, both channels per
// dword), normalised there with the statistics of crop_kernel (same thread mapping and
// summation order, so the BF16 values are identical), and the 2 -> 16 convolution of
// conv_fwd_kernel<1, 2> runs on them from LDS (same lane mapping and k order, so the
// output is identical too).  out: [n][16^3][16] BF16 after bias and ReLU.
__global__ void __launch_bounds__(kCropThreads) crop_conv1_kernel(const uint8_t *v0, const uint8_t *v1, int bpv,
                                                                  int W, int H, int T, const double *pts,
                                                                  const int *idx, int n, const uint16_t *Wb,
                                                                  const float *bias, uint16_t *out) {
    extern __shared__ __align__(16) uint32_t box[];       // 32768 dwords: channel 0 low half, channel 1 high
    __shared__ double red[4][16];
    __shared__ float stat[4];
    const int p = blockIdx.x;
    if (p >= n) return;
    const int pt = idx ? idx[p] : p;
    const int ox = (int)floor(pts[3 * (size_t)pt + 0]) - 15;
    const int oy = (int)floor(pts[3 * (size_t)pt + 1]) - 15;
    const int oz = (int)floor(pts[3 * (size_t)pt + 2]) - 15;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    constexpr int PER = 32768 / kCropThreads;
    double s0 = 0.0, q0 = 0.0, s1 = 0.0, q1 = 0.0;
    for (int b0 = 0; b0 < PER; b0 += kCropBatch) {
        uint32_t ra[kCropBatch], rb[kCropBatch];
#pragma unroll
        for (int k = 0; k < kCropBatch; k++) {
            ra[k] = crop_voxel(v0, bpv, W, H, T, ox, oy, oz, 0, tid + (b0 + k) * kCropThreads);
            rb[k] = crop_voxel(v1, bpv, W, H, T, ox, oy, oz, 0, tid + (b0 + k) * kCropThreads);
        }
#pragma unroll
        for (int k = 0; k < kCropBatch; k++) {
            const float a = (float)ra[k], b = (float)rb[k];
            s0 += a; q0 += (double)a * a;
            s1 += b; q1 += (double)b * b;
            box[tid + (b0 + k) * kCropThreads] = ra[k] | (rb[k] << 16);      // raw, both channels
        }
    }
    crop_stats(s0, q0, s1, q1, red, stat);        // ends with a barrier: box complete
    {
        const float m0 = stat[0], i0 = stat[1], m1 = stat[2], i1 = stat[3];
        for (int k = 0; k < PER; k++) {
            const uint32_t r = box[tid + k * kCropThreads];        // this thread's own entries
            const uint32_t a = f2bf(((float)(r & 0xffffu) - m0) * i0), b = f2bf(((float)(r >> 16) - m1) * i1);
            box[tid + k * kCropThreads] = a | (b << 16);
        }
    }
    __syncthreads();
    // conv 2 -> 16, stride 2, pad 1: 4096 output rows, 16 columns, K = 54 padded to 64
    const int kq = (lane >> 4) * 8;
    uint4 bw[2];
#pragma unroll
    for (int ks = 0; ks < 2; ks++) bw[ks] = *(const uint4 *)(Wb + (lane & 15) * 64 + ks * 32 + kq);
    const float bn = bias[lane & 15];
    uint16_t *o = out + (size_t)p * 4096 * 16;
    for (int tile = wid; tile < 256; tile += kCropThreads / 64) {
        const int m = tile * 16 + (lane & 15);
        const int qz = m >> 8, qy = (m >> 4) & 15, qx = m & 15;
        floatx4 acc = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            uint32_t w4[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int tap = (ks * 32 + kq) / 2 + j;
                const int kz = tap / 9, ky = (tap / 3) % 3, kx = tap % 3;
                const int z = 2 * qz + kz - 1, y = 2 * qy + ky - 1, x = 2 * qx + kx - 1;
                const bool ok = tap < 27 && z >= 0 && z < 32 && y >= 0 && y < 32 && x >= 0 && x < 32;
                w4[j] = ok ? box[(z * 32 + y) * 32 + x] : 0u;
            }
            const uint4 av = uint4{w4[0], w4[1], w4[2], w4[3]};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av),
                                                          __builtin_bit_cast(bf16x8, bw[ks]), acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = tile * 16 + 4 * (lane >> 4) + r;
            const float v = acc[r] + bn;
            o[row * 16 + (lane & 15)] = f2bf(v > 0.0f ? v : 0.0f);
        }
    }
}

// ------------------------------------------------------------------ im2col

// col[m][k], m = ((b*Do + oz)*Do + oy)*Do + ox, k = tap*C + c (tap = (kz*3 + ky)*3 + kx),
// zero for k >= 27*C and for taps outside the input (padding 1, stride 2).  V channels
// per work item (8 when C is a multiple of 8: one 16-byte load and store).
template <int V>
__global__ void im2col_kernel(const uint16_t *in, int B, int D, int C, int Kp, uint16_t *col) {
    const int Do = D >> 1;
    const long long M = (long long)B * Do * Do * Do;
    const int kv = Kp / V;
    const long long total = M * kv;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x; it < total; it += stride) {
        const long long m = it / kv;
        const int k = (int)(it - m * kv) * V;
        uint16_t v[V];
#pragma unroll
        for (int j = 0; j < V; j++) v[j] = 0;
        if (k < 27 * C) {
            const int tap = k / C, c = k - tap * C;
            const int kz = tap / 9, ky = (tap / 3) % 3, kx = tap % 3;
            const int d3 = Do * Do * Do;
            const int b = (int)(m / d3);
            const int r = (int)(m - (long long)b * d3);
            const int oz = r / (Do * Do), oy = (r / Do) % Do, ox = r % Do;
            const int z = 2 * oz + kz - 1, y = 2 * oy + ky - 1, x = 2 * ox + kx - 1;
            if (z >= 0 && z < D && y >= 0 && y < D && x >= 0 && x < D) {
                const uint16_t *src = in + ((((long long)b * D + z) * D + y) * D + x) * C + c;
                if (V == 8) {
                    *(uint4 *)v = *(const uint4 *)src;
                } else {
#pragma unroll
                    for (int j = 0; j < V; j++) v[j] = src[j];
                }
            }
        }
        uint16_t *dst = col + m * Kp + k;
        if (V == 8) {
            *(uint4 *)dst = *(const uint4 *)v;
        } else {
#pragma unroll
            for (int j = 0; j < V; j++) dst[j] = v[j];
        }
    }
}

// ------------------------------------------------------------------ GEMM

// C[m][n] = sum_k A[m*lda + k] * B[n*ldb + k] over k in this workgroup's split
// [z*klen, min(K, (z+1)*klen)), on v_mfma_f32_16x16x32_bf16.  Four waves per workgroup,
// each owning 16 rows and NB 16-column tiles: lane l reads 8 consecutive k of row
// (l & 15) of A and of row (l & 15) of each B tile (one 16-byte load each), the same
// k mapping for both operands, and receives C rows 4*(l >> 4) + r, column l & 15.
// K and klen are multiples of 32, lda and ldb multiples of 8, A and B 16-byte aligned.
// With part != null the raw sums go to part[z][m][n]; otherwise bias, ReLU and the
// FP32 / BF16 outputs are applied.
struct GemmArgs {
    const uint16_t *A; int lda;
    const uint16_t *Bm; int ldb;
    int M, N, K, klen;
    const float *bias; int relu;
    float *outf; int ldf;
    uint16_t *outb; int ldb16;
    float *part;
};

template <int NB>
__global__ void __launch_bounds__(256) gemm_nt_kernel(GemmArgs g) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = blockIdx.x * 64 + wave * 16;
    const int col0 = blockIdx.y * (16 * NB);
    const int kbeg = blockIdx.z * g.klen;
    const int kend = min(g.K, kbeg + g.klen);
    const int kq = (lane >> 4) * 8;
    const int arow = row0 + (lane & 15);
    floatx4 acc[NB];
#pragma unroll
    for (int t = 0; t < NB; t++) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint4 zero = {0u, 0u, 0u, 0u};
    const uint16_t *ap = g.A + (long long)arow * g.lda + kq;
    const uint16_t *bp[NB];
    bool bok[NB];
#pragma unroll
    for (int t = 0; t < NB; t++) {
        const int ncol = col0 + t * 16 + (lane & 15);
        bok[t] = ncol < g.N;
        bp[t] = g.Bm + (long long)(bok[t] ? ncol : 0) * g.ldb + kq;
    }
    const bool aok = arow < g.M;
#pragma unroll 4
    for (int k = kbeg; k < kend; k += 32) {
        const uint4 av = aok ? *(const uint4 *)(ap + k) : zero;
        const bf16x8 a8 = __builtin_bit_cast(bf16x8, av);
#pragma unroll
        for (int t = 0; t < NB; t++) {
            const uint4 bv = bok[t] ? *(const uint4 *)(bp[t] + k) : zero;
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a8, __builtin_bit_cast(bf16x8, bv), acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NB; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = row0 + 4 * (lane >> 4) + r;
            const int n = col0 + t * 16 + (lane & 15);
            if (m >= g.M || n >= g.N) continue;
            float v = acc[t][r];
            if (g.part) {
                g.part[((long long)blockIdx.z * g.M + m) * g.N + n] = v;
                continue;
            }
            if (g.bias) v += g.bias[n];
            if (g.relu) v = v > 0.0f ? v : 0.0f;
            if (g.outf) g.outf[(long long)m * g.ldf + n] = v;
            if (g.outb) g.outb[(long long)m * g.ldb16 + n] = f2bf(v);
        }
}

// The forward convolution as an implicit GEMM: the same tiles and lane mapping as
// gemm_nt_kernel, but each lane gathers its 8 patch values of A straight from the
// channels-last input [B][D][D][D][C] (stride 2, padding 1), so the patch matrix is
// never stored.  C a multiple of 8: one 16-byte load per lane and k-step (8 channels
// of one tap); C = 2: four 4-byte loads (2 channels of 4 taps).  Taps past 27 and
// voxels outside the input read as zero.  Each wave computes RT consecutive 16-row
// tiles, so a k-step's weight fragments serve RT products and RT gathers are in flight
// together; every output still sums its k-steps in the same order.  Epilogue: bias,
// ReLU, BF16.
template <int NB, int C, int RT = 1>
__global__ void __launch_bounds__(256) conv_fwd_kernel(const uint16_t *in, int B, int D, const uint16_t *Wb, int Kp,
                                                       int N, const float *bias, uint16_t *out) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Do = D >> 1;
    const int M = B * Do * Do * Do;
    const int row0 = blockIdx.x * 64 * RT + wave * 16 * RT;
    const int col0 = blockIdx.y * (16 * NB);
    const int kq = (lane >> 4) * 8;
    bool aok[RT];
    int oz[RT], oy[RT], ox[RT];
    const uint16_t *base[RT];
#pragma unroll
    for (int rt = 0; rt < RT; rt++) {
        const int arow = row0 + rt * 16 + (lane & 15);
        aok[rt] = arow < M;
        int b = 0;
        oz[rt] = oy[rt] = ox[rt] = 0;
        if (aok[rt]) {
            const int d3 = Do * Do * Do;
            b = arow / d3;
            const int r = arow - b * d3;
            oz[rt] = r / (Do * Do); oy[rt] = (r / Do) % Do; ox[rt] = r % Do;
        }
        base[rt] = in + (long long)b * D * D * D * C;
    }
    floatx4 acc[RT][NB];
#pragma unroll
    for (int rt = 0; rt < RT; rt++)
#pragma unroll
        for (int t = 0; t < NB; t++) acc[rt][t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint4 zero = {0u, 0u, 0u, 0u};
    const uint16_t *bp[NB];
    bool bok[NB];
#pragma unroll
    for (int t = 0; t < NB; t++) {
        const int ncol = col0 + t * 16 + (lane & 15);
        bok[t] = ncol < N;
        bp[t] = Wb + (long long)(bok[t] ? ncol : 0) * Kp + kq;
    }
    auto voxel_ptr = [&](int rt, int tap, bool &ok) -> const uint16_t * {
        const int kz = tap / 9, ky = (tap / 3) % 3, kx = tap % 3;
        const int z = 2 * oz[rt] + kz - 1, y = 2 * oy[rt] + ky - 1, x = 2 * ox[rt] + kx - 1;
        ok = aok[rt] && tap < 27 && z >= 0 && z < D && y >= 0 && y < D && x >= 0 && x < D;
        return base[rt] + (((long long)z * D + y) * D + x) * C;
    };
#pragma unroll 4
    for (int k = 0; k < Kp; k += 32) {
        uint4 bv[NB];
#pragma unroll
        for (int t = 0; t < NB; t++) bv[t] = bok[t] ? *(const uint4 *)(bp[t] + k) : zero;
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
            uint4 av;
            if (C % 8 == 0) {
                const int kk = k + kq;
                bool ok;
                const uint16_t *ptr = voxel_ptr(rt, kk / C, ok);
                av = ok ? *(const uint4 *)(ptr + (kk % C)) : zero;
            } else {
                uint32_t w[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    bool ok;
                    const uint16_t *ptr = voxel_ptr(rt, (k + kq) / 2 + j, ok);
                    w[j] = ok ? *(const uint32_t *)ptr : 0u;
                }
                av = uint4{w[0], w[1], w[2], w[3]};
            }
            const bf16x8 a8 = __builtin_bit_cast(bf16x8, av);
#pragma unroll
            for (int t = 0; t < NB; t++)
                acc[rt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a8, __builtin_bit_cast(bf16x8, bv[t]), acc[rt][t],
                                                                     0, 0, 0);
        }
    }
#pragma unroll
    for (int rt = 0; rt < RT; rt++)
#pragma unroll
        for (int t = 0; t < NB; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = row0 + rt * 16 + 4 * (lane >> 4) + r;
                const int n = col0 + t * 16 + (lane & 15);
                if (m >= M || n >= N) continue;
                float v = acc[rt][t][r] + bias[n];
                out[(long long)m * N + n] = f2bf(v > 0.0f ? v : 0.0f);
            }
}

// dst[o + m*ldd + n] = sum_z part[z][m][n] for n < ncols (drops padding columns)
__global__ void reduce_part_kernel(const float *part, int splits, int M, int N, int ncols, float *dst, int ldd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * ncols) return;
    const int m = i / ncols, n = i - m * ncols;
    float s = 0.0f;
    for (int z = 0; z < splits; z++) s += part[((long long)z * M + m) * N + n];
    dst[(long long)m * ldd + n] = s;
}

// out[c][r] = in[r][c] for r < R, c < Cc (BF16), 32 x 32 tiles through LDS
__global__ void transpose_kernel(const uint16_t *in, int R, int Cc, int ldi, uint16_t *out, int ldo) {
    __shared__ uint16_t tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;     // 32 x 8
    for (int j = ty; j < 32; j += 8) {
        const int r = r0 + j, c = c0 + tx;
        tile[j][tx] = (r < R && c < Cc) ? in[(long long)r * ldi + c] : (uint16_t)0;
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int c = c0 + j, r = r0 + tx;
        if (c < Cc && r < R) out[(long long)c * ldo + r] = tile[tx][j];
    }
}

// ------------------------------------------------------------------ backward pieces

// dZ = dY where the forward output was positive (ReLU), as BF16
__global__ void relu_grad_kernel(const float *dY, const uint16_t *act, long long n, uint16_t *dZ) {
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st)
        dZ[i] = bf2f(act[i]) > 0.0f ? f2bf(dY[i]) : (uint16_t)0;
}

// part[blockIdx.x][n] = sum over this block's rows of dY[m][n] * (act[m][n] > 0);
// N divides 256
__global__ void __launch_bounds__(256) bias_grad_kernel(const float *dY, const uint16_t *act, int M, int N,
                                                        int rows_per_block, float *part) {
    __shared__ float red[256];
    const int tid = threadIdx.x;
    const int n = tid % N, r = tid / N, rstep = 256 / N;
    const int m0 = blockIdx.x * rows_per_block;
    const int m1 = min(M, m0 + rows_per_block);
    float s = 0.0f;
    for (int m = m0 + r; m < m1; m += rstep) {
        const long long i = (long long)m * N + n;
        if (bf2f(act[i]) > 0.0f) s += dY[i];
    }
    red[tid] = s;
    __syncthreads();
    if (tid < N) {
        float t = 0.0f;
        for (int j = tid; j < 256; j += N) t += red[j];
        part[(long long)blockIdx.x * N + tid] = t;
    }
}

// dA[b][z][y][x][c] = sum of dcol[m][tap*C + c] over the (output, tap) pairs that read
// input voxel (z, y, x) in the forward im2col (stride 2, pad 1)
__global__ void col2im_kernel(const float *dcol, int B, int D, int C, int Kp, float *dA) {
    const int Do = D >> 1;
    const long long total = (long long)B * D * D * D * C;
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += st) {
        const int c = (int)(i % C);
        long long v = i / C;
        const int x = (int)(v % D); v /= D;
        const int y = (int)(v % D); v /= D;
        const int z = (int)(v % D);
        const int b = (int)(v / D);
        float s = 0.0f;
        for (int kz = 0; kz < 3; kz++) {
            const int tz = z + 1 - kz;
            if (tz < 0 || (tz & 1) || (tz >> 1) >= Do) continue;
            for (int ky = 0; ky < 3; ky++) {
                const int ty = y + 1 - ky;
                if (ty < 0 || (ty & 1) || (ty >> 1) >= Do) continue;
                for (int kx = 0; kx < 3; kx++) {
                    const int tx = x + 1 - kx;
                    if (tx < 0 || (tx & 1) || (tx >> 1) >= Do) continue;
                    const long long m = (((long long)b * Do + (tz >> 1)) * Do + (ty >> 1)) * Do + (tx >> 1);
                    s += dcol[m * Kp + ((kz * 3 + ky) * 3 + kx) * C + c];
                }
            }
        }
        dA[i] = s;
    }
}

// ------------------------------------------------------------------ head

constexpr int kFeat = 128, kOut = 12;

// The per-parameter scale of the network output: voxels, radians, strain.
__host__ __device__ __forceinline__ float out_scale(int j) { return j < 3 ? 8.0f : (j < 6 ? 0.05f : 0.02f); }

// One workgroup (128 threads) per sample: feat = mean of the last activation over its
// 8 voxels, y = Wfc feat + bfc.  par (optional) receives y * scale in FP64 at
// par[idx[b]*12 + j] (idx null: b).
__global__ void __launch_bounds__(128) head_forward_kernel(const uint16_t *act4, const float *Wfc, const float *bfc,
                                                           int B, float *feat, float *y, double *par, const int *idx) {
    __shared__ float f[kFeat];
    const int b = blockIdx.x, c = threadIdx.x;
    if (b >= B) return;
    float s = 0.0f;
    for (int v = 0; v < 8; v++) s += bf2f(act4[((long long)b * 8 + v) * kFeat + c]);
    s *= 0.125f;
    f[c] = s;
    if (feat) feat[(long long)b * kFeat + c] = s;
    __syncthreads();
    if (c < kOut) {
        float o = bfc[c];
        for (int q = 0; q < kFeat; q++) o += Wfc[c * kFeat + q] * f[q];
        if (y) y[b * kOut + c] = o;
        if (par) {
            const long long row = idx ? idx[b] : b;
            par[row * 12 + c] = (double)(o * out_scale(c));
        }
    }
}

// Mean-squared loss of y against label/scale, and its gradient back to the last
// activation: dy = 2 (y - label/scale) / B; dA4[b][v][c] = (Wfc^T dy)[c] / 8.
// loss_part[b] = sum_j (y - label/scale)^2 / B.
__global__ void __launch_bounds__(128) head_backward_kernel(const float *y, const float *label, const float *Wfc, int B,
                                                            float *dy, float *dA4, float *loss_part) {
    __shared__ float d[kOut];
    const int b = blockIdx.x, c = threadIdx.x;
    if (b >= B) return;
    if (c < kOut) {
        const float e = y[b * kOut + c] - label[b * kOut + c] / out_scale(c);
        d[c] = 2.0f * e / (float)B;
        dy[b * kOut + c] = d[c];
    }
    __syncthreads();
    if (c == 0) {
        float l = 0.0f;
        for (int j = 0; j < kOut; j++) {
            const float e = y[b * kOut + j] - label[b * kOut + j] / out_scale(j);
            l += e * e;
        }
        loss_part[b] = l / (float)B;
    }
    float s = 0.0f;
    for (int j = 0; j < kOut; j++) s += Wfc[j * kFeat + c] * d[j];
    s *= 0.125f;
    for (int v = 0; v < 8; v++) dA4[((long long)b * 8 + v) * kFeat + c] = s;
}

// gW[j][c] = sum_b dy[b][j] feat[b][c]; gb[j] = sum_b dy[b][j]; loss = sum_b loss_part[b]
__global__ void __launch_bounds__(128) fc_grad_kernel(const float *dy, const float *feat, const float *loss_part, int B,
                                                      float *gW, float *gb, float *loss) {
    const int j = blockIdx.x, c = threadIdx.x;
    float s = 0.0f;
    for (int b = 0; b < B; b++) s += dy[b * kOut + j] * feat[(long long)b * kFeat + c];
    gW[j * kFeat + c] = s;
    if (c == 0) {
        float t = 0.0f;
        for (int b = 0; b < B; b++) t += dy[b * kOut + j];
        gb[j] = t;
        if (j == 0 && loss) {
            float l = 0.0f;
            for (int b = 0; b < B; b++) l += loss_part[b];
            *loss = l;
        }
    }
}

// ------------------------------------------------------------------ optimiser

__global__ void adam_kernel(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                            float c1, float c2, float eps) {
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// W [Co][K] FP32 -> Wb [Co][Kp] BF16 (zero padded) and WbT [Kp][Co]
__global__ void pack_kernel(const float *W, int Co, int K, int Kp, uint16_t *Wb, uint16_t *WbT) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Co * Kp) return;
    const int co = i / Kp, k = i - co * Kp;
    const uint16_t h = k < K ? f2bf(W[(long long)co * K + k]) : (uint16_t)0;
    Wb[i] = h;
    WbT[(long long)k * Co + co] = h;
}

// ------------------------------------------------------------------ training pairs

// Target volume of a training pair: every tile x tile x tile block t of the volume is
// the reference resampled under that tile's own affine map, out(x) = ref(M_t x + o_t)
// (maps[t] = M row-major, then o), by separable Catmull-Rom (the interior form of the
// path's tricubic interpolant) with indices clamped to the volume, rounded to the
// nearest integer and clamped to the voxel type's range.  Tiles are numbered x fastest.
template <int BPV>
__global__ void warp_tiles_kernel(const uint8_t *src, uint8_t *dst, int W, int H, int T, int tile, int ntx, int nty,
                                  const double *maps) {
    const long long n = (long long)W * H * T;
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
        const int x = (int)(i % W), y = (int)((i / W) % H), z = (int)(i / ((long long)W * H));
        const int t = (z / tile * nty + y / tile) * ntx + x / tile;
        const double *m = maps + (long long)t * 12;
        const double X = m[0] * x + m[1] * y + m[2] * z + m[9];
        const double Y = m[3] * x + m[4] * y + m[5] * z + m[10];
        const double Z = m[6] * x + m[7] * y + m[8] * z + m[11];
        const double fx = floor(X), fy = floor(Y), fz = floor(Z);
        const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
        double w[3][4];
        const double tt[3] = {X - fx, Y - fy, Z - fz};
        for (int a = 0; a < 3; a++) {
            const double u = tt[a], u2 = u * u, u3 = u2 * u;
            w[a][0] = 0.5 * (-u3 + 2 * u2 - u);
            w[a][1] = 0.5 * (3 * u3 - 5 * u2 + 2);
            w[a][2] = 0.5 * (-3 * u3 + 4 * u2 + u);
            w[a][3] = 0.5 * (u3 - u2);
        }
        double acc = 0.0;
        for (int c = 0; c < 4; c++) {
            const int zz = min(max(iz - 1 + c, 0), T - 1);
            for (int b = 0; b < 4; b++) {
                const int yy = min(max(iy - 1 + b, 0), H - 1);
                const long long row = ((long long)zz * H + yy) * W;
                double r = 0.0;
                for (int a = 0; a < 4; a++) {
                    const int xx = min(max(ix - 1 + a, 0), W - 1);
                    const double v = BPV == 1 ? (double)src[row + xx] : (double)((const uint16_t *)src)[row + xx];
                    r += w[0][a] * v;
                }
                acc += w[2][c] * w[1][b] * r;
            }
        }
        const double top = BPV == 1 ? 255.0 : 65535.0;
        double v = rint(acc);
        v = v < 0.0 ? 0.0 : (v > top ? top : v);
        if (BPV == 1) dst[i] = (uint8_t)v;
        else ((uint16_t *)dst)[i] = (uint16_t)v;
    }
}

// ------------------------------------------------------------------ layout

constexpr int kLayers = 4;
constexpr int kCin[kLayers] = {2, 16, 32, 64};
constexpr int kCout[kLayers] = {16, 32, 64, 128};
constexpr int kDin[kLayers] = {32, 16, 8, 4};
constexpr int kK(int l) { return 27 * kCin[l]; }
constexpr int kKp(int l) { return (kK(l) + 31) / 32 * 32; }

struct Layout {
    long long w[kLayers], b[kLayers], fcw, fcb, total;
    Layout() {
        long long o = 0;
        for (int l = 0; l < kLayers; l++) {
            w[l] = o; o += (long long)kCout[l] * kK(l);
            b[l] = o; o += kCout[l];
        }
        fcw = o; o += kOut * kFeat;
        fcb = o; o += kOut;
        total = o;
    }
};

}  // namespace net
}  // namespace dvc

using namespace dvc;
using namespace dvc::net;

struct dvc_net {
    dvc_ctx *ctx = nullptr;
    int cap = 0;                       // batch capacity
    int training = 0;
    Layout L;
    long long step = 0;
    std::string err;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;      // FP32 params, grads, Adam moments
    uint16_t *wb[kLayers] = {}, *wbT[kLayers] = {};                     // packed BF16 weights
    uint16_t *in = nullptr;                                             // crops [cap][32^3][2]
    uint16_t *act[kLayers + 1] = {};                                    // act[0] aliases the input
    uint16_t *col[kLayers] = {};
    float *feat = nullptr, *y = nullptr, *dy = nullptr, *loss_part = nullptr, *loss = nullptr, *label = nullptr;
    // training workspace
    float *dA = nullptr, *dA2 = nullptr;   // activation gradients (FP32), ping-pong
    uint16_t *dZ = nullptr, *dZT = nullptr, *colT = nullptr;
    float *dcol = nullptr, *part = nullptr, *bpart = nullptr;
    long long part_cap = 0, bpart_cap = 0;
    std::vector<void *> allocs;
};

namespace {

long long layer_rows(int l, int B) {
    const long long d = kDin[l] / 2;
    return (long long)B * d * d * d;
}

int net_fail(dvc_net *n, int code, const std::string &msg) {
    if (n) n->err = msg;
    return code;
}

template <typename T>
int net_alloc(dvc_net *n, T **ptr, size_t count) {
    if (count == 0) count = 1;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) return net_fail(n, DVC_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    n->allocs.push_back(q);
    *ptr = (T *)q;
    return 0;
}

int launch_ok(dvc_net *n, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return net_fail(n, DVC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return 0;
}

unsigned grid_for(long long items, int block, long long cap = 1 << 16) {
    long long b = (items + block - 1) / block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

int gemm(dvc_net *n, hipStream_t st, GemmArgs a, int splits) {
    if (a.K % 32 || a.klen % 32 || a.lda % 8 || a.ldb % 8 || a.N % 16 ||
        ((uintptr_t)a.A & 15) || ((uintptr_t)a.Bm & 15))
        return net_fail(n, DVC_ERR_ARG, "gemm operand shape or alignment");
    const int nb = a.N >= 64 ? 4 : (a.N >= 32 ? 2 : 1);
    dim3 grid((unsigned)((a.M + 63) / 64), (unsigned)((a.N + 16 * nb - 1) / (16 * nb)), (unsigned)splits);
    if (nb == 4) hipLaunchKernelGGL(gemm_nt_kernel<4>, grid, dim3(256), 0, st, a);
    else if (nb == 2) hipLaunchKernelGGL(gemm_nt_kernel<2>, grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(gemm_nt_kernel<1>, grid, dim3(256), 0, st, a);
    return launch_ok(n, "gemm");
}

int im2col(dvc_net *n, hipStream_t st, const uint16_t *in, int B, int l, uint16_t *col) {
    const int C = kCin[l], D = kDin[l], Kp = kKp(l);
    const long long items = layer_rows(l, B) * Kp;
    if (C % 8 == 0)
        hipLaunchKernelGGL(im2col_kernel<8>, dim3(grid_for(items / 8, 256)), dim3(256), 0, st, in, B, D, C, Kp, col);
    else
        hipLaunchKernelGGL(im2col_kernel<1>, dim3(grid_for(items, 256)), dim3(256), 0, st, in, B, D, C, Kp, col);
    return launch_ok(n, "im2col");
}

int transpose(dvc_net *n, hipStream_t st, const uint16_t *in, long long R, int Cc, uint16_t *out) {
    dim3 grid((unsigned)((Cc + 31) / 32), (unsigned)((R + 31) / 32));
    if (grid.y > 2147483647u) return net_fail(n, DVC_ERR_ARG, "transpose too large");
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, st, in, (int)R, Cc, Cc, out, (int)R);
    return launch_ok(n, "transpose");
}

int conv_fwd(dvc_net *n, hipStream_t st, int l, int B) {
    const int D = kDin[l], N = kCout[l], Kp = kKp(l);
    const long long M = layer_rows(l, B);
    if (M > 2147483647LL) return net_fail(n, DVC_ERR_ARG, "batch too large");
    const int nb = N >= 64 ? 4 : (N >= 32 ? 2 : 1);
    const int rt = l == 1 ? 4 : (l == 2 ? 2 : 1);   // row tiles per wave (8^3 and 4^3 outputs per crop)
    dim3 grid((unsigned)((M + 64 * rt - 1) / (64 * rt)), (unsigned)((N + 16 * nb - 1) / (16 * nb)));
    const uint16_t *in = n->act[l];
    const float *bias = n->p + n->L.b[l];
    uint16_t *out = n->act[l + 1];
    switch (l) {
        case 0: hipLaunchKernelGGL((conv_fwd_kernel<1, 2>), grid, dim3(256), 0, st, in, B, D, n->wb[l], Kp, N, bias, out); break;
        case 1: hipLaunchKernelGGL((conv_fwd_kernel<2, 16, 4>), grid, dim3(256), 0, st, in, B, D, n->wb[l], Kp, N, bias, out); break;
        case 2: hipLaunchKernelGGL((conv_fwd_kernel<4, 32, 2>), grid, dim3(256), 0, st, in, B, D, n->wb[l], Kp, N, bias, out); break;
        default: hipLaunchKernelGGL((conv_fwd_kernel<4, 64>), grid, dim3(256), 0, st, in, B, D, n->wb[l], Kp, N, bias, out); break;
    }
    return launch_ok(n, "conv forward");
}

int forward(dvc_net *n, hipStream_t st, int B, double *par, const int *idx) {
    n->act[0] = n->in;
    for (int l = 0; l < kLayers; l++) {
        if (!n->training) {
            // inference: implicit GEMM, no patch matrix
            if (int e = conv_fwd(n, st, l, B)) return e;
            continue;
        }
        if (int e = im2col(n, st, n->act[l], B, l, n->col[l])) return e;
        GemmArgs a{};
        a.A = n->col[l]; a.lda = kKp(l);
        a.Bm = n->wb[l]; a.ldb = kKp(l);
        a.M = (int)layer_rows(l, B); a.N = kCout[l]; a.K = kKp(l); a.klen = kKp(l);
        a.bias = n->p + n->L.b[l]; a.relu = 1;
        a.outb = n->act[l + 1]; a.ldb16 = kCout[l];
        if (int e = gemm(n, st, a, 1)) return e;
    }
    hipLaunchKernelGGL(head_forward_kernel, dim3(B), dim3(128), 0, st, n->act[kLayers], n->p + n->L.fcw,
                       n->p + n->L.fcb, B, n->feat, n->y, par, idx);
    return launch_ok(n, "head");
}

// inference from points: fused crop + first convolution, then layers 2-4 and the head;
// par[idx[b]*12..] receives the proposals
int propose_fused(dvc_net *n, hipStream_t st, const double *pts, const int *idx, int B, double *par) {
    const uint8_t *v0, *v1;
    int bpv, W, H, T;
    if (ctx_volumes(n->ctx, &v0, &v1, &bpv, &W, &H, &T) != 0) return net_fail(n, DVC_ERR_STATE, "no volume pair");
    hipLaunchKernelGGL(crop_conv1_kernel, dim3(B), dim3(kCropThreads), 32768 * 4, st, v0, v1, bpv, W, H, T, pts, idx,
                       B, n->wb[0], n->p + n->L.b[0], n->act[1]);
    if (int e = launch_ok(n, "crop + first convolution")) return e;
    for (int l = 1; l < kLayers; l++)
        if (int e = conv_fwd(n, st, l, B)) return e;
    hipLaunchKernelGGL(head_forward_kernel, dim3(B), dim3(128), 0, st, n->act[kLayers], n->p + n->L.fcw,
                       n->p + n->L.fcb, B, n->feat, n->y, par, idx);
    return launch_ok(n, "head");
}

int pack(dvc_net *n, hipStream_t st) {
    for (int l = 0; l < kLayers; l++) {
        const int items = kCout[l] * kKp(l);
        hipLaunchKernelGGL(pack_kernel, dim3(grid_for(items, 256)), dim3(256), 0, st, n->p + n->L.w[l], kCout[l], kK(l),
                           kKp(l), n->wb[l], n->wbT[l]);
        if (int e = launch_ok(n, "pack")) return e;
    }
    return 0;
}

// gradients of the loss over a batch already run forward (act/col/feat/y filled)
int backward(dvc_net *n, hipStream_t st, int B) {
    hipLaunchKernelGGL(head_backward_kernel, dim3(B), dim3(128), 0, st, n->y, n->label, n->p + n->L.fcw, B, n->dy,
                       n->dA, n->loss_part);
    if (int e = launch_ok(n, "head backward")) return e;
    hipLaunchKernelGGL(fc_grad_kernel, dim3(kOut), dim3(128), 0, st, n->dy, n->feat, n->loss_part, B, n->g + n->L.fcw,
                       n->g + n->L.fcb, n->loss);
    if (int e = launch_ok(n, "fc grad")) return e;
    float *dY = n->dA, *dNext = n->dA2;
    for (int l = kLayers - 1; l >= 0; l--) {
        const long long M = layer_rows(l, B);
        const int Co = kCout[l], Kp = kKp(l), K = kK(l);
        const uint16_t *act = n->act[l + 1];
        hipLaunchKernelGGL(relu_grad_kernel, dim3(grid_for(M * Co, 256)), dim3(256), 0, st, dY, act, M * Co, n->dZ);
        if (int e = launch_ok(n, "relu grad")) return e;
        // bias gradient: column sums of the masked gradient
        const int rows = 1024;
        const int nblk = (int)((M + rows - 1) / rows);
        if ((long long)nblk * Co > n->bpart_cap) return net_fail(n, DVC_ERR_ARG, "bias workspace");
        hipLaunchKernelGGL(bias_grad_kernel, dim3(nblk), dim3(256), 0, st, dY, act, (int)M, Co, rows, n->bpart);
        if (int e = launch_ok(n, "bias grad")) return e;
        hipLaunchKernelGGL(reduce_part_kernel, dim3(grid_for(Co, 256)), dim3(256), 0, st, n->bpart, nblk, 1, Co, Co,
                           n->g + n->L.b[l], Co);
        if (int e = launch_ok(n, "bias reduce")) return e;
        // weight gradient: dZ^T (Co x M) times col (M x Kp), split over M
        if (int e = transpose(n, st, n->dZ, M, Co, n->dZT)) return e;
        if (int e = transpose(n, st, n->col[l], M, Kp, n->colT)) return e;
        const int klen = 4096;
        const int splits = (int)((M + klen - 1) / klen);
        if ((long long)splits * Co * Kp > n->part_cap) return net_fail(n, DVC_ERR_ARG, "split workspace");
        GemmArgs w{};
        w.A = n->dZT; w.lda = (int)M;
        w.Bm = n->colT; w.ldb = (int)M;
        w.M = Co; w.N = Kp; w.K = (int)M; w.klen = klen;
        w.part = n->part;
        if (int e = gemm(n, st, w, splits)) return e;
        hipLaunchKernelGGL(reduce_part_kernel, dim3(grid_for((long long)Co * K, 256)), dim3(256), 0, st, n->part, splits,
                           Co, Kp, K, n->g + n->L.w[l], K);
        if (int e = launch_ok(n, "weight reduce")) return e;
        if (l == 0) break;
        // input gradient: dZ (M x Co) times W (Co x Kp), then col2im
        GemmArgs d{};
        d.A = n->dZ; d.lda = Co;
        d.Bm = n->wbT[l]; d.ldb = Co;
        d.M = (int)M; d.N = Kp; d.K = Co; d.klen = Co;
        d.outf = n->dcol; d.ldf = Kp;
        if (int e = gemm(n, st, d, 1)) return e;
        const long long nin = (long long)B * kDin[l] * kDin[l] * kDin[l] * kCin[l];
        hipLaunchKernelGGL(col2im_kernel, dim3(grid_for(nin, 256)), dim3(256), 0, st, n->dcol, B, kDin[l], kCin[l], Kp,
                           dNext);
        if (int e = launch_ok(n, "col2im")) return e;
        float *t = dY; dY = dNext; dNext = t;
    }
    return 0;
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

int dvc_net_create(dvc_ctx *ctx, int max_batch, int training, dvc_net **out) {
    if (!ctx || !out || max_batch < 1 || max_batch > 4096) return DVC_ERR_ARG;
    if (training && (max_batch % 4)) return DVC_ERR_ARG;
    *out = nullptr;
    dvc_net *n = new dvc_net();
    n->ctx = ctx;
    n->cap = max_batch;
    n->training = training ? 1 : 0;
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) { delete n; return DVC_ERR_HIP; }
    if (hipFuncSetAttribute((const void *)crop_conv1_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 32768 * 4) !=
        hipSuccess) {
        delete n;
        return DVC_ERR_HIP;
    }
    const int B = max_batch;
    int e = 0;
    e = e ? e : net_alloc(n, &n->p, (size_t)n->L.total);
    e = e ? e : net_alloc(n, &n->in, (size_t)B * 65536);
    for (int l = 0; l < kLayers && !e; l++) {
        e = net_alloc(n, &n->wb[l], (size_t)kCout[l] * kKp(l));
        e = e ? e : net_alloc(n, &n->wbT[l], (size_t)kCout[l] * kKp(l));
        if (training) e = e ? e : net_alloc(n, &n->col[l], (size_t)layer_rows(l, B) * kKp(l));
        e = e ? e : net_alloc(n, &n->act[l + 1], (size_t)layer_rows(l, B) * kCout[l]);
    }
    e = e ? e : net_alloc(n, &n->feat, (size_t)B * kFeat);
    e = e ? e : net_alloc(n, &n->y, (size_t)B * kOut);
    if (!e && training) {
        long long maxMC = 0, maxMKp = 0, maxIn = 0, maxPart = 0, maxB = 0;
        for (int l = 0; l < kLayers; l++) {
            const long long M = layer_rows(l, B);
            maxMC = std::max(maxMC, M * kCout[l]);
            maxMKp = std::max(maxMKp, M * kKp(l));
            maxIn = std::max(maxIn, (long long)B * kDin[l] * kDin[l] * kDin[l] * kCin[l]);
            maxPart = std::max(maxPart, (M + 4095) / 4096 * kCout[l] * kKp(l));
            maxB = std::max(maxB, (M + 1023) / 1024 * kCout[l]);
        }
        n->part_cap = maxPart;
        n->bpart_cap = maxB;
        e = e ? e : net_alloc(n, &n->g, (size_t)n->L.total);
        e = e ? e : net_alloc(n, &n->m, (size_t)n->L.total);
        e = e ? e : net_alloc(n, &n->v, (size_t)n->L.total);
        e = e ? e : net_alloc(n, &n->dA, (size_t)std::max(maxMC, maxIn));
        e = e ? e : net_alloc(n, &n->dA2, (size_t)std::max(maxMC, maxIn));
        e = e ? e : net_alloc(n, &n->dZ, (size_t)maxMC);
        e = e ? e : net_alloc(n, &n->dZT, (size_t)maxMC);
        e = e ? e : net_alloc(n, &n->colT, (size_t)maxMKp);
        e = e ? e : net_alloc(n, &n->dcol, (size_t)maxMKp);
        e = e ? e : net_alloc(n, &n->part, (size_t)maxPart);
        e = e ? e : net_alloc(n, &n->bpart, (size_t)maxB);
        e = e ? e : net_alloc(n, &n->dy, (size_t)B * kOut);
        e = e ? e : net_alloc(n, &n->loss_part, (size_t)B);
        e = e ? e : net_alloc(n, &n->loss, 1);
        e = e ? e : net_alloc(n, &n->label, (size_t)B * kOut);
        if (!e) {
            hipMemset(n->m, 0, sizeof(float) * n->L.total);
            hipMemset(n->v, 0, sizeof(float) * n->L.total);
            hipMemset(n->g, 0, sizeof(float) * n->L.total);
        }
    }
    if (!e) e = hipMemset(n->p, 0, sizeof(float) * n->L.total) == hipSuccess ? 0 : DVC_ERR_HIP;
    if (e) {
        for (void *q : n->allocs) hipFree(q);
        delete n;
        return e;
    }
    *out = n;
    return DVC_OK;
}

void dvc_net_destroy(dvc_net *n) {
    if (!n) return;
    hipSetDevice(ctx_device(n->ctx));
    hipStreamSynchronize((hipStream_t)ctx_stream(n->ctx));
    for (void *q : n->allocs) hipFree(q);
    delete n;
}

const char *dvc_net_last_error(const dvc_net *n) { return n ? n->err.c_str() : "null network"; }

int64_t dvc_net_num_params(void) { return Layout().total; }

int dvc_net_params(dvc_net *n, float *host, int64_t count, int set) {
    if (!n || !host) return DVC_ERR_ARG;
    if (count != n->L.total) return net_fail(n, DVC_ERR_ARG, "parameter count mismatch");
    hipStream_t st = (hipStream_t)ctx_stream(n->ctx);
    if (hipSetDevice(ctx_device(n->ctx)) != hipSuccess) return DVC_ERR_HIP;
    if (set) {
        if (hipMemcpyAsync(n->p, host, sizeof(float) * count, hipMemcpyHostToDevice, st) != hipSuccess)
            return net_fail(n, DVC_ERR_HIP, "parameter upload");
        if (int e = pack(n, st)) return e;
        if (n->training) {
            hipMemsetAsync(n->m, 0, sizeof(float) * count, st);
            hipMemsetAsync(n->v, 0, sizeof(float) * count, st);
            n->step = 0;
        }
    } else if (hipMemcpyAsync(host, n->p, sizeof(float) * count, hipMemcpyDeviceToHost, st) != hipSuccess) {
        return net_fail(n, DVC_ERR_HIP, "parameter download");
    }
    return hipStreamSynchronize(st) == hipSuccess ? DVC_OK : net_fail(n, DVC_ERR_HIP, "synchronise");
}

int dvc_net_crop(dvc_net *n, const double *points_dev, const int32_t *idx_dev, int32_t count, void *out_dev,
                 const int32_t *slot_dev, const int32_t *flip_dev) {
    if (!n || !points_dev || count < 0) return DVC_ERR_ARG;
    if (count == 0) return DVC_OK;
    const uint8_t *v0, *v1;
    int bpv, W, H, T;
    if (ctx_volumes(n->ctx, &v0, &v1, &bpv, &W, &H, &T) != 0) return net_fail(n, DVC_ERR_STATE, "no volume pair");
    hipStream_t st = (hipStream_t)ctx_stream(n->ctx);
    uint16_t *out = out_dev ? (uint16_t *)out_dev : n->in;
    if (!out_dev && (count > n->cap || slot_dev)) return net_fail(n, DVC_ERR_ARG, "more crops than the batch capacity");
    hipLaunchKernelGGL(crop_kernel, dim3(count), dim3(kCropThreads), 0, st, v0, v1, bpv, W, H, T, points_dev, idx_dev,
                       count, out, slot_dev, flip_dev);
    return launch_ok(n, "crop");
}

static int stage_input(dvc_net *n, const void *in_dev, int B, hipStream_t st) {
    if (in_dev && in_dev != n->in &&
        hipMemcpyAsync(n->in, in_dev, (size_t)B * 65536 * 2, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return net_fail(n, DVC_ERR_HIP, "input copy");
    return 0;
}

int dvc_net_forward(dvc_net *n, const void *in_dev, int32_t B, double *par_dev, float *y_dev) {
    if (!n || B < 1) return DVC_ERR_ARG;
    if (B > n->cap) return net_fail(n, DVC_ERR_ARG, "batch above capacity");
    hipStream_t st = (hipStream_t)ctx_stream(n->ctx);
    if (int e = stage_input(n, in_dev, B, st)) return e;
    if (int e = forward(n, st, B, par_dev, nullptr)) return e;
    if (y_dev && hipMemcpyAsync(y_dev, n->y, sizeof(float) * B * kOut, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return net_fail(n, DVC_ERR_HIP, "output copy");
    return DVC_OK;
}

int dvc_net_gradients(dvc_net *n, const void *in_dev, const float *label_dev, int32_t B, float *grad_host,
                      float *loss_host) {
    if (!n || !label_dev || B < 1) return DVC_ERR_ARG;
    if (!n->training) return net_fail(n, DVC_ERR_STATE, "network created without training buffers");
    if (B > n->cap || B % 4) return net_fail(n, DVC_ERR_ARG, "training batches are multiples of 4 within capacity");
    hipStream_t st = (hipStream_t)ctx_stream(n->ctx);
    if (int e = stage_input(n, in_dev, B, st)) return e;
    if (hipMemcpyAsync(n->label, label_dev, sizeof(float) * B * kOut, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return net_fail(n, DVC_ERR_HIP, "label copy");
    if (int e = forward(n, st, B, nullptr, nullptr)) return e;
    if (int e = backward(n, st, B)) return e;
    if (grad_host &&
        hipMemcpyAsync(grad_host, n->g, sizeof(float) * n->L.total, hipMemcpyDeviceToHost, st) != hipSuccess)
        return net_fail(n, DVC_ERR_HIP, "gradient download");
    if (loss_host && hipMemcpyAsync(loss_host, n->loss, sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess)
        return net_fail(n, DVC_ERR_HIP, "loss download");
    if ((grad_host || loss_host) && hipStreamSynchronize(st) != hipSuccess)
        return net_fail(n, DVC_ERR_HIP, "synchronise");
    return DVC_OK;
}

int dvc_net_train_step(dvc_net *n, const void *in_dev, const float *label_dev, int32_t B, float lr, float *loss_host) {
    if (int e = dvc_net_gradients(n, in_dev, label_dev, B, nullptr, nullptr)) return e;
    hipStream_t st = (hipStream_t)ctx_stream(n->ctx);
    n->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)n->step), c2 = 1.0f - powf(b2, (float)n->step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n->L.total, 256, 1024)), dim3(256), 0, st, n->p, n->g, n->m, n->v,
                       n->L.total, lr, b1, b2, c1, c2, 1e-8f);
    if (int e = launch_ok(n, "adam")) return e;
    if (int e = pack(n, st)) return e;
    if (loss_host) {
        if (hipMemcpyAsync(loss_host, n->loss, sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return net_fail(n, DVC_ERR_HIP, "loss download");
    }
    return DVC_OK;
}

int dvc_synth_warp_tiles(dvc_ctx *ctx, int tile, const double *maps, int64_t nmaps) {
    if (!ctx || !maps || tile < 1) return DVC_ERR_ARG;
    const uint8_t *v0;
    int bpv, W, H, T;
    if (ctx_volume_info(ctx, 0, &v0, &bpv, &W, &H, &T) != 0) return DVC_ERR_STATE;
    const int ntx = (W + tile - 1) / tile, nty = (H + tile - 1) / tile, ntz = (T + tile - 1) / tile;
    if (nmaps != (int64_t)ntx * nty * ntz) return DVC_ERR_ARG;
    hipStream_t st = (hipStream_t)ctx_stream(ctx);
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return DVC_ERR_HIP;
    uint8_t *v1 = ctx_volume_alloc(ctx, 1, bpv, W, H, T);     // the target, same shape and type
    if (!v1) return DVC_ERR_HIP;
    double *d = nullptr;
    if (hipMalloc(&d, sizeof(double) * 12 * nmaps) != hipSuccess) return DVC_ERR_HIP;
    int rc = DVC_OK;
    if (hipMemcpyAsync(d, maps, sizeof(double) * 12 * nmaps, hipMemcpyHostToDevice, st) != hipSuccess) rc = DVC_ERR_HIP;
    const long long n = (long long)W * H * T;
    if (rc == DVC_OK) {
        if (bpv == 1)
            hipLaunchKernelGGL(warp_tiles_kernel<1>, dim3(grid_for(n, 256, 1 << 20)), dim3(256), 0, st, v0, v1, W,
                               H, T, tile, ntx, nty, d);
        else
            hipLaunchKernelGGL(warp_tiles_kernel<2>, dim3(grid_for(n, 256, 1 << 20)), dim3(256), 0, st, v0, v1, W,
                               H, T, tile, ntx, nty, d);
        if (hipGetLastError() != hipSuccess) rc = DVC_ERR_HIP;
    }
    if (hipStreamSynchronize(st) != hipSuccess) rc = DVC_ERR_HIP;
    hipFree(d);
    return rc;
}

int dvc_search_propose(dvc_ctx *ctx, dvc_net *n) {
    if (!ctx || !n || n->ctx != ctx) return DVC_ERR_ARG;
    const double *pts;
    const int *order;
    double *start;
    int npts, nproc;
    if (ctx_prepared(ctx, &pts, &order, &npts, &nproc, &start) != 0)
        return net_fail(n, DVC_ERR_STATE, "prepare a given-start or judged search first");
    hipStream_t st = (hipStream_t)ctx_stream(ctx);
    for (int i0 = 0; i0 < nproc; i0 += n->cap) {
        const int B = std::min(n->cap, nproc - i0);
        if (n->training) {
            if (int e = dvc_net_crop(n, pts, order + i0, B, nullptr, nullptr, nullptr)) return e;
            if (int e = forward(n, st, B, start, order + i0)) return e;
        } else if (int e = propose_fused(n, st, pts, order + i0, B, start)) {
            return e;
        }
    }
    ctx_mark_starts(ctx);
    return DVC_OK;
}

}  // extern "C"
