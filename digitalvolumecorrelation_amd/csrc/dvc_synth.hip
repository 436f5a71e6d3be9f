// dvc_synth.hip -- seeded synthetic volume pairs made on the device, for workloads too
// large to generate on the host (1536^3 uint16: 6.75 GiB per volume).  The reference
// ships no generator; its test data are files.  The construction follows the host
// generator (digitalvolumecorrelation_amd/synth.py): uniform noise, separable Gaussian
// smoothing (truncated at 4 sigma, mirror-reflected at the faces), min-max rescaling
// to the full integer range; the target is the reference resampled by separable
// Catmull-Rom at the inverse of a smooth deformation.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <vector>
#include "dvc_internal.h"

namespace dvc {
namespace synth {

// splitmix64 finaliser (Steele, Lea, Flood 2014): 64 well-mixed bits per index
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void noise_kernel(float *v, long long n, uint64_t seed) {
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st)
        v[i] = (float)(mix64(seed * 0x100000001B3ull + (uint64_t)i) >> 40) * (1.0f / 16777216.0f);
}

// one axis of the separable Gaussian: out(c) = sum_k w[k+R] in(reflect(c + k)), with
// scipy's 'reflect' rule (d c b a | a b c d | d c b a); R < the axis length
__global__ void blur_kernel(const float *in, float *out, int W, int H, int T, int axis, const float *w, int R) {
    const long long n = (long long)W * H * T;
    const long long st = (long long)gridDim.x * blockDim.x;
    const long long stride = axis == 0 ? 1 : (axis == 1 ? (long long)W : (long long)W * H);
    const int len = axis == 0 ? W : (axis == 1 ? H : T);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
        const int c = axis == 0 ? (int)(i % W) : (axis == 1 ? (int)((i / W) % H) : (int)(i / ((long long)W * H)));
        const long long base = i - (long long)c * stride;
        float s = 0.0f;
        for (int k = -R; k <= R; k++) {
            int q = c + k;
            if (q < 0) q = -q - 1;
            if (q >= len) q = 2 * len - q - 1;
            s += w[k + R] * in[base + (long long)q * stride];
        }
        out[i] = s;
    }
}

// per-block minimum and maximum, then a one-block pass over the block results
__global__ void __launch_bounds__(256) minmax_kernel(const float *v, long long n, float *part) {
    __shared__ float lo[256], hi[256];
    float a = 3.0e38f, b = -3.0e38f;
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
        a = fminf(a, v[i]);
        b = fmaxf(b, v[i]);
    }
    lo[threadIdx.x] = a;
    hi[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            lo[threadIdx.x] = fminf(lo[threadIdx.x], lo[threadIdx.x + s]);
            hi[threadIdx.x] = fmaxf(hi[threadIdx.x], hi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = lo[0];
        part[2 * blockIdx.x + 1] = hi[0];
    }
}

// out = rint((v - lo) / (hi - lo) * top), lo/hi from the nblk block results in part
__global__ void quantize_kernel(const float *v, long long n, const float *part, int nblk, int bpv, uint8_t *out) {
    __shared__ float mm[2];
    if (threadIdx.x == 0) {
        float a = part[0], b = part[1];
        for (int k = 1; k < nblk; k++) { a = fminf(a, part[2 * k]); b = fmaxf(b, part[2 * k + 1]); }
        mm[0] = a;
        mm[1] = b;
    }
    __syncthreads();
    const float lo = mm[0], span = mm[1] - mm[0];
    const float top = bpv == 1 ? 255.0f : 65535.0f;
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
        float q = rintf(span > 0.0f ? (v[i] - lo) / span * top : 0.0f);
        q = fminf(fmaxf(q, 0.0f), top);
        if (bpv == 1) out[i] = (uint8_t)q;
        else ((uint16_t *)out)[i] = (uint16_t)q;
    }
}

// Displacement field u(X) = sum_b a_b exp(-|X - c_b|^2 / (2 w_b^2))
//                        + sum_s a_s sin(k_s . X + phi_s) + e (X - C):
// Gaussian bumps (bumps[7b..]: centre xyz, width, amplitude xyz), plane sinusoids
// (waves[7s..]: wave vector xyz, phase, amplitude xyz) and a uniform strain e about C.
struct Field {
    const double *bumps;
    int nb;
    const double *waves;
    int nw;
    double e, C[3];
};

__device__ __forceinline__ void disp(const Field &f, const double X[3], double u[3]) {
    u[0] = f.e * (X[0] - f.C[0]);
    u[1] = f.e * (X[1] - f.C[1]);
    u[2] = f.e * (X[2] - f.C[2]);
    for (int b = 0; b < f.nb; b++) {
        const double *p = f.bumps + 7 * b;
        const double dx = X[0] - p[0], dy = X[1] - p[1], dz = X[2] - p[2];
        const double g = exp(-(dx * dx + dy * dy + dz * dz) / (2.0 * p[3] * p[3]));
        u[0] += p[4] * g;
        u[1] += p[5] * g;
        u[2] += p[6] * g;
    }
    for (int s = 0; s < f.nw; s++) {
        const double *p = f.waves + 7 * s;
        const double g = sin(p[0] * X[0] + p[1] * X[1] + p[2] * X[2] + p[3]);
        u[0] += p[4] * g;
        u[1] += p[5] * g;
        u[2] += p[6] * g;
    }
}

// target(x) = ref(X) with X + u(X) = x (fixed-point iterations X <- x - u(X)), separable
// Catmull-Rom with indices clamped to the volume, rounded and clamped to the voxel range
template <int BPV>
__global__ void warp_field_kernel(const uint8_t *src, uint8_t *dst, int W, int H, int T, Field f, int iters) {
    const long long n = (long long)W * H * T;
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
        const double x[3] = {(double)(i % W), (double)((i / W) % H), (double)(i / ((long long)W * H))};
        double X[3] = {x[0], x[1], x[2]}, u[3];
        for (int it = 0; it < iters; it++) {
            disp(f, X, u);
            X[0] = x[0] - u[0];
            X[1] = x[1] - u[1];
            X[2] = x[2] - u[2];
        }
        const int ix = (int)floor(X[0]), iy = (int)floor(X[1]), iz = (int)floor(X[2]);
        double w[3][4];
        const double tt[3] = {X[0] - ix, X[1] - iy, X[2] - iz};
        for (int a = 0; a < 3; a++) {
            const double t = tt[a], t2 = t * t, t3 = t2 * t;
            w[a][0] = 0.5 * (-t3 + 2 * t2 - t);
            w[a][1] = 0.5 * (3 * t3 - 5 * t2 + 2);
            w[a][2] = 0.5 * (-3 * t3 + 4 * t2 + t);
            w[a][3] = 0.5 * (t3 - t2);
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
                    r += w[0][a] * (BPV == 1 ? (double)src[row + xx] : (double)((const uint16_t *)src)[row + xx]);
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

unsigned blocks_for(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

}  // namespace synth
}  // namespace dvc

using namespace dvc;
using namespace dvc::synth;

extern "C" {

int dvc_synth_noise(dvc_ctx *ctx, int slot, int bytes_per, int64_t wide, int64_t high, int64_t tall, double sigma,
                    uint64_t seed) {
    if (!ctx || (slot != 0 && slot != 1) || (bytes_per != 1 && bytes_per != 2) || wide < 4 || high < 4 || tall < 4 ||
        wide > 65536 || high > 65536 || tall > 65536 || !(sigma > 0.0) || sigma > 100.0)
        return DVC_ERR_ARG;
    const int R = (int)(4.0 * sigma + 0.5);          // scipy: radius = int(truncate * sigma + 0.5)
    if (R >= wide || R >= high || R >= tall) return DVC_ERR_ARG;
    hipStream_t st = (hipStream_t)ctx_stream(ctx);
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return DVC_ERR_HIP;
    const long long n = (long long)wide * high * tall;
    std::vector<float> w(2 * R + 1);
    double tot = 0.0;
    for (int k = -R; k <= R; k++) tot += exp(-0.5 * (double)k * k / (sigma * sigma));
    for (int k = -R; k <= R; k++) w[k + R] = (float)(exp(-0.5 * (double)k * k / (sigma * sigma)) / tot);
    float *a = nullptr, *b = nullptr, *dw = nullptr, *part = nullptr;
    const unsigned nblk = 1024;
    int rc = DVC_OK;
    if (hipMalloc(&a, sizeof(float) * n) != hipSuccess || hipMalloc(&b, sizeof(float) * n) != hipSuccess ||
        hipMalloc(&dw, sizeof(float) * w.size()) != hipSuccess || hipMalloc(&part, sizeof(float) * 2 * nblk) != hipSuccess)
        rc = DVC_ERR_HIP;
    uint8_t *out = rc == DVC_OK ? ctx_volume_alloc(ctx, slot, bytes_per, wide, high, tall) : nullptr;
    if (!out) rc = DVC_ERR_HIP;
    if (rc == DVC_OK) {
        hipMemcpyAsync(dw, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice, st);
        hipLaunchKernelGGL(noise_kernel, dim3(blocks_for(n)), dim3(256), 0, st, a, n, seed);
        hipLaunchKernelGGL(blur_kernel, dim3(blocks_for(n)), dim3(256), 0, st, a, b, (int)wide, (int)high, (int)tall, 0, dw, R);
        hipLaunchKernelGGL(blur_kernel, dim3(blocks_for(n)), dim3(256), 0, st, b, a, (int)wide, (int)high, (int)tall, 1, dw, R);
        hipLaunchKernelGGL(blur_kernel, dim3(blocks_for(n)), dim3(256), 0, st, a, b, (int)wide, (int)high, (int)tall, 2, dw, R);
        hipLaunchKernelGGL(minmax_kernel, dim3(nblk), dim3(256), 0, st, b, n, part);
        hipLaunchKernelGGL(quantize_kernel, dim3(blocks_for(n)), dim3(256), 0, st, b, n, part, (int)nblk, bytes_per, out);
        if (hipGetLastError() != hipSuccess) rc = DVC_ERR_HIP;
    }
    if (hipStreamSynchronize(st) != hipSuccess) rc = DVC_ERR_HIP;
    hipFree(a); hipFree(b); hipFree(dw); hipFree(part);
    return rc;
}

int dvc_synth_warp_field(dvc_ctx *ctx, const double *bumps, int32_t nbumps, const double *waves, int32_t nwaves,
                         double strain, const double center[3], int32_t iterations) {
    if (!ctx || nbumps < 0 || nbumps > 1024 || (nbumps > 0 && !bumps) || nwaves < 0 || nwaves > 1024 ||
        (nwaves > 0 && !waves) || !center || iterations < 1 || iterations > 50)
        return DVC_ERR_ARG;
    for (int b = 0; b < nbumps; b++)
        if (!(bumps[7 * b + 3] > 0.0)) return DVC_ERR_ARG;
    const uint8_t *v0;
    int bpv, W, H, T;
    if (ctx_volume_info(ctx, 0, &v0, &bpv, &W, &H, &T) != 0) return DVC_ERR_STATE;
    hipStream_t st = (hipStream_t)ctx_stream(ctx);
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return DVC_ERR_HIP;
    uint8_t *v1 = ctx_volume_alloc(ctx, 1, bpv, W, H, T);     // the target, same shape and type
    if (!v1) return DVC_ERR_HIP;
    double *d = nullptr;
    if (hipMalloc(&d, sizeof(double) * 7 * (nbumps + nwaves + 1)) != hipSuccess) return DVC_ERR_HIP;
    int rc = DVC_OK;
    if (nbumps > 0 && hipMemcpyAsync(d, bumps, sizeof(double) * 7 * nbumps, hipMemcpyHostToDevice, st) != hipSuccess)
        rc = DVC_ERR_HIP;
    if (nwaves > 0 &&
        hipMemcpyAsync(d + 7 * nbumps, waves, sizeof(double) * 7 * nwaves, hipMemcpyHostToDevice, st) != hipSuccess)
        rc = DVC_ERR_HIP;
    Field f;
    f.bumps = d; f.nb = nbumps; f.waves = d + 7 * nbumps; f.nw = nwaves; f.e = strain;
    for (int q = 0; q < 3; q++) f.C[q] = center[q];
    const long long n = (long long)W * H * T;
    if (rc == DVC_OK) {
        if (bpv == 1)
            hipLaunchKernelGGL(warp_field_kernel<1>, dim3(blocks_for(n)), dim3(256), 0, st, v0, v1, W, H, T, f,
                               iterations);
        else
            hipLaunchKernelGGL(warp_field_kernel<2>, dim3(blocks_for(n)), dim3(256), 0, st, v0, v1, W, H, T, f,
                               iterations);
        if (hipGetLastError() != hipSuccess) rc = DVC_ERR_HIP;
    }
    if (hipStreamSynchronize(st) != hipSuccess) rc = DVC_ERR_HIP;
    hipFree(d);
    return rc;
}

}  // extern "C"
