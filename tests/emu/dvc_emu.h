// dvc_emu.h -- CPU build of digitalvolumecorrelation_amd/csrc/dvc_kernels.hip.
//
// Test infrastructure.  Every thread of a workgroup is a host thread; several
// workgroups can run at once, each with its own LDS (DVC_SHARED slots and dynamic
// LDS), its own __syncthreads barrier, and one barrier per wave through which wave
// shuffles and the 16x16x4 FP32 MFMA exchange lane values.
#pragma once

#include <atomic>
#include <barrier>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#define __device__
#define __global__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(x) alignas(x)

namespace emu {
struct Dim { unsigned x = 0, y = 0, z = 0; };

struct Block {
    unsigned id = 0, threads = 0;
    std::unique_ptr<std::barrier<>> bar;
    std::vector<std::unique_ptr<std::barrier<>>> wbar;
    std::vector<double> xd;
    std::vector<float> xa, xb;
    std::mutex mu;
    std::map<int, std::shared_ptr<unsigned char[]>> slots;
    std::vector<unsigned char> dyn;
};

inline thread_local Dim thread_idx, block_idx;
inline thread_local Block *cur = nullptr;
inline Dim block_dim, grid_dim;
inline int lds_fill = 0xA5;

template <typename T>
T *shared_slot(int key) {
    std::lock_guard<std::mutex> g(cur->mu);
    auto it = cur->slots.find(key);
    if (it == cur->slots.end()) {
        std::shared_ptr<unsigned char[]> p(new unsigned char[sizeof(T) + 16]);
        std::memset(p.get(), lds_fill, sizeof(T) + 16);   // LDS starts with stale contents
        it = cur->slots.emplace(key, p).first;
    }
    return reinterpret_cast<T *>(it->second.get());
}

inline void wave_sync() { cur->wbar[thread_idx.x / 64]->arrive_and_wait(); }

inline double shfl_xor(double v, int mask) {
    unsigned t = thread_idx.x, w = t / 64, l = t % 64;
    cur->xd[w * 64 + l] = v;
    wave_sync();
    double r = cur->xd[w * 64 + (l ^ (unsigned)mask)];
    wave_sync();
    return r;
}

// value of lane `src` of this wave (src may differ per lane)
inline double shfl_idx(double v, int src) {
    unsigned t = thread_idx.x, w = t / 64, l = t % 64;
    cur->xd[w * 64 + l] = v;
    wave_sync();
    double r = cur->xd[w * 64 + ((unsigned)src & 63u)];
    wave_sync();
    return r;
}

// __ballot: bit i set when lane i of this wave passes true
inline unsigned long long ballot(bool b) {
    unsigned t = thread_idx.x, w = t / 64, l = t % 64;
    cur->xd[w * 64 + l] = b ? 1.0 : 0.0;
    wave_sync();
    unsigned long long m = 0;
    for (unsigned i = 0; i < 64 && w * 64 + i < cur->threads; i++)
        if (cur->xd[w * 64 + i] != 0.0) m |= 1ull << i;
    wave_sync();
    return m;
}

typedef float f4 __attribute__((vector_size(16)));
inline f4 mfma16x16x4(float a, float b, f4 c) {
    unsigned t = thread_idx.x, w = t / 64, l = t % 64;
    cur->xa[w * 64 + l] = a;
    cur->xb[w * 64 + l] = b;
    wave_sync();
    // A[i][k] is held by lane i + 16k, B[k][j] by lane j + 16k,
    // D[i][j] is register i % 4 of lane j + 16 (i / 4).
    f4 d = c;
    unsigned j = l % 16;
    for (int r = 0; r < 4; r++) {
        unsigned i = 4 * (l / 16) + r;
        float s = 0.0f;
        for (unsigned k = 0; k < 4; k++) s += cur->xa[w * 64 + i + 16 * k] * cur->xb[w * 64 + j + 16 * k];
        d[r] += s;
    }
    wave_sync();
    return d;
}

// Run `blocks` workgroups of `threads` host threads each, all at once, starting at
// block index `first`.  Each workgroup gets `dyn_bytes` of dynamic LDS.
template <typename F>
void run_grid(unsigned first, unsigned blocks, unsigned threads, size_t dyn_bytes, F body) {
    std::vector<std::unique_ptr<Block>> bl;
    for (unsigned b = 0; b < blocks; b++) {
        auto B = std::make_unique<Block>();
        B->id = first + b;
        B->threads = threads;
        B->bar = std::make_unique<std::barrier<>>(threads);
        for (unsigned w = 0; w < (threads + 63) / 64; w++) B->wbar.push_back(std::make_unique<std::barrier<>>(64));
        B->xd.assign(threads, 0.0);
        B->xa.assign(threads, 0.0f);
        B->xb.assign(threads, 0.0f);
        B->dyn.assign(dyn_bytes + 64, (unsigned char)lds_fill);
        bl.push_back(std::move(B));
    }
    block_dim.x = threads;
    std::vector<std::thread> th;
    for (unsigned b = 0; b < blocks; b++)
        for (unsigned t = 0; t < threads; t++) {
            Block *B = bl[b].get();
            th.emplace_back([t, B, &body] {
                cur = B;
                thread_idx.x = t;
                block_idx.x = B->id;
                body();
            });
        }
    for (auto &x : th) x.join();
}
}  // namespace emu

#define DVC_SHARED(type, name) \
    using name##_t = type;     \
    name##_t &name = *emu::shared_slot<name##_t>(__LINE__)
#define DVC_DYN_LDS(type, name) type *name = (type *)emu::cur->dyn.data()

#define threadIdx (emu::thread_idx)
#define blockIdx (emu::block_idx)
#define blockDim (emu::block_dim)
#define gridDim (emu::grid_dim)

inline void __syncthreads() { emu::cur->bar->arrive_and_wait(); }
inline double __shfl_xor(double v, int mask, int) { return emu::shfl_xor(v, mask); }
#define __builtin_amdgcn_wave_barrier() emu::wave_sync()
#define __ballot(x) emu::ballot(x)
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)
#define __builtin_amdgcn_s_sleep(n) std::this_thread::yield()
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, x, y, z) emu::mfma16x16x4((a), (b), (c))
// a reciprocal estimate of float precision over the whole double range, as the hardware's
inline double __builtin_amdgcn_rcp(double x) {
    int e = 0;
    const double m = std::frexp(x, &e);
    return std::ldexp((double)(1.0f / (float)m), -e);
}
inline uint32_t __builtin_amdgcn_alignbyte(uint32_t hi, uint32_t lo, uint32_t s) {
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (8 * (s & 3)));
}
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), (order))
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), (order))
inline int atomicAdd(int *p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicOr(int *p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
inline long long wall_clock64() {
    return (long long)(std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() * 1e8);
}
