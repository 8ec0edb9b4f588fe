// CPU stand-ins for the HIP device features used by the generated coset-group kernels, so that the
// exact kernel source the library compiles with hipRTC can also be compiled with g++ and run on the
// host: one std::thread per work item, a std::barrier per workgroup, per-wave (64 work items)
// global-to-LDS loads done synchronously, scheduling hints as no-ops.
//
// Test infrastructure only (tests/test_codegen_emulation.py): nothing in the library includes it.
#pragma once
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

struct float2 { float x, y; };
struct double2 { double x, y; };
inline float2 make_float2(float a, float b) { return {a, b}; }
inline double2 make_double2(double a, double b) { return {a, b}; }

#define __global__
#define __device__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __restrict__
#define __align__(x) alignas(x)
#define __shared__ static

struct Dim3 { unsigned x, y, z; };
inline thread_local Dim3 threadIdx, blockIdx;
inline Dim3 gridDim, blockDim;
inline std::barrier<> *g_bar = nullptr;
inline void __syncthreads() { g_bar->arrive_and_wait(); }

inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline double __longlong_as_double(long long b) { double d; memcpy(&d, &b, 8); return d; }
inline long long __double_as_longlong(double d) { long long b; memcpy(&b, &d, 8); return b; }

// butterfly exchange inside a 64-wide wave (the callers reach it with the whole workgroup)
inline double g_shfl[1024];
inline double __shfl_xor(double v, int off) {
    g_shfl[threadIdx.x] = v;
    __syncthreads();
    const double r = g_shfl[threadIdx.x ^ off];
    __syncthreads();
    return r;
}
inline std::mutex g_atomic_mu;
inline unsigned long long atomicMax(unsigned long long *p, unsigned long long v) {
    std::lock_guard<std::mutex> l(g_atomic_mu);
    const auto o = *p;
    if (v > o) *p = v;
    return o;
}
using std::fma;
using std::frexp;
using std::ldexp;

// Run f on a gx x gy grid of workgroups of bx work items, one workgroup after another.
template <class F>
void launch(F f, unsigned gx, unsigned gy, unsigned bx) {
    gridDim = {gx, gy, 1};
    blockDim = {bx, 1, 1};
    for (unsigned by = 0; by < gy; ++by)
        for (unsigned b = 0; b < gx; ++b) {
            std::barrier<> bar(bx);
            g_bar = &bar;
            std::vector<std::thread> th;
            for (unsigned t = 0; t < bx; ++t)
                th.emplace_back([&, t] {
                    threadIdx = {t, 0, 0};
                    blockIdx = {b, by, 0};
                    f();
                });
            for (auto &x : th) x.join();
        }
}

// binary16 with round-to-nearest-even conversions from float (the host compiler may lack _Float16)
struct _Float16 {
    unsigned short b;
    _Float16() = default;
    _Float16(float f) {
        unsigned x;
        memcpy(&x, &f, 4);
        const unsigned short sign = (x >> 16) & 0x8000u;
        const unsigned ax = x & 0x7fffffffu;
        if (ax >= 0x7f800000u) {  // inf, NaN (quiet, payload kept where it fits)
            b = sign | 0x7c00u | (ax > 0x7f800000u ? (0x200u | ((ax >> 13) & 0x3ffu)) : 0u);
            return;
        }
        if (ax >= 0x477ff000u) {  // rounds past the largest finite binary16
            b = sign | 0x7c00u;
            return;
        }
        if (ax < 0x38800000u) {  // binary16 subnormal or zero
            if (ax < 0x33000000u) {
                b = sign;
                return;
            }
            const unsigned e = ax >> 23, mant = (ax & 0x7fffffu) | 0x800000u, shift = 126u - e;
            unsigned h = mant >> shift;
            const unsigned rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1u);
            if (rem > half || (rem == half && (h & 1u))) ++h;
            b = sign | h;
            return;
        }
        unsigned r = ax - 0x38000000u;
        r += 0xfffu + ((r >> 13) & 1u);
        b = sign | (r >> 13);
    }
    operator float() const {
        const unsigned sign = (unsigned)(b & 0x8000u) << 16, e = (b >> 10) & 0x1fu, m = b & 0x3ffu;
        unsigned x;
        if (e == 0) {
            if (m == 0) {
                x = sign;
            } else {
                float f = (float)m * 5.9604644775390625e-08f;  // m * 2^-24
                memcpy(&x, &f, 4);
                x |= sign;
            }
        } else if (e == 31) {
            x = sign | 0x7f800000u | (m << 13);
        } else {
            x = sign | ((e + 112u) << 23) | (m << 13);
        }
        float f;
        memcpy(&f, &x, 4);
        return f;
    }
};

// global-to-LDS load: work item (threadIdx.x & 63) of the issuing wave writes `size` bytes at the
// wave's LDS base + its lane * size (the hardware's M0-relative placement)
inline void emu_global_load_lds(const void *src, void *dst, int size, int off) {
    memcpy((char *)dst + off + (threadIdx.x & 63u) * (unsigned)size, src, size);
}
#define __builtin_amdgcn_global_load_lds(s, d, n, o, a) emu_global_load_lds((const void *)(s), (void *)(d), n, o)
#define __builtin_amdgcn_s_waitcnt(x) ((void)0)
#define __builtin_amdgcn_s_setprio(x) ((void)0)
inline float emu_maximum(float a, float b) { return (a != a || b != b) ? a + b : (a > b ? a : b); }
#define __builtin_elementwise_maximum(a, b) emu_maximum(a, b)
