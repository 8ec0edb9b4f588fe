// Device-side arithmetic and the orbit update formulas shared by every frontier kernel.
//
// Arithmetic convention (used identically by the C oracle): complex products put the left operand
// first and fuse its real part, x*y = (fma(xr, yr, -(xi*yi)), fma(xr, yi, xi*yr)); everything else
// is unfused (the library is compiled with -ffp-contract=off).  Operand order follows the complex128
// orbit executor kernel_opt/inplace_plan.py:97-127.
//
// The same text is compiled into the runtime-generated group kernels (hipRTC), which define
// DCS_RTC and the fixed-width integer types before it.
#pragma once
#ifndef DCS_RTC
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

typedef unsigned long long u64;

struct cd {
    double r, i;
};

__device__ __forceinline__ cd cmul(cd x, cd y) {
    return cd{fma(x.r, y.r, -(x.i * y.i)), fma(x.r, y.i, x.i * y.r)};
}
__device__ __forceinline__ cd cadd(cd a, cd b) { return cd{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return cd{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cd cneg_if(cd a, int neg) { return neg ? cd{-a.r, -a.i} : a; }

// Sign bit of the orbit formulas: 1 when (-1)^{popcount(c & beta) + ov} is negative.
__device__ __forceinline__ int sign_bit(u64 c, u64 beta, int ov) { return (__popcll(c & beta) & 1) ^ ov; }

// Storage traits: load to FP64, round-to-storage on store, and |stored|^2 in storage precision.
template <class S>
struct Store;

template <>
struct Store<float2> {
    __device__ static __forceinline__ cd load(float2 v) { return cd{(double)v.x, (double)v.y}; }
    __device__ static __forceinline__ float2 round(cd v) { return make_float2((float)v.r, (float)v.i); }
    __device__ static __forceinline__ double norm2(float2 v) { return (double)(v.x * v.x + v.y * v.y); }
    __device__ static __forceinline__ float2 zero() { return make_float2(0.f, 0.f); }
};

// IEEE binary16 storage of the real and imaginary parts (4 bytes per entry), the reference's
// half-storage engine (kernel_opt/inplace_half_gpu.py:14-28): a value is loaded as half -> float ->
// double and stored as double -> float (round to nearest) -> half (round to nearest even); |stored|^2
// is formed in float from the stored halves.
struct __align__(4) h16x2 {
    unsigned short x, y;
};

__device__ __forceinline__ float h16_to_f32(unsigned short h) { return (float)__builtin_bit_cast(_Float16, h); }
__device__ __forceinline__ unsigned short f32_to_h16(float f) {
    return __builtin_bit_cast(unsigned short, (_Float16)f);
}

template <>
struct Store<h16x2> {
    __device__ static __forceinline__ cd load(h16x2 v) { return cd{(double)h16_to_f32(v.x), (double)h16_to_f32(v.y)}; }
    __device__ static __forceinline__ h16x2 round(cd v) {
        h16x2 o;
        o.x = f32_to_h16((float)v.r);
        o.y = f32_to_h16((float)v.i);
        return o;
    }
    __device__ static __forceinline__ double norm2(h16x2 v) {
        const float r = h16_to_f32(v.x), i = h16_to_f32(v.y);
        return (double)(r * r + i * i);
    }
    __device__ static __forceinline__ h16x2 zero() { return h16x2{0, 0}; }
};

template <>
struct Store<double2> {
    __device__ static __forceinline__ cd load(double2 v) { return cd{v.x, v.y}; }
    __device__ static __forceinline__ double2 round(cd v) { return make_double2(v.r, v.i); }
    __device__ static __forceinline__ double norm2(double2 v) { return v.x * v.x + v.y * v.y; }
    __device__ static __forceinline__ double2 zero() { return make_double2(0.0, 0.0); }
};

// NaN-propagating maximum of non-negative values (a NaN anywhere must reach the host check).
__device__ __forceinline__ double max_nan(double a, double b) {
    const double m = a > b ? a : b;
    return (a != a || b != b) ? a + b : m;   // a + b is NaN whenever either operand is; no branches
}

// Scale of the next unit from the previous maximum |out|^2:  2^-e with e = frexp(sqrt(m2)).exp,
// as the reference host driver computes it (kernel_opt/inplace_gpu.py:115,123).  With m2 = M 2^E2,
// M in [0.5, 1): sqrt(m2) has frexp exponent ceil(E2 / 2), exactly, for a correctly rounded sqrt.
__device__ __forceinline__ double scale_from_m2(double m2) {
    int E2 = 0;
    (void)frexp(m2, &E2);
    const int e = (m2 == 0.0) ? 0 : ((E2 + 1) >> 1);
    return ldexp(1.0, -e);
}

// Per-lane, per-step parameters of a contraction: the step vertex's local vector (f0, f1), the offset
// parity bit and the forced value of a structurally fixed vertex.  g0, g1 serve the f0-factored group
// kernels: the ratio f1 / f0 in g1 for a factored step, and (P f0, P f1) for the last step of a group,
// P the product of the f0s its factored steps left pending.
struct Dyn {
    double f0r, f0i, f1r, f1i;
    int32_t ov, y;
    int32_t pad[2];
    double g0r, g0i, g1r, g1i;
};

__device__ __forceinline__ double bits2d(u64 b) { return __longlong_as_double((long long)b); }

// Wave-level max, then one atomic per wave (values are >= 0 or NaN, so the u64 order of the
// float64 bit patterns is their numeric order, NaN above +inf).
__device__ __forceinline__ void commit_max(double lm, u64 *slot) {
    for (int off = 32; off > 0; off >>= 1) lm = max_nan(lm, __shfl_xor(lm, off));
    if ((threadIdx.x & 63) == 0) atomicMax(slot, (u64)__double_as_longlong(lm));
}

// Scale, round to storage, store, and fold |stored|^2 (in storage precision) into the running max.
template <class S>
__device__ __forceinline__ void put_scaled(S *E, u64 c, cd v, double scale, double &lm) {
    const S o = Store<S>::round(cd{v.r * scale, v.i * scale});
    E[c] = o;
    lm = max_nan(lm, Store<S>::norm2(o));
}

__device__ __forceinline__ u64 pdep_loop(u64 x, u64 mask) {
    u64 r = 0;
    for (u64 m = mask; m; m &= m - 1) {
        if (x & 1) r |= m & (~m + 1);
        x >>= 1;
    }
    return r;
}

// Insert a zero bit at position p of x.
__device__ __forceinline__ uint32_t insert0(uint32_t x, int p) {
    const uint32_t lo = x & ((1u << p) - 1);
    return ((x ^ lo) << 1) | lo;
}
__device__ __forceinline__ u64 insert0_64(u64 x, int p) {
    const u64 lo = x & ((1ull << p) - 1);
    return ((x ^ lo) << 1) | lo;
}

// ---- the eight orbit updates (kernel_opt/inplace_plan.py:97-127, inplace_gpu.py:28-59) --------
// sc, sd, sc1, sd1: sign bits at c, d = c^Ka, c1 = c^K, d1 = d^K.
__device__ __forceinline__ void orb_grow(cd a, cd f0, cd f1, int sc, cd &o0, cd &o1) {
    o0 = cmul(f0, a);
    o1 = cneg_if(cmul(f1, a), sc);
}
__device__ __forceinline__ void orb_mix_pair(cd a, cd b, cd f0, cd f1, int sc, int sd, cd &o0, cd &o1) {
    o0 = cadd(cmul(f0, a), cmul(cneg_if(f1, sd), b));
    o1 = cadd(cmul(f0, b), cmul(cneg_if(f1, sc), a));
}
__device__ __forceinline__ cd orb_mix_diag(cd a, cd f0, cd f1, int sc) { return cmul(a, cadd(f0, cneg_if(f1, sc))); }
__device__ __forceinline__ void orb_replace(cd a, cd b, cd f0, cd f1, int sc, int sd, cd &o0, cd &o1) {
    o0 = cmul(f0, cadd(a, b));
    o1 = cmul(f1, cadd(cneg_if(a, sc), cneg_if(b, sd)));
}
__device__ __forceinline__ cd orb_fixed(cd a, cd f0, cd f1, int sc, int y) {
    return cmul(a, y ? cneg_if(f1, sc) : f0);
}
__device__ __forceinline__ cd orb_fixed_shrink(cd a, cd b, cd f0, cd f1, int sc, int sd, int y) {
    return y ? cmul(f1, cadd(cneg_if(a, sc), cneg_if(b, sd))) : cmul(f0, cadd(a, b));
}
__device__ __forceinline__ void orb_shrink_quad(cd a, cd b, cd e, cd g, cd f0, cd f1, int sc, int sc1, int sd,
                                                int sd1, cd &o0, cd &o1) {
    o0 = cadd(cmul(f0, cadd(a, b)), cmul(f1, cadd(cneg_if(e, sd), cneg_if(g, sd1))));
    o1 = cadd(cmul(f0, cadd(e, g)), cmul(f1, cadd(cneg_if(a, sc), cneg_if(b, sc1))));
}
__device__ __forceinline__ cd orb_shrink_pair(cd a, cd b, cd f0, cd f1, int sc, int sd) {
    return cadd(cmul(cadd(f0, cneg_if(f1, sc)), a), cmul(cadd(f0, cneg_if(f1, sd)), b));
}
