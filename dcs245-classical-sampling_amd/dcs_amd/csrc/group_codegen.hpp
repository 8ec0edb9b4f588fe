// Source generator of plan-specialised coset-group kernels (compiled at plan set-up with hipRTC).
//
// The generic k_coset kernel reads every group, pass and step descriptor at run time and selects
// the orbit pattern, the role of each element and the uniform part of every sign with scalar
// branches.  For a fixed plan all of these are constants, so this generator writes one kernel per
// fused group in which they are literals: each step becomes straight-line FP64 code on registers,
// signs become negations folded into the arithmetic, and only the per-slice parameters (f0, f1, the
// offset bit ov, the forced value y) and the lane-dependent part of the sign stay dynamic.  The
// per-step formulas are those of dcs_kernels.hpp (orb_*), with the same operand order, so a step's
// result is bit-identical to the per-step kernel's before rounding to storage.
//
// Pipelines of one workgroup (persistent over the tiles of its group):
// Two LDS tiles filled by global-to-LDS loads (lds_buffers 3, complex64 tiles): run the first register
//   pass on the current tile, store the previous tile's results from the other LDS tile, synchronise,
//   issue the next tile's loads into that tile (HBM -> LDS, no registers), run the remaining passes,
//   then wait for the loads.
// Two LDS tiles filled from registers (lds_buffers 2): store the previous tile's results (LDS -> HBM),
//   issue the next tile's loads (HBM -> registers), run the register passes on the current tile (LDS),
//   then move the loaded tile into the other LDS buffer.  Stores and loads are in flight during the
//   passes; the only wait on global memory comes after them.
// Waves raise their priority (s_setprio) while they issue a tile's loads and stores and drop it for the
// register passes, so that memory work of one wave is not held back behind the passes of others.
// One LDS tile (half the LDS, twice the workgroups per CU): issue the next tile's loads, run the
//   passes, then each thread stores its finished entries and puts the loaded ones in their place
//   (the same LDS addresses, so no barrier is needed between the two); the stores drain during the
//   next tile's passes.
//
// Host-side code (included by dcs_amd.hip after the descriptor types).
#pragma once
#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

struct CodegenOptions {
    int storage;       // DCS_STORAGE_C64 / DCS_STORAGE_C128 / DCS_STORAGE_F16X2
    int tile_log2;     // log2 of the largest tile (entries)
    bool exact;        // round to storage precision after every step
    bool lds_storage;  // LDS tiles in storage precision (rounded once per register pass) instead of FP64
    int lds_buffers;   // LDS tiles per workgroup: 2 (double-buffered) or 1; 3: two tiles filled by
                       // asynchronous global-to-LDS loads where a group allows it (else as 1)
    // Per-step coefficient classes for f0-factored steps (nullptr: none): 0 the step runs with (f0, f1);
    // 1, 2, 3 the ratio f1 / f0 is real, imaginary or complex and the step runs with (1, f1 / f0).  Used
    // only where factored() holds.
    const int8_t *cls = nullptr;
    // f0 factoring: wherever the FP64 arithmetic between storage roundings is free to change, i.e. not
    // the per-step rounding mode; complex64 and binary16 storage also need tiles held in storage
    // precision in LDS (rounded once per register pass), complex128 holds FP64 tiles either way.
    bool factored() const { return cls && !exact && (storage == 1 || lds_storage); }
};

namespace codegen {

inline void emitf(std::string &o, const char *fmt, ...) {
    char buf[4096];
    va_list ap;
    va_start(ap, fmt);
    const int n = vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (n >= (int)sizeof(buf)) {
        std::vector<char> big(n + 1);
        va_start(ap, fmt);
        vsnprintf(big.data(), big.size(), fmt, ap);
        va_end(ap);
        o.append(big.data(), n);
    } else {
        o.append(buf, n);
    }
}

inline int par32(uint32_t x) { return __builtin_popcount(x) & 1; }
inline int top32(uint32_t x) { return 31 - __builtin_clz(x); }

// Device-side definitions shared by every generated kernel of one module.
inline std::string prelude(const CodegenOptions &opt, const char *kernels_hpp) {
    std::string o;
    o += "#define DCS_RTC 1\n"
         "typedef unsigned int uint32_t;\ntypedef int int32_t;\n"
         "typedef unsigned long long uint64_t;\ntypedef long long int64_t;\n";
    o += kernels_hpp;
    o += "\n";
    // LDS tiles: complex64 storage with lds_storage holds tiles as complex64, complex128 as itself;
    // binary16 storage holds them as complex64 (lds_storage) or FP64, never as binary16
    const char *st = opt.storage == 0 ? "float2" : opt.storage == 1 ? "double2" : "h16x2";
    const char *lt = opt.storage == 2 ? (opt.lds_storage ? "float2" : "double2")
                                      : ((opt.lds_storage || opt.storage == 1) ? st : "double2");
    emitf(o, "typedef %s St;\ntypedef %s Lt;\n", st, lt);
    o += "typedef __attribute__((address_space(1))) const void *dcs_gptr;\n"
         "typedef __attribute__((address_space(3))) void *dcs_lptr;\n";
    // Workgroup barrier that waits for this wave's LDS instructions only.  __syncthreads() carries a
    // workgroup release fence, which also waits for global-to-LDS loads still in flight; between the
    // register passes of the asynchronous-load pipeline those target the other tile, which nobody
    // reads before the explicit wait after the passes.
    o += "__device__ __forceinline__ void dcs_lds_barrier() { asm volatile(\"s_waitcnt lgkmcnt(0)\\n\\ts_barrier\" ::: \"memory\"); }\n";
    o += R"(
__device__ __forceinline__ cd cneg(cd a) { return cd{-a.r, -a.i}; }
template <int U> __device__ __forceinline__ cd sgn(cd a) { if constexpr (U) return cneg(a); else return a; }
// (-1)^UA a + (-1)^UB b, evaluated as the per-element sign formulas do (negation is exact)
template <int UA, int UB> __device__ __forceinline__ cd psum(cd a, cd b) {
    if constexpr (!UA && !UB) return cadd(a, b);
    else if constexpr (!UA && UB) return csub(a, b);
    else if constexpr (UA && !UB) return csub(b, a);
    else return cneg(cadd(a, b));
}
__device__ __forceinline__ cd LG(Lt v) { return Store<Lt>::load(v); }
__device__ __forceinline__ Lt LP(cd v) { return Store<Lt>::round(v); }
__device__ __forceinline__ cd RS(cd v) { return Store<St>::load(Store<St>::round(v)); }
// p*a + q*b, each component one chain of fused multiply-adds
__device__ __forceinline__ cd cfma2(cd p, cd a, cd q, cd b) {
    double r = q.r * b.r;
    r = fma(-q.i, b.i, r);
    r = fma(p.r, a.r, r);
    r = fma(-p.i, a.i, r);
    double i = q.r * b.i;
    i = fma(q.i, b.r, i);
    i = fma(p.r, a.i, i);
    i = fma(p.i, a.r, i);
    return cd{r, i};
}
// f0-factored steps: p is the lane-signed ratio f1 / f0 of a step whose class C says which parts are
// nonzero (1: real, 2: imaginary, 3: both); rmul<C>(p, b) = p b and radd<C, U>(a, p, b) = a + (-1)^U p b,
// one fused multiply-add per component for C 1 and 2
template <int C> __device__ __forceinline__ cd rmul(cd p, cd b) {
    if constexpr (C == 1) return cd{p.r * b.r, p.r * b.i};
    else if constexpr (C == 2) return cd{-(p.i * b.i), p.i * b.r};
    else return cmul(p, b);
}
template <int C, int U> __device__ __forceinline__ cd radd(cd a, cd p, cd b) {
    const double kr = U ? -p.r : p.r, ki = U ? -p.i : p.i;
    if constexpr (C == 1) return cd{fma(kr, b.r, a.r), fma(kr, b.i, a.i)};
    else if constexpr (C == 2) return cd{fma(-ki, b.i, a.r), fma(ki, b.r, a.i)};
    else return cd{fma(-ki, b.i, fma(kr, b.r, a.r)), fma(ki, b.r, fma(kr, b.i, a.i))};
}
// A complex64 result stored non-temporally (a group's output is not read again before the next group
// runs over the whole frontier; 2% faster on the 64-qubit plan, as is the non-temporal bit of the
// global-to-LDS loads)
__device__ __forceinline__ void dcs_store_nt(float2 *p, float2 r) { typedef float v2 __attribute__((ext_vector_type(2))); __builtin_nontemporal_store(v2{r.x, r.y}, (v2 *)p); }
// Two neighbouring complex64 results in one 16-byte non-temporal store (p 16-byte aligned)
__device__ __forceinline__ void dcs_store_nt2(float2 *p, float2 a, float2 b) { typedef float v4 __attribute__((ext_vector_type(4))); __builtin_nontemporal_store(v4{a.x, a.y, b.x, b.y}, (v4 *)p); }
// NaN-propagating maximum (llvm.maximum: one v_maximum3_f32 on gfx950)
__device__ __forceinline__ float max_nan_f(float a, float b) { return __builtin_elementwise_maximum(a, b); }
)";
    return o;
}

// Bit deposit of a run-time value into the set bits of `mask` (lowest `nbits` bits of the value).
inline std::string deposit_expr(const char *x, u64 mask, int nbits) {
    std::string e = "(0ull";
    int j = 0;
    for (int p = 0; p < 64 && j < nbits; ++p)
        if ((mask >> p) & 1) {
            char buf[96];
            snprintf(buf, sizeof(buf), " | ((((u64)(%s) >> %d) & 1ull) << %d)", x, j, p);
            e += buf;
            ++j;
        }
    return e + ")";
}

struct GroupView {
    const GroupDev *g;
    const PassDev *passes;   // g->npasses entries
    const PStepDev *psteps;  // indexed by PassDev::first_pstep
    const u64 *addrv;        // 2^d entries: physical offset of coset coordinate t
};

inline u64 addr_of(const GroupView &v, uint32_t t) { return v.addrv[t]; }

// Per-element lane deposit (constant part): pdep of a constant lane coordinate.
inline u64 lane_dep_const(const GroupDev &g, uint32_t l) {
    if (g.lane_identity) return l;
    u64 r = 0;
    int j = 0;
    for (int p = 0; p < 64 && j < g.wbits; ++p)
        if ((g.rep_low_mask >> p) & 1) {
            if ((l >> j) & 1) r |= 1ull << p;
            ++j;
        }
    return r;
}

// One register pass.  `T` names the current LDS buffer pointer in the generated code.
// m0: coset coordinates whose address carries bit 0; the entry (t, l) then sits at LDS index
// (t W + l) ^ parity(t & m0), a linear relabelling (zero outside the asynchronous-load mode).
// fold: the group's power-of-two scale multiplies this pass's first step's coefficients (the scale is
// exact, so every later value carries it and the stores need no multiply).  direct: this is the group's
// last pass, and its results go from registers straight to HBM (rounded to storage, dead outputs
// skipped, the maximum |stored|^2 kept) instead of back into the LDS tile; gb<m> / dlo<m> are the
// per-thread address and liveness values set up by emit_direct_setup.
// Bits of `t` at the set positions of `mask`, packed to the low end in order (constant t).
inline uint32_t pext32(uint32_t t, uint32_t mask) {
    uint32_t r = 0;
    int j = 0;
    for (int p = 0; p < 32; ++p)
        if ((mask >> p) & 1) {
            if ((t >> p) & 1) r |= 1u << j;
            ++j;
        }
    return r;
}

// The same as an expression over a run-time value named x.
inline std::string pext_expr(const char *x, uint32_t mask) {
    std::string e = "(0u";
    int j = 0;
    for (int p = 0; p < 32; ++p)
        if ((mask >> p) & 1) {
            char buf[96];
            snprintf(buf, sizeof(buf), " | (((%s >> %d) & 1u) << %d)", x, p, j);
            e += buf;
            ++j;
        }
    return e + ")";
}

// csrc (the ring pipeline's first pass): the pass reads its inputs from the compact buffer C, which
// holds only the entries live before the group (coset coordinates t with t & dead_in_t == 0, at index
// (pext(t, live) W + l) ^ parity(t & m0)), takes dead inputs as zero, and writes every entry of T.
inline void emit_pass(std::string &o, const CodegenOptions &opt, const GroupView &v, const PassDev &ps, int threads,
                      uint32_t m0 = 0, bool fold = false, bool direct = false, bool last = false, bool csrc = false) {
    const GroupDev &g = *v.g;
    const int J = ps.nmask, NE = 1 << J, WB = g.wbits, D = g.d;
    const uint32_t norb = (1u << (D - J)) << WB;
    const int OPT = norb >= (uint32_t)threads ? (int)(norb / threads) : 1;
    uint32_t TE[16];
    for (int E = 0; E < NE; ++E) {
        uint32_t t = 0;
        for (int l = 0; l < J; ++l)
            if ((E >> l) & 1) t ^= ps.mw[l];
        TE[E] = t;
    }
    // mask over the orbit number o whose parity is parity(t & bvec) ^ parity(l & beta_lane)
    auto orbit_sign_mask = [&](uint32_t bvec, uint32_t beta_lane) {
        uint32_t mx = 0;
        for (int j = 0; j < D - J; ++j) {
            uint32_t pos = 1u << j;
            for (int i = 0; i < J; ++i) {
                const uint32_t lo = pos & ((1u << ps.piv[i]) - 1);
                pos = ((pos ^ lo) << 1) | lo;  // insert0(pos, piv[i])
            }
            if (pos & bvec) mx |= 1u << j;
        }
        return (mx << WB) | (beta_lane & ((1u << WB) - 1));
    };
    emitf(o, "    { // pass: span %d, %d steps\n", J, ps.nsteps);
    if (norb < (uint32_t)threads) emitf(o, "    if (tid < %uu) {\n", norb);
    for (int m = 0; m < OPT; ++m) {
        emitf(o, "    const uint32_t o%d = tid + %uu;\n", m, (uint32_t)(m * threads));
        emitf(o, "    const uint32_t l%d = o%d & %uu;\n", m, m, (1u << WB) - 1);
        emitf(o, "    uint32_t t%d = o%d >> %d;\n", m, m, WB);
        for (int i = 0; i < J; ++i) emitf(o, "    t%d = insert0(t%d, %d);\n", m, m, ps.piv[i]);
        if (m0) emitf(o, "    const uint32_t b%d = ((t%d << %d) | l%d) ^ (__popc(t%d & %uu) & 1u);\n", m, m, WB, m, m, m0);
        else emitf(o, "    const uint32_t b%d = (t%d << %d) | l%d;\n", m, m, WB, m);
        if (!csrc) {
            for (int E = 0; E < NE; ++E)
                emitf(o, "    cd x%d_%d = LG(T[b%d ^ %uu]);\n", m, E, m, (TE[E] << WB) ^ (uint32_t)par32(TE[E] & m0));
            continue;
        }
        uint32_t pivm = 0;
        for (int i = 0; i < J; ++i) pivm |= 1u << ps.piv[i];
        const uint32_t DI = v.g->dead_in_t, LM = ((1u << D) - 1) & ~DI, DM = DI & ~pivm;
        char tn[16];
        snprintf(tn, sizeof(tn), "t%d", m);
        emitf(o, "    const uint32_t bc%d = ((%s << %d) | l%d) ^ (__popc(t%d & %uu) & 1u);\n", m, pext_expr(tn, LM).c_str(), WB,
              m, m, m0);
        if (DM) emitf(o, "    const uint32_t dc%d = t%d & %uu;\n", m, m, DM);
        for (int E = 0; E < NE; ++E) {
            const uint32_t te = TE[E];
            const uint32_t ce = (pext32(te, LM) << WB) ^ (uint32_t)par32(te & m0);
            if (te & DI & pivm) emitf(o, "    cd x%d_%d = cd{0.0, 0.0};\n", m, E);          // dead for every orbit
            else if (DM) emitf(o, "    cd x%d_%d = (dc%d == %uu) ? LG(C[bc%d ^ %uu]) : cd{0.0, 0.0};\n", m, E, m, te & DM, m, ce);
            else emitf(o, "    cd x%d_%d = LG(C[bc%d ^ %uu]);\n", m, E, m, ce);
        }
    }
    for (int s = 0; s < ps.nsteps; ++s) {
        const PStepDev &q = v.psteps[ps.first_pstep + s];
        // f0 factoring: a step of class C > 0 (not the group's last) applies (1, rho), rho = f1 / f0, and
        // leaves the factor f0 pending; the pending factors of a group are one complex number per lane
        // (they multiply every entry alike), so the group's last step applies them at once with the
        // coefficients (P f0, P f1), P the product of the pending f0s (g0, g1 of its Dyn entry)
        const bool fac = opt.factored();
        const bool fin = fac && last && s + 1 == ps.nsteps;
        const int C = (fac && !fin) ? opt.cls[q.step] : 0;
        emitf(o, "    { // step %d kind %d%s\n", q.step, q.kind, C ? " (f0 factored)" : fin ? " (group's pending factors)" : "");
        emitf(o, "      const Dyn &q = dl[%d];\n", q.step);
        // the store-free pipeline folds the group's power-of-two scale into a step's coefficients: the
        // first step, or with factoring the last (the factored steps carry no f0 to scale)
        const bool sc = fac ? (fin && direct) : (fold && s == 0);
        if (C == 0) {
            const char *c0 = fin ? "g0" : "f0", *c1 = fin ? "g1" : "f1";
            if (sc) {
                emitf(o, "      const double q_f1r = q.%sr * scale, q_f1i = q.%si * scale;\n", c1, c1);
                emitf(o, "      const cd f0{q.%sr * scale, q.%si * scale};\n", c0, c0);
            } else {
                emitf(o, "      const double q_f1r = q.%sr, q_f1i = q.%si;\n", c1, c1);
                emitf(o, "      const cd f0{q.%sr, q.%si};\n", c0, c0);
            }
        }
        emitf(o, "      const int sg = (__popcll(tb & 0x%llxull) + q.ov) & 1;\n", (unsigned long long)q.beta);
        auto u = [&](int E) { return par32(TE[E] & q.bvec); };
        for (int m = 0; m < OPT; ++m) {
            // parity(t & bvec) + parity(l & beta_lane) is parity(o & M): t and l are bit moves of the
            // orbit number o (l its low WB bits, t the rest with zeros inserted at the pass pivots).
            // Only bit 0 of sb is used.
            emitf(o, "      const int sb%d = __popc(o%d & %uu) + sg;\n", m, m, orbit_sign_mask(q.bvec, q.beta_lane));
            // +-f1 by flipping the sign bits (bit-identical to negation): no divergent branch around the
            // parameter loads
            emitf(o, "      const long long sm%d = (long long)((unsigned long long)sb%d << 63);\n", m, m);
            if (C) {
                // the lane-signed ratio; only the parts its class uses
                emitf(o, "      const cd P%d{%s, %s};\n", m,
                      C & 1 ? (std::string("__longlong_as_double(__double_as_longlong(q.g1r) ^ sm") + std::to_string(m) + ")").c_str() : "0.0",
                      C & 2 ? (std::string("__longlong_as_double(__double_as_longlong(q.g1i) ^ sm") + std::to_string(m) + ")").c_str() : "0.0");
                continue;
            }
            emitf(o, "      const cd F%d{__longlong_as_double(__double_as_longlong(q_f1r) ^ sm%d), "
                     "__longlong_as_double(__double_as_longlong(q_f1i) ^ sm%d)};\n", m, m, m);
            if (q.kind == 2 || q.kind == 7)
                emitf(o, "      const cd W%d_0 = cadd(f0, F%d), W%d_1 = csub(f0, F%d);\n", m, m, m, m);
            if (q.kind == 4) emitf(o, "      const cd C%d_0 = q.y ? F%d : f0, C%d_1 = q.y ? cneg(F%d) : f0;\n", m, m, m, m);
        }
        // Elementary updates of this step, each keyed by the element playing "c" (an update is needed
        // only when that element is live).  With complex64 or binary16 storage outside the per-step
        // rounding mode, two-product sums use one chain of fused multiply-adds.
        const bool fma2 = opt.storage != 1 && !opt.exact;
        for (int m = 0; m < OPT; ++m) {
            std::vector<std::pair<int, std::string>> ops;
            char buf[512];
            auto add = [&](int A, const char *fmt, ...) {
                va_list ap;
                va_start(ap, fmt);
                vsnprintf(buf, sizeof(buf), fmt, ap);
                va_end(ap);
                ops.push_back({A, std::string(buf)});
            };
            if (C) {
                // the orbit formulas with (f0, F) replaced by (1, P): W_u = 1 + (-1)^u P, C_u = y ? (-1)^u P : 1,
                // and f0 a + F b = a + P b
                if (q.kind == 2 || q.kind == 4) {
                    for (int E = 0; E < NE; ++E) {
                        if (q.kind == 2) add(E, "x%d_%d = radd<%d, %d>(x%d_%d, P%d, x%d_%d);", m, E, C, u(E), m, E, m, m, E);
                        else add(E, "if (q.y) x%d_%d = sgn<%d>(rmul<%d>(P%d, x%d_%d));", m, E, u(E), C, m, m, E);
                    }
                } else if (q.kind == 6) {
                    const uint32_t b1 = q.pat, b2 = q.pat2;
                    for (int E = 0; E < NE; ++E) {
                        if ((E & b1) || (E & b2)) continue;
                        int c = -1;
                        const int mem[4] = {E, (int)(E ^ b1), (int)(E ^ b2), (int)(E ^ b1 ^ b2)};
                        for (int X : mem)
                            if (!par32(X & q.pm) && !par32(X & q.pma)) c = X;
                        const int c1 = c ^ b1, dd = c ^ b2, d1 = dd ^ b1;
                        add(c, "{ const cd a = x%d_%d, b = x%d_%d, e = x%d_%d, g = x%d_%d; "
                               "x%d_%d = radd<%d, 0>(cadd(a, b), P%d, psum<%d, %d>(e, g)); "
                               "x%d_%d = radd<%d, 0>(cadd(e, g), P%d, psum<%d, %d>(a, b)); }",
                            m, c, m, c1, m, dd, m, d1, m, c, C, m, u(dd), u(d1), m, dd, C, m, u(c), u(c1));
                    }
                } else {
                    const uint32_t PI = q.pat;
                    const int tp = top32(PI);
                    for (int E = 0; E < NE; ++E) {
                        if ((E >> tp) & 1) continue;
                        const int EP = E ^ PI;
                        const int A = par32(E & q.pm) ? EP : E, B = A ^ PI;
                        switch (q.kind) {
                            case 0:
                                add(A, "x%d_%d = sgn<%d>(rmul<%d>(P%d, x%d_%d));", m, B, u(A), C, m, m, A);
                                break;
                            case 1:
                                add(A, "{ const cd a = x%d_%d, b = x%d_%d; x%d_%d = radd<%d, %d>(a, P%d, b); "
                                       "x%d_%d = radd<%d, %d>(b, P%d, a); }",
                                    m, A, m, B, m, A, C, u(B), m, m, B, C, u(A), m);
                                break;
                            case 3:
                                add(A, "{ const cd a = x%d_%d, b = x%d_%d; x%d_%d = cadd(a, b); x%d_%d = rmul<%d>(P%d, psum<%d, %d>(a, b)); }",
                                    m, A, m, B, m, A, m, B, C, m, u(A), u(B));
                                break;
                            case 5:
                                add(A, "{ const cd a = x%d_%d, b = x%d_%d; x%d_%d = q.y ? rmul<%d>(P%d, psum<%d, %d>(a, b)) : cadd(a, b); }",
                                    m, A, m, B, m, A, C, m, u(A), u(B));
                                break;
                            default:  // 7
                                add(A, "{ const cd a = x%d_%d, b = x%d_%d; x%d_%d = radd<%d, 0>(cadd(a, b), P%d, psum<%d, %d>(a, b)); }",
                                    m, A, m, B, m, A, C, m, u(A), u(B));
                                break;
                        }
                    }
                }
            } else if (q.kind == 2 || q.kind == 4) {
                for (int E = 0; E < NE; ++E)
                    add(E, "x%d_%d = cmul(x%d_%d, %c%d_%d);", m, E, m, E, q.kind == 2 ? 'W' : 'C', m, u(E));
            } else if (q.kind == 6) {
                const uint32_t b1 = q.pat, b2 = q.pat2;
                for (int E = 0; E < NE; ++E) {
                    if ((E & b1) || (E & b2)) continue;
                    int c = -1;
                    const int mem[4] = {E, (int)(E ^ b1), (int)(E ^ b2), (int)(E ^ b1 ^ b2)};
                    for (int X : mem)
                        if (!par32(X & q.pm) && !par32(X & q.pma)) c = X;
                    const int c1 = c ^ b1, dd = c ^ b2, d1 = dd ^ b1;
                    if (fma2)
                        add(c, "{ const cd a = x%d_%d, b = x%d_%d, e = x%d_%d, g = x%d_%d; "
                               "x%d_%d = cfma2(f0, cadd(a, b), F%d, psum<%d, %d>(e, g)); "
                               "x%d_%d = cfma2(f0, cadd(e, g), F%d, psum<%d, %d>(a, b)); }",
                            m, c, m, c1, m, dd, m, d1, m, c, m, u(dd), u(d1), m, dd, m, u(c), u(c1));
                    else
                        add(c, "{ const cd a = x%d_%d, b = x%d_%d, e = x%d_%d, g = x%d_%d; "
                               "x%d_%d = cadd(cmul(f0, cadd(a, b)), cmul(F%d, psum<%d, %d>(e, g))); "
                               "x%d_%d = cadd(cmul(f0, cadd(e, g)), cmul(F%d, psum<%d, %d>(a, b))); }",
                            m, c, m, c1, m, dd, m, d1, m, c, m, u(dd), u(d1), m, dd, m, u(c), u(c1));
                }
            } else {
                const uint32_t PI = q.pat;
                const int tp = top32(PI);
                for (int E = 0; E < NE; ++E) {
                    if ((E >> tp) & 1) continue;
                    const int EP = E ^ PI;
                    const int A = par32(E & q.pm) ? EP : E, B = A ^ PI;
                    switch (q.kind) {
                        case 0:
                            add(A, "{ const cd a = x%d_%d; x%d_%d = cmul(f0, a); x%d_%d = sgn<%d>(cmul(F%d, a)); }", m, A, m, A, m, B,
                                u(A), m);
                            break;
                        case 1:
                            if (fma2)
                                add(A, "{ const cd a = x%d_%d, b = x%d_%d; x%d_%d = cfma2(f0, a, sgn<%d>(F%d), b); "
                                       "x%d_%d = cfma2(f0, b, sgn<%d>(F%d), a); }",
                                    m, A, m, B, m, A, u(B), m, m, B, u(A), m);
                            else
                                add(A, "{ const cd a = x%d_%d, b = x%d_%d; x%d_%d = cadd(cmul(f0, a), sgn<%d>(cmul(F%d, b))); "
                                       "x%d_%d = cadd(cmul(f0, b), sgn<%d>(cmul(F%d, a))); }",
                                    m, A, m, B, m, A, u(B), m, m, B, u(A), m);
                            break;
                        case 3:
                            add(A, "{ const cd a = x%d_%d, b = x%d_%d; x%d_%d = cmul(f0, cadd(a, b)); "
                                   "x%d_%d = cmul(F%d, psum<%d, %d>(a, b)); }",
                                m, A, m, B, m, A, m, B, m, u(A), u(B));
                            break;
                        case 5:
                            add(A, "{ const cd a = x%d_%d, b = x%d_%d; x%d_%d = q.y ? cmul(F%d, psum<%d, %d>(a, b)) : cmul(f0, cadd(a, b)); }",
                                m, A, m, B, m, A, m, u(A), u(B));
                            break;
                        default:  // 7
                            if (fma2)
                                add(A, "x%d_%d = cfma2(W%d_%d, x%d_%d, W%d_%d, x%d_%d);", m, A, m, u(A), m, A, m, u(B), m, B);
                            else
                                add(A, "x%d_%d = cadd(cmul(W%d_%d, x%d_%d), cmul(W%d_%d, x%d_%d));", m, A, m, u(A), m, A, m, u(B),
                                    m, B);
                            break;
                    }
                }
            }
            // Liveness: element E of this orbit is live before the step iff (t ^ TE[E]) & dead_t == 0.
            // The pass pivots of t are zero, so keys with a pass-pivot bit are dead for every orbit.
            uint32_t pivm = 0;
            for (int i = 0; i < J; ++i) pivm |= 1u << ps.piv[i];
            const uint32_t DT = q.dead_t;
            std::map<uint32_t, std::vector<std::string>> byKey;
            for (auto &op : ops) {
                const uint32_t key = TE[op.first] & DT;
                if (key & pivm) continue;  // statically dead
                std::string code = op.second;
                if (opt.exact) {
                    // round what the update wrote (the outputs of the op) to storage precision
                    for (int E = 0; E < NE; ++E) {
                        char pat[32];
                        snprintf(pat, sizeof(pat), "x%d_%d = ", m, E);
                        if (code.find(pat) != std::string::npos) {
                            snprintf(buf, sizeof(buf), " x%d_%d = RS(x%d_%d);", m, E, m, E);
                            code += buf;
                        }
                    }
                }
                byKey[key].push_back(code);
            }
            const bool runtime = (DT & ~pivm) != 0;
            if (runtime) emitf(o, "      const uint32_t ct%d = t%d & %uu;\n", m, m, DT);
            for (auto &kv : byKey) {
                if (runtime) emitf(o, "      if (ct%d == %uu) {\n", m, kv.first);
                for (auto &c : kv.second) emitf(o, "      %s\n", c.c_str());
                if (runtime) emitf(o, "      }\n");
            }
        }
        emitf(o, "    }\n");
    }
    if (direct) {
        // element E of orbit m is coset coordinate t_m ^ TE[E] (t_m with zeros at the pass pivots) and lane
        // l_m: its address is tb ^ gb_m ^ addrv(TE[E]); it is dead after the group when its coordinate
        // meets dead_out_t (statically when TE[E] does so at a pass pivot)
        uint32_t pivm = 0;
        for (int i = 0; i < J; ++i) pivm |= 1u << ps.piv[i];
        const uint32_t DO = v.g->dead_out_t;
        for (int m = 0; m < OPT; ++m)
            for (int E = 0; E < NE; ++E) {
                if (TE[E] & DO & pivm) continue;
                const unsigned long long C = (unsigned long long)addr_of(v, TE[E]);
                const bool rt = (DO & ~pivm) != 0;
                if (rt) emitf(o, "    if (dlo%d == %uu) {\n", m, TE[E] & DO & ~pivm);
                emitf(o, "    { const float2 r = LP(x%d_%d); dcs_store_nt(&E[tb ^ gb%d ^ 0x%llxull], r); "
                         "lmf = max_nan_f(lmf, r.x * r.x + r.y * r.y); }\n", m, E, m, C);
                if (rt) emitf(o, "    }\n");
            }
    } else {
        for (int m = 0; m < OPT; ++m)
            for (int E = 0; E < NE; ++E)
                emitf(o, "    T[b%d ^ %uu] = LP(x%d_%d);\n", m, (TE[E] << WB) ^ (uint32_t)par32(TE[E] & m0), m, E);
    }
    if (norb < (uint32_t)threads) emitf(o, "    }\n");
    emitf(o, "    }\n");
}

// Per-thread constants of a group's last pass when it stores straight to HBM (emit_pass direct): for
// each of the thread's orbits m, gb<m> = lane l_m ^ addrv(t_m) (the address part that does not change
// from tile to tile; the lane deposit is the identity in the pipelines that use this) and dlo<m> =
// t_m & dead_out_t outside the pass pivots.
inline void emit_direct_setup(std::string &o, const GroupView &v, const PassDev &ps, int threads) {
    const GroupDev &g = *v.g;
    const int J = ps.nmask, WB = g.wbits, D = g.d;
    const uint32_t norb = (1u << (D - J)) << WB;
    const int OPT = norb >= (uint32_t)threads ? (int)(norb / threads) : 1;
    uint32_t pivm = 0;
    for (int i = 0; i < J; ++i) pivm |= 1u << ps.piv[i];
    for (int m = 0; m < OPT; ++m) {
        emitf(o, "  u64 gb%d;\n  uint32_t dlo%d;\n  {\n", m, m);
        emitf(o, "    const uint32_t o = tid + %uu;\n", (uint32_t)(m * threads));
        emitf(o, "    uint32_t t = o >> %d;\n", WB);
        for (int i = 0; i < J; ++i) emitf(o, "    t = insert0(t, %d);\n", ps.piv[i]);
        emitf(o, "    u64 a = (u64)(o & %uu);\n", (1u << WB) - 1);
        for (int i = 0; i < D; ++i)
            if (!((pivm >> i) & 1))
                emitf(o, "    a ^= (0ull - (u64)((t >> %d) & 1u)) & 0x%llxull;\n", i, (unsigned long long)addr_of(v, 1u << i));
        emitf(o, "    gb%d = a;\n    dlo%d = t & %uu;\n  }\n", m, m, g.dead_out_t & ~pivm);
    }
}

// Workgroup size of a group's kernel: one thread per 8 tile entries (256 .. 1024 threads), so that
// larger tiles bring more waves rather than more registers per thread.
inline int group_threads(const GroupDev &g) {
    const uint32_t N = (1u << g.d) << g.wbits;
    int t = 256;
    while (t < 1024 && (uint32_t)t * 8 < N) t *= 2;
    return t;
}

// Kernel `dcs_group_<gi>` of one fused group.
//
// Load / store mapping: thread tid owns elements e = tid + T k of every tile, T threads (element e = t * W + l
// sits at LDS index e).  Deposit and addrv are linear, so an element's address splits into a
// per-thread part (pth, computed once) and a per-k constant.  (Pairing adjacent lanes into 16-byte
// accesses measured slower on the 64-qubit plan and is not used.)
inline std::string group_kernel(const CodegenOptions &opt, int gi, const GroupView &v) {
    const GroupDev &g = *v.g;
    const int threads = group_threads(g);
    const int D = g.d, WB = g.wbits;
    const uint32_t W = 1u << WB, N = (1u << D) << WB;
    const bool c64 = opt.storage == 0;
    const bool lds_is_storage = opt.lds_storage || opt.storage == 1;  // Lt == St (complex64 / complex128)
    const int EB = __builtin_ctz((unsigned)threads);  // bits of the per-thread element offset tid
    const uint32_t span = (uint32_t)threads;
    const int PER = N >= span ? (int)(N / span) : 1;
    const bool guard = N < (uint32_t)threads;
    std::string o;
    // at least 4 waves per SIMD (<= 128 VGPRs): two workgroups of 512 threads per CU
    emitf(o, "extern \"C\" __global__ void __launch_bounds__(%d) __attribute__((amdgpu_waves_per_eu(4))) "
             "dcs_group_%d(St *__restrict__ base, u64 lane_stride,\n", threads, gi);
    o += "    const int *__restrict__ lane_ids, const Dyn *__restrict__ dyn, int nsteps, u64 *__restrict__ dmax, int nunits,\n"
         "    int unit) {\n";
    // Asynchronous tile loads (lds_buffers 3): every lane moves 16 bytes, two complex64 entries that
    // are neighbours in memory (the tile row holds the lowest address bits), straight into the other
    // LDS tile; no registers hold the next tile and there is no put phase.
    // The pair x, x+1 (x even, lane bit 0 clear) sits at addresses a, a ^ 1: in memory order when a is
    // even, swapped when the coset part of a carries bit 0 (the tile base never does, bit 0 being a lane
    // bit).  The LDS index permutation with m0 = the coordinates carrying bit 0 stores both
    // orientations where the 16-byte load puts them.
    const bool dma = opt.lds_buffers >= 3 && c64 && lds_is_storage && g.lane_identity && WB >= 1 && !guard && PER >= 2;
    // lds_buffers 4: the asynchronous-load pipeline whose last register pass stores to HBM from registers
    // (5: the same where a group does not take the ring pipeline below)
    const bool direct = dma && opt.lds_buffers >= 4;
    // lds_buffers 5, ring pipeline: for a group whose input is a half or a quarter live (one or two coset
    // coordinates born inside it), the live inputs of the next two tiles are in flight at once, in a ring
    // of three compact LDS buffers beside the working tile (80 KiB or less, two workgroups per CU).  Each
    // thread issues the same loads and stores every tile (no runtime liveness in the first and last
    // passes), so the wait for a tile's loads counts exactly the memory operations issued after them.
    const PassDev &lastp = v.passes[g.npasses - 1];
    uint32_t last_pivm = 0;
    for (int i = 0; i < lastp.nmask; ++i) last_pivm |= 1u << lastp.piv[i];
    const int Dn = __builtin_popcount(g.dead_in_t);
    const uint32_t SC = Dn <= D ? ((1u << (D - Dn)) << WB) : 0;  // compact entries per tile
    // per wave and tile: LC loads of compact pairs, ST stores of the last pass (its statically live outputs)
    const uint32_t LC = SC / 2 / (uint32_t)threads;
    int ST = 0;
    {
        const int J = lastp.nmask, NE = 1 << J;
        const uint32_t norb_last = (1u << (D - J)) << WB;
        const int OPT = norb_last >= (uint32_t)threads ? (int)(norb_last / (uint32_t)threads) : 0;
        for (int E = 0; E < NE; ++E) {
            uint32_t t = 0;
            for (int l = 0; l < J; ++l)
                if ((E >> l) & 1) t ^= lastp.mw[l];
            if (!(t & g.dead_out_t & last_pivm)) ST += OPT;
        }
    }
    const bool ring = direct && opt.lds_buffers == 5 && Dn >= 1 && Dn <= 2 && (SC / 2) % (uint32_t)threads == 0 &&
                      (g.dead_out_t & ~last_pivm) == 0 &&
                      (((1u << (D - lastp.nmask)) << WB) >= (uint32_t)threads) &&
                      (N + 3 * SC) * 8 <= 80 * 1024 && 2 * (int)LC + 2 * ST <= 63;
    uint32_t m0 = 0;
    if (dma)
        for (int i = 0; i < D; ++i)
            if (addr_of(v, 1u << i) & 1ull) m0 |= 1u << i;
    auto lidx = [&](uint32_t e) { return e ^ (uint32_t)par32((e >> WB) & m0); };
    const bool single = opt.lds_buffers == 1 || (opt.lds_buffers >= 3 && !dma);
    emitf(o, "  __shared__ Lt buf[%u];\n", ring ? N + 3 * SC : single ? N : 2 * N);
    emitf(o, "  const u64 NT = %lluull;\n", (unsigned long long)g.ntiles);
    o += "  u64 ti = blockIdx.x;\n  if (ti >= NT) return;\n"
         "  const uint32_t tid = threadIdx.x;\n"
         "  const int li = blockIdx.y;\n"
         "  St *__restrict__ E = base + (u64)lane_ids[li] * lane_stride;\n"
         "  const Dyn *__restrict__ dl = dyn + (u64)li * nsteps;\n"
         "  const double scale = scale_from_m2(unit ? bits2d(dmax[(u64)li * nunits + unit - 1]) : 1.0);\n";
    if (c64 && lds_is_storage)
        o += "  const float sf = (float)scale;\n"
             "  const bool sfast = scale >= 0x1p-120 && scale <= 0x1p120;  // float product exact and equal\n"
             "  float lmf = 0.0f;\n";
    else
        o += "  double lm = 0.0;\n";
    o += "  const uint32_t vt = tid;\n";
    std::vector<u64> Ck(PER);
    std::vector<int> din_k(PER), dout_k(PER);
    if (WB >= EB || guard) {
        const int vt_lane_bits = std::min(WB, EB);
        const int vt_t_bits = guard ? std::max(0, EB - WB) : 0;
        std::string lp = g.lane_identity ? std::string("(u64)(vt & ") + std::to_string((1u << vt_lane_bits) - 1) + "u)"
                                         : deposit_expr("vt", g.rep_low_mask, vt_lane_bits);
        emitf(o, "  u64 pth = %s;\n", lp.c_str());
        if (vt_t_bits > 0) {
            for (int i = 0; i < std::min(vt_t_bits, D); ++i)
                emitf(o, "  pth ^= (0ull - (u64)((vt >> %d) & 1u)) & 0x%llxull;\n", WB + i, (unsigned long long)addr_of(v, 1u << i));
            emitf(o, "  const bool tdin = ((vt >> %d) & %uu) != 0;\n", WB, g.dead_in_t);
            emitf(o, "  const bool tdout = ((vt >> %d) & %uu) != 0;\n", WB, g.dead_out_t);
        } else {
            o += "  const bool tdin = false, tdout = false;\n";
        }
        for (int k = 0; k < PER; ++k) {
            const uint32_t e0 = (uint32_t)k * span;
            const uint32_t t = guard ? 0u : (e0 >> WB), l = guard ? 0u : (e0 & (W - 1));
            Ck[k] = lane_dep_const(g, l) ^ addr_of(v, t);
            din_k[k] = (t & g.dead_in_t) != 0;
            dout_k[k] = (t & g.dead_out_t) != 0;
        }
    } else {
        const int tb_t = EB - WB;  // t bits carried by vt
        std::string lp = g.lane_identity ? std::string("(u64)(vt & ") + std::to_string(W - 1) + "u)"
                                         : deposit_expr("vt", g.rep_low_mask, WB);
        emitf(o, "  u64 pth = %s;\n", lp.c_str());
        for (int i = 0; i < std::min(tb_t, D); ++i)
            emitf(o, "  pth ^= (0ull - (u64)((vt >> %d) & 1u)) & 0x%llxull;\n", WB + i, (unsigned long long)addr_of(v, 1u << i));
        emitf(o, "  const bool tdin = ((vt >> %d) & %uu) != 0;\n", WB, g.dead_in_t & ((1u << tb_t) - 1));
        emitf(o, "  const bool tdout = ((vt >> %d) & %uu) != 0;\n", WB, g.dead_out_t & ((1u << tb_t) - 1));
        for (int k = 0; k < PER; ++k) {
            const uint32_t t = (uint32_t)k << tb_t;
            Ck[k] = addr_of(v, t);
            din_k[k] = (t & g.dead_in_t) != 0;
            dout_k[k] = (t & g.dead_out_t) != 0;
        }
    }
    // tile base
    {
        std::string dep;
        if (g.high.n < 0) {
            dep = "pdep_loop(x, " + std::to_string((unsigned long long)g.high.mask) + "ull)";
        } else {
            dep = "(0ull";
            for (int r = 0; r < g.high.n; ++r) {
                char buf[128];
                snprintf(buf, sizeof(buf), " | (((x >> %d) & 0x%llxull) << %d)", g.high.src[r],
                         (unsigned long long)((g.high.len[r] >= 64) ? ~0ull : ((1ull << g.high.len[r]) - 1)), g.high.dst[r]);
                dep += buf;
            }
            dep += ")";
        }
        emitf(o, "  auto tile_base = [](u64 x) -> u64 { return %s; };\n", dep.c_str());
    }
    const std::string gexpr = guard ? (std::string(" && tid < ") + std::to_string(N) + "u") : "";
    // loads into registers
    emitf(o, "  St R[%d];\n", PER);
    o += "  auto issue = [&](u64 tbx) {\n";
    for (int k = 0; k < PER; ++k) {
        if (din_k[k]) emitf(o, "    R[%d] = Store<St>::zero();\n", k);
        else
            emitf(o, "    R[%d] = (!tdin%s) ? E[tbx ^ pth ^ 0x%llxull] : Store<St>::zero();\n", k, gexpr.c_str(),
                  (unsigned long long)Ck[k]);
    }
    o += "  };\n";
    o += "  auto put = [&](Lt *P) {\n";
    for (int k = 0; k < PER; ++k) {
        const uint32_t off = (uint32_t)k * span;
        if (guard) {
            emitf(o, "    if (tid < %uu) P[tid] = LP(Store<St>::load(R[%d]));\n", N, k);
        } else {
            emitf(o, "    P[vt + %u] = LP(Store<St>::load(R[%d]));\n", off, k);
        }
    }
    o += "  };\n";
    // fm: 1 the float product (scale exact in float), 2 the FP64 product; the caller branches on the
    // wave-uniform sfast once around a whole store instead of once per entry
    auto store_one = [&](int k, int fm) {
        if (dout_k[k]) return;
        const uint32_t off = (uint32_t)k * span;
        const unsigned long long C = (unsigned long long)Ck[k];
        if (c64 && lds_is_storage) {
            // complex64 values from LDS: the power-of-two scale is exact in float when representable
            const std::string at = guard ? std::string("tid")
                                 : m0 ? std::string("lvt ^ ") + std::to_string(lidx(off)) + "u"
                                      : std::string("vt + ") + std::to_string(off);
            emitf(o, "    if (!tdout%s) { const float2 a = P[%s]; float2 r;\n", gexpr.c_str(), at.c_str());
            if (fm == 1) o += "      r = make_float2(a.x * sf, a.y * sf);\n";
            else o += "      r = make_float2((float)((double)a.x * scale), (float)((double)a.y * scale));\n";
            emitf(o, "      dcs_store_nt(&E[tbx ^ pth ^ 0x%llxull], r);\n", C);
            o += "      lmf = max_nan_f(lmf, r.x * r.x + r.y * r.y); }\n";
        } else {
            emitf(o, "    if (!tdout%s) put_scaled(E, tbx ^ pth ^ 0x%llxull, LG(P[%s]), scale, lm);\n", gexpr.c_str(), C,
                  guard ? "tid" : (std::string("vt + ") + std::to_string(off)).c_str());
        }
    };
    if (m0) emitf(o, "  const uint32_t lvt = vt ^ (__popc((vt >> %d) & %uu) & 1u);\n", WB, m0);
    const bool split = c64 && lds_is_storage;
    o += "  auto store = [&](Lt *P, u64 tbx) {\n";
    if (split) {
        o += "   if (sfast) {\n";
        for (int k = 0; k < PER; ++k) store_one(k, 1);
        o += "   } else {\n";
        for (int k = 0; k < PER; ++k) store_one(k, 2);
        o += "   }\n";
    } else {
        for (int k = 0; k < PER; ++k) store_one(k, 0);
    }
    o += "  };\n";
    if (single) {
        // each thread reads its finished entry, stores it, then writes the loaded entry at the same
        // LDS address: per-thread program order keeps the read before the write
        o += "  auto store_put = [&](Lt *P, u64 tbx) {\n";
        for (int fm : split ? std::vector<int>{1, 2} : std::vector<int>{0}) {
            if (split) o += fm == 1 ? "   if (sfast) {\n" : "   } else {\n";
            for (int k = 0; k < PER; ++k) {
                store_one(k, fm);
                const uint32_t off = (uint32_t)k * span;
                if (guard) emitf(o, "    if (tid < %uu) P[tid] = LP(Store<St>::load(R[%d]));\n", N, k);
                else emitf(o, "    P[vt + %u] = LP(Store<St>::load(R[%d]));\n", off, k);
            }
        }
        if (split) o += "   }\n";
        o += "  };\n";
    }
    if (dma) {
        // lane element x = 2 (j threads + tid): address = tile base ^ pa ^ D_j (linear deposit)
        std::string pa = "0ull";
        for (int p = 1; p <= EB; ++p) {
            const u64 c = (p < WB) ? lane_dep_const(g, 1u << p) : addr_of(v, 1u << (p - WB));
            char b[160];
            snprintf(b, sizeof(b), " ^ ((0ull - (u64)((vt >> %d) & 1u)) & 0x%llxull)", p - 1, (unsigned long long)c);
            pa += b;
        }
        emitf(o, "  const u64 pa = %s;\n", pa.c_str());
        emitf(o, "  const bool tdin2 = ((((uint32_t)vt << 1) >> %d) & %uu) != 0;\n", WB, g.dead_in_t);
        // dead entries are zero-filled first: an LDS store after a global-to-LDS load into the same
        // tile would make the compiler wait for the load
        o += "  auto dma_fill = [&](Lt *P, u64 tbx) {\n";
        for (int j = 0; j < PER / 2; ++j) {
            const uint32_t x = 2u * (uint32_t)j * span;
            const uint32_t t = x >> WB;
            if (t & g.dead_in_t)
                emitf(o, "    P[%uu + 2u * vt] = Store<St>::zero(); P[%uu + 2u * vt + 1u] = Store<St>::zero();\n", x, x);
            else
                emitf(o, "    if (tdin2) { P[%uu + 2u * vt] = Store<St>::zero(); P[%uu + 2u * vt + 1u] = Store<St>::zero(); }\n",
                      x, x);
        }
        for (int j = 0; j < PER / 2; ++j) {
            const uint32_t x = 2u * (uint32_t)j * span;
            const uint32_t t = x >> WB, l = x & (W - 1);
            const unsigned long long Dj = (unsigned long long)(lane_dep_const(g, l) ^ addr_of(v, t));
            if (!(t & g.dead_in_t))
                emitf(o, "    if (!tdin2) __builtin_amdgcn_global_load_lds((dcs_gptr)(E + ((tbx ^ pa ^ 0x%llxull) & ~1ull)), "
                         "(dcs_lptr)(P + %uu + 2u * (vt & ~63u)), 16, 0, 2);\n", Dj, x);
        }
        o += "  };\n";
        // vmcnt(0) with the other counters at their maximum (gfx9 encoding): every load into LDS landed
        o += "  auto dma_wait = [&]() { __builtin_amdgcn_s_waitcnt(0x0F70); };\n";
        if (!direct) {
            // The stores of the asynchronous-load pipeline mirror its loads: each lane stores the 16 bytes
            // of LDS entries x, x + 1, which hold the aligned pair of addresses in memory order, to that
            // pair (scaled, with their maxima), skipping pairs dead at the group's end.  One 16-byte store
            // per two entries instead of two 8-byte ones: 7% faster than those on the 64-qubit plan.
            emitf(o, "  const bool tdout2 = ((((uint32_t)vt << 1) >> %d) & %uu) != 0;\n", WB, g.dead_out_t);
            o += "  auto store_pairs = [&](Lt *P, u64 tbx) {\n   if (sfast) {\n";
            for (int fm = 1; fm <= 2; ++fm) {
                if (fm == 2) o += "   } else {\n";
                for (int j = 0; j < PER / 2; ++j) {
                    const uint32_t x = 2u * (uint32_t)j * span;
                    const uint32_t t = x >> WB, l = x & (W - 1);
                    if (t & g.dead_out_t) continue;
                    const unsigned long long Dj = (unsigned long long)(lane_dep_const(g, l) ^ addr_of(v, t));
                    emitf(o, "    if (!tdout2) { const float2 a = P[%uu + 2u * vt], b = P[%uu + 2u * vt + 1u]; float2 ra, rb;\n",
                          x, x);
                    if (fm == 1)
                        o += "      ra = make_float2(a.x * sf, a.y * sf); rb = make_float2(b.x * sf, b.y * sf);\n";
                    else
                        o += "      ra = make_float2((float)((double)a.x * scale), (float)((double)a.y * scale));\n"
                             "      rb = make_float2((float)((double)b.x * scale), (float)((double)b.y * scale));\n";
                    emitf(o, "      dcs_store_nt2(E + ((tbx ^ pa ^ 0x%llxull) & ~1ull), ra, rb);\n", Dj);
                    o += "      lmf = max_nan_f(lmf, max_nan_f(ra.x * ra.x + ra.y * ra.y, rb.x * rb.x + rb.y * rb.y)); }\n";
                }
            }
            o += "   }\n  };\n";
        }
    }
    // pipeline
    if (ring) {
        // Each tile k (the workgroup's k-th): issue the loads of tile k + 2 into ring slot (c + 2) % 3 (read
        // by tile k - 1's first pass, which every wave has finished: a barrier follows each pass), wait
        // for tile k's loads, synchronise, run the first pass from ring slot c into T, then the other
        // passes on T, the last one storing to HBM from registers.  Per wave and tile: LC loads, ST
        // stores, all unconditional; memory operations complete in issue order, so tile k's loads have
        // landed once at most LC (tile k + 1) + LC (tile k + 2) + ST (tile k - 1) + ST (tile k - 2) of
        // the operations issued after them are outstanding (fewer near the start and the end).
        const uint32_t LM = ((1u << D) - 1) & ~g.dead_in_t;
        // compact pair x (even) of tile base tbx: entries x, x + 1 at t = pdep(x >> WB, LM), l = x & (W - 1)
        auto ccontrib = [&](int p) -> u64 {  // address part of bit p of a compact entry index
            if (p < WB) return lane_dep_const(g, 1u << p);
            int j = p - WB, q = 0;
            for (int b = 0; b < D; ++b)
                if ((LM >> b) & 1) {
                    if (q == j) return addr_of(v, 1u << b);
                    ++q;
                }
            return 0;
        };
        std::string pc = "0ull";
        for (int p = 1; p <= EB; ++p) {
            char b[160];
            snprintf(b, sizeof(b), " ^ ((0ull - (u64)((vt >> %d) & 1u)) & 0x%llxull)", p - 1, (unsigned long long)ccontrib(p));
            pc += b;
        }
        emitf(o, "  const u64 pc = %s;\n", pc.c_str());
        o += "  auto ring_fill = [&](Lt *__restrict__ P, u64 tbx) {\n";
        for (uint32_t k = 0; k < LC; ++k) {
            const uint32_t x = 2u * k * span;
            u64 Dk = 0;
            for (int p = 0; p < 32; ++p)
                if ((x >> p) & 1) Dk ^= ccontrib(p);
            emitf(o, "    __builtin_amdgcn_global_load_lds((dcs_gptr)(E + ((tbx ^ pc ^ 0x%llxull) & ~1ull)), "
                     "(dcs_lptr)(P + %uu + 2u * (vt & ~63u)), 16, 0, 2);\n", (unsigned long long)Dk, x);
        }
        o += "  };\n";
        // s_waitcnt vmcnt(n), the other counters at their maximum (gfx9 encoding: vmcnt bits 3:0 and 15:14)
        auto wait_imm = [](int n) { return (n & 15) | (((n >> 4) & 3) << 14) | 0x0F70; };
        emit_direct_setup(o, v, lastp, threads);
        emitf(o, "  const u64 G = gridDim.x;\n");
        o += "  u64 tb = tile_base(ti);\n";
        o += "  auto passes = [&](Lt *__restrict__ T, Lt *__restrict__ C, Lt *__restrict__ F, bool more2, u64 tn2, int k, bool more) {\n"
             "    if (more2) { __builtin_amdgcn_s_setprio(2); ring_fill(F, tile_base(tn2)); __builtin_amdgcn_s_setprio(0); }\n";
        // the wait: k in {0, 1, >= 2} times (more, more2) in {(1, 1), (1, 0), (0, 0)}
        o += "    {\n      const int ks = k < 2 ? k : 2;\n";
        for (int ks = 0; ks <= 2; ++ks)
            for (int mm = 0; mm <= 2; ++mm) {  // mm = more + more2
                const int n = (int)LC * mm + ST * ks;
                emitf(o, "      %sif (ks == %d && (int)more + (int)more2 == %d) __builtin_amdgcn_s_waitcnt(0x%x);\n",
                      (ks || mm) ? "else " : "", ks, mm, wait_imm(n));
            }
        o += "    }\n    dcs_lds_barrier();\n";
        for (int pi = 0; pi < g.npasses; ++pi) {
            emit_pass(o, opt, v, v.passes[pi], threads, m0, pi == 0, pi + 1 == g.npasses, pi + 1 == g.npasses, pi == 0);
            if (pi + 1 < g.npasses) o += "    dcs_lds_barrier();\n";
        }
        o += "  };\n";
        emitf(o, "  Lt *const R0 = buf + %u, *const R1 = buf + %u, *const R2 = buf + %u;\n", N, N + SC, N + 2 * SC);
        o += "  ring_fill(R0, tb);\n"
             "  if (ti + G < NT) ring_fill(R1, tile_base(ti + G));\n"
             "  int k = 0;\n"
             "  for (;;) {\n"
             "    const u64 tn = ti + G, tn2 = ti + 2 * G;\n"
             "    const bool more = tn < NT, more2 = tn2 < NT;\n"
             "    const int c = k % 3;\n"
             "    Lt *const C = c == 0 ? R0 : c == 1 ? R1 : R2;\n"
             "    Lt *const F = c == 0 ? R2 : c == 1 ? R0 : R1;\n"
             "    passes(buf, C, F, more2, tn2, k, more);\n"
             "    if (!more) break;\n"
             "    ++k;\n    ti = tn;\n    tb = tile_base(ti);\n"
             "  }\n";
        o += "  commit_max((double)lmf, &dmax[(u64)li * nunits + unit]);\n";
        o += "}\n";
        return o;
    }
    if (direct) {
        // Each tile: issue the next tile's loads into U (free: the previous tile's last pass stored its
        // results from registers), run the register passes on T, the last one storing to HBM; then wait
        // for the loads and synchronise (every wave is done reading T before it becomes U).  The group's
        // scale is folded into its first step, so stored values need no multiply.
        emit_direct_setup(o, v, v.passes[g.npasses - 1], threads);
        o += "  u64 tb = tile_base(ti);\n";
        o += "  auto passes = [&](Lt *__restrict__ T, Lt *__restrict__ U, bool more, u64 tn) {\n"
             "    if (more) { __builtin_amdgcn_s_setprio(2); dma_fill(U, tile_base(tn)); __builtin_amdgcn_s_setprio(0); }\n";
        for (int pi = 0; pi < g.npasses; ++pi) {
            emit_pass(o, opt, v, v.passes[pi], threads, m0, pi == 0, pi + 1 == g.npasses, pi + 1 == g.npasses);
            if (pi + 1 < g.npasses) o += "    dcs_lds_barrier();\n";
        }
        o += "  };\n";
        o += "  dma_fill(buf, tb);\n  dma_wait();\n  __syncthreads();\n"
             "  int c = 0;\n"
             "  for (;;) {\n";
        emitf(o, "    Lt *T = buf + c * %u;\n    Lt *U = buf + (c ^ 1) * %u;\n", N, N);
        o += "    const u64 tn = ti + gridDim.x;\n"
             "    const bool more = tn < NT;\n"
             "    passes(T, U, more, tn);\n"
             "    if (!more) break;\n"
             "    dma_wait();\n"
             "    __syncthreads();\n"
             "    c ^= 1;\n    ti = tn;\n    tb = tile_base(ti);\n"
             "  }\n";
        o += "  commit_max((double)lmf, &dmax[(u64)li * nunits + unit]);\n";
        o += "}\n";
        return o;
    }
    if (dma) {
        // the tile body takes the two tiles as restrict pointers: the loads into U and the passes'
        // accesses to T get distinct alias scopes, so no wait for the loads is placed before the passes.
        // U holds the previous tile's results until they are stored: each wave stores them (16-byte
        // pairs) after its part of the first register pass (mixing the stores with other waves' FP64
        // work), and the next tile's loads into U follow the barrier that ends that pass.
        o += "  u64 tb = tile_base(ti);\n";
        o += "  auto passes = [&](Lt *__restrict__ T, Lt *__restrict__ U, bool more, u64 tn, bool prev, u64 tprev) {\n"
             "    const u64 tbn = more ? tile_base(tn) : 0ull;\n"
             "    if (more && !prev) { __builtin_amdgcn_s_setprio(2); dma_fill(U, tbn); __builtin_amdgcn_s_setprio(0); }\n";
        for (int pi = 0; pi < g.npasses; ++pi) {
            emit_pass(o, opt, v, v.passes[pi], threads, m0, false, false, pi + 1 == g.npasses);
            if (pi == 0) {
                o += "    if (prev) { __builtin_amdgcn_s_setprio(2); store_pairs(U, tprev); __builtin_amdgcn_s_setprio(0); }\n"
                     "    dcs_lds_barrier();\n"
                     "    if (more && prev) { __builtin_amdgcn_s_setprio(2); dma_fill(U, tbn); __builtin_amdgcn_s_setprio(0); }\n";
            } else if (pi + 1 < g.npasses) {
                o += "    dcs_lds_barrier();\n";
            }
        }
        o += "  };\n";
        o += "  dma_fill(buf, tb);\n  dma_wait();\n  __syncthreads();\n"
             "  int c = 0;\n  bool prev = false;\n  u64 tprev = 0;\n"
             "  for (;;) {\n";
        emitf(o, "    Lt *T = buf + c * %u;\n    Lt *U = buf + (c ^ 1) * %u;\n", N, N);
        o += "    const u64 tn = ti + gridDim.x;\n"
             "    const bool more = tn < NT;\n"
             "    passes(T, U, more, tn, prev, tprev);\n";
        o += "    dma_wait();\n"
             "    __syncthreads();\n"
             "    if (!more) { __builtin_amdgcn_s_setprio(2); store_pairs(T, tb); break; }\n"
             "    prev = true;\n    tprev = tb;\n"
             "    c ^= 1;\n    ti = tn;\n    tb = tile_base(ti);\n"
             "  }\n";
        if (c64 && lds_is_storage) o += "  commit_max((double)lmf, &dmax[(u64)li * nunits + unit]);\n";
        else o += "  commit_max(lm, &dmax[(u64)li * nunits + unit]);\n";
        o += "}\n";
        return o;
    }
    emitf(o, "  u64 tb = tile_base(ti);\n  issue(tb);\n  put(buf);\n  __syncthreads();\n");
    if (single) {
        o += "  for (;;) {\n"
             "    Lt *T = buf;\n"
             "    const u64 tn = ti + gridDim.x;\n"
             "    const bool more = tn < NT;\n"
             "    if (more) { __builtin_amdgcn_s_setprio(2); issue(tile_base(tn)); __builtin_amdgcn_s_setprio(0); }\n";
        for (int pi = 0; pi < g.npasses; ++pi) {
            emit_pass(o, opt, v, v.passes[pi], threads, 0, false, false, pi + 1 == g.npasses);
            o += "    __syncthreads();\n";
        }
        o += "    __builtin_amdgcn_s_setprio(2);\n"
             "    if (!more) { store(T, tb); break; }\n"
             "    store_put(T, tb);\n"
             "    __builtin_amdgcn_s_setprio(0);\n"
             "    __syncthreads();\n"
             "    ti = tn;\n    tb = tile_base(ti);\n"
             "  }\n";
    } else {
        emitf(o, "  int c = 0;\n  u64 tprev = 0;\n  bool have_prev = false;\n");
        emitf(o, "  for (;;) {\n");
        emitf(o, "    Lt *T = buf + c * %u;\n    Lt *U = buf + (c ^ 1) * %u;\n", N, N);
        o += "    if (have_prev) store(U, tprev);\n"
             "    const u64 tn = ti + gridDim.x;\n"
             "    const bool more = tn < NT;\n"
             "    if (more) issue(tile_base(tn));\n";
        for (int pi = 0; pi < g.npasses; ++pi) {
            emit_pass(o, opt, v, v.passes[pi], threads, 0, false, false, pi + 1 == g.npasses);
            if (pi + 1 < g.npasses) o += "    __syncthreads();\n";
        }
        o += "    if (more) put(U);\n"
             "    __syncthreads();\n"
             "    if (!more) { store(T, tb); break; }\n"
             "    have_prev = true;\n    tprev = tb;\n    c ^= 1;\n    ti = tn;\n    tb = tile_base(ti);\n"
             "  }\n";
    }
    if (c64 && lds_is_storage) o += "  commit_max((double)lmf, &dmax[(u64)li * nunits + unit]);\n";
    else o += "  commit_max(lm, &dmax[(u64)li * nunits + unit]);\n";
    o += "}\n";
    return o;
}

}  // namespace codegen
