// Register passes of the fused coset kernel.
//
// A group's steps are split into passes: runs of consecutive steps whose orbit masks (in the group's
// coset coordinates t) span at most J <= 4 dimensions.  With a basis b_0..b_{J-1} of that span, the
// elements t0 ^ sum_l E_l b_l (E in [0, 2^J)) of a "pass orbit" are closed under every step of the
// pass, so one thread loads them from LDS into registers once, applies all steps with compile-time
// register indices (a step's mask is the local pattern PI = its coordinates), and stores them once.
//
// Roles.  The reference formulas name the orbit element whose physical pivot bit is zero ("c").  In
// t coordinates that bit is t bit P = top bit of the step's mask; P is a pivot of the pass span, so it
// is zero in t0, and element E has t bit P = parity(E & pm), pm = {l : b_l holds bit P}.  The parity is
// wave-uniform (E is a compile-time constant), so a role choice is a scalar branch.  The mix-pair
// update is symmetric in its two elements and needs no role at all.
//
// Signs.  The sign of element E at a step is (-1)^(sb0 ^ u(E)): sb0 is per lane (tile base, lane
// coordinate, orbit base, offset bit), u(E) = parity(E & lbv) is wave-uniform.  The lane part is folded
// once per step into F1 = (-1)^sb0 f1; the uniform part becomes a +-1.0 scalar operand.  Products with
// +-1 are exact, so every result is bit-identical to the reference formulas evaluated with the
// per-element signs.
#pragma once
#include "dcs_kernels.hpp"

constexpr int kPassMax = 4;

struct PassDev {
    int32_t nmask, nsteps, first_pstep, pad;
    int32_t piv[kPassMax];   // t-space pivots of the pass span, ascending (zero in the orbit base t0)
    uint32_t mw[kPassMax];   // basis of the span (t coordinates)
};

struct PStepDev {
    u64 beta;                // physical sign mask (for the tile base)
    int32_t kind, step;
    uint32_t pat, pat2;      // local pattern of the step's mask; quad: single-bit patterns of K, Ka
    uint32_t pm, pma;        // role masks (basis vectors holding the K / Ka pivot t bit)
    uint32_t lbv;            // bit l = parity(b_l & bvec)
    uint32_t bvec, beta_lane;
    uint32_t dead_t;         // t-space basis indices whose pivot slot is not live before the step
};

template <int E>
struct Ctz {
    static constexpr int v = (E & 1) ? 0 : 1 + Ctz<(E >> 1)>::v;
};
template <>
struct Ctz<0> {
    static constexpr int v = 0;
};
template <int E>
struct Top {
    static constexpr int v = (E >> 1) ? 1 + Top<(E >> 1)>::v : 0;
};
template <>
struct Top<0> {
    static constexpr int v = -1;
};

__device__ __forceinline__ int par(uint32_t x) { return __popc(x) & 1; }
__device__ __forceinline__ double pm1(int s) { return s ? -1.0 : 1.0; }

// x * su for su = +-1 (exact)
__device__ __forceinline__ cd cscale(cd x, double su) { return cd{x.r * su, x.i * su}; }
// r + su * p, one rounding per component (the product by +-1 is exact)
__device__ __forceinline__ cd cfold(double su, cd p, cd r) { return cd{fma(su, p.r, r.r), fma(su, p.i, r.i)}; }

// ---- pair kinds -----------------------------------------------------------------------------------
// A = element playing "c", B = its partner; uA, uB: uniform sign parts; F1 = (-1)^sb0 f1.
template <int KIND>
__device__ __forceinline__ void pair_apply(cd &A, cd &B, double uA, double uB, cd f0, cd F1, int y) {
    if constexpr (KIND == 0) {  // grow: B <- s_c f1 a
        const cd a = A;
        A = cmul(f0, a);
        B = cscale(cmul(F1, a), uA);
    } else if constexpr (KIND == 1) {  // mix pair (symmetric)
        const cd a = A, b = B;
        A = cfold(uB, cmul(F1, b), cmul(f0, a));
        B = cfold(uA, cmul(F1, a), cmul(f0, b));
    } else if constexpr (KIND == 3) {  // replace
        const cd a = A, b = B;
        A = cmul(f0, cadd(a, b));
        B = cmul(F1, cfold(uA, a, cscale(b, uB)));
    } else if constexpr (KIND == 5) {  // fixed shrink
        const cd a = A, b = B;
        A = y ? cmul(F1, cfold(uA, a, cscale(b, uB))) : cmul(f0, cadd(a, b));
    } else {  // shrink pair
        const cd a = A, b = B;
        const cd wa{fma(uA, F1.r, f0.r), fma(uA, F1.i, f0.i)};
        const cd wb{fma(uB, F1.r, f0.r), fma(uB, F1.i, f0.i)};
        A = cadd(cmul(wa, a), cmul(wb, b));
    }
}

// Role variants are distinct functions with identical arguments, each closed by a distinct empty asm
// marker, so the two branches are never merged into one sequence on a runtime-selected element
// (which would move the register array to scratch memory).
template <int KIND, int SW>
__device__ __forceinline__ void pair_role(cd &lo, cd &hi, double ulo, double uhi, cd f0, cd F1, int y) {
    if constexpr (SW == 0) {
        pair_apply<KIND>(lo, hi, ulo, uhi, f0, F1, y);
        asm volatile("; pair role 0");
    } else {
        pair_apply<KIND>(hi, lo, uhi, ulo, f0, F1, y);
        asm volatile("; pair role 1");
    }
}

template <int J, int PI, int KIND, int E = 0>
__device__ __forceinline__ void pair_loop(cd (&x)[1 << J], const PStepDev &st, cd f0, cd F1, int y) {
    if constexpr (E < (1 << J)) {
        if constexpr (((E >> Top<PI>::v) & 1) == 0) {
            constexpr int EP = E ^ PI;
            const double ulo = pm1(par(E & st.lbv)), uhi = pm1(par(EP & st.lbv));
            if constexpr (KIND == 1) {
                pair_apply<1>(x[E], x[EP], ulo, uhi, f0, F1, y);
            } else {
                if (par(E & st.pm)) pair_role<KIND, 1>(x[E], x[EP], ulo, uhi, f0, F1, y);
                else pair_role<KIND, 0>(x[E], x[EP], ulo, uhi, f0, F1, y);
            }
        }
        pair_loop<J, PI, KIND, E + 1>(x, st, f0, F1, y);
    }
}

template <int J, int PI>
__device__ __forceinline__ void pair_kind(cd (&x)[1 << J], const PStepDev &st, cd f0, cd F1, int y) {
    switch (st.kind) {
        case 0: pair_loop<J, PI, 0>(x, st, f0, F1, y); break;
        case 1: pair_loop<J, PI, 1>(x, st, f0, F1, y); break;
        case 3: pair_loop<J, PI, 3>(x, st, f0, F1, y); break;
        case 5: pair_loop<J, PI, 5>(x, st, f0, F1, y); break;
        default: pair_loop<J, PI, 7>(x, st, f0, F1, y); break;
    }
}

template <int J, int PI = 1>
__device__ __forceinline__ void pair_dispatch(cd (&x)[1 << J], const PStepDev &st, cd f0, cd F1, int y) {
    if constexpr (PI < (1 << J)) {
        if (st.pat == PI) pair_kind<J, PI>(x, st, f0, F1, y);
        else pair_dispatch<J, PI + 1>(x, st, f0, F1, y);
    }
}

// ---- quad (kind 6): K at local bit B1, Ka at local bit B2 -------------------------------------------
// v00..v11 are e, e^K, e^Ka, e^K^Ka; V = 2*alpha + beta names the element e^alpha K^beta Ka that plays
// "c"; c1 = c^K, d = c^Ka, d1 = d^K.
__device__ __forceinline__ void quad_apply(cd &c, cd &c1, cd &d, cd &d1, double uc, double uc1, double ud, double ud1,
                                           cd f0, cd F1) {
    const cd a = c, b = c1, e = d, g = d1;
    c = cadd(cmul(f0, cadd(a, b)), cmul(F1, cfold(ud, e, cscale(g, ud1))));
    d = cadd(cmul(f0, cadd(e, g)), cmul(F1, cfold(uc, a, cscale(b, uc1))));
}

template <int V>
__device__ __forceinline__ void quad_role(cd &v00, cd &v10, cd &v01, cd &v11, double u00, double u10, double u01,
                                          double u11, cd f0, cd F1) {
    if constexpr (V == 0) {
        quad_apply(v00, v10, v01, v11, u00, u10, u01, u11, f0, F1);
        asm volatile("; quad role 0");
    } else if constexpr (V == 1) {
        quad_apply(v01, v11, v00, v10, u01, u11, u00, u10, f0, F1);
        asm volatile("; quad role 1");
    } else if constexpr (V == 2) {
        quad_apply(v10, v00, v11, v01, u10, u00, u11, u01, f0, F1);
        asm volatile("; quad role 2");
    } else {
        quad_apply(v11, v01, v10, v00, u11, u01, u10, u00, f0, F1);
        asm volatile("; quad role 3");
    }
}

template <int J, int B1, int B2, int E = 0>
__device__ __forceinline__ void quad_loop(cd (&x)[1 << J], const PStepDev &st, cd f0, cd F1) {
    if constexpr (E < (1 << J)) {
        if constexpr (((E >> B1) & 1) == 0 && ((E >> B2) & 1) == 0) {
            constexpr int b1 = 1 << B1, b2 = 1 << B2;
            const int al = par(E & st.pm);
            const int be = par(E & st.pma) ^ (al & (int)((st.pma >> B1) & 1));
            const double u00 = pm1(par(E & st.lbv)), u10 = pm1(par((E ^ b1) & st.lbv));
            const double u01 = pm1(par((E ^ b2) & st.lbv)), u11 = pm1(par((E ^ b1 ^ b2) & st.lbv));
            switch (al * 2 + be) {
                case 0: quad_role<0>(x[E], x[E ^ b1], x[E ^ b2], x[E ^ b1 ^ b2], u00, u10, u01, u11, f0, F1); break;
                case 1: quad_role<1>(x[E], x[E ^ b1], x[E ^ b2], x[E ^ b1 ^ b2], u00, u10, u01, u11, f0, F1); break;
                case 2: quad_role<2>(x[E], x[E ^ b1], x[E ^ b2], x[E ^ b1 ^ b2], u00, u10, u01, u11, f0, F1); break;
                default: quad_role<3>(x[E], x[E ^ b1], x[E ^ b2], x[E ^ b1 ^ b2], u00, u10, u01, u11, f0, F1); break;
            }
        }
        quad_loop<J, B1, B2, E + 1>(x, st, f0, F1);
    }
}

template <int J, int B1 = 0, int B2 = 0>
__device__ __forceinline__ void quad_dispatch(cd (&x)[1 << J], const PStepDev &st, cd f0, cd F1) {
    if constexpr (B1 < J) {
        if constexpr (B2 >= J) {
            quad_dispatch<J, B1 + 1, 0>(x, st, f0, F1);
        } else if constexpr (B1 == B2) {
            quad_dispatch<J, B1, B2 + 1>(x, st, f0, F1);
        } else {
            if (st.pat == (1u << B1) && st.pat2 == (1u << B2)) quad_loop<J, B1, B2>(x, st, f0, F1);
            else quad_dispatch<J, B1, B2 + 1>(x, st, f0, F1);
        }
    }
}

// ---- diagonal kinds (2 mix without partner, 4 fixed) ------------------------------------------------
template <int J, int E = 0>
__device__ __forceinline__ void diag_loop(cd (&x)[1 << J], const PStepDev &st, cd f0, cd F1, int y, bool fixed) {
    if constexpr (E < (1 << J)) {
        const double u = pm1(par(E & st.lbv));
        if (fixed) {
            x[E] = y ? cmul(x[E], cscale(F1, u)) : cmul(x[E], f0);
        } else {
            const cd w{fma(u, F1.r, f0.r), fma(u, F1.i, f0.i)};
            x[E] = cmul(x[E], w);
        }
        diag_loop<J, E + 1>(x, st, f0, F1, y, fixed);
    }
}

template <int J, int E = 1>
__device__ __forceinline__ void make_index(uint32_t (&idx)[1 << J], const uint32_t *mw) {
    if constexpr (E < (1 << J)) {
        idx[E] = idx[E & (E - 1)] ^ mw[Ctz<E>::v];
        make_index<J, E + 1>(idx, mw);
    }
}

// One pass over the tile (held in FP64 in LDS).  S is the frontier storage type: with EXACT every
// step's results are rounded to it, as the per-step kernel does; otherwise values stay FP64 until the
// group is written back.  tb: physical tile base; dyn_lane: this lane's per-step parameters.
template <class S, int J, bool EXACT>
__device__ __forceinline__ void run_pass(double2 *tile, const PassDev &ps, const PStepDev *__restrict__ psteps,
                                         const void *dyn_lane, int dyn_stride, int d, int wb, u64 tb) {
    constexpr int NE = 1 << J;
    const uint32_t W = 1u << wb;
    const uint32_t norb = (1u << (d - J)) << wb;
    uint32_t mw[kPassMax];
#pragma unroll
    for (int i = 0; i < kPassMax; ++i) mw[i] = (i < J) ? (ps.mw[i] << wb) : 0u;
    for (uint32_t o = threadIdx.x; o < norb; o += blockDim.x) {
        const uint32_t l = o & (W - 1);
        uint32_t t0 = o >> wb;
#pragma unroll
        for (int i = 0; i < J; ++i) t0 = insert0(t0, ps.piv[i]);
        uint32_t idx[NE];
        idx[0] = (t0 << wb) | l;
        make_index<J>(idx, mw);
        cd x[NE];
#pragma unroll
        for (int e = 0; e < NE; ++e) x[e] = Store<double2>::load(tile[idx[e]]);
        for (int s = 0; s < ps.nsteps; ++s) {
            const PStepDev &st = psteps[ps.first_pstep + s];
            const double *dd = (const double *)((const char *)dyn_lane + (size_t)st.step * dyn_stride);
            const cd f0{dd[0], dd[1]};
            const int ov = ((const int32_t *)(dd + 4))[0], y = ((const int32_t *)(dd + 4))[1];
            const int sb0 = (__popc(t0 & st.bvec) + __popc(l & st.beta_lane) + __popcll(tb & st.beta) + ov) & 1;
            const double sl = pm1(sb0);
            const cd F1{dd[2] * sl, dd[3] * sl};
            const int kind = st.kind;
            if (kind == 2 || kind == 4) {
                diag_loop<J>(x, st, f0, F1, y, kind == 4);
            } else if (kind == 6) {
                if constexpr (J >= 2) quad_dispatch<J>(x, st, f0, F1);
            } else {
                pair_dispatch<J>(x, st, f0, F1, y);
            }
            if constexpr (EXACT) {
#pragma unroll
                for (int e = 0; e < NE; ++e) x[e] = Store<S>::load(Store<S>::round(x[e]));
            }
        }
#pragma unroll
        for (int e = 0; e < NE; ++e) tile[idx[e]] = Store<double2>::round(x[e]);
    }
}
