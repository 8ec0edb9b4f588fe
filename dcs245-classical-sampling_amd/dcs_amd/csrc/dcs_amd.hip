// libdcs_amd: MI355X (gfx950) executor of the preserved-coordinate frontier contraction, the
// coherent conditional block and its CDF draw.  C ABI in include/dcs_amd.h.
//
// Kernels
//   k_step      one in-place step over the whole frontier, one thread per orbit; fused
//               power-of-two scale and max |out|^2 (kernel_opt/inplace_gpu.py:20-63).
//   k_resident  a run of consecutive narrow steps whose live slots all sit below 2^RB: the frontier
//               stays in LDS, one workgroup per lane, the exact per-step scale chain.
//   k_coset     a fused group of consecutive steps.  The XOR masks of the group span a subspace V
//               (dim <= 8) of the address space; every coset c0 + V is closed under all steps of
//               the group, so a workgroup loads a tile of cosets from HBM once, applies every step
//               in LDS and writes it back once.  Values are rounded to storage precision after
//               each step exactly as in k_step; only the power-of-two scale is applied once per
//               group, so the stored mantissas equal those of the per-step schedule.
//   k_scatter, k_wht_pass, k_phase, k_u_pass
//               coherent slice sum and the Walsh-Hadamard / phase / local-Clifford tail
//               (sampling/conditional_block_gpu.py:155-171).
//   k_weights, k_chain, k_search
//               |a|^2, the sequential float64 cumsum, normalisation and searchsorted(side='right')
//               of draw_checked (sampling/sampling_protocol.py:56-65); the running sum is one
//               dependent chain in numpy's accumulation order, so every cumulative value is
//               bit-identical to cumsum's.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/dcs_amd.h"
#include "dcs_kernels.hpp"
#include "coset_pass.hpp"

namespace {

thread_local std::string g_error;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(DCS_ERR_RUNTIME, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                  \
    } while (0)

constexpr int kThreads = 256;
constexpr int kMaxRuns = 16;
constexpr int kMaxDim = 8;          // coset dimension of the generic k_coset kernel
constexpr int kMaxDimCompiled = 12; // coset dimension of generated group kernels

struct Runs {
    int n;  // number of runs; -1: use pdep over `mask`
    uint8_t src[kMaxRuns], len[kMaxRuns], dst[kMaxRuns];
    u64 mask;
};

Runs make_runs(u64 mask) {
    Runs r;
    memset(&r, 0, sizeof(r));
    r.mask = mask;
    int src = 0, n = 0;
    for (int p = 0; p < 64;) {
        if (!((mask >> p) & 1)) {
            ++p;
            continue;
        }
        int q = p;
        while (q < 64 && ((mask >> q) & 1)) ++q;
        if (n == kMaxRuns) {
            r.n = -1;
            return r;
        }
        r.src[n] = (uint8_t)src;
        r.len[n] = (uint8_t)(q - p);
        r.dst[n] = (uint8_t)p;
        src += q - p;
        ++n;
        p = q;
    }
    r.n = n;
    return r;
}

__device__ __forceinline__ u64 deposit(u64 x, const Runs &r) {
    if (r.n < 0) return pdep_loop(x, r.mask);
    u64 c = 0;
    for (int i = 0; i < r.n; ++i) c |= ((x >> r.src[i]) & ((1ull << r.len[i]) - 1)) << r.dst[i];
    return c;
}

struct StepArgs {
    u64 norbits, K, Ka, beta;
    int32_t kind, pivot;
    Runs runs;
};

struct StepDev {
    u64 K, Ka, beta, slots, in_mask, out_mask;
    int32_t kind, pivot, pivot_a, nslots;
};

struct GroupDev {
    int32_t d, wbits, lane_identity, npasses;
    int32_t first_pass, addrv_offset;
    uint32_t dead_in_t, dead_out_t;
    u64 rep_low_mask;
    u64 ntiles;
    Runs high;
};

struct DrawOut {
    int64_t index;
    double weight, total, lower, upper, margin;
    double mass_check, lower_check, upper_check;  // the independent compensated-sum interval
    int32_t status;
    int32_t pad;
};

// DrawOut status beyond dcs_status: the independent interval does not contain the uniform.
constexpr int32_t kDrawIntervalDisagrees = 100;
// Workgroups of the draw's independent interval check (k_interval_part).
constexpr int kIntervalBlocks = 256;

// ------------------------------------------------------------------------------------------------
// k_step: one step, one thread per orbit, orbit index deposited into the enumerated slots.
template <class S, int KIND>
__global__ void __launch_bounds__(kThreads) k_step(S *__restrict__ base, u64 lane_stride,
                                                   const int *__restrict__ lane_ids, StepArgs a,
                                                   const Dyn *__restrict__ dyn, int nsteps, int step,
                                                   u64 *__restrict__ dmax, int nunits, int unit) {
    const int li = blockIdx.y;
    S *__restrict__ E = base + (u64)lane_ids[li] * lane_stride;
    const Dyn dd = dyn[(u64)li * nsteps + step];
    const double prev = unit ? bits2d(dmax[(u64)li * nunits + unit - 1]) : 1.0;
    const double scale = scale_from_m2(prev);
    const cd f0{dd.f0r, dd.f0i}, f1{dd.f1r, dd.f1i};
    const int ov = dd.ov, y = dd.y;
    const u64 beta = a.beta;
    double lm = 0.0;
    const u64 stride = (u64)blockDim.x * gridDim.x;
    for (u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x; x < a.norbits; x += stride) {
        const u64 c = deposit(x, a.runs);
        const cd va = Store<S>::load(E[c]);
        const int sc = sign_bit(c, beta, ov);
        if (KIND == 0) {
            cd o0, o1;
            orb_grow(va, f0, f1, sc, o0, o1);
            put_scaled(E, c, o0, scale, lm);
            put_scaled(E, c | (1ull << a.pivot), o1, scale, lm);
        } else if (KIND == 1) {
            const u64 d = c ^ a.Ka;
            const cd vb = Store<S>::load(E[d]);
            cd o0, o1;
            orb_mix_pair(va, vb, f0, f1, sc, sign_bit(d, beta, ov), o0, o1);
            put_scaled(E, c, o0, scale, lm);
            put_scaled(E, d, o1, scale, lm);
        } else if (KIND == 2) {
            put_scaled(E, c, orb_mix_diag(va, f0, f1, sc), scale, lm);
        } else if (KIND == 3) {
            const u64 d = c ^ a.K;
            const cd vb = Store<S>::load(E[d]);
            cd o0, o1;
            orb_replace(va, vb, f0, f1, sc, sign_bit(d, beta, ov), o0, o1);
            put_scaled(E, c, o0, scale, lm);
            put_scaled(E, d, o1, scale, lm);
        } else if (KIND == 4) {
            put_scaled(E, c, orb_fixed(va, f0, f1, sc, y), scale, lm);
        } else if (KIND == 5) {
            const u64 d = c ^ a.K;
            const cd vb = Store<S>::load(E[d]);
            put_scaled(E, c, orb_fixed_shrink(va, vb, f0, f1, sc, sign_bit(d, beta, ov), y), scale, lm);
        } else if (KIND == 6) {
            const u64 c1 = c ^ a.K, d = c ^ a.Ka, d1 = d ^ a.K;
            const cd vb = Store<S>::load(E[c1]), ve = Store<S>::load(E[d]), vg = Store<S>::load(E[d1]);
            cd o0, o1;
            orb_shrink_quad(va, vb, ve, vg, f0, f1, sc, sign_bit(c1, beta, ov), sign_bit(d, beta, ov),
                            sign_bit(d1, beta, ov), o0, o1);
            put_scaled(E, c, o0, scale, lm);
            put_scaled(E, d, o1, scale, lm);
        } else {
            const u64 d = c ^ a.K;
            const cd vb = Store<S>::load(E[d]);
            put_scaled(E, c, orb_shrink_pair(va, vb, f0, f1, sc, sign_bit(d, beta, ov)), scale, lm);
        }
    }
    commit_max(lm, &dmax[(u64)li * nunits + unit]);
}

// ------------------------------------------------------------------------------------------------
// Apply one step of kind `kind` to a frontier held in LDS (physical address == LDS index).
template <class S>
__device__ __forceinline__ void resident_step(S *tile, const StepDev &st, const Dyn &dd, double scale,
                                              double &lm) {
    const cd f0{dd.f0r, dd.f0i}, f1{dd.f1r, dd.f1i};
    const int ov = dd.ov, y = dd.y;
    const u64 beta = st.beta;
    const u64 norb = 1ull << st.nslots;
    for (u64 x = threadIdx.x; x < norb; x += blockDim.x) {
        const u64 c = pdep_loop(x, st.slots);
        const cd va = Store<S>::load(tile[c]);
        const int sc = sign_bit(c, beta, ov);
        switch (st.kind) {
            case 0: {
                cd o0, o1;
                orb_grow(va, f0, f1, sc, o0, o1);
                put_scaled(tile, c, o0, scale, lm);
                put_scaled(tile, c | (1ull << st.pivot), o1, scale, lm);
            } break;
            case 1: {
                const u64 d = c ^ st.Ka;
                const cd vb = Store<S>::load(tile[d]);
                cd o0, o1;
                orb_mix_pair(va, vb, f0, f1, sc, sign_bit(d, beta, ov), o0, o1);
                put_scaled(tile, c, o0, scale, lm);
                put_scaled(tile, d, o1, scale, lm);
            } break;
            case 2:
                put_scaled(tile, c, orb_mix_diag(va, f0, f1, sc), scale, lm);
                break;
            case 3: {
                const u64 d = c ^ st.K;
                const cd vb = Store<S>::load(tile[d]);
                cd o0, o1;
                orb_replace(va, vb, f0, f1, sc, sign_bit(d, beta, ov), o0, o1);
                put_scaled(tile, c, o0, scale, lm);
                put_scaled(tile, d, o1, scale, lm);
            } break;
            case 4:
                put_scaled(tile, c, orb_fixed(va, f0, f1, sc, y), scale, lm);
                break;
            case 5: {
                const u64 d = c ^ st.K;
                const cd vb = Store<S>::load(tile[d]);
                put_scaled(tile, c, orb_fixed_shrink(va, vb, f0, f1, sc, sign_bit(d, beta, ov), y), scale, lm);
            } break;
            case 6: {
                const u64 c1 = c ^ st.K, d = c ^ st.Ka, d1 = d ^ st.K;
                const cd vb = Store<S>::load(tile[c1]), ve = Store<S>::load(tile[d]),
                         vg = Store<S>::load(tile[d1]);
                cd o0, o1;
                orb_shrink_quad(va, vb, ve, vg, f0, f1, sc, sign_bit(c1, beta, ov), sign_bit(d, beta, ov),
                                sign_bit(d1, beta, ov), o0, o1);
                put_scaled(tile, c, o0, scale, lm);
                put_scaled(tile, d, o1, scale, lm);
            } break;
            default: {
                const u64 d = c ^ st.K;
                const cd vb = Store<S>::load(tile[d]);
                put_scaled(tile, c, orb_shrink_pair(va, vb, f0, f1, sc, sign_bit(d, beta, ov)), scale, lm);
            } break;
        }
    }
}

__device__ __forceinline__ double block_max(double lm, double *red) {
    for (int off = 32; off > 0; off >>= 1) lm = max_nan(lm, __shfl_xor(lm, off));
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wave] = lm;
    __syncthreads();
    double m = red[0];
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 1; w < nw; ++w) m = max_nan(m, red[w]);
    __syncthreads();
    return m;
}

// A run of resident steps s0..s1-1 on lane `li`: the frontier of the lane is loaded into LDS once, every
// step runs there with the exact per-step scale chain, and the live slots are written back.
template <class S, int RB>
__device__ __forceinline__ void resident_run(S *__restrict__ tile, double *__restrict__ red, S *__restrict__ base,
                                             u64 lane_stride, const int *__restrict__ lane_ids,
                                             const StepDev *__restrict__ steps, int s0, int s1,
                                             const Dyn *__restrict__ dyn, int nsteps, u64 *__restrict__ dmax,
                                             int nunits, int unit0, int li) {
    S *__restrict__ E = base + (u64)lane_ids[li] * lane_stride;
    const u64 live_in = steps[s0].in_mask;
    for (u64 e = threadIdx.x; e < (1ull << RB); e += blockDim.x)
        tile[e] = (e & ~live_in) ? Store<S>::zero() : E[e];
    double prev = unit0 ? bits2d(dmax[(u64)li * nunits + unit0 - 1]) : 1.0;
    __syncthreads();
    for (int s = s0; s < s1; ++s) {
        const StepDev st = steps[s];
        const Dyn dd = dyn[(u64)li * nsteps + s];
        double lm = 0.0;
        resident_step(tile, st, dd, scale_from_m2(prev), lm);
        const double m2 = block_max(lm, red);  // contains the barrier that orders this step's writes
        if (threadIdx.x == 0) dmax[(u64)li * nunits + unit0 + (s - s0)] = (u64)__double_as_longlong(m2);
        prev = m2;
    }
    const u64 live_out = steps[s1 - 1].out_mask;
    for (u64 e = threadIdx.x; e < (1ull << RB); e += blockDim.x)
        if (!(e & ~live_out)) E[e] = tile[e];
}

template <class S, int RB>
__global__ void __launch_bounds__(kThreads) k_resident(S *__restrict__ base, u64 lane_stride,
                                                       const int *__restrict__ lane_ids,
                                                       const StepDev *__restrict__ steps, int s0, int s1,
                                                       const Dyn *__restrict__ dyn, int nsteps,
                                                       u64 *__restrict__ dmax, int nunits, int unit0) {
    __shared__ S tile[1 << RB];
    __shared__ double red[kThreads / 64];
    resident_run<S, RB>(tile, red, base, lane_stride, lane_ids, steps, s0, s1, dyn, nsteps, dmax, nunits, unit0,
                        blockIdx.x);
}

// ------------------------------------------------------------------------------------------------
// k_coset: fused group.  Tile layout: element (t, l) at LDS index t*W + l, t in [0, 2^d) the
// coordinate in the RREF basis of V (address = c0(l) ^ addrv[t]), l the lane coordinate.
// Tiles slot, slot + nslots, ... of group gi on lane `li` (the persistent loop of one workgroup).
template <class S, int TL, bool EXACT, int JMAX>
__device__ __forceinline__ void coset_run(double2 *__restrict__ tile, u64 *__restrict__ addrv, S *__restrict__ base,
                                          u64 lane_stride, const int *__restrict__ lane_ids,
                                          const GroupDev *__restrict__ groups, int gi,
                                          const PassDev *__restrict__ passes, const PStepDev *__restrict__ psteps,
                                          const u64 *__restrict__ addrv_all, const Dyn *__restrict__ dyn, int nsteps,
                                          u64 *__restrict__ dmax, int nunits, int unit, int li, u64 slot, u64 nslots) {
    const GroupDev &g = groups[gi];
    S *__restrict__ E = base + (u64)lane_ids[li] * lane_stride;
    const Dyn *dyn_lane = dyn + (u64)li * nsteps;
    const int d = g.d, wb = g.wbits;
    const uint32_t W = 1u << wb;
    const uint32_t n = (1u << d) << wb;
    for (uint32_t t = threadIdx.x; t < (1u << d); t += blockDim.x) addrv[t] = addrv_all[g.addrv_offset + t];
    const double prev = unit ? bits2d(dmax[(u64)li * nunits + unit - 1]) : 1.0;
    const double scale = scale_from_m2(prev);
    double lm = 0.0;
    __syncthreads();
    // Software pipeline: the next tile's HBM loads are issued into registers before the current
    // tile's passes run, so each workgroup overlaps its own memory traffic with its LDS/VALU work.
    constexpr int PER = (1 << TL) / kThreads;
    S R[PER];
    auto issue = [&](u64 tbase) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t e = threadIdx.x + k * kThreads;
            S v = Store<S>::zero();
            if (e < n) {
                const uint32_t t = e >> wb, l = e & (W - 1);
                if (!(t & g.dead_in_t)) {
                    const u64 lp = g.lane_identity ? (u64)l : pdep_loop(l, g.rep_low_mask);
                    v = E[tbase ^ lp ^ addrv[t]];
                }
            }
            R[k] = v;
        }
    };
    u64 ti = slot;
    u64 tb = 0;
    if (ti < g.ntiles) {
        tb = deposit(ti, g.high);
        issue(tb);
    }
    while (ti < g.ntiles) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t e = threadIdx.x + k * kThreads;
            if (e < n) {
                const cd v = Store<S>::load(R[k]);
                tile[e] = make_double2(v.r, v.i);
            }
        }
        __syncthreads();
        const u64 tn = ti + nslots;
        u64 tb_next = 0;
        if (tn < g.ntiles) {
            tb_next = deposit(tn, g.high);
            issue(tb_next);
        }
        for (int pi = 0; pi < g.npasses; ++pi) {
            const PassDev &ps = passes[g.first_pass + pi];
            switch (ps.nmask) {
                case 0: run_pass<S, 0, EXACT>(tile, ps, psteps, dyn_lane, sizeof(Dyn), d, wb, tb); break;
                case 1: run_pass<S, 1, EXACT>(tile, ps, psteps, dyn_lane, sizeof(Dyn), d, wb, tb); break;
                case 2: run_pass<S, 2, EXACT>(tile, ps, psteps, dyn_lane, sizeof(Dyn), d, wb, tb); break;
                case 3: if constexpr (JMAX >= 3) run_pass<S, 3, EXACT>(tile, ps, psteps, dyn_lane, sizeof(Dyn), d, wb, tb); break;
                default: if constexpr (JMAX >= 4) run_pass<S, 4, EXACT>(tile, ps, psteps, dyn_lane, sizeof(Dyn), d, wb, tb); break;
            }
            __syncthreads();
        }
        for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
            const uint32_t t = e >> wb, l = e & (W - 1);
            if (t & g.dead_out_t) continue;
            const u64 lp = g.lane_identity ? (u64)l : pdep_loop(l, g.rep_low_mask);
            put_scaled(E, tb ^ lp ^ addrv[t], Store<double2>::load(tile[e]), scale, lm);
        }
        __syncthreads();
        ti = tn;
        tb = tb_next;
    }
    commit_max(lm, &dmax[(u64)li * nunits + unit]);
}

template <class S, int TL, bool EXACT, int JMAX>
__global__ void __launch_bounds__(kThreads) k_coset(S *__restrict__ base, u64 lane_stride,
                                                    const int *__restrict__ lane_ids,
                                                    const GroupDev *__restrict__ groups, int gi,
                                                    const PassDev *__restrict__ passes,
                                                    const PStepDev *__restrict__ psteps,
                                                    const u64 *__restrict__ addrv_all,
                                                    const Dyn *__restrict__ dyn, int nsteps,
                                                    u64 *__restrict__ dmax, int nunits, int unit) {
    __shared__ double2 tile[1 << TL];
    __shared__ u64 addrv[1 << kMaxDim];
    coset_run<S, TL, EXACT, JMAX>(tile, addrv, base, lane_stride, lane_ids, groups, gi, passes, psteps, addrv_all, dyn,
                                  nsteps, dmax, nunits, unit, blockIdx.y, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------------------------------------------
// Launches shared by several plans (dcs_contract_batch).  One entry per plan: its frontier, lanes,
// tables, parameters and maxima, and the segment it runs in this launch.  Workgroups
// [begin, begin + nlanes * blocks) of the launch belong to the entry; workgroup begin + r serves lane
// r / blocks at tile slot r % blocks.  The plans differ in width, group shapes and tile counts; each
// entry's segment runs through the same device code as the single-plan kernels.
struct BatchEntry {
    void *base;
    u64 lane_stride;
    const int *lanes;
    const StepDev *steps;
    const GroupDev *groups;
    const PassDev *passes;
    const PStepDev *psteps;
    const u64 *addrv;
    const Dyn *dyn;
    u64 *dmax;
    int32_t nsteps, nunits, unit;
    int32_t first, last;  // coset: group index (first); resident run: steps first .. last-1
    int32_t nlanes;
    uint32_t blocks, begin;
};

// Entry of this workgroup: the last one whose first workgroup is not after it (entries are few).
__device__ __forceinline__ int batch_entry(const BatchEntry *__restrict__ e, int n) {
    int i = 0;
    while (i + 1 < n && blockIdx.x >= e[i + 1].begin) ++i;
    return i;
}

template <class S, int RB>
__global__ void __launch_bounds__(kThreads) k_resident_batch(const BatchEntry *__restrict__ ents, int n) {
    __shared__ S tile[1 << RB];
    __shared__ double red[kThreads / 64];
    const BatchEntry &b = ents[batch_entry(ents, n)];
    resident_run<S, RB>(tile, red, (S *)b.base, b.lane_stride, b.lanes, b.steps, b.first, b.last, b.dyn, b.nsteps,
                        b.dmax, b.nunits, b.unit, (int)(blockIdx.x - b.begin));
}

template <class S, int TL, bool EXACT, int JMAX>
__global__ void __launch_bounds__(kThreads) k_coset_batch(const BatchEntry *__restrict__ ents, int n) {
    __shared__ double2 tile[1 << TL];
    __shared__ u64 addrv[1 << kMaxDim];
    const BatchEntry &b = ents[batch_entry(ents, n)];
    const uint32_t r = blockIdx.x - b.begin;
    coset_run<S, TL, EXACT, JMAX>(tile, addrv, (S *)b.base, b.lane_stride, b.lanes, b.groups, b.first, b.passes,
                                  b.psteps, b.addrv, b.dyn, b.nsteps, b.dmax, b.nunits, b.unit, (int)(r / b.blocks),
                                  r % b.blocks, b.blocks);
}

// ------------------------------------------------------------------------------------------------
// k_gather: one step of the canonical gather executor (dcs_rankwidth/gpu_kernel.py:58-101) with the
// fused power-of-two scale and max |out|^2.  The step's byte tables (<= 5 x 256 u64) and kernel
// combinations are staged in LDS.  Sum over combinations in their order, products e * w with the
// library's complex-product convention.
struct GatherArgs {
    u64 n_out, mask, beta;
    const u64 *T, *combos;
    int32_t r_in, nbytes, ncombos, is_free;
};

template <class S>
__global__ void __launch_bounds__(kThreads) k_gather(const S *__restrict__ Ein, S *__restrict__ Eout, GatherArgs a,
                                                     const Dyn *__restrict__ dyn, int step, u64 *__restrict__ dmax,
                                                     int unit) {
    __shared__ u64 tab[5 * 256];
    __shared__ u64 cmb[8];
    for (int i = threadIdx.x; i < a.nbytes * 256; i += blockDim.x) tab[i] = a.T[i];
    if ((int)threadIdx.x < a.ncombos) cmb[threadIdx.x] = a.combos[threadIdx.x];
    __syncthreads();
    const Dyn dd = dyn[step];
    const cd f0{dd.f0r, dd.f0i}, f1{dd.f1r, dd.f1i};
    const double scale = scale_from_m2(unit ? bits2d(dmax[unit - 1]) : 1.0);
    double lm = 0.0;
    for (u64 cp = (u64)blockIdx.x * blockDim.x + threadIdx.x; cp < a.n_out; cp += (u64)blockDim.x * gridDim.x) {
        u64 p0 = 0;
        for (int b = 0; b < a.nbytes; ++b) p0 ^= tab[b * 256 + ((cp >> (8 * b)) & 255)];
        cd acc{0.0, 0.0};
        for (int k = 0; k < a.ncombos; ++k) {
            const u64 p = p0 ^ cmb[k];
            const u64 c = p & a.mask;
            const int par = (__popcll(c & a.beta) & 1) ^ dd.ov;
            const int yv = a.is_free ? (int)((p >> a.r_in) & 1) : dd.y;
            acc = cadd(acc, cmul(Store<S>::load(Ein[c]), cneg_if(yv ? f1 : f0, yv & par)));
        }
        put_scaled(Eout, cp, acc, scale, lm);
    }
    commit_max(lm, &dmax[unit]);
}

// ------------------------------------------------------------------------------------------------
__global__ void k_fill_u64(u64 *p, u64 n, u64 v) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)blockDim.x * gridDim.x) p[i] = v;
}

template <class S>
__global__ void k_init_lanes(S *base, u64 lane_stride, const int *lane_ids, int nlanes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nlanes) base[(u64)lane_ids[i] * lane_stride] = Store<S>::round(cd{1.0, 0.0});
}

template <class S>
__global__ void k_scatter(const S *__restrict__ E, Runs term, u64 n, const u64 *__restrict__ smap, u64 xmask,
                          cd coef, double2 *__restrict__ total) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)blockDim.x * gridDim.x) {
        const cd p = cmul(Store<S>::load(E[deposit(i, term)]), coef);
        const u64 j = smap[i] ^ xmask;
        const double2 t = total[j];
        total[j] = make_double2(t.x + p.r, t.y + p.i);
    }
}

template <class S>
__global__ void k_gather_terminal(const S *__restrict__ E, Runs term, u64 n, S *__restrict__ out) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)blockDim.x * gridDim.x)
        out[i] = E[deposit(i, term)];
}

__global__ void k_wht_pass(double2 *v, u64 half, int q) {
    for (u64 p = (u64)blockIdx.x * blockDim.x + threadIdx.x; p < half; p += (u64)blockDim.x * gridDim.x) {
        const u64 i0 = insert0_64(p, q), i1 = i0 | (1ull << q);
        const double2 lo = v[i0], hi = v[i1];
        v[i0] = make_double2(lo.x + hi.x, lo.y + hi.y);
        v[i1] = make_double2(lo.x - hi.x, lo.y - hi.y);
    }
}

__global__ void k_phase(double2 *v, const double *phase, u64 n) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)blockDim.x * gridDim.x) {
        const cd r = cmul(cd{v[i].x, v[i].y}, cd{phase[i], 0.0});
        v[i] = make_double2(r.r, r.i);
    }
}

__global__ void k_u_pass(double2 *v, u64 half, int q, cd u00, cd u01, cd u10, cd u11) {
    for (u64 p = (u64)blockIdx.x * blockDim.x + threadIdx.x; p < half; p += (u64)blockDim.x * gridDim.x) {
        const u64 i0 = insert0_64(p, q), i1 = i0 | (1ull << q);
        const cd lo{v[i0].x, v[i0].y}, hi{v[i1].x, v[i1].y};
        const cd a = cadd(cmul(u00, lo), cmul(u01, hi));
        const cd b = cadd(cmul(u10, lo), cmul(u11, hi));
        v[i0] = make_double2(a.r, a.i);
        v[i1] = make_double2(b.r, b.i);
    }
}

// Weights w = re*re + im*im of the block (numpy's real**2 + imag**2, unfused).
__global__ void k_weights(const double2 *__restrict__ v, u64 n, double *__restrict__ w) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)blockDim.x * gridDim.x) {
        const double2 a = v[i];
        w[i] = a.x * a.x + a.y * a.y;
    }
}

// The strictly sequential float64 running sum of numpy's cumsum, as one dependent chain of adds in a
// single wavefront.  Chunks of 512 weights are staged in LDS (the next chunk is loaded into registers
// while the chain runs over the current one) and read back with wave-uniform addresses.  Only every
// 64th running sum is kept,
//   ck[j] = w[0] + ... + w[64 j + 63]  (j < n / 64),  ck[n / 64] = the total;
// k_search recomputes the 64 sums of one segment from ck[j - 1] in the same order, so every
// cumulative value it uses is bit-identical to cumsum's.
//
// W is the weight source: WeightArray reads the weights k_weights wrote, BlockWeights forms them from
// the block itself (the same unfused re*re + im*im, so the same bits) and needs no weight buffer.
struct WeightArray {
    const double *__restrict__ w;
    __device__ __forceinline__ double operator()(u64 i) const { return w[i]; }
};

struct BlockWeights {
    const double2 *__restrict__ v;
    __device__ __forceinline__ double operator()(u64 i) const {
        const double2 a = v[i];
        return a.x * a.x + a.y * a.y;
    }
};

template <class W>
__global__ void __launch_bounds__(64) k_chain(W w, u64 n, double *__restrict__ ck) {
    __shared__ double seg[512];
    const int lane = threadIdx.x;
    const u64 nchunk = (n + 511) >> 9;
    double r[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const u64 i = (u64)k * 64 + lane;
        r[k] = i < n ? w(i) : 0.0;
    }
    double acc = 0.0;
    for (u64 c = 0; c < nchunk; ++c) {
        const u64 base = c << 9;
#pragma unroll
        for (int k = 0; k < 8; ++k) seg[k * 64 + lane] = r[k];
        __syncthreads();
        if (c + 1 < nchunk) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const u64 i = base + 512 + (u64)k * 64 + lane;
                r[k] = i < n ? w(i) : 0.0;
            }
        }
        if (base + 512 <= n) {
            // register double buffer of 32 operands: the LDS reads of the next 32 are issued before
            // the 32 dependent adds of the current ones (the scheduling barrier keeps that order)
            double cur[32], nxt[32];
#pragma unroll
            for (int t = 0; t < 32; ++t) cur[t] = seg[t];
#pragma unroll
            for (int h = 0; h < 16; ++h) {
                if (h + 1 < 16) {
#pragma unroll
                    for (int t = 0; t < 32; ++t) nxt[t] = seg[(h + 1) * 32 + t];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int t = 0; t < 32; ++t) acc = acc + cur[t];
                if (h & 1) ck[(base >> 6) + (h >> 1)] = acc;
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int t = 0; t < 32; ++t) cur[t] = nxt[t];
            }
        } else {
            const int cnt = (int)(n - base);
            for (int t = 0; t < cnt; ++t) {
                acc = acc + seg[t];
                if (((t + 1) & 63) == 0) ck[(base + t) >> 6] = acc;
            }
        }
        __syncthreads();
    }
    ck[n >> 6] = acc;
}

// searchsorted(cdf, u, side='right') on cdf = cumsum(w) / total with cdf[-1] forced to 1
// (sampling/sampling_protocol.py:56-65): binary search over the segment ends, then the segment's
// running sums in order.
template <class W>
__global__ void k_search(W w, const double *__restrict__ ck, u64 n, double u, DrawOut *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    DrawOut r;
    memset(&r, 0, sizeof(r));
    const u64 nseg = n >> 6;
    const double total = ck[nseg];
    r.total = total;
    if (!(total > 0.0) || !isfinite(total)) {
        r.status = isfinite(total) ? DCS_ERR_ZERO : DCS_ERR_NONFINITE;
        *out = r;
        return;
    }
    auto cdf_end = [&](u64 j) { return ((j << 6) + 63 == n - 1) ? 1.0 : ck[j] / total; };
    u64 lo = 0, hi = nseg;
    while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        if (cdf_end(mid) <= u) lo = mid + 1;
        else hi = mid;
    }
    const u64 s0 = lo << 6, s1 = (s0 + 64 < n) ? s0 + 64 : n;
    double acc = lo ? ck[lo - 1] : 0.0;
    double lower = lo ? acc / total : 0.0;
    u64 idx = n;
    for (u64 i = s0; i < s1; ++i) {
        acc = acc + w(i);
        const double c = (i == n - 1) ? 1.0 : acc / total;
        if (c > u) {
            idx = i;
            r.upper = c;
            break;
        }
        lower = c;
    }
    r.index = (int64_t)idx;
    if (idx >= n) {
        r.status = DCS_ERR_ZERO;
        *out = r;
        return;
    }
    r.weight = w(idx);
    r.lower = idx ? lower : 0.0;
    r.margin = fmin(u - r.lower, r.upper - u);
    r.status = (r.weight > 0.0) ? DCS_OK : DCS_ERR_ZERO;
    *out = r;
}

// Device scratch of dcs_draw_noalloc: k_chain's segment sums of a block of up to 2^DCS_DRAW_NOALLOC_MAX_K
// entries (2^(max - 6) + 1 doubles, 512 KiB), then one DrawOut, then the interval check's partial sums
// (8 KiB).  Every device the library runs on
// holds one copy for the life of the process; host code reaches it only through the address that
// hipGetSymbolAddress returned on that device (dcs_context::scratch).
constexpr u64 kNoallocSums = (1ull << (DCS_DRAW_NOALLOC_MAX_K - 6)) + 1;
constexpr u64 kNoallocDrawOut = kNoallocSums + 3;    // 32-byte aligned offset of the DrawOut (in doubles)
constexpr u64 kNoallocParts = kNoallocDrawOut + 16;  // the interval check's partial sums
static_assert(sizeof(DrawOut) <= 16 * sizeof(double), "DrawOut must fit its scratch slot");
static __device__ __align__(256) double scratch[kNoallocParts + 4 * kIntervalBlocks];

// Independent check of the drawn interval (draw_checked's math.fsum interval,
// sampling/sampling_protocol.py:66-72): the prefix mass below the drawn index and the total mass are
// summed again, in a different order and with compensation, and [prefix, prefix + w) / total must
// contain the uniform.  Each sum is Ogita, Rump and Oishi's Sum2 ("Accurate sum and dot product",
// SIAM J. Sci. Comput. 26, 2005): Knuth's error-free TwoSum keeps every rounding error, the errors
// are added up separately and folded in at the end, so the result is as accurate as a sum in twice
// the working precision (within one unit in the last place of the exactly rounded sum that math.fsum
// returns).  Partial sums of kIntervalBlocks workgroups are merged in a fixed order by one wave.

struct Sum2 {
    double s, c;  // value s + c
};

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    const double bv = s - a;
    e = (a - (s - bv)) + (b - bv);
}

__device__ __forceinline__ void sum2_add(Sum2 &a, double x) {
    double s, e;
    two_sum(a.s, x, s, e);
    a.s = s;
    a.c += e;
}

__device__ __forceinline__ Sum2 sum2_merge(const Sum2 &a, const Sum2 &b) {
    double s, e;
    two_sum(a.s, b.s, s, e);
    return Sum2{s, (a.c + b.c) + e};
}

__device__ __forceinline__ Sum2 sum2_wave(Sum2 v) {
    for (int off = 32; off > 0; off >>= 1) {
        const Sum2 o{__shfl_xor(v.s, off, 64), __shfl_xor(v.c, off, 64)};
        v = sum2_merge(v, o);
    }
    return v;
}

// part[4 b .. 4 b + 3] = prefix (s, c), total (s, c) of workgroup b; kIntervalBlocks x 256 threads.
template <class W>
__global__ void __launch_bounds__(256) k_interval_part(W w, u64 n, const DrawOut *__restrict__ d,
                                                       double *__restrict__ part) {
    __shared__ double red[4][4];
    const u64 end = d->status == DCS_OK ? (u64)d->index : 0;
    Sum2 pre{0.0, 0.0}, tot{0.0, 0.0};
    for (u64 i = (u64)blockIdx.x * 256 + threadIdx.x; i < n; i += (u64)256 * gridDim.x) {
        const double x = w(i);
        sum2_add(tot, x);
        if (i < end) sum2_add(pre, x);
    }
    pre = sum2_wave(pre);
    tot = sum2_wave(tot);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave][0] = pre.s;
        red[wave][1] = pre.c;
        red[wave][2] = tot.s;
        red[wave][3] = tot.c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Sum2 p{red[0][0], red[0][1]}, t{red[0][2], red[0][3]};
        for (int k = 1; k < 4; ++k) {
            p = sum2_merge(p, Sum2{red[k][0], red[k][1]});
            t = sum2_merge(t, Sum2{red[k][2], red[k][3]});
        }
        double *o = part + 4 * blockIdx.x;
        o[0] = p.s;
        o[1] = p.c;
        o[2] = t.s;
        o[3] = t.c;
    }
}

// Offset (in doubles) of the interval check's partial sums in a plan's draw scratch, after the n / 64 + 1
// segment sums of k_chain; and the scratch bytes a block of n entries needs.
inline u64 draw_parts_offset(u64 n) { return ((n >> 6) + 1 + 3) & ~3ull; }
inline u64 draw_scratch_bytes(u64 n) { return std::max<u64>(16 * n, 8 * (draw_parts_offset(n) + 4 * kIntervalBlocks)); }

// One wave: merge the partial sums and check the interval of the drawn index.
template <class W>
__global__ void __launch_bounds__(64) k_interval_check(W w, const double *__restrict__ part, double u,
                                                       DrawOut *__restrict__ d) {
    Sum2 pre{0.0, 0.0}, tot{0.0, 0.0};
    for (int b = threadIdx.x; b < kIntervalBlocks; b += 64) {
        pre = sum2_merge(pre, Sum2{part[4 * b], part[4 * b + 1]});
        tot = sum2_merge(tot, Sum2{part[4 * b + 2], part[4 * b + 3]});
    }
    pre = sum2_wave(pre);
    tot = sum2_wave(tot);
    if (threadIdx.x != 0 || d->status != DCS_OK) return;
    const double prefix = pre.s + pre.c, total = tot.s + tot.c;
    const double lower = prefix / total, upper = (prefix + w((u64)d->index)) / total;
    d->mass_check = total;
    d->lower_check = lower;
    d->upper_check = upper;
    if (!(lower <= u && u < upper)) d->status = kDrawIntervalDisagrees;
}

// Bandwidth ceiling probe: a copy with U independent 16-byte loads in flight per thread, issued before
// any of their stores, non-temporal (streamed past the caches), grid-stride over chunks of U * 256
// entries; coalesced (consecutive threads take consecutive 16-byte entries).  n is a multiple of U * 256.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int U>
__global__ void __launch_bounds__(256) k_stream_copy(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst, u64 n) {
    for (u64 base = (u64)blockIdx.x * (256 * U); base < n; base += (u64)gridDim.x * (256 * U)) {
        u32x4 v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) v[k] = __builtin_nontemporal_load(src + base + k * 256 + threadIdx.x);
#pragma unroll
        for (int k = 0; k < U; ++k) __builtin_nontemporal_store(v[k], dst + base + k * 256 + threadIdx.x);
    }
}

// ------------------------------------------------------------------------------------------------
// Host runtime.

struct Seg {
    int type;  // 0 step, 1 resident run, 2 coset group
    int first, count, dim, group, unit;
};

struct StepHost {
    int kind, pivot, pivot_a;
    u64 K, Ka, beta, in_mask, out_mask, slots;
};

inline int top_bit(u64 x) { return 63 - __builtin_clzll(x); }
inline int popc(u64 x) { return __builtin_popcountll(x); }

}  // namespace

#include "group_codegen.hpp"

namespace {
// Text of dcs_kernels.hpp, compiled into every generated module.
const char kKernelsHpp[] = {
#embed "dcs_kernels.hpp"
    , 0};
}  // namespace

struct dcs_context {
    int device;
    hipStream_t stream;
    int cus;
    void *scratch;      // this device's address of the static scratch (dcs_draw_noalloc)
    DrawOut *h_draw;    // pinned result of dcs_draw_noalloc
    // entries of dcs_contract_batch's launches: pinned staging and device copy, grown on demand
    BatchEntry *h_batch = nullptr, *d_batch = nullptr;
    size_t batch_cap = 0;
};

namespace {
// Device address of `scratch`, looked up once per device with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices] = {};

int scratch_address(int device, void **out) {
    if (device < 0 || device >= kMaxDevices) return fail(DCS_ERR_INVALID, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(DCS_ERR_RUNTIME, "no device address for the draw scratch on device %d", device);
        g_scratch_addr[device] = ptr;
    }
    *out = g_scratch_addr[device];
    return DCS_OK;
}

// Fill a dcs_draw_result from the device's DrawOut and map its status.
int draw_result(const DrawOut &r, dcs_draw_result *out) {
    out->index = r.index;
    out->weight = r.weight;
    out->total = r.total;
    out->cdf_lower = r.lower;
    out->cdf_upper = r.upper;
    out->margin = r.margin;
    out->mass_check = r.mass_check;
    out->cdf_lower_check = r.lower_check;
    out->cdf_upper_check = r.upper_check;
    if (r.status == kDrawIntervalDisagrees)
        return fail(DCS_ERR_ZERO, "independent compensated CDF interval [%.17g, %.17g) of index %lld does not contain "
                    "the uniform; the draw is refused, not repeated", r.lower_check, r.upper_check, (long long)r.index);
    if (r.status == DCS_ERR_NONFINITE) return fail(DCS_ERR_NONFINITE, "nonfinite block mass");
    if (r.status != DCS_OK) return fail(DCS_ERR_ZERO, "numerically zero block or invalid CDF outcome");
    return DCS_OK;
}
}  // namespace

struct PendingEvent {
    int family;
    hipEvent_t a, b;
};

struct dcs_plan {
    dcs_context *ctx;
    int storage, maxslots, lanes, nsteps, k, blocks;
    int fuse_max_dim, resident_bits, tile_log2, exact_rounding, pass_max_dim, lds_storage, group_policy, lds_buffers;
    size_t esz;
    u64 lane_stride;
    u64 terminal_mask;
    std::vector<StepHost> steps;
    std::vector<StepArgs> step_args;
    std::vector<double> unitaries;  // k*8
    std::vector<Seg> sched_steps, sched_fused;
    std::vector<u64> group_ntiles;
    // host copies of the fused-schedule tables (source of the generated kernels)
    std::vector<GroupDev> h_groups;
    std::vector<PassDev> h_passes;
    std::vector<PStepDev> h_psteps;
    std::vector<u64> h_addrv;
    // canonical gather executor (dcs_gather_plan_create): frontier holds the current table
    int executor = 0;  // 0 in-place, 1 gather
    u64 device_bytes = 0;
    std::vector<GatherArgs> gsteps;
    void *frontier2 = nullptr;
    u64 *d_gtables = nullptr, *d_gcombos = nullptr;
    // DCS_MODE_COMPILED
    bool compiled = false;
    std::vector<hipFunction_t> group_fn;
    std::vector<unsigned> group_grid;
    std::vector<int> group_threads;
    std::vector<int> group_pass_span;  // register-pass span each compiled group kernel was built with
    // f0 factoring of the group kernels: per-step coefficient classes (empty: off); when on, group_fn holds
    // the factored kernels and group_fn_plain the unfactored ones, compiled when a contraction's
    // parameters first do not fit the classes; call_factored says which set the staged call runs
    std::vector<int8_t> coef_class;
    bool factored_module = false;
    bool compiled_plain = false;
    std::vector<hipFunction_t> group_fn_plain;
    std::vector<unsigned> group_grid_plain;
    bool call_factored = false;
    int units_steps, units_fused;
    // device memory
    void *frontier = nullptr;
    Dyn *d_dyn = nullptr;
    int *d_lanes = nullptr;
    u64 *d_dmax = nullptr;
    StepDev *d_steps = nullptr;
    GroupDev *d_groups = nullptr;
    PassDev *d_passes = nullptr;
    PStepDev *d_psteps = nullptr;
    u64 *d_addrv = nullptr;
    u64 *d_smap = nullptr;
    double *d_phase = nullptr;
    double2 *d_blocks = nullptr;
    double *d_cum = nullptr;
    void *d_scratch = nullptr;
    DrawOut *d_draw = nullptr;
    // pinned staging
    Dyn *h_dyn = nullptr;
    int *h_lanes = nullptr;
    u64 *h_dmax = nullptr;
    DrawOut *h_draw = nullptr;
    // profiling
    bool profiling = false;
    double prof_ms[5] = {0, 0, 0, 0, 0};
    int64_t prof_n[5] = {0, 0, 0, 0, 0};
    std::vector<PendingEvent> pending;
    // captured launch sequences of dcs_contract, one per (mode, lane count, factored kernels or not)
    bool use_graphs = false;
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;
};

namespace {

int ev_begin(dcs_plan *p, int family, PendingEvent &pe) {
    if (!p->profiling) return DCS_OK;
    pe.family = family;
    HIPCHK(hipEventCreate(&pe.a));
    HIPCHK(hipEventCreate(&pe.b));
    HIPCHK(hipEventRecord(pe.a, p->ctx->stream));
    return DCS_OK;
}
int ev_end(dcs_plan *p, PendingEvent &pe) {
    if (!p->profiling) return DCS_OK;
    HIPCHK(hipEventRecord(pe.b, p->ctx->stream));
    p->pending.push_back(pe);
    return DCS_OK;
}
int ev_resolve(dcs_plan *p) {
    for (auto &pe : p->pending) {
        float ms = 0.f;
        HIPCHK(hipEventSynchronize(pe.b));
        HIPCHK(hipEventElapsedTime(&ms, pe.a, pe.b));
        p->prof_ms[pe.family] += ms;
        p->prof_n[pe.family] += 1;
        (void)hipEventDestroy(pe.a);
        (void)hipEventDestroy(pe.b);
    }
    p->pending.clear();
    return DCS_OK;
}

#define PROF(fam, launch)                              \
    do {                                               \
        PendingEvent pe_;                              \
        int rc_ = ev_begin(p, fam, pe_);               \
        if (rc_) return rc_;                           \
        launch;                                        \
        HIPCHK(hipGetLastError());                     \
        rc_ = ev_end(p, pe_);                          \
        if (rc_) return rc_;                           \
    } while (0)

size_t storage_bytes(int storage) {
    return storage == DCS_STORAGE_C64 ? sizeof(float2) : storage == DCS_STORAGE_C128 ? sizeof(double2) : sizeof(h16x2);
}

unsigned grid_for(u64 work, int cus, unsigned per_cu) {
    const u64 blocks = (work + kThreads - 1) / kThreads;
    const u64 cap = (u64)cus * per_cu;
    return (unsigned)std::max<u64>(1, std::min(blocks, cap));
}

u64 orbit_slots(const StepHost &s) {
    u64 m = s.in_mask;
    if (s.kind == 1) m &= ~(1ull << s.pivot_a);
    else if (s.kind == 3 || s.kind == 5 || s.kind == 7) m &= ~(1ull << s.pivot);
    else if (s.kind == 6) m &= ~((1ull << s.pivot) | (1ull << s.pivot_a));
    return m;
}

std::vector<u64> step_masks(const StepHost &s) {
    switch (s.kind) {
        case 0: return {1ull << s.pivot};
        case 1: return {s.Ka};
        case 3: case 5: case 7: return {s.K};
        case 6: return {s.K, s.Ka};
        default: return {};
    }
}

// Insert v into an RREF basis keyed by highest bit (pivots distinct, every other vector zero there).
bool rref_insert(std::vector<u64> &basis, u64 v) {
    for (u64 b : basis)
        if ((v >> top_bit(b)) & 1) v ^= b;
    if (!v) return false;
    const int p = top_bit(v);
    for (u64 &b : basis)
        if ((b >> p) & 1) b ^= v;
    basis.push_back(v);
    std::sort(basis.begin(), basis.end(), [](u64 a, u64 b) { return top_bit(a) < top_bit(b); });
    return true;
}

bool step_is_resident(const StepHost &s, int rb) {
    if (rb <= 0) return false;
    const u64 lim = (rb >= 64) ? ~0ull : ((1ull << rb) - 1);
    return !(s.in_mask & ~lim) && !(s.out_mask & ~lim);
}

// Ends of the coset groups that may start at step s: every e such that steps [s, e) span an orbit-mask
// subspace of dimension <= D (a single step always qualifies), every slot grown inside the run is a
// pivot of that subspace, and every slot dead at the run's start or end is one too (the coset
// kernels zero-fill / skip exactly those coordinates).
std::vector<int> group_ends_from(const dcs_plan *p, int s, int D, std::vector<int> *dims = nullptr) {
    std::vector<int> ends;
    std::vector<u64> V;
    u64 Aun = p->steps[s].in_mask;
    for (int e = s; e < p->nsteps && !step_is_resident(p->steps[e], p->resident_bits); ++e) {
        for (u64 m : step_masks(p->steps[e])) rref_insert(V, m);
        if ((int)V.size() > D && e > s) break;
        if (p->steps[e].kind == 0) Aun |= 1ull << p->steps[e].pivot;
        u64 piv = 0;
        for (u64 v : V) piv |= 1ull << top_bit(v);
        const u64 dead = (Aun & ~p->steps[s].in_mask) | (Aun & ~p->steps[e].out_mask);
        if (!(piv & ~Aun) && !(dead & ~piv)) {
            ends.push_back(e + 1);
            if (dims) dims->push_back((int)V.size());
        }
    }
    return ends;
}

// Codegen options of the plan's group kernels; with factored, the f0 factoring of the plan's coefficient
// classes where the schedule allows it.
CodegenOptions codegen_options(const dcs_plan *p, bool factored) {
    CodegenOptions opt{p->storage, p->tile_log2, p->exact_rounding != 0, p->lds_storage != 0, p->lds_buffers};
    if (factored && !p->coef_class.empty()) opt.cls = p->coef_class.data();
    return opt;
}

// Whether the plan's compiled schedule runs f0-factored group kernels by default.
bool factoring_on(const dcs_plan *p) { return codegen_options(p, true).factored(); }

// Tile size of a group of coset dimension d: 2^TL entries, except that a group wider than the generic
// kernels take (only generated kernels run it) keeps rows of at least 16 entries (128 bytes in complex64)
// with tiles of up to 2^13 entries when its LDS tiles hold 8 bytes an entry (complex64 or binary16
// storage in LDS: two such tiles take 128 of the 160 KiB); 64-byte rows cost more than the larger tile
// (dimension 9 on 2^12 tiles: 405 ms of group kernels per 64-qubit sample against 358 ms with 2^13
// tiles for those groups).
int group_tile_log2(const dcs_plan *p, int d) {
    const int TL = p->tile_log2;
    const bool lds8 = p->lds_storage != 0 && p->storage != DCS_STORAGE_C128;
    return (d > kMaxDim && TL < 13 && lds8) ? std::max(TL, std::min(13, d + 4)) : TL;
}

// HBM cost factor of a group on a wider tile than the plan's (one 1024-thread workgroup per CU instead of
// two of 512): dimension 9 on 2^13 tiles moved 6% fewer bytes in the same time as dimension 8 on 2^12
// tiles on the 64-qubit plan; 1.05 and 1.1 both give 350 ms against 354 ms for 1.0.
constexpr double kWideTileCost = 1.1;

// Group boundaries over the non-resident runs of the plan.  Policy 0: greedy maximal groups (extend
// while the span stays <= D).  Policy 1: the partition with the fewest HBM bytes, by dynamic
// programming over group ends: a group reads its live entries once and writes them once, so its cost
// is (2^|live before| + 2^|live after|) entries plus a fixed launch allowance of 2^20 entries.
// Policy 2: the same dynamic programme with the register passes priced too: a group runs each of its
// steps over every entry of its tile span, live or not, 2^|slots live anywhere in the group| per step,
// weighted 0.31 against 1.73 per entry moved (the measured cost of an element-step and of an HBM entry
// of the generated kernels on MI355X, from compute-only and memory-only variants).
void choose_groups(const dcs_plan *p, int D, std::vector<int> &group_end) {
    const int n = p->nsteps;
    group_end.assign(n, -1);
    int s = 0;
    while (s < n) {
        if (step_is_resident(p->steps[s], p->resident_bits)) {
            ++s;
            continue;
        }
        int b = s;
        while (b < n && !step_is_resident(p->steps[b], p->resident_bits)) ++b;
        auto greedy = [&]() {
            for (int i = s; i < b;) {
                std::vector<u64> V;
                int e = i + 1;
                for (int j = i; j < b; ++j) {
                    for (u64 m : step_masks(p->steps[j])) rref_insert(V, m);
                    if ((int)V.size() > D && j > i) break;
                    e = j + 1;
                }
                group_end[i] = e;
                i = e;
            }
        };
        if (p->group_policy == 0) {
            greedy();
        } else {
            const double inf = 1e300;
            std::vector<double> best(b - s + 1, inf);
            std::vector<int> from(b - s + 1, -1);
            best[0] = 0;
            auto live = [&](u64 m) { return std::ldexp(1.0, popc(m)); };
            // policy 1: HBM entries moved (+ a launch allowance of 2^20 entries); 2: entries at 1.73
            // and element-steps over the group's span at 0.31 (measured costs, summed); 3: the same two
            // times overlapped, max(memory, compute) + a quarter of the smaller one (a group's loads and
            // stores run during its passes, three quarters of the shorter phase hidden as measured).
            // f0-factored kernels issue 0.61 times the VALU instructions per element-step (7.1 against
            // 11.7 lane instructions, SQ_INSTS_VALU on the 64-qubit plan), so their steps cost 0.19;
            // their compute-only and memory-only variants give 0.16, and 0.16, 0.19 and 0.25 all measure
            // within 0.5% of each other on that plan, 0.19 the fastest.
            const double w_mem = p->group_policy >= 2 ? 1.73 : 1.0;
            const double w_cmp = p->group_policy >= 2 ? (factoring_on(p) ? 0.19 : 0.31) : 0.0;
            const bool overlap = p->group_policy == 3;
            for (int i = s; i < b; ++i) {
                if (best[i - s] >= inf) continue;
                u64 span = p->steps[i].in_mask;
                int last = i;
                std::vector<int> dims;
                const std::vector<int> ends = group_ends_from(p, i, D, &dims);
                for (size_t k = 0; k < ends.size(); ++k) {
                    const int e = ends[k];
                    for (; last < e; ++last) span |= p->steps[last].out_mask;
                    const double wm = group_tile_log2(p, dims[k]) > p->tile_log2 ? w_mem * kWideTileCost : w_mem;
                    const double tm = wm * (live(p->steps[i].in_mask) + live(p->steps[e - 1].out_mask));
                    const double tc = w_cmp * (double)(e - i) * live(span);
                    const double c = best[i - s] + w_mem * std::ldexp(1.0, 20) +
                                     (overlap ? std::max(tm, tc) + 0.25 * std::min(tm, tc) : tm + tc);
                    if (c < best[e - s]) {
                        best[e - s] = c;
                        from[e - s] = i;
                    }
                }
            }
            if (best[b - s] >= inf) {
                greedy();  // no admissible partition: the greedy groups report the offending step
            } else {
                for (int e = b; e > s;) {
                    const int i = from[e - s];
                    group_end[i] = e;
                    e = i;
                }
            }
        }
        s = b;
    }
}

// pass_span > 0 forms the register passes with that span instead of the plan's pass_max_dim (the
// schedules and groups do not depend on it).
int build_schedules(dcs_plan *p, std::vector<GroupDev> &groups, std::vector<PassDev> &passes,
                    std::vector<PStepDev> &psteps, std::vector<u64> &addrv, int pass_span = 0) {
    const int n = p->nsteps;
    const int JP = pass_span > 0 ? pass_span : p->pass_max_dim;
    p->sched_steps.clear();
    for (int s = 0; s < n; ++s) p->sched_steps.push_back(Seg{0, s, 1, 0, -1, s});
    p->units_steps = n;

    p->sched_fused.clear();
    int unit = 0;
    const int D = std::max(1, std::min(kMaxDimCompiled, p->fuse_max_dim));
    const int TL = p->tile_log2;
    std::vector<int> group_end;
    if (p->fuse_max_dim > 0) choose_groups(p, D, group_end);
    int s = 0;
    while (s < n) {
        if (step_is_resident(p->steps[s], p->resident_bits)) {
            int e = s;
            while (e < n && step_is_resident(p->steps[e], p->resident_bits)) ++e;
            p->sched_fused.push_back(Seg{1, s, e - s, 0, -1, unit});
            unit += e - s;
            s = e;
            continue;
        }
        if (p->fuse_max_dim <= 0) {
            p->sched_fused.push_back(Seg{0, s, 1, 0, -1, unit});
            ++unit;
            ++s;
            continue;
        }
        const int e = group_end[s];
        if (e <= s) return fail(DCS_ERR_INVALID, "no group boundary chosen at step %d", s);
        std::vector<u64> V;
        u64 Aun = p->steps[s].in_mask;
        for (int j = s; j < e; ++j) {
            for (u64 m : step_masks(p->steps[j])) rref_insert(V, m);
            if (p->steps[j].kind == 0) Aun |= 1ull << p->steps[j].pivot;
        }
        const int d = (int)V.size();
        u64 piv = 0;
        for (u64 v : V) piv |= 1ull << top_bit(v);
        if (piv & ~Aun) return fail(DCS_ERR_INVALID, "group pivots outside live slots at step %d", s);
        // Coordinate order: pivots whose slot is live throughout the group first, those that die or are
        // born inside it last, so that the latter map to the high orbit-index bits (uniform per
        // wavefront) when a register pass does not hold them.
        {
            u64 always = ~0ull;
            for (int j = s; j < e; ++j) always &= p->steps[j].in_mask & p->steps[j].out_mask;
            std::stable_partition(V.begin(), V.end(), [&](u64 v) { return (always >> top_bit(v)) & 1; });
        }
        const u64 R = Aun & ~piv;
        const int rbits = popc(R);
        int wb = std::min(group_tile_log2(p, d) - d, rbits);
        if (wb < 0) wb = 0;
        u64 low = 0;
        {
            u64 r = R;
            for (int j = 0; j < wb; ++j) {
                low |= r & (~r + 1);
                r &= r - 1;
            }
        }
        GroupDev g;
        memset(&g, 0, sizeof(g));
        g.d = d;
        g.wbits = wb;
        g.lane_identity = (low == ((wb == 64) ? ~0ull : ((1ull << wb) - 1)));
        g.addrv_offset = (int)addrv.size();
        g.rep_low_mask = low;
        g.ntiles = 1ull << (rbits - wb);
        g.high = make_runs(R & ~low);
        auto tcoords = [&](u64 m, uint32_t &out) -> bool {
            uint32_t c = 0;
            u64 chk = 0;
            for (int i = 0; i < d; ++i)
                if ((m >> top_bit(V[i])) & 1) {
                    c |= 1u << i;
                    chk ^= V[i];
                }
            out = c;
            return chk == m;
        };
        auto tindex = [&](int slot) -> int {
            for (int i = 0; i < d; ++i)
                if (top_bit(V[i]) == slot) return i;
            return -1;
        };
        const u64 dead_in = Aun & ~p->steps[s].in_mask;
        const u64 dead_out = Aun & ~p->steps[e - 1].out_mask;
        for (int i = 0; i < d; ++i) {
            const int slot = top_bit(V[i]);
            if ((dead_in >> slot) & 1) g.dead_in_t |= 1u << i;
            if ((dead_out >> slot) & 1) g.dead_out_t |= 1u << i;
        }
        if ((dead_in | dead_out) & ~piv)
            return fail(DCS_ERR_INVALID, "group %d: a dead slot is not a coset pivot", (int)groups.size());
        for (uint32_t t = 0; t < (1u << d); ++t) {
            u64 a = 0;
            for (int i = 0; i < d; ++i)
                if ((t >> i) & 1) a ^= V[i];
            addrv.push_back(a);
        }
        // Split the group into register passes: consecutive steps whose t-space masks stay linearly
        // independent, at most three masks per pass.
        // m: t-space masks of the step (K or Ka); r: t-coordinate of each mask's slot pivot, the bit that
        // is zero on the element playing "c".  With the always-live-first coordinate order this need not
        // be the mask's highest t-bit.
        struct TStep {
            int j;
            uint32_t m[2];
            int r[2];
            int nm;
        };
        std::vector<TStep> ts;
        for (int j = s; j < e; ++j) {
            const StepHost &st = p->steps[j];
            TStep t{j, {0u, 0u}, {-1, -1}, 0};
            uint32_t c = 0;
            auto add_mask = [&](uint32_t m, int slot) -> bool {
                const int r = tindex(slot);
                if (r < 0 || !((m >> r) & 1)) return false;
                t.m[t.nm] = m;
                t.r[t.nm++] = r;
                return true;
            };
            if (st.kind == 0) {
                const int i = tindex(st.pivot);
                if (i < 0 || !add_mask(1u << i, st.pivot))
                    return fail(DCS_ERR_INVALID, "step %d: grow slot outside its group subspace", j);
            } else if (st.kind == 1) {
                if (!tcoords(st.Ka, c) || !add_mask(c, st.pivot_a))
                    return fail(DCS_ERR_INVALID, "step %d: Ka outside its group subspace", j);
            } else if (st.kind == 3 || st.kind == 5 || st.kind == 6 || st.kind == 7) {
                if (!tcoords(st.K, c) || !add_mask(c, st.pivot))
                    return fail(DCS_ERR_INVALID, "step %d: K outside its group subspace", j);
                if (st.kind == 6) {
                    if (!tcoords(st.Ka, c) || !add_mask(c, st.pivot_a))
                        return fail(DCS_ERR_INVALID, "step %d: Ka outside its group subspace", j);
                }
            }
            ts.push_back(t);
        }
        auto top32 = [](uint32_t x) { return 31 - __builtin_clz(x); };
        // Pass state: basis vectors (insertion order) and a fully reduced copy in which every reduced
        // vector owns a pivot bit that no other reduced vector has, remembering which basis vectors it
        // combines.  Orbit representatives are zero on the pivots, so a step's role coordinate (the
        // t-bit that is zero on its "c" element) must be a pivot for the role to be a compile-time
        // property of the element index: pivots are chosen as the role coordinates.
        struct Red {
            uint32_t vec, combo;
            int piv;
        };
        struct PassState {
            std::vector<uint32_t> basis;
            std::vector<Red> red;
            bool is_pivot(int b) const {
                for (const Red &rc : red)
                    if (rc.piv == b) return true;
                return false;
            }
        };
        auto reduce = [&](const PassState &P, uint32_t v, uint32_t &combo) -> uint32_t {
            combo = 0;
            for (const Red &rc : P.red)
                if ((v >> rc.piv) & 1) {
                    v ^= rc.vec;
                    combo ^= rc.combo;
                }
            return v;
        };
        // Express mask m (role coordinate r) in the pass basis, adding it as a new basis vector when it
        // is outside the span; false when r cannot be a pivot or the span would exceed pass_max_dim.
        auto express = [&](PassState &P, uint32_t m, int r, uint32_t &c) -> bool {
            const uint32_t rem = reduce(P, m, c);
            if (!rem) return P.is_pivot(r);
            if ((int)P.basis.size() >= JP) return false;
            int piv;
            if ((rem >> r) & 1) piv = r;
            else if (P.is_pivot(r)) piv = top32(rem);
            else return false;
            const int idx = (int)P.basis.size();
            const uint32_t combo = c | (1u << idx);
            for (Red &rc : P.red)
                if ((rc.vec >> piv) & 1) {
                    rc.vec ^= rem;
                    rc.combo ^= combo;
                }
            P.basis.push_back(m);
            P.red.push_back({rem, combo, piv});
            c = 1u << idx;
            return true;
        };
        // Place one step into a pass state; returns false when it does not fit (span > pass_max_dim, a
        // role coordinate that cannot be a pivot, or a quad whose K / Ka are not single basis vectors).
        auto place = [&](PassState &P, const TStep &t, uint32_t &pat, uint32_t &pat2) -> bool {
            pat = pat2 = 0;
            for (int q = 0; q < t.nm; ++q) {
                uint32_t c = 0;
                if (!express(P, t.m[q], t.r[q], c)) return false;
                if (t.nm == 2 && __builtin_popcount(c) != 1) return false;
                (q == 0 ? pat : pat2) = c;
            }
            return true;
        };
        g.first_pass = (int)passes.size();
        size_t k0 = 0;
        while (k0 < ts.size()) {
            PassState P;
            std::vector<std::pair<uint32_t, uint32_t>> pats;
            size_t k1 = k0;
            while (k1 < ts.size()) {
                PassState trial = P;
                uint32_t pat = 0, pat2 = 0;
                if (!place(trial, ts[k1], pat, pat2)) {
                    if (k1 == k0) return fail(DCS_ERR_INVALID, "step %d: orbit masks do not fit a pass", ts[k1].j);
                    break;
                }
                P = trial;
                pats.push_back({pat, pat2});
                ++k1;
            }
            PassDev ps;
            memset(&ps, 0, sizeof(ps));
            ps.nmask = (int)P.basis.size();
            ps.nsteps = (int)(k1 - k0);
            ps.first_pstep = (int)psteps.size();
            std::vector<int> piv;
            for (const Red &rc : P.red) piv.push_back(rc.piv);
            std::sort(piv.begin(), piv.end());
            for (size_t i = 0; i < P.basis.size(); ++i) {
                ps.piv[i] = piv[i];
                ps.mw[i] = P.basis[i];
            }
            auto rolemask = [&](int bit) {
                uint32_t m = 0;
                for (size_t l = 0; l < P.basis.size(); ++l)
                    if ((P.basis[l] >> bit) & 1) m |= 1u << l;
                return m;
            };
            for (size_t k = k0; k < k1; ++k) {
                const StepHost &st = p->steps[ts[k].j];
                PStepDev q;
                memset(&q, 0, sizeof(q));
                q.beta = st.beta;
                q.kind = st.kind;
                q.step = ts[k].j;
                q.pat = pats[k - k0].first;
                q.pat2 = pats[k - k0].second;
                for (int i = 0; i < d; ++i)
                    if (popc(V[i] & st.beta) & 1) q.bvec |= 1u << i;
                for (size_t l = 0; l < P.basis.size(); ++l)
                    if (__builtin_popcount(P.basis[l] & q.bvec) & 1) q.lbv |= 1u << l;
                if (ts[k].nm >= 1) q.pm = rolemask(ts[k].r[0]);
                if (ts[k].nm == 2) {
                    q.pma = rolemask(ts[k].r[1]);
                    if ((ts[k].m[1] >> ts[k].r[0]) & 1)
                        return fail(DCS_ERR_INVALID, "step %d: Ka holds the K pivot", ts[k].j);
                }
                {
                    u64 r = low;
                    for (int jb = 0; jb < wb; ++jb) {
                        const u64 bit = r & (~r + 1);
                        if (st.beta & bit) q.beta_lane |= 1u << jb;
                        r &= r - 1;
                    }
                }
                for (int i = 0; i < d; ++i)
                    if (!((st.in_mask >> top_bit(V[i])) & 1)) q.dead_t |= 1u << i;
                psteps.push_back(q);
            }
            passes.push_back(ps);
            k0 = k1;
        }
        g.npasses = (int)passes.size() - g.first_pass;
        p->sched_fused.push_back(Seg{2, s, e - s, d, (int)groups.size(), unit});
        groups.push_back(g);
        ++unit;
        s = e;
    }
    p->units_fused = unit;
    return DCS_OK;
}

template <class S>
int launch_step(dcs_plan *p, int nl, int step, int nunits, int unit) {
    const StepArgs &a = p->step_args[step];
    dim3 grid(grid_for(a.norbits, p->ctx->cus, 32), nl);
    S *E = (S *)p->frontier;
#define L(K)                                                                                       \
    case K:                                                                                        \
        PROF(0, (k_step<S, K><<<grid, kThreads, 0, p->ctx->stream>>>(E, p->lane_stride, p->d_lanes, a, p->d_dyn, \
                                                                     p->nsteps, step, p->d_dmax, nunits, unit))); \
        break;
    switch (a.kind) {
        L(0) L(1) L(2) L(3) L(4) L(5) L(6) L(7)
        default: return fail(DCS_ERR_INVALID, "bad kind");
    }
#undef L
    return DCS_OK;
}

template <class S>
int launch_resident(dcs_plan *p, int nl, const Seg &sg, int nunits) {
    S *E = (S *)p->frontier;
    const int rb = p->resident_bits;
#define R(B)                                                                                         \
    case B:                                                                                          \
        PROF(1, (k_resident<S, B><<<nl, kThreads, 0, p->ctx->stream>>>(E, p->lane_stride, p->d_lanes, p->d_steps, \
                                                                       sg.first, sg.first + sg.count, p->d_dyn, \
                                                                       p->nsteps, p->d_dmax, nunits, sg.unit))); \
        break;
    switch (rb) {
        R(6) R(7) R(8) R(9) R(10) R(11) R(12)
        default: return fail(DCS_ERR_INVALID, "resident_bits must be 6..12");
    }
#undef R
    return DCS_OK;
}

template <class S>
int launch_coset(dcs_plan *p, int nl, const Seg &sg, int nunits, u64 ntiles) {
    if (sg.dim > kMaxDim)
        return fail(DCS_ERR_INVALID, "coset dimension %d exceeds %d: use DCS_MODE_COMPILED or a smaller fuse_max_dim", sg.dim,
                    kMaxDim);
    S *E = (S *)p->frontier;
    dim3 grid((unsigned)std::max<u64>(1, std::min<u64>(ntiles, (u64)p->ctx->cus * 8)), nl);
#define C(T, X, JM)                                                                                     \
    PROF(2, (k_coset<S, T, X, JM><<<grid, kThreads, 0, p->ctx->stream>>>(E, p->lane_stride, p->d_lanes,      \
                                                                         p->d_groups, sg.group, p->d_passes, \
                                                                         p->d_psteps, p->d_addrv, p->d_dyn,  \
                                                                         p->nsteps, p->d_dmax, nunits, sg.unit)))
    const bool x = p->exact_rounding != 0;
    const int key = p->tile_log2 * 100 + (x ? 10 : 0) + p->pass_max_dim;
    switch (key) {
        case 1203: C(12, false, 3); break;
        case 1204: C(12, false, 4); break;
        case 1213: C(12, true, 3); break;
        case 1214: C(12, true, 4); break;
        case 1103: C(11, false, 3); break;
        case 1104: C(11, false, 4); break;
        case 1113: C(11, true, 3); break;
        case 1114: C(11, true, 4); break;
        default: return fail(DCS_ERR_INVALID, "unsupported tile_log2 %d / pass_max_dim %d", p->tile_log2, p->pass_max_dim);
    }
#undef C
    return DCS_OK;
}

// ---- DCS_MODE_COMPILED: generated group kernels -------------------------------------------------
struct ModuleCache {
    std::mutex mu;
    std::map<std::pair<int, std::string>, hipModule_t> mods;  // (device, source) -> loaded module
};
ModuleCache &module_cache() {
    static ModuleCache c;
    return c;
}

// Source of the module holding one generated kernel per fused group of the plan's schedule.
std::string group_module_source(const dcs_plan *p, const std::vector<GroupDev> &groups, const std::vector<PassDev> &passes,
                                const std::vector<PStepDev> &psteps, const std::vector<u64> &addrv, bool factored) {
    const CodegenOptions opt = codegen_options(p, factored);
    std::string src = codegen::prelude(opt, kKernelsHpp);
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        const GroupDev &g = groups[gi];
        codegen::GroupView v{&g, passes.data() + g.first_pass, psteps.data(), addrv.data() + g.addrv_offset};
        src += codegen::group_kernel(opt, (int)gi, v);
    }
    return src;
}

// Compile a module source for the plan's device with hipRTC, or find it in the module cache.
int load_module(dcs_plan *p, const std::string &src, hipModule_t &mod, double *seconds) {
    HIPCHK(hipSetDevice(p->ctx->device));
    mod = nullptr;
    {
        ModuleCache &mc = module_cache();
        std::lock_guard<std::mutex> lk(mc.mu);
        auto key = std::make_pair(p->ctx->device, src);
        auto it = mc.mods.find(key);
        if (it != mc.mods.end()) {
            mod = it->second;
        } else {
            const auto t0 = std::chrono::steady_clock::now();
            hiprtcProgram prog;
            if (hiprtcCreateProgram(&prog, src.c_str(), "dcs_groups.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
                return fail(DCS_ERR_RUNTIME, "hiprtcCreateProgram failed");
            const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                                  "-Wno-pragma-once-outside-header", "-Xclang", "-target-feature", "-Xclang",
                                  "-packed-fp32-ops"};
            const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
            if (rc != HIPRTC_SUCCESS) {
                size_t ls = 0;
                hiprtcGetProgramLogSize(prog, &ls);
                std::string log(ls, '\0');
                if (ls) hiprtcGetProgramLog(prog, &log[0]);
                hiprtcDestroyProgram(&prog);
                return fail(DCS_ERR_UNSUPPORTED, "hipRTC compile of the group kernels failed: %.900s", log.c_str());
            }
            size_t cs = 0;
            hiprtcGetCodeSize(prog, &cs);
            std::vector<char> code(cs);
            hiprtcGetCode(prog, code.data());
            hiprtcDestroyProgram(&prog);
            HIPCHK(hipModuleLoadData(&mod, code.data()));
            mc.mods.emplace(std::move(key), mod);
            if (seconds) *seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    }
    return DCS_OK;
}

// Compile (or find in the module cache) the plan's group kernels, factored or not, and size their
// persistent grids from their occupancy.  With register passes of span 4, a kernel that spills
// registers to scratch memory is replaced by the same group's kernel with passes of span 3 (a second
// module; the groups, their tiles and their launches are the same): on the 64-qubit plan span 4 is 1.8%
// faster over all, but a spilling kernel ran up to 27% slower than its span-3 version.
int compile_module(dcs_plan *p, bool factored, std::vector<hipFunction_t> &fn, std::vector<unsigned> &grid,
                   double *seconds) {
    const std::string src = group_module_source(p, p->h_groups, p->h_passes, p->h_psteps, p->h_addrv, factored);
    hipModule_t mod = nullptr;
    {
        const int rc = load_module(p, src, mod, seconds);
        if (rc) return rc;
    }
    fn.assign(p->h_groups.size(), nullptr);
    grid.assign(p->h_groups.size(), 1);
    p->group_threads.assign(p->h_groups.size(), kThreads);
    std::vector<size_t> spilled;
    for (size_t gi = 0; gi < p->h_groups.size(); ++gi) {
        const std::string name = "dcs_group_" + std::to_string(gi);
        HIPCHK(hipModuleGetFunction(&fn[gi], mod, name.c_str()));
        int scratch = 0;
        HIPCHK(hipFuncGetAttribute(&scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, fn[gi]));
        if (scratch > 0) spilled.push_back(gi);
    }
    p->group_pass_span.assign(p->h_groups.size(), p->pass_max_dim);
    if (!spilled.empty() && p->pass_max_dim > 3) {
        std::vector<GroupDev> groups;
        std::vector<PassDev> passes;
        std::vector<PStepDev> psteps;
        std::vector<u64> addrv;
        int rc = build_schedules(p, groups, passes, psteps, addrv, 3);
        if (rc) return rc;
        if (groups.size() != p->h_groups.size()) return fail(DCS_ERR_RUNTIME, "span-3 schedule has other groups");
        hipModule_t mod3 = nullptr;
        rc = load_module(p, group_module_source(p, groups, passes, psteps, addrv, factored), mod3, seconds);
        if (rc) return rc;
        for (size_t gi : spilled) {
            const std::string name = "dcs_group_" + std::to_string(gi);
            HIPCHK(hipModuleGetFunction(&fn[gi], mod3, name.c_str()));
            p->group_pass_span[gi] = 3;
        }
    }
    for (size_t gi = 0; gi < p->h_groups.size(); ++gi) {
        p->group_threads[gi] = codegen::group_threads(p->h_groups[gi]);
        int nb = 0;
        HIPCHK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn[gi], p->group_threads[gi], 0));
        if (nb < 1) return fail(DCS_ERR_UNSUPPORTED, "group kernel %zu cannot be resident (registers or LDS)", gi);
        const u64 cap = (u64)p->ctx->cus * (u64)std::max(1, nb);
        grid[gi] = (unsigned)std::max<u64>(1, std::min<u64>(p->h_groups[gi].ntiles, cap));
    }
    return DCS_OK;
}

int compile_groups(dcs_plan *p, double *seconds) {
    if (seconds) *seconds = 0.0;
    if (p->executor != 0) return fail(DCS_ERR_INVALID, "a gather plan has no coset groups to compile");
    if (p->compiled) return DCS_OK;
    p->factored_module = factoring_on(p);
    const int rc = compile_module(p, p->factored_module, p->group_fn, p->group_grid, seconds);
    if (rc) return rc;
    p->compiled = true;
    return DCS_OK;
}

// The unfactored group kernels of a plan whose default kernels are f0-factored (compiled on first use).
int compile_plain(dcs_plan *p) {
    if (p->compiled_plain) return DCS_OK;
    const int rc = compile_module(p, false, p->group_fn_plain, p->group_grid_plain, nullptr);
    if (rc) return rc;
    p->compiled_plain = true;
    return DCS_OK;
}

template <class S>
int launch_compiled(dcs_plan *p, int nl, const Seg &sg, int nunits) {
    S *E = (S *)p->frontier;
    u64 ls = p->lane_stride;
    const int *lanes = p->d_lanes;
    const Dyn *dyn = p->d_dyn;
    int nsteps = p->nsteps, nu = nunits, unit = sg.unit;
    u64 *dmax = p->d_dmax;
    void *args[] = {&E, &ls, &lanes, &dyn, &nsteps, &dmax, &nu, &unit};
    const bool fac = !p->factored_module || p->call_factored;
    PROF(2, HIPCHK(hipModuleLaunchKernel(fac ? p->group_fn[sg.group] : p->group_fn_plain[sg.group],
                                         fac ? p->group_grid[sg.group] : p->group_grid_plain[sg.group], (unsigned)nl, 1,
                                         (unsigned)p->group_threads[sg.group], 1, 1, 0,
                                         p->ctx->stream, args, nullptr)));
    return DCS_OK;
}

template <class S>
int run_gather(dcs_plan *p, int mode) {
    const int n = p->nsteps;
    k_fill_u64<<<grid_for(n, p->ctx->cus, 4), kThreads, 0, p->ctx->stream>>>(p->d_dmax, n, 0ull);
    HIPCHK(hipGetLastError());
    k_init_lanes<S><<<1, 64, 0, p->ctx->stream>>>((S *)p->frontier, p->lane_stride, p->d_lanes, 1);
    HIPCHK(hipGetLastError());
    for (int s = 0; s < n; ++s) {
        const GatherArgs &a = p->gsteps[s];
        const unsigned g = grid_for(a.n_out, p->ctx->cus, 32);
        PROF(0, (k_gather<S><<<g, kThreads, 0, p->ctx->stream>>>((const S *)p->frontier, (S *)p->frontier2, a, p->d_dyn, s,
                                                                 p->d_dmax, s)));
        std::swap(p->frontier, p->frontier2);
        if (mode == DCS_MODE_STEPS_HOST) {
            // the reference reads the maximum back after every step (conditional_block_gpu.py:118-122)
            HIPCHK(hipMemcpyAsync(p->h_dmax, p->d_dmax + s, sizeof(u64), hipMemcpyDeviceToHost, p->ctx->stream));
            HIPCHK(hipStreamSynchronize(p->ctx->stream));
            double m2;
            memcpy(&m2, p->h_dmax, 8);
            if (!std::isfinite(m2) || m2 == 0.0) {
                // leave the table that holds the failing step's output in place, as the reference does
                break;
            }
        }
    }
    return DCS_OK;
}

template <class S>
int run_contract(dcs_plan *p, int mode, int nl, std::vector<u64> &ntiles_of_group) {
    const bool fused = mode == DCS_MODE_FUSED || mode == DCS_MODE_COMPILED;
    const std::vector<Seg> &sched = fused ? p->sched_fused : p->sched_steps;
    const int nunits = fused ? p->units_fused : p->units_steps;
    {
        const u64 cnt = (u64)nl * nunits;
        k_fill_u64<<<grid_for(cnt, p->ctx->cus, 4), kThreads, 0, p->ctx->stream>>>(p->d_dmax, cnt, 0ull);
        HIPCHK(hipGetLastError());
        k_init_lanes<S><<<1, 64, 0, p->ctx->stream>>>((S *)p->frontier, p->lane_stride, p->d_lanes, nl);
        HIPCHK(hipGetLastError());
    }
    for (const Seg &sg : sched) {
        int rc = DCS_OK;
        if (sg.type == 0) rc = launch_step<S>(p, nl, sg.first, nunits, sg.unit);
        else if (sg.type == 1) rc = launch_resident<S>(p, nl, sg, nunits);
        else if (mode == DCS_MODE_COMPILED) rc = launch_compiled<S>(p, nl, sg, nunits);
        else rc = launch_coset<S>(p, nl, sg, nunits, ntiles_of_group[sg.group]);
        if (rc) return rc;
        if (mode == DCS_MODE_STEPS_HOST) {
            // The reference driver reads the maximum back after every step (inplace_gpu.py:120-123).
            HIPCHK(hipMemcpyAsync(p->h_dmax, p->d_dmax, sizeof(u64) * (size_t)nl * nunits, hipMemcpyDeviceToHost,
                                  p->ctx->stream));
            HIPCHK(hipStreamSynchronize(p->ctx->stream));
            bool all_zero = true, bad = false;
            for (int li = 0; li < nl; ++li) {
                double m2;
                const u64 b = p->h_dmax[(size_t)li * nunits + sg.unit];
                memcpy(&m2, &b, 8);
                if (!std::isfinite(m2)) bad = true;
                if (m2 != 0.0) all_zero = false;
            }
            if (bad || all_zero) break;
        }
    }
    return DCS_OK;
}

int enqueue_contract(dcs_plan *p, int mode, int nl, int nunits_out);

// Lane ids of a contraction: within the plan's lanes and distinct.
int check_lanes(const dcs_plan *p, int nl, const int32_t *lanes) {
    if (nl < 1 || nl > p->lanes) return fail(DCS_ERR_INVALID, "nlanes %d outside 1..%d", nl, p->lanes);
    for (int i = 0; i < nl; ++i) {
        if (lanes[i] < 0 || lanes[i] >= p->lanes) return fail(DCS_ERR_INVALID, "lane %d out of range", lanes[i]);
        for (int j = 0; j < i; ++j)
            if (lanes[j] == lanes[i]) return fail(DCS_ERR_INVALID, "lane %d repeated", lanes[i]);
    }
    return DCS_OK;
}

// The g0, g1 parameters of the f0-factored group kernels for the staged lanes.  Within each coset group
// a step of class c > 0 (not the last) gets g1 = f1 / f0 (the part its class uses) and leaves f0
// pending; the last step gets (g0, g1) = (P f0, P f1), P the product of the pending f0s in step order.
// False when some lane does not fit the classes: f0 == 0, a ratio not of its class (relative 1e-12)
// or beyond 2^30, or P not finite and nonzero.
bool stage_factors(dcs_plan *p, int nl) {
    const int n = p->nsteps;
    for (int li = 0; li < nl; ++li) {
        Dyn *dl = p->h_dyn + (size_t)li * n;
        for (const Seg &sg : p->sched_fused) {
            if (sg.type != 2) continue;
            double Pr = 1.0, Pi = 0.0;
            const int last = sg.first + sg.count - 1;
            for (int s = sg.first; s < last; ++s) {
                const int c = p->coef_class[s];
                if (!c) continue;
                Dyn &d = dl[s];
                const double m2 = d.f0r * d.f0r + d.f0i * d.f0i;
                if (!(m2 > 0.0) || !std::isfinite(m2)) return false;
                const double rr = (d.f1r * d.f0r + d.f1i * d.f0i) / m2, ri = (d.f1i * d.f0r - d.f1r * d.f0i) / m2;
                const double mag = std::hypot(rr, ri);
                if (!std::isfinite(mag) || mag > 0x1p30) return false;
                if (c == 1) {
                    if (std::fabs(ri) > 1e-12 * mag) return false;
                    d.g1r = rr;
                } else if (c == 2) {
                    if (std::fabs(rr) > 1e-12 * mag) return false;
                    d.g1i = ri;
                } else {
                    d.g1r = rr;
                    d.g1i = ri;
                }
                const double nr = Pr * d.f0r - Pi * d.f0i, ni = Pr * d.f0i + Pi * d.f0r;
                Pr = nr;
                Pi = ni;
            }
            const double pm = std::hypot(Pr, Pi);
            if (!(pm > 0.0) || !std::isfinite(pm)) return false;
            Dyn &d = dl[last];
            d.g0r = Pr * d.f0r - Pi * d.f0i;
            d.g0i = Pr * d.f0i + Pi * d.f0r;
            d.g1r = Pr * d.f1r - Pi * d.f1i;
            d.g1i = Pr * d.f1i + Pi * d.f1r;
        }
    }
    return true;
}

// Per-lane step parameters and lane ids into the plan's pinned staging buffers.
void stage_parameters(dcs_plan *p, int nl, const int32_t *lanes, const double *coeffs, const int32_t *ov,
                      const int32_t *y) {
    const int n = p->nsteps;
    for (int li = 0; li < nl; ++li)
        for (int s = 0; s < n; ++s) {
            Dyn &d = p->h_dyn[(size_t)li * n + s];
            const double *c = coeffs + ((size_t)li * n + s) * 4;
            d.f0r = c[0];
            d.f0i = c[1];
            d.f1r = c[2];
            d.f1i = c[3];
            d.ov = ov[(size_t)li * n + s] & 1;
            d.y = y[(size_t)li * n + s] & 1;
            d.pad[0] = d.pad[1] = 0;
            d.g0r = d.g0i = d.g1r = d.g1i = 0.0;
        }
    for (int i = 0; i < nl; ++i) p->h_lanes[i] = lanes[i];
    p->call_factored = factoring_on(p) && stage_factors(p, nl);
}

// One launch of dcs_contract_batch: entries [first, first + count) of the round, all of one segment type.
struct BatchLaunch {
    int type, first, count;
    uint32_t workgroups;
};

// The fused schedules of several plans in lockstep: round r runs segment r of every plan that has one,
// the resident runs of the round in one k_resident_batch launch and the coset groups in one
// k_coset_batch launch.  A plan's segments stay in order (one stream), and the plans of a round touch
// disjoint memory, so the two launches of a round may run in either order.
template <class S>
int run_batch(dcs_context *ctx, int n, const dcs_batch_item *it, int *launches) {
    hipStream_t st = ctx->stream;
    for (int i = 0; i < n; ++i) {
        dcs_plan *p = it[i].plan;
        const int nl = it[i].nlanes;
        HIPCHK(hipMemcpyAsync(p->d_dyn, p->h_dyn, sizeof(Dyn) * (size_t)nl * p->nsteps, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(p->d_lanes, p->h_lanes, sizeof(int) * nl, hipMemcpyHostToDevice, st));
        const u64 cnt = (u64)nl * p->units_fused;
        k_fill_u64<<<grid_for(cnt, ctx->cus, 4), kThreads, 0, st>>>(p->d_dmax, cnt, 0ull);
        HIPCHK(hipGetLastError());
        k_init_lanes<S><<<1, 64, 0, st>>>((S *)p->frontier, p->lane_stride, p->d_lanes, nl);
        HIPCHK(hipGetLastError());
    }
    std::vector<BatchEntry> ents;
    std::vector<BatchLaunch> plan_launches;
    for (size_t r = 0;; ++r) {
        bool any = false;
        for (int type = 1; type <= 2; ++type) {
            BatchLaunch L{type, (int)ents.size(), 0, 0};
            for (int i = 0; i < n; ++i) {
                const dcs_plan *p = it[i].plan;
                if (r >= p->sched_fused.size() || p->sched_fused[r].type != type) continue;
                const Seg &sg = p->sched_fused[r];
                BatchEntry e;
                memset(&e, 0, sizeof(e));
                e.base = p->frontier;
                e.lane_stride = p->lane_stride;
                e.lanes = p->d_lanes;
                e.steps = p->d_steps;
                e.groups = p->d_groups;
                e.passes = p->d_passes;
                e.psteps = p->d_psteps;
                e.addrv = p->d_addrv;
                e.dyn = p->d_dyn;
                e.dmax = p->d_dmax;
                e.nsteps = p->nsteps;
                e.nunits = p->units_fused;
                e.unit = sg.unit;
                e.nlanes = it[i].nlanes;
                if (type == 1) {
                    e.first = sg.first;
                    e.last = sg.first + sg.count;
                    e.blocks = 1;
                } else {
                    if (sg.dim > kMaxDim)
                        return fail(DCS_ERR_INVALID, "item %d: coset dimension %d exceeds %d (fuse_max_dim)", i, sg.dim,
                                    kMaxDim);
                    e.first = sg.group;
                    e.blocks = (uint32_t)std::max<u64>(1, std::min<u64>(p->group_ntiles[sg.group], (u64)ctx->cus * 8));
                }
                e.begin = L.workgroups;
                const u64 total = (u64)L.workgroups + (u64)e.blocks * (u64)e.nlanes;
                if (total > 0x7fffffffull) return fail(DCS_ERR_INVALID, "batched launch exceeds the grid limit");
                L.workgroups = (uint32_t)total;
                ents.push_back(e);
                ++L.count;
                any = true;
            }
            if (L.count) plan_launches.push_back(L);
        }
        if (!any) break;
    }
    if (ents.size() > ctx->batch_cap) {
        if (ctx->h_batch) (void)hipHostFree(ctx->h_batch);
        if (ctx->d_batch) (void)hipFree(ctx->d_batch);
        ctx->h_batch = ctx->d_batch = nullptr;
        ctx->batch_cap = 0;
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipHostMalloc((void **)&ctx->h_batch, sizeof(BatchEntry) * ents.size(), hipHostMallocDefault));
        HIPCHK(hipMalloc((void **)&ctx->d_batch, sizeof(BatchEntry) * ents.size()));
        ctx->batch_cap = ents.size();
    }
    // the staging buffer may still feed an earlier batch's copy
    HIPCHK(hipStreamSynchronize(st));
    if (!ents.empty()) {
        memcpy(ctx->h_batch, ents.data(), sizeof(BatchEntry) * ents.size());
        HIPCHK(hipMemcpyAsync(ctx->d_batch, ctx->h_batch, sizeof(BatchEntry) * ents.size(), hipMemcpyHostToDevice, st));
    }
    const dcs_plan *p0 = it[0].plan;
    for (const BatchLaunch &L : plan_launches) {
        const BatchEntry *e = ctx->d_batch + L.first;
        // the shapes dcs_contract_batch accepts: resident_bits 12, tile_log2 12 or 11, pass_max_dim 3, no
        // per-step rounding (one instantiation each, to keep the library small)
        if (L.type == 1) k_resident_batch<S, 12><<<L.workgroups, kThreads, 0, st>>>(e, L.count);
        else if (p0->tile_log2 == 12) k_coset_batch<S, 12, false, 3><<<L.workgroups, kThreads, 0, st>>>(e, L.count);
        else k_coset_batch<S, 11, false, 3><<<L.workgroups, kThreads, 0, st>>>(e, L.count);
        HIPCHK(hipGetLastError());
    }
    for (int i = 0; i < n; ++i) {
        dcs_plan *p = it[i].plan;
        HIPCHK(hipMemcpyAsync(p->h_dmax, p->d_dmax, sizeof(u64) * (size_t)it[i].nlanes * p->units_fused,
                              hipMemcpyDeviceToHost, st));
    }
    if (launches) *launches = (int)plan_launches.size();
    return DCS_OK;
}

}  // namespace

// ================================================================================================
extern "C" {

const char *dcs_last_error(void) { return g_error.c_str(); }

int dcs_get_device_count(int *count) {
    if (!count) return fail(DCS_ERR_INVALID, "null count");
    HIPCHK(hipGetDeviceCount(count));
    return DCS_OK;
}

int dcs_get_device_info(int device, dcs_device_info *out) {
    if (!out) return fail(DCS_ERR_INVALID, "null out");
    memset(out, 0, sizeof(*out));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    snprintf(out->name, sizeof(out->name), "%s", prop.name);
    snprintf(out->arch, sizeof(out->arch), "%s", prop.gcnArchName);
    out->compute_units = prop.multiProcessorCount;
    HIPCHK(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    out->free_bytes = fr;
    out->total_bytes = tot;
    HIPCHK(hipRuntimeGetVersion(&out->runtime_version));
    HIPCHK(hipDriverGetVersion(&out->driver_version));
    out->device = device;
    return DCS_OK;
}

int dcs_context_create(int device, dcs_context **out) {
    if (!out) return fail(DCS_ERR_INVALID, "null out");
    *out = nullptr;
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(DCS_ERR_INVALID, "device %d out of range (%d devices)", device, n);
    HIPCHK(hipSetDevice(device));
    dcs_context *c = new dcs_context();
    std::unique_ptr<dcs_context> guard(c);
    c->device = device;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    c->cus = prop.multiProcessorCount;
    int rc = scratch_address(device, &c->scratch);
    if (rc) return rc;
    HIPCHK(hipHostMalloc((void **)&c->h_draw, sizeof(DrawOut), hipHostMallocDefault));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipHostFree(c->h_draw);
        return fail(DCS_ERR_RUNTIME, "hipStreamCreateWithFlags failed on device %d", device);
    }
    *out = guard.release();
    return DCS_OK;
}

void dcs_context_destroy(dcs_context *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipHostFree(ctx->h_draw);
    if (ctx->h_batch) (void)hipHostFree(ctx->h_batch);
    if (ctx->d_batch) (void)hipFree(ctx->d_batch);
    delete ctx;
}

int dcs_synchronize(dcs_context *ctx) {
    if (!ctx) return fail(DCS_ERR_INVALID, "null context");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DCS_OK;
}

void dcs_plan_destroy(dcs_plan *p) {
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    for (auto &kv : p->graphs) (void)hipGraphExecDestroy(kv.second);
    p->graphs.clear();
    for (auto &pe : p->pending) {
        (void)hipEventDestroy(pe.a);
        (void)hipEventDestroy(pe.b);
    }
    void *dev[] = {p->frontier, p->d_dyn, p->d_lanes, p->d_dmax, p->d_steps, p->d_groups, p->d_passes, p->d_psteps,
                   p->d_addrv, p->d_smap, p->d_phase, p->d_blocks, p->d_cum, p->d_scratch, p->d_draw, p->frontier2,
                   p->d_gtables, p->d_gcombos};
    for (void *q : dev)
        if (q) (void)hipFree(q);
    void *host[] = {p->h_dyn, p->h_lanes, p->h_dmax, p->h_draw};
    for (void *q : host)
        if (q) (void)hipHostFree(q);
    delete p;
}

// Checks of a plan description that need no device.
int check_plan_desc(const dcs_plan_desc *d) {
    if (d->storage != DCS_STORAGE_C64 && d->storage != DCS_STORAGE_C128 && d->storage != DCS_STORAGE_F16X2)
        return fail(DCS_ERR_INVALID, "storage must be 0 (complex64), 1 (complex128) or 2 (binary16 pairs)");
    if (d->maxslots < 0 || d->maxslots > 40) return fail(DCS_ERR_INVALID, "maxslots %d out of range", d->maxslots);
    if (d->lanes < 1 || d->nsteps < 1 || d->blocks < 1) return fail(DCS_ERR_INVALID, "lanes, nsteps, blocks must be >= 1");
    if (d->k < 0 || d->k > 30 || popc(d->terminal_mask) > d->k)
        return fail(DCS_ERR_INVALID, "terminal mask holds more than k slots (k=%d)", d->k);
    if (d->tile_log2 < 11 || d->tile_log2 > 13)
        return fail(DCS_ERR_INVALID, "tile_log2 must be 11 or 12 (13: DCS_MODE_COMPILED only)");
    if (d->group_policy < 0 || d->group_policy > 3) return fail(DCS_ERR_INVALID, "group_policy must be 0..3");
    if (d->lds_buffers < 0 || d->lds_buffers > 5) return fail(DCS_ERR_INVALID, "lds_buffers must be 0..5");
    if (d->resident_bits != 0 && (d->resident_bits < 6 || d->resident_bits > 12))
        return fail(DCS_ERR_INVALID, "resident_bits must be 0 or 6..12");
    if (d->fuse_max_dim < 0 || d->fuse_max_dim > kMaxDimCompiled)
        return fail(DCS_ERR_INVALID, "fuse_max_dim must be 0..%d (0..%d for DCS_MODE_FUSED)", kMaxDimCompiled, kMaxDim);
    if (d->pass_max_dim < 3 || d->pass_max_dim > kPassMax) return fail(DCS_ERR_INVALID, "pass_max_dim must be 3 or 4");
    const u64 lim = (d->maxslots == 0) ? 1ull : ((1ull << d->maxslots) - 1);
    if (d->terminal_mask & ~lim) return fail(DCS_ERR_INVALID, "terminal slots beyond maxslots");
    return DCS_OK;
}

// Plan fields and validated host steps from a checked description (no device work).
int ingest_plan_steps(const dcs_plan_desc *d, dcs_plan *p, std::vector<StepDev> &sdev) {
    const u64 lim = (d->maxslots == 0) ? 1ull : ((1ull << d->maxslots) - 1);
    p->storage = d->storage;
    p->maxslots = d->maxslots;
    p->lanes = d->lanes;
    p->nsteps = d->nsteps;
    p->k = d->k;
    p->blocks = d->blocks;
    p->fuse_max_dim = d->fuse_max_dim;
    p->resident_bits = d->resident_bits;
    p->tile_log2 = d->tile_log2;
    p->exact_rounding = d->exact_rounding;
    p->pass_max_dim = d->pass_max_dim;
    p->lds_storage = d->lds_storage ? 1 : 0;
    p->group_policy = d->group_policy;
    p->lds_buffers = d->lds_buffers ? d->lds_buffers : 2;
    p->coef_class.clear();
    if (d->coef_class) {
        for (int s = 0; s < d->nsteps; ++s)
            if (d->coef_class[s] < 0 || d->coef_class[s] > 3)
                return fail(DCS_ERR_INVALID, "coef_class[%d] = %d outside 0..3", s, d->coef_class[s]);
        p->coef_class.assign(d->coef_class, d->coef_class + d->nsteps);
    }
    p->esz = storage_bytes(d->storage);
    p->lane_stride = 1ull << d->maxslots;
    p->terminal_mask = d->terminal_mask;
    p->unitaries.assign(d->open_unitaries, d->open_unitaries + (size_t)d->k * 8);
    sdev.assign(d->nsteps, StepDev{});
    for (int s = 0; s < d->nsteps; ++s) {
        StepHost h;
        h.kind = d->kind[s];
        h.pivot = d->pivot[s];
        h.pivot_a = d->pivot_a[s];
        h.K = d->K[s];
        h.Ka = d->Ka[s];
        h.beta = d->beta[s];
        h.in_mask = d->in_mask[s];
        h.out_mask = d->out_mask[s];
        if (h.kind < 0 || h.kind > 7) return (fail(DCS_ERR_INVALID, "step %d: kind %d", s, h.kind));
        if ((h.in_mask | h.out_mask | h.K | h.Ka | h.beta) & ~lim)
            return (fail(DCS_ERR_INVALID, "step %d: mask beyond maxslots", s));
        const bool needK = h.kind == 3 || h.kind == 5 || h.kind == 6 || h.kind == 7;
        const bool needKa = h.kind == 1 || h.kind == 6;
        if (needK && (!h.K || top_bit(h.K) != h.pivot || (h.K & ~h.in_mask)))
            return (fail(DCS_ERR_INVALID, "step %d: K / pivot inconsistent", s));
        if (needKa && (!h.Ka || top_bit(h.Ka) != h.pivot_a || (h.Ka & ~h.in_mask)))
            return (fail(DCS_ERR_INVALID, "step %d: Ka / pivot_a inconsistent", s));
        if (h.kind == 6 && ((h.Ka >> h.pivot) & 1)) return (fail(DCS_ERR_INVALID, "step %d: Ka holds the K pivot", s));
        if (h.kind == 0 && (h.pivot < 0 || h.pivot >= d->maxslots || ((h.in_mask >> h.pivot) & 1)))
            return (fail(DCS_ERR_INVALID, "step %d: grow slot invalid", s));
        if (s > 0 && h.in_mask != p->steps.back().out_mask)
            return (fail(DCS_ERR_INVALID, "step %d: live slots do not chain", s));
        if (s == 0 && h.in_mask != 0) return (fail(DCS_ERR_INVALID, "first step must start from the empty basis"));
        h.slots = orbit_slots(h);
        p->steps.push_back(h);
        StepArgs a;
        memset(&a, 0, sizeof(a));
        a.norbits = 1ull << popc(h.slots);
        a.K = h.K;
        a.Ka = h.Ka;
        a.beta = h.beta;
        a.kind = h.kind;
        a.pivot = h.pivot;
        a.runs = make_runs(h.slots);
        p->step_args.push_back(a);
        StepDev &sd = sdev[s];
        sd.K = h.K;
        sd.Ka = h.Ka;
        sd.beta = h.beta;
        sd.slots = h.slots;
        sd.in_mask = h.in_mask;
        sd.out_mask = h.out_mask;
        sd.kind = h.kind;
        sd.pivot = h.pivot;
        sd.pivot_a = h.pivot_a;
        sd.nslots = popc(h.slots);
    }
    if (p->steps.back().out_mask != d->terminal_mask)
        return (fail(DCS_ERR_INVALID, "terminal slots differ from the live slots after the last step"));
    return DCS_OK;
}

int dcs_plan_create(dcs_context *ctx, const dcs_plan_desc *d, dcs_plan **out) {
    if (!ctx || !d || !out) return fail(DCS_ERR_INVALID, "null argument");
    *out = nullptr;
    {
        const int rc = check_plan_desc(d);
        if (rc) return rc;
    }
    HIPCHK(hipSetDevice(ctx->device));

    dcs_plan *p = new dcs_plan();
    p->ctx = ctx;
    auto bail = [&](int rc) {
        dcs_plan_destroy(p);
        return rc;
    };
    std::vector<StepDev> sdev;
    {
        const int rc = ingest_plan_steps(d, p, sdev);
        if (rc) return bail(rc);
    }

    std::vector<GroupDev> groups;
    std::vector<PassDev> passes;
    std::vector<PStepDev> psteps;
    std::vector<u64> addrv;
    {
        const int rc = build_schedules(p, groups, passes, psteps, addrv);
        if (rc) return bail(rc);
    }

    // Device memory admission: the frontier must fit what is free now.
    const u64 frontier_bytes = (u64)d->lanes * p->lane_stride * p->esz;
    const u64 block_len = 1ull << d->k;
    const u64 aux_bytes = block_len * (8 + 8 + 16 * (u64)d->blocks + 8 + 16) + (64ull << 20);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return bail(fail(DCS_ERR_RUNTIME, "hipMemGetInfo failed"));
    if (frontier_bytes + aux_bytes > fr)
        return bail(fail(DCS_ERR_NOMEM, "frontier needs %llu bytes (+%llu auxiliary), %llu free", (unsigned long long)frontier_bytes,
                         (unsigned long long)aux_bytes, (unsigned long long)fr));
#define DALLOC(ptr, bytes)                                                                       \
    do {                                                                                         \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess)                                   \
            return bail(fail(DCS_ERR_NOMEM, "hipMalloc of %llu bytes failed", (unsigned long long)(bytes))); \
        p->device_bytes += (u64)(bytes);                                                         \
    } while (0)
#define HALLOC(ptr, bytes)                                                                       \
    do {                                                                                         \
        if (hipHostMalloc((void **)&(ptr), (bytes), hipHostMallocDefault) != hipSuccess)         \
            return bail(fail(DCS_ERR_NOMEM, "hipHostMalloc of %llu bytes failed", (unsigned long long)(bytes))); \
    } while (0)
    const int max_units = std::max(p->units_steps, p->units_fused);
    DALLOC(p->frontier, frontier_bytes);
    DALLOC(p->d_dyn, sizeof(Dyn) * (size_t)d->lanes * d->nsteps);
    DALLOC(p->d_lanes, sizeof(int) * (size_t)d->lanes);
    DALLOC(p->d_dmax, sizeof(u64) * (size_t)d->lanes * max_units);
    DALLOC(p->d_steps, sizeof(StepDev) * sdev.size());
    DALLOC(p->d_groups, sizeof(GroupDev) * std::max<size_t>(1, groups.size()));
    DALLOC(p->d_passes, sizeof(PassDev) * std::max<size_t>(1, passes.size()));
    DALLOC(p->d_psteps, sizeof(PStepDev) * std::max<size_t>(1, psteps.size()));
    DALLOC(p->d_addrv, sizeof(u64) * std::max<size_t>(1, addrv.size()));
    DALLOC(p->d_smap, sizeof(u64) * block_len);
    DALLOC(p->d_phase, sizeof(double) * block_len);
    DALLOC(p->d_blocks, sizeof(double2) * block_len * d->blocks);
    DALLOC(p->d_cum, sizeof(double) * block_len);
    DALLOC(p->d_scratch, draw_scratch_bytes(block_len));
    DALLOC(p->d_draw, sizeof(DrawOut));
    HALLOC(p->h_dyn, sizeof(Dyn) * (size_t)d->lanes * d->nsteps);
    HALLOC(p->h_lanes, sizeof(int) * (size_t)d->lanes);
    HALLOC(p->h_dmax, sizeof(u64) * (size_t)d->lanes * max_units);
    HALLOC(p->h_draw, sizeof(DrawOut));
#undef DALLOC
#undef HALLOC
    hipStream_t st = ctx->stream;
#define H2D(dst, src, bytes)                                                                   \
    do {                                                                                       \
        if (hipMemcpyAsync((dst), (src), (bytes), hipMemcpyHostToDevice, st) != hipSuccess)    \
            return bail(fail(DCS_ERR_RUNTIME, "upload failed"));                               \
    } while (0)
    H2D(p->d_steps, sdev.data(), sizeof(StepDev) * sdev.size());
    if (!groups.empty()) H2D(p->d_groups, groups.data(), sizeof(GroupDev) * groups.size());
    if (!passes.empty()) H2D(p->d_passes, passes.data(), sizeof(PassDev) * passes.size());
    if (!psteps.empty()) H2D(p->d_psteps, psteps.data(), sizeof(PStepDev) * psteps.size());
    if (!addrv.empty()) H2D(p->d_addrv, addrv.data(), sizeof(u64) * addrv.size());
    H2D(p->d_smap, d->syndrome_map, sizeof(u64) * ((u64)1 << popc(d->terminal_mask)));
    H2D(p->d_phase, d->phase, sizeof(double) * block_len);
#undef H2D
    if (hipStreamSynchronize(st) != hipSuccess) return bail(fail(DCS_ERR_RUNTIME, "plan upload sync failed"));
    p->group_ntiles.resize(groups.size());
    for (size_t i = 0; i < groups.size(); ++i) p->group_ntiles[i] = groups[i].ntiles;
    p->h_groups = std::move(groups);
    p->h_passes = std::move(passes);
    p->h_psteps = std::move(psteps);
    p->h_addrv = std::move(addrv);
    *out = p;
    return DCS_OK;
}


int dcs_gather_plan_create(dcs_context *ctx, const dcs_gather_desc *d, dcs_plan **out) {
    if (!ctx || !d || !out) return fail(DCS_ERR_INVALID, "null argument");
    *out = nullptr;
    if (d->storage != DCS_STORAGE_C64 && d->storage != DCS_STORAGE_C128)
        return fail(DCS_ERR_INVALID, "storage must be 0 (complex64) or 1 (complex128)");
    if (d->rmax < 0 || d->rmax > 40) return fail(DCS_ERR_INVALID, "rmax %d out of range", d->rmax);
    if (d->nsteps < 1 || d->blocks < 1) return fail(DCS_ERR_INVALID, "nsteps and blocks must be >= 1");
    if (d->k < 0 || d->k > 30 || d->terminal_rank < 0 || d->terminal_rank > d->k)
        return fail(DCS_ERR_INVALID, "terminal rank %d outside 0..k (k=%d)", d->terminal_rank, d->k);
    if (!d->r_in || !d->r_out || !d->is_free || !d->nbytes || !d->ncombos || !d->tab_off || !d->comb_off || !d->tables ||
        !d->combos || !d->beta || !d->syndrome_map || !d->phase || !d->open_unitaries)
        return fail(DCS_ERR_INVALID, "null table");
    int64_t ntab = 0, ncomb = 0;
    for (int s = 0; s < d->nsteps; ++s) {
        if (d->r_in[s] < 0 || d->r_in[s] > d->rmax || d->r_out[s] < 0 || d->r_out[s] > d->rmax)
            return fail(DCS_ERR_INVALID, "step %d: table bits outside 0..rmax", s);
        if (s > 0 && d->r_in[s] != d->r_out[s - 1]) return fail(DCS_ERR_INVALID, "step %d: tables do not chain", s);
        if (s == 0 && d->r_in[s] != 0) return fail(DCS_ERR_INVALID, "the first step must read the one-entry table");
        if (d->nbytes[s] < 0 || d->nbytes[s] > 5 || d->nbytes[s] * 8 < d->r_out[s])
            return fail(DCS_ERR_INVALID, "step %d: %d byte tables for %d output bits", s, d->nbytes[s], d->r_out[s]);
        if (d->ncombos[s] < 1 || d->ncombos[s] > 8) return fail(DCS_ERR_INVALID, "step %d: 1..8 kernel combinations", s);
        if (d->tab_off[s] < 0 || d->comb_off[s] < 0) return fail(DCS_ERR_INVALID, "step %d: negative offset", s);
        ntab = std::max<int64_t>(ntab, d->tab_off[s] + (int64_t)d->nbytes[s] * 256);
        ncomb = std::max<int64_t>(ncomb, d->comb_off[s] + d->ncombos[s]);
    }
    if (d->r_out[d->nsteps - 1] != d->terminal_rank)
        return fail(DCS_ERR_INVALID, "the last table has %d bits, terminal rank %d", d->r_out[d->nsteps - 1], d->terminal_rank);
    HIPCHK(hipSetDevice(ctx->device));
    dcs_plan *p = new dcs_plan();
    auto bail = [&](int rc) {
        dcs_plan_destroy(p);
        return rc;
    };
    p->ctx = ctx;
    p->executor = 1;
    p->storage = d->storage;
    p->maxslots = d->rmax;
    p->lanes = 1;
    p->nsteps = d->nsteps;
    p->k = d->k;
    p->blocks = d->blocks;
    p->esz = storage_bytes(d->storage);
    p->lane_stride = 1ull << d->rmax;
    p->terminal_mask = (d->terminal_rank == 0) ? 0ull : ((1ull << d->terminal_rank) - 1);
    p->unitaries.assign(d->open_unitaries, d->open_unitaries + (size_t)d->k * 8);
    for (int s = 0; s < d->nsteps; ++s) p->sched_steps.push_back(Seg{0, s, 1, 0, -1, s});
    p->sched_fused = p->sched_steps;
    p->units_steps = p->units_fused = d->nsteps;
    const u64 table_bytes = p->lane_stride * p->esz;
    const u64 block_len = 1ull << d->k;
    const u64 aux_bytes = block_len * (8 + 8 + 16 * (u64)d->blocks + 8 + 16) + (64ull << 20) + 8ull * (ntab + ncomb);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return bail(fail(DCS_ERR_RUNTIME, "hipMemGetInfo failed"));
    if (2 * table_bytes + aux_bytes > fr)
        return bail(fail(DCS_ERR_NOMEM, "two tables need %llu bytes (+%llu auxiliary), %llu free",
                         (unsigned long long)(2 * table_bytes), (unsigned long long)aux_bytes, (unsigned long long)fr));
#define DALLOC(ptr, bytes)                                                                       \
    do {                                                                                         \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess)                                   \
            return bail(fail(DCS_ERR_NOMEM, "hipMalloc of %llu bytes failed", (unsigned long long)(bytes))); \
        p->device_bytes += (u64)(bytes);                                                         \
    } while (0)
#define HALLOC(ptr, bytes)                                                                       \
    do {                                                                                         \
        if (hipHostMalloc((void **)&(ptr), (bytes), hipHostMallocDefault) != hipSuccess)         \
            return bail(fail(DCS_ERR_NOMEM, "hipHostMalloc of %llu bytes failed", (unsigned long long)(bytes))); \
    } while (0)
    DALLOC(p->frontier, table_bytes);
    DALLOC(p->frontier2, table_bytes);
    DALLOC(p->d_gtables, sizeof(u64) * std::max<int64_t>(1, ntab));
    DALLOC(p->d_gcombos, sizeof(u64) * std::max<int64_t>(1, ncomb));
    DALLOC(p->d_dyn, sizeof(Dyn) * (size_t)d->nsteps);
    DALLOC(p->d_lanes, sizeof(int));
    DALLOC(p->d_dmax, sizeof(u64) * (size_t)d->nsteps);
    DALLOC(p->d_smap, sizeof(u64) * block_len);
    DALLOC(p->d_phase, sizeof(double) * block_len);
    DALLOC(p->d_blocks, sizeof(double2) * block_len * d->blocks);
    DALLOC(p->d_cum, sizeof(double) * block_len);
    DALLOC(p->d_scratch, draw_scratch_bytes(block_len));
    DALLOC(p->d_draw, sizeof(DrawOut));
    HALLOC(p->h_dyn, sizeof(Dyn) * (size_t)d->nsteps);
    HALLOC(p->h_lanes, sizeof(int));
    HALLOC(p->h_dmax, sizeof(u64) * (size_t)d->nsteps);
    HALLOC(p->h_draw, sizeof(DrawOut));
#undef DALLOC
#undef HALLOC
    hipStream_t st = ctx->stream;
    if (hipMemcpyAsync(p->d_gtables, d->tables, sizeof(u64) * ntab, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(p->d_gcombos, d->combos, sizeof(u64) * ncomb, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(p->d_smap, d->syndrome_map, sizeof(u64) * ((u64)1 << d->terminal_rank), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(p->d_phase, d->phase, sizeof(double) * block_len, hipMemcpyHostToDevice, st) != hipSuccess)
        return bail(fail(DCS_ERR_RUNTIME, "upload failed"));
    if (hipStreamSynchronize(st) != hipSuccess) return bail(fail(DCS_ERR_RUNTIME, "plan upload sync failed"));
    for (int s = 0; s < d->nsteps; ++s) {
        GatherArgs a;
        memset(&a, 0, sizeof(a));
        a.n_out = 1ull << d->r_out[s];
        a.mask = (d->r_in[s] == 0) ? 0ull : ((1ull << d->r_in[s]) - 1);
        a.beta = d->beta[s];
        a.T = p->d_gtables + d->tab_off[s];
        a.combos = p->d_gcombos + d->comb_off[s];
        a.r_in = d->r_in[s];
        a.nbytes = d->nbytes[s];
        a.ncombos = d->ncombos[s];
        a.is_free = d->is_free[s] ? 1 : 0;
        p->gsteps.push_back(a);
    }
    *out = p;
    return DCS_OK;
}

int dcs_plan_compile(dcs_plan *p, double *seconds) {
    if (!p) return fail(DCS_ERR_INVALID, "null plan");
    return compile_groups(p, seconds);
}

int dcs_plan_factoring(dcs_plan *p, int32_t *module_factored, int32_t *last_call_factored) {
    if (!p || !module_factored || !last_call_factored) return fail(DCS_ERR_INVALID, "null argument");
    *module_factored = p->factored_module ? 1 : 0;
    *last_call_factored = (p->factored_module && p->call_factored) ? 1 : 0;
    return DCS_OK;
}

int dcs_plan_kernel_source(const dcs_plan_desc *d, char *src, uint64_t src_cap, uint64_t *src_len, int32_t *group,
                           int32_t *first, int32_t *count, int32_t *threads, int cap, int *ngroups) {
    if (!d || !src_len || !ngroups) return fail(DCS_ERR_INVALID, "null argument");
    {
        const int rc = check_plan_desc(d);
        if (rc) return rc;
    }
    std::unique_ptr<dcs_plan> p(new dcs_plan());  // no context: nothing here touches a device
    std::vector<StepDev> sdev;
    std::vector<GroupDev> groups;
    std::vector<PassDev> passes;
    std::vector<PStepDev> psteps;
    std::vector<u64> addrv;
    {
        int rc = ingest_plan_steps(d, p.get(), sdev);
        if (!rc) rc = build_schedules(p.get(), groups, passes, psteps, addrv);
        if (rc) return rc;
    }
    const std::string text = group_module_source(p.get(), groups, passes, psteps, addrv, true);
    *src_len = text.size();
    if (src && src_cap) {
        const size_t n = std::min<size_t>(text.size(), (size_t)src_cap - 1);
        memcpy(src, text.data(), n);
        src[n] = '\0';
    }
    int n = 0;
    for (const Seg &sg : p->sched_fused) {
        if (sg.type != 2) continue;
        if (n < cap) {
            if (group) group[n] = sg.group;
            if (first) first[n] = sg.first;
            if (count) count[n] = sg.count;
            if (threads) threads[n] = codegen::group_threads(groups[sg.group]);
        }
        ++n;
    }
    *ngroups = n;
    return DCS_OK;
}

int dcs_plan_memory(dcs_plan *p, uint64_t *bytes) {
    if (!p || !bytes) return fail(DCS_ERR_INVALID, "null argument");
    *bytes = p->device_bytes;
    return DCS_OK;
}

int dcs_plan_units(dcs_plan *p, int mode, int *nunits) {
    if (!p || !nunits) return fail(DCS_ERR_INVALID, "null argument");
    if (mode < 0 || mode > 3) return fail(DCS_ERR_INVALID, "mode %d", mode);
    *nunits = (mode >= DCS_MODE_FUSED) ? p->units_fused : p->units_steps;
    return DCS_OK;
}

int dcs_plan_schedule(dcs_plan *p, int mode, int cap, int32_t *seg_kind, int32_t *first, int32_t *count,
                      int32_t *dim, int *nseg) {
    if (!p || !nseg) return fail(DCS_ERR_INVALID, "null argument");
    if (mode < 0 || mode > 3) return fail(DCS_ERR_INVALID, "mode %d", mode);
    const std::vector<Seg> &s = (mode >= DCS_MODE_FUSED) ? p->sched_fused : p->sched_steps;
    *nseg = (int)s.size();
    for (int i = 0; i < (int)s.size() && i < cap; ++i) {
        if (seg_kind) seg_kind[i] = s[i].type;
        if (first) first[i] = s[i].first;
        if (count) count[i] = s[i].count;
        if (dim) dim[i] = s[i].dim;
    }
    return DCS_OK;
}

int dcs_contract(dcs_plan *p, int mode, int nl, const int32_t *lanes, const double *coeffs, const int32_t *ov,
                 const int32_t *y, double *unit_m2) {
    if (!p || !lanes || !coeffs || !ov || !y || !unit_m2) return fail(DCS_ERR_INVALID, "null argument");
    if (mode < 0 || mode > 3) return fail(DCS_ERR_INVALID, "mode %d", mode);
    if (p->executor == 1 && mode != DCS_MODE_STEPS && mode != DCS_MODE_STEPS_HOST)
        return fail(DCS_ERR_INVALID, "a gather plan runs DCS_MODE_STEPS or DCS_MODE_STEPS_HOST");
    {
        const int rc = check_lanes(p, nl, lanes);
        if (rc) return rc;
    }
    HIPCHK(hipSetDevice(p->ctx->device));
    if (mode == DCS_MODE_COMPILED && !p->compiled) {
        const int rc = compile_groups(p, nullptr);
        if (rc) return rc;
    }
    stage_parameters(p, nl, lanes, coeffs, ov, y);
    if (mode == DCS_MODE_COMPILED && p->factored_module && !p->call_factored) {
        const int rc = compile_plain(p);
        if (rc) return rc;
    }
    hipStream_t st = p->ctx->stream;
    const int nunits_out = (mode >= DCS_MODE_FUSED) ? p->units_fused : p->units_steps;
    // Graph replay: the whole sequence below (parameter upload from the pinned staging buffers, maxima
    // cleared by a fill kernel, every launch, maxima read back into pinned memory) is captured once per
    // (mode, lane count) and replayed; the staging buffers are read when a replay runs.
    const bool graphable = p->use_graphs && !p->profiling && p->executor == 0 && mode != DCS_MODE_STEPS_HOST;
    if (graphable) {
        // the factored and unfactored group kernels are captured apart
        const bool plain = mode == DCS_MODE_COMPILED && p->factored_module && !p->call_factored;
        const auto key = std::make_pair(mode + (plain ? 16 : 0), nl);
        auto it = p->graphs.find(key);
        if (it == p->graphs.end()) {
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_contract(p, mode, nl, nunits_out);
            hipGraph_t graph = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (ec != hipSuccess) return fail(DCS_ERR_RUNTIME, "graph capture: %s", hipGetErrorString(ec));
            hipGraphExec_t exec = nullptr;
            const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) return fail(DCS_ERR_RUNTIME, "graph instantiate: %s", hipGetErrorString(ei));
            it = p->graphs.emplace(key, exec).first;
        }
        HIPCHK(hipGraphLaunch(it->second, st));
        HIPCHK(hipStreamSynchronize(st));
        memcpy(unit_m2, p->h_dmax, sizeof(double) * (size_t)nl * nunits_out);
        return DCS_OK;
    }
    int rc = enqueue_contract(p, mode, nl, nunits_out);
    if (rc) {
        (void)hipStreamSynchronize(st);
        return rc;
    }
    HIPCHK(hipStreamSynchronize(st));
    memcpy(unit_m2, p->h_dmax, sizeof(double) * (size_t)nl * nunits_out);
    return ev_resolve(p);
}

int dcs_contract_batch(int n, const dcs_batch_item *items, int *launches) {
    if (!items || n < 1 || n > DCS_BATCH_MAX) return fail(DCS_ERR_INVALID, "batch of %d items outside 1..%d", n, DCS_BATCH_MAX);
    const dcs_plan *p0 = items[0].plan;
    if (!p0) return fail(DCS_ERR_INVALID, "item 0: null plan");
    for (int i = 0; i < n; ++i) {
        const dcs_batch_item &b = items[i];
        const dcs_plan *p = b.plan;
        if (!p || !b.lanes || !b.coeffs || !b.ov || !b.y || !b.unit_m2)
            return fail(DCS_ERR_INVALID, "item %d: null argument", i);
        if (p->executor != 0) return fail(DCS_ERR_INVALID, "item %d: a gather plan cannot be batched", i);
        if (p->ctx != p0->ctx) return fail(DCS_ERR_INVALID, "item %d: batched plans must share one context", i);
        if (p->storage != p0->storage || p->tile_log2 != p0->tile_log2 || p->pass_max_dim != p0->pass_max_dim ||
            p->exact_rounding != p0->exact_rounding || p->resident_bits != p0->resident_bits)
            return fail(DCS_ERR_INVALID,
                        "item %d: batched plans must share storage, tile_log2, pass_max_dim, exact_rounding and "
                        "resident_bits", i);
        for (int j = 0; j < i; ++j)
            if (items[j].plan == p) return fail(DCS_ERR_INVALID, "item %d: plan repeated (give it several lanes)", i);
        const int rc = check_lanes(p, b.nlanes, b.lanes);
        if (rc) return rc;
        for (const Seg &sg : p->sched_fused)
            if (sg.type == 2 && sg.dim > kMaxDim)
                return fail(DCS_ERR_INVALID, "item %d: coset dimension %d exceeds %d: use a smaller fuse_max_dim", i,
                            sg.dim, kMaxDim);
    }
    if ((p0->tile_log2 != 11 && p0->tile_log2 != 12) || p0->pass_max_dim != 3 || p0->exact_rounding ||
        p0->resident_bits != 12 || p0->storage == DCS_STORAGE_F16X2)
        return fail(DCS_ERR_INVALID,
                    "batched plans: complex64 or complex128 storage, tile_log2 11 or 12, pass_max_dim 3, "
                    "resident_bits 12, no exact_rounding");
    dcs_context *ctx = p0->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    for (int i = 0; i < n; ++i)
        stage_parameters(items[i].plan, items[i].nlanes, items[i].lanes, items[i].coeffs, items[i].ov, items[i].y);
    int rc;
    if (p0->storage == DCS_STORAGE_C64) rc = run_batch<float2>(ctx, n, items, launches);
    else rc = run_batch<double2>(ctx, n, items, launches);
    if (rc) {
        (void)hipStreamSynchronize(ctx->stream);
        return rc;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i)
        memcpy(items[i].unit_m2, items[i].plan->h_dmax, sizeof(double) * (size_t)items[i].nlanes * items[i].plan->units_fused);
    return DCS_OK;
}

int dcs_plan_use_graphs(dcs_plan *p, int enable) {
    if (!p) return fail(DCS_ERR_INVALID, "null plan");
    HIPCHK(hipSetDevice(p->ctx->device));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    if (!enable) {
        for (auto &kv : p->graphs) (void)hipGraphExecDestroy(kv.second);
        p->graphs.clear();
    }
    p->use_graphs = enable != 0;
    return DCS_OK;
}

}  // extern "C"

namespace {

int enqueue_contract(dcs_plan *p, int mode, int nl, int nunits_out) {
    const int n = p->nsteps;
    hipStream_t st = p->ctx->stream;
    HIPCHK(hipMemcpyAsync(p->d_dyn, p->h_dyn, sizeof(Dyn) * (size_t)nl * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->d_lanes, p->h_lanes, sizeof(int) * nl, hipMemcpyHostToDevice, st));
    int rc;
    if (p->executor == 1)
        rc = (p->storage == DCS_STORAGE_C64) ? run_gather<float2>(p, mode) : run_gather<double2>(p, mode);
    else if (p->storage == DCS_STORAGE_C64)
        rc = run_contract<float2>(p, mode, nl, p->group_ntiles);
    else if (p->storage == DCS_STORAGE_C128)
        rc = run_contract<double2>(p, mode, nl, p->group_ntiles);
    else
        rc = run_contract<h16x2>(p, mode, nl, p->group_ntiles);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(p->h_dmax, p->d_dmax, sizeof(u64) * (size_t)nl * nunits_out, hipMemcpyDeviceToHost, st));
    return DCS_OK;
}

}  // namespace

extern "C" {

int dcs_block_zero(dcs_plan *p, int block) {
    if (!p || block < 0 || block >= p->blocks) return fail(DCS_ERR_INVALID, "block out of range");
    HIPCHK(hipSetDevice(p->ctx->device));
    const u64 n = 2ull << p->k;  // doubles
    PROF(3, (k_fill_u64<<<grid_for(n, p->ctx->cus, 8), kThreads, 0, p->ctx->stream>>>(
                 (u64 *)(p->d_blocks + ((u64)block << p->k)), n, 0ull)));
    return DCS_OK;
}

int dcs_block_accumulate(dcs_plan *p, int block, int lane, uint64_t xmask, double cr, double ci) {
    if (!p || block < 0 || block >= p->blocks) return fail(DCS_ERR_INVALID, "block out of range");
    if (lane < 0 || lane >= p->lanes) return fail(DCS_ERR_INVALID, "lane out of range");
    if (xmask >= (1ull << p->k)) return fail(DCS_ERR_INVALID, "xmask outside the block");
    const u64 n = 1ull << popc(p->terminal_mask);
    HIPCHK(hipSetDevice(p->ctx->device));
    const Runs term = make_runs(p->terminal_mask);
    double2 *tot = p->d_blocks + ((u64)block << p->k);
    const unsigned g = grid_for(n, p->ctx->cus, 8);
    if (p->storage == DCS_STORAGE_C64) {
        const float2 *E = (const float2 *)p->frontier + (u64)lane * p->lane_stride;
        PROF(3, (k_scatter<float2><<<g, kThreads, 0, p->ctx->stream>>>(E, term, n, p->d_smap, xmask, cd{cr, ci}, tot)));
    } else if (p->storage == DCS_STORAGE_F16X2) {
        const h16x2 *E = (const h16x2 *)p->frontier + (u64)lane * p->lane_stride;
        PROF(3, (k_scatter<h16x2><<<g, kThreads, 0, p->ctx->stream>>>(E, term, n, p->d_smap, xmask, cd{cr, ci}, tot)));
    } else {
        const double2 *E = (const double2 *)p->frontier + (u64)lane * p->lane_stride;
        PROF(3, (k_scatter<double2><<<g, kThreads, 0, p->ctx->stream>>>(E, term, n, p->d_smap, xmask, cd{cr, ci}, tot)));
    }
    return DCS_OK;
}

int dcs_block_transform(dcs_plan *p, int block) {
    if (!p || block < 0 || block >= p->blocks) return fail(DCS_ERR_INVALID, "block out of range");
    HIPCHK(hipSetDevice(p->ctx->device));
    const u64 n = 1ull << p->k;
    double2 *v = p->d_blocks + ((u64)block << p->k);
    const unsigned gh = grid_for(std::max<u64>(1, n / 2), p->ctx->cus, 8);
    const unsigned gn = grid_for(n, p->ctx->cus, 8);
    for (int q = 0; q < p->k; ++q) PROF(3, (k_wht_pass<<<gh, kThreads, 0, p->ctx->stream>>>(v, n / 2, q)));
    PROF(3, (k_phase<<<gn, kThreads, 0, p->ctx->stream>>>(v, p->d_phase, n)));
    for (int q = 0; q < p->k; ++q) {
        const double *u = &p->unitaries[(size_t)q * 8];
        PROF(3, (k_u_pass<<<gh, kThreads, 0, p->ctx->stream>>>(v, n / 2, q, cd{u[0], u[1]}, cd{u[2], u[3]},
                                                               cd{u[4], u[5]}, cd{u[6], u[7]})));
    }
    return DCS_OK;
}

int dcs_block_read(dcs_plan *p, int block, double *host_out) {
    if (!p || !host_out || block < 0 || block >= p->blocks) return fail(DCS_ERR_INVALID, "bad argument");
    HIPCHK(hipSetDevice(p->ctx->device));
    HIPCHK(hipMemcpyAsync(host_out, p->d_blocks + ((u64)block << p->k), sizeof(double2) << p->k, hipMemcpyDeviceToHost,
                          p->ctx->stream));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    return ev_resolve(p);
}

int dcs_block_device_pointer(dcs_plan *p, int block, void **ptr) {
    if (!p || !ptr || block < 0 || block >= p->blocks) return fail(DCS_ERR_INVALID, "bad argument");
    *ptr = (void *)(p->d_blocks + ((u64)block << p->k));
    return DCS_OK;
}

int dcs_draw(dcs_plan *p, int block, double u, dcs_draw_result *out) {
    if (!p || !out || block < 0 || block >= p->blocks) return fail(DCS_ERR_INVALID, "bad argument");
    if (!(u >= 0.0 && u < 1.0)) return fail(DCS_ERR_INVALID, "uniform must lie in [0, 1)");
    HIPCHK(hipSetDevice(p->ctx->device));
    const u64 n = 1ull << p->k;
    const double2 *v = p->d_blocks + ((u64)block << p->k);
    double *ck = (double *)p->d_scratch;  // n / 64 + 1 doubles
    PROF(4, (k_weights<<<grid_for(n, p->ctx->cus, 4), kThreads, 0, p->ctx->stream>>>(v, n, p->d_cum)));
    PROF(4, (k_chain<<<1, 64, 0, p->ctx->stream>>>(WeightArray{p->d_cum}, n, ck)));
    PROF(4, (k_search<<<1, 64, 0, p->ctx->stream>>>(WeightArray{p->d_cum}, ck, n, u, p->d_draw)));
    double *part = ck + draw_parts_offset(n);
    PROF(4, (k_interval_part<<<kIntervalBlocks, 256, 0, p->ctx->stream>>>(WeightArray{p->d_cum}, n, p->d_draw, part)));
    PROF(4, (k_interval_check<<<1, 64, 0, p->ctx->stream>>>(WeightArray{p->d_cum}, part, u, p->d_draw)));
    HIPCHK(hipMemcpyAsync(p->h_draw, p->d_draw, sizeof(DrawOut), hipMemcpyDeviceToHost, p->ctx->stream));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    int rc = ev_resolve(p);
    if (rc) return rc;
    return draw_result(*p->h_draw, out);
}

int dcs_draw_noalloc(dcs_context *ctx, const void *block, int k, double u, dcs_draw_result *out) {
    if (!ctx || !block || !out) return fail(DCS_ERR_INVALID, "bad argument");
    if (k < 0 || k > DCS_DRAW_NOALLOC_MAX_K)
        return fail(DCS_ERR_INVALID, "block of 2^%d entries: dcs_draw_noalloc's scratch holds blocks of at most 2^%d",
                    k, DCS_DRAW_NOALLOC_MAX_K);
    if (!(u >= 0.0 && u < 1.0)) return fail(DCS_ERR_INVALID, "uniform must lie in [0, 1)");
    HIPCHK(hipSetDevice(ctx->device));
    const u64 n = 1ull << k;
    // every segment sum k_search reads (ck[0 .. n/64]), the whole DrawOut and every partial sum of the
    // interval check are written by this call's own kernels, so nothing left by an earlier call is read
    double *ck = (double *)ctx->scratch;
    DrawOut *d = (DrawOut *)(ck + kNoallocDrawOut);
    double *part = ck + kNoallocParts;
    const BlockWeights w{(const double2 *)block};
    k_chain<<<1, 64, 0, ctx->stream>>>(w, n, ck);
    k_search<<<1, 64, 0, ctx->stream>>>(w, ck, n, u, d);
    k_interval_part<<<kIntervalBlocks, 256, 0, ctx->stream>>>(w, n, d, part);
    k_interval_check<<<1, 64, 0, ctx->stream>>>(w, part, u, d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ctx->h_draw, d, sizeof(DrawOut), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return draw_result(*ctx->h_draw, out);
}

int dcs_frontier_read(dcs_plan *p, int lane, void *host_out) {
    if (!p || !host_out || lane < 0 || lane >= p->lanes) return fail(DCS_ERR_INVALID, "bad argument");
    HIPCHK(hipSetDevice(p->ctx->device));
    const u64 n = 1ull << popc(p->terminal_mask);
    const Runs term = make_runs(p->terminal_mask);
    const unsigned g = grid_for(n, p->ctx->cus, 8);
    if (p->storage == DCS_STORAGE_C64) {
        const float2 *E = (const float2 *)p->frontier + (u64)lane * p->lane_stride;
        k_gather_terminal<float2><<<g, kThreads, 0, p->ctx->stream>>>(E, term, n, (float2 *)p->d_scratch);
    } else if (p->storage == DCS_STORAGE_F16X2) {
        const h16x2 *E = (const h16x2 *)p->frontier + (u64)lane * p->lane_stride;
        k_gather_terminal<h16x2><<<g, kThreads, 0, p->ctx->stream>>>(E, term, n, (h16x2 *)p->d_scratch);
    } else {
        const double2 *E = (const double2 *)p->frontier + (u64)lane * p->lane_stride;
        k_gather_terminal<double2><<<g, kThreads, 0, p->ctx->stream>>>(E, term, n, (double2 *)p->d_scratch);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host_out, p->d_scratch, p->esz * n, hipMemcpyDeviceToHost, p->ctx->stream));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    return DCS_OK;
}

int dcs_plan_profile(dcs_plan *p, int enable, double *ms, int64_t *launches) {
    if (!p) return fail(DCS_ERR_INVALID, "null plan");
    HIPCHK(hipSetDevice(p->ctx->device));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    int rc = ev_resolve(p);
    if (rc) return rc;
    for (int i = 0; i < 5; ++i) {
        if (ms) ms[i] = p->prof_ms[i];
        if (launches) launches[i] = p->prof_n[i];
        p->prof_ms[i] = 0;
        p->prof_n[i] = 0;
    }
    p->profiling = enable != 0;
    return DCS_OK;
}

int dcs_stream_copy(dcs_context *ctx, uint64_t bytes, int reps, double *ms) {
    if (!ctx || !ms || reps < 1 || bytes < 16 * 2048 || (bytes % (16 * 2048)))
        return fail(DCS_ERR_INVALID, "bytes must be a positive multiple of 32 KiB");
    HIPCHK(hipSetDevice(ctx->device));
    u32x4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void **)&a, bytes) != hipSuccess) return fail(DCS_ERR_NOMEM, "stream copy source");
    if (hipMalloc((void **)&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return fail(DCS_ERR_NOMEM, "stream copy destination");
    }
    const u64 n = bytes / 16;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipMemsetAsync(a, 0, bytes, ctx->stream));
    // launch shapes: entries in flight per thread x resident workgroups per CU; the fastest is the ceiling
    struct Shape {
        int u;
        unsigned per_cu;
    };
    const Shape shapes[] = {{2, 8}, {4, 8}, {8, 4}, {8, 8}, {4, 16}};
    double best = 1e30;
    for (const Shape &sh : shapes) {
        const u64 chunks = n / (256ull * sh.u);
        const unsigned g = (unsigned)std::max<u64>(1, std::min<u64>(chunks, (u64)ctx->cus * sh.per_cu));
        auto launch = [&]() {
            switch (sh.u) {
                case 2: k_stream_copy<2><<<g, 256, 0, ctx->stream>>>(a, b, n); break;
                case 4: k_stream_copy<4><<<g, 256, 0, ctx->stream>>>(a, b, n); break;
                default: k_stream_copy<8><<<g, 256, 0, ctx->stream>>>(a, b, n); break;
            }
        };
        launch();  // warm-up
        HIPCHK(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < reps; ++r) launch();
        HIPCHK(hipEventRecord(e1, ctx->stream));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipGetLastError());
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        best = std::min(best, (double)t / reps);
    }
    *ms = best;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return DCS_OK;
}

}  // extern "C"
