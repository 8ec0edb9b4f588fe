/*
 * CPU oracle of the sampling hot path.  TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the
 * CPU-baseline leg of the benchmark use it as the checker; the product never links or calls it.
 *
 * It restates, in plain C with OpenMP over orbits:
 *   oracle_contract        kernel_opt/inplace_plan.py:90-135 (execute_inplace, the eight orbit
 *                          updates) with the scale chain of the GPU driver
 *                          kernel_opt/inplace_gpu.py:107-123 (step j multiplied by 2^-e,
 *                          e = frexp(sqrt(max |out_{j-1}|^2)).exp; max of the stored values); storage
 *                          complex64, complex128 or binary16 pairs (kernel_opt/inplace_half_gpu.py:14-78)
 *   oracle_f32_to_h16,     the binary16 conversions of that storage
 *   oracle_h16_to_f32
 *   oracle_gather          dcs_rankwidth/scan_plan.py:161-195 and dcs_rankwidth/gpu_kernel.py:58-101
 *                          (the canonical gather executor: out[c'] = sum over kernel combinations of
 *                          E[p & mask] * (+-f_y), p = XOR of byte tables ^ combination) with the
 *                          fused power-of-two scale of sampling/conditional_block_gpu.py:92-126
 *   oracle_accumulate      sampling/conditional_block_gpu.py:155-158 (total[smap ^ x] += env * coef)
 *   oracle_transform       sampling/conditional_block_gpu.py:159-171 (Walsh-Hadamard, phase, U)
 *   oracle_draw            sampling/sampling_protocol.py:56-65,81-82 (weights, sequential cumsum,
 *                          normalisation with cdf[-1] = 1, searchsorted side='right')
 *
 * Arithmetic convention: complex products put the left operand first and fuse its real part,
 * x*y = (fma(xr, yr, -(xi*yi)), fma(xr, yi, xi*yr)) -- the rounding numpy 2.2 applies to complex128
 * products on FMA hosts -- and operand order follows the numpy expressions of execute_inplace.
 * Build with -ffp-contract=off so nothing else is fused.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef struct { double r, i; } cd;

static inline cd cmul(cd x, cd y) { cd o = {fma(x.r, y.r, -(x.i * y.i)), fma(x.r, y.i, x.i * y.r)}; return o; }
static inline cd cadd(cd a, cd b) { cd o = {a.r + b.r, a.i + b.i}; return o; }
static inline cd cneg_if(cd a, int s) { if (s) { a.r = -a.r; a.i = -a.i; } return a; }
static inline int sbit(u64 c, u64 beta, int ov) { return (__builtin_popcountll(c & beta) & 1) ^ ov; }

static inline u64 pdep(u64 x, u64 mask) {
    u64 r = 0;
    for (u64 m = mask; m; m &= m - 1) {
        if (x & 1) r |= m & (~m + 1);
        x >>= 1;
    }
    return r;
}

/* IEEE binary16 <-> binary32, bit for bit (the conversions of kernel_opt/inplace_half_gpu.py:14-21:
 * half -> float is exact; float -> half rounds to nearest, ties to even, with subnormal halves and
 * overflow to infinity). */
static inline float h16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) x = sign;
        else {
            const float f = (float)m * 5.9604644775390625e-08f; /* m * 2^-24, exact */
            memcpy(&x, &f, 4);
            x |= sign;
        }
    } else if (e == 31) x = sign | 0x7f800000u | (m << 13);
    else x = sign | ((e + 112u) << 23) | (m << 13);
    float f;
    memcpy(&f, &x, 4);
    return f;
}
static inline uint16_t f32_to_h16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? (0x200u | ((ax >> 13) & 0x3ffu)) : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 rounds to infinity */
    if (ax < 0x38800000u) {                                   /* below 2^-14: subnormal half or zero */
        if (ax < 0x33000000u) return sign;                    /* below 2^-25 (2^-25 itself ties to 0) */
        const uint32_t e = ax >> 23, mant = (ax & 0x7fffffu) | 0x800000u;
        const uint32_t shift = 126u - e; /* value = mant * 2^(e-150) = h * 2^-24 */
        uint32_t h = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1u), halfway = 1u << (shift - 1u);
        if (rem > halfway || (rem == halfway && (h & 1u))) ++h;
        return (uint16_t)(sign | h);
    }
    uint32_t r = ax - 0x38000000u; /* exponent rebias 127 -> 15 */
    r += 0xfffu + ((r >> 13) & 1u);
    return (uint16_t)(sign | (r >> 13));
}

/* storage access: storage 0 = complex64 (float pairs), 1 = complex128, 2 = binary16 pairs */
static inline cd ld(const void *E, int st, u64 c) {
    cd o;
    if (st == 1) { const double *p = (const double *)E + 2 * c; o.r = p[0]; o.i = p[1]; }
    else if (st == 0) { const float *p = (const float *)E + 2 * c; o.r = p[0]; o.i = p[1]; }
    else { const uint16_t *p = (const uint16_t *)E + 2 * c; o.r = h16_to_f32(p[0]); o.i = h16_to_f32(p[1]); }
    return o;
}
/* store v * scale rounded to storage; returns |stored|^2 in storage precision (float for binary16,
 * inplace_half_gpu.py:23-26) */
static inline double put(void *E, int st, u64 c, cd v, double scale) {
    if (st == 1) {
        double *p = (double *)E + 2 * c;
        double r = v.r * scale, i = v.i * scale;
        p[0] = r; p[1] = i;
        return r * r + i * i;
    } else if (st == 0) {
        float *p = (float *)E + 2 * c;
        float r = (float)(v.r * scale), i = (float)(v.i * scale);
        p[0] = r; p[1] = i;
        return (double)(r * r + i * i);
    } else {
        uint16_t *p = (uint16_t *)E + 2 * c;
        p[0] = f32_to_h16((float)(v.r * scale));
        p[1] = f32_to_h16((float)(v.i * scale));
        const float r = h16_to_f32(p[0]), i = h16_to_f32(p[1]);
        return (double)(r * r + i * i);
    }
}

void oracle_f32_to_h16(const float *in, uint16_t *out, u64 n) {
    for (u64 j = 0; j < n; ++j) out[j] = f32_to_h16(in[j]);
}
void oracle_h16_to_f32(const uint16_t *in, float *out, u64 n) {
    for (u64 j = 0; j < n; ++j) out[j] = h16_to_f32(in[j]);
}

static inline double maxnan(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    return a > b ? a : b;
}

/* frexp(sqrt(m2)).exp, as the reference host computes it */
static inline double scale_of(double m2) {
    int e;
    frexp(sqrt(m2), &e);
    return ldexp(1.0, -e);
}

/* contiguous runs of a slot mask: bits [src, src+len) of the orbit index go to [dst, dst+len) */
typedef struct { int n; int src[64], len[64], dst[64]; } runs_t;
static void make_runs(u64 mask, runs_t *r) {
    int src = 0;
    r->n = 0;
    for (int p = 0; p < 64;) {
        if (!((mask >> p) & 1)) { ++p; continue; }
        int q = p;
        while (q < 64 && ((mask >> q) & 1)) ++q;
        r->src[r->n] = src; r->len[r->n] = q - p; r->dst[r->n] = p;
        src += q - p; r->n++; p = q;
    }
}
static inline u64 deposit(u64 x, const runs_t *r) {
    u64 c = 0;
    for (int i = 0; i < r->n; ++i) c |= ((x >> r->src[i]) & ((1ull << r->len[i]) - 1)) << r->dst[i];
    return c;
}

static u64 orbit_slots(int kind, u64 in_mask, int pivot, int pivot_a) {
    u64 m = in_mask;
    if (kind == 1) m &= ~(1ull << pivot_a);
    else if (kind == 3 || kind == 5 || kind == 7) m &= ~(1ull << pivot);
    else if (kind == 6) m &= ~((1ull << pivot) | (1ull << pivot_a));
    return m;
}

/*
 * Run every step over E (2^maxslots entries, caller-owned; only E[0] is read before it is written).
 * coeffs[s*4..] = f0r, f0i, f1r, f1i.  unit_m2[s] receives max |out|^2 of step s.  Returns the index
 * of the first step whose maximum is zero or nonfinite, or nsteps.
 */
int oracle_contract(int storage, int nsteps, const int32_t *kind, const u64 *K, const u64 *Ka,
                    const int32_t *pivot, const int32_t *pivot_a, const u64 *beta, const u64 *in_mask,
                    const double *coeffs, const int32_t *ov, const int32_t *y, void *E, double *unit_m2) {
    if (storage == 1) { ((double *)E)[0] = 1.0; ((double *)E)[1] = 0.0; }
    else if (storage == 0) { ((float *)E)[0] = 1.0f; ((float *)E)[1] = 0.0f; }
    else { ((uint16_t *)E)[0] = 0x3c00u; ((uint16_t *)E)[1] = 0u; }
    double prev = 1.0;
    for (int s = 0; s < nsteps; ++s) {
        const int kd = kind[s];
        const u64 slots = orbit_slots(kd, in_mask[s], pivot[s], pivot_a[s]);
        const long long norb = 1ll << __builtin_popcountll(slots);
        const cd f0 = {coeffs[4 * s], coeffs[4 * s + 1]}, f1 = {coeffs[4 * s + 2], coeffs[4 * s + 3]};
        const int o = ov[s] & 1, yy = y[s] & 1;
        const u64 b = beta[s], k = K[s], ka = Ka[s];
        const int pv = pivot[s];
        const double scale = scale_of(prev);
        runs_t runs;
        make_runs(slots, &runs);
        double m = 0.0;
#pragma omp parallel for schedule(static) reduction(max : m)
        for (long long x = 0; x < norb; ++x) {
            const u64 c = deposit((u64)x, &runs);
            const cd a = ld(E, storage, c);
            const int sc = sbit(c, b, o);
            double lm = 0.0;
            cd o0, o1;
            switch (kd) {
                case 0:
                    o0 = cmul(f0, a);
                    o1 = cneg_if(cmul(f1, a), sc);
                    lm = maxnan(lm, put(E, storage, c, o0, scale));
                    lm = maxnan(lm, put(E, storage, c | (1ull << pv), o1, scale));
                    break;
                case 1: {
                    const u64 d = c ^ ka;
                    const cd bb = ld(E, storage, d);
                    o0 = cadd(cmul(f0, a), cmul(cneg_if(f1, sbit(d, b, o)), bb));
                    o1 = cadd(cmul(f0, bb), cmul(cneg_if(f1, sc), a));
                    lm = maxnan(lm, put(E, storage, c, o0, scale));
                    lm = maxnan(lm, put(E, storage, d, o1, scale));
                } break;
                case 2:
                    lm = put(E, storage, c, cmul(a, cadd(f0, cneg_if(f1, sc))), scale);
                    break;
                case 3: {
                    const u64 d = c ^ k;
                    const cd bb = ld(E, storage, d);
                    o0 = cmul(f0, cadd(a, bb));
                    o1 = cmul(f1, cadd(cneg_if(a, sc), cneg_if(bb, sbit(d, b, o))));
                    lm = maxnan(lm, put(E, storage, c, o0, scale));
                    lm = maxnan(lm, put(E, storage, d, o1, scale));
                } break;
                case 4:
                    lm = put(E, storage, c, cmul(a, yy ? cneg_if(f1, sc) : f0), scale);
                    break;
                case 5: {
                    const u64 d = c ^ k;
                    const cd bb = ld(E, storage, d);
                    o0 = yy ? cmul(f1, cadd(cneg_if(a, sc), cneg_if(bb, sbit(d, b, o)))) : cmul(f0, cadd(a, bb));
                    lm = put(E, storage, c, o0, scale);
                } break;
                case 6: {
                    const u64 c1 = c ^ k, d = c ^ ka, d1 = d ^ k;
                    const cd bb = ld(E, storage, c1), e = ld(E, storage, d), g = ld(E, storage, d1);
                    o0 = cadd(cmul(f0, cadd(a, bb)), cmul(f1, cadd(cneg_if(e, sbit(d, b, o)), cneg_if(g, sbit(d1, b, o)))));
                    o1 = cadd(cmul(f0, cadd(e, g)), cmul(f1, cadd(cneg_if(a, sc), cneg_if(bb, sbit(c1, b, o)))));
                    lm = maxnan(lm, put(E, storage, c, o0, scale));
                    lm = maxnan(lm, put(E, storage, d, o1, scale));
                } break;
                default: {
                    const u64 d = c ^ k;
                    const cd bb = ld(E, storage, d);
                    o0 = cadd(cmul(cadd(f0, cneg_if(f1, sc)), a), cmul(cadd(f0, cneg_if(f1, sbit(d, b, o))), bb));
                    lm = put(E, storage, c, o0, scale);
                } break;
            }
            if (lm != lm) lm = INFINITY; /* the reduction below cannot carry NaN; inf still fails closed */
            if (lm > m) m = lm;
        }
        unit_m2[s] = m;
        if (!(m > 0.0) || !isfinite(m)) {
            for (int t = s + 1; t < nsteps; ++t) unit_m2[t] = 0.0;
            return s;
        }
        prev = m;
    }
    return nsteps;
}

/*
 * Canonical gather executor over two caller-owned buffers A, B of 2^rmax entries.  Step s reads the
 * 2^r_in[s] entries of the current table and writes 2^r_out[s] entries of the other one:
 *   p0 = XOR_b tables[b][byte_b(c')],  for each combination k: p = p0 ^ combos[k], c = p & (2^r_in - 1),
 *   w = f_{y} with y = bit r_in of p (free vertex) or y[s] (fixed), negated when y and
 *   parity(c & beta) ^ ov is odd,  out[c'] = round(scale * sum_k E[c] * w)  (sum in combination order).
 * Tables and combinations are concatenated; tab_off / comb_off index them per step.  Returns the
 * number of the table holding the result (0: A, 1: B) or -1 - s when step s produced a zero or
 * nonfinite maximum.
 */
int oracle_gather(int storage, int nsteps, const int32_t *r_in, const int32_t *r_out, const int32_t *is_free,
                  const int32_t *nbytes, const int64_t *tab_off, const u64 *tables, const int32_t *ncombos,
                  const int64_t *comb_off, const u64 *combos, const u64 *beta, const double *coeffs,
                  const int32_t *ov, const int32_t *y, void *A, void *B, double *unit_m2) {
    void *buf[2] = {A, B};
    int cur = 0;
    if (storage != 0 && storage != 1) return -1; /* the canonical executor has no binary16 storage */
    if (storage) { ((double *)A)[0] = 1.0; ((double *)A)[1] = 0.0; }
    else { ((float *)A)[0] = 1.0f; ((float *)A)[1] = 0.0f; }
    double prev = 1.0;
    for (int s = 0; s < nsteps; ++s) {
        const long long nout = 1ll << r_out[s];
        const u64 mask = (r_in[s] >= 64) ? ~0ull : ((1ull << r_in[s]) - 1);
        const u64 *T = tables + tab_off[s];
        const u64 *cb = combos + comb_off[s];
        const int nb = nbytes[s], nc = ncombos[s], fr = is_free[s], ri = r_in[s], o = ov[s] & 1, yy = y[s] & 1;
        const u64 bt = beta[s];
        const cd f0 = {coeffs[4 * s], coeffs[4 * s + 1]}, f1 = {coeffs[4 * s + 2], coeffs[4 * s + 3]};
        const double scale = scale_of(prev);
        const void *Ein = buf[cur];
        void *Eout = buf[cur ^ 1];
        double m = 0.0;
#pragma omp parallel for schedule(static) reduction(max : m)
        for (long long cpl = 0; cpl < nout; ++cpl) {
            const u64 cp = (u64)cpl;
            u64 p0 = 0;
            for (int b = 0; b < nb; ++b) p0 ^= T[b * 256 + ((cp >> (8 * b)) & 255)];
            cd acc = {0.0, 0.0};
            for (int k = 0; k < nc; ++k) {
                const u64 p = p0 ^ cb[k];
                const u64 c = p & mask;
                const int par = (__builtin_popcountll(c & bt) & 1) ^ o;
                const int yv = fr ? (int)((p >> ri) & 1) : yy;
                const cd w = cneg_if(yv ? f1 : f0, yv & par);
                acc = cadd(acc, cmul(ld(Ein, storage, c), w));
            }
            double lm = put(Eout, storage, cp, acc, scale);
            if (lm != lm) lm = INFINITY;
            if (lm > m) m = lm;
        }
        cur ^= 1;
        unit_m2[s] = m;
        if (!(m > 0.0) || !isfinite(m)) {
            for (int t = s + 1; t < nsteps; ++t) unit_m2[t] = 0.0;
            return -1 - s;
        }
        prev = m;
    }
    return cur;
}

/* total[smap[i] ^ xmask] += env(i) * coef over the 2^r terminal coordinates, env(i) = E[pdep(i, mask)] */
void oracle_accumulate(int storage, const void *E, u64 terminal_mask, int k, const u64 *smap, u64 xmask,
                       double cr, double ci, double *total) {
    const cd coef = {cr, ci};
    (void)k;
    const long long n = 1ll << __builtin_popcountll(terminal_mask);
    for (long long i = 0; i < n; ++i) {
        const cd p = cmul(ld(E, storage, pdep((u64)i, terminal_mask)), coef);
        const u64 j = smap[i] ^ xmask;
        total[2 * j] = total[2 * j] + p.r;
        total[2 * j + 1] = total[2 * j + 1] + p.i;
    }
}

/* Walsh-Hadamard over the k bits (bit 0 first), the +-1 phase, then the open-vertex unitaries
 * U[q] = u00r,u00i,u01r,u01i,u10r,u10i,u11r,u11i applied to bit q (q = 0 first). */
void oracle_transform(int k, double *v, const double *phase, const double *U) {
    const u64 n = 1ull << k;
    for (int q = 0; q < k; ++q)
        for (u64 i0 = 0; i0 < n; ++i0) {
            if (i0 & (1ull << q)) continue;
            const u64 i1 = i0 | (1ull << q);
            const double lr = v[2 * i0], li = v[2 * i0 + 1], hr = v[2 * i1], hi = v[2 * i1 + 1];
            v[2 * i0] = lr + hr; v[2 * i0 + 1] = li + hi;
            v[2 * i1] = lr - hr; v[2 * i1 + 1] = li - hi;
        }
    for (u64 i = 0; i < n; ++i) {
        const cd r = cmul((cd){v[2 * i], v[2 * i + 1]}, (cd){phase[i], 0.0});
        v[2 * i] = r.r; v[2 * i + 1] = r.i;
    }
    for (int q = 0; q < k; ++q) {
        const double *u = U + 8 * q;
        const cd u00 = {u[0], u[1]}, u01 = {u[2], u[3]}, u10 = {u[4], u[5]}, u11 = {u[6], u[7]};
        for (u64 i0 = 0; i0 < n; ++i0) {
            if (i0 & (1ull << q)) continue;
            const u64 i1 = i0 | (1ull << q);
            const cd lo = {v[2 * i0], v[2 * i0 + 1]}, hi = {v[2 * i1], v[2 * i1 + 1]};
            const cd a = cadd(cmul(u00, lo), cmul(u01, hi)), b = cadd(cmul(u10, lo), cmul(u11, hi));
            v[2 * i0] = a.r; v[2 * i0 + 1] = a.i;
            v[2 * i1] = b.r; v[2 * i1 + 1] = b.i;
        }
    }
}

/* out: index, weight, total, lower, upper.  Returns 0, or 5 for a zero / invalid outcome, 4 for a
 * nonfinite total. */
int oracle_draw(const double *v, u64 n, double u, double *cum, double *out) {
    double acc = 0.0;
    for (u64 i = 0; i < n; ++i) {
        const double w = v[2 * i] * v[2 * i] + v[2 * i + 1] * v[2 * i + 1];
        acc = acc + w;
        cum[i] = acc;
    }
    const double total = acc;
    out[2] = total;
    if (!isfinite(total)) return 4;
    if (!(total > 0.0)) return 5;
    u64 lo = 0, hi = n;
    while (lo < hi) {
        const u64 mid = (lo + hi) >> 1;
        const double c = (mid == n - 1) ? 1.0 : cum[mid] / total;
        if (c <= u) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= n) return 5;
    out[0] = (double)lo;
    out[1] = v[2 * lo] * v[2 * lo] + v[2 * lo + 1] * v[2 * lo + 1];
    out[3] = lo ? ((lo - 1 == n - 1) ? 1.0 : cum[lo - 1] / total) : 0.0;
    out[4] = (lo == n - 1) ? 1.0 : cum[lo] / total;
    return out[1] > 0.0 ? 0 : 5;
}

int oracle_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
