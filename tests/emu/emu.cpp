// Runs the generated coset-group kernels of a plan on the CPU (through shim.hpp) and compares them
// with the per-step orbit formulas of dcs_kernels.hpp applied one step at a time.
//
//   emu plan.txt groups.txt gridcap [classes.txt tol]
//
// plan.txt: "nsteps maxslots terminal_mask", then one line per step: kind K Ka pivot pivot_a beta
// in_mask out_mask.  groups.txt: one line per group segment: kernel first count threads.  The
// including build provides kernels.inc (the generated source, adapted by the test) and table.inc
// (the kernel entry points in index order).
//
// Per group: a random frontier, random step parameters; the kernel's entries (live after the group)
// and its recorded maximum |a|^2 must equal the step-by-step result scaled by the kernel's unit scale.
// Chain (when the groups cover every step): the whole contraction from |0...0> with the scale chain
// carried from group to group, as dcs_contract runs it.  Exit status 0 iff everything is equal.
//
// With classes.txt (one coefficient class per step, for f0-factored kernels) the step parameters are
// drawn to fit the classes (f1 = f0 r with r real, imaginary or complex; class 0 steps unconstrained,
// f0 sometimes zero), the factored parameters g0, g1 are derived here from the definition (each group:
// r = f1 / f0 for its class > 0 steps but the last, whose coefficients become (P f0, P f1) with P the
// product of the other f0s), and entries and maxima must agree within tol of the largest magnitude:
// those kernels round to storage once per register pass, not per step.
#include "shim.hpp"

#include <array>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>

#include "kernels.inc"

typedef void (*GroupFn)(St *, u64, const int *, const Dyn *, int, u64 *, int, int);

static std::mt19937_64 &rng_global() {
    static std::mt19937_64 rng(7);
    return rng;
}
static GroupFn kTable[] = {
#include "table.inc"
};

struct StepH {
    int kind, pivot, pivot_a;
    u64 K, Ka, beta, in_mask, out_mask, slots;
};

static u64 orbit_slots_of(const StepH &s) {
    u64 m = s.in_mask;
    if (s.kind == 1) m &= ~(1ull << s.pivot_a);
    else if (s.kind == 3 || s.kind == 5 || s.kind == 7) m &= ~(1ull << s.pivot);
    else if (s.kind == 6) m &= ~((1ull << s.pivot) | (1ull << s.pivot_a));
    return m;
}

// one step over every orbit, rounding to storage after the step
static void ref_step(St *E, const StepH &a, const Dyn &dd) {
    const cd f0{dd.f0r, dd.f0i}, f1{dd.f1r, dd.f1i};
    const int ov = dd.ov, y = dd.y;
    const u64 n = 1ull << __builtin_popcountll(a.slots);
    auto put = [&](u64 c, cd v) { E[c] = Store<St>::round(v); };
    for (u64 x = 0; x < n; ++x) {
        const u64 c = pdep_loop(x, a.slots);
        const cd va = Store<St>::load(E[c]);
        const int sc = sign_bit(c, a.beta, ov);
        switch (a.kind) {
            case 0: {
                cd o0, o1;
                orb_grow(va, f0, f1, sc, o0, o1);
                put(c, o0);
                put(c | (1ull << a.pivot), o1);
            } break;
            case 1: {
                const u64 d = c ^ a.Ka;
                cd o0, o1;
                orb_mix_pair(va, Store<St>::load(E[d]), f0, f1, sc, sign_bit(d, a.beta, ov), o0, o1);
                put(c, o0);
                put(d, o1);
            } break;
            case 2: put(c, orb_mix_diag(va, f0, f1, sc)); break;
            case 3: {
                const u64 d = c ^ a.K;
                cd o0, o1;
                orb_replace(va, Store<St>::load(E[d]), f0, f1, sc, sign_bit(d, a.beta, ov), o0, o1);
                put(c, o0);
                put(d, o1);
            } break;
            case 4: put(c, orb_fixed(va, f0, f1, sc, y)); break;
            case 5: {
                const u64 d = c ^ a.K;
                put(c, orb_fixed_shrink(va, Store<St>::load(E[d]), f0, f1, sc, sign_bit(d, a.beta, ov), y));
            } break;
            case 6: {
                const u64 c1 = c ^ a.K, d = c ^ a.Ka, d1 = d ^ a.K;
                cd o0, o1;
                orb_shrink_quad(va, Store<St>::load(E[c1]), Store<St>::load(E[d]), Store<St>::load(E[d1]), f0, f1, sc,
                                sign_bit(c1, a.beta, ov), sign_bit(d, a.beta, ov), sign_bit(d1, a.beta, ov), o0, o1);
                put(c, o0);
                put(d, o1);
            } break;
            default: {
                const u64 d = c ^ a.K;
                put(c, orb_shrink_pair(va, Store<St>::load(E[d]), f0, f1, sc, sign_bit(d, a.beta, ov)));
            } break;
        }
    }
}

static void random_dyn(std::vector<Dyn> &dyn, std::mt19937_64 &rng) {
    std::uniform_real_distribution<double> U(-1, 1);
    for (auto &d : dyn) {
        d.f0r = U(rng);
        d.f0i = U(rng);
        d.f1r = U(rng);
        d.f1i = U(rng);
        d.ov = (int)(rng() & 1);
        d.y = (int)(rng() & 1);
    }
}

// coefficients fitting the classes (cls empty: random_dyn)
static void class_dyn(std::vector<Dyn> &dyn, const std::vector<int> &cls, std::mt19937_64 &rng) {
    random_dyn(dyn, rng);
    std::uniform_real_distribution<double> U(-1, 1);
    for (size_t s = 0; s < dyn.size(); ++s) {
        Dyn &d = dyn[s];
        const double a = U(rng), b = U(rng);
        cd r{a, b};
        if (cls[s] == 1) r = cd{a, 0.0};
        else if (cls[s] == 2) r = cd{0.0, b};
        else if (cls[s] == 0) {
            if ((rng() & 3) == 0) d.f0r = d.f0i = 0.0;  // an unfactored step may have f0 == 0
            continue;
        }
        d.f1r = d.f0r * r.r - d.f0i * r.i;
        d.f1i = d.f0r * r.i + d.f0i * r.r;
    }
}

// g0, g1 of the factored kernels for the group segment [first, first + count)
static void factor_params(std::vector<Dyn> &dyn, const std::vector<int> &cls, int first, int count) {
    std::complex<double> P(1.0, 0.0);
    const int last = first + count - 1;
    for (int s = first; s < last; ++s) {
        if (!cls[s]) continue;
        const std::complex<double> f0(dyn[s].f0r, dyn[s].f0i), f1(dyn[s].f1r, dyn[s].f1i);
        const std::complex<double> r = f1 / f0;
        dyn[s].g1r = cls[s] == 2 ? 0.0 : r.real();
        dyn[s].g1i = cls[s] == 1 ? 0.0 : r.imag();
        P *= f0;
    }
    const std::complex<double> f0(dyn[last].f0r, dyn[last].f0i), f1(dyn[last].f1r, dyn[last].f1i);
    const std::complex<double> a = P * f0, b = P * f1;
    dyn[last].g0r = a.real();
    dyn[last].g0i = a.imag();
    dyn[last].g1r = b.real();
    dyn[last].g1i = b.imag();
}

// entries that differ (tol 0: bit for bit; else beyond tol times the largest reference magnitude)
static long compare(const std::vector<St> &E, const std::vector<St> &R, u64 live, double &m2, double tol = 0.0) {
    long nbad = 0;
    m2 = 0;
    for (u64 a = 0; a < (u64)E.size(); ++a)
        if (!(a & ~live)) m2 = std::max(m2, (double)Store<St>::norm2(R[a]));
    const double lim = tol * std::sqrt(m2);
    for (u64 a = 0; a < (u64)E.size(); ++a) {
        if (a & ~live) continue;
        const cd x = Store<St>::load(E[a]), y = Store<St>::load(R[a]);
        if (tol == 0.0 ? (x.r != y.r || x.i != y.i) : !(std::hypot(x.r - y.r, x.i - y.i) <= lim)) ++nbad;
    }
    return nbad;
}

static bool max_ok(u64 dmax, double m2, double tol) {
    const double got = __longlong_as_double((long long)dmax);
    return tol == 0.0 ? got == m2 : std::fabs(got - m2) <= 2.0 * tol * m2;
}

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: emu plan.txt groups.txt gridcap\n");
        return 2;
    }
    std::ifstream f(argv[1]);
    const unsigned gridcap = (unsigned)atoi(argv[3]);
    int n, M;
    unsigned long long tmask;
    f >> n >> M >> tmask;
    std::vector<StepH> steps(n);
    for (auto &h : steps) {
        unsigned long long K, Ka, b, im, om;
        f >> h.kind >> K >> Ka >> h.pivot >> h.pivot_a >> b >> im >> om;
        h.K = K;
        h.Ka = Ka;
        h.beta = b;
        h.in_mask = im;
        h.out_mask = om;
        h.slots = orbit_slots_of(h);
    }
    std::vector<std::array<int, 4>> gs;
    {
        std::ifstream gf(argv[2]);
        int g, first, count, thr;
        while (gf >> g >> first >> count >> thr) gs.push_back({g, first, count, thr});
    }
    std::vector<int> cls;
    double tol = 0.0;
    if (argc >= 6) {
        std::ifstream cf(argv[4]);
        int c;
        while (cf >> c) cls.push_back(c);
        if ((int)cls.size() != n) {
            fprintf(stderr, "classes.txt holds %zu classes for %d steps\n", cls.size(), n);
            return 2;
        }
        tol = atof(argv[5]);
    }
    auto make_dyn = [&](std::vector<Dyn> &dyn) {
        if (cls.empty()) {
            random_dyn(dyn, rng_global());
            return;
        }
        class_dyn(dyn, cls, rng_global());
        for (const auto &s : gs) factor_params(dyn, cls, s[1], s[2]);
    };
    const u64 L = 1ull << M;
    std::mt19937_64 &rng = rng_global();
    int bad = 0;
    // each group on its own, from a random frontier (dead entries hold values the kernel must ignore)
    for (const auto &s : gs) {
        const int g = s[0], first = s[1], count = s[2], thr = s[3];
        std::uniform_real_distribution<double> U(-1, 1);
        std::vector<St> E(L);
        for (auto &e : E) e = Store<St>::round(cd{U(rng), U(rng)});
        std::vector<Dyn> dyn(n);
        make_dyn(dyn);
        std::vector<St> R = E;
        for (int k = first; k < first + count; ++k) ref_step(R.data(), steps[k], dyn[k]);
        // unit 0 takes the scale of a unit maximum
        const double scale = scale_from_m2(1.0);
        for (auto &r : R) {
            const cd v = Store<St>::load(r);
            r = Store<St>::round(cd{v.r * scale, v.i * scale});
        }
        int lane = 0;
        u64 dmax = 0;
        launch([&] { kTable[g](E.data(), L, &lane, dyn.data(), n, &dmax, 1, 0); }, gridcap, 1, thr);
        double m2;
        const long nbad = compare(E, R, steps[first + count - 1].out_mask, m2, tol);
        const bool maxok = max_ok(dmax, m2, tol);
        printf("group %d steps %d..%d: mismatches %ld, max %s\n", g, first, first + count - 1, nbad, maxok ? "ok" : "BAD");
        if (nbad || !maxok) ++bad;
    }
    // the chain from |0...0>, when the groups cover the steps in order
    bool covers = !gs.empty() && gs[0][1] == 0;
    for (size_t i = 0; i < gs.size(); ++i)
        covers = covers && (i + 1 == gs.size() ? gs[i][1] + gs[i][2] == n : gs[i][1] + gs[i][2] == gs[i + 1][1]);
    if (covers) {
        std::vector<St> E(L);
        for (auto &e : E) e = Store<St>::round(cd{NAN, NAN});  // every entry not yet born is dead
        E[0] = Store<St>::round(cd{1.0, 0.0});
        std::vector<St> R = E;
        std::vector<Dyn> dyn(n);
        make_dyn(dyn);
        const int G = (int)gs.size();
        std::vector<u64> dmax(G, 0);
        int lane = 0;
        double scale = scale_from_m2(1.0);
        for (int u = 0; u < G; ++u) {
            launch([&] { kTable[gs[u][0]](E.data(), L, &lane, dyn.data(), n, dmax.data(), G, u); }, gridcap, 1, gs[u][3]);
            for (int k = gs[u][1]; k < gs[u][1] + gs[u][2]; ++k) ref_step(R.data(), steps[k], dyn[k]);
            const u64 live = steps[gs[u][1] + gs[u][2] - 1].out_mask;
            for (u64 a = 0; a < L; ++a)
                if (!(a & ~live)) {
                    const cd v = Store<St>::load(R[a]);
                    R[a] = Store<St>::round(cd{v.r * scale, v.i * scale});
                }
            double m2;
            const long nbad = compare(E, R, live, m2, tol);
            const bool maxok = max_ok(dmax[u], m2, tol);
            printf("chain unit %d group %d: mismatches %ld, max %s\n", u, gs[u][0], nbad, maxok ? "ok" : "BAD");
            if (nbad || !maxok) ++bad;
            // the kernels' own maximum sets the next scale (within tol of the reference's it could
            // cross a power of four where the reference's does not)
            scale = scale_from_m2(tol == 0.0 ? m2 : __longlong_as_double((long long)dmax[u]));
        }
    } else {
        printf("chain skipped: the groups do not cover every step\n");
    }
    printf("RESULT %s (%d bad)\n", bad ? "BAD" : "ok", bad);
    return bad ? 1 : 0;
}
