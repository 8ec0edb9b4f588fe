"""The oracle pinned: dense state vectors, the independent gather executor, brute-force graph sums,
the reference's draw self-test vectors, and the scale rule."""
import math

import numpy as np
import pytest

from dcs_amd.block_plan import BlockPlan
from dcs_amd.scan_plan import local_vectors, structurally_fixed
from dcs_amd.stabilizer import graph_form, inject, ring_circuit, synthetic_instance
from oracle import oracle


def _block_vs_statevector(n, depth, t, seed, open_vertices, prefixes):
    inst = synthetic_instance(n, depth, t, seed)
    psi = oracle.statevector(inst['ops'], n)
    bp = BlockPlan(inst['graph'], inst['unitaries'], n, inst['natural_order'], [], open_vertices)
    k = len(open_vertices)
    D = 2.0 ** (n - k)
    phase = None
    closed = [q for q in range(n) if q not in open_vertices]
    for pref in prefixes:
        x = np.zeros(n, np.uint8)
        x[closed] = pref
        v = oracle.block(bp, x)
        idx = np.zeros(1 << k, np.int64)
        for A in range(1 << k):
            full = 0
            for q in closed:
                full |= int(x[q]) << q
            for j, q in enumerate(open_vertices):
                full |= ((A >> j) & 1) << q
            idx[A] = full
        want = math.sqrt(D) * psi[idx]
        if phase is None:
            i = int(np.argmax(np.abs(want)))
            phase = v[i] / want[i]
            assert abs(abs(phase) - 1) < 1e-12
        err = np.max(np.abs(v - phase * want)) / np.max(np.abs(want))
        assert err < 1e-12, err


def test_c1_all_amplitudes_match_statevector():
    """16 qubits, depth 12, 20 T: all 2^16 amplitudes (no prefix), complex128."""
    _block_vs_statevector(16, 12, 20, 0, list(range(16)), [np.zeros(0, np.uint8)])


@pytest.mark.parametrize('seed', [1, 2])
def test_prefix_blocks_match_statevector(seed):
    rng = np.random.default_rng(seed)
    opn = [9, 2, 7, 11, 4, 0]
    prefixes = [rng.integers(0, 2, 12 - len(opn), dtype=np.uint8) for _ in range(4)]
    _block_vs_statevector(12, 10, 14, seed, opn, prefixes)


def test_block_norm_is_one_c1():
    inst = synthetic_instance(16, 12, 20, 0)
    bp = BlockPlan(inst['graph'], inst['unitaries'], 16, inst['natural_order'], [], list(range(16)))
    v = oracle.block(bp, np.zeros(16, np.uint8))
    assert abs(np.vdot(v, v).real - 1) < 1e-12


def test_graph_form_against_dense_injected_state():
    """(U_1 x ... x U_N)|G> equals the injected stabiliser state up to one phase (N = 10)."""
    ops = ring_circuit(6, 3, 4, seed=5)
    tab, N, natural, _ = inject(ops, 6)
    edges, U = graph_form(tab)
    # dense injected circuit on N qubits
    injected = []
    k = 6
    for op in ops:
        if op[0] == 'T':
            injected.append(('c1', k, 'H'))     # CNOT(q, k) = H_k CZ(q, k) H_k
            injected.append(('cz', op[1], k))
            injected.append(('c1', k, 'H'))
            k += 1
        else:
            injected.append(op)
    S = oracle.statevector(injected, N)
    A = np.zeros((N, N), np.uint8)
    for a, b in edges:
        A[a, b] = A[b, a] = 1
    ph = None
    for x in range(0, 1 << N, 7):
        bras = [np.eye(2)[(x >> q) & 1] for q in range(N)]
        g = oracle.graph_amplitude(A, U, bras)
        if abs(S[x]) > 1e-9:
            if ph is None:
                ph = g / S[x]
            assert abs(g - ph * S[x]) < 1e-12
        else:
            assert abs(g) < 1e-12


def test_inplace_oracle_equals_gather_executor():
    """Single amplitude (no open vertices) by the in-place oracle and by the gather executor."""
    inst = synthetic_instance(10, 8, 12, 4)
    n = 10
    x = np.array([1, 0, 1, 1, 0, 0, 1, 0, 1, 1], np.uint8)
    bp = BlockPlan(inst['graph'], inst['unitaries'], n, inst['natural_order'], [], [])
    f = bp.local_vectors(x)
    fr = {bp.relabel[v]: f[v] for v in bp.keep}
    yfix = {q: (0 if abs(fr[q][0]) > 1e-12 else 1) for q in bp.fixed}
    r = oracle.contract_inplace(bp, fr, yfix)
    val_inplace = complex(r['E'][0]) * math.exp(r['ls'])
    g, ls = oracle.gather_amplitude(bp.N, bp.edges, bp.order, bp.fixed, fr, yfix)
    val_gather = g * math.exp(ls)
    assert abs(val_inplace - val_gather) <= 1e-13 * abs(val_gather)
    psi = oracle.statevector(inst['ops'], n)
    want = psi[sum(int(b) << q for q, b in enumerate(x))]
    assert abs(abs(val_inplace) * 2 ** (-bp.N_full / 2) - abs(want)) < 1e-13


def test_gather_oracle_block_c1_matches_statevector_and_inplace():
    """The canonical gather executor's block (C1, all 2^16 amplitudes, complex128) equals the in-place
    block to 1e-14 and the dense state vector up to one global phase; complex64 tables stay within 1e-6."""
    inst = synthetic_instance(16, 12, 20, 0)
    bp = BlockPlan(inst['graph'], inst['unitaries'], 16, inst['natural_order'], [], list(range(16)))
    x = np.zeros(16, np.uint8)
    g = oracle.gather_block(bp, x)
    a = oracle.block(bp, x)
    assert np.max(np.abs(g - a)) / np.max(np.abs(a)) < 1e-14
    psi = oracle.statevector(inst['ops'], 16)
    i = int(np.argmax(np.abs(psi)))
    assert np.max(np.abs(g - (g[i] / psi[i]) * psi)) / np.max(np.abs(g)) < 1e-12
    g64 = oracle.gather_block(bp, x, storage=0)
    assert np.max(np.abs(g64 - a)) / np.max(np.abs(a)) < 1e-6


def test_gather_oracle_prefix_and_parity_slices_match_inplace():
    """Prefix blocks (6 open of 12) and a parity-sliced graph (Y, Z and X operations): gather and in-place
    executors agree to 1e-13."""
    inst = synthetic_instance(12, 10, 14, 1)
    bp = BlockPlan(inst['graph'], inst['unitaries'], 12, inst['natural_order'], [], [9, 2, 7, 11, 4, 0])
    for seed in range(3):
        x = np.random.default_rng(seed).integers(0, 2, 12, dtype=np.uint8)
        a = oracle.block(bp, x)
        assert np.max(np.abs(oracle.gather_block(bp, x) - a)) / np.max(np.abs(a)) < 1e-13
    inst = synthetic_instance(10, 10, 12, seed=2)
    refs = [v for v in range(10, inst['N']) if inst['graph'][v].any()]
    nb = [int(u) for u in np.flatnonzero(inst['graph'][refs[2]])]
    ops = [('Y', refs[0], None), ('Z', refs[1], None), ('X', refs[2], nb[0])]
    bp = BlockPlan(inst['graph'], inst['unitaries'], 10, inst['natural_order'], ops, [0, 3, 6, 8])
    x = np.array([1, 0, 0, 1, 1, 0, 0, 1, 0, 1], np.uint8)
    a = oracle.block(bp, x)
    assert len(bp.slices) == 3
    assert np.max(np.abs(oracle.gather_block(bp, x) - a)) / np.max(np.abs(a)) < 1e-13


def test_reference_draw_selftest_vectors():
    """The CDF self-test of sampling_protocol.py:103-111: midpoint of every nonzero interval."""
    from dcs_amd.sampling_protocol import draw_checked
    for weights in [np.array([0., 1., 0., 2., 3., 0., 4., 0.]), np.array([1e-100, 1., .5, 0., 7., 2e-100, 1., .1])]:
        v = np.sqrt(weights) * np.exp(1j * np.arange(len(weights)))
        cdf = np.cumsum(weights) / np.sum(weights)
        for idx in np.flatnonzero(weights):
            lo = cdf[idx - 1] if idx else 0.
            if cdf[idx] == lo:
                continue
            u = (lo + cdf[idx]) / 2
            assert draw_checked(v, u)['A_index'] == idx
            assert oracle.draw(v, u)['A_index'] == idx


def test_oracle_draw_equals_host_draw_random():
    from dcs_amd.sampling_protocol import draw_checked
    rng = np.random.default_rng(7)
    for n in (1, 2, 64, 1000, 4096):
        m = 1 << int(np.ceil(np.log2(n))) if n > 1 else 1
        v = rng.standard_normal(m) + 1j * rng.standard_normal(m)
        v[rng.random(m) < 0.2] = 0
        if not np.any(v):
            v[0] = 1
        for u in rng.random(20):
            h = draw_checked(v, u)
            o = oracle.draw(v, u)
            assert o['A_index'] == h['A_index']
            assert o['cdf_upper_numpy'] == h['cdf_upper_numpy'] and o['cdf_lower_numpy'] == h['cdf_lower_numpy']
            assert o['scaled_mass_cdf'] == h['scaled_mass_cdf']


def test_device_scale_rule_equals_host_frexp_sqrt():
    """The kernels take e = ceil(E2/2) from m2 = M 2^E2; the host takes frexp(sqrt(m2)).exp."""
    rng = np.random.default_rng(3)
    vals = list(np.exp(rng.uniform(-700, 700, 20000))) + [1.0, 0.5, 2.0, 4.0, 0.25, 5e-324, 1e-310,
                                                           float(np.nextafter(1, 0)), float(np.nextafter(4, 0)),
                                                           float(np.nextafter(4, 8)), 1.7976931348623157e308]
    vals += [float(np.float32(v)) for v in rng.uniform(0, 10, 1000)]
    for m2 in vals:
        E2 = math.frexp(m2)[1]
        assert (E2 + 1) >> 1 == math.frexp(math.sqrt(m2))[1], m2


def test_structurally_fixed_reference_rule():
    inst = synthetic_instance(8, 6, 6, 2)
    U = {q: inst['unitaries'][q] for q in range(inst['N'])}
    fx = structurally_fixed(U, inst['N'], 8)
    f = local_vectors(U, inst['N'], 8, np.zeros(8, np.uint8))
    for q in fx:
        assert min(abs(f[q][0]), abs(f[q][1])) < 1e-12


def test_binary16_conversions_match_numpy():
    """The oracle's binary16 conversions (the half-storage engine's cvt.f32.f16 / cvt.rn.f16.f32,
    kernel_opt/inplace_half_gpu.py:14-21) against numpy's float16: every half exactly, and float32 ->
    half on every tie and neighbour of every half interval plus a strided sweep of all float32 patterns
    (normal, subnormal, overflow, infinities, NaN)."""
    h = np.arange(1 << 16, dtype=np.uint32).astype(np.uint16)
    got = oracle.h16_to_f32(h)
    want = h.view(np.float16).astype(np.float32)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got.view(np.uint32)[~nan], want.view(np.uint32)[~nan])
    # midpoints between consecutive finite halves (exact in float32) and one float32 ulp either side
    fin = want[np.isfinite(want) & (want >= 0)]
    fin = np.unique(fin)
    mids = ((fin[:-1].astype(np.float64) + fin[1:]) / 2).astype(np.float32)
    probes = np.concatenate([mids, np.nextafter(mids, np.float32(0)), np.nextafter(mids, np.float32(np.inf)),
                             fin, np.float32([65504, 65519.996, 65520, 65536, 1e-8, 2.0 ** -25, 2.0 ** -24,
                                              3 * 2.0 ** -26, 0.0, np.inf, np.nan])])
    strided = (np.arange(0, 1 << 32, 65537, dtype=np.uint64)).astype(np.uint32).view(np.float32)
    x = np.concatenate([probes, -probes, strided])
    with np.errstate(over='ignore', invalid='ignore'):
        want16 = x.astype(np.float16)
    got16 = oracle.f32_to_h16(x)
    wn = np.isnan(want16)
    assert np.array_equal(np.isnan(got16.view(np.float16)), wn)
    assert np.array_equal(got16[~wn], want16.view(np.uint16)[~wn])


def test_binary16_storage_block_tracks_statevector_c1():
    """Half-storage contraction of C1 (16q/d12/20T, all 2^16 amplitudes): within binary16 accuracy of
    the dense state vector (rounding 2^-11 per stored part per step, about 20 steps deep), unit total
    probability within the same accuracy, and the complex64 path strictly closer."""
    inst = synthetic_instance(16, 12, 20, seed=0)
    psi = oracle.statevector(inst['ops'], 16)
    bp = BlockPlan(inst['graph'], inst['unitaries'], 16, inst['natural_order'], [], list(range(16)))
    x = np.zeros(16, np.uint8)
    errs = {}
    for storage in (0, 2):
        v = oracle.block(bp, x, storage=storage)
        i = int(np.argmax(np.abs(psi)))
        ph = v[i] / psi[i]
        errs[storage] = np.max(np.abs(v - ph * psi)) / np.max(np.abs(psi))
        assert abs(np.vdot(v, v).real - 1) < (1e-3 if storage == 2 else 1e-6)
    assert errs[2] < 1e-3, errs          # observed 1.1e-4
    assert errs[0] < 1e-6 and errs[0] < errs[2] / 100, errs
