"""The C oracle's reduced-precision storage modes against an independent NumPy restatement.

The restatement follows the reference's NumPy orbit executor (kernel_opt/inplace_plan.py:80-135: the
same addresses, signs and per-kind expressions, evaluated by numpy on complex128 values) and replaces its
renormalisation by the GPU engines' storage and scale chain:

  * every output is multiplied by scale = 2^-e, e = frexp(sqrt(previous max))[1]
    (kernel_opt/inplace_gpu.py:115-123), then rounded to storage:
      complex64  -- (float)(x * scale) per part (kernel_opt/inplace_gpu.py:18);
      binary16   -- (float)(x * scale), then float -> half with round to nearest even, per part
                    (kernel_opt/inplace_half_gpu.py:23-24: cvt.rn.f16.f32 after the float cast);
  * the step's maximum is |stored|^2 = r*r + i*i evaluated in float32 on the stored parts
    (inplace_half_gpu.py:25; inplace_gpu.py:18 for complex64);
  * reads widen the stored parts exactly (half -> float -> double).

The float32 and float16 roundings are numpy's astype conversions, which share no code with the oracle's
bit-level conversions.  Misplaced roundings -- the maximum taken before rounding, binary16 entries left at
float precision, a complex64 |stored|^2 summed in FP64 -- change the bits on these plans (checked
below), so bit-for-bit agreement pins where the oracle rounds.  (Rounding binary16 straight from FP64 instead of through float differs only where the two
roundings meet a halfway case, which these plans never hit; the float step is fixed by the conversion
tests, which check the oracle's float -> half against numpy's on every half and every tie.)"""
import math

import numpy as np
import pytest

from dcs_amd.block_plan import BlockPlan
from dcs_amd.stabilizer import synthetic_instance
from oracle import oracle


def addresses(active, omit=()):
    slots = [p for p in active if p not in omit]
    x = np.arange(1 << len(slots), dtype=np.uint64)
    out = np.zeros_like(x)
    for j, p in enumerate(slots):
        out |= ((x >> np.uint64(j)) & np.uint64(1)) << np.uint64(p)
    return out


def signs(x, beta, ov):
    par = np.array([bin(int(v) & beta).count('1') & 1 for v in x], np.int64)
    return 1 - 2 * (par ^ ov)


def store(v, scale, storage):
    """Round v * scale to storage; returns (stored values widened to complex128, |stored|^2 in float32)."""
    re = (v.real * scale).astype(np.float32)
    im = (v.imag * scale).astype(np.float32)
    if storage == 2:
        re = re.astype(np.float16).astype(np.float32)
        im = im.astype(np.float16).astype(np.float32)
    m2 = re * re + im * im          # float32 products and sum, each rounded
    return re.astype(np.float64) + 1j * im.astype(np.float64), m2


def restated_contract(bp, f, yfix, storage):
    """Terminal frontier (physical entries), per-step maxima and final offset of one slice."""
    st_list = bp.inplace_steps
    E = np.zeros(1 << bp.inplace_maxslots, np.complex128)
    E[0] = 1
    offset, prev, maxima = 0, 1.0, []
    for st in st_list:
        ov = (offset >> st.j) & 1
        offset &= ~(1 << st.j)
        f0, f1 = map(complex, f[st.v])
        y = int(yfix[st.v]) if not st.free else 0
        if y:
            offset ^= bp.fut[st.j]
        scale = math.ldexp(1., -math.frexp(math.sqrt(prev))[1])
        out = []                         # (addresses, values) written by this step
        kind = st.kind
        if kind == 'grow':
            c = addresses(st.active_in)
            a = E[c].copy()
            out += [(c, f0 * a), (c | np.uint64(1 << st.pivot), f1 * a * signs(c, st.beta, ov))]
        elif kind == 'mix' and st.Ka:
            c = addresses(st.active_in, (st.pivot_a,))
            d = c ^ np.uint64(st.Ka)
            a, b = E[c].copy(), E[d].copy()
            s0, s1 = signs(c, st.beta, ov), signs(d, st.beta, ov)
            out += [(c, f0 * a + f1 * s1 * b), (d, f0 * b + f1 * s0 * a)]
        elif kind == 'mix':
            c = addresses(st.active_in)
            out.append((c, E[c] * (f0 + f1 * signs(c, st.beta, ov))))
        elif kind == 'replace':
            c = addresses(st.active_in, (st.pivot,))
            d = c ^ np.uint64(st.K)
            a, b = E[c].copy(), E[d].copy()
            s0, s1 = signs(c, st.beta, ov), signs(d, st.beta, ov)
            out += [(c, f0 * (a + b)), (d, f1 * (s0 * a + s1 * b))]
        elif kind == 'fixed':
            c = addresses(st.active_in)
            out.append((c, E[c] * (f1 * signs(c, st.beta, ov) if y else f0)))
        elif kind == 'fixed_shrink':
            c = addresses(st.active_in, (st.pivot,))
            d = c ^ np.uint64(st.K)
            a, b = E[c].copy(), E[d].copy()
            out.append((c, f1 * (signs(c, st.beta, ov) * a + signs(d, st.beta, ov) * b) if y else f0 * (a + b)))
        elif kind == 'shrink' and st.Ka:
            c = addresses(st.active_in, (st.pivot, st.pivot_a))
            d = c ^ np.uint64(st.Ka)
            c1, d1 = c ^ np.uint64(st.K), d ^ np.uint64(st.K)
            a, b, e, g = E[c].copy(), E[c1].copy(), E[d].copy(), E[d1].copy()
            out += [(c, f0 * (a + b) + f1 * (signs(d, st.beta, ov) * e + signs(d1, st.beta, ov) * g)),
                    (d, f0 * (e + g) + f1 * (signs(c, st.beta, ov) * a + signs(c1, st.beta, ov) * b))]
        elif kind == 'shrink':
            c = addresses(st.active_in, (st.pivot,))
            d = c ^ np.uint64(st.K)
            a, b = E[c].copy(), E[d].copy()
            out.append((c, (f0 + f1 * signs(c, st.beta, ov)) * a + (f0 + f1 * signs(d, st.beta, ov)) * b))
        else:
            raise ValueError(kind)
        m = np.float32(0)
        for addr, val in out:
            stored, m2 = store(val, scale, storage)
            E[addr] = stored
            m = max(m, m2.max())
        maxima.append(float(m))
        if not m > 0:
            break
        prev = float(m)
    return E, np.array(maxima), offset


CASES = [(10, 10, 14, 3, [0, 4, 7]), (11, 12, 16, 1, [1, 5, 9, 10]), (12, 8, 10, 3, list(range(12)))]


@pytest.mark.parametrize('storage', [2, 0])
@pytest.mark.parametrize('n,depth,t,seed,opn', CASES)
def test_oracle_reduced_storage_equals_numpy_restatement(n, depth, t, seed, opn, storage):
    inst = synthetic_instance(n, depth, t, seed=seed)
    bp = BlockPlan(inst['graph'], inst['unitaries'], n, inst['natural_order'], [], opn)
    assert bp.inplace_maxslots <= 12
    rng = np.random.default_rng(seed)
    checked = 0
    for _ in range(3):
        x = rng.integers(0, 2, n, dtype=np.uint8)
        for _, _, f, yfix, _ in bp.slice_assignments(bp.local_vectors(x)):
            if f is None:
                continue
            want_E, want_m2, want_off = restated_contract(bp, f, yfix, storage)
            got = oracle.contract_inplace(bp, f, yfix, storage=storage)
            assert got['offset'] == want_off
            n_ok = len(want_m2)
            assert np.array_equal(got['m2'][:n_ok], want_m2)
            live = addresses(bp.inplace_steps[-1].active_out).astype(np.int64)
            E = oracle.as_complex(got['E'], storage).astype(np.complex128)
            assert np.array_equal(E[live].view(np.float64), want_E[live].view(np.float64))
            checked += 1
    assert checked >= 3


@pytest.mark.parametrize('storage,mistake', [(2, 'max_before_rounding'), (2, 'float_only'), (0, 'max_in_fp64')])
def test_misplaced_rounding_is_detected(storage, mistake):
    """Plausible misplacements of the storage rounding change bits the oracle produces: the step maximum
    taken on the unrounded FP64 values, binary16 entries kept at float precision, and the complex64
    maximum |stored|^2 summed in FP64 instead of float32."""
    n, depth, t, seed, opn = CASES[0]
    inst = synthetic_instance(n, depth, t, seed=seed)
    bp = BlockPlan(inst['graph'], inst['unitaries'], n, inst['natural_order'], [], opn)
    global store
    good = store

    def wrong(v, scale, st):
        stored, m2 = good(v, scale, st)
        if mistake == 'max_before_rounding':
            w = v * scale
            return stored, (w.real * w.real + w.imag * w.imag).astype(np.float32)
        if mistake == 'float_only':
            return good(v, scale, 0)
        return stored, stored.real * stored.real + stored.imag * stored.imag

    rng = np.random.default_rng(0)
    differs = False
    try:
        store = wrong
        for _ in range(4):
            x = rng.integers(0, 2, n, dtype=np.uint8)
            for _, _, f, yfix, _ in bp.slice_assignments(bp.local_vectors(x)):
                if f is None:
                    continue
                want_E, want_m2, _ = restated_contract(bp, f, yfix, storage)
                got = oracle.contract_inplace(bp, f, yfix, storage=storage)
                if not np.array_equal(got['m2'][:len(want_m2)], want_m2):
                    differs = True
    finally:
        store = good
    assert differs
