"""The per-prefix host bookkeeping of BlockPlan (local vectors, slice assignments, step parameters) is
tabulated per plan; every value must stay bit-identical to the direct per-prefix computation, restated
here as the loops of sampling/conditional_block_gpu.py:139-155 and kernel_opt/inplace_gpu.py:110-115."""
import itertools

import numpy as np
import pytest

from dcs_amd.block_plan import BlockPlan
from dcs_amd.sampling_protocol import make_protocol
from dcs_amd.scan_plan import local_vectors
from dcs_amd.stabilizer import synthetic_instance
from dcs_amd.width_planner import best_rotation, fit_width


def direct_slices(bp, f_full, part=0, parts=1):
    count = 0
    for idx, bits in enumerate(itertools.product((0, 1), repeat=len(bp.slices))):
        if idx % parts != part:
            continue
        count += 1
        values = dict(zip(bp.slices, bits))
        coef = complex(np.prod([f_full[s][values[s]] for s in bp.slices]))
        coef *= (-1) ** sum(values[a] * values[b] for a, b in bp.slice_edges)
        if not coef:
            yield count, coef, None, None, 0
            continue
        flips = np.zeros(bp.N, np.uint8)
        for s in bp.slices:
            if values[s]:
                for v in bp.nbrs_in_keep[s]:
                    flips[v] ^= 1
        f = {bp.relabel[v]: np.asarray(f_full[v]) * np.array([1, 1 - 2 * int(flips[bp.relabel[v]])]) for v in bp.keep}
        yfix = {q: (0 if abs(f[q][0]) > 1e-12 else 1) for q in bp.fixed}
        freeflip = sum(int(flips[q]) << j for j, q in enumerate(bp.open_red))
        yield count, coef, f, yfix, freeflip


def direct_steps(bp, f, yfix):
    m = len(bp.inplace_steps)
    coeffs = np.zeros((m, 4), np.float64)
    ov = np.zeros(m, np.int32)
    ys = np.zeros(m, np.int32)
    offset = 0
    for i, st in enumerate(bp.inplace_steps):
        ov[i] = (offset >> st.j) & 1
        offset &= ~(1 << st.j)
        if not st.free:
            y = int(yfix[st.v])
            ys[i] = y
            if y:
                offset ^= bp.fut[st.j]
        f0, f1 = map(complex, f[st.v])
        coeffs[i] = (f0.real, f0.imag, f1.real, f1.imag)
    return coeffs, ov, ys, offset


def bits_equal(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def _plans(ibm64):
    out = []
    for n, d, t, seed, k, sliced in ((16, 12, 20, 0, 8, False), (24, 24, 80, 3, 8, False), (28, 32, 64, 0, 8, True)):
        inst = synthetic_instance(n, d, t, seed)
        if sliced:
            _, _, width = best_rotation(inst['graph'], inst['unitaries'], n, inst['natural_order'], k, stride=2)
            fit = fit_width(inst['graph'], inst['unitaries'], n, inst['natural_order'], k, width - 1, max_slices=2)
            out.append(BlockPlan(inst['graph'], inst['unitaries'], n, fit['order'], fit['operations'],
                                 fit['open_vertices']))
        else:
            order, opn, _ = best_rotation(inst['graph'], inst['unitaries'], n, inst['natural_order'], k, stride=2)
            out.append(BlockPlan(inst['graph'], inst['unitaries'], n, order, [], opn))
    w = ibm64
    out.append(BlockPlan(w['graph'], w['unitaries'], 64, w['order'], w['plans']['inplace_k2_fp32']['operations'],
                         w['open_vertices']))
    return out


def test_tabulated_bookkeeping_is_bit_identical(ibm64):
    for bp in _plans(ibm64):
        assert bp.slices or bp.n < 64
        prot = make_protocol(7, 6, bp.n, bp.open_full, 1.0)
        xs = [np.array([int(c) for c in r['input_data_bits']], np.uint8) for r in prot['records']]
        for x in xs:
            ref_f = local_vectors(bp.U_full, bp.N_full, bp.n, x)
            f_full = bp.local_vectors(x)
            assert sorted(f_full) == sorted(ref_f)
            assert all(bits_equal(f_full[q], ref_f[q]) for q in ref_f)
            for parts, part in ((1, 0), (2, 1)):
                got = list(bp.slice_assignments(f_full, part, parts))
                want = list(direct_slices(bp, ref_f, part, parts))
                assert len(got) == len(want)
                for (c1, k1, f1, y1, ff1), (c2, k2, f2, y2, ff2) in zip(got, want):
                    assert (c1, ff1) == (c2, ff2) and bits_equal(np.complex128(k1), np.complex128(k2))
                    if f2 is None:
                        assert f1 is None
                        continue
                    assert y1 == y2
                    assert all(bits_equal(f1[v], f2[v]) for v in f2)
                    a, b = bp.step_parameters(f1, y1), direct_steps(bp, f2, y2)
                    assert all(bits_equal(p, q) for p, q in zip(a[:3], b[:3])) and a[3] == b[3]
                    # a dict argument (callers that build f themselves) gives the same parameters
                    c = bp.step_parameters(f2, y2)
                    assert all(bits_equal(p, q) for p, q in zip(c[:3], b[:3])) and c[3] == b[3]


def test_local_vectors_with_a_functional_and_read_only_tables():
    inst = synthetic_instance(16, 12, 20, 0)
    bp = BlockPlan(inst['graph'], inst['unitaries'], 16, inst['natural_order'], [], list(range(8, 16)))
    x = np.zeros(16, np.uint8)
    w = np.array([0.6, 0.8j])
    got, want = bp.local_vectors(x, w=w), local_vectors(bp.U_full, bp.N_full, bp.n, x, w=w)
    assert all(bits_equal(got[q], want[q]) for q in want)
    f = bp.local_vectors(x)
    with pytest.raises(ValueError):
        f[0][0] = 0
