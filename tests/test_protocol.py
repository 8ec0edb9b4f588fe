"""Host protocol functions against the frozen protocol records and the reference self-tests."""
import math
from fractions import Fraction

import numpy as np

from dcs_amd.sampling_protocol import actual_cdf_law_vector, assemble_output, draw_checked, make_protocol


def test_make_protocol_replays_frozen_generation(records):
    g = records['generation']
    p = make_protocol(g['seed'], 16, g['n'], g['free_vertices'], g['mixture_weight'])
    for got, want in zip(p['records'], g['first_records']):
        assert got['input_data_bits'] == want['input_data_bits']
        assert got['draw_uniform'] == want['draw_uniform']
        assert got['branch_uniform'] == want['branch_uniform']


def test_make_protocol_uniforms_match_run_records(records):
    g = records['generation']
    p = make_protocol(g['seed'], g['count'], g['n'], g['free_vertices'], g['mixture_weight'])
    for rec in p['records']:
        s = records['samples'][str(rec['sample_id'])]
        assert s['bits'] == rec['input_data_bits'] and s['u'] == rec['draw_uniform']


def test_protocol_selftest():
    m = make_protocol(733245, 113, 6, [5, 1, 3], .45)
    assert m == make_protocol(733245, 113, 6, [5, 1, 3], .45)
    assert any(r['conditional_branch'] for r in m['records'])
    assert any(not r['conditional_branch'] for r in m['records'])
    for i in range(8):
        bits, qiskit = assemble_output('000000', [5, 1, 3], i, [5, 0, 4, 1, 3, 2])
        assert sum(int(bits[q]) << j for j, q in enumerate([5, 1, 3])) == i
        assert len(qiskit) == 6


def test_draw_checked_errors():
    import pytest
    with pytest.raises(ValueError):
        draw_checked(np.ones(3), .5)
    with pytest.raises(ValueError):
        draw_checked(np.array([1, np.nan, 0, 0]), .5)
    with pytest.raises(ValueError):
        draw_checked(np.ones(4), 1.0)
    with pytest.raises(ArithmeticError):
        draw_checked(np.zeros(4), .5)


def test_actual_cdf_law_exact_grid_counts():
    """53-bit grid counts equal exact rational ceilings (production/test_protocol_cpu.py:22-42)."""
    rng = np.random.default_rng(11)
    a = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    a[::7] = 0
    law, meta = actual_cdf_law_vector(a)
    w = a.real * a.real + a.imag * a.imag
    cum = np.cumsum(w)
    cum /= cum[-1]
    cum[-1] = 1.
    ends = [math.ceil(Fraction(float(c)) * 2 ** 53) for c in cum]
    counts = np.diff([0] + ends)
    assert np.allclose(np.abs(law) ** 2, np.array(counts, float) / 2 ** 53, rtol=0, atol=4e-16)
    assert meta['grid_total'] == 1 << 53 and abs(meta['law_norm_squared'] - 1) < 1e-12


def test_frozen_draw_intervals_are_consistent(records):
    """Every recorded c64 draw satisfies lo <= u < hi (the interval the verifier replays)."""
    for s in records['samples'].values():
        assert s['lo'] <= s['u'] < s['hi']
