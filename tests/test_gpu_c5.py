"""The serving mix on one device: 40q/d48/180T ring circuits with 2^16 blocks, and the 64-qubit IBM QAT
#245 graph with 2^20 and 2^22 blocks (20 or 22 open data qubits, plan inplace_k2_fp32, complex64).

The 52q/d60/250T member is refused at admission: its frontier (width 39 for the best ring rotation with
18 open qubits, 2^39 complex64 entries) exceeds the whole device.  The two 64-qubit circuits (128 GiB
each) and the 40q circuit (16 GiB per lane, four lanes) do not fit together, so requests for a circuit
that is not resident wait until an idle one is evicted.  Requests arrive as a seeded Poisson stream and
are served by one worker thread per resident circuit; every block equals serving its request alone, bit
for bit, and every draw equals draw_checked on the host copy of its block.

The 2^22 block is tied to the 2^20 block: with the two extra open qubits fixed to the bits a 2^20
request closes, it equals that request's block scaled by 1/2 (D halves twice), within the complex64
contraction error (the two plans keep their open qubits last, so they differ)."""
import math

import numpy as np
import pytest

from dcs_amd.sampling_protocol import draw_checked
from dcs_amd.serving import Request, Server, _synthetic, latency_stats, poisson_workload

pytestmark = pytest.mark.gpu


def served_alone(srv, r):
    one = Request(r.rid, r.circuit, r.x_bits, r.uniform, 0.0)
    srv.serve([one], realtime=False, keep_blocks=True)
    return one


def check_draw(d, block, u):
    h = draw_checked(block, u)
    assert d['A_index'] == h['A_index']
    for k in ('cdf_lower_numpy', 'cdf_upper_numpy', 'scaled_mass_cdf'):
        assert d[k] == h[k], k
    for dk, hk in (('cdf_lower_compensated', 'cdf_lower_fsum'), ('cdf_upper_compensated', 'cdf_upper_fsum')):
        assert abs(d[dk] - h[hk]) <= math.ulp(h[hk]), (dk, d[dk], h[hk])


def test_c5_mix_admits_as_memory_frees_and_equals_alone(ibm64):
    print('case c5 serving mix: 40q 2^16, 64q 2^20, 64q 2^22 (52q refused)', flush=True)
    srv = Server()
    args40, w40 = _synthetic(40, 48, 180, 0, 16)
    assert w40 == 31
    srv.add('40q', *args40, lanes=4, max_batch=4, schedule='compiled')
    args52, w52 = _synthetic(52, 60, 250, 0, 18)
    with pytest.raises(MemoryError):
        srv.add('52q', *args52, lanes=1, max_batch=2, schedule='compiled')
    order = list(ibm64['order'])
    opn20 = list(ibm64['open_vertices'])
    extra = [v for v in order if v < 64 and v not in opn20][::-1][:2]
    opn22 = opn20 + extra
    ops = ibm64['plans']['inplace_k2_fp32']['operations']
    srv.add('64q-2^20', ibm64['graph'], ibm64['unitaries'], 64, order, ops, opn20, lanes=1, max_batch=2,
            schedule='compiled')
    srv.add('64q-2^22', ibm64['graph'], ibm64['unitaries'], 64, order, ops, opn22, lanes=1, max_batch=2,
            schedule='compiled')
    assert srv.catalog['64q-2^22'].engine.bp.k == 22 and srv.catalog['64q-2^22'].engine.bp.inplace_maxslots == 34
    mix = {'40q': (3, 40, args40[5]), '64q-2^20': (0.5, 64, opn20), '64q-2^22': (0.5, 64, opn22)}
    reqs = poisson_workload(mix, rate=6.0, count=16, seed=11)
    counts = {k: sum(r.circuit == k for r in reqs) for k in mix}
    print('requests', counts, flush=True)
    srv.serve(reqs, concurrent=True, keep_blocks=True)
    st = latency_stats(reqs)
    print({k: (v['count'], round(v['p50_ms'], 1), round(v['p99_ms'], 1)) for k, v in st.items()},
          'evictions', srv.evictions, flush=True)
    assert srv.used <= srv.budget
    if counts['64q-2^20'] and counts['64q-2^22']:
        assert srv.evictions >= 1
    for name in mix:                  # alone, one circuit at a time (admissions re-allocate as needed)
        for r in (r for r in reqs if r.circuit == name):
            one = served_alone(srv, r)
            assert np.array_equal(one.block.view(np.float64), r.block.view(np.float64)), r.rid
            assert one.draw == r.draw, r.rid
            check_draw(r.draw, r.block, r.uniform)
    # the 2^22 block of a 2^20 request's prefix, restricted to the two extra qubits' values
    r20 = next((r for r in reqs if r.circuit == '64q-2^20'), None)
    if r20 is not None:
        x22 = r20.x_bits.copy()
        bits = [int(r20.x_bits[q]) for q in extra]
        x22[extra] = 0
        one = served_alone(srv, Request(-1, '64q-2^22', x22, 0.5, 0.0))
        # bit j of a block index is open qubit opn22[j]: the extra qubits are bits 20 and 21
        sub = one.block[np.arange(1 << 20) | (bits[0] << 20) | (bits[1] << 21)]
        err = float(np.max(np.abs(2 * sub - r20.block)) / np.max(np.abs(r20.block)))
        print(f'2^22 block restricted to the 2^20 prefix: max rel err {err:.3e}', flush=True)
        assert err <= 1e-5
    srv.close()
