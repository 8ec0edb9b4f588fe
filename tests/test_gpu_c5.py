"""The serving mix on one device: 40q/d48/180T ring circuits with 2^16 blocks, 52q/d60/250T with 2^18
blocks, and the 64-qubit IBM QAT #245 graph with 2^20 and 2^22 blocks (20 or 22 open data qubits, plan
inplace_k2_fp32, complex64), added to a server through dcs_amd.serving.c5_mix.

The 52q member (circuit seed 1) is the best ring rotation with three greedy parity slices of T
references (width 36 -> 33: a 64 GiB complex64 frontier and 8 contractions per block); unsliced it
is refused as larger than the device.  The two 64-qubit circuits (128 GiB each) do not fit together
with the rest, so requests for a circuit that is not resident wait until an idle one is evicted.
Requests arrive as a seeded Poisson stream and are served by one worker thread per resident circuit;
every block equals serving its request alone, bit for bit, and every draw equals draw_checked on the
host copy of its block.

The 2^22 block is tied to the 2^20 block: with the two extra open qubits fixed to the bits a 2^20
request closes, it equals that request's block scaled by 1/2 (D halves twice), within the complex64
contraction error (the two plans keep their open qubits last, so they differ)."""
import math

import numpy as np
import pytest

from dcs_amd.sampling_protocol import draw_checked
from dcs_amd.serving import Request, Server, c5_mix, latency_stats, poisson_workload
from dcs_amd.stabilizer import synthetic_instance
from dcs_amd.width_planner import best_rotation

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
    print('case c5 serving mix: 40q 2^16, 52q 2^18 (3 parity slices), 64q 2^20 and 2^22', flush=True)
    srv = Server()
    mix, widths = c5_mix(srv)
    assert widths['40q/d48/180T'] == 31 and widths['52q/d60/250T'] == 33 and widths['52q_slices'] == 3
    assert srv.catalog['64q/2^22'].engine.bp.k == 22 and srv.catalog['64q/2^22'].engine.bp.inplace_maxslots == 34
    assert srv.catalog['52q/d60/250T'].engine.bp.k == 18
    # the unsliced 52q circuit (best rotation, width 36 for this seed) would need 512 GiB
    with pytest.raises(MemoryError):
        inst = synthetic_instance(52, 60, 250, 1)
        order, opn, w = best_rotation(inst['graph'], inst['unitaries'], 52, inst['natural_order'], 18, stride=2)
        srv.add('52q unsliced', inst['graph'], inst['unitaries'], 52, order, [], opn, lanes=1, max_batch=2)
    reqs = poisson_workload(mix, rate=4.0, count=16, seed=11)
    counts = {k: sum(r.circuit == k for r in reqs) for k in mix}
    print('requests', counts, flush=True)
    srv.serve(reqs, concurrent=True, keep_blocks=True)
    st = latency_stats(reqs)
    print({k: (v['count'], round(v['p50_ms'], 1), round(v['p99_ms'], 1)) for k, v in st.items()},
          'evictions', srv.evictions, flush=True)
    assert srv.used <= srv.budget
    if counts['64q/2^20'] and counts['64q/2^22']:
        assert srv.evictions >= 1
    for name in mix:                  # alone, one circuit at a time (admissions re-allocate as needed)
        for r in (r for r in reqs if r.circuit == name):
            one = served_alone(srv, r)
            assert np.array_equal(one.block.view(np.float64), r.block.view(np.float64)), r.rid
            assert one.draw == r.draw, r.rid
            check_draw(r.draw, r.block, r.uniform)
    # the 2^22 block of a 2^20 request's prefix, restricted to the two extra qubits' values
    opn20, opn22 = mix['64q/2^20'][2], mix['64q/2^22'][2]
    extra = opn22[len(opn20):]
    r20 = next((r for r in reqs if r.circuit == '64q/2^20'), None)
    if r20 is None:
        anchor = np.array([int(c) for c in ibm64['protocol']['anchor_input_data_bits']], np.uint8)
        r20 = served_alone(srv, Request(-2, '64q/2^20', anchor, 0.5, 0.0))
    x22 = r20.x_bits.copy()
    bits = [int(r20.x_bits[q]) for q in extra]
    x22[extra] = 0
    one = served_alone(srv, Request(-1, '64q/2^22', x22, 0.5, 0.0))
    # bit j of a block index is open qubit opn22[j]: the extra qubits are bits 20 and 21
    sub = one.block[np.arange(1 << 20) | (bits[0] << 20) | (bits[1] << 21)]
    err = float(np.max(np.abs(2 * sub - r20.block)) / np.max(np.abs(r20.block)))
    print(f'2^22 block restricted to the 2^20 prefix: max rel err {err:.3e}', flush=True)
    assert err <= 1e-5
    srv.close()


def test_c5_52q_sliced_plan_complex64_against_complex128():
    """The 52q member's sliced plan (8 contractions of a width-33 frontier per block): the complex64
    block of one prefix within 1e-5 of the block maximum of the complex128 block, the same draw unless
    the uniform is within 1e-7 of a CDF edge."""
    from dcs_amd.conditional_block import InplaceConditionalBlockGPU
    from dcs_amd.sampling_protocol import make_protocol
    from dcs_amd.serving import RING_PLANS
    from dcs_amd.width_planner import fit_width
    inst = synthetic_instance(52, 60, 250, 1)
    fit = fit_width(inst['graph'], inst['unitaries'], 52, inst['natural_order'], 18, 33, max_slices=4,
                    cache_path=RING_PLANS)
    args = (inst['graph'], inst['unitaries'], 52, fit['order'], fit['operations'], fit['open_vertices'])
    rec = make_protocol(5, 1, 52, fit['open_vertices'], 1.0)['records'][0]
    x = np.array([int(c) for c in rec['input_data_bits']], np.uint8)
    out = {}
    for dt in (np.complex64, np.complex128):
        print(f'case 52q sliced {np.dtype(dt).name}', flush=True)
        e = InplaceConditionalBlockGPU(*args)
        e.alloc(dtype=dt)
        vec, meta = e.block(x)
        assert meta['slices_evaluated'] == 8 and e.inplace_rmax == 33
        out[dt] = (np.asarray(vec), vec.draw(rec['draw_uniform']), meta['seconds'])
        e.close()
    a, r = out[np.complex64][0], out[np.complex128][0]
    err = float(np.max(np.abs(a - r)) / np.max(np.abs(r)))
    d64, d128 = out[np.complex64][1], out[np.complex128][1]
    print(f'52q: complex64 {out[np.complex64][2]:.2f} s, complex128 {out[np.complex128][2]:.2f} s, err {err:.3e}, '
          f'index {d64["A_index"]} vs {d128["A_index"]}', flush=True)
    assert err <= 1e-5
    u = rec['draw_uniform']
    if min(u - d128['cdf_lower_numpy'], d128['cdf_upper_numpy'] - u) >= 1e-7:
        assert d64['A_index'] == d128['A_index']
