"""Serving loop on the device: a Poisson mix of three resident circuits; every request's block and draw
equal serving it alone (bit for bit), and the draws agree with the host draw_checked."""
import math

import numpy as np
import pytest

from dcs_amd.sampling_protocol import draw_checked
from dcs_amd.serving import Server, latency_stats, poisson_workload
from dcs_amd.stabilizer import synthetic_instance

pytestmark = pytest.mark.gpu


def assert_draw_matches_host(d, block, u):
    """A served draw against draw_checked on the host copy of its block: the same index and primary
    (numpy) interval bit for bit, and the device's compensated interval within one unit in the last place
    of the math.fsum interval."""
    h = draw_checked(block, u)
    assert d['A_index'] == h['A_index']
    for k in ('cdf_lower_numpy', 'cdf_upper_numpy', 'scaled_mass_cdf', 'selected_scaled_weight'):
        assert d[k] == h[k], k
    for dk, hk in (('scaled_mass_compensated', 'scaled_mass_fsum'), ('cdf_lower_compensated', 'cdf_lower_fsum'),
                   ('cdf_upper_compensated', 'cdf_upper_fsum')):
        assert abs(d[dk] - h[hk]) <= math.ulp(h[hk]), (dk, d[dk], h[hk])
    assert d['independent_cdf_check_passed'] and h['independent_cdf_check_passed']


@pytest.mark.parametrize('concurrent', [False, True])
def test_mixed_serving_equals_alone(concurrent):
    """Sequential loop, and one worker thread per circuit (the circuits' batches in flight together on
    their own streams)."""
    print(f'case serving mix concurrent={concurrent}', flush=True)
    a = synthetic_instance(12, 10, 14, seed=1)
    b = synthetic_instance(16, 12, 20, seed=0)
    c = synthetic_instance(28, 32, 64, seed=0)
    circuits = {
        'q12': ((a['graph'], a['unitaries'], 12, a['natural_order'], [], [9, 2, 7, 11, 4, 0]), 4),
        'q16': ((b['graph'], b['unitaries'], 16, b['natural_order'], [], list(range(10, 16))), 2),
        'q28': ((c['graph'], c['unitaries'], 28, c['natural_order'], [], list(range(20, 28))), 3),
    }
    srv = Server(hbm_budget=32 << 30)
    mix = {}
    for name, (args, lanes) in circuits.items():
        srv.register(name, *args, lanes=lanes, max_batch=3, schedule='compiled', lds_storage=True)
        mix[name] = (1, args[2], args[5])
    reqs = poisson_workload(mix, rate=200.0, count=24, seed=5)
    srv.serve(reqs, realtime=False, keep_blocks=True, concurrent=concurrent)
    st = latency_stats(reqs)
    print({k: (v['count'], round(v['p50_ms'], 2), round(v['p99_ms'], 2)) for k, v in st.items()}, flush=True)
    assert st['all']['count'] == 24
    assert max(r.batch for r in reqs) < 24  # some requests were batched together
    for r in reqs:
        eng = srv.circuits[r.circuit].engine
        vec, _ = eng.block(r.x_bits)
        alone = np.asarray(vec)
        assert np.array_equal(alone, r.block), r.rid
        d = vec.draw(r.uniform)
        assert d == r.draw
        assert_draw_matches_host(r.draw, alone, r.uniform)
    srv.close()


def test_autotuned_schedule_is_correct():
    """The tuner times three tilings of a width-20 plan on the device and the engine built with the
    winner stays within 1e-5 of the complex128 oracle."""
    import numpy as np
    from dcs_amd.autotune import autotune
    from dcs_amd.block_plan import BlockPlan
    from dcs_amd.conditional_block import InplaceConditionalBlockGPU
    from dcs_amd.stabilizer import synthetic_instance
    from oracle import oracle
    print('case autotune w20', flush=True)
    inst = synthetic_instance(28, 32, 64, seed=0)
    args = (inst['graph'], inst['unitaries'], 28, inst['natural_order'], [], list(range(20, 28)))
    x = np.random.default_rng(11).integers(0, 2, 28, dtype=np.uint8)
    res = autotune(InplaceConditionalBlockGPU, args, x=x, candidates=(
        {'fuse_max_dim': 8, 'tile_log2': 12, 'group_policy': 1, 'lds_buffers': 1},
        {'fuse_max_dim': 6, 'tile_log2': 11, 'group_policy': 0, 'lds_buffers': 2},
        {'fuse_max_dim': 9, 'tile_log2': 13, 'group_policy': 1, 'lds_buffers': 1}))
    print(res['best'], [round(r['seconds'] * 1e3, 3) for r in res['table']], flush=True)
    assert all(r['seconds'] and r['seconds'] > 0 for r in res['table'])
    e = InplaceConditionalBlockGPU(*args)
    e.alloc(dtype=np.complex64, **res['best'])
    v = np.asarray(e.block(x)[0])
    e.close()
    ref = oracle.block(BlockPlan(*args), x)
    assert np.max(np.abs(v - ref)) <= 1e-5 * np.max(np.abs(ref))
