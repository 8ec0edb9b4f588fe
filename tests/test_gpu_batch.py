"""Circuits of different frontier widths contracted in shared launches (dcs_contract_batch through
block_batch): every block equals the same engine's block() alone, bit for bit, and one of them equals the
oracle's block."""
import numpy as np
import pytest

from dcs_amd import _native
from dcs_amd.conditional_block import InplaceConditionalBlockGPU, block_batch
from dcs_amd.sampling_protocol import make_protocol
from dcs_amd.stabilizer import synthetic_instance
from dcs_amd.width_planner import best_rotation, fit_width

pytestmark = pytest.mark.gpu

# (qubits, depth, T count, seed, open qubits) of seeded ring circuits; frontier widths 7, 10, 16, 18, 19 with
# the best ring rotation (the narrow ones LDS-resident throughout), and 17 for the last with a parity slice
SPECS = [(16, 12, 20, 0, 8), (20, 16, 40, 1, 10), (24, 24, 80, 3, 8), (28, 32, 96, 1, 10), (32, 30, 100, 0, 12),
         (28, 32, 64, 0, 8)]


def _engine_args(n, depth, t, seed, k, sliced=False):
    inst = synthetic_instance(n, depth, t, seed)
    if sliced:
        order, opn, width = best_rotation(inst['graph'], inst['unitaries'], n, inst['natural_order'], k, stride=2)
        fit = fit_width(inst['graph'], inst['unitaries'], n, inst['natural_order'], k, width - 1, max_slices=2)
        return (inst['graph'], inst['unitaries'], n, fit['order'], fit['operations'], fit['open_vertices'])
    order, opn, _ = best_rotation(inst['graph'], inst['unitaries'], n, inst['natural_order'], k, stride=2)
    return (inst['graph'], inst['unitaries'], n, order, [], opn)


def _prefixes(args, count, seed):
    prot = make_protocol(seed, count, args[2], args[5], 1.0)
    return [np.array([int(c) for c in r['input_data_bits']], np.uint8) for r in prot['records']]


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
def test_batch_equals_alone(dtype):
    print(f'case block_batch {np.dtype(dtype).name}', flush=True)
    ctx = _native.Context(0)
    engines, requests, segments = [], [], 0
    for i, spec in enumerate(SPECS):
        args = _engine_args(*spec, sliced=(i == len(SPECS) - 1))
        e = InplaceConditionalBlockGPU(*args)
        # uneven lane counts: some engines need more rounds than others
        e.alloc(dtype=dtype, schedule='fused', lanes=1 + i % 2, blocks=3, context=ctx)
        engines.append(e)
        segments += len(e.schedule())
        for b, x in enumerate(_prefixes(args, 2, seed=i)):
            requests.append((e, x, b))
    blocks, meta = block_batch(requests)
    got = [np.asarray(v) for v in blocks]
    print('launches', meta['launches'], 'rounds', meta['rounds'], 'segments alone per slice', segments, flush=True)
    assert meta['launches'] < segments
    for (e, x, b), g in zip(requests, got):
        v, _ = e.block(x, block_index=2)
        alone = np.asarray(v)
        assert np.array_equal(alone.view(np.float64), g.view(np.float64)), (e.n, b)
        assert np.isfinite(g).all() and np.abs(g).max() > 0
    for e in engines:
        e.close()
    ctx.close()


def test_batch_matches_oracle():
    """One batched block against the oracle's contraction of the same prefix (complex128 storage)."""
    from oracle import oracle
    print('case block_batch oracle', flush=True)
    ctx = _native.Context(0)
    args_a = _engine_args(*SPECS[2])
    args_b = _engine_args(*SPECS[4])
    ea, eb = InplaceConditionalBlockGPU(*args_a), InplaceConditionalBlockGPU(*args_b)
    for e in (ea, eb):
        e.alloc(dtype=np.complex128, schedule='fused', lanes=1, context=ctx)
    xa, xb = _prefixes(args_a, 1, 5)[0], _prefixes(args_b, 1, 6)[0]
    blocks, _ = block_batch([(ea, xa, 0), (eb, xb, 0)])
    for e, x, v in ((ea, xa, blocks[0]), (eb, xb, blocks[1])):
        ref = oracle.block(e.bp, x, storage=1)
        got = np.asarray(v)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    for e in (ea, eb):
        e.close()
    ctx.close()


def test_batch_rejects_mixed_contexts():
    ctx = _native.Context(0)
    args = _engine_args(*SPECS[0])
    ea, eb = InplaceConditionalBlockGPU(*args), InplaceConditionalBlockGPU(*args)
    ea.alloc(dtype=np.complex64, schedule='fused', context=ctx)
    eb.alloc(dtype=np.complex64, schedule='fused')
    x = _prefixes(args, 1, 0)[0]
    with pytest.raises(ValueError, match='share one context'):
        block_batch([(ea, x, 0), (eb, x, 0)])
    ec = InplaceConditionalBlockGPU(*args)
    ec.alloc(dtype=np.complex64, schedule='compiled', context=ctx)
    with pytest.raises(ValueError, match='fused'):
        block_batch([(ea, x, 0), (ec, x, 0)])
    for e in (ea, eb, ec):
        e.close()
    ctx.close()


@pytest.mark.parametrize('concurrent', [False, True])
def test_serving_narrow_circuits_in_shared_launches(concurrent):
    """Server(narrow_width=20): four narrow circuits served through shared launches; every request's block
    equals its circuit's fused engine alone, bit for bit, and its draw equals draw_checked on that block."""
    from dcs_amd.serving import Request, Server, latency_stats
    from test_gpu_serving import assert_draw_matches_host
    print(f'case narrow serving concurrent={concurrent}', flush=True)
    srv = Server(hbm_budget=16 << 30, narrow_width=20)
    specs = {f'q{s[0]}w{i}': _engine_args(*s) for i, s in enumerate(SPECS[:5]) if i != 1}
    for name, args in specs.items():
        srv.add(name, *args, lanes=2, max_batch=2)
    rng = np.random.default_rng(2)
    names = list(specs)
    reqs = []
    for i in range(24):
        name = names[int(rng.integers(len(names)))]
        x = _prefixes(specs[name], 1, 100 + i)[0]
        reqs.append(Request(i, name, x, float(rng.random()), 0.002 * (i // 3)))
    srv.serve(reqs, realtime=concurrent, keep_blocks=True, concurrent=concurrent)
    assert srv.narrow_batches >= 1 and all(c.narrow for c in srv.circuits.values())
    print('narrow batches', srv.narrow_batches, latency_stats(reqs)['all'], flush=True)
    alone = {}
    for name, args in specs.items():
        e = InplaceConditionalBlockGPU(*args)
        e.alloc(dtype=np.complex64, schedule='fused')
        alone[name] = e
    for r in reqs:
        v, _ = alone[r.circuit].block(r.x_bits)
        ref = np.asarray(v)
        assert np.array_equal(ref.view(np.float64), r.block.view(np.float64)), r.rid
        assert_draw_matches_host(r.draw, r.block, r.uniform)
    for e in alone.values():
        e.close()
    srv.close()
