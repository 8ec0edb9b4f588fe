"""Serving loop host logic (admission, continuous batching, workload generator) with a stand-in engine,
and the width planner (cut ranks, rotations, greedy parity slicing) against brute force and the oracle."""
import math

import numpy as np
import pytest

from dcs_amd.block_plan import BlockPlan
from dcs_amd.serving import Request, Server, footprint_bytes, latency_stats, poisson_workload
from dcs_amd.stabilizer import synthetic_instance
from dcs_amd.width_planner import best_rotation, choose_parity_slices, order_cost, prefix_cut_ranks
from oracle import oracle


class _BP:
    def __init__(self, k, width):
        self.k, self.inplace_maxslots = k, width


class _Vec:
    def __init__(self, x):
        self.x = x

    def draw(self, u):
        return {'A_index': int(self.x.sum()) * 1000 + int(u * 1000)}


class FakeEngine:
    """Stand-in with the engine surface the server uses; records every batch."""
    log = []

    def __init__(self, graph, unitaries, n, order, ops, opn):
        self.bp = _BP(len(opn), int(graph))   # `graph` carries the width in these tests
        self.inplace_rmax = int(graph)
        self.lanes = self.blocks = None

    def alloc(self, dtype, device, lanes, blocks, **kw):
        self.lanes, self.blocks, self.dtype = lanes, blocks, dtype

    def device_bytes(self):
        return footprint_bytes(self.bp, self.dtype, self.lanes, self.blocks)

    def block_many(self, xs):
        FakeEngine.log.append((self.inplace_rmax, len(xs)))
        return [_Vec(x) for x in xs], {}

    def close(self):
        pass


def _server(budget):
    FakeEngine.log = []
    return Server(hbm_budget=budget, engine_factory=FakeEngine)


def test_admission_reduces_lanes_and_refuses():
    srv = _server(budget=3 * (1 << 30) * 8 + (512 << 20))
    c = srv.register('a', 30, None, 8, None, [], list(range(4)), lanes=8, max_batch=2)   # 8 GiB per lane
    assert c.lanes == 2
    assert srv.used == footprint_bytes(_BP(4, 30), np.complex64, 2, 2)
    c2 = srv.register('b', 29, None, 8, None, [], list(range(4)), lanes=4, max_batch=2)   # 4 GiB per lane
    assert c2.lanes == 2 and srv.used <= srv.budget
    with pytest.raises(MemoryError):
        srv.register('c', 31, None, 8, None, [], list(range(4)), lanes=1)
    srv.unregister('a')
    assert srv.register('c', 31, None, 8, None, [], list(range(4)), lanes=1).lanes == 1


def test_continuous_batching_order_and_bounds():
    srv = _server(budget=1 << 40)
    srv.register('small', 10, None, 8, None, [], [0, 1], lanes=4, max_batch=3)
    srv.register('big', 12, None, 8, None, [], [0, 1], lanes=1, max_batch=2)
    x = np.zeros(8, np.uint8)
    reqs = [Request(0, 'big', x, .1, 0.0), Request(1, 'small', x, .2, 0.001), Request(2, 'small', x, .3, 0.002),
            Request(3, 'small', x, .4, 0.003), Request(4, 'small', x, .5, 0.004), Request(5, 'big', x, .6, 0.005)]
    srv.serve(reqs, realtime=False)
    assert all(r.index >= 0 and r.done >= r.arrival for r in reqs)
    assert sorted({r.batch for r in reqs}) == list(range(len(FakeEngine.log)))
    for b in range(len(FakeEngine.log)):
        members = [r for r in reqs if r.batch == b]
        assert len({r.circuit for r in members}) == 1
        limit = 3 if members[0].circuit == 'small' else 2
        assert len(members) <= limit
    # FIFO within a circuit: batch numbers never decrease with arrival
    for name in ('small', 'big'):
        seq = [r.batch for r in sorted(reqs, key=lambda r: r.arrival) if r.circuit == name]
        assert seq == sorted(seq)
    st = latency_stats(reqs)
    assert st['all']['count'] == 6 and st['small']['count'] == 4 and st['all']['p99_ms'] >= st['all']['p50_ms']


class SlowEngine(FakeEngine):
    """Stand-in whose batches take time proportional to the width (released interpreter lock)."""
    active, peak = 0, 0

    def block_many(self, xs):
        import time
        SlowEngine.active += 1
        SlowEngine.peak = max(SlowEngine.peak, SlowEngine.active)
        time.sleep(0.002 * (self.inplace_rmax - 9))
        SlowEngine.active -= 1
        return super().block_many(xs)


def test_concurrent_serving_overlaps_circuits_and_keeps_results():
    SlowEngine.log, SlowEngine.active, SlowEngine.peak = [], 0, 0
    srv = Server(hbm_budget=1 << 50, engine_factory=SlowEngine)
    srv.register('small', 10, None, 8, None, [], [0, 1], lanes=4, max_batch=2)
    srv.register('big', 40, None, 8, None, [], [0, 1], lanes=1, max_batch=2)
    rng = np.random.default_rng(3)
    reqs = [Request(i, 'big' if i % 4 == 0 else 'small', rng.integers(0, 2, 8, dtype=np.uint8), float(rng.random()),
                    0.005 * i) for i in range(24)]
    want = {r.rid: r.x_bits.sum() * 1000 + int(r.uniform * 1000) for r in reqs}
    srv.serve(reqs, concurrent=True)
    assert all(r.index == want[r.rid] and r.done >= r.arrival >= 0 for r in reqs)
    assert SlowEngine.peak == 2            # a small batch ran while a big one was in flight
    for name in ('small', 'big'):          # FIFO per circuit, batches within max_batch
        mine = sorted((r for r in reqs if r.circuit == name), key=lambda r: r.arrival)
        assert [r.batch for r in mine] == sorted(r.batch for r in mine)
        assert max(collections_count(r.batch for r in mine).values()) <= 2
    st = latency_stats(reqs)
    assert st['small']['p99_ms'] < st['big']['p50_ms'] + 100   # small requests do not wait behind big ones


def collections_count(it):
    import collections
    return collections.Counter(it)


def test_concurrent_serving_propagates_engine_errors():
    class Broken(FakeEngine):
        def block_many(self, xs):
            raise ArithmeticError('numerically zero block')
    srv = Server(hbm_budget=1 << 40, engine_factory=Broken)
    srv.register('a', 10, None, 8, None, [], [0, 1], lanes=1, max_batch=2)
    x = np.zeros(8, np.uint8)
    with pytest.raises(ArithmeticError):
        srv.serve([Request(i, 'a', x, .5, 0.0) for i in range(3)], concurrent=True)


class TrackedEngine(SlowEngine):
    """Stand-in that records the device bytes held at once (through the server's accounting)."""
    server, peak_used, live = None, 0, set()

    def alloc(self, dtype, device, lanes, blocks, **kw):
        super().alloc(dtype, device, lanes, blocks, **kw)
        TrackedEngine.live.add(id(self))

    def block_many(self, xs):
        TrackedEngine.peak_used = max(TrackedEngine.peak_used, TrackedEngine.server.used)
        assert id(self) in TrackedEngine.live        # never a batch on an evicted engine
        return super().block_many(xs)

    def close(self):
        TrackedEngine.live.discard(id(self))


@pytest.mark.parametrize('concurrent', [False, True])
def test_requests_admit_circuits_as_memory_frees(concurrent):
    """Three known circuits, of which only two fit the budget at once: requests for a circuit that is not
    resident wait, an idle circuit (nothing queued, no batch in flight) is evicted least recently used
    first, every request is served with the right result, and the budget is never exceeded."""
    SlowEngine.log, SlowEngine.active, SlowEngine.peak = [], 0, 0
    per_lane = footprint_bytes(_BP(2, 30), np.complex64, 1, 2)
    srv = Server(hbm_budget=2 * per_lane + (1 << 20), engine_factory=TrackedEngine)
    TrackedEngine.server, TrackedEngine.peak_used, TrackedEngine.live = srv, 0, set()
    for name in ('a', 'b', 'c'):
        srv.add(name, 30, None, 8, None, [], [0, 1], lanes=4, max_batch=2)
    assert srv.used == 0 and not srv.circuits
    with pytest.raises(MemoryError, match='whole'):
        srv.add('huge', 32, None, 8, None, [], [0, 1], lanes=1, max_batch=2)
    rng = np.random.default_rng(5)
    names = ['a', 'b', 'c', 'a', 'c', 'b', 'b', 'a', 'c', 'c', 'a', 'b']
    reqs = [Request(i, nm, rng.integers(0, 2, 8, dtype=np.uint8), float(rng.random()), 0.004 * i)
            for i, nm in enumerate(names)]
    want = {r.rid: r.x_bits.sum() * 1000 + int(r.uniform * 1000) for r in reqs}
    srv.serve(reqs, realtime=concurrent, concurrent=concurrent)
    assert all(r.index == want[r.rid] and r.done >= r.arrival for r in reqs)
    assert srv.evictions >= 1 and len(srv.circuits) <= 2
    assert TrackedEngine.peak_used <= srv.budget and srv.used <= srv.budget
    assert all(srv.catalog[nm].admissions >= 1 for nm in 'abc')
    srv.close()
    assert srv.used == 0


class _FakeCtx:
    made = 0

    def __init__(self, device):
        _FakeCtx.made += 1
        self.closed = False

    def close(self):
        self.closed = True


class NarrowEngine(TrackedEngine):
    """Stand-in recording how it was allocated (schedule, shared context)."""

    def alloc(self, dtype, device, lanes, blocks, **kw):
        super().alloc(dtype, device, lanes, blocks, **kw)
        self.kw = kw


def _fake_batch(log):
    import time

    def batch(work):
        engines = {id(e) for e, _, _ in work}
        for e, _, b in work:
            assert id(e) in TrackedEngine.live and 0 <= b < e.blocks
        log.append(sorted(e.inplace_rmax for e in {id(e): e for e, _, _ in work}.values()))
        time.sleep(0.002 * len(engines))
        return [_Vec(x) for _, x, _ in work], {'launches': 1}
    return batch


@pytest.mark.parametrize('concurrent', [False, True])
def test_narrow_circuits_share_batches(concurrent):
    """Server(narrow_width=w): circuits of width <= w are allocated in the fused schedule on one shared
    context and served together (one batch takes the waiting requests of all of them); wider circuits
    keep their own batches; results, FIFO order and the budget hold, and a narrow circuit is evicted
    and re-admitted like any other."""
    log = []
    _FakeCtx.made = 0
    per_lane = footprint_bytes(_BP(2, 30), np.complex64, 1, 2)
    small = footprint_bytes(_BP(2, 11), np.complex64, 4, 2)
    srv = Server(hbm_budget=per_lane + 2 * small + (1 << 20), engine_factory=NarrowEngine, narrow_width=12,
                 batch_fn=_fake_batch(log), context_factory=_FakeCtx)
    TrackedEngine.server, TrackedEngine.peak_used, TrackedEngine.live = srv, 0, set()
    SlowEngine.log, SlowEngine.active, SlowEngine.peak = [], 0, 0
    for name, w in (('n10', 10), ('n11', 11), ('n12', 12)):
        srv.add(name, w, None, 8, None, [], [0, 1], lanes=4, max_batch=2, schedule='compiled', lds_buffers=3)
    srv.add('wide', 30, None, 8, None, [], [0, 1], lanes=1, max_batch=2)
    rng = np.random.default_rng(11)
    names = ['n10', 'n11', 'wide', 'n10', 'n11', 'n10', 'n12', 'wide', 'n11', 'n12', 'n10', 'n11', 'n12', 'wide']
    # requests arrive in pairs, so that two narrow circuits often wait at the same time
    reqs = [Request(i, nm, rng.integers(0, 2, 8, dtype=np.uint8), float(rng.random()), 0.004 * (i // 2))
            for i, nm in enumerate(names)]
    want = {r.rid: r.x_bits.sum() * 1000 + int(r.uniform * 1000) for r in reqs}
    srv.serve(reqs, realtime=concurrent, concurrent=concurrent)
    assert all(r.index == want[r.rid] and r.done >= r.arrival for r in reqs)
    assert srv.narrow_batches == len(log) >= 1
    assert any(len(g) >= 2 for g in log)                 # at least one batch shared by two narrow circuits
    assert all(w <= 12 for g in log for w in g)          # the wide circuit never joins one
    assert all(e == 30 for e, _ in SlowEngine.log)      # block_many served only the wide circuit
    for nm in ('n10', 'n11', 'n12'):
        kw = srv.catalog[nm].engine.kw
        assert kw['schedule'] == 'fused' and kw['context'] is srv._shared_ctx and 'lds_buffers' not in kw
        seq = [r.batch for r in sorted(reqs, key=lambda r: r.arrival) if r.circuit == nm]
        assert seq == sorted(seq)
    assert srv.catalog['wide'].engine.kw.get('context') is None
    assert _FakeCtx.made == 1
    assert TrackedEngine.peak_used <= srv.budget and srv.evictions >= 1
    ctx = srv._shared_ctx
    srv.close()
    assert ctx.closed and srv.used == 0


def test_poisson_workload_is_seeded_and_weighted():
    mix = {'a': (3, 10, [8, 9]), 'b': (1, 12, [0])}
    r1 = poisson_workload(mix, rate=50.0, count=400, seed=7)
    r2 = poisson_workload(mix, rate=50.0, count=400, seed=7)
    assert [(r.circuit, r.uniform, r.arrival, r.x_bits.tobytes()) for r in r1] == \
           [(r.circuit, r.uniform, r.arrival, r.x_bits.tobytes()) for r in r2]
    frac = sum(r.circuit == 'a' for r in r1) / len(r1)
    assert 0.68 < frac < 0.82
    gaps = np.diff([0.0] + [r.arrival for r in r1])
    assert abs(gaps.mean() - 1 / 50.0) < 0.2 / 50.0
    assert all(r.x_bits[8] == 0 and r.x_bits[9] == 0 for r in r1 if r.circuit == 'a')


# ---- width planner ---------------------------------------------------------------------------------
def _gf2rank(M):
    M = (M % 2).astype(np.uint8)
    r = 0
    for c in range(M.shape[1]):
        p = [i for i in range(r, M.shape[0]) if M[i, c]]
        if not p:
            continue
        M[[r, p[0]]] = M[[p[0], r]]
        for i in range(M.shape[0]):
            if i != r and M[i, c]:
                M[i] ^= M[r]
        r += 1
        if r == M.shape[0]:
            break
    return r


def test_prefix_cut_ranks_match_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(6):
        N = 15
        A = np.triu((rng.random((N, N)) < 0.3).astype(np.uint8), 1)
        A = A | A.T
        order = [int(v) for v in rng.permutation(N)[:13]]
        want = [_gf2rank(A[np.ix_(order[:i], order[i:])]) for i in range(1, len(order))]
        assert list(prefix_cut_ranks(A, order)) == want


def test_ibm_order_cut_rank_is_36(ibm64):
    """The unsliced IBM graph has cut rank 36 along its frozen order."""
    assert int(prefix_cut_ranks(ibm64['graph'], ibm64['order']).max()) == 36


def test_best_rotation_not_wider_than_natural():
    inst = synthetic_instance(16, 12, 20, seed=0)
    order, opn, w = best_rotation(inst['graph'], inst['unitaries'], 16, inst['natural_order'], 6)
    nat_data = [v for v in inst['natural_order'] if v < 16]
    w_nat = BlockPlan(inst['graph'], inst['unitaries'], 16, inst['natural_order'], [], nat_data[-6:]).inplace_maxslots
    assert w <= w_nat and len(opn) == 6 and sorted(order) == sorted(inst['natural_order'])


def test_greedy_parity_slices_lower_width_and_keep_the_block():
    """Two greedy slices on a 10-qubit ring: the max cut rank does not grow, and the sliced block equals
    the unsliced one up to one global phase (oracle)."""
    inst = synthetic_instance(10, 10, 14, seed=3)
    n, N = 10, inst['N']
    edges = inst['edges']
    refs = set(range(n, N))
    ops, removed = choose_parity_slices(N, edges, inst['natural_order'], 2, seed=0, allowed=refs)
    assert len(ops) == 2 and len(set(removed)) == 2 and set(removed) <= refs
    A = inst['graph']
    before = order_cost(A, inst['natural_order'])[0]
    opn = [0, 4, 7]
    x = np.array([1, 0, 1, 1, 0, 1, 0, 0, 1, 1], np.uint8)
    full = oracle.block(BlockPlan(inst['graph'], inst['unitaries'], n, inst['natural_order'], [], opn), x)
    bp = BlockPlan(inst['graph'], inst['unitaries'], n, inst['natural_order'], ops, opn)
    sl = oracle.block(bp, x)
    i = int(np.argmax(np.abs(full)))
    assert np.max(np.abs(sl - (sl[i] / full[i]) * full)) / np.max(np.abs(full)) < 1e-12
    assert bp.rmax <= before + 1


def test_greedy_slicing_reproduces_the_frozen_ibm_k5_operations(ibm64):
    """The frozen canonical_k5_fp64 plan of the 64-qubit graph is the greedy's choice with seed 0 and
    k = 5 along the frozen order (inputs/plans.json)."""
    A = ibm64['graph']
    edges = [(int(a), int(b)) for a, b in zip(*np.nonzero(np.triu(A, 1)))]
    ops, _ = choose_parity_slices(len(A), edges, ibm64['order'], 5, seed=0)
    want = [tuple(None if x is None else (x if isinstance(x, str) else int(x)) for x in op)
            for op in ibm64['plans']['canonical_k5_fp64']['operations']]
    assert ops == want


def test_fit_width_slices_to_the_budget_and_keeps_the_block(tmp_path):
    """A 16-qubit ring whose best rotation has width 7: fit_width reaches width 5 with greedy slices of T
    references, the sliced block equals the unsliced one up to one global phase (oracle, complex128),
    and a second call is served from the plan cache."""
    from dcs_amd.width_planner import fit_width
    n, k = 16, 4
    inst = synthetic_instance(n, 14, 28, seed=2)
    order, opn, w = best_rotation(inst['graph'], inst['unitaries'], n, inst['natural_order'], k, stride=2)
    assert w == 7
    cache = tmp_path / 'plans.json'
    r = fit_width(inst['graph'], inst['unitaries'], n, inst['natural_order'], k, 5, max_slices=4, cache_path=cache)
    assert r['width'] <= 5 and r['slices'] >= 1 and all(op[1] >= n for op in r['operations'])
    bp = BlockPlan(inst['graph'], inst['unitaries'], n, r['order'], r['operations'], r['open_vertices'])
    assert bp.inplace_maxslots == r['width']
    x = np.random.default_rng(4).integers(0, 2, n, dtype=np.uint8)
    full = oracle.block(BlockPlan(inst['graph'], inst['unitaries'], n, r['order'], [], r['open_vertices']), x)
    sl = oracle.block(bp, x)
    i = int(np.argmax(np.abs(full)))
    assert np.max(np.abs(sl - (sl[i] / full[i]) * full)) / np.max(np.abs(full)) < 1e-12
    again = fit_width(inst['graph'], inst['unitaries'], n, inst['natural_order'], k, 5, max_slices=4, cache_path=cache)
    assert again == r
    assert fit_width(inst['graph'], inst['unitaries'], n, inst['natural_order'], k, 3, max_slices=2) is None
