"""Serving loop: continuous batching of sampling requests over several device-resident circuits.

The reference serves samples one at a time per worker process (production/worker.py:106-133: build one
engine, then for every assigned record contract its block, copy it to the host and draw).  Here one
process keeps several circuits resident on the GPU and serves a stream of requests:

  * admission -- a circuit is registered only if its frontier(s), plan tables and block accumulators fit
    the HBM budget left by the circuits already resident; the number of lanes (slice frontiers held at
    once) is reduced until it fits, and registration fails with MemoryError if even one lane does not;
  * continuous batching -- requests queue per circuit as they arrive; the loop repeatedly takes the
    circuit whose oldest waiting request arrived first and serves up to `max_batch` of its waiting
    requests together (slices of all of them share lane launches, block_many), without waiting for a
    batch to fill;
  * every block is bit-identical to serving its request alone, and so is its draw (block_many);
  * per-request latency (arrival to draw) gives p50 / p99 per circuit;
  * serve(concurrent=True) runs one worker thread per resident circuit, each batching its own queue on
    its engine's HIP stream (default priority), so requests for different circuits are on the GPU at
    the same time (the library releases the interpreter lock during device calls) and a short request
    no longer waits behind another circuit's batch; arrivals are released in real time.

  * Server(narrow_width=w) serves circuits whose frontier width is at most w together: they are
    allocated on one shared context in the fused schedule, and one batch takes the waiting requests of
    all of them through block_batch (their schedules' segments in shared launches), since a narrow
    circuit alone leaves most of the GPU idle.  Off by default.

`poisson_workload` generates seeded arrivals (exponential gaps) over a weighted circuit mix, with
prefixes and uniforms drawn from PCG64 in the order of the reference protocol (sampling_protocol.py:21-33).
"""
from __future__ import annotations

import collections
import json
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .conditional_block import InplaceConditionalBlockGPU


@dataclass
class Request:
    rid: int
    circuit: str
    x_bits: np.ndarray
    uniform: float
    arrival: float                  # seconds after the start of serve()
    start: float = math.nan         # when its batch started
    done: float = math.nan          # when its draw completed
    batch: int = -1
    index: int = -1                 # drawn block index (A_index)
    draw: dict = field(default_factory=dict)
    block: np.ndarray = None        # host copy of the block (serve(keep_blocks=True))

    @property
    def latency(self):
        return self.done - self.arrival


def footprint_bytes(bp, dtype, lanes, blocks):
    """Device bytes an in-place engine will hold: lanes frontiers of 2^maxslots entries, the complex128
    block accumulators and the per-block tables (syndrome map, phase, draw scratch), plus a fixed
    allowance for plan tables and the runtime."""
    esz = np.dtype(dtype).itemsize
    blen = 1 << bp.k
    return (lanes << bp.inplace_maxslots) * esz + blen * (16 * blocks + 8 + 8 + 8 + 16) + (64 << 20)


class Circuit:
    """A resident circuit: its engine holds frontiers, plan tables and block accumulators on the device.
    narrow: served with the other narrow circuits in shared launches (Server(narrow_width=...))."""

    def __init__(self, name, engine, lanes, max_batch, nbytes, narrow=False):
        self.name, self.engine, self.lanes, self.max_batch, self.nbytes = name, engine, lanes, max_batch, nbytes
        self.narrow = narrow
        self.served = 0
        self.last_used = time.perf_counter()


@dataclass
class CircuitSpec:
    """A circuit the server knows: its host plan is built, its device memory is not."""
    name: str
    args: tuple                     # engine arguments (graph, unitaries, n, order, parity ops, open vertices)
    engine: object                  # the engine object (allocated while resident)
    dtype: object
    lanes: int
    max_batch: int
    alloc_kw: dict
    autotune: bool = False
    tune_cache: object = None
    admissions: int = 0


class Server:
    """Circuits and the batching loop on one device.

    add() makes a circuit known (host plan only); it becomes resident -- frontiers, tables and blocks
    allocated -- when a request for it is admitted, and an idle resident circuit (nothing queued, no batch
    in flight) is evicted, least recently used first, when a waiting circuit needs its memory.
    register() adds a circuit and admits it at once."""

    def __init__(self, device=0, hbm_budget=None, reserve=8 << 30, engine_factory=None, narrow_width=0,
                 batch_fn=None, context_factory=None):
        """narrow_width: circuits whose in-place frontier width is at most this are served together in
        shared launches (block_batch; 0 = never).  batch_fn and context_factory replace block_batch and
        the shared device context (stand-ins for host tests)."""
        self.device = device
        if hbm_budget is None:
            hbm_budget = _native.device_info(device)['free_bytes'] - reserve
        self.budget = int(hbm_budget)
        self.used = 0
        self.catalog = collections.OrderedDict()
        self.circuits = collections.OrderedDict()
        self.evictions = 0
        self._factory = engine_factory or InplaceConditionalBlockGPU
        self.narrow_width = int(narrow_width)
        self.narrow_batches = 0
        self._batch_fn = batch_fn
        self._context_factory = context_factory or _native.Context
        self._shared_ctx = None

    def _narrow(self, spec):
        return self.narrow_width > 0 and spec.engine.bp.inplace_maxslots <= self.narrow_width

    def _shared_context(self):
        if self._shared_ctx is None:
            self._shared_ctx = self._context_factory(self.device)
        return self._shared_ctx

    # ---- admission ---------------------------------------------------------------------------
    def add(self, name, graph, unitaries, n, order, parity_operations, open_vertices, dtype=np.complex64,
            lanes=4, max_batch=4, autotune=False, tune_cache=None, **alloc_kw):
        """Make a circuit known without allocating device memory.  Raises MemoryError when even one lane
        exceeds the whole budget (it could never be admitted)."""
        if name in self.catalog:
            raise ValueError(f'circuit {name!r} already known')
        args = (graph, unitaries, n, order, parity_operations, open_vertices)
        engine = self._factory(*args)
        need = footprint_bytes(engine.bp, dtype, 1, max_batch)
        if need > self.budget:
            raise MemoryError(f'circuit {name!r} needs {need} device bytes with one lane, more than the whole '
                              f'{self.budget}-byte budget')
        spec = CircuitSpec(name, args, engine, dtype, max(1, int(lanes)), int(max_batch), dict(alloc_kw), autotune,
                           tune_cache)
        self.catalog[name] = spec
        return spec

    def register(self, name, graph, unitaries, n, order, parity_operations, open_vertices, dtype=np.complex64,
                 lanes=4, max_batch=4, autotune=False, tune_cache=None, **alloc_kw):
        """Add a circuit and admit it now.  The lane count is halved until it fits what the resident
        circuits leave of the budget; MemoryError if one lane does not fit.  autotune=True first picks the
        compiled schedule's tiling for this plan by measurement (dcs_amd.autotune; cached per plan
        fingerprint, optionally in tune_cache)."""
        self.add(name, graph, unitaries, n, order, parity_operations, open_vertices, dtype=dtype, lanes=lanes,
                 max_batch=max_batch, autotune=autotune, tune_cache=tune_cache, **alloc_kw)
        c = self._admit(name)
        if c is None:
            spec = self.catalog.pop(name)
            need = footprint_bytes(spec.engine.bp, dtype, 1, max_batch)
            raise MemoryError(f'circuit {name!r} needs {need} device bytes with one lane; '
                              f'{self.budget - self.used} of the {self.budget}-byte budget are left')
        return c

    def _fitting_lanes(self, spec, left):
        lanes = spec.lanes
        while lanes >= 1 and footprint_bytes(spec.engine.bp, spec.dtype, lanes, spec.max_batch) > left:
            lanes //= 2
        return lanes

    def _victims(self, name, idle):
        """Idle resident circuits to evict (least recently used first) so that `name` fits with one lane,
        or None when evicting all of them is not enough."""
        spec = self.catalog[name]
        need = footprint_bytes(spec.engine.bp, spec.dtype, 1, spec.max_batch)
        left = self.budget - self.used
        out = []
        for c in sorted((self.circuits[v] for v in idle), key=lambda c: c.last_used):
            if left >= need:
                break
            out.append(c.name)
            left += c.nbytes
        return out if left >= need else None

    def _admit(self, name, idle=()):
        """Allocate a known circuit, evicting idle circuits (names in `idle`) if it needs their memory.
        Returns the resident Circuit, or None when it does not fit."""
        spec = self.catalog[name]
        if name in self.circuits:
            return self.circuits[name]
        victims = self._victims(name, idle)
        if victims is None:
            return None
        for v in victims:
            self._evict(v)
        lanes = self._fitting_lanes(spec, self.budget - self.used)
        alloc_kw = dict(spec.alloc_kw)
        narrow = self._narrow(spec)
        if narrow:
            # shared launches run the fused schedule's generic kernels with their default tile shapes
            alloc_kw = {k: v for k, v in alloc_kw.items() if k == 'group_policy'}
            alloc_kw.update(schedule='fused', context=self._shared_context())
        elif spec.autotune:
            # candidates are timed with the lane and block counts this circuit is admitted with
            from .autotune import autotune as tune
            shared = {k: v for k, v in alloc_kw.items() if k in ('lds_storage', 'pass_max_dim', 'resident_bits')}
            shared.update(lanes=lanes, blocks=spec.max_batch)
            res = tune(self._factory, spec.args, dtype=spec.dtype, device=self.device, cache_path=spec.tune_cache,
                       alloc_kw=shared)
            alloc_kw = {**alloc_kw, **res['best'], 'schedule': 'compiled'}
        spec.engine.alloc(dtype=spec.dtype, device=self.device, lanes=lanes, blocks=spec.max_batch, **alloc_kw)
        nbytes = spec.engine.device_bytes() or footprint_bytes(spec.engine.bp, spec.dtype, lanes, spec.max_batch)
        self.used += nbytes
        spec.admissions += 1
        c = Circuit(name, spec.engine, lanes, spec.max_batch, nbytes, narrow=narrow)
        self.circuits[name] = c
        return c

    def _evict(self, name):
        c = self.circuits.pop(name)
        c.engine.close()
        self.used -= c.nbytes
        self.evictions += 1

    def unregister(self, name):
        if name in self.circuits:
            self._evict(name)
        self.catalog.pop(name, None)

    def close(self):
        for name in list(self.catalog):
            self.unregister(name)
        if self._shared_ctx is not None:
            self._shared_ctx.close()
            self._shared_ctx = None

    # ---- the loop ------------------------------------------------------------------------------
    def serve(self, requests, realtime=True, keep_blocks=False, concurrent=False):
        """Serve requests (any order; each has an arrival time in seconds from now).  With realtime=False
        idle gaps are skipped instead of slept through (latencies are then measured on a clock that
        jumps to the next arrival when nothing is waiting).  concurrent=True serves every resident circuit
        from its own worker thread (arrivals in real time).  Requests for a circuit that is not resident
        wait until it is admitted.  Returns the requests, filled in."""
        for r in requests:
            if r.circuit not in self.catalog:
                raise KeyError(f'request {r.rid}: unknown circuit {r.circuit!r}')
        if concurrent:
            return self._serve_concurrent(requests, keep_blocks)
        todo = sorted(requests, key=lambda r: (r.arrival, r.rid))
        queues = {name: collections.deque() for name in self.catalog}
        t0 = time.perf_counter()
        skew = 0.0  # idle time skipped (realtime=False)
        clock = lambda: time.perf_counter() - t0 + skew
        i, nbatch = 0, 0
        while i < len(todo) or any(queues.values()):
            now = clock()
            while i < len(todo) and todo[i].arrival <= now:
                queues[todo[i].circuit].append(todo[i])
                i += 1
            waiting = sorted((q for q in queues.values() if q), key=lambda q: (q[0].arrival, q[0].rid))
            if not waiting:
                gap = todo[i].arrival - now
                if realtime:
                    time.sleep(max(0.0, gap))
                else:
                    skew += max(0.0, gap)
                continue
            c = None
            for q in waiting:       # the oldest waiting request whose circuit is, or can be made, resident
                idle = [name for name in self.circuits if not queues[name]]
                c = self._admit(q[0].circuit, idle)
                if c is not None:
                    break
            if c is None:
                raise RuntimeError('no waiting circuit can be admitted although every resident one is idle')
            if c.narrow:
                # the waiting requests of every resident narrow circuit, in shared launches
                groups = []
                for name, nc in self.circuits.items():
                    if nc.narrow and queues[name]:
                        qn = queues[name]
                        groups.append((nc, [qn.popleft() for _ in range(min(nc.max_batch, len(qn)))]))
                self._run_narrow(groups, clock, nbatch, keep_blocks)
            else:
                batch = [q.popleft() for _ in range(min(c.max_batch, len(q)))]
                self._run_batch(c, batch, clock, nbatch, keep_blocks)
            nbatch += 1
        return requests

    def _run_batch(self, c, batch, clock, nbatch, keep_blocks):
        """Contract a batch of one circuit's requests together and draw each one on the device: the draw
        follows draw_checked (sampling/sampling_protocol.py:44-88) -- numpy's cumsum order for the index,
        and an independent, compensated sum of the drawn interval that must contain the uniform, else
        ArithmeticError and no result (never a second draw)."""
        start = clock()
        vecs, _ = c.engine.block_many([r.x_bits for r in batch])
        for r, v in zip(batch, vecs):
            r.draw = v.draw(r.uniform)
            if keep_blocks:
                r.block = np.asarray(v)
            r.index = r.draw['A_index']
            r.start, r.done, r.batch = start, clock(), nbatch
        c.served += len(batch)
        c.last_used = time.perf_counter()

    def _run_narrow(self, groups, clock, nbatch, keep_blocks):
        """One batch over several narrow circuits: their requests' blocks by block_batch (every circuit's
        schedule segments in shared launches; each block bit-identical to its engine's block() alone),
        then each request's device draw as in _run_batch."""
        batch_fn = self._batch_fn
        if batch_fn is None:
            from .conditional_block import block_batch as batch_fn
        start = clock()
        work = [(c.engine, r.x_bits, i) for c, batch in groups for i, r in enumerate(batch)]
        vecs, _ = batch_fn(work)
        flat = [r for _, batch in groups for r in batch]
        for r, v in zip(flat, vecs):
            r.draw = v.draw(r.uniform)
            if keep_blocks:
                r.block = np.asarray(v)
            r.index = r.draw['A_index']
            r.start, r.done, r.batch = start, clock(), nbatch
        now = time.perf_counter()
        for c, batch in groups:
            c.served += len(batch)
            c.last_used = now
        self.narrow_batches += 1

    def _serve_concurrent(self, requests, keep_blocks):
        todo = sorted(requests, key=lambda r: (r.arrival, r.rid))
        cv = threading.Condition()
        queues = {name: collections.deque() for name in self.catalog}
        running, stop, workers = set(), set(), {}
        state = {'released': 0, 'batches': 0, 'error': None, 'narrow_alive': False, 'narrow_threads': 0}
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0

        def worker(name, c):
            while True:
                with cv:
                    cv.wait_for(lambda: queues[name] or name in stop or state['error']
                                or state['released'] == len(todo))
                    if state['error'] or name in stop or not queues[name]:
                        return
                    q = queues[name]
                    batch = [q.popleft() for _ in range(min(c.max_batch, len(q)))]
                    nb = state['batches']
                    state['batches'] += 1
                    running.add(name)
                try:
                    self._run_batch(c, batch, clock, nb, keep_blocks)
                except BaseException as exc:     # stop every worker, re-raise in the caller
                    with cv:
                        state['error'] = exc
                        running.discard(name)
                        cv.notify_all()
                    return
                with cv:
                    running.discard(name)
                    cv.notify_all()

        def narrow_ready():
            return [n for n, c in self.circuits.items() if c.narrow and queues[n] and n not in stop and n not in running]

        def narrow_worker():
            # one thread serves every resident narrow circuit: each batch takes the waiting requests of all
            # of them (shared launches); it leaves when nothing narrow is waiting and every request has been
            # released, and the dispatcher starts a new one if a narrow circuit is admitted after that
            while True:
                with cv:
                    cv.wait_for(lambda: narrow_ready() or state['error'] or state['released'] == len(todo))
                    names = narrow_ready()
                    if state['error'] or not names:
                        state['narrow_alive'] = False
                        return
                    groups = []
                    for n in names:
                        c, q = self.circuits[n], queues[n]
                        groups.append((c, [q.popleft() for _ in range(min(c.max_batch, len(q)))]))
                    nb = state['batches']
                    state['batches'] += 1
                    running.update(names)
                try:
                    self._run_narrow(groups, clock, nb, keep_blocks)
                except BaseException as exc:
                    with cv:
                        state['error'] = exc
                        running.difference_update(names)
                        state['narrow_alive'] = False
                        cv.notify_all()
                    return
                with cv:
                    running.difference_update(names)
                    cv.notify_all()

        def start(name):
            # called with the lock held, or before any worker runs
            if self.circuits[name].narrow:
                if not state['narrow_alive']:
                    state['narrow_alive'] = True
                    state['narrow_threads'] += 1
                    t = threading.Thread(target=narrow_worker, daemon=True)
                    workers[('narrow', state['narrow_threads'])] = t
                    t.start()
                return
            t = threading.Thread(target=worker, args=(name, self.circuits[name]), daemon=True)
            workers[name] = t
            t.start()

        for name in self.circuits:
            start(name)
        i = 0
        try:
            while True:
                with cv:
                    if state['error']:
                        break
                    now = clock()
                    while i < len(todo) and todo[i].arrival <= now:
                        queues[todo[i].circuit].append(todo[i])
                        state['released'] += 1
                        i += 1
                        cv.notify_all()
                    pending = sorted((q for name, q in queues.items() if q and name not in self.circuits),
                                     key=lambda q: (q[0].arrival, q[0].rid))
                    if i == len(todo) and not pending:
                        break
                    idle = [n for n in self.circuits if not queues[n] and n not in running]
                    plan = None
                    for q in pending:
                        victims = self._victims(q[0].circuit, idle)
                        if victims is not None:
                            plan = (q[0].circuit, victims)
                            break
                    if plan is None:
                        # wait for the next arrival or for a worker to finish a batch
                        gap = todo[i].arrival - now if i < len(todo) else 0.05
                        cv.wait(timeout=max(1e-4, min(gap, 0.05)))
                        continue
                    stop.update(plan[1])
                    cv.notify_all()
                # outside the lock: let the evicted circuits' workers leave, free them, admit the new one
                name, victims = plan
                for v in victims:
                    t = workers.pop(v, None)    # narrow circuits have no worker of their own
                    if t is not None:
                        t.join()
                    self._evict(v)
                    stop.discard(v)
                self._admit(name)
                with cv:
                    start(name)
                    cv.notify_all()
        finally:
            with cv:
                state['released'] = len(todo)
                cv.notify_all()
            for t in list(workers.values()):
                t.join()
        if state['error'] is not None:
            raise state['error']
        return requests


def latency_stats(requests):
    """Per-circuit and overall count, p50 / p99 latency (ms) and makespan."""
    out = {}
    groups = collections.defaultdict(list)
    for r in requests:
        groups[r.circuit].append(r.latency)
        groups['all'].append(r.latency)
    for name, lat in groups.items():
        a = np.array(lat) * 1e3
        out[name] = {'count': len(a), 'p50_ms': float(np.percentile(a, 50)), 'p99_ms': float(np.percentile(a, 99)),
                     'mean_ms': float(a.mean())}
    span = max(r.done for r in requests) - min(r.arrival for r in requests)
    out['all']['samples_per_s'] = len(requests) / span if span > 0 else math.inf
    return out


def poisson_workload(mix, rate, count, seed):
    """Seeded request stream.  mix: {name: (weight, n, open_vertices)}; rate: mean arrivals per second.
    Each request draws its circuit by weight, its closed data bits (open ones are 0) and its uniform
    from one PCG64 stream, as the reference protocol draws prefixes and uniforms (sampling_protocol.py:21-33)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    names = list(mix)
    w = np.array([mix[k][0] for k in names], np.float64)
    w /= w.sum()
    t, reqs = 0.0, []
    for rid in range(count):
        t += rng.exponential(1.0 / rate)
        name = names[int(rng.choice(len(names), p=w))]
        _, n, opn = mix[name]
        closed = [q for q in range(n) if q not in set(opn)]
        x = np.zeros(n, np.uint8)
        x[closed] = rng.integers(0, 2, len(closed), dtype=np.uint8)
        reqs.append(Request(rid, name, x, float(rng.random()), t))
    return reqs


# ---- the serving benchmark (C5-style mix on one GPU) --------------------------------------------
def _synthetic(n, depth, t, seed, k):
    """A seeded ring circuit with the order rotation of least frontier width and its last k data
    qubits open."""
    from .width_planner import best_rotation
    from .stabilizer import synthetic_instance
    inst = synthetic_instance(n, depth, t, seed)
    order, opn, width = best_rotation(inst['graph'], inst['unitaries'], n, inst['natural_order'], k, stride=2)
    return (inst['graph'], inst['unitaries'], n, order, [], opn), width


RING_PLANS = __import__('pathlib').Path(__file__).resolve().parent / 'data' / 'ring_plans.json'


def c5_mix(srv, plan_cache=RING_PLANS, with_2_22=True, log=None):
    """Add the serving mix of ring Clifford+T circuits to a server (nothing is allocated until a request
    is admitted): 40q/d48/180T with 2^16 blocks (best ring rotation, width 31), 52q/d60/250T with 2^18
    blocks (circuit seed 1: best rotation plus greedy parity slices of T references down to width 33,
    8 slices per block) and the 64-qubit IBM QAT #245 graph with 2^20 and 2^22 blocks (plan
    inplace_k2_fp32, width 34; the 2^22 circuit opens the two next data qubits of the order), weighted
    3 : 2 : 1 (the 64-qubit weight split between the two block sizes).  Returns the Poisson mix table."""
    from .stabilizer import synthetic_instance
    from .width_planner import fit_width
    from .workloads import load_ibm64
    kw = dict(schedule='compiled', lds_storage=True)
    args40, w40 = _synthetic(40, 48, 180, 0, 16)
    srv.add('40q/d48/180T', *args40, lanes=4, max_batch=4, **kw)
    inst = synthetic_instance(52, 60, 250, 1)
    fit = fit_width(inst['graph'], inst['unitaries'], 52, inst['natural_order'], 18, 33, max_slices=4,
                    cache_path=plan_cache)
    srv.add('52q/d60/250T', inst['graph'], inst['unitaries'], 52, fit['order'], fit['operations'],
            fit['open_vertices'], lanes=1, max_batch=2, **kw)
    w = load_ibm64()
    ops = w['plans']['inplace_k2_fp32']['operations']
    opn20 = [int(v) for v in w['open_vertices']]
    srv.add('64q/2^20', w['graph'], w['unitaries'], 64, w['order'], ops, opn20, lanes=1, max_batch=2, **kw)
    mix = {'40q/d48/180T': (3, 40, args40[5]), '52q/d60/250T': (2, 52, fit['open_vertices']),
           '64q/2^20': (0.5 if with_2_22 else 1, 64, opn20)}
    if with_2_22:
        extra = [int(v) for v in w['order'] if v < 64 and int(v) not in set(opn20)][::-1][:2]
        srv.add('64q/2^22', w['graph'], w['unitaries'], 64, w['order'], ops, opn20 + extra, lanes=1, max_batch=2,
                **kw)
        mix['64q/2^22'] = (0.5, 64, opn20 + extra)
    if log:
        log(f'widths: 40q {w40}, 52q {fit["width"]} with {fit["slices"]} slices, 64q 34')
    return mix, {'40q/d48/180T': w40, '52q/d60/250T': fit['width'], '52q_slices': fit['slices'], '64q': 34}


# Seeded ring circuits of 16-32 qubits (qubits, depth, T count, seed, open qubits): frontier widths 7-19
# with the best ring rotation, 2^8 - 2^12 blocks.
NARROW_MIX = [(16, 12, 20, 0, 8), (20, 16, 40, 1, 10), (24, 24, 80, 3, 8), (28, 32, 96, 1, 10),
              (32, 30, 100, 0, 12), (24, 20, 60, 2, 10), (28, 24, 80, 2, 10), (32, 28, 90, 1, 12)]


def narrow_mix(srv, lanes=4, max_batch=4):
    """Add the eight narrow ring circuits of NARROW_MIX, equally weighted; returns the Poisson mix table."""
    mix = {}
    for n, d, t, seed, k in NARROW_MIX:
        args, w = _synthetic(n, d, t, seed, k)
        name = f'{n}q/d{d}/{t}T/s{seed}'
        srv.add(name, *args, lanes=lanes, max_batch=max_batch, schedule='compiled', lds_storage=True)
        mix[name] = (1, n, args[5])
    return mix


def _serve_narrow_mix(a, log):
    """The narrow mix served twice from the same seeded stream: one compiled engine and worker per
    circuit, then every circuit in shared launches (narrow_width 20).  One JSON line."""
    out = {'metric': 'serving narrow circuits: p50/p99 ms, per-circuit engines vs shared launches',
           'requests': a.requests, 'rate_per_s': a.rate, 'seed': a.seed, 'concurrent': not a.sequential,
           'circuits': [f'{n}q/d{d}/{t}T seed {s}, 2^{k} block' for n, d, t, s, k in NARROW_MIX]}
    for label, width in (('per_circuit', 0), ('shared_launches', 20)):
        srv = Server(narrow_width=width)
        mix = narrow_mix(srv)
        for j, name in enumerate(mix):          # warm-up outside the statistics
            srv.serve([Request(-1 - j, name, np.zeros(mix[name][1], np.uint8), 0.5, 0.0)], realtime=False)
        reqs = poisson_workload(mix, a.rate, a.requests, a.seed)
        t1 = time.perf_counter()
        srv.serve(reqs, realtime=True, concurrent=not a.sequential)
        span = time.perf_counter() - t1
        st = latency_stats(reqs)['all']
        out[label] = {'serve_s': span, 'samples_per_s': len(reqs) / span, 'p50_ms': st['p50_ms'],
                      'p99_ms': st['p99_ms'], 'mean_ms': st['mean_ms'], 'shared_batches': srv.narrow_batches}
        log(label, out[label])
        srv.close()
    print(json.dumps(out), flush=True)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description='Serving benchmark: Poisson mix of ring circuits on one GPU')
    ap.add_argument('--requests', type=int, default=36)
    ap.add_argument('--rate', type=float, default=2.0, help='mean arrivals per second')
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--no-2-22', action='store_true', help='64-qubit requests with 2^20 blocks only')
    ap.add_argument('--sequential', action='store_true', help='one loop instead of one worker thread per circuit')
    ap.add_argument('--narrow-mix', action='store_true',
                    help='instead: eight narrow circuits, per-circuit engines against shared launches')
    a = ap.parse_args(argv)
    log = lambda *m: print(*m, file=__import__('sys').stderr, flush=True)
    if a.narrow_mix:
        _serve_narrow_mix(a, log)
        return
    srv = Server()
    t0 = time.perf_counter()
    mix, widths = c5_mix(srv, with_2_22=not a.no_2_22, log=log)
    # warm-up: one request per circuit, alone (admission, hipRTC compile, first launches), outside the stats
    for j, name in enumerate(mix):
        srv.serve([Request(-1 - j, name, np.zeros(mix[name][1], np.uint8), 0.5, 0.0)], realtime=False)
        log(f'warm {name}: resident {list(srv.circuits)}')
    setup = time.perf_counter() - t0
    evictions0 = srv.evictions
    reqs = poisson_workload(mix, a.rate, a.requests, a.seed)
    t1 = time.perf_counter()
    srv.serve(reqs, realtime=True, concurrent=not a.sequential)
    span = time.perf_counter() - t1
    stats = latency_stats(reqs)
    line = {'metric': 'serving: continuous batching with admission by HBM footprint, p50/p99 ms per circuit',
            'requests': a.requests, 'rate_per_s': a.rate, 'seed': a.seed, 'concurrent': not a.sequential,
            'setup_s': setup, 'serve_s': span, 'samples_per_s': len(reqs) / span,
            'evictions_during_serving': srv.evictions - evictions0, 'hbm_budget_bytes': srv.budget,
            'admissions': {name: spec.admissions for name, spec in srv.catalog.items()},
            'widths': widths, 'stats': stats}
    print(json.dumps(line), flush=True)
    srv.close()


if __name__ == '__main__':
    main()
