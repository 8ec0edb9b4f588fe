"""Benchmark: samples/s of the 64-qubit, 2^20-block conditional sampler on one MI355X.

Workload: the frozen 378-vertex injected graph of the 64-qubit, depth-73, 314-T ring circuit (IBM
QAT #245), 20 free output bits (2^20 complex128 block per sample), 44-bit prefixes and draw uniforms
replayed from the frozen PCG64 protocol (its first samples).  One step = one sample:
every slice contraction, the coherent slice sum, the Walsh-Hadamard / phase / Clifford tail and the
device CDF draw.  Inputs (graph, plan, frontier buffer) are resident before the timed region.

Fast path (value): complex64 storage, plan inplace_k2_fp32 (4 slices, rmax 34, 128 GiB frontier per
lane; two prefixes at a time by default, their slices sharing each launch; each slice pair's launch
sequence replayed from a captured HIP graph),
compiled schedule (LDS-resident runs + one generated kernel per coset group, complex64 LDS tiles
filled by global-to-LDS loads; coset groups of dimension <= 8 on 2^12-entry tiles with the
fewest-HBM-bytes group boundaries).
Plain path (reported beside it): complex128
storage, plan inplace_k3_fp64 (8 slices, rmax 33), one launch per step and a host read of every
maximum, one prefix at a time -- the reference method.

The line also carries the roofline of the dominant kernels (the coset-group kernels, HBM-bound)
measured with HIP events on the library's stream, and a CPU baseline: the C oracle (OpenMP) on a bounded sample of the
same workload (the first steps of one slice, live widths <= 2^32), extrapolated by bytes.
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / 'dcs245-classical-sampling_amd'))
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12
# Generation seed of the frozen 64-qubit protocol (its 1,389 prefixes and uniforms are replayed).
PROTOCOL_SEED = 10636914113228816500957915140926462448


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def prefixes(count):
    from dcs_amd.sampling_protocol import make_protocol
    from dcs_amd.workloads import load_ibm64
    w = load_ibm64()
    p = make_protocol(PROTOCOL_SEED, count, 64, w['open_vertices'], 1.0)
    return [(np.array([int(c) for c in r['input_data_bits']], np.uint8), r['draw_uniform']) for r in p['records']]


# FP64 flops per live input entry of a step, by orbit code (kernel_opt/inplace_gpu.py:29-59): a complex
# product is 6 flops, a complex sum 2.  0 grow: f0 a and f1 a; 1 mix: f0 a + f1 b per output; 2, 4: one
# product; 3 replace: f0 (a + b) and f1 (a +- b) per pair; 5: f (a +- b) per pair; 6 shrink: two outputs
# f0 (a + b) + f1 (e +- g) per quad; 7: W0 a + W1 b per pair.
FLOPS_PER_INPUT = {0: 12, 1: 14, 2: 6, 3: 8, 4: 6, 5: 4, 6: 9, 7: 7}
# The same for an f0-factored step (coefficients (1, r), r = f1 / f0) whose ratio r is real or imaginary
# (a product by r is 2 flops, a + r b is 4) and for a complex r (6 and 8): 0: r a; 1: a + r b per output;
# 2: x + r x; 3: a + b and r (a +- b) per pair; 4: r x; 5: r (a +- b) per pair; 6: (a + b) + r (e +- g) per
# output of a quad; 7: (a + b) + r (a +- b) per pair.
FLOPS_PER_INPUT_FACTORED = {1: {0: 2, 1: 4, 2: 4, 3: 3, 4: 2, 5: 2, 6: 4, 7: 4},
                            3: {0: 6, 1: 8, 2: 8, 3: 5, 4: 6, 5: 4, 6: 6, 7: 6}}
FP64_PEAK = 78.6e12          # MI355X FP64 vector, FLOP/s
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4   # wave64 VALU instructions per second: 4 SIMDs per CU, 4 cycles each


def fused_flops_per_sample(engine, factored=False):
    """Algorithmic FP64 flops of the steps inside coset groups, per sample (live inputs only); with
    factored, those of the f0-factored kernels (every group's last step and the steps of class 0 keep
    the unfactored formulas)."""
    bp = engine.bp
    kinds = bp.step_table()['kind']
    cls = bp.coefficient_classes()
    st = bp.inplace_steps
    total = 0
    for kind, first, count, _ in engine.schedule():
        if kind != 2:
            continue
        for i in range(first, first + count):
            k = int(kinds[i])
            c = int(cls[i]) if factored and i < first + count - 1 else 0
            per = FLOPS_PER_INPUT[k] if c == 0 else FLOPS_PER_INPUT_FACTORED[3 if c == 3 else 1][k]
            total += per << len(st[i].active_in)
    return total << len(bp.slices)


def fused_bytes_per_sample(engine, elem):
    """Algorithmic HBM bytes of the fused schedule per sample, per segment kind."""
    bp = engine.bp
    st = bp.inplace_steps
    out = {0: 0, 1: 0, 2: 0}
    for kind, first, count, _ in engine.schedule():
        last = first + count - 1
        out[kind] += ((1 << len(st[first].active_in)) + (1 << len(st[last].active_out))) * elem
    nslices = 1 << len(bp.slices)
    return {k: v * nslices for k, v in out.items()}


def run_path(plan_name, dtype, schedule, nsamples, warmup, profile=False, lanes=1, **alloc_kw):
    """Contract nsamples prefixes `lanes` at a time (block_many: the slices of `lanes` prefixes share
    each launch), after `warmup` untimed ones.  With lanes > 1 a request's latency is its batch's."""
    from dcs_amd.conditional_block import InplaceConditionalBlockGPU
    from dcs_amd.workloads import load_ibm64
    w = load_ibm64()
    e = InplaceConditionalBlockGPU(w['graph'], w['unitaries'], 64, w['order'], w['plans'][plan_name]['operations'],
                                   w['open_vertices'])
    t0 = time.perf_counter()
    e.alloc(dtype=dtype, schedule=schedule, lanes=lanes, blocks=lanes, **alloc_kw)
    setup = time.perf_counter() - t0
    work = prefixes(warmup + nsamples + lanes)

    kept = {}

    def run(batch, keep=False):
        vecs, _ = e.block_many([x for x, _ in batch])
        out = []
        for v, (_, u) in zip(vecs, batch):
            d = v.draw(u)
            if keep and not kept:     # the first prefix's block and draw, for the parity check
                kept.update(block=np.asarray(v), draw=d)
            out.append(d['A_index'])
        return out
    for i in range(0, warmup, lanes):
        run(work[i:min(warmup, i + lanes)], keep=True)
    e.ctx.synchronize()
    lat = []
    draws = []
    t0 = time.perf_counter()
    for i in range(warmup, warmup + nsamples, lanes):
        batch = work[i:min(warmup + nsamples, i + lanes)]
        s = time.perf_counter()
        draws += run(batch, keep=warmup == 0 and i == 0)
        lat += [time.perf_counter() - s] * len(batch)
    e.ctx.synchronize()
    elapsed = time.perf_counter() - t0
    prof = None
    if profile:
        # one more batch outside the timed region with HIP events around every launch, per sample
        e._plan.profile(True)
        run(work[warmup + nsamples:warmup + nsamples + lanes])
        ms, n = e._plan.profile(False)
        prof = (ms / lanes, n / lanes)
    info = {'engine': e, 'options': dict(e.schedule_options), 'elapsed': elapsed, 'lat': lat, 'setup': setup,
            'draws': draws, 'profile': prof, 'lanes': lanes, 'first': kept,
            'samples_run': warmup + nsamples + (lanes if profile else 0)}
    return info


def first_prefix_parity(fast, plain, uniform, tie=1e-7):
    """The fast path's complex64 block of the first protocol prefix against the plain complex128 block
    of the same prefix (different slice plans: one global phase aligned), and the two draws."""
    a, r = fast['block'], plain['block']
    ratio = np.vdot(r, a) / np.vdot(r, r)
    err = float(np.max(np.abs(a - (ratio / abs(ratio)) * r)) / np.max(np.abs(r)))
    lo, hi = plain['draw']['cdf_lower_numpy'], plain['draw']['cdf_upper_numpy']
    margin = min(uniform - lo, hi - uniform)
    same = fast['draw']['A_index'] == plain['draw']['A_index']
    return {'prefixes': 1, 'max_rel_amp_error_vs_c128': err, 'draw_identical': same,
            'cdf_edge_margin': margin, 'tie': bool(margin < tie)}


def cpu_baseline(max_width=32):
    """C oracle, complex64 storage, on the first steps of one k2 slice whose live width is <= max_width."""
    from oracle import oracle
    from dcs_amd.block_plan import BlockPlan
    from dcs_amd.workloads import load_ibm64
    import ctypes as C
    w = load_ibm64()
    bp = BlockPlan(w['graph'], w['unitaries'], 64, w['order'], w['plans']['inplace_k2_fp32']['operations'],
                   w['open_vertices'])
    x, _ = prefixes(1)[0]
    f_full = bp.local_vectors(x)
    _, coef, f, yfix, _ = next(bp.slice_assignments(f_full))
    t = bp.step_table()
    coeffs, ov, ys, _ = bp.step_parameters(f, yfix)
    m = 0
    while m < len(bp.inplace_steps) and max(len(bp.inplace_steps[m].active_in), len(bp.inplace_steps[m].active_out)) <= max_width:
        m += 1
    nbytes = sum(((1 << len(s.active_in)) + (1 << len(s.active_out))) * 8 for s in bp.inplace_steps[:m])
    E = np.zeros(1 << max_width, np.complex64)
    E[:] = 0  # touch every page before the timed region
    m2 = np.zeros(m, np.float64)
    arr = {k: np.ascontiguousarray(t[k][:m]) for k in ('kind', 'K', 'Ka', 'pivot', 'pivot_a', 'beta', 'in_mask')}
    P = lambda a, ct: a.ctypes.data_as(C.POINTER(ct))
    L = oracle.lib()
    t0 = time.perf_counter()
    L.oracle_contract(0, m, P(arr['kind'], C.c_int32), P(arr['K'], C.c_uint64), P(arr['Ka'], C.c_uint64),
                      P(arr['pivot'], C.c_int32), P(arr['pivot_a'], C.c_int32), P(arr['beta'], C.c_uint64),
                      P(arr['in_mask'], C.c_uint64), P(np.ascontiguousarray(coeffs[:m]), C.c_double),
                      P(np.ascontiguousarray(ov[:m]), C.c_int32), P(np.ascontiguousarray(ys[:m]), C.c_int32),
                      E.ctypes.data_as(C.c_void_p), P(m2, C.c_double))
    dt = time.perf_counter() - t0
    per_sample = bp.bytes_per_slice(8) * (1 << len(bp.slices))
    rate = nbytes / dt
    return {'value': rate / per_sample, 'unit': 'samples/s', 'cores': int(L.oracle_threads()), 'kind': 'port',
            'sample': f'C oracle (OpenMP, complex64 storage, FP64 arithmetic) over the first {m} steps of one '
                      f'k2 slice (live width <= 2^{max_width}), {nbytes / 1e9:.1f} GB in {dt:.1f} s '
                      f'= {rate / 1e9:.2f} GB/s, extrapolated by the {per_sample:.4g} algorithmic bytes of a sample',
            'seconds': dt}


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def config_lines(repeats=3):
    """The smaller configurations measured directly on both sides: the 16q/d12/20T ring (all 2^16
    amplitudes), the 32q/d40/120T ring (width 29, 2^16 blocks) and the 40q/d48/180T ring (width 31, 2^16
    blocks), each on the GPU (complex64 storage, compiled schedule) and with the CPU oracle (OpenMP,
    complex64 storage, FP64 arithmetic) over whole samples of the same seeded prefixes.  One JSON line per
    configuration."""
    from oracle import oracle
    from dcs_amd.conditional_block import InplaceConditionalBlockGPU
    from dcs_amd.sampling_protocol import make_protocol
    from dcs_amd.serving import _synthetic
    from dcs_amd.stabilizer import synthetic_instance
    c1 = synthetic_instance(16, 12, 20, seed=0)
    cfgs = [('16q/d12/20T ring, all 2^16 amplitudes', (c1['graph'], c1['unitaries'], 16, c1['natural_order'], [],
                                                       list(range(16))), 1, 1)]
    for n, d, t, lanes in ((32, 40, 120, 8), (40, 48, 180, 4)):
        args, w = _synthetic(n, d, t, 0, 16)
        cfgs.append((f'{n}q/d{d}/{t}T ring, width {w}, 2^16 block', args, 8, lanes))
    threads = int(oracle.lib().oracle_threads())
    for name, args, nprefix, lanes in cfgs:
        eng = InplaceConditionalBlockGPU(*args)
        bp = eng.bp
        prot = make_protocol(0, nprefix, args[2], args[5], 1.0)
        xs = [np.array([int(c) for c in r['input_data_bits']], np.uint8) for r in prot['records']]
        us = [r['draw_uniform'] for r in prot['records']]
        nbytes = bp.bytes_per_slice(8) * (1 << len(bp.slices))
        eng.alloc(dtype=np.complex64, lanes=lanes, blocks=lanes)
        for i in range(0, len(xs), lanes):                      # warm-up: first launches
            eng.block_many(xs[i:i + lanes])
        eng.ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            for i in range(0, len(xs), lanes):
                vecs, _ = eng.block_many(xs[i:i + lanes])
                for v, u in zip(vecs, us[i:i + lanes]):
                    v.draw(u)
        gpu_s = (time.perf_counter() - t0) / (repeats * len(xs))
        first = np.asarray(eng.block(xs[0])[0])
        eng.close()
        t0 = time.perf_counter()
        ref = oracle.block(bp, xs[0], storage=0)
        oracle.draw(ref, us[0])
        cpu_s = time.perf_counter() - t0
        err = float(np.max(np.abs(first - ref)) / np.max(np.abs(ref)))
        print(json.dumps({'config': name, 'width': bp.inplace_maxslots, 'block': 1 << bp.k,
                          'algorithmic_bytes_per_sample': nbytes,
                          'gpu': {'samples_per_s': 1 / gpu_s, 'ms_per_sample': 1e3 * gpu_s, 'GBps': nbytes / gpu_s / 1e9,
                                  'prefixes': len(xs), 'lanes': lanes, 'storage': 'complex64', 'schedule': 'compiled'},
                          'cpu': {'samples_per_s': 1 / cpu_s, 'ms_per_sample': 1e3 * cpu_s, 'GBps': nbytes / cpu_s / 1e9,
                                  'threads': threads, 'nproc': os.cpu_count(), 'model': cpu_model(),
                                  'kind': 'port', 'samples': 1},
                          'gpu_over_cpu': cpu_s / gpu_s, 'max_rel_amp_diff_gpu_vs_cpu_c64': err}), flush=True)




def batch_line(prefixes=4, repeats=5):
    """Small circuits of uneven frontier widths (seeded ring circuits, widths 7 to 19, 2^8 - 2^12
    blocks), `prefixes` requests each: one JSON line with the throughput of block_batch (every circuit's
    segments in shared launches, generic kernels) against the same engines one request at a time, and
    against each circuit's own compiled engine with its requests batched in lanes."""
    from dcs_amd import _native
    from dcs_amd.block_plan import BlockPlan
    from dcs_amd.conditional_block import InplaceConditionalBlockGPU, block_batch
    from dcs_amd.sampling_protocol import make_protocol
    from dcs_amd.serving import NARROW_MIX
    from dcs_amd.stabilizer import synthetic_instance
    from dcs_amd.width_planner import best_rotation
    cases = []
    for n, d, t, seed, k in NARROW_MIX:
        inst = synthetic_instance(n, d, t, seed)
        order, opn, _ = best_rotation(inst['graph'], inst['unitaries'], n, inst['natural_order'], k, stride=2)
        args = (inst['graph'], inst['unitaries'], n, order, [], opn)
        prot = make_protocol(seed, prefixes, n, opn, 1.0)
        xs = [np.array([int(c) for c in r['input_data_bits']], np.uint8) for r in prot['records']]
        cases.append((args, xs, BlockPlan(*args).inplace_maxslots))
    ctx = _native.Context(0)
    fused = []
    for args, xs, _ in cases:
        e = InplaceConditionalBlockGPU(*args)
        e.alloc(dtype=np.complex64, schedule='fused', lanes=prefixes, blocks=prefixes + 1, context=ctx)
        fused.append(e)
    requests = [(e, x, b) for e, (_, xs, _) in zip(fused, cases) for b, x in enumerate(xs)]
    nreq = len(requests)

    def timed(fn, sync):
        fn()
        sync()
        t0 = time.perf_counter()
        for _ in range(repeats):
            fn()
        sync()
        return (time.perf_counter() - t0) / repeats

    meta = {}

    def batched():
        meta.update(block_batch(requests)[1])

    t_batch = timed(batched, ctx.synchronize)
    first = np.asarray(block_batch(requests[:1])[0][0])
    t_alone = timed(lambda: [e.block(x, block_index=b) for e, x, b in requests], ctx.synchronize)
    same = bool(np.array_equal(np.asarray(requests[0][0].block(requests[0][1], block_index=prefixes)[0]), first))
    segments = sum(len(e.schedule()) for e in fused)
    for e in fused:
        e.close()
    ctx.close()
    compiled = []
    for args, xs, _ in cases:
        e = InplaceConditionalBlockGPU(*args)
        e.alloc(dtype=np.complex64, lanes=prefixes, blocks=prefixes)
        compiled.append((e, xs))
    t_comp = timed(lambda: [e.block_many(xs) for e, xs in compiled],
                   lambda: [e.ctx.synchronize() for e, _ in compiled])
    for e, _ in compiled:
        e.close()
    print(json.dumps({'config': 'batch of small circuits of uneven width',
                      'circuits': [f'{n}q/d{d}/{t}T seed {s}, width {w}, 2^{k} block'
                                   for (n, d, t, s, k), (_, _, w) in zip(NARROW_MIX, cases)],
                      'requests': nreq, 'storage': 'complex64',
                      'block_batch': {'samples_per_s': nreq / t_batch, 'ms_per_round': 1e3 * t_batch,
                                      'launches': meta.get('launches'), 'rounds': meta.get('rounds')},
                      'fused_one_at_a_time': {'samples_per_s': nreq / t_alone, 'ms': 1e3 * t_alone,
                                              'segment_launches': segments * prefixes},
                      'compiled_per_circuit_lanes': {'samples_per_s': nreq / t_comp, 'ms': 1e3 * t_comp},
                      'batch_equals_alone_first_request': same}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=4)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--plain-samples', type=int, default=1)
    ap.add_argument('--schedule', default='compiled', choices=['compiled', 'fused', 'steps', 'steps_host'])
    ap.add_argument('--lds-fp64', action='store_true', help='compiled schedule: FP64 LDS tiles')
    ap.add_argument('--lds-buffers', type=int, default=None, help='compiled schedule: tile pipeline (engine default if unset)')
    ap.add_argument('--no-factor', action='store_true', help='compiled schedule: group kernels without f0 factoring')
    ap.add_argument('--no-graphs', action='store_true',
                    help='launch every kernel of a slice directly instead of replaying its captured HIP graph')
    ap.add_argument('--lanes', type=int, default=2,
                    help='prefixes contracted together (frontiers held at once); falls back to 1 if they do not fit')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-plain', action='store_true')
    ap.add_argument('--configs', action='store_true',
                    help='instead: the 16q, 32q and 40q ring configurations, GPU and CPU oracle, one line each')
    ap.add_argument('--batch-mix', action='store_true',
                    help='instead: eight small circuits of uneven width, shared launches against one at a time')
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({'error': 'this benchmark runs on exactly one GPU; the sampler shards by sample '
                          'and needs no collectives', 'n_gpus': a.gpus}))
        sys.exit(2)
    if a.configs:
        config_lines()
        return
    if a.batch_mix:
        batch_line()
        return
    from dcs_amd import _native
    dev = _native.device_info(0)
    log('device', dev)
    # streaming copy of 4 GiB (read + write): the practical HBM ceiling, and the calibration kernel
    # (known bytes) for the FETCH_SIZE / WRITE_SIZE counter runs
    ctx = _native.Context(0)
    copy_bytes = 4 << 30
    copy_ms = ctx.stream_copy_ms(copy_bytes, 5)
    ctx.close()
    stream_GBps = 2 * copy_bytes / (copy_ms / 1e3) / 1e9
    log(f'stream copy (best launch shape): {stream_GBps:.1f} GB/s')

    log('fast path: complex64, k2 plan, schedule', a.schedule)
    kw = {'lds_storage': not a.lds_fp64, 'factor': not a.no_factor} if a.schedule == 'compiled' else {}
    if a.lds_buffers is not None:
        kw['lds_buffers'] = a.lds_buffers
    if a.schedule != 'steps_host':
        kw['graphs'] = not a.no_graphs
    try:
        fast = run_path('inplace_k2_fp32', np.complex64, a.schedule, a.steps, a.warmup, profile=True, lanes=a.lanes, **kw)
    except MemoryError as exc:
        log(f'{a.lanes} lanes do not fit ({exc}); one lane')
        fast = run_path('inplace_k2_fp32', np.complex64, a.schedule, a.steps, a.warmup, profile=True, lanes=1, **kw)
    eng = fast['engine']
    ms, launches = fast['profile']
    sps = a.steps / fast['elapsed']
    bp = eng.bp
    unfused_bytes = bp.bytes_per_slice(8) * (1 << len(bp.slices))
    fb = fused_bytes_per_sample(eng, 8)
    factored = eng.factoring()[1] if a.schedule == 'compiled' else False
    fused_flops = fused_flops_per_sample(eng, factored)
    coset_ms = ms[2]
    coset_bytes = fb[2]
    achieved = coset_bytes / (coset_ms / 1e3) if coset_ms > 0 else None
    log(f'fast: {sps:.4f} samples/s, kernel ms/sample: step {ms[0]:.1f} resident {ms[1]:.1f} '
        f'coset {coset_ms:.1f} block {ms[3]:.1f} draw {ms[4]:.1f}; launches {launches.tolist()}')
    eng.close()
    del eng, fast['engine']

    plain = None
    parity = None
    if not a.no_plain and a.plain_samples > 0:
        log('plain path: complex128, k3 plan, per-step launches with host maxima')
        p = run_path('inplace_k3_fp64', np.complex128, 'steps_host', a.plain_samples, 0)
        pb = p['engine'].bp
        pbytes = pb.bytes_per_slice(16) * (1 << len(pb.slices))
        plain = {'samples_per_s': a.plain_samples / p['elapsed'], 'ms_per_sample': 1e3 * p['elapsed'] / a.plain_samples,
                 'algorithmic_bytes_per_sample': pbytes,
                 'hbm_GBps': pbytes * a.plain_samples / p['elapsed'] / 1e9,
                 'roofline_frac': pbytes * a.plain_samples / p['elapsed'] / HBM_PEAK}
        log('plain', plain)
        p['engine'].close()
        if fast['first'] and p['first']:
            parity = first_prefix_parity(fast['first'], p['first'], prefixes(1)[0][1])
            log('parity', parity)

    cpu = None
    if not a.no_cpu:
        log('cpu baseline ...')
        cpu = cpu_baseline()
        log('cpu', cpu)

    traffic = valu = None
    kernel = 'dcs_group_*' if a.schedule == 'compiled' else 'k_coset'
    tf = ROOT / 'profiles' / 'coset_groups_ibm64_k2_c64_counters.json'
    if a.schedule == 'compiled' and tf.exists():
        rec = json.loads(tf.read_text())
        # counter readings belong to one schedule: use them only for the same launches and bytes
        tiling = {k: fast['options'][k] for k in ('fuse_max_dim', 'tile_log2', 'pass_max_dim', 'group_policy',
                                                   'lds_buffers')}
        if (rec.get('launches_per_sample') == float(launches[2]) and rec.get('lanes_per_launch', 1) == fast['lanes']
                and rec.get('algorithmic_bytes_per_sample') == coset_bytes
                and bool(rec.get('f0_factored', False)) == bool(factored)
                and rec.get('tiling', tiling) == tiling):
            traffic = rec.get('hbm_bytes_per_sample')
            valu = rec.get('valu_instructions_per_sample')
    # Three ceilings of the group kernels, each over their own time per sample: HBM bytes (algorithmic),
    # FP64 flops (algorithmic), and VALU instruction issue (counted by SQ_INSTS_VALU in the committed
    # counter profile).  `bound` names the resource closest to its ceiling.
    sec = coset_ms / 1e3 if coset_ms > 0 else None
    hbm_frac = achieved / HBM_PEAK if achieved else None
    fp64_tflops = fused_flops / sec / 1e12 if sec else None
    fp64_frac = fp64_tflops * 1e12 / FP64_PEAK if sec else None
    valu_frac = valu / sec / VALU_ISSUE_PEAK if (valu and sec) else None
    fracs = {'hbm': hbm_frac, 'fp64': fp64_frac, 'valu_issue': valu_frac}
    bound = max((k for k in fracs if fracs[k] is not None), key=lambda k: fracs[k], default='hbm')
    roofline = {'bound': bound, 'achieved': achieved / 1e9 if achieved else None, 'peak': HBM_PEAK / 1e9,
                'unit': 'GB/s', 'frac': hbm_frac, 'traffic': traffic, 'traffic_unit': 'HBM bytes per sample',
                'launches_per_sample': float(launches[2]), 'lanes_per_launch': fast['lanes'],
                'kernel': kernel, 'bytes_per_sample': coset_bytes, 'ms_per_sample': coset_ms,
                'fp64': {'achieved': fp64_tflops, 'peak': FP64_PEAK / 1e12, 'unit': 'TFLOP/s', 'frac': fp64_frac,
                         'flops_per_sample': fused_flops},
                'valu_issue': {'instructions_per_sample': valu, 'peak_per_s': VALU_ISSUE_PEAK, 'frac': valu_frac}}
    lat = np.array(fast['lat']) * 1e3
    line = {
        'metric': 'samples/s @64q 2^20 block: batched c64 vs plain c128, HBM GB/s; p50/p99 ms',
        'value': sps, 'unit': 'samples/s', 'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup,
        'ms_per_sample': 1e3 * fast['elapsed'] / a.steps, 'ms_per_step': 1e3 * fast['elapsed'] / a.steps,
        'p50_ms': float(np.percentile(lat, 50)), 'p99_ms': float(np.percentile(lat, 99)),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'c64 storage / f64 arithmetic', 'data': 'frozen IBM QAT #245 graph, protocol-replayed prefixes',
        'config': {'workload': 'IBM QAT #245 64q/d73/314T injected graph (N=378), plan inplace_k2_fp32 '
                               '(4 parity slices, rmax 34, 128 GiB c64 frontier), 2^20 block, 44-bit prefixes',
                   'schedule': a.schedule, 'lds_tiles': ('fp64' if a.lds_fp64 else 'storage') if a.schedule == 'compiled' else None,
                   'fusion': {**{k: fast['options'][k] for k in ('fuse_max_dim', 'tile_log2', 'pass_max_dim',
                                                                  'group_policy', 'lds_buffers', 'graphs')},
                              'f0_factored': factored},
                   'batch': fast['lanes'],
                   'block': 1 << 20},
        'setup_s': fast['setup'], 'fast_samples_run': fast['samples_run'],
        'effective_hbm_GBps': unfused_bytes * sps / 1e9,
        'effective_roofline_frac': unfused_bytes * sps / HBM_PEAK,
        'algorithmic_bytes_per_sample': {'unfused': unfused_bytes, 'fused': sum(fb.values())},
        'stream_copy_GBps': stream_GBps,
        'plain_c128': plain,
        'speedup_vs_plain_c128': (sps / plain['samples_per_s']) if plain else None,
        'parity_first_prefix': parity,
        'roofline': roofline,
        'cpu_baseline': cpu,
        'device': {k: dev[k] for k in ('device_name', 'arch', 'compute_units', 'total_bytes')},
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
