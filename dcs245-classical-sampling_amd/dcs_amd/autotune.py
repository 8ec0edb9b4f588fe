"""Schedule autotuner: pick the coset-group tiling of the compiled schedule per plan by measurement.

The compiled schedule has four knobs that trade HBM bytes against register passes and occupancy:
fuse_max_dim (largest coset dimension of a group), tile_log2 (entries per LDS tile), group_policy
(greedy, fewest-bytes, bytes-and-passes summed or overlapped boundaries) and lds_buffers (one LDS tile per workgroup, two,
or two filled by asynchronous global-to-LDS loads).  Their best
setting depends on the plan's width profile, so for a given plan the tuner allocates an engine per
candidate, times its contraction of one prefix (after a warm-up contraction, with the schedule's own
device timers summed over the group, resident and per-step kernels), and keeps the fastest.  Results
are cached per plan fingerprint (step table, storage, device architecture), in memory and optionally
in a JSON file, so a serving process tunes a circuit once.
"""
from __future__ import annotations

import hashlib
import json
import time
from pathlib import Path

import numpy as np

DEFAULT_CANDIDATES = (
    {'fuse_max_dim': 9, 'tile_log2': 12, 'group_policy': 3, 'lds_buffers': 3},
    {'fuse_max_dim': 8, 'tile_log2': 12, 'group_policy': 3, 'lds_buffers': 3},
    {'fuse_max_dim': 8, 'tile_log2': 12, 'group_policy': 3, 'lds_buffers': 4},
    {'fuse_max_dim': 8, 'tile_log2': 12, 'group_policy': 3, 'lds_buffers': 5},
    {'fuse_max_dim': 8, 'tile_log2': 12, 'group_policy': 2, 'lds_buffers': 3},
    {'fuse_max_dim': 8, 'tile_log2': 12, 'group_policy': 1, 'lds_buffers': 3},
    {'fuse_max_dim': 8, 'tile_log2': 12, 'group_policy': 1, 'lds_buffers': 4},
    {'fuse_max_dim': 8, 'tile_log2': 12, 'group_policy': 1, 'lds_buffers': 1},
    {'fuse_max_dim': 9, 'tile_log2': 13, 'group_policy': 3, 'lds_buffers': 4},
    {'fuse_max_dim': 9, 'tile_log2': 13, 'group_policy': 2, 'lds_buffers': 3},
    {'fuse_max_dim': 9, 'tile_log2': 13, 'group_policy': 1, 'lds_buffers': 1},
    {'fuse_max_dim': 7, 'tile_log2': 12, 'group_policy': 1, 'lds_buffers': 3},
    {'fuse_max_dim': 6, 'tile_log2': 11, 'group_policy': 1, 'lds_buffers': 3},
    {'fuse_max_dim': 6, 'tile_log2': 11, 'group_policy': 0, 'lds_buffers': 2},
)

_memory_cache = {}


def plan_fingerprint(engine_cls, args, dtype, arch, alloc_kw=None):
    """SHA-256 of what fixes the schedule's behaviour: the in-place step table, the storage, the device
    architecture and the allocation options every candidate shares (lane count, LDS storage, register
    pass span, resident bits)."""
    eng = engine_cls(*args)
    t = eng.bp.step_table()
    h = hashlib.sha256()
    for k in sorted(t):
        h.update(k.encode())
        h.update(np.ascontiguousarray(t[k]).tobytes())
    h.update(f'{engine_cls.__name__}|{np.dtype(dtype).name}|{arch}|{eng.bp.inplace_maxslots}'.encode())
    h.update(json.dumps(sorted((alloc_kw or {}).items()), default=str).encode())
    return h.hexdigest()


def _time_engine(engine, x, repeats):
    engine.block(x)                      # warm-up: first launches of the compiled module
    best = float('inf')
    for _ in range(repeats):
        engine._plan.profile(True)
        t0 = time.perf_counter()
        engine.block(x)
        wall = time.perf_counter() - t0
        ms, _ = engine._plan.profile(False)
        device = float(ms[0] + ms[1] + ms[2]) / 1e3
        best = min(best, device if device > 0 else wall)
    return best


def autotune(engine_cls, args, dtype=np.complex64, x=None, candidates=DEFAULT_CANDIDATES, repeats=2,
             cache_path=None, device=0, arch=None, alloc_kw=None, log=None):
    """Fastest candidate for the plan of engine_cls(*args).  Returns dict(best, table, fingerprint,
    cached).  x: prefix bits to time (default all zeros); alloc_kw: options shared by every candidate
    (e.g. lanes, lds_storage); arch: device architecture string for the fingerprint (default: read from
    the device)."""
    if arch is None:
        from . import _native
        arch = _native.device_info(device)['arch']
    fp = plan_fingerprint(engine_cls, args, dtype, arch, alloc_kw)
    disk = {}
    if cache_path is not None and Path(cache_path).exists():
        disk = json.loads(Path(cache_path).read_text())
    hit = _memory_cache.get(fp) or disk.get(fp)
    if hit is not None:
        _memory_cache[fp] = hit
        return {**hit, 'fingerprint': fp, 'cached': True}
    n = args[2]
    x = np.zeros(n, np.uint8) if x is None else np.asarray(x, np.uint8)
    table = []
    for cand in candidates:
        eng = engine_cls(*args)
        try:
            eng.alloc(dtype=dtype, device=device, schedule='compiled', **{**(alloc_kw or {}), **cand})
            seconds = _time_engine(eng, x, repeats)
        except (MemoryError, ValueError) as exc:
            # a candidate the plan or device refuses: no memory, an invalid tiling, or a schedule that does
            # not compile or cannot be resident (ScheduleUnsupported); a RuntimeError is a device failure
            # and ends the tuning
            table.append({'options': dict(cand), 'seconds': None, 'error': f'{type(exc).__name__}: {exc}'})
            continue
        finally:
            eng.close()
        table.append({'options': dict(cand), 'seconds': seconds})
        if log:
            log(json.dumps(table[-1]))
    timed = [row for row in table if row['seconds'] is not None]
    if not timed:
        raise RuntimeError('no candidate schedule could be allocated')
    best = min(timed, key=lambda row: row['seconds'])
    result = {'best': best['options'], 'best_seconds': best['seconds'], 'table': table}
    _memory_cache[fp] = result
    if cache_path is not None:
        disk[fp] = result
        Path(cache_path).write_text(json.dumps(disk, indent=1))
    return {**result, 'fingerprint': fp, 'cached': False}


def main(argv=None):
    """Tune the 64-qubit generation plan (or another plan of the frozen workload) and print the table."""
    import argparse
    from .conditional_block import InplaceConditionalBlockGPU
    from .workloads import load_ibm64
    ap = argparse.ArgumentParser(description='Autotune the compiled schedule of a frozen 64-qubit plan')
    ap.add_argument('--plan', default='inplace_k2_fp32')
    ap.add_argument('--c128', action='store_true')
    ap.add_argument('--cache', default=None)
    a = ap.parse_args(argv)
    w = load_ibm64()
    args = (w['graph'], w['unitaries'], 64, w['order'], w['plans'][a.plan]['operations'], w['open_vertices'])
    x = np.array([int(c) for c in w['protocol']['anchor_input_data_bits']], np.uint8)
    res = autotune(InplaceConditionalBlockGPU, args, dtype=np.complex128 if a.c128 else np.complex64, x=x,
                   cache_path=a.cache, log=lambda s: print(s, flush=True))
    print(json.dumps({'plan': a.plan, 'best': res['best'], 'best_seconds': res['best_seconds']}), flush=True)


if __name__ == '__main__':
    main()
