"""Autotuner search logic with a stand-in engine (no device): the fastest candidate wins, refused
candidates are recorded and skipped, results are cached per plan fingerprint in memory and on disk, and
the fingerprint changes with the plan, the storage and the architecture."""
import json

import numpy as np
import pytest

from dcs_amd import autotune as at
from dcs_amd.block_plan import BlockPlan
from dcs_amd.stabilizer import synthetic_instance


class FakePlan:
    def __init__(self, cost):
        self.cost, self.on = cost, False

    def profile(self, enable):
        self.on = enable
        return np.array([0.0, 0.0, self.cost * 1e3, 0.0, 0.0]), np.zeros(5, np.int64)


class FakeEngine:
    allocs = []

    def __init__(self, *args):
        self.bp = BlockPlan(*args)

    def alloc(self, dtype, device, schedule, **kw):
        FakeEngine.allocs.append(kw)
        if kw.get('tile_log2') == 13:
            raise MemoryError('does not fit')
        # pretend the cost is smallest at dimension 7 on 2^12 tiles
        self._plan = FakePlan(abs(kw['fuse_max_dim'] - 7) + abs(kw['tile_log2'] - 12) + 0.1 * kw['lds_buffers'])

    def block(self, x):
        return None, {}

    def close(self):
        pass


@pytest.fixture()
def args():
    inst = synthetic_instance(12, 8, 10, seed=4)
    return (inst['graph'], inst['unitaries'], 12, inst['natural_order'], [], [8, 9, 10, 11])


def test_autotune_picks_fastest_and_caches(args, tmp_path):
    at._memory_cache.clear()
    FakeEngine.allocs.clear()
    cache = tmp_path / 'tune.json'
    res = at.autotune(FakeEngine, args, arch='gfx950', cache_path=cache)
    assert res['best'] == {'fuse_max_dim': 7, 'tile_log2': 12, 'group_policy': 1, 'lds_buffers': 3}
    assert not res['cached'] and len(res['table']) == len(at.DEFAULT_CANDIDATES)
    refused = [r for r in res['table'] if r['seconds'] is None]
    assert len(refused) == sum(c['tile_log2'] == 13 for c in at.DEFAULT_CANDIDATES)
    assert all('does not fit' in r['error'] for r in refused)
    n_alloc = len(FakeEngine.allocs)
    again = at.autotune(FakeEngine, args, arch='gfx950', cache_path=cache)
    assert again['cached'] and again['best'] == res['best'] and len(FakeEngine.allocs) == n_alloc
    at._memory_cache.clear()
    from_disk = at.autotune(FakeEngine, args, arch='gfx950', cache_path=cache)
    assert from_disk['cached'] and from_disk['best'] == res['best']
    assert res['fingerprint'] in json.loads(cache.read_text())


def test_fingerprint_depends_on_plan_storage_and_arch(args):
    base = at.plan_fingerprint(FakeEngine, args, np.complex64, 'gfx950')
    assert base == at.plan_fingerprint(FakeEngine, args, np.complex64, 'gfx950')
    assert base != at.plan_fingerprint(FakeEngine, args, np.complex128, 'gfx950')
    assert base != at.plan_fingerprint(FakeEngine, args, np.complex64, 'gfx942')
    other = args[:5] + ([7, 9, 10, 11],)
    assert base != at.plan_fingerprint(FakeEngine, other, np.complex64, 'gfx950')


class FaultingEngine(FakeEngine):
    def alloc(self, dtype, device, schedule, **kw):
        if kw.get('group_policy') == 2:
            raise RuntimeError('hipModuleLaunchKernel failed: an illegal memory access was encountered')
        if kw.get('fuse_max_dim') == 9:
            from dcs_amd._native import ScheduleUnsupported
            raise ScheduleUnsupported('group kernel 3 cannot be resident (registers or LDS)')
        super().alloc(dtype, device, schedule, **kw)


def test_device_failure_ends_tuning_and_refusals_are_recorded(args):
    """A RuntimeError (a device failure) propagates instead of being logged as a refused candidate; a
    schedule the device cannot build (ScheduleUnsupported) is recorded and skipped."""
    at._memory_cache.clear()
    with pytest.raises(RuntimeError, match='illegal memory access'):
        at.autotune(FaultingEngine, args, arch='gfx950')
    at._memory_cache.clear()
    cands = [c for c in at.DEFAULT_CANDIDATES if c['group_policy'] != 2]
    res = at.autotune(FaultingEngine, args, arch='gfx950', candidates=cands)
    refused = [r for r in res['table'] if r['seconds'] is None]
    assert refused and all(r['error'].startswith(('ScheduleUnsupported', 'MemoryError')) for r in refused)


def test_fingerprint_depends_on_shared_allocation_options(args):
    one = at.plan_fingerprint(FakeEngine, args, np.complex64, 'gfx950', {'lanes': 1, 'blocks': 1})
    two = at.plan_fingerprint(FakeEngine, args, np.complex64, 'gfx950', {'lanes': 2, 'blocks': 1})
    assert one != two
    assert one == at.plan_fingerprint(FakeEngine, args, np.complex64, 'gfx950', {'blocks': 1, 'lanes': 1})
    assert one != at.plan_fingerprint(FakeEngine, args, np.complex64, 'gfx950', {'lanes': 1, 'blocks': 1,
                                                                                 'lds_storage': False})
