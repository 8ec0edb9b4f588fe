"""Production driver and record verifier on the CPU: a worker run over a small ring circuit whose engine
is the oracle (complex64 storage for generation, complex128 for audits), then the verifier on its
directory, and the verifier catching planted record changes (a replayed uniform, a vector file, an
output bit).  The worker files follow production/worker.py:106-195; the checks verify.py:300-328,
426-537."""
import json
import time

import numpy as np
import pytest

from dcs_amd.block_plan import BlockPlan
from dcs_amd.driver import run_worker, verify_worker_dir
from dcs_amd.stabilizer import synthetic_instance
from oracle import oracle


class OracleEngine:
    """Engine interface of the worker (block, inplace_rmax, close) computed by the oracle."""

    def __init__(self, bp, storage):
        self.bp, self.storage = bp, storage
        self.inplace_rmax = bp.inplace_maxslots

    def block(self, x):
        t = time.perf_counter()
        v = oracle.block(self.bp, x, storage=self.storage)
        return v, {'seconds': time.perf_counter() - t, 'prefix_dtype': 'complex128' if self.storage else 'complex64',
                   'open_vertices': self.bp.open_full, 'rmax': self.inplace_rmax}

    def close(self):
        pass


@pytest.fixture(scope='module')
def small():
    n = 12
    inst = synthetic_instance(n, 10, 14, seed=3)
    order = list(inst['natural_order'])
    data = [v for v in order if v < n]
    free = data[-5:]
    bp = BlockPlan(inst['graph'], inst['unitaries'], n, order, [], free)
    plans = {'gen': {'operations': [], 'rmax': bp.inplace_maxslots},
             'ref': {'operations': [], 'rmax': bp.inplace_maxslots}}
    protocol = {'schema_version': 'test', 'engine': 'fp32', 'worker_count': 2,
                'generation': {'seed': 1234, 'count': 9, 'n': n, 'free_vertices': free, 'mixture_weight': 1.0},
                'generation_plan': 'gen', 'reference_plan': 'ref',
                'audit': {'ids_by_worker': {'0': [0, 4, 6], '1': [3]}},
                'anchor_input_data_bits': '0' * n}
    workload = {'graph': inst['graph'], 'unitaries': inst['unitaries'], 'order': order, 'ring': list(range(n)),
                'open_vertices': free, 'plans': plans, 'protocol': protocol}

    def make_engine(plan, dtype):
        return OracleEngine(bp, 0 if dtype == np.complex64 else 1)
    return workload, make_engine


def run(tmp_path, small, **kw):
    workload, make_engine = small
    out = tmp_path / 'w0'
    summary = run_worker(out, 0, workload=workload, make_engine=make_engine, hardware=False,
                         log=lambda *a, **k: None, **kw)
    return out, summary


def test_worker_writes_reference_files_and_verifies(tmp_path, small):
    out, summary = run(tmp_path, small)
    assert summary['sample_ids'] == [0, 2, 4, 6, 8] and summary['audit_ids'] == [0, 4, 6]
    for name in ('source.json', 'anchor.json', 'samples.jsonl', 'audits.jsonl', 'summary.json', 'artifacts.json',
                 'completed.json', 'approx_000004.npy', 'reference_000006.npy'):
        assert (out / name).exists(), name
    assert not (out / 'failure.json').exists()
    rows = [json.loads(l) for l in (out / 'samples.jsonl').read_text().splitlines()]
    assert [r['is_preselected_audit'] for r in rows] == [True, False, True, True, False]
    res = verify_worker_dir(out, workload=small[0])
    assert res['passed'], res['failures'][:5]
    assert res['checks'] > 100
    cert = res['certificate']
    assert cert['audit_count'] == 3 and cert['audit_ids'] == [0, 4, 6]
    # complex64 storage against complex128: witness losses are tiny but the law grid makes them nonzero
    assert 0 <= cert['observed_mean_loss'] < 1e-5
    audits = [json.loads(l) for l in (out / 'audits.jsonl').read_text().splitlines()]
    for a in audits:
        assert a['raw_vector_witness']['normalized_conditional_overlap_abs'] > 1 - 1e-6


def test_worker_subset_and_no_audit(tmp_path, small):
    out, summary = run(tmp_path, small, sample_ids=[2, 4], audit=False)
    assert summary['sample_ids'] == [2, 4] and summary['audit_ids'] == [] and summary['audit_count'] == 0
    assert verify_worker_dir(out, workload=small[0])['passed']
    with pytest.raises(ValueError, match='not all assigned'):
        run_worker(tmp_path / 'bad', 0, sample_ids=[3], workload=small[0], make_engine=small[1], hardware=False)


def test_worker_anchor_mismatch_writes_failure(tmp_path, small):
    with pytest.raises(ArithmeticError, match='anchor'):
        run(tmp_path, small, anchor_expected_sha256='0' * 64)
    fail = json.loads((tmp_path / 'w0' / 'failure.json').read_text())
    assert fail['phase'] == 'fixed_anchor' and fail['completed'] is False
    anchor = json.loads((tmp_path / 'w0' / 'anchor.json').read_text())
    assert anchor['passed'] is False


def _rewrite_rows(path, fn):
    rows = [json.loads(l) for l in path.read_text().splitlines()]
    fn(rows)
    path.write_text(''.join(json.dumps(r) + '\n' for r in rows))


def test_verifier_catches_changed_uniform(tmp_path, small):
    out, _ = run(tmp_path, small)

    def change(rows):
        rows[1]['draw_uniform'] = 0.5
    _rewrite_rows(out / 'samples.jsonl', change)
    res = verify_worker_dir(out, workload=small[0])
    assert not res['passed']
    assert {'sample_rng_draw_uniform', 'draw_uniform'} <= {f['code'] for f in res['failures']}


def test_verifier_catches_changed_output_bit(tmp_path, small):
    out, _ = run(tmp_path, small)

    def change(rows):
        s = rows[0]['bitstring_qiskit']
        rows[0]['bitstring_qiskit'] = s[:-1] + ('1' if s[-1] == '0' else '0')
    _rewrite_rows(out / 'samples.jsonl', change)
    codes = {f['code'] for f in verify_worker_dir(out, workload=small[0])['failures']}
    assert codes == {'output_bit_mapping', 'summary_samples_sha256', 'artifact_sha256'}


def test_verifier_catches_changed_reference_vector(tmp_path, small):
    out, _ = run(tmp_path, small)
    path = out / 'reference_000004.npy'
    v = np.load(path)
    v[3] *= 1.001
    path.unlink()
    np.save(path, v)
    codes = {f['code'] for f in verify_worker_dir(out, workload=small[0])['failures']}
    assert 'vector_payload_hash' in codes and 'witness_reference_scaled_mass' in codes


def test_run_verification_over_every_worker(tmp_path, small):
    """Two workers of a protocol whose audit ids come from an audit seed: the run-level checks pass over
    both directories (selection replayed, every id once, one anchor hash, homogeneous hardware) and the
    certificate pools every audit; a missing worker or a changed selection fails them."""
    import copy
    import hashlib
    from dcs_amd.driver import replay_audit_selection, verify_run
    workload, make_engine = small
    w = copy.deepcopy(workload)
    prot = w['protocol']
    prot['audit'] = {'seed': 99, 'per_worker': 2, 'alpha': 0.05}
    prot['audit']['ids_by_worker'] = replay_audit_selection(prot['audit'], 9, 2)
    prot['audit']['ids'] = sorted(i for v in prot['audit']['ids_by_worker'].values() for i in v)
    x = np.zeros(12, np.uint8)
    anchor = hashlib.sha256(make_engine('gen', np.complex64).block(x)[0].astype('<c16').tobytes()).hexdigest()
    dirs, results = [], {}
    for wid in (0, 1):
        out = tmp_path / f'worker_{wid}'
        run_worker(out, wid, workload=w, make_engine=make_engine, hardware=False, anchor_expected_sha256=anchor,
                   log=lambda *a, **k: None)
        (out / 'hardware.json').write_text(json.dumps({'device_name': 'cpu', 'arch': 'x', 'numpy': '2',
                                                       'hip_runtime_version': 1, 'hip_driver_version': 1,
                                                       'library_sha256': 'abc'}))
        dirs.append(out)
        results[wid] = verify_worker_dir(out, workload=w)
    res = verify_run(dirs, results, workload=w)
    assert res['passed'], res['failures'][:5]
    assert res['sample_count'] == 9 and res['audit_count'] == 4 and res['quality']['audit_count'] == 4
    assert res['anchor_sha256'] == anchor
    bad = verify_run(dirs[:1], {0: results[0]}, workload=w)
    assert {'missing_worker_directories', 'all_sample_ids_exactly_once'} <= {f['code'] for f in bad['failures']}
    w2 = copy.deepcopy(w)
    w2['protocol']['audit']['seed'] = 100
    assert 'audit_selection_replay' in {f['code'] for f in verify_run(dirs, results, workload=w2)['failures']}


def test_compare_draws_counts_ties(tmp_path, small):
    from dcs_amd.driver import compare_draws
    out, _ = run(tmp_path, small)
    rows = [json.loads(l) for l in (out / 'samples.jsonl').read_text().splitlines()]
    ref = {str(r['sample_id']): {'A_index': r['draw']['A_index'], 'margin': 1.0} for r in rows}
    assert compare_draws([out], ref) == {'compared': 5, 'identical': 5, 'ties': 0, 'differ': 0, 'differing_ids': [],
                                         'tie_margin': 1e-7}
    ref['0'] = {'A_index': rows[0]['draw']['A_index'] ^ 1, 'margin': 1e-9}
    ref['2'] = {'A_index': rows[1]['draw']['A_index'] ^ 1, 'margin': 0.5}
    res = compare_draws([out], ref)
    assert (res['identical'], res['ties'], res['differ'], res['differing_ids']) == (3, 1, 1, [2])
