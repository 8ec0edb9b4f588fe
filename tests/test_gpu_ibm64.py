"""The frozen 64-qubit QAT #245 graph on the device against the reference's run records.

* complex128, plan inplace_k3_fp64 (8 slices, rmax 33, 128 GiB frontier): the amplitude at the drawn
  index agrees with fullscore/analysis/scores.jsonl to 1e-10 relative (the reference's own c128
  reruns agree to 2.2e-16; the difference here is the FMA placement of the CUDA build), and the block
  mass with the audit record to 1e-10 relative.
* complex64, plan inplace_k2_fp32 (4 slices, rmax 34): the draw index equals the recorded r1 draw
  unless the recorded CDF margin is below 1e-7 (a reported tie); block masses agree with the r1 masses
  to 1e-5; amplitudes at the drawn index agree with the c128 scores to 1e-5 relative up to one global
  phase per plan (the k2 and k3 parity operations differ by a pivot, whose local-complementation
  identity holds up to a phase; the reference's audits compare |overlap| for the same reason)."""
import json

import numpy as np
import pytest

from dcs_amd.conditional_block import InplaceConditionalBlockGPU

pytestmark = pytest.mark.gpu

TIE = 1e-7


def make(ibm64, plan, dtype, **kw):
    e = InplaceConditionalBlockGPU(ibm64['graph'], ibm64['unitaries'], 64, ibm64['order'],
                                   ibm64['plans'][plan]['operations'], ibm64['open_vertices'])
    e.alloc(dtype=dtype, **kw)
    assert e.inplace_rmax == ibm64['plans'][plan]['rmax']
    return e


def bits(s):
    return np.array([int(c) for c in s], np.uint8)


@pytest.mark.parametrize('sched', ['fused', 'compiled'])
def test_ibm64_c128_k3_matches_scores(ibm64, records, sched):
    print(f'case ibm64 c128 k3 {sched}', flush=True)
    e = make(ibm64, 'inplace_k3_fp64', np.complex128, schedule=sched)
    audited = sorted(int(k) for k in records['audits'])[:1]
    for sid in [0] + audited:
        s = records['samples'][str(sid)]
        vec, meta = e.block(bits(s['bits']))
        v = np.asarray(vec)
        A, re, im, _ = records['scores'][str(sid)]
        want = complex(re, im)
        print(f'sample {sid}: {meta["seconds"]:.2f} s, amp {v[A]} vs {want}', flush=True)
        assert abs(v[A] - want) <= 1e-10 * abs(want)
        assert meta['slices_evaluated'] == 8 and meta['rmax'] == 33 and meta['work_per_slice'] == 315122088218
        if str(sid) in records['audits']:
            mass = float(np.vdot(v, v).real)
            assert abs(mass - records['audits'][str(sid)]['ref_mass']) <= 1e-10
    e.close()


@pytest.mark.parametrize('sched,lds', [('fused', False), ('compiled', False), ('compiled', True)])
def test_ibm64_c64_k2_draws_match_r1(ibm64, records, sched, lds):
    print(f'case ibm64 c64 k2 {sched} lds_storage={lds}', flush=True)
    kw = {'lds_storage': True} if lds else {}
    e = make(ibm64, 'inplace_k2_fp32', np.complex64, schedule=sched, **kw)
    ties = 0
    phase = None
    for sid in range(3):
        s = records['samples'][str(sid)]
        vec, meta = e.block(bits(s['bits']))
        d = vec.draw(s['u'])
        v = np.asarray(vec)
        A, re, im, _ = records['scores'][str(sid)]
        print(f'sample {sid}: {meta["seconds"]:.2f} s, index {d["A_index"]} vs {s["A_index"]}, '
              f'mass {d["scaled_mass_cdf"]} vs {s["mass"]}', flush=True)
        if s['margin'] < TIE:
            ties += 1
        else:
            assert d['A_index'] == s['A_index']
        want = complex(re, im)
        ratio = v[A] / want
        assert abs(abs(ratio) - 1) <= 1e-5
        phase = ratio if phase is None else phase
        assert abs(ratio - phase) <= 1e-5
        assert abs(d['scaled_mass_cdf'] - s['mass']) <= 1e-5
    e.close()


def test_ibm64_driver_worker_with_audit_matches_r1(ibm64, records, tmp_path):
    """The production driver on the first preselected audit prefix: complex64 generation (k2, compiled
    schedule), draw and records, complex128 reference (k3), the 53-bit CDF law witness; the record
    verifier passes on the directory, and draw index, masses, witness and loss equal the r1 records."""
    from dcs_amd.driver import run_worker, verify_worker_dir
    sid = min(int(k) for k in records['audits'])
    s, a = records['samples'][str(sid)], records['audits'][str(sid)]
    worker = sid % ibm64['protocol']['worker_count']
    print(f'case ibm64 driver worker {worker} sample {sid} with audit', flush=True)
    out = tmp_path / f'worker_{worker}'
    summary = run_worker(out, worker, sample_ids=[sid], workload=ibm64)
    assert summary['audit_ids'] == [sid]
    res = verify_worker_dir(out, workload=ibm64)
    print(json.dumps({k: res[k] for k in ('passed', 'checks')}), res['failures'][:3], flush=True)
    assert res['passed'], res['failures'][:5]
    row = json.loads((out / 'samples.jsonl').read_text().splitlines()[0])
    assert row['input_data_bits'] == s['bits'] and row['draw_uniform'] == s['u']
    if s['margin'] >= TIE:
        assert row['draw']['A_index'] == s['A_index']
    assert abs(row['draw']['scaled_mass_fsum'] - s['mass']) <= 1e-5
    audit = json.loads((out / 'audits.jsonl').read_text().splitlines()[0])
    w = audit['witness']
    assert abs(w['reference_scaled_mass'] - a['ref_mass']) <= 1e-10
    assert abs(w['approximate_input_mass'] - 1.0) <= 1e-12
    assert abs(w['uncapped_overlap_witness'] - a['witness']) <= 2e-6
    assert abs(w['loss'] - a['loss']) <= 2e-6
    hw = json.loads((out / 'hardware.json').read_text())
    assert hw['arch'].startswith('gfx950')


def test_ibm64_half_k1_whole_vector_against_c128(ibm64, records):
    """Binary16 storage on the frozen half_k1 plan (2 parity slices, rmax 35: a 2^35-entry, 128 GiB
    frontier of binary16 pairs), compiled schedule, first protocol prefix: the whole 2^20 block against
    the complex128 k3 block (the FP64 validation the reference's half engine asks for,
    conditional_block_half_gpu.py:3-4 and 31-32): normalised overlap, block mass and the largest
    deviation after aligning the global phase."""
    from dcs_amd.conditional_block import HalfConditionalBlockGPU
    s = records['samples']['0']
    print('case ibm64 half_k1 compiled sample 0', flush=True)
    e = HalfConditionalBlockGPU(ibm64['graph'], ibm64['unitaries'], 64, ibm64['order'],
                                ibm64['plans']['half_k1']['operations'], ibm64['open_vertices'])
    e.alloc()
    assert e.inplace_rmax == ibm64['plans']['half_k1']['rmax'] == 35
    vec, meta = e.block(bits(s['bits']))
    half = np.asarray(vec)
    e.close()
    assert meta['slices_evaluated'] == 2 and meta['coherent_accumulation'] == 'complex128'
    e = make(ibm64, 'inplace_k3_fp64', np.complex128)
    ref = np.asarray(e.block(bits(s['bits']))[0])
    e.close()
    overlap = abs(np.vdot(ref, half)) / np.sqrt(np.vdot(ref, ref).real * np.vdot(half, half).real)
    A = records['scores']['0'][0]
    mass_h, mass_r = float(np.vdot(half, half).real), float(np.vdot(ref, ref).real)
    print(f'half: {meta["seconds"]:.2f} s, overlap {overlap:.8f}, mass {mass_h:.6f} vs {mass_r:.6f}, '
          f'|amp| {abs(half[A]):.6e} vs {abs(ref[A]):.6e}, max rel err {rel_err(half, ref):.3e}', flush=True)
    # observed: overlap 0.99999813, masses 1.000012 / 0.999888, largest deviation 1.8e-3 of the block max
    assert overlap > 0.9999
    assert abs(mass_h - mass_r) < 1e-3
    assert rel_err(half, ref) < 1e-2


def rel_err(a, b):
    return float(np.max(np.abs(a - b * (np.vdot(b, a) / np.vdot(b, b)))) / np.max(np.abs(b)))
