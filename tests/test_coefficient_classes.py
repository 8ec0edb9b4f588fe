"""Coefficient classes of the f0-factored group kernels (BlockPlan.coefficient_classes), on the CPU.

A step of class 1 or 2 must have f0 != 0 and a ratio f1 / f0 that is real (1) or imaginary (2), to the
1e-12 relative tolerance the library checks, for every prefix and every slice; a step of class 3 must
have f0 != 0.  Checked here against the actual step parameters (BlockPlan.step_parameters of every
slice assignment) of the 64-qubit workload's generation plan on protocol prefixes and of seeded ring
circuits on random prefixes, where data vertices, reference vertices and parity slices all occur.
The derivation must also not give up on the structure: every reference (T) vertex step factors.
"""
import numpy as np
import pytest

from dcs_amd.block_plan import BlockPlan
from dcs_amd.sampling_protocol import make_protocol
from dcs_amd.stabilizer import synthetic_instance
from dcs_amd.workloads import load_ibm64


def _check(bp, prefixes):
    cls = bp.coefficient_classes()
    assert cls.shape == (len(bp.inplace_steps),) and set(np.unique(cls)) <= {0, 1, 2, 3}
    inverse = {bp.relabel[v]: v for v in bp.keep}
    ref_steps = [i for i, s in enumerate(bp.inplace_steps) if inverse.get(int(s.v), -1) >= bp.n]
    assert all(cls[i] > 0 for i in ref_steps)
    checked = 0
    for x in prefixes:
        for _, _, f, yfix, _ in bp.slice_assignments(bp.local_vectors(x)):
            if f is None:
                continue
            c, _, _, _ = bp.step_parameters(f, yfix)
            f0 = c[:, 0] + 1j * c[:, 1]
            f1 = c[:, 2] + 1j * c[:, 3]
            fac = cls > 0
            assert np.all(f0[fac] != 0)
            r = f1[fac] / f0[fac]
            k = cls[fac]
            assert np.all(np.abs(r.imag[k == 1]) <= 1e-12 * np.abs(r[k == 1]))
            assert np.all(np.abs(r.real[k == 2]) <= 1e-12 * np.abs(r[k == 2]))
            checked += 1
    assert checked > 0
    return cls


def test_classes_of_the_64_qubit_generation_plan():
    w = load_ibm64()
    bp = BlockPlan(w['graph'], w['unitaries'], 64, w['order'], w['plans']['inplace_k2_fp32']['operations'],
                   w['open_vertices'])
    prot = make_protocol(w['protocol']['generation']['seed'], 24, 64, w['open_vertices'], 1.0)
    cls = _check(bp, [np.array([int(c) for c in r['input_data_bits']], np.uint8) for r in prot['records']])
    # 353 of 356 steps factor; the other three are data vertices whose local Clifford has a zero entry
    assert (cls > 0).sum() == 353 and set(np.unique(cls[cls > 0])) == {1, 2}


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_classes_of_ring_circuits(seed):
    inst = synthetic_instance(20, 16, 30, seed=seed)
    bp = BlockPlan(inst['graph'], inst['unitaries'], 20, inst['natural_order'], [], list(range(12, 20)))
    rng = np.random.default_rng(seed)
    _check(bp, [rng.integers(0, 2, 20).astype(np.uint8) for _ in range(8)])
