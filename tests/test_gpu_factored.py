"""f0-factored group kernels on the device, through the C ABI.

The compiled complex64 schedule runs each step whose f0 cannot be zero with (1, f1 / f0) and applies the
pending f0s at each group's last step (BlockPlan.coefficient_classes).  Checks: the factored blocks stay
within 1e-5 of the complex128 oracle and no further from it than twice the unfactored kernels' error;
the plan reports that the factored kernels ran; and a contraction whose parameters do not fit the
classes it was compiled for (classes forced wrong) runs the unfactored kernels of the same schedule,
bit for bit, with and without graph replay."""
import numpy as np
import pytest

from dcs_amd.block_plan import BlockPlan
from dcs_amd.conditional_block import InplaceConditionalBlockGPU
from dcs_amd.stabilizer import synthetic_instance
from oracle import oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


@pytest.fixture(scope='module')
def cases():
    out = []
    inst = synthetic_instance(16, 12, 20, seed=0)
    args = (inst['graph'], inst['unitaries'], 16, inst['natural_order'], [], list(range(16)))
    out.append(('c1', args, [np.zeros(16, np.uint8)]))
    inst = synthetic_instance(28, 32, 64, seed=0)
    args = (inst['graph'], inst['unitaries'], 28, inst['natural_order'], [], list(range(20, 28)))
    rng = np.random.default_rng(3)
    out.append(('w20', args, [rng.integers(0, 2, 28).astype(np.uint8) for _ in range(3)]))
    return out


def engine(args, **kw):
    e = InplaceConditionalBlockGPU(*args)
    e.alloc(dtype=np.complex64, schedule='compiled', resident_bits=6, **kw)
    return e


def test_factored_blocks_against_oracle(cases):
    for name, args, xs in cases:
        bp = BlockPlan(*args)
        cls = bp.coefficient_classes()
        assert (cls > 0).sum() > 0.5 * len(cls)      # every T reference vertex factors
        fac, plain = engine(args), engine(args, factor=False)
        for i, x in enumerate(xs):
            print(f'case {name} prefix {i}', flush=True)
            ref = oracle.block(bp, x)
            vf, _ = fac.block(x)
            assert fac.factoring() == (True, True)
            vp, _ = plain.block(x)
            assert plain.factoring() == (False, False)
            ef, ep = rel(vf, ref), rel(vp, ref)
            print(f'  error vs FP64: factored {ef:.3e} unfactored {ep:.3e}', flush=True)
            assert ef < 1e-5 and ef <= 2 * ep + 1e-12, (ef, ep)
        fac.close()
        plain.close()


@pytest.mark.parametrize('graphs', [False, True])
def test_misfit_classes_run_unfactored(cases, graphs):
    """Classes forced wrong (real and imaginary swapped): every contraction falls back to the unfactored
    kernels of the same schedule, whose blocks equal those of an engine whose classes are all 0 (the
    same group boundaries, every step with (f0, f1), the last with (1 f0, 1 f1)) bit for bit."""
    name, args, xs = cases[1]
    bp = BlockPlan(*args)
    cls = bp.coefficient_classes()
    wrong = np.where(cls == 2, 1, np.where(cls == 1, 2, cls)).astype(np.int32)   # real <-> imaginary
    print(f'case {name} misfit classes graphs={graphs}', flush=True)
    mis = engine(args, factor=wrong, graphs=graphs)
    zero = engine(args, factor=np.zeros_like(cls), graphs=graphs)
    good = engine(args, graphs=graphs)
    assert mis.schedule_options == zero.schedule_options == good.schedule_options
    for x in xs[:2]:
        vm, _ = mis.block(x)
        assert mis.factoring() == (True, False)
        vz, _ = zero.block(x)
        assert zero.factoring() == (True, True)
        assert np.array_equal(np.asarray(vm), np.asarray(vz))
        vg, _ = good.block(x)
        assert good.factoring() == (True, True)
        assert rel(vg, vz) < 1e-5
    for e in (mis, zero, good):
        e.close()


@pytest.mark.parametrize('factor', [False, True])
def test_ring_pipeline_equals_direct_store(factor):
    """The three asynchronous-load pipelines run the same arithmetic: the ring of compact input buffers
    (lds_buffers=5: the live inputs of the next two tiles in flight, waits counted per wave), the
    direct-store pipeline (4: the last pass stores one entry per lane from registers) and the default
    (3: finished tiles stored from LDS two entries, 16 bytes, per lane).  On the 32q/d40/120T ring plan
    (width 29, 4 GiB complex64 frontier, six of its groups take the ring) every block is the same bit
    for bit, per-step rounding and f0-factored, over four protocol prefixes two lanes at a time."""
    from dcs_amd.sampling_protocol import make_protocol
    from dcs_amd.width_planner import best_rotation
    inst = synthetic_instance(32, 40, 120, seed=0)
    order, opn, width = best_rotation(inst['graph'], inst['unitaries'], 32, inst['natural_order'], 16, stride=2)
    args = (inst['graph'], inst['unitaries'], 32, order, [], opn)
    xs = [np.array([int(c) for c in r['input_data_bits']], np.uint8) for r in make_protocol(2, 4, 32, opn, 1.0)['records']]
    kw = dict(dtype=np.complex64, schedule='compiled', lanes=2, blocks=2, group_policy=1, exact_rounding=not factor,
              factor=factor)
    out = {}
    for buffers in (3, 4, 5):
        print(f'case 32q ring plan lds_buffers={buffers} factor={factor}', flush=True)
        e = InplaceConditionalBlockGPU(*args)
        e.alloc(lds_buffers=buffers, **kw)
        out[buffers] = []
        for i in (0, 2):
            vecs, _ = e.block_many(xs[i:i + 2])
            out[buffers] += [np.asarray(v) for v in vecs]
        e.close()
    for a, b, c in zip(out[3], out[4], out[5]):
        assert np.array_equal(a, b) and np.array_equal(b, c)
