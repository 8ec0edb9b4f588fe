"""dcs_draw_noalloc (the draw that allocates nothing) against dcs_draw on the same device blocks.

Two calls in a row through the C ABI, the second on a smaller block (2^16 entries, then 2^12), each
checked field for field against dcs_draw of the block's own engine at several uniforms, including
the first and last CDF intervals.  The static scratch keeps the first call's sums; the second call
must not read them.  Sizes beyond the scratch and bad uniforms are refused before any launch."""
import numpy as np
import pytest

from dcs_amd import _native
from dcs_amd.conditional_block import InplaceConditionalBlockGPU
from dcs_amd.stabilizer import synthetic_instance

pytestmark = pytest.mark.gpu


def outcome(call):
    try:
        return ('ok', call())
    except ArithmeticError as exc:
        return ('refused', str(exc))


def engine(n, depth, t, seed):
    inst = synthetic_instance(n, depth, t, seed=seed)
    e = InplaceConditionalBlockGPU(inst['graph'], inst['unitaries'], n, inst['natural_order'], [], list(range(n)))
    e.alloc(dtype=np.complex128, schedule='steps', resident_bits=6)
    return e


def test_draw_noalloc_equals_dcs_draw_twice_smaller_second():
    ctx = _native.Context(0)
    try:
        for n, depth, t, seed in ((16, 12, 20, 0), (12, 8, 10, 3)):
            print(f'case draw_noalloc {n}q block 2^{n}', flush=True)
            e = engine(n, depth, t, seed)
            vec, _ = e.block(np.zeros(n, np.uint8))
            host = np.asarray(vec)
            w = host.real ** 2 + host.imag ** 2
            cdf = np.cumsum(w) / np.cumsum(w)[-1]
            us = [0.0, 0.37, 0.999999, float(np.nextafter(cdf[0], 0.0)), float(cdf[0]),
                  float(np.nextafter(cdf[-2], 1.0))]
            for u in us:
                # uniforms one float away from a CDF edge may be refused by the independent interval
                # check; both entry points must then refuse alike
                assert outcome(lambda: ctx.draw_noalloc(vec.ptr, e.k, u)) == outcome(lambda: vec.draw(u)), (n, u)
            assert ctx.draw_noalloc(vec.ptr, e.k, 0.37) == vec.draw(0.37)
            e.close()
    finally:
        ctx.close()


def test_draw_noalloc_refuses_oversized_blocks_and_bad_uniforms():
    ctx = _native.Context(0)
    e = engine(12, 8, 10, 3)
    vec, _ = e.block(np.zeros(12, np.uint8))
    try:
        with pytest.raises(ValueError, match='at most 2\\^22'):
            ctx.draw_noalloc(vec.ptr, _native.DRAW_NOALLOC_MAX_K + 1, 0.5)
        with pytest.raises(ValueError):
            ctx.draw_noalloc(vec.ptr, -1, 0.5)
        for u in (1.0, -0.25, float('nan')):
            with pytest.raises(ValueError):
                ctx.draw_noalloc(vec.ptr, e.k, u)
        assert ctx.draw_noalloc(vec.ptr, e.k, 0.5) == vec.draw(0.5)
    finally:
        e.close()
        ctx.close()
