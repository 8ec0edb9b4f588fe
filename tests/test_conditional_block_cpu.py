"""Engine-class behaviour that needs no device."""
import numpy as np
import pytest


def test_half_engine_exports_complex64_only():
    """HalfConditionalBlockGPU keeps the reference's dtype rule (inplace_half_gpu.py:40-42): complex128
    is refused before any device work."""
    from dcs_amd.conditional_block import HalfConditionalBlockGPU
    from dcs_amd.stabilizer import synthetic_instance
    inst = synthetic_instance(10, 6, 4, seed=0)
    e = HalfConditionalBlockGPU(inst['graph'], inst['unitaries'], 10, inst['natural_order'], [], [8, 9])
    with pytest.raises(ValueError, match='complex64'):
        e.alloc(dtype=np.complex128)
