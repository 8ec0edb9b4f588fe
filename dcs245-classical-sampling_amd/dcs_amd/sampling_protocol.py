"""Sampling protocol and CDF helpers (host side of the draw).

Same functions, arguments, outputs and errors as sampling/sampling_protocol.py:14-100 and
production/protocol_common.py:48-70, so records written with them replay under the reference's
verifier.  The primary CDF of draw_checked (weights, sequential float64 cumsum, normalisation,
searchsorted side='right') is also computed on the device by dcs_draw; both give the same index and
bounds for the same vector.
"""
from __future__ import annotations

import hashlib
import math

import numpy as np


def make_protocol(seed, count, n, free_vertices, mixture_weight=1.):
    """Deterministic per-sample prefixes and uniforms from one PCG64 stream
    (sampling_protocol.py:14-41): branch uniform, then either the closed bits (uint8) and the draw
    uniform, or all n bits."""
    free = [int(q) for q in free_vertices]
    if len(set(free)) != len(free) or not all(0 <= q < n for q in free):
        raise ValueError('Free vertices must be distinct data indices')
    if not 0 <= mixture_weight <= 1 or count < 1:
        raise ValueError('Invalid mixture weight or sample count')
    closed = [q for q in range(n) if q not in set(free)]
    rng = np.random.Generator(np.random.PCG64(seed))
    records = []
    for sample_id in range(count):
        before = rng.bit_generator.state
        branch_u = float(rng.random())
        expensive = branch_u < mixture_weight
        x = np.zeros(n, np.uint8)
        if expensive:
            x[closed] = rng.integers(0, 2, len(closed), dtype=np.uint8)
            draw_u = float(rng.random())
        else:
            x[:] = rng.integers(0, 2, n, dtype=np.uint8)
            draw_u = None
        records.append({'sample_id': sample_id, 'conditional_branch': expensive, 'branch_uniform': branch_u,
                        'draw_uniform': draw_u, 'input_data_bits': ''.join(map(str, x)),
                        'rng_before': before, 'rng_after': rng.bit_generator.state})
    return {'seed': int(seed), 'count': count, 'n': n, 'free_vertices': free, 'closed_vertices': closed,
            'mixture_weight': mixture_weight,
            'rng': 'NumPy PCG64; branch float64, uint8 vector bits, conditional float64',
            'numpy_version': np.__version__, 'records': records}


def cdf_primary(amplitudes, uniform):
    """The primary CDF draw only: (index, weights, normalised cdf) (sampling_protocol.py:56-65)."""
    v = np.asarray(amplitudes, dtype=np.complex128)
    if v.ndim != 1 or not len(v) or len(v) & (len(v) - 1):
        raise ValueError('Expected a power-of-two one-dimensional amplitude vector')
    if not np.isfinite(v).all() or not 0 <= uniform < 1:
        raise ValueError('Nonfinite vector or invalid uniform')
    weights = np.asarray(v.real * v.real + v.imag * v.imag, dtype='<f8', order='C')
    cumulative = np.cumsum(weights, dtype=np.float64)
    total = float(cumulative[-1])
    if not total > 0 or not np.isfinite(total):
        raise ArithmeticError('Numerically zero/nonfinite block: do not replace the query')
    cumulative /= total
    cumulative[-1] = 1.
    index = int(np.searchsorted(cumulative, uniform, side='right'))
    if index >= len(weights) or weights[index] <= 0:
        raise ArithmeticError('CDF produced an invalid outcome')
    return index, weights, cumulative, total


def draw_checked(amplitudes, uniform):
    """Draw once and check the whole CDF interval with math.fsum (sampling_protocol.py:44-88).

    Never resamples: a disagreement between the numpy CDF and the accurately summed interval raises."""
    v = np.asarray(amplitudes, dtype=np.complex128)
    index, weights, cumulative, total_cdf = cdf_primary(v, uniform)
    independent_mass = math.fsum(map(float, weights))
    prefix_mass = math.fsum(map(float, weights[:index]))
    lo, hi = prefix_mass / independent_mass, (prefix_mass + float(weights[index])) / independent_mass
    if not lo <= uniform < hi:
        raise ArithmeticError('Independent accurately summed CDF interval disagrees; do not resample')
    grid_ends = np.ceil(cumulative * 2. ** 53)
    grid_masses = np.diff(grid_ends, prepend=0.) / 2. ** 53
    draw_tv = float(.5 * np.sum(np.abs(grid_masses - weights / independent_mass), dtype=np.float64))
    return {'A_index': index, 'uniform': float(uniform), 'scaled_mass_cdf': total_cdf,
            'scaled_mass_fsum': independent_mass, 'selected_scaled_weight': float(weights[index]),
            'selected_conditional_probability': float(weights[index] / independent_mass),
            'cdf_lower_numpy': float(cumulative[index - 1]) if index else 0.,
            'cdf_upper_numpy': float(cumulative[index]), 'cdf_lower_fsum': lo, 'cdf_upper_fsum': hi,
            'uniform_to_nearest_boundary': min(uniform - lo, hi - uniform),
            'independent_cdf_check_passed': True,
            'finite_53bit_draw_TV_against_computed_normalized_weights': draw_tv,
            'weights_sha256': hashlib.sha256(weights.tobytes()).hexdigest(),
            'amplitudes_sha256': hashlib.sha256(v.astype('<c16', copy=False).tobytes()).hexdigest(),
            'replay_scope': 'Full CDF checked in memory; hashes/interval retained. Independent '
                            'regeneration requires rerunning the frozen block contraction.'}


def assemble_output(input_bits, free_vertices, A_index, ring=None):
    """Bit j of A_index goes to free vertex free_vertices[j]; then the ring permutation and the
    Qiskit (reversed) string (sampling_protocol.py:91-100)."""
    bits = [int(b) for b in input_bits]
    if not 0 <= A_index < 1 << len(free_vertices):
        raise ValueError('Conditional index outside free-output range')
    for j, q in enumerate(free_vertices):
        bits[q] = (A_index >> j) & 1
    if ring is None:
        ring = list(range(len(bits)))
    if sorted(ring) != list(range(len(bits))):
        raise ValueError('Invalid ring permutation')
    original = [0] * len(bits)
    for q, actual in enumerate(ring):
        original[actual] = bits[q]
    return ''.join(map(str, bits)), ''.join(map(str, original[::-1]))


def actual_cdf_law_vector(amplitudes):
    """Amplitudes whose probabilities equal the sampler's 53-bit grid law, source phases kept
    (production/protocol_common.py:48-70)."""
    a = np.asarray(amplitudes, dtype=np.complex128)
    if a.ndim != 1 or not len(a) or not np.isfinite(a).all():
        raise ValueError('Invalid vector')
    weights = np.asarray(a.real * a.real + a.imag * a.imag, dtype='<f8', order='C')
    cumulative = np.cumsum(weights, dtype=np.float64)
    if not cumulative[-1] > 0:
        raise ArithmeticError('Zero conditional vector; no silent replacement')
    total = float(cumulative[-1])
    cumulative /= total
    cumulative[-1] = 1.
    endpoints = np.ceil(cumulative * 2. ** 53).astype(np.uint64)
    counts = np.diff(np.concatenate((np.zeros(1, np.uint64), endpoints)))
    if int(counts.sum(dtype=np.uint64)) != 1 << 53:
        raise AssertionError('Grid counts do not normalize')
    phases = np.ones(len(a), np.complex128)
    nonzero = np.abs(a) > 0
    phases[nonzero] = a[nonzero] / np.abs(a[nonzero])
    law = phases * np.sqrt(counts.astype(np.float64) / 2. ** 53)

    def sha(b):
        return hashlib.sha256(b).hexdigest()
    return law, {'grid_bits': 53, 'grid_total': 1 << 53,
                 'grid_counts_sha256': sha(counts.astype('<u8', copy=False).tobytes()),
                 'actual_law_vector_sha256': sha(law.astype('<c16', copy=False).tobytes()),
                 'source_amplitudes_sha256': sha(a.astype('<c16', copy=False).tobytes()),
                 'source_weights_sha256': sha(weights.tobytes()), 'source_cdf_total': total,
                 'law_norm_squared': float(np.vdot(law, law).real)}
