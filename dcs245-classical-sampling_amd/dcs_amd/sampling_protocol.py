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
    """Per-sample prefixes and uniforms from one PCG64 stream (sampling_protocol.py:14-41).

    The stream is consumed in a fixed order per sample -- one float for the branch, then either the
    closed data bits as uint8 integers and one float for the draw (expensive branch), or all n bits
    (uniform branch) -- so the same seed replays the same records."""
    free = [int(q) for q in free_vertices]
    if len(set(free)) != len(free) or any(not 0 <= q < n for q in free):
        raise ValueError(f'free vertices {free} must be distinct data qubits in 0..{n - 1}')
    if not 0 <= mixture_weight <= 1:
        raise ValueError(f'mixture weight {mixture_weight!r} outside [0, 1]')
    if count < 1:
        raise ValueError(f'sample count {count!r} must be positive')
    free_set = set(free)
    closed = [q for q in range(n) if q not in free_set]
    gen = np.random.Generator(np.random.PCG64(seed))
    rows = []
    for sid in range(count):
        state0 = gen.bit_generator.state
        u_branch = float(gen.random())
        conditional = u_branch < mixture_weight
        bits = np.zeros(n, np.uint8)
        if conditional:
            bits[closed] = gen.integers(0, 2, len(closed), dtype=np.uint8)
            u_draw = float(gen.random())
        else:
            bits[:] = gen.integers(0, 2, n, dtype=np.uint8)
            u_draw = None
        rows.append({'sample_id': sid, 'conditional_branch': conditional, 'branch_uniform': u_branch,
                     'draw_uniform': u_draw, 'input_data_bits': ''.join(str(int(b)) for b in bits),
                     'rng_before': state0, 'rng_after': gen.bit_generator.state})
    return {'seed': int(seed), 'count': count, 'n': n, 'free_vertices': free, 'closed_vertices': closed,
            'mixture_weight': mixture_weight,
            'rng': 'NumPy PCG64; branch float64, uint8 vector bits, conditional float64',
            'numpy_version': np.__version__, 'records': rows}


def cdf_primary(amplitudes, uniform):
    """The primary CDF draw only (sampling_protocol.py:56-65): (index, weights, normalised cdf, total).

    weights = re^2 + im^2 in float64, cdf = numpy's sequential cumsum divided by its last entry, which is
    then set to exactly 1, and the index is searchsorted(cdf, u, side='right')."""
    amp = np.asarray(amplitudes, dtype=np.complex128)
    size = amp.shape[0] if amp.ndim == 1 else 0
    if size == 0 or size & (size - 1):
        raise ValueError(f'amplitude vector of shape {amp.shape}: one dimension of power-of-two length expected')
    if not 0 <= uniform < 1:
        raise ValueError(f'uniform {uniform!r} outside [0, 1)')
    if not np.isfinite(amp).all():
        raise ValueError('amplitude vector holds a nonfinite entry')
    w = np.asarray(amp.real * amp.real + amp.imag * amp.imag, dtype='<f8', order='C')
    cdf = np.cumsum(w, dtype=np.float64)
    mass = float(cdf[-1])
    if not (np.isfinite(mass) and mass > 0):
        raise ArithmeticError(f'block mass {mass!r}: the block is zero or nonfinite; the query is not replaced')
    cdf /= mass
    cdf[-1] = 1.
    k = int(np.searchsorted(cdf, uniform, side='right'))
    if k >= size or not w[k] > 0:
        raise ArithmeticError(f'the CDF selected index {k}, which has no positive weight')
    return k, w, cdf, mass


def draw_checked(amplitudes, uniform):
    """One draw, fail-closed (sampling_protocol.py:44-88): the primary CDF picks the index, then the
    interval of that index is recomputed from exactly rounded sums (math.fsum) of the weights; if it does
    not contain the uniform the draw raises instead of being repeated.  Also reports the total variation
    between the sampler's finite 53-bit grid law and the normalised weights, and SHA-256 digests of the
    weights and amplitudes."""
    amp = np.asarray(amplitudes, dtype=np.complex128)
    k, w, cdf, mass = cdf_primary(amp, uniform)
    exact_mass = math.fsum(w)
    below = math.fsum(w[:k])
    lower, upper = below / exact_mass, (below + float(w[k])) / exact_mass
    if not lower <= uniform < upper:
        raise ArithmeticError(f'the exactly summed interval [{lower!r}, {upper!r}) of index {k} does not contain '
                              f'the uniform {uniform!r}; the draw is refused, not repeated')
    # random() yields multiples of 2^-53: bin i receives the grid points in [cdf[i-1], cdf[i])
    grid = np.diff(np.ceil(cdf * 2. ** 53), prepend=0.) / 2. ** 53
    tv = float(.5 * np.sum(np.abs(grid - w / exact_mass), dtype=np.float64))
    return {'A_index': k, 'uniform': float(uniform), 'scaled_mass_cdf': mass,
            'scaled_mass_fsum': exact_mass, 'selected_scaled_weight': float(w[k]),
            'selected_conditional_probability': float(w[k] / exact_mass),
            'cdf_lower_numpy': float(cdf[k - 1]) if k else 0., 'cdf_upper_numpy': float(cdf[k]),
            'cdf_lower_fsum': lower, 'cdf_upper_fsum': upper,
            'uniform_to_nearest_boundary': min(uniform - lower, upper - uniform),
            'independent_cdf_check_passed': True,
            'finite_53bit_draw_TV_against_computed_normalized_weights': tv,
            'weights_sha256': hashlib.sha256(w.tobytes()).hexdigest(),
            'amplitudes_sha256': hashlib.sha256(amp.astype('<c16', copy=False).tobytes()).hexdigest(),
            'replay_scope': 'The whole CDF was checked in memory and only its digests and the drawn interval '
                            'are kept; replaying the draw needs the block contracted again.'}


def assemble_output(input_bits, free_vertices, A_index, ring=None):
    """Output bitstrings of one sample (sampling_protocol.py:91-100): bit j of A_index is written to free
    data qubit free_vertices[j]; then the ring permutation maps logical qubit q to physical ring[q], and
    the physical string is reversed (Qiskit order).  Returns (logical bits, Qiskit string)."""
    logical = [int(b) for b in input_bits]
    nfree = len(free_vertices)
    if not 0 <= A_index < 1 << nfree:
        raise ValueError(f'index {A_index} outside the {1 << nfree} outcomes of {nfree} free qubits')
    for j, q in enumerate(free_vertices):
        logical[q] = (A_index >> j) & 1
    perm = list(range(len(logical))) if ring is None else list(ring)
    if sorted(perm) != list(range(len(logical))):
        raise ValueError(f'ring {perm} is not a permutation of the {len(logical)} qubits')
    physical = [0] * len(logical)
    for q, p in enumerate(perm):
        physical[p] = logical[q]
    return ''.join(map(str, logical)), ''.join(map(str, reversed(physical)))


def actual_cdf_law_vector(amplitudes):
    """The law the sampler really draws from, as amplitudes (production/protocol_common.py:48-70): outcome
    i receives the multiples of 2^-53 in [cdf[i-1], cdf[i]), so its probability is that count over 2^53;
    each amplitude keeps the phase of its source entry (phase 1 for a zero entry)."""
    amp = np.asarray(amplitudes, dtype=np.complex128)
    if amp.ndim != 1 or amp.size == 0 or not np.isfinite(amp).all():
        raise ValueError(f'expected a nonempty, finite one-dimensional vector, got shape {amp.shape}')
    w = np.asarray(amp.real * amp.real + amp.imag * amp.imag, dtype='<f8', order='C')
    cdf = np.cumsum(w, dtype=np.float64)
    mass = float(cdf[-1])
    if not mass > 0:
        raise ArithmeticError('the block is zero: there is no law to report and none is substituted')
    cdf /= mass
    cdf[-1] = 1.
    ends = np.ceil(cdf * 2. ** 53).astype(np.uint64)
    counts = np.diff(ends, prepend=np.uint64(0))
    if int(counts.sum(dtype=np.uint64)) != 1 << 53:
        raise AssertionError(f'grid counts sum to {int(counts.sum(dtype=np.uint64))}, not 2^53')
    phase = np.ones(amp.size, np.complex128)
    nz = np.abs(amp) > 0
    phase[nz] = amp[nz] / np.abs(amp[nz])
    law = phase * np.sqrt(counts.astype(np.float64) / 2. ** 53)

    def digest(b):
        return hashlib.sha256(b).hexdigest()
    return law, {'grid_bits': 53, 'grid_total': 1 << 53,
                 'grid_counts_sha256': digest(counts.astype('<u8', copy=False).tobytes()),
                 'actual_law_vector_sha256': digest(law.astype('<c16', copy=False).tobytes()),
                 'source_amplitudes_sha256': digest(amp.astype('<c16', copy=False).tobytes()),
                 'source_weights_sha256': digest(w.tobytes()), 'source_cdf_total': mass,
                 'law_norm_squared': float(np.vdot(law, law).real)}
