"""Audit witnesses and the fidelity certificate of the conditional sampler.

The sampler outputs, with probability lambda, a bitstring drawn as (uniform prefix y, then the free bits
from the normalised block of y), and otherwise a uniform bitstring.  For an audited prefix y let r be
the complex128 reference block scaled by sqrt(D) (D = number of prefixes, so ||r||^2 = D p_y) and a the
vector the sampler actually drew from.  The witness

    s_y = |<r, a>| / ||a||

is the overlap of the sampler's normalised conditional state with the true sub-normalised one, and
averaging s_y over uniform y gives |<psi, phi>| for the state phi the sampler realises (a phase per
prefix chosen to align).  The loss l_y = 1 - min(s_y, 1) lies in [0, 1], so the mean loss over m
preselected, iid uniform prefixes bounds the population mean by the Chernoff inequality for bounded
variables: P(mean <= mu - e) <= exp(-m KL(mu - e || mu)), i.e. with confidence 1 - alpha

    E[l] <= b*,  the root in [mean, 1] of  m KL(mean || b) = log(1 / alpha)

(KL between Bernoulli laws, in nats).  Then |<psi, phi>| >= 1 - b* (less any stated numerical error
bounds), the conditional state fidelity is at least (1 - b*)^2, and the lambda-mixture with the
maximally mixed state has fidelity at least lambda (1 - b*)^2 + (1 - lambda) 2^-n.

Record keys follow the reference's audit rows (review/full_vector_audit_certificate.py:15-43) and
certificate (review/conditional_mixture_certificate.py:59-98) so that records replay under its
verifier; the numbers are pinned by the frozen 64-qubit run's recorded certificate.
"""
from __future__ import annotations

import math

import numpy as np

_ASSUMPTIONS = (
    'The number of outputs and the mixture weight are fixed before any block is computed.',
    'Each audited prefix is a fresh uniform draw chosen before its mass is known; none is dropped.',
    'Masses and witnesses are taken before the conditional block is normalised.',
    'Only sampling error is bounded here; floating-point error needs its own bound.')


def audit_vector_pair(reference_scaled, approximate_vector, audit_id):
    """Witness row of one audited prefix: s = |<r, a>| / ||a||, weight min(s, 1), loss 1 - weight, and the
    normalised overlap s / ||r|| (capped at 1) with its square."""
    r = np.asarray(reference_scaled, dtype=np.complex128)
    a = np.asarray(approximate_vector, dtype=np.complex128)
    if r.ndim != 1 or r.shape != a.shape or r.size == 0:
        raise ValueError(f'audit {audit_id}: the two vectors must be one-dimensional, nonempty and of one length')
    if not (np.isfinite(r).all() and np.isfinite(a).all()):
        raise ValueError(f'audit {audit_id}: a vector holds a nonfinite entry (the audit counts, it is not skipped)')
    r_mass = float(np.vdot(r, r).real)
    a_mass = float(np.vdot(a, a).real)
    if not (math.isfinite(r_mass) and math.isfinite(a_mass)) or a_mass <= 0:
        raise ValueError(f'audit {audit_id}: the sampled vector has no usable norm')
    s = float(abs(np.vdot(r, a)) / math.sqrt(a_mass))
    weight = min(s, 1.)
    overlap = min(1., s / math.sqrt(r_mass)) if r_mass > 0 else None
    return {'audit_id': audit_id, 'vector_length': int(r.size), 'reference_scaled_mass': r_mass,
            'approximate_input_mass': a_mass, 'normalized_conditional_overlap_abs': overlap,
            'normalized_conditional_fidelity': None if overlap is None else overlap * overlap,
            'uncapped_overlap_witness': s, 'clipped_weight': weight, 'loss': 1 - weight,
            'reference_arithmetic': 'complex128 reference block (floating point, not interval arithmetic)'}


def kl_bernoulli(a, b):
    """KL(Bernoulli(a) || Bernoulli(b)) in nats, with the limits 0 log 0 = 0 and +inf where b hits 0 or 1."""
    for v in (a, b):
        if not (math.isfinite(v) and 0 <= v <= 1):
            raise ValueError(f'Bernoulli parameter {v!r} outside [0, 1]')
    if a == b:
        return 0.
    if b == 0. or b == 1.:
        return math.inf
    ka = 0. if a == 0 else a * math.log(a / b)
    kb = 0. if a == 1 else (1 - a) * (math.log1p(-a) - math.log1p(-b))
    return ka + kb


def _smallest_at_least(g, lo, hi, target):
    """Smallest float x in (lo, hi] with g(x) >= target, for g nondecreasing on [lo, hi], g(lo) < target and
    g(hi) >= target: halve the bracket until its ends are adjacent floats."""
    while True:
        mid = lo + (hi - lo) / 2
        if mid <= lo or mid >= hi:
            return hi
        if g(mid) >= target:
            hi = mid
        else:
            lo = mid


def bounded_mean_upper(mean, n, alpha=.05):
    """Upper confidence bound b* on the mean of n iid [0, 1] variables with observed mean `mean`: the
    smallest b >= mean with n KL(mean || b) >= log(1 / alpha).  No observations give 1."""
    if isinstance(n, bool) or not isinstance(n, int) or n < 0:
        raise ValueError(f'observation count {n!r} is not a nonnegative integer')
    if not (math.isfinite(alpha) and 0 < alpha < 1):
        raise ValueError(f'confidence parameter alpha={alpha!r} outside (0, 1)')
    if n == 0:
        return 1.
    if mean is None or not (math.isfinite(mean) and 0 <= mean <= 1):
        raise ValueError(f'observed mean {mean!r} outside [0, 1]')
    t = math.log(1 / alpha) / n
    if mean == 1:
        return 1.
    if mean == 0:
        return -math.expm1(-t)          # KL(0 || b) = -log(1 - b)
    return _smallest_at_least(lambda b: kl_bernoulli(mean, b), mean, 1., t)


def certificate_from_losses(losses, mixture_weight, n_qubits=64, alpha=.05, mean_weight_error_bound=0.,
                            conditional_state_l2_error_bound=0.):
    """Fidelity certificate of lambda |phi><phi| + (1 - lambda) I / 2^n from the audit losses."""
    if not (math.isfinite(mixture_weight) and 0 <= mixture_weight <= 1):
        raise ValueError(f'mixture weight {mixture_weight!r} outside [0, 1]')
    if isinstance(n_qubits, bool) or not isinstance(n_qubits, int) or n_qubits < 1:
        raise ValueError(f'qubit count {n_qubits!r} is not a positive integer')
    errs = (mean_weight_error_bound, conditional_state_l2_error_bound)
    if any(not math.isfinite(e) or e < 0 for e in errs):
        raise ValueError('numerical error bounds must be finite and nonnegative')
    ls = [float(v) for v in losses]
    bad = [v for v in ls if not (math.isfinite(v) and 0 <= v <= 1)]
    if bad:
        raise ValueError(f'loss {bad[0]!r} outside [0, 1]')
    m = len(ls)
    mean = math.fsum(ls) / m if m else None
    upper = bounded_mean_upper(mean, m, alpha)
    overlap = max(0., 1 - upper - sum(errs))
    fid = overlap * overlap
    uniform = math.ldexp(1 - mixture_weight, -n_qubits)
    return {'statistic': 'one-sided bounded-loss KL-Chernoff', 'alpha': alpha, 'expensive_branch_count': m,
            'observed_mean_loss': mean, 'population_mean_loss_upper': upper, 'overlap_lower': overlap,
            'conditional_state_fidelity_lower': fid, 'mixture_weight': mixture_weight, 'n_qubits': n_qubits,
            'uniform_state_fidelity_contribution': uniform,
            'mixture_state_fidelity_lower': mixture_weight * fid + uniform,
            'mean_weight_error_bound': mean_weight_error_bound,
            'conditional_state_l2_error_bound': conditional_state_l2_error_bound,
            'scope': 'A statistical bound in exact arithmetic; floating-point error enters only through the '
                     'error bounds supplied.',
            'requirements': list(_ASSUMPTIONS)}


def certificate_from_scaled_masses(scaled_masses, mixture_weight, **kwargs):
    """The certificate from block masses z_y = D p_y alone: the witness of a block against itself is
    sqrt(z_y), so the loss is 1 - min(sqrt z_y, 1)."""
    z = [float(v) for v in scaled_masses]
    if any(not math.isfinite(v) or v < 0 for v in z):
        raise ValueError('block masses must be finite and nonnegative')
    out = certificate_from_losses([1 - min(math.sqrt(v), 1.) for v in z], mixture_weight, **kwargs)
    out['observed_min_scaled_mass'] = min(z, default=None)
    out['observed_max_scaled_mass'] = max(z, default=None)
    return out


def certificate_from_audits(audits, expected_audit_ids, mixture_weight=1., total_expensive_branches=None,
                            n_qubits=64, alpha=.05, reference_mean_witness_error_bound=0.):
    """The certificate over audit rows, which must be exactly the preselected audits: each expected id once,
    no other id (the same prefix may occur under two ids)."""
    rows = list(audits)
    ids = [row['audit_id'] for row in rows]
    want = list(expected_audit_ids)
    for label, seq in (('preselected', want), ('supplied', ids)):
        if len(set(seq)) != len(seq):
            raise ValueError(f'an audit id occurs twice among the {label} audits')
    if sorted(ids) != sorted(want):
        missing, extra = sorted(set(want) - set(ids)), sorted(set(ids) - set(want))
        raise ValueError(f'audit rows do not match the preselection (missing {missing[:5]}, extra {extra[:5]})')
    if total_expensive_branches is not None and (isinstance(total_expensive_branches, bool)
                                                 or not isinstance(total_expensive_branches, int)
                                                 or total_expensive_branches < len(rows)):
        raise ValueError(f'expensive branch count {total_expensive_branches!r} is below the audit count')
    losses = []
    for row in rows:
        s = float(row['uncapped_overlap_witness'])
        if not (math.isfinite(s) and s >= 0):
            raise ValueError(f'audit {row["audit_id"]}: witness {s!r} is not a finite nonnegative number')
        losses.append(1 - min(s, 1.))
    out = certificate_from_losses(losses, mixture_weight, n_qubits=n_qubits, alpha=alpha,
                                  mean_weight_error_bound=reference_mean_witness_error_bound)
    out['audit_count'] = out.pop('expensive_branch_count')
    out['total_expensive_branches'] = total_expensive_branches
    out['audit_ids'] = ids
    return out


def required_mean_loss(mixture_weight, n, target, alpha=.05, n_qubits=64):
    """Largest observed mean loss over n audits whose certificate still exceeds `target` fidelity; None
    when zero loss cannot reach it."""
    uniform = math.ldexp(1 - mixture_weight, -n_qubits)

    def fidelity(mean):
        return mixture_weight * (1 - bounded_mean_upper(mean, n, alpha)) ** 2 + uniform

    if fidelity(0.) <= target:
        return None
    # fidelity falls as the mean rises: the first mean at which it no longer exceeds target, one float down
    first_fail = _smallest_at_least(lambda m: -fidelity(m), 0., 1., -target) if fidelity(1.) <= target else 1.
    return first_fail if fidelity(first_fail) > target else math.nextafter(first_fail, 0.)
