"""Full-vector audits and the KL-Chernoff fidelity certificate of the conditional sampler.

Restates review/full_vector_audit_certificate.py:15-88 and review/conditional_mixture_certificate.py:
14-125 of the reference (host-side arithmetic on blocks the engines return):

  * audit_vector_pair -- witness of one audited prefix: the complex128 reference block scaled by sqrt(D)
    against the complete approximate vector the sampler drew from; uncapped witness
    s = |<ref, approx>| / ||approx||, clipped weight min(s, 1), loss 1 - weight;
  * bounded_mean_upper -- one-sided KL-Chernoff upper bound on the mean of iid [0, 1] losses: the
    largest b with n * KL(mean || b) < log(1 / alpha), found by bisection on [mean, 1];
  * certificate_from_losses -- overlap lower bound 1 - upper - error terms, its square (conditional state
    fidelity) and the fidelity of the lambda-mixture with the maximally mixed state;
  * certificate_from_audits / certificate_from_scaled_masses -- the same from audit rows (all preselected
    audits, identity-checked) or from block masses z_y (loss 1 - min(sqrt z, 1)).

The bisection and the clipping rules are the reference's, so a certificate computed here from the same
losses is the same number (pinned against the frozen 64-qubit run's recorded certificate).
"""
from __future__ import annotations

import math

import numpy as np

_REQUIREMENTS = [
    'Fixed total output count; independent Bernoulli labels with the stated mixture_weight.',
    'Every expensive branch uses a fresh iid uniform prefix and one conditional output.',
    'Original full-circuit block masses before conditional normalization; no filtering.',
    'Partition, confidence function and algorithm fixed independently of production masses.',
    'Numerical correctness and RNG/CDF replay require separate validation.']


def audit_vector_pair(reference_scaled, approximate_vector, audit_id):
    """Scalar audit metrics of one prefix (full_vector_audit_certificate.py:15-43)."""
    ref = np.asarray(reference_scaled, dtype=np.complex128)
    approx = np.asarray(approximate_vector, dtype=np.complex128)
    if ref.ndim != 1 or ref.shape != approx.shape or not len(ref):
        raise ValueError('Need two nonempty full vectors with identical shape')
    if not np.isfinite(ref).all() or not np.isfinite(approx).all():
        raise ValueError('Nonfinite vector values; do not filter this audit')
    ref_mass = float(np.vdot(ref, ref).real)
    approx_mass = float(np.vdot(approx, approx).real)
    if not math.isfinite(ref_mass) or not math.isfinite(approx_mass) or approx_mass <= 0:
        raise ValueError('Invalid norm; supply the actual declared fallback if used')
    score = float(abs(np.vdot(ref, approx)) / math.sqrt(approx_mass))
    overlap = min(1., score / math.sqrt(ref_mass)) if ref_mass > 0 else None
    weight = min(score, 1.)
    return {'audit_id': audit_id, 'vector_length': len(ref), 'reference_scaled_mass': ref_mass,
            'approximate_input_mass': approx_mass, 'normalized_conditional_overlap_abs': overlap,
            'normalized_conditional_fidelity': overlap ** 2 if overlap is not None else None,
            'uncapped_overlap_witness': score, 'clipped_weight': weight, 'loss': 1 - weight,
            'reference_arithmetic': 'complex128 numerical reference; not an interval-arithmetic proof'}


def kl_bernoulli(a, b):
    """KL(Bernoulli(a) || Bernoulli(b)) in nats (conditional_mixture_certificate.py:14-25)."""
    if not (math.isfinite(a) and math.isfinite(b) and 0 <= a <= 1 and 0 <= b <= 1):
        raise ValueError('KL arguments must be finite and in [0,1]')
    if a == b:
        return 0.
    if b in (0., 1.):
        return math.inf
    if a == 0:
        return -math.log1p(-b)
    if a == 1:
        return -math.log(b)
    return a * math.log(a / b) + (1 - a) * (math.log1p(-a) - math.log1p(-b))


def bounded_mean_upper(mean, n, alpha=.05):
    """One-sided KL-Chernoff upper bound for the mean of n iid [0, 1] observations
    (conditional_mixture_certificate.py:28-55).  1 for n = 0."""
    if not isinstance(n, int) or isinstance(n, bool) or n < 0:
        raise ValueError('n must be a nonnegative integer')
    if not math.isfinite(alpha) or not 0 < alpha < 1:
        raise ValueError('alpha must be in (0,1)')
    if n == 0:
        return 1.
    if mean is None or not math.isfinite(mean) or not 0 <= mean <= 1:
        raise ValueError('mean must be finite and in [0,1]')
    if mean == 1:
        return 1.
    if mean == 0:
        return -math.expm1(math.log(alpha) / n)
    threshold = math.log(1 / alpha) / n
    lo, hi = mean, 1.
    for _ in range(100):
        mid = (lo + hi) / 2
        if mid == lo or mid == hi:
            break
        if kl_bernoulli(mean, mid) >= threshold:
            hi = mid
        else:
            lo = mid
    return hi


def certificate_from_losses(losses, mixture_weight, n_qubits=64, alpha=.05, mean_weight_error_bound=0.,
                            conditional_state_l2_error_bound=0.):
    """Certificate of lambda |phi><phi| + (1 - lambda) I / 2^n (conditional_mixture_certificate.py:58-98)."""
    if not math.isfinite(mixture_weight) or not 0 <= mixture_weight <= 1:
        raise ValueError('mixture_weight must be in [0,1]')
    if not isinstance(n_qubits, int) or isinstance(n_qubits, bool) or n_qubits < 1:
        raise ValueError('n_qubits must be a positive integer')
    for value in (mean_weight_error_bound, conditional_state_l2_error_bound):
        if not math.isfinite(value) or value < 0:
            raise ValueError('Error bounds must be finite and nonnegative')
    losses = [float(v) for v in losses]
    if any(not math.isfinite(v) or not 0 <= v <= 1 for v in losses):
        raise ValueError('Each loss must be finite and in [0,1]')
    n = len(losses)
    mean = math.fsum(losses) / n if n else None
    upper = bounded_mean_upper(mean, n, alpha)
    overlap = max(0., 1 - upper - mean_weight_error_bound - conditional_state_l2_error_bound)
    pure = overlap ** 2
    floor = math.ldexp(1 - mixture_weight, -n_qubits)
    return {'statistic': 'one-sided bounded-loss KL-Chernoff', 'alpha': alpha, 'expensive_branch_count': n,
            'observed_mean_loss': mean, 'population_mean_loss_upper': upper, 'overlap_lower': overlap,
            'conditional_state_fidelity_lower': pure, 'mixture_weight': mixture_weight, 'n_qubits': n_qubits,
            'uniform_state_fidelity_contribution': floor,
            'mixture_state_fidelity_lower': mixture_weight * pure + floor,
            'mean_weight_error_bound': mean_weight_error_bound,
            'conditional_state_l2_error_bound': conditional_state_l2_error_bound,
            'scope': 'Exact-arithmetic statistical theorem unless separately justified global numerical error '
                     'bounds are supplied.',
            'requirements': list(_REQUIREMENTS)}


def certificate_from_scaled_masses(scaled_masses, mixture_weight, **kwargs):
    """Certificate from block masses z_y = D p_y (conditional_mixture_certificate.py:101-108)."""
    masses = [float(v) for v in scaled_masses]
    if any(not math.isfinite(v) or v < 0 for v in masses):
        raise ValueError('Scaled masses must be finite and nonnegative')
    out = certificate_from_losses([1 - min(math.sqrt(v), 1.) for v in masses], mixture_weight, **kwargs)
    out['observed_min_scaled_mass'] = min(masses) if masses else None
    out['observed_max_scaled_mass'] = max(masses) if masses else None
    return out


def certificate_from_audits(audits, expected_audit_ids, mixture_weight=1., total_expensive_branches=None,
                            n_qubits=64, alpha=.05, reference_mean_witness_error_bound=0.):
    """Certificate from ALL preselected audit rows (full_vector_audit_certificate.py:46-88): the audit ids
    must be exactly the frozen selection; losses come from the uncapped witnesses, clipped at 1."""
    rows = list(audits)
    expected = list(expected_audit_ids)
    got = [row['audit_id'] for row in rows]
    if len(set(expected)) != len(expected) or len(set(got)) != len(got):
        raise ValueError('Repeated audit IDs are invalid; repeated PREFIXES are allowed')
    if set(got) != set(expected) or len(got) != len(expected):
        raise ValueError('Missing, extra or substituted preselected audit IDs')
    if total_expensive_branches is not None and (not isinstance(total_expensive_branches, int)
                                                 or total_expensive_branches < len(rows)):
        raise ValueError('Invalid total expensive branch count')
    losses = []
    for row in rows:
        score = float(row['uncapped_overlap_witness'])
        if not math.isfinite(score) or score < 0:
            raise ValueError('Invalid audit witness; do not filter this audit')
        losses.append(1 - min(score, 1.))
    out = certificate_from_losses(losses, mixture_weight, n_qubits=n_qubits, alpha=alpha,
                                  mean_weight_error_bound=reference_mean_witness_error_bound)
    out['audit_count'] = out.pop('expensive_branch_count')
    out['total_expensive_branches'] = total_expensive_branches
    out['audit_ids'] = got
    return out


def required_mean_loss(mixture_weight, n, target, alpha=.05, n_qubits=64):
    """Largest observed mean loss that still certifies `target` (conditional_mixture_certificate.py:
    111-131); None when even zero loss cannot."""
    floor = math.ldexp(1 - mixture_weight, -n_qubits)

    def bound(mean):
        return mixture_weight * (1 - bounded_mean_upper(mean, n, alpha)) ** 2 + floor

    if bound(0.) <= target:
        return None
    lo, hi = 0., 1.
    for _ in range(100):
        mid = (lo + hi) / 2
        if bound(mid) > target:
            lo = mid
        else:
            hi = mid
    return lo
