"""Audit witness and KL-Chernoff certificate, pinned by the frozen 64-qubit run's recorded certificate
(run_r1/analysis/verify_records.json quality) over its 64 recorded audit losses, and by the algebraic
identities of the witness."""
import math

import numpy as np
import pytest

from dcs_amd.audit import (audit_vector_pair, bounded_mean_upper, certificate_from_audits,
                           certificate_from_losses, certificate_from_scaled_masses, kl_bernoulli,
                           required_mean_loss)


def test_recorded_64q_certificate_from_recorded_losses(records):
    losses = [a['loss'] for a in records['audits'].values()]
    want = records['certificate']
    assert len(losses) == want['audit_count'] == 64
    got = certificate_from_losses(losses, 1.0, n_qubits=64, alpha=want['alpha'])
    assert abs(got['observed_mean_loss'] - want['observed_mean_loss']) < 1e-15
    assert abs(got['population_mean_loss_upper'] - want['population_mean_loss_upper']) < 1e-12
    assert abs(got['overlap_lower'] - want['overlap_lower']) < 1e-12
    assert abs(got['conditional_state_fidelity_lower'] - want['state_fidelity_lower']) < 1e-12


def test_recorded_witness_losses_are_clipped_witnesses(records):
    for a in records['audits'].values():
        assert a['loss'] == pytest.approx(1 - min(a['witness'], 1.0), abs=1e-15)
        assert a['approx_mass'] == pytest.approx(1.0, abs=1e-12)   # the law vector is normalised


def test_kl_chernoff_bound_properties():
    for mean, n in [(0.0002, 64), (0.01, 100), (0.3, 10)]:
        up = bounded_mean_upper(mean, n, 0.05)
        assert mean < up < 1
        assert kl_bernoulli(mean, up) == pytest.approx(math.log(20) / n, rel=1e-9)
        assert bounded_mean_upper(mean, 4 * n, 0.05) < up
    assert bounded_mean_upper(0.0, 10, 0.05) == pytest.approx(1 - 0.05 ** 0.1, rel=1e-15)
    assert bounded_mean_upper(None, 0) == 1.0
    lam, n = 0.5, 200
    ell = required_mean_loss(lam, n, 0.3)
    c = certificate_from_losses([ell * 0.999] * n, lam, n_qubits=20)
    assert c['mixture_state_fidelity_lower'] > 0.3
    assert certificate_from_losses([min(1.0, ell * 1.01)] * n, lam, n_qubits=20)['mixture_state_fidelity_lower'] <= 0.3


def test_witness_identity_on_synthetic_blocks():
    """sum_y witness_y / D = |<psi, phi>| for phi_y = phase_y * approx_y / (||approx_y|| sqrt D), the state
    the sampler's output distribution realises; a zero reference block has witness 0 and loss 1."""
    rng = np.random.default_rng(11)
    for D, K in [(2, 4), (4, 8), (8, 16)]:
        psi = rng.normal(size=(D, K)) + 1j * rng.normal(size=(D, K))
        psi[0] = 0
        psi /= np.linalg.norm(psi)
        approx = psi + 0.03 * (rng.normal(size=(D, K)) + 1j * rng.normal(size=(D, K)))
        rows, phi = [], np.empty_like(psi)
        for y in range(D):
            a = approx[y] / np.linalg.norm(approx[y])
            inner = np.vdot(psi[y], a)
            phase = np.conj(inner) / abs(inner) if inner else 1.0
            phi[y] = phase * a / math.sqrt(D)
            rows.append(audit_vector_pair(math.sqrt(D) * psi[y], approx[y], y))
        assert abs(math.fsum(r['uncapped_overlap_witness'] for r in rows) / D - abs(np.vdot(psi, phi))) < 1e-14
        assert rows[0]['uncapped_overlap_witness'] == 0 and rows[0]['loss'] == 1
        cert = certificate_from_audits(rows, list(range(D)), 0.45, total_expensive_branches=D, n_qubits=10)
        assert cert['audit_count'] == D
        with pytest.raises(ValueError):
            certificate_from_audits(rows[:-1], list(range(D)), 0.45)
        with pytest.raises(ValueError):
            certificate_from_audits(rows, list(range(1, D + 1)), 0.45)


def test_masses_certificate_and_input_checks():
    c = certificate_from_scaled_masses([0.99, 1.01, 1.0], 1.0, n_qubits=8)
    assert c['observed_min_scaled_mass'] == 0.99 and c['observed_max_scaled_mass'] == 1.01
    with pytest.raises(ValueError):
        audit_vector_pair(np.ones(4), np.zeros(4), 0)
    with pytest.raises(ValueError):
        certificate_from_losses([1.2], 1.0)


def test_bound_is_the_smallest_float_reaching_the_threshold():
    """b* is where n KL(mean || b) first reaches log(1 / alpha): the float below it falls short, and it
    agrees with an independent root finder (scipy's Brent method on the same function) to 1e-14."""
    from scipy.optimize import brentq
    for mean, n, alpha in [(1e-4, 64, 0.05), (0.0216, 64, 0.05), (0.2, 7, 0.01), (0.9, 1000, 0.1)]:
        t = math.log(1 / alpha) / n
        b = bounded_mean_upper(mean, n, alpha)
        assert kl_bernoulli(mean, b) >= t > kl_bernoulli(mean, math.nextafter(b, 0.))
        root = brentq(lambda x: kl_bernoulli(mean, x) - t, mean, 1 - 1e-300, xtol=1e-16, rtol=1e-15)
        assert abs(b - root) <= 1e-14
