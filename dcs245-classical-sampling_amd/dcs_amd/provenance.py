"""Provenance of the injected graph: the published circuit's injected stabilizer state against the stored
graph and local Cliffords.

Restates provenance/verify_graph_provenance.py:24-88 without qiskit or stim:

  * injected_circuit_tableau -- the circuit on a stabilizer tableau, each T replaced by a CNOT onto a
    fresh reference qubit (numbered n, n+1, ... in gate order), physical qubit q on tableau qubit
    ring.index(q); sx is H S H (equal as matrices);
  * graph_state_tableau -- H on every vertex, CZ on every edge, then each vertex's 2x2 local unitary as a
    single-qubit Clifford (its action on X, Y, Z read from U P U^dagger);
  * canonical_stabilizers -- the reduced row echelon form of the stabilizer group over the columns
    x_0, z_0, x_1, z_1, ... with signs, rows in pivot order, written '+X_Z...' ('_' for identity): the
    canonical form stim documents for TableauSimulator.canonical_stabilizers (pivot on X0, Z0, X1, Z1,
    ...; multiply the pivot row into every other row holding that generator; swap it into place).  The
    reduced echelon form of a group is unique, so the list and its SHA-256 equal stim's for the same
    state;
  * verify_ibm64 -- both lists equal, and their hash equal to the one the reference recorded
    (inputs/graph_provenance_validation.json).
"""
from __future__ import annotations

import hashlib

import numpy as np

from .stabilizer import Tableau

_PAULI = {'X': np.array([[0, 1], [1, 0]], complex), 'Y': np.array([[0, -1j], [1j, 0]], complex),
          'Z': np.array([[1, 0], [0, -1]], complex)}


def injected_circuit_tableau(gates, n, ring):
    """Tableau of the circuit (rows (code h=0 s=1 sx=2 t=3 cz=4, qubit, qubit)) with T injection."""
    to_tab = {int(q): i for i, q in enumerate(ring)}
    t = int(np.sum(np.asarray(gates)[:, 0] == 3))
    tab = Tableau(n + t)
    ref = n
    for code, a, b in np.asarray(gates, np.int64):
        qa = to_tab[int(a)]
        if code == 0:
            tab.h(qa)
        elif code == 1:
            tab.s(qa)
        elif code == 2:
            tab.h(qa); tab.s(qa); tab.h(qa)
        elif code == 3:
            tab.cnot(qa, ref)
            ref += 1
        elif code == 4:
            tab.cz(qa, to_tab[int(b)])
        else:
            raise ValueError(f'unknown gate code {code}')
    return tab


def clifford_images(U):
    """{P: (image Pauli, sign)} for P in X, Y, Z with U P U^dagger = sign * image; ValueError if U is not a
    single-qubit Clifford (up to phase)."""
    U = np.asarray(U, complex)
    if np.max(np.abs(U.conj().T @ U - np.eye(2))) > 1e-12:
        raise ValueError('local operator is not unitary')
    out = {}
    for p, P in _PAULI.items():
        img = U @ P @ U.conj().T
        for q, Q in _PAULI.items():
            for sign in (1, -1):
                if np.max(np.abs(img - sign * Q)) < 1e-9:
                    out[p] = (q, sign)
        if p not in out:
            raise ValueError('local operator is not a Clifford')
    return out


def apply_single_clifford(tab, q, images):
    """Conjugate every stabiliser row by the Clifford with the given Pauli images on qubit q."""
    x, z = tab.x[:, q].copy(), tab.z[:, q].copy()
    bits = {'X': (1, 0), 'Y': (1, 1), 'Z': (0, 1)}
    nx, nz, flip = np.zeros_like(x), np.zeros_like(z), np.zeros_like(x)
    for p, (xb, zb) in bits.items():
        sel = (x == xb) & (z == zb)
        img, sign = images[p]
        nx[sel], nz[sel] = bits[img]
        if sign < 0:
            flip[sel] = 1
    tab.x[:, q], tab.z[:, q] = nx, nz
    tab.r ^= flip


def graph_state_tableau(graph, unitaries):
    A = np.asarray(graph)
    N = A.shape[0]
    tab = Tableau(N)
    for q in range(N):
        tab.h(q)
    for a, b in zip(*np.nonzero(np.triu(A, 1))):
        tab.cz(int(a), int(b))
    for q in range(N):
        apply_single_clifford(tab, q, clifford_images(unitaries[q]))
    return tab


def _product_signs(x, z, r, px, pz, pr, rows):
    """Signs of pivot * row for the selected rows (Aaronson-Gottesman g function)."""
    x1, z1 = px.astype(np.int64)[None, :], pz.astype(np.int64)[None, :]
    x2, z2 = x[rows].astype(np.int64), z[rows].astype(np.int64)
    g = np.where((x1 == 1) & (z1 == 1), z2 - x2,
                 np.where((x1 == 1) & (z1 == 0), z2 * (2 * x2 - 1),
                          np.where((x1 == 0) & (z1 == 1), x2 * (1 - 2 * z2), 0)))
    tot = (2 * r[rows].astype(np.int64) + 2 * int(pr) + g.sum(axis=1)) % 4
    if np.any(tot % 2):
        raise AssertionError('stabilisers do not commute')
    return (tot // 2).astype(np.uint8)


def canonical_stabilizers(tab):
    """Canonical generator list of the tableau's stabiliser group (see the module docstring)."""
    x, z, r = tab.x.copy(), tab.z.copy(), tab.r.copy()
    N = x.shape[0]
    top = 0
    for q in range(x.shape[1]):
        for M in (x, z):
            cand = np.nonzero(M[top:, q])[0]
            if not len(cand):
                continue
            p = top + int(cand[0])
            rows = np.nonzero(M[:, q])[0]
            rows = rows[rows != p]
            if len(rows):
                r[rows] = _product_signs(x, z, r, x[p], z[p], r[p], rows)
                x[rows] ^= x[p]
                z[rows] ^= z[p]
            if p != top:
                for arr in (x, z, r):
                    arr[[p, top]] = arr[[top, p]]
            top += 1
    if top != N:
        raise AssertionError('stabiliser generators are not independent')
    chars = np.array(['_', 'X', 'Z', 'Y'])
    table = chars[x.astype(np.int64) + 2 * z.astype(np.int64)]
    return [('-' if r[i] else '+') + ''.join(table[i]) for i in range(N)]


def stabilizers_sha256(stabilizers):
    return hashlib.sha256('\n'.join(stabilizers).encode()).hexdigest()


def verify_ibm64(workload=None):
    """The provenance check of the frozen 64-qubit graph; returns the reference's result fields."""
    from .workloads import load_ibm64
    w = workload or load_ibm64()
    gates, prov = w['circuit_gates'], w['provenance']
    n, N = int(w['n']), int(w['graph'].shape[0])
    counts = {name: int(np.sum(gates[:, 0] == c)) for c, name in enumerate(('h', 's', 'sx', 't', 'cz'))}
    if counts['t'] != N - n:
        raise AssertionError('T count differs from the number of reference vertices')
    actual = canonical_stabilizers(injected_circuit_tableau(gates, n, w['ring']))
    rebuilt = canonical_stabilizers(graph_state_tableau(w['graph'], w['unitaries']))
    if actual != rebuilt:
        raise AssertionError('Injected circuit state differs from graph/local-Clifford state')
    digest = stabilizers_sha256(actual)
    return {'passed': digest == prov['canonical_stabilizers_sha256'], 'target': prov['target'],
            'qasm_sha256': prov['qasm_sha256'], 'logical_qubits': n, 'injected_qubits': N,
            'edges': int(np.triu(w['graph'], 1).sum()), 'gate_counts': counts,
            'canonical_stabilizers_checked': len(actual), 'canonical_stabilizers_sha256': digest,
            'expected_canonical_stabilizers_sha256': prov['canonical_stabilizers_sha256']}
