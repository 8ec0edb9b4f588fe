"""Provenance of the frozen 64-qubit graph (provenance/verify_graph_provenance.py:24-88): the injected
circuit's canonical stabilisers equal those of the stored graph and local Cliffords, and their SHA-256
equals the hash the reference recorded with stim 1.16 (inputs/graph_provenance_validation.json).  The
canonical form is pinned by closed forms on small states; planted changes to the graph or a local
Clifford must fail the check."""
import numpy as np
import pytest

from dcs_amd.provenance import (canonical_stabilizers, clifford_images, graph_state_tableau,
                                injected_circuit_tableau, stabilizers_sha256, verify_ibm64)
from dcs_amd.stabilizer import Tableau


def test_canonical_form_closed_forms():
    assert canonical_stabilizers(Tableau(3)) == ['+Z__', '+_Z_', '+__Z']
    t = Tableau(2)
    t.h(0); t.cnot(0, 1)                      # Bell state: XX, ZZ
    assert canonical_stabilizers(t) == ['+XX', '+ZZ']
    t = Tableau(1)
    t.h(0); t.s(0); t.s(0); t.h(0)            # X|0> = |1>: -Z
    assert canonical_stabilizers(t) == ['-Z']
    t = Tableau(1)
    t.h(0); t.s(0)                            # S|+>: +Y
    assert canonical_stabilizers(t) == ['+Y']
    t = Tableau(3)
    t.h(0); t.cnot(0, 1); t.cnot(1, 2)        # GHZ: XXX, ZZ_, _ZZ; fully reduced, the Z0-pivot row
    assert canonical_stabilizers(t) == ['+XXX', '+Z_Z', '+_ZZ']   # loses Z1: ZZ_ * _ZZ = Z_Z
    # the form depends only on the group: the same state prepared another way
    u = Tableau(3)
    u.h(2); u.cnot(2, 1); u.cnot(1, 0)
    assert canonical_stabilizers(u) == canonical_stabilizers(t)


def test_clifford_images_of_h_and_s():
    H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    S = np.diag([1, 1j])
    assert clifford_images(H) == {'X': ('Z', 1), 'Y': ('Y', -1), 'Z': ('X', 1)}
    assert clifford_images(S) == {'X': ('Y', 1), 'Y': ('X', -1), 'Z': ('Z', 1)}
    with pytest.raises(ValueError):
        clifford_images(np.diag([1, np.exp(1j * np.pi / 4)]))   # T is not Clifford


def test_ibm64_injected_circuit_matches_graph_and_frozen_hash(ibm64):
    res = verify_ibm64(ibm64)
    assert res['passed'] and res['canonical_stabilizers_checked'] == 378 and res['edges'] == 30078
    assert res['canonical_stabilizers_sha256'] == ibm64['provenance']['canonical_stabilizers_sha256']
    assert res['gate_counts'] == {'h': 125, 's': 2395, 'sx': 4641, 't': 314, 'cz': 2336}


def test_ibm64_planted_changes_fail(ibm64):
    circuit = canonical_stabilizers(injected_circuit_tableau(ibm64['circuit_gates'], 64, ibm64['ring']))
    g = ibm64['graph'].copy()
    a, b = 3, 200
    g[a, b] ^= 1; g[b, a] ^= 1
    assert canonical_stabilizers(graph_state_tableau(g, ibm64['unitaries'])) != circuit
    U = ibm64['unitaries'].copy()
    U[7] = U[7] @ np.diag([1, 1j])
    assert canonical_stabilizers(graph_state_tableau(ibm64['graph'], U)) != circuit
    gates = ibm64['circuit_gates'].copy()
    i = int(np.nonzero(gates[:, 0] == 1)[0][10])
    gates[i, 0] = 0                                               # one S turned into H
    assert stabilizers_sha256(canonical_stabilizers(injected_circuit_tableau(gates, 64, ibm64['ring']))) != \
        ibm64['provenance']['canonical_stabilizers_sha256']
