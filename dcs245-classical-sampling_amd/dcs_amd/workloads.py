"""Workload inputs: the frozen 64-qubit QAT #245 injected graph shipped with the package."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

DATA = Path(__file__).resolve().parent / 'data'


def load_ibm64():
    """The 378-vertex injected graph of IBM QAT #245 (64 qubits, depth 73, 314 T) with its slice plans.

    Returns a dict: graph u8[378,378], unitaries c128[378,2,2], order, ring, natural_order,
    open_vertices (20 free outputs), n = 64, plans {name: {'operations', 'rmax', 'buffer_GiB'}},
    protocol (generation seed / count / free and closed vertices / mixture weight, worker count, plan
    names, audit selection), circuit_gates (the published circuit, rows (code h=0 s=1 sx=2 t=3 cz=4,
    qubit, qubit or -1)) and provenance (its SHA-256 and the frozen canonical-stabilizer hash)."""
    z = np.load(DATA / 'ibm_qat245_64q.npz', allow_pickle=False)
    N = int(z['N'])
    graph = np.unpackbits(z['graph_packed'], axis=1)[:, :N].astype(np.uint8)
    plans = {p['name']: p for p in json.loads(bytes(z['plans_json']).decode())}
    protocol = json.loads(bytes(z['protocol_json']).decode()) if 'protocol_json' in z.files else None
    provenance = json.loads(bytes(z['provenance_json']).decode()) if 'provenance_json' in z.files else None
    return {'graph': graph, 'unitaries': z['unitaries'], 'order': z['order'], 'ring': z['ring'],
            'natural_order': z['natural_order'], 'open_vertices': [int(q) for q in z['open_vertices']],
            'n': 64, 'plans': plans, 'protocol': protocol, 'provenance': provenance,
            'circuit_gates': z['circuit_gates'] if 'circuit_gates' in z.files else None}
