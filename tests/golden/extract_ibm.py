"""Write ibm_qat245_64q.npz: the frozen 64-qubit QAT #245 injected graph and its slice plans.

Source files (read-only, in the reference checkout): inputs/injected_graph.npz (graph u8[378,378],
unitaries c128[378,2,2], order, ring, natural_order), inputs/plans.json (open vertices and the
parity-operation lists of every plan) and protocol.json (the frozen sampling protocol's parameters:
generation seed, count, free / closed vertices, mixture weight, worker count, plan names, the audit
selection; not its 1,389 per-record RNG states, which make_protocol regenerates from the seed).  The
adjacency is stored bit-packed; nothing else changes.  For the provenance check it also stores the gate
list of inputs/circuit.qasm (one row per gate: code h=0 s=1 sx=2 t=3 cz=4, qubit, second qubit or -1),
the circuit file's SHA-256 and the frozen canonical-stabilizer hash of
inputs/graph_provenance_validation.json.

Usage (build container): python tests/golden/extract_ibm.py /root/reference
"""
import hashlib
import json
import re
import sys
from pathlib import Path

import numpy as np


def main(ref):
    ref = Path(ref)
    src = np.load(ref / 'inputs/injected_graph.npz', allow_pickle=False)
    plans = json.loads((ref / 'inputs/plans.json').read_text())
    prot = json.loads((ref / 'protocol.json').read_text())
    g = prot['generation']
    protocol = {
        'schema_version': prot['schema_version'], 'engine': prot['engine'], 'worker_count': prot['worker_count'],
        'generation': {k: g[k] for k in ('seed', 'count', 'n', 'free_vertices', 'closed_vertices', 'mixture_weight',
                                         'rng', 'numpy_version')},
        'generation_plan': prot['generation_plan'], 'reference_plan': prot['reference_plan'],
        'audit': prot['audit'], 'anchor_input_data_bits': prot['anchor']['input_data_bits'], 'law': prot['law']}
    qasm = (ref / 'inputs/circuit.qasm').read_bytes()
    codes = {'h': 0, 's': 1, 'sx': 2, 't': 3, 'cz': 4}
    gates = []
    for line in qasm.decode().splitlines()[3:]:
        m = re.fullmatch(r'(h|s|sx|t) q\[(\d+)\];|cz q\[(\d+)\],q\[(\d+)\];', line.strip())
        if m is None:
            raise ValueError(f'unexpected circuit line {line!r}')
        gates.append((codes[m.group(1)], int(m.group(2)), -1) if m.group(1) else (4, int(m.group(3)), int(m.group(4))))
    validation = json.loads((ref / 'inputs/graph_provenance_validation.json').read_text())
    provenance = {'qasm_sha256': hashlib.sha256(qasm).hexdigest(), 'target': validation['target'],
                  'canonical_stabilizers_sha256': validation['canonical_stabilizers_sha256'],
                  'canonical_stabilizers_checked': validation['canonical_stabilizers_checked'],
                  'edges': validation['edges'], 'qiskit_gate_counts': validation['qiskit_gate_counts'],
                  'stim_version': validation['stim_version']}
    graph = src['graph'].astype(np.uint8)
    assert (graph == graph.T).all() and not graph.diagonal().any()
    out = {
        'graph_packed': np.packbits(graph, axis=1),
        'N': np.int64(graph.shape[0]),
        'unitaries': src['unitaries'].astype(np.complex128),
        'order': src['order'].astype(np.int64),
        'ring': src['ring'].astype(np.int64),
        'natural_order': src['natural_order'].astype(np.int64),
        'open_vertices': np.asarray(plans['open_vertices'], np.int64),
        'plans_json': np.frombuffer(json.dumps(plans['plans']).encode(), np.uint8),
        'protocol_json': np.frombuffer(json.dumps(protocol).encode(), np.uint8),
        'circuit_gates': np.asarray(gates, np.int16),
        'provenance_json': np.frombuffer(json.dumps(provenance).encode(), np.uint8),
    }
    # the workload file ships with the package (dcs_amd/data), which loads it at run time
    dst = Path(__file__).resolve().parents[2] / 'dcs245-classical-sampling_amd' / 'dcs_amd' / 'data' / 'ibm_qat245_64q.npz'
    np.savez_compressed(dst, **out)
    print(dst, dst.stat().st_size)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
