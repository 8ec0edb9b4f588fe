"""Write plan fixtures: the reference's own make_plan and compile_inplace run on reduced graphs.

Runs only in a container that holds the reference checkout (argument 1, default /root/reference).  The
two planner files are numpy-only and are loaded by file path; no other reference code is imported.
The reduced graph, order, future masks and fixed set are built by this repository's BlockPlan (they are
pinned separately by the frozen rmax / work checksums), then handed to the reference planners.  Output:
tests/golden/plan_fixtures.json with every Step / InplaceStep field per case.
"""
import importlib.util
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / 'dcs245-classical-sampling_amd'))


def load(path, name):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def cases():
    from dcs_amd.block_plan import BlockPlan
    from dcs_amd.stabilizer import synthetic_instance
    from dcs_amd.workloads import load_ibm64
    w = load_ibm64()
    for name in ('inplace_k2_fp32', 'inplace_k3_fp64'):
        bp = BlockPlan(w['graph'], w['unitaries'], w['n'], w['order'], w['plans'][name]['operations'],
                       w['open_vertices'])
        yield 'ibm64_' + name, bp
    inst = synthetic_instance(16, 12, 20, seed=0)
    yield 'ring16_d12_t20_s0', BlockPlan(inst['graph'], inst['unitaries'], 16, inst['natural_order'], [], list(range(16)))
    inst = synthetic_instance(32, 40, 120, seed=0)
    yield 'ring32_d40_t120_s0', BlockPlan(inst['graph'], inst['unitaries'], 32, inst['natural_order'], [],
                                          list(range(16, 32)))


def main(ref):
    ref = Path(ref)
    sp = load(ref / 'source/dcs_rankwidth/scan_plan.py', 'ref_scan_plan')
    ip = load(ref / 'source/kernel_opt/inplace_plan.py', 'ref_inplace_plan')
    out = {}
    for name, bp in cases():
        steps, profile = sp.make_plan(bp.N, bp.edges, bp.order, bp.fixed)
        isteps, basis, maxslots = ip.compile_inplace(bp.order, bp.fut, bp.fixed, bp.stop)
        out[name] = {
            'inputs': {'N': bp.N, 'order': list(map(int, bp.order)), 'fixed': sorted(map(int, bp.fixed)),
                       'stop': bp.stop, 'n_edges': len(bp.edges)},
            'make_plan': [{'v': int(s.v), 'free': bool(s.free), 'r_in': int(s.r_in), 'r_out': int(s.r_out),
                           'beta': int(s.beta), 'Lcols': list(map(int, s.Lcols)), 'kernel': list(map(int, s.kernel))}
                          for s in steps],
            'profile': list(map(int, profile)),
            'compile_inplace': [{'v': int(s.v), 'j': int(s.j), 'kind': s.kind, 'active_in': list(map(int, s.active_in)),
                                 'active_out': list(map(int, s.active_out)), 'beta': int(s.beta), 'K': int(s.K),
                                 'Ka': int(s.Ka), 'pivot': int(s.pivot), 'pivot_a': int(s.pivot_a),
                                 'free': bool(s.free)} for s in isteps],
            'terminal_basis': {str(k): int(v) for k, v in sorted(basis.items())},
            'maxslots': int(maxslots),
        }
        print(name, len(steps), len(isteps), maxslots, flush=True)
    dst = Path(__file__).resolve().parent / 'plan_fixtures.json'
    dst.write_text(json.dumps(out, separators=(',', ':')))
    print(dst, dst.stat().st_size)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
