"""Width planning: cut ranks of an elimination order, order rotation, greedy parity slicing.

The frontier of a plan has 2^w entries with w the largest cut rank along the order (plus the effect of
structurally fixed vertices), so w decides whether a circuit fits HBM at all.  Two levers:

  * best_rotation -- for a ring circuit, start the natural order (data qubit, its references, next
    qubit, ...) at every data qubit (or a stride of them), put the last k data qubits of that order
    open, and keep the rotation whose in-place plan is narrowest (the reference ships its 64-qubit
    order but no search);
  * choose_parity_slices -- the reference's greedy (dcs_rankwidth/parity_slicing.py:60-103): remove
    k vertices one at a time, each by the Z, Y (local complementation) or X (pivot) measurement that
    most lowers (max cut rank, plateau), trying vertices near the cuts of maximal rank.  Each slice
    doubles the number of contractions per block.

prefix_cut_ranks computes, for every prefix P of the order, rank_GF(2) A[P, order \\ P], from one
column-ordered elimination of [e_0, a_0, e_1, a_1, ...] (a_j the adjacency column of order[j]): after
the first 2i columns the rank is i + cutrank(order[:i]) (parity_slicing.py:35-52 computes the same
quantity by row reduction of the interleaved matrix).
"""
from __future__ import annotations

import numpy as np

from .block_plan import BlockPlan
from .graph import lc


def prefix_cut_ranks(A, order):
    """r[i-1] = cut rank of order[:i] against the rest of `order`, i = 1 .. len(order) - 1."""
    order = [int(v) for v in order]
    pos = {v: j for j, v in enumerate(order)}
    M = len(order)
    cols = []
    for v in order:
        m = 0
        for u in np.flatnonzero(A[v]):
            j = pos.get(int(u))
            if j is not None:
                m |= 1 << j
        cols.append(m)
    basis = {}   # pivot bit -> vector with that highest bit
    rank = 0
    out = []
    for i, v in enumerate(order):
        for vec in (1 << i, cols[i]):
            while vec:
                t = vec.bit_length() - 1
                b = basis.get(t)
                if b is None:
                    basis[t] = vec
                    rank += 1
                    break
                vec ^= b
        if i + 1 < M:
            out.append(rank - (i + 1))
    return np.array(out, np.int64)


def order_cost(A, order):
    """(max cut rank, plateau = sum 2^(r - max)) of an order (parity_slicing.py:55-57)."""
    r = prefix_cut_ranks(A, order)
    if r.size == 0:
        return 0, 0.0
    m = int(r.max())
    return m, float(np.sum(2.0 ** (r - m)))


def choose_parity_slices(N, edges, order, k, window_pad=15, max_cand=60, n_pivots=4, seed=0, allowed=None):
    """Greedy parity slicing (dcs_rankwidth/parity_slicing.py:60-103).  Returns (operations
    [(kind, w, u)], removed vertices).  The candidate window, the random subsampling and the trial order
    follow the reference so that the same seed makes the same choices.  `allowed` (a set of vertices,
    default all) restricts the removable vertices, e.g. to the T references, the only ones the block
    adapter slices (sampling/conditional_block_gpu.py:33-34)."""
    rng = np.random.default_rng(seed)
    A = np.zeros((N, N), dtype=np.uint8)
    for a, b in edges:
        A[a, b] = A[b, a] = 1
    ops, removed = [], []
    alive = [int(v) for v in order]
    for _ in range(k):
        r = prefix_cut_ranks(A, alive)
        m = int(r.max())
        hot = [i for i, x in enumerate(r) if x == m]
        lo, hi = max(0, min(hot) - window_pad), min(len(alive), max(hot) + window_pad + 2)
        cand = alive[lo:hi]
        if allowed is not None:
            cand = [v for v in cand if v in allowed] or [v for v in alive if v in allowed]
        if len(cand) > max_cand:
            cand = list(rng.choice(cand, max_cand, replace=False))
        best = None
        alive_set = set(alive)
        for w in cand:
            w = int(w)
            rest = [x for x in alive if x != w]
            trials = [('Z', w, None), ('Y', w, None)]
            nb = [int(u) for u in np.flatnonzero(A[w]) if int(u) in alive_set]
            for u in (rng.choice(nb, min(n_pivots, len(nb)), replace=False) if nb else []):
                trials.append(('X', w, int(u)))
            for kind, ww, u in trials:
                B = A.copy()
                if kind == 'Y':
                    lc(B, None, ww)
                elif kind == 'X':
                    lc(B, None, u)
                    lc(B, None, ww)
                    lc(B, None, u)
                c = order_cost(B, rest)
                if best is None or c < best[0]:
                    best = (c, kind, ww, u)
        _, kind, w, u = best
        if kind == 'Y':
            lc(A, None, w)
        elif kind == 'X':
            lc(A, None, u)
            lc(A, None, w)
            lc(A, None, u)
        ops.append((kind, w, u))
        removed.append(w)
        alive = [x for x in alive if x != w]
    return ops, removed


def best_rotation(graph, unitaries, n, natural_order, k, parity_operations=(), stride=1):
    """Rotation of a ring circuit's natural order (start at data qubit s, every `stride`-th s) whose
    in-place plan is narrowest with the last k data qubits of the rotated order open.  Returns
    (order, open_vertices, width)."""
    nat = [int(v) for v in natural_order]
    best = None
    for s in range(0, n, max(1, int(stride))):
        p = nat.index(s)
        order = nat[p:] + nat[:p]
        data = [v for v in order if v < n]
        opn = data[len(data) - k:] if k else []
        w = BlockPlan(graph, unitaries, n, order, list(parity_operations), opn).inplace_maxslots
        if best is None or w < best[2]:
            best = (order, opn, w)
    return best


def fit_width(graph, unitaries, n, natural_order, k, max_width, max_slices=6, stride=2, greedy_seeds=(0,),
              cache_path=None):
    """A plan of a ring circuit whose in-place frontier has at most 2^max_width entries, with the fewest
    algorithmic bytes per sample among those tried: the narrowest rotation with the last k data qubits
    open (best_rotation), then, if that is too wide, greedy parity slices of T references
    (choose_parity_slices along the order the plan eliminates, open qubits last), keeping the shortest
    prefix of slices that fits.  Each slice doubles the contractions per block, so fewer slices win
    when both fit.  Returns dict(order, open_vertices, operations, width, slices, bytes_per_sample_c64)
    or None when nothing tried fits.  With cache_path the result is kept in a JSON file under a digest of
    the circuit and the arguments, so a server plans each circuit once."""
    import hashlib
    import json
    from pathlib import Path
    key = None
    if cache_path is not None:
        h = hashlib.sha256()
        for a in (np.ascontiguousarray(graph, np.uint8), np.ascontiguousarray(unitaries, np.complex128),
                  np.asarray(natural_order, np.int64)):
            h.update(a.tobytes())
        h.update(json.dumps([n, k, max_width, max_slices, stride, list(greedy_seeds)]).encode())
        key = h.hexdigest()
        path = Path(cache_path)
        table = json.loads(path.read_text()) if path.exists() else {}
        if key in table:
            hit = table[key]
            if hit is not None:
                hit['operations'] = [tuple(op) for op in hit['operations']]
            return hit
    order, opn, width = best_rotation(graph, unitaries, n, natural_order, k, stride=stride)
    best = None
    if width <= max_width:
        bp = BlockPlan(graph, unitaries, n, order, [], opn)
        best = {'order': [int(v) for v in order], 'open_vertices': [int(v) for v in opn], 'operations': [],
                'width': int(width), 'slices': 0, 'bytes_per_sample_c64': int(bp.bytes_per_slice(8))}
    else:
        A = np.asarray(graph, np.uint8)
        N = len(A)
        edges = [(int(a), int(b)) for a, b in zip(*np.nonzero(np.triu(A, 1)))]
        opn_set = set(opn)
        eliminated = [v for v in order if v not in opn_set] + [v for v in order if v in opn_set]
        for seed in greedy_seeds:
            ops, _ = choose_parity_slices(N, edges, eliminated, max_slices, seed=seed, allowed=set(range(n, N)))
            for j in range(1, len(ops) + 1):
                bp = BlockPlan(graph, unitaries, n, order, ops[:j], opn)
                if bp.inplace_maxslots <= max_width:
                    nbytes = int(bp.bytes_per_slice(8)) << j
                    if best is None or nbytes < best['bytes_per_sample_c64']:
                        best = {'order': [int(v) for v in order], 'open_vertices': [int(v) for v in opn],
                                'operations': [tuple(op) for op in ops[:j]], 'width': int(bp.inplace_maxslots),
                                'slices': j, 'bytes_per_sample_c64': nbytes}
                    break
    if key is not None:
        table[key] = best
        Path(cache_path).write_text(json.dumps(table, indent=1) + '\n')
    return best
