"""Input producer: seeded ring Clifford+T circuits, T-gadget injection and graph-state form.

Every T gate on qubit q becomes a fresh reference qubit r in |0>, a CNOT(q -> r), and a final
contraction of r with <w_T| = (1, e^{i pi/4}); since CNOT|b>|0> = |b>|b> this reproduces
diag(1, e^{i pi/4}) exactly, and <x|C|0> = (<x| (x) <w_T| ...) |S> with |S> a stabiliser state on
n + t qubits (dcs_rankwidth/t_injection.py:6-14,21-39).

|S> is written as a graph state with one local Clifford per vertex, |S> = (U_1 (x) ... (x) U_N)|G>
(the role of stim's to_circuit(method="graph_state") in t_injection.py:62-79).  The conversion here is
the textbook one on a CHP tableau (Aaronson-Gottesman update rules):
  1. row-reduce the X block; apply H to every qubit outside its pivot columns (the X block is then
     invertible, because the Z rows of the X-free generators span the orthogonal complement of
     the X row space);
  2. row-reduce the X block to the identity (phases tracked by Pauli row products);
  3. a Y on the diagonal is turned into X by S^dagger, a minus sign is removed by Z;
  4. the Z block is now the adjacency matrix; U_q = H^[q in Q] S^[Y_q] Z^[sign_q].
The graph is not unique (any local-complementation-equivalent one represents the same state), so it
need not equal stim's; amplitudes agree up to a global phase, probabilities exactly.
"""
from __future__ import annotations

import math

import numpy as np

ONE_QUBIT_LAYER = ('I', 'H', 'S', 'HS', 'SH', 'HSH')

_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2)
_S = np.diag([1, 1j]).astype(np.complex128)
_Z = np.diag([1, -1]).astype(np.complex128)


def ring_circuit(n, depth, t_count, seed):
    """Seeded ring Clifford+T circuit as an op list.

    Layer L (0-based): every qubit draws one of I, H, S, HS, SH, HSH (letters applied left to right in
    time); then T on the qubits whose (L, q) slot was drawn; then CZ on the ring edges (q, q+1 mod n)
    with q even for even L and q odd for odd L.  Exactly `t_count` distinct (layer, qubit) slots
    carry a T, drawn uniformly without replacement.  Ops: ('c1', q, name) | ('T', q) | ('cz', a, b)."""
    if t_count > n * depth:
        raise ValueError('more T gates than slots')
    rng = np.random.default_rng(seed)
    picks = rng.integers(0, len(ONE_QUBIT_LAYER), size=(depth, n))
    slots = rng.choice(n * depth, size=t_count, replace=False)
    has_t = np.zeros(n * depth, bool)
    has_t[slots] = True
    has_t = has_t.reshape(depth, n)
    ops = []
    for L in range(depth):
        for q in range(n):
            name = ONE_QUBIT_LAYER[picks[L, q]]
            for g in name:
                if g != 'I':
                    ops.append(('c1', q, g))
        for q in range(n):
            if has_t[L, q]:
                ops.append(('T', q))
        for q in range(L % 2, n - (1 if n > 2 else 0) if n % 2 else n, 2):
            a, b = q, (q + 1) % n
            if a != b:
                ops.append(('cz', a, b))
    return ops


class Tableau:
    """Stabiliser generators of an N-qubit state: rows are generators, x/z bit matrices, r sign bits
    (Aaronson & Gottesman, PRA 70, 052328 (2004))."""

    def __init__(self, N):
        self.N = N
        self.x = np.zeros((N, N), np.uint8)
        self.z = np.eye(N, dtype=np.uint8)
        self.r = np.zeros(N, np.uint8)

    def h(self, q):
        x, z = self.x[:, q].copy(), self.z[:, q].copy()
        self.r ^= x & z
        self.x[:, q], self.z[:, q] = z, x

    def s(self, q):
        self.r ^= self.x[:, q] & self.z[:, q]
        self.z[:, q] ^= self.x[:, q]

    def sdg(self, q):
        self.s(q); self.s(q); self.s(q)

    def zgate(self, q):
        self.s(q); self.s(q)

    def cnot(self, a, b):
        xa, xb, za, zb = self.x[:, a], self.x[:, b], self.z[:, a], self.z[:, b]
        self.r ^= xa & zb & (xb ^ za ^ 1)
        self.x[:, b] ^= xa
        self.z[:, a] ^= zb

    def cz(self, a, b):
        self.h(b); self.cnot(a, b); self.h(b)

    def rowmul(self, h, i):
        """Generator h <- generator i * generator h, with the product's sign."""
        x1, z1, x2, z2 = (self.x[i].astype(np.int64), self.z[i].astype(np.int64),
                          self.x[h].astype(np.int64), self.z[h].astype(np.int64))
        g = np.where((x1 == 1) & (z1 == 1), z2 - x2,
                     np.where((x1 == 1) & (z1 == 0), z2 * (2 * x2 - 1),
                              np.where((x1 == 0) & (z1 == 1), x2 * (1 - 2 * z2), 0)))
        tot = (2 * int(self.r[h]) + 2 * int(self.r[i]) + int(g.sum())) % 4
        if tot not in (0, 2):
            raise AssertionError('generators do not commute')
        self.r[h] = tot // 2
        self.x[h] ^= self.x[i]
        self.z[h] ^= self.z[i]


def inject(ops, n):
    """Clifford tableau of the injected circuit (t_injection.py:21-39).

    Returns (tableau, N, natural_order, refs_of) with references numbered n, n+1, ... in the order
    the T gates appear; natural_order puts each reference right after its data qubit."""
    t = sum(op[0] == 'T' for op in ops)
    N = n + t
    tab = Tableau(N)
    refs_of = {q: [] for q in range(n)}
    k = n
    for op in ops:
        if op[0] == 'c1':
            {'H': tab.h, 'S': tab.s}[op[2]](op[1])
        elif op[0] == 'cz':
            tab.cz(op[1], op[2])
        elif op[0] == 'T':
            tab.cnot(op[1], k)
            refs_of[op[1]].append(k)
            k += 1
        else:
            raise ValueError(op)
    natural = []
    for q in range(n):
        natural.append(q)
        natural.extend(refs_of[q])
    return tab, N, natural, refs_of


def _rref_rows(tab, cols_x=True, rows=None, start=0):
    """Row-reduce the X block (or Z block) over `rows`; returns (pivot columns, next free row)."""
    N = tab.N
    M = tab.x if cols_x else tab.z
    piv = []
    row = start
    rows = list(range(start, N)) if rows is None else rows
    for c in range(N):
        sel = [i for i in range(row, N) if M[i, c]]
        if not sel:
            continue
        p = sel[0]
        if p != row:
            for arr in (tab.x, tab.z):
                arr[[row, p]] = arr[[p, row]]
            tab.r[[row, p]] = tab.r[[p, row]]
        for i in range(N):
            if i != row and M[i, c]:
                tab.rowmul(i, row)
        piv.append(c)
        row += 1
        if row == N:
            break
    return piv, row


def graph_form(tab):
    """(edges, U): graph edges and exact 2x2 local Cliffords with |S> = (U_1 (x) ... U_N)|G>, up to a
    global phase (t_injection.py:62-79)."""
    tab = _copy(tab)
    N = tab.N
    piv, k = _rref_rows(tab, cols_x=True)
    Q = [q for q in range(N) if q not in set(piv)]
    for q in Q:
        tab.h(q)
    piv2, k2 = _rref_rows(tab, cols_x=True)
    if k2 != N:
        raise AssertionError('X block not invertible after Hadamards')
    # rows are now ordered with pivot i at column i (full-rank RREF = identity)
    if not (tab.x == np.eye(N, dtype=np.uint8)).all():
        raise AssertionError('X block is not the identity')
    ydiag = [q for q in range(N) if tab.z[q, q]]
    for q in ydiag:
        tab.sdg(q)
    signs = [q for q in range(N) if tab.r[q]]
    for q in signs:
        tab.zgate(q)
    if tab.r.any() or np.diag(tab.z).any():
        raise AssertionError('graph normal form failed')
    A = tab.z
    if not (A == A.T).all():
        raise AssertionError('adjacency not symmetric')
    edges = [(int(a), int(b)) for a, b in zip(*np.nonzero(np.triu(A, 1)))]
    Qs, Ys, Zs = set(Q), set(ydiag), set(signs)
    U = {}
    for q in range(N):
        m = np.eye(2, dtype=np.complex128)
        if q in Zs:
            m = _Z @ m
        if q in Ys:
            m = _S @ m
        if q in Qs:
            m = _H @ m
        U[q] = snap(m)
    return edges, U


def _copy(tab):
    t = Tableau.__new__(Tableau)
    t.N, t.x, t.z, t.r = tab.N, tab.x.copy(), tab.z.copy(), tab.r.copy()
    return t


_EXACT = np.array([0.0, 0.5, -0.5, 1.0, -1.0, 2 ** -0.5, -(2 ** -0.5)])


def snap(M):
    """Project each entry onto the exact Clifford values {0, +-1/2, +-1, +-1/sqrt 2}
    (t_injection.py:52-59)."""
    M = np.asarray(M, dtype=np.complex128)

    def f(x):
        return _EXACT[np.argmin(np.abs(_EXACT[None, :] - x.reshape(-1, 1)), axis=1)].reshape(x.shape)
    return f(M.real) + 1j * f(M.imag)


def synthetic_instance(n, depth, t_count, seed):
    """Everything the engine needs for a seeded ring circuit: dict with graph u8[N,N], unitaries
    c128[N,2,2], N, n, natural_order, ops."""
    ops = ring_circuit(n, depth, t_count, seed)
    tab, N, natural, _ = inject(ops, n)
    edges, U = graph_form(tab)
    A = np.zeros((N, N), np.uint8)
    for a, b in edges:
        A[a, b] = A[b, a] = 1
    return {'graph': A, 'unitaries': np.array([U[q] for q in range(N)]), 'N': N, 'n': n,
            'natural_order': natural, 'ops': ops, 'edges': edges}
