"""GF(2) scan plans for the vertex-by-vertex contraction of a graph state.

A graph state with one local Clifford per vertex, contracted against a product bra, is

    <v|S> = 2^{-N/2} sum_y prod_k f_k(y_k) (-1)^{sum_{(a,b) in E} y_a y_b},   f_k = v_k U_k.

Vertices are summed in a fixed order.  After position j the processed vertices enter the rest of the
sum only through their syndrome (for every later vertex, the parity of its processed neighbours set to
one).  The syndrome lives in an affine space o + span(basis); the environment is a table of 2^r
complex entries indexed by coordinates in that basis, r being the cut rank at j.

This module compiles the per-step linear algebra once, in exact integer arithmetic (Python ints are
bit vectors over positions).  Parity with dcs_rankwidth/scan_plan.py:45-158 is field-for-field: the
same basis insertion order, pivots (lowest set bit), right inverse and kernel ordering, so a gather
executor driven by either plan reads the same addresses.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Contraction functional of a T reference qubit: <w| = (1, e^{i pi/4}).  scan_plan.py:28
W_T = np.array([1.0, np.exp(1j * math.pi / 4)])


@dataclass(slots=True)
class Step:
    """One vertex of the scan (scan_plan.py:31-32).

    v       vertex label processed at this position
    free    False when the vertex value y is forced (structurally fixed vertex)
    r_in    rank of the syndrome space before the step
    r_out   rank after the step
    beta    bit i set when basis vector i has the current position set (sign mask)
    Lcols   right inverse P of the coordinate map M: M P[k] = e_k
    kernel  basis of ker M (input combinations that land on the same output)
    """
    v: int
    free: bool
    r_in: int
    r_out: int = 0
    beta: int = 0
    Lcols: list = field(default_factory=list)
    kernel: list = field(default_factory=list)


def future_masks(N, edges, order):
    """Bit mask over positions of the later neighbours of the vertex at each position
    (scan_plan.py:45-55)."""
    where = [0] * N
    for j, q in enumerate(order):
        where[q] = j
    fut = [0] * N
    for a, b in edges:
        pa, pb = where[a], where[b]
        lo, hi = (pa, pb) if pa < pb else (pb, pa)
        fut[lo] |= 1 << hi
    return fut


def _reduce(vec, basis, pivots):
    """Reduce vec against an echelon basis in insertion order (scan_plan.py:35-42).

    Each basis vector owns its lowest set bit as pivot.  Returns (residual, coefficient mask)."""
    coef = 0
    for k in range(len(basis)):
        if vec >> pivots[k] & 1:
            vec ^= basis[k]
            coef |= 1 << k
    return vec, coef


def _insert_columns(columns):
    """Gaussian insertion of `columns` (scan_plan.py:69-78).

    Returns (new_basis, pivots, M) where M[i] is the coordinate mask of column i in the new basis."""
    nb, piv, M = [], [], []
    for col in columns:
        res, coef = _reduce(col, nb, piv)
        if res:
            piv.append((res & -res).bit_length() - 1)
            nb.append(res)
            coef |= 1 << (len(nb) - 1)
        M.append(coef)
    return nb, piv, M


def _right_inverse_and_kernel(M, rank):
    """Right inverse P (M P[k] = e_k) and kernel basis of the column map M (scan_plan.py:79-102).

    The pivot column of output coordinate k is the first column whose highest coordinate is k."""
    first = {}
    for i, m in enumerate(M):
        if m:
            top = m.bit_length() - 1
            if top not in first:
                first[top] = i
    P = [0] * rank
    for k in range(rank):
        i = first[k]
        m = M[i]
        p = 1 << i
        low = m & ((1 << k) - 1)
        while low:
            l = (low & -low).bit_length() - 1
            p ^= P[l]
            low &= low - 1
        P[k] = p
    pivcols = set(first.values())
    K = []
    for i, m in enumerate(M):
        if i in pivcols:
            continue
        kv = 1 << i
        mm = m
        while mm:
            k = (mm & -mm).bit_length() - 1
            kv ^= P[k]
            mm &= mm - 1
        K.append(kv)
    return P, K


def make_plan(N, edges, order, fixed=frozenset()):
    """Compile the scan along `order` (scan_plan.py:58-105).

    `fixed` holds vertices whose value is deterministic; they add no basis vector.  Returns
    (steps, rank profile).  The plan does not depend on the contracted bitstring."""
    fut = future_masks(N, edges, order)
    basis = []
    steps, profile = [], []
    for j in range(N):
        v = int(order[j])
        st = Step(v=v, free=v not in fixed, r_in=len(basis))
        beta = 0
        for i, b in enumerate(basis):
            if b >> j & 1:
                beta |= 1 << i
        st.beta = beta
        clear = ~(1 << j)
        columns = [b & clear for b in basis]
        if st.free:
            columns.append(fut[j])
        nb, _, M = _insert_columns(columns)
        st.r_out = len(nb)
        st.Lcols, st.kernel = _right_inverse_and_kernel(M, st.r_out)
        basis = nb
        steps.append(st)
        profile.append(st.r_out)
    return steps, profile


def terminal_basis_after(steps, fut, stop):
    """Syndrome basis after the first `stop` steps, shifted down by `stop` bits
    (conditional_block_gpu.py:57-67)."""
    basis = []
    for j, st in enumerate(steps[:stop]):
        cols = [b & ~(1 << j) for b in basis] + ([fut[j]] if st.free else [])
        basis, _, _ = _insert_columns(cols)
    return [b >> stop for b in basis]


def local_vectors(U, N, n, x_bits, w=None):
    """Contraction vectors f_q = bra_q U_q (scan_plan.py:108-120).

    Data vertices (q < n) use <x_q|; reference vertices use w: None (T functional everywhere), one
    2-vector, or a dict {reference: 2-vector}."""
    f = {}
    eye = np.eye(2)
    for q in range(N):
        if q < n:
            bra = eye[int(x_bits[q])]
        elif isinstance(w, dict):
            bra = np.asarray(w[q], dtype=np.complex128)
        elif w is None:
            bra = W_T
        else:
            bra = np.asarray(w)
        f[q] = bra @ U[q]
    return f


def structurally_fixed(U, N, n):
    """Vertices whose value is forced for every bitstring: every possible contraction vector has an
    entry below 1e-12 in modulus (scan_plan.py:123-130)."""
    eye = np.eye(2)
    out = set()
    for q in range(N):
        vecs = [eye[0] @ U[q], eye[1] @ U[q]] if q < n else [W_T @ U[q]]
        if all(min(abs(x[0]), abs(x[1])) < 1e-12 for x in vecs):
            out.add(q)
    return out


def byte_tables(cols, nbits):
    """XOR lookup tables: T[b][val] = XOR of cols[8b+t] over the set bits t of val, so that
    c -> XOR_i c_i cols[i] is nb table reads (scan_plan.py:133-145)."""
    nb = max(1, (nbits + 7) // 8)
    T = np.zeros((nb, 256), dtype=np.uint64)
    for b in range(nb):
        base = [cols[8 * b + t] if 8 * b + t < nbits else 0 for t in range(8)]
        row = [0] * 256
        for val in range(1, 256):
            low = (val & -val).bit_length() - 1
            row[val] = row[val & (val - 1)] ^ base[low]
        T[b] = np.array(row, dtype=np.uint64)
    return T


def kernel_combos(kernel):
    """All XOR combinations of the kernel vectors, in binary-counter order (scan_plan.py:148-152)."""
    combos = [0]
    for kv in kernel:
        combos.extend([c ^ kv for c in combos])
    return combos


def plan_work(steps):
    """Gather work of a plan: sum over steps of 2^r_out * 2^|kernel| (conditional_block_gpu.py:55)."""
    return sum((1 << s.r_out) << len(s.kernel) for s in steps)
