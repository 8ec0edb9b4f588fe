"""Graph-state transforms on the host: local complementation, parity slicing, vertex slicing.

Measuring a reference vertex w in the Z basis removes it (vertex slicing).  Measuring it in Y or X
fixes a parity of w and its neighbours: a local complementation at w (Y) or a pivot on an edge
(w, u) (X), followed by the removal of w.  With
    |G*v> = exp(+i pi/4 X_v) prod_{u in N(v)} exp(-i pi/4 Z_u) |G>   (up to a phase)
the local Cliffords update as U_v <- U_v exp(-i pi/4 X), U_u <- U_u exp(+i pi/4 Z)
(dcs_rankwidth/parity_slicing.py:3-11,23-32).
"""
from __future__ import annotations

import numpy as np

from .scan_plan import structurally_fixed

_C = np.cos(np.pi / 4)
_S = np.sin(np.pi / 4)
# exp(-i pi/4 X) and exp(+i pi/4 Z), built from the same float constants cos(pi/4), sin(pi/4).
RX_MINUS = np.array([[_C, -1j * _S], [-1j * _S, _C]], dtype=complex)
RZ_PLUS = np.array([[_C + 1j * _S, 0], [0, _C - 1j * _S]], dtype=complex)


def lc(A, U, v):
    """Local complementation at v, in place, on adjacency A (uint8 N x N) and the local Cliffords
    U (dict or list of 2x2) (parity_slicing.py:23-32)."""
    nb = np.flatnonzero(A[v])
    if len(nb) > 1:
        A[np.ix_(nb, nb)] ^= 1
        A[nb, nb] = 0
    if U is not None:
        U[v] = U[v] @ RX_MINUS
        for u in nb:
            U[u] = U[u] @ RZ_PLUS


def apply_parity_operations(A, U, operations):
    """Apply [(kind, w, u)] in order: 'Y' = lc(w); 'X' = pivot lc(u) lc(w) lc(u); 'Z' = nothing
    (conditional_block_gpu.py:35-41; parity_slicing.py:115-119)."""
    for kind, w, u in operations:
        if kind == 'Y':
            lc(A, U, int(w))
        elif kind == 'X':
            lc(A, U, int(u))
            lc(A, U, int(w))
            lc(A, U, int(u))
        elif kind != 'Z':
            raise ValueError(kind)


def edges_of(A):
    """Upper-triangle edge list in row-major order (conditional_block_gpu.py:42)."""
    a, b = np.nonzero(np.triu(A, 1))
    return [(int(x), int(y)) for x, y in zip(a, b)]


class SliceReduction:
    """Remove the slice vertices S from the graph (dcs_rankwidth/slicing.py:75-109).

    Fixing slice vertex s to value a contributes f_s(a), turns each edge (s, u) into a sign
    (-1)^{a y_u} absorbed into f_u, and each edge inside S into a constant sign.  Attributes:
        keep, relabel       surviving vertices and their new labels 0..N-1
        nbrs_in_keep[s]     new labels of the surviving neighbours of s
        slice_edges         edges with both ends in S
        N, edges, U         reduced graph and local Cliffords
        fixed               structurally fixed vertices of the reduced graph (new labels)
        order               reduced order (new labels)
    """

    def __init__(self, N, n, order, slices, edges, U):
        self.N_full, self.n, self.slices = N, n, [int(s) for s in slices]
        self.U_full, self.edges_full = U, edges
        S = set(self.slices)
        self.keep = [v for v in range(N) if v not in S]
        self.relabel = {v: i for i, v in enumerate(self.keep)}
        self.nbrs_in_keep = {s: [] for s in self.slices}
        self.slice_edges = []
        red = []
        for a, b in edges:
            if a in S and b in S:
                self.slice_edges.append((a, b))
            elif a in S:
                self.nbrs_in_keep[a].append(self.relabel[b])
            elif b in S:
                self.nbrs_in_keep[b].append(self.relabel[a])
            else:
                red.append((self.relabel[a], self.relabel[b]))
        self.N = len(self.keep)
        self.edges = red
        self.U = {self.relabel[v]: U[v] for v in self.keep}
        fx = structurally_fixed(U, N, n)
        self.fixed = {self.relabel[v] for v in fx if v in self.relabel}
        self.order = [self.relabel[int(v)] for v in order if int(v) in self.relabel]
