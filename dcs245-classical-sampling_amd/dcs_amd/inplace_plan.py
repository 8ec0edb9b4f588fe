"""Preserved-coordinate (one-buffer) plan for the in-place frontier executor.

The syndrome basis is kept in physical slots: slot p of the frontier address is the coordinate of
basis vector `basis[p]`.  Slots are reused but never compacted, so every step touches disjoint orbits
of 1, 2 or 4 addresses (generated by the XOR masks K and Ka), and an executor that loads a whole orbit
before writing it is race-free on a single buffer.

Compiler parity with kernel_opt/inplace_plan.py:12-78: the same elimination order (slots ascending,
pivot = highest set bit), the same dependency K, the same lowest-free-slot rule for growth, so every
InplaceStep field (kind, K, Ka, pivot, pivot_a, beta, active sets) is identical.
"""
from __future__ import annotations

from dataclasses import dataclass

# Orbit codes shared with the HIP kernels and the C oracle (kernel_opt/inplace_gpu.py:74-82).
KIND_GROW, KIND_MIX_PAIR, KIND_MIX_DIAG, KIND_REPLACE = 0, 1, 2, 3
KIND_FIXED, KIND_FIXED_SHRINK, KIND_SHRINK_QUAD, KIND_SHRINK_PAIR = 4, 5, 6, 7


@dataclass(slots=True)
class InplaceStep:
    """One in-place step (kernel_opt/inplace_plan.py:12-24)."""
    v: int
    j: int
    kind: str
    active_in: tuple
    active_out: tuple
    beta: int
    K: int = 0
    Ka: int = 0
    pivot: int = -1
    pivot_a: int = -1
    free: bool = True


def eliminate(columns):
    """Echelon form of slot -> vector, slots visited in ascending order (inplace_plan.py:26-36).

    Returns (piv, deps): piv maps a highest bit to (reduced vector, slot combination); deps lists the
    slot combinations that reduce to zero."""
    piv, deps = {}, []
    for slot in sorted(columns):
        x = int(columns[slot])
        coef = 1 << slot
        while x:
            top = x.bit_length() - 1
            hit = piv.get(top)
            if hit is None:
                piv[top] = (x, coef)
                break
            x ^= hit[0]
            coef ^= hit[1]
        if not x:
            deps.append(coef)
    return piv, deps


def solve(vec, piv):
    """Express vec in the echelon basis; returns (residual, slot combination) (inplace_plan.py:38-44)."""
    x = int(vec)
    coef = 0
    while x:
        hit = piv.get(x.bit_length() - 1)
        if hit is None:
            break
        x ^= hit[0]
        coef ^= hit[1]
    return x, coef


def compile_inplace(order, future, fixed, stop=None):
    """Compile the preserved-coordinate plan over order[:stop] (inplace_plan.py:46-78).

    Returns (steps, terminal basis {slot: vector}, maxslots)."""
    basis = {}
    steps = []
    maxslots = 0
    for j, v in enumerate(order[:stop]):
        v = int(v)
        beta = 0
        for p, b in basis.items():
            if b >> j & 1:
                beta |= 1 << p
        active = tuple(sorted(basis))
        clear = ~(1 << j)
        restricted = {p: b & clear for p, b in basis.items()}
        piv, deps = eliminate(restricted)
        if len(deps) > 1:
            raise AssertionError('Deleting one vertex cannot drop more than one rank')
        K = deps[0] if deps else 0
        pk = K.bit_length() - 1
        free = v not in fixed
        st = InplaceStep(v=v, j=j, kind='', active_in=active, active_out=(), beta=beta, K=K,
                         pivot=pk, free=free)
        if not free:
            st.kind = 'fixed_shrink' if K else 'fixed'
            if K:
                del restricted[pk]
        else:
            residual, Ka = solve(future[j], piv)
            if residual:
                if K:
                    st.kind = 'replace'
                    restricted[pk] ^= future[j]
                else:
                    st.kind = 'grow'
                    slot = 0
                    while slot in restricted:
                        slot += 1
                    st.pivot = slot
                    restricted[slot] = future[j]
            elif K:
                st.kind = 'shrink'
                if Ka >> pk & 1:
                    Ka ^= K
                st.Ka = Ka
                st.pivot_a = Ka.bit_length() - 1
                del restricted[pk]
            else:
                st.kind = 'mix'
                st.Ka = Ka
                st.pivot_a = Ka.bit_length() - 1
        basis = restricted
        st.active_out = tuple(sorted(basis))
        steps.append(st)
        if basis:
            maxslots = max(maxslots, max(basis) + 1)
    return steps, basis, maxslots


def orbit_descriptor(st):
    """(orbit code, omitted slots) of a step (inplace_gpu.py:74-82).

    The omitted slots are the coordinates that are fixed to zero in the orbit representative."""
    kind = st.kind
    if kind == 'grow':
        return KIND_GROW, ()
    if kind == 'mix':
        return (KIND_MIX_PAIR, (st.pivot_a,)) if st.Ka else (KIND_MIX_DIAG, ())
    if kind == 'replace':
        return KIND_REPLACE, (st.pivot,)
    if kind == 'fixed':
        return KIND_FIXED, ()
    if kind == 'fixed_shrink':
        return KIND_FIXED_SHRINK, (st.pivot,)
    if kind == 'shrink':
        return (KIND_SHRINK_QUAD, (st.pivot, st.pivot_a)) if st.Ka else (KIND_SHRINK_PAIR, (st.pivot,))
    raise ValueError(kind)


def orbit_slots(st):
    """Physical slots enumerated by the orbit index, ascending (inplace_gpu.py:98)."""
    _, omit = orbit_descriptor(st)
    return [p for p in st.active_in if p not in omit]
