"""Oracle of the sampling hot path.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py import this module, and only
as the checker (or, for cpu_baseline, as the timed CPU port).  The product path never calls it.

What each function restates and what pins it:

  contract_inplace   C restatement (liboracle.so) of execute_inplace (kernel_opt/inplace_plan.py:90-135)
                     with the GPU driver's scale chain (kernel_opt/inplace_gpu.py:107-123); storage 2 is
                     the binary16-pair storage of kernel_opt/inplace_half_gpu.py:14-78, whose
                     conversions (f32_to_h16, h16_to_f32) are pinned against numpy's float16 over every
                     half and over float32 boundary and random patterns.  The complex64 and binary16
                     contractions are pinned bit for bit by an independent NumPy restatement of the
                     reference's executor (inplace_plan.py:80-135) with the engines' storage rounding and
                     scale chain (tests/test_storage_restatement.py; misplaced roundings are shown to
                     break it), and to binary16 accuracy by the dense state vector (C1).
                     Pinned by: the dense state vector of the Clifford+T circuit (statevector() below,
                     closed-form gate matrices) through block(); by gather_amplitude(), an independent
                     restatement of the other reference executor (dcs_rankwidth/scan_plan.py:161-195)
                     on the same graph; by the frozen IBM plan checksums (rmax, work per slice).
  block              slice loop, scatter, Walsh-Hadamard / phase / U tail
                     (sampling/conditional_block_gpu.py:128-177).  Pinned by statevector() (probabilities
                     and amplitudes up to one global phase, every prefix).
  draw               weights, sequential cumsum, searchsorted(side='right') (sampling_protocol.py:56-65).
                     Pinned by the reference self-test vectors (sampling_protocol.py:103-117) and brute
                     force interval search.
  statevector        dense simulation of the ring circuit (H, S, CZ, T = diag(1, e^{i pi/4})); the
                     mathematical ground truth, pinned by gate identities in the tests.
  contract_gather,   C restatement (oracle_gather) of the canonical gather executor (dcs_rankwidth/scan_plan.py:
  gather_block       161-195, dcs_rankwidth/gpu_kernel.py:58-101) with the reference's scale chain
                     (sampling/conditional_block_gpu.py:92-126), and its block.  Pinned by the dense state
                     vector (C1, every amplitude) and by block(): the two executors share nothing below the
                     plan's step list, so their agreement to 1e-12 on several graphs checks both.
  gather_amplitude   numpy restatement of sweep_reference / amplitude_reference (scan_plan.py:161-207).
  graph_amplitude    brute-force graph-state sum (scan_plan.py:6-8) for N <= 20.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / 'liboracle.so'
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f'{LIB_PATH} missing: run make -C oracle')
        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        u64p, i32p, dp = P(C.c_uint64), P(C.c_int32), P(C.c_double)
        L.oracle_contract.restype = C.c_int
        L.oracle_contract.argtypes = [C.c_int, C.c_int, i32p, u64p, u64p, i32p, i32p, u64p, u64p, dp, i32p, i32p,
                                      C.c_void_p, dp]
        i64p = P(C.c_int64)
        L.oracle_gather.restype = C.c_int
        L.oracle_gather.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, i32p, i64p, u64p, i32p, i64p, u64p, u64p,
                                    dp, i32p, i32p, C.c_void_p, C.c_void_p, dp]
        L.oracle_accumulate.restype = None
        L.oracle_accumulate.argtypes = [C.c_int, C.c_void_p, C.c_uint64, C.c_int, u64p, C.c_uint64, C.c_double,
                                        C.c_double, dp]
        L.oracle_transform.restype = None
        L.oracle_transform.argtypes = [C.c_int, dp, dp, dp]
        L.oracle_draw.restype = C.c_int
        L.oracle_draw.argtypes = [dp, C.c_uint64, C.c_double, dp, dp]
        L.oracle_threads.restype = C.c_int
        L.oracle_f32_to_h16.restype = None
        L.oracle_f32_to_h16.argtypes = [P(C.c_float), P(C.c_uint16), C.c_uint64]
        L.oracle_h16_to_f32.restype = None
        L.oracle_h16_to_f32.argtypes = [P(C.c_uint16), P(C.c_float), C.c_uint64]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def frexp_scale_chain(m2):
    """ls and liveness of one slice from its per-step maxima (inplace_gpu.py:115-123)."""
    ls, maximum = 0., 1.
    for m in m2:
        ls += math.frexp(maximum)[1] * math.log(2.)
        if not np.isfinite(m):
            raise ArithmeticError('Nonfinite in-place environment')
        if m == 0:
            return ls, False
        maximum = math.sqrt(float(m))
    return ls, True


STORAGE_DTYPES = {0: np.complex64, 1: np.complex128, 2: np.dtype((np.uint16, 2))}


def frontier_buffer(storage, width):
    """Zeroed frontier of 2^width entries: complex64 (storage 0), complex128 (1) or binary16 pairs (2,
    uint16[2^width, 2])."""
    if storage == 2:
        return np.zeros((1 << width, 2), np.uint16)
    return np.zeros(1 << width, STORAGE_DTYPES[storage])


def f32_to_h16(x):
    """float32 -> binary16 bits (round to nearest even), as the oracle stores binary16 entries."""
    a = np.ascontiguousarray(x, np.float32)
    out = np.empty(a.shape, np.uint16)
    lib().oracle_f32_to_h16(_p(a, C.c_float), _p(out, C.c_uint16), a.size)
    return out


def h16_to_f32(h):
    a = np.ascontiguousarray(h, np.uint16)
    out = np.empty(a.shape, np.float32)
    lib().oracle_h16_to_f32(_p(a, C.c_uint16), _p(out, C.c_float), a.size)
    return out


def as_complex(E, storage):
    """Complex view of a frontier in storage precision (binary16 pairs widened to complex64)."""
    if storage != 2:
        return E
    f = h16_to_f32(E)
    return (f[..., 0] + 1j * f[..., 1]).astype(np.complex64)


def contract_inplace(bp, f, yfix, storage=1, buffer=None):
    """Run the in-place plan of BlockPlan `bp` for one slice.  Returns dict(E, m2, ls, alive, offset).
    storage: 0 complex64, 1 complex128, 2 binary16 pairs (kernel_opt/inplace_half_gpu.py)."""
    t = bp.step_table()
    coeffs, ov, ys, offset = bp.step_parameters(f, yfix)
    n = len(t['kind'])
    E = buffer if buffer is not None else frontier_buffer(storage, bp.inplace_maxslots)
    m2 = np.zeros(n, np.float64)
    arrs = {k: np.ascontiguousarray(t[k]) for k in ('kind', 'K', 'Ka', 'pivot', 'pivot_a', 'beta', 'in_mask')}
    lib().oracle_contract(int(storage), n, _p(arrs['kind'], C.c_int32), _p(arrs['K'], C.c_uint64),
                          _p(arrs['Ka'], C.c_uint64), _p(arrs['pivot'], C.c_int32), _p(arrs['pivot_a'], C.c_int32),
                          _p(arrs['beta'], C.c_uint64), _p(arrs['in_mask'], C.c_uint64),
                          _p(np.ascontiguousarray(coeffs), C.c_double), _p(np.ascontiguousarray(ov), C.c_int32),
                          _p(np.ascontiguousarray(ys), C.c_int32), E.ctypes.data_as(C.c_void_p), _p(m2, C.c_double))
    ls, alive = frexp_scale_chain(m2)
    return {'E': E, 'm2': m2, 'ls': ls, 'alive': alive, 'offset': offset}


def terminal_env(bp, E):
    idx = np.zeros(1 << bp.terminal_rank, np.uint64)
    ids = np.arange(1 << bp.terminal_rank, dtype=np.uint64)
    for j, p in enumerate(bp.terminal_slots):
        idx |= ((ids >> np.uint64(j)) & np.uint64(1)) << np.uint64(p)
    return E[idx.astype(np.int64)]


def block(bp, x_bits, w=None, storage=1, part=0, parts=1):
    """Complex128 block sqrt(D) psi[prefix, :] by the oracle path (conditional_block_gpu.py:128-177)."""
    k = bp.k
    total = np.zeros(1 << k, np.complex128)
    tv = total.view(np.float64)
    smap = np.ascontiguousarray(bp.syndrome_map())
    f_full = bp.local_vectors(x_bits, w=w)
    E = frontier_buffer(storage, bp.inplace_maxslots)
    for _, coef, f, yfix, freeflip in bp.slice_assignments(f_full, part, parts):
        if f is None:
            continue
        r = contract_inplace(bp, f, yfix, storage, buffer=E)
        if not r['alive']:
            continue
        c = coef * bp.slice_scale(r['ls'])
        xm = (r['offset'] >> bp.stop) ^ freeflip
        lib().oracle_accumulate(int(storage), E.ctypes.data_as(C.c_void_p), C.c_uint64(bp.terminal_mask()), k,
                                _p(smap, C.c_uint64), C.c_uint64(xm), c.real, c.imag, _p(tv, C.c_double))
    ph = np.ascontiguousarray(bp.phase())
    U = np.ascontiguousarray(bp.open_unitaries().reshape(-1, 4).view(np.float64))
    lib().oracle_transform(k, _p(tv, C.c_double), _p(ph, C.c_double), _p(U, C.c_double))
    return total


def gather_tables(bp):
    """Per-step arrays of the canonical gather plan (byte tables, kernel combinations), concatenated."""
    import sys
    sys.path.insert(0, str(HERE.parent / 'dcs245-classical-sampling_amd'))
    from dcs_amd.scan_plan import byte_tables, kernel_combos
    steps = bp.steps
    tabs = [byte_tables(s.Lcols, s.r_out) for s in steps]
    combs = [np.asarray(kernel_combos(s.kernel), np.uint64) for s in steps]
    out = {
        'r_in': np.array([s.r_in for s in steps], np.int32), 'r_out': np.array([s.r_out for s in steps], np.int32),
        'free': np.array([1 if s.free else 0 for s in steps], np.int32),
        'beta': np.array([s.beta for s in steps], np.uint64),
        'nbytes': np.array([t.shape[0] for t in tabs], np.int32),
        'ncombos': np.array([len(c) for c in combs], np.int32),
        'tables': np.ascontiguousarray(np.concatenate([t.reshape(-1) for t in tabs]), np.uint64),
        'combos': np.ascontiguousarray(np.concatenate(combs), np.uint64)}
    out['tab_off'] = np.concatenate([[0], np.cumsum(out['nbytes'].astype(np.int64) * 256)[:-1]]).astype(np.int64)
    out['comb_off'] = np.concatenate([[0], np.cumsum(out['ncombos'].astype(np.int64))[:-1]]).astype(np.int64)
    return out


def contract_gather(bp, f, yfix, storage=1, tables=None):
    """One slice through the canonical gather executor (oracle_gather): dict(env = the final table's
    first 2^r entries, ls, alive, offset, m2)."""
    g = gather_tables(bp) if tables is None else tables
    coeffs, ov, ys, offset = bp.step_parameters(f, yfix)
    m = len(bp.steps)
    dt = np.complex128 if storage else np.complex64
    A = np.zeros(1 << bp.rmax, dt)
    B = np.zeros(1 << bp.rmax, dt)
    m2 = np.zeros(m, np.float64)
    coeffs = np.ascontiguousarray(coeffs)
    rc = lib().oracle_gather(int(storage), m, _p(g['r_in'], C.c_int32), _p(g['r_out'], C.c_int32),
                             _p(g['free'], C.c_int32), _p(g['nbytes'], C.c_int32), _p(g['tab_off'], C.c_int64),
                             _p(g['tables'], C.c_uint64), _p(g['ncombos'], C.c_int32), _p(g['comb_off'], C.c_int64),
                             _p(g['combos'], C.c_uint64), _p(g['beta'], C.c_uint64), _p(coeffs, C.c_double),
                             _p(ov, C.c_int32), _p(ys, C.c_int32), A.ctypes.data_as(C.c_void_p),
                             B.ctypes.data_as(C.c_void_p), _p(m2, C.c_double))
    ls, alive = frexp_scale_chain(m2)
    r = len(bp.gather_terminal_basis)
    env = ((A, B)[rc][:1 << r].copy()) if rc >= 0 else None
    return {'env': env, 'ls': ls, 'alive': alive and rc >= 0, 'offset': offset, 'm2': m2}


def gather_block(bp, x_bits, w=None, storage=1, part=0, parts=1):
    """Complex128 block by the canonical gather executor (conditional_block_gpu.py:128-177 with
    _prefix :92-126)."""
    k = bp.k
    total = np.zeros(1 << k, np.complex128)
    tv = total.view(np.float64)
    smap = np.ascontiguousarray(bp.syndrome_map(bp.gather_terminal_basis))
    r = len(bp.gather_terminal_basis)
    g = gather_tables(bp)
    for _, coef, f, yfix, freeflip in bp.slice_assignments(bp.local_vectors(x_bits, w=w), part, parts):
        if f is None:
            continue
        res = contract_gather(bp, f, yfix, storage, tables=g)
        if not res['alive']:
            continue
        c = coef * bp.slice_scale(res['ls'])
        xm = (res['offset'] >> bp.stop) ^ freeflip
        env = np.ascontiguousarray(res['env'])
        lib().oracle_accumulate(int(storage), env.ctypes.data_as(C.c_void_p), C.c_uint64((1 << r) - 1), k,
                                _p(smap, C.c_uint64), C.c_uint64(xm), c.real, c.imag, _p(tv, C.c_double))
    ph = np.ascontiguousarray(bp.phase())
    U = np.ascontiguousarray(bp.open_unitaries().reshape(-1, 4).view(np.float64))
    lib().oracle_transform(k, _p(tv, C.c_double), _p(ph, C.c_double), _p(U, C.c_double))
    return total


def draw(v, u):
    """Primary CDF draw; returns dict(A_index, selected_scaled_weight, scaled_mass_cdf, cdf_lower_numpy,
    cdf_upper_numpy) or raises ArithmeticError."""
    v = np.ascontiguousarray(v, dtype=np.complex128)
    n = len(v)
    cum = np.zeros(n, np.float64)
    out = np.zeros(5, np.float64)
    rc = lib().oracle_draw(_p(v.view(np.float64), C.c_double), n, float(u), _p(cum, C.c_double), _p(out, C.c_double))
    if rc:
        raise ArithmeticError(f'oracle draw failed ({rc})')
    return {'A_index': int(out[0]), 'selected_scaled_weight': out[1], 'scaled_mass_cdf': out[2],
            'cdf_lower_numpy': out[3], 'cdf_upper_numpy': out[4]}


# ---- mathematical ground truth ---------------------------------------------------------------
_GATES = {
    'H': np.array([[1, 1], [1, -1]], np.complex128) / math.sqrt(2),
    'S': np.diag([1, 1j]).astype(np.complex128),
    'T': np.diag([1, np.exp(1j * math.pi / 4)]).astype(np.complex128),
}


def statevector(ops, n):
    """<x|C|0> for every x (bit q of the index is qubit q)."""
    psi = np.zeros((2,) * n, np.complex128)
    psi[(0,) * n] = 1
    for op in ops:
        if op[0] in ('c1', 'T'):
            q = op[1]
            g = _GATES['T'] if op[0] == 'T' else _GATES[op[2]]
            ax = n - 1 - q
            psi = np.moveaxis(np.tensordot(g, psi, axes=([1], [ax])), 0, ax)
        elif op[0] == 'cz':
            a, b = n - 1 - op[1], n - 1 - op[2]
            idx = [slice(None)] * n
            idx[a] = 1
            idx[b] = 1
            psi[tuple(idx)] *= -1
        else:
            raise ValueError(op)
    return psi.reshape(-1)


def graph_amplitude(A, U, bras):
    """Brute-force <v|U|G> = 2^{-N/2} sum_y prod f_k(y_k) (-1)^{sum_E y_a y_b}, f_k = bra_k U_k."""
    N = len(A)
    f = np.array([bras[q] @ U[q] for q in range(N)])
    total = 0j
    iu = np.triu_indices(N, 1)
    for yint in range(1 << N):
        y = np.array([(yint >> q) & 1 for q in range(N)])
        w = np.prod(f[np.arange(N), y])
        if w == 0:
            continue
        s = int((np.outer(y, y)[iu] * A[iu]).sum()) & 1
        total += w * (-1 if s else 1)
    return total * 2 ** (-N / 2)


def gather_amplitude(N, edges, order, fixed, f, yfix):
    """Full-scan amplitude (up to 2^{-N/2}) by the gather executor: returns (value, logscale)
    (scan_plan.py:161-195), using this repository's make_plan for the plan."""
    import sys
    sys.path.insert(0, str(HERE.parent / 'dcs245-classical-sampling_amd'))
    from dcs_amd.scan_plan import byte_tables, future_masks, kernel_combos, make_plan
    steps, _ = make_plan(N, edges, order, fixed)
    fut = future_masks(N, edges, order)
    E = np.ones(1, np.complex128)
    o = 0
    ls = 0.
    for j, st in enumerate(steps):
        f0, f1 = complex(f[st.v][0]), complex(f[st.v][1])
        ov = (o >> j) & 1
        o &= ~(1 << j)
        y = 0
        if not st.free:
            y = int(yfix[st.v])
            if y:
                o ^= fut[j]
        T = byte_tables(st.Lcols, st.r_out)
        cp = np.arange(1 << st.r_out, dtype=np.uint64)
        p0 = np.zeros_like(cp)
        for b in range(T.shape[0]):
            p0 ^= T[b][((cp >> np.uint64(8 * b)) & np.uint64(255)).astype(np.int64)]
        mask = np.uint64((1 << st.r_in) - 1)
        acc = np.zeros(1 << st.r_out, np.complex128)
        for kc in kernel_combos(st.kernel):
            p = p0 ^ np.uint64(kc)
            c = p & mask
            par = np.array([bin(int(v) & st.beta).count('1') & 1 for v in c], np.int64) ^ ov
            val = E[c.astype(np.int64)]
            if st.free:
                yb = ((p >> np.uint64(st.r_in)) & np.uint64(1)).astype(np.int64)
                sg = 1 - 2 * (yb & par)
                acc += val * np.where(yb == 1, f1, f0) * sg
            else:
                sg = 1 - 2 * (y & par)
                acc += val * (f1 if y else f0) * sg
        E = acc
        m = float(np.max(np.abs(E))) if E.size else 0.
        if m > 0:
            E = E / m
            ls += math.log(m)
    return complex(E[0]), ls
