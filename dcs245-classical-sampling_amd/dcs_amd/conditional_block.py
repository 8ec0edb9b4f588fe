"""Coherent, parity-sliced conditional output blocks on one MI355X.

Drop-in for the reference engine classes (sampling/conditional_block_gpu.py:24-177,
sampling/conditional_block_inplace_gpu.py:17-38, kernel_opt/inplace_gpu.py:84-128):

    engine = InplaceConditionalBlockGPU(graph, unitaries, n, order, parity_operations, open_vertices)
    engine.alloc(dtype=np.complex64)            # or np.complex128 for audit blocks
    vector, meta = engine.block(x_bits)         # vector: DeviceBlock (np.asarray(vector) copies it)

ConditionalBlockGPU is the canonical engine of the reference (two tables, the gather executor of
dcs_rankwidth/gpu_kernel.py); InplaceConditionalBlockGPU replaces its prefix with the one-buffer
preserved-coordinate executor, as the reference's subclass does.

The host keeps the reference's bookkeeping (local vectors, slice coefficients, offsets, the log-scale
sum); every pass over a frontier, the slice scatter, the Walsh-Hadamard / phase / Clifford tail and
the CDF draw run in libdcs_amd.so kernels.  There is no CPU fallback: alloc() raises when the library
or the device is missing.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import _native
from .block_plan import BlockPlan

SCHEDULES = {'steps': _native.MODE_STEPS, 'steps_host': _native.MODE_STEPS_HOST, 'fused': _native.MODE_FUSED,
             'compiled': _native.MODE_COMPILED}


def logscale_from_maxima(m2):
    """Replay the reference's host scale chain over the unit maxima of one slice
    (kernel_opt/inplace_gpu.py:108-123): before unit u, e = frexp(sqrt(previous max))[1] and
    ls += e ln 2.  Returns (ls, alive); alive is False when a unit's maximum is exactly zero
    (slice skipped).  A nonfinite maximum raises ArithmeticError."""
    ls = 0.
    maximum = 1.
    for m in m2:
        e = math.frexp(maximum)[1]
        ls += e * math.log(2.)
        m = float(m)
        if not np.isfinite(m):
            raise ArithmeticError('Nonfinite in-place environment')
        if m == 0:
            return ls, False
        maximum = math.sqrt(m)
    return ls, True


class DeviceBlock:
    """The 2^k complex128 block left on the device by block().  np.asarray(), .get() or
    cupy-style asnumpy-by-copy return a host copy; .ptr is the device address."""

    def __init__(self, engine, index):
        self._engine = engine
        self._index = index
        self.shape = (1 << engine.k,)
        self.dtype = np.dtype(np.complex128)
        self.ptr = engine._plan.block_pointer(index)

    def __len__(self):
        return self.shape[0]

    def get(self):
        return self._engine._plan.block_read(self._index)

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a if dtype is None else a.astype(dtype)

    def draw(self, uniform):
        """Device CDF draw (primary numpy semantics of draw_checked) on this block."""
        return self._engine._plan.draw(self._index, uniform)


class ConditionalBlockGPU:
    """Canonical engine (sampling/conditional_block_gpu.py:24-177): the gather executor over two
    tables of 2^rmax entries, one launch per step, the reference's scale chain."""

    def __init__(self, graph, unitaries, n, order, parity_operations, open_vertices):
        self.bp = bp = BlockPlan(graph, unitaries, n, order, parity_operations, open_vertices)
        # attributes the reference exposes
        self.N_full, self.N, self.n = bp.N_full, bp.N, bp.n
        self.open_full, self.open_red = bp.open_full, bp.open_red
        self.slices, self.slice_edges = bp.slices, bp.slice_edges
        self.order, self.fixed, self.fut = bp.order, bp.fixed, bp.fut
        self.steps, self.stop, self.rank_profile = bp.steps, bp.stop, bp.rank_profile
        self.rmax, self.work = bp.rmax, bp.work
        self.terminal_basis, self.terminal_rank = bp.gather_terminal_basis, len(bp.gather_terminal_basis)
        self.open_edges = bp.open_edges
        self.U_full = bp.U_full
        self.k = bp.k
        self.profile = False
        self._plan = None

    def alloc(self, dtype=np.complex64, device=0, fp64_kernel_path=None, schedule='steps', blocks=1):
        """Allocate the two gather tables and upload the plan (dcs_rankwidth/gpu_kernel.py:106-121,
        conditional_block_gpu.py:71-90).

        dtype      np.complex64 or np.complex128 storage (FP64 arithmetic in both)
        schedule   'steps' (scale chain on the device, no host synchronisation per step) or 'steps_host'
                   (the host reads every maximum, as the reference does)
        The fp64_kernel_path argument is accepted for interface compatibility and ignored."""
        dt = np.dtype(dtype)
        if dt not in (np.dtype(np.complex64), np.dtype(np.complex128)):
            raise ValueError(dtype)
        if schedule not in ('steps', 'steps_host'):
            raise ValueError(f'the gather executor runs schedule steps or steps_host, not {schedule!r}')
        bp = self.bp
        self.use_fp64 = dt == np.dtype(np.complex128)
        self.mode = SCHEDULES[schedule]
        self.lanes = 1
        self.terminal_basis = bp.gather_terminal_basis
        self.terminal_rank = len(self.terminal_basis)
        if self.terminal_rank > bp.k:
            raise ValueError(f'terminal rank {self.terminal_rank} exceeds the {bp.k} open vertices')
        self.ctx = _native.Context(device)
        storage = _native.STORAGE_C128 if self.use_fp64 else _native.STORAGE_C64
        self._plan = _native.Plan.gather(self.ctx, bp.steps, bp.rmax, storage, bp.k, self.terminal_rank,
                                         bp.syndrome_map(self.terminal_basis), bp.phase(), bp.open_unitaries(),
                                         blocks=blocks)
        self.compile_seconds = 0.0
        self.buf = True

    # ---- reference mixin hooks (sampling/conditional_block_gpu.py:92-126) -----------------------
    def _prefix(self, f, yfix):
        """(env or None, logscale, offset >> stop) of one slice; env is the host copy of the final
        table in storage precision."""
        coeffs, ov, ys, offset = self.bp.step_parameters(f, yfix)
        m2 = self._plan.contract(self.mode, [0], coeffs, ov, ys)[0]
        ls, alive = logscale_from_maxima(m2)
        if not alive:
            return None, ls, 0
        return self._plan.frontier_terminal(0), ls, offset >> self.stop

    def _sweep(self, f, yfix, renorm_every=None):
        env, ls, _ = self._prefix(f, yfix)
        return (0j if env is None else complex(env[0])), ls

    # ---- the block ---------------------------------------------------------------------------
    def _slice_jobs(self, x_bits, w, part, parts):
        f_full = self.bp.local_vectors(x_bits, w=w)
        jobs, count = [], 0
        for count, coef, f, yfix, freeflip in self.bp.slice_assignments(f_full, part, parts):
            if f is None:
                continue
            coeffs, ov, ys, offset = self.bp.step_parameters(f, yfix)
            jobs.append((coef, coeffs, ov, ys, offset, freeflip))
        return jobs, count

    def block(self, x_bits, w=None, part=0, parts=1, block_index=0):
        """Scaled conditional amplitudes sqrt(D) psi[prefix, :] (conditional_block_gpu.py:128-177).

        Slices are contracted `lanes` at a time; each one's environment is scattered into the
        complex128 accumulator with its coefficient exp(ls - (N_full - n + k) ln2 / 2), then the
        tail is applied once."""
        if self._plan is None:
            raise RuntimeError('call alloc() first')
        start = time.perf_counter()
        plan = self._plan
        jobs, count = self._slice_jobs(x_bits, w, part, parts)
        plan.block_zero(block_index)
        for i in range(0, len(jobs), self.lanes):
            batch = jobs[i:i + self.lanes]
            lanes = np.arange(len(batch), dtype=np.int32)
            m2 = plan.contract(self.mode, lanes, np.stack([j[1] for j in batch]), np.stack([j[2] for j in batch]),
                               np.stack([j[3] for j in batch]))
            for lane, (coef, _, _, _, offset, freeflip) in enumerate(batch):
                ls, alive = logscale_from_maxima(m2[lane])
                if not alive:
                    continue
                c = coef * self.bp.slice_scale(ls)
                plan.block_accumulate(block_index, lane, (offset >> self.stop) ^ freeflip, c)
        plan.block_transform(block_index)
        self.ctx.synchronize()
        vec = DeviceBlock(self, block_index)
        return vec, {'seconds': time.perf_counter() - start, 'part': part, 'parts': parts,
                     'slices_evaluated': count,
                     'prefix_dtype': 'complex128' if self.use_fp64 else 'complex64',
                     'rmax': self.rmax, 'terminal_rank': self.terminal_rank,
                     'work_per_slice': self.work, 'open_vertices': self.open_full,
                     'vector_scale': 'sqrt(2^(n-k)); sum all complex parts before probabilities'}

    def block_many(self, prefixes, w=None, block_indices=None):
        """Blocks of several prefixes at once: the slices of all of them are contracted `lanes` at a time
        (one launch covers slices of different prefixes) and each is scattered into its own complex128
        accumulator.  Every block equals what block() returns for its prefix alone, bit for bit: a slice's
        contraction does not depend on the lane it runs in, and each block receives its slices in the
        same order.  Returns ([DeviceBlock], meta)."""
        if self._plan is None:
            raise RuntimeError('call alloc() first')
        if block_indices is None:
            block_indices = list(range(len(prefixes)))
        if len(block_indices) != len(prefixes) or len(set(block_indices)) != len(block_indices):
            raise ValueError('one distinct block index per prefix')
        if any(not 0 <= b < self._plan.blocks for b in block_indices):
            raise ValueError(f'block indices must lie in 0..{self._plan.blocks - 1}')
        start = time.perf_counter()
        plan = self._plan
        jobs, counts = [], []
        for b, x in zip(block_indices, prefixes):
            js, count = self._slice_jobs(x, w, 0, 1)
            jobs.extend((b,) + j for j in js)
            counts.append(count)
            plan.block_zero(b)
        launches = 0
        for i in range(0, len(jobs), self.lanes):
            batch = jobs[i:i + self.lanes]
            lanes = np.arange(len(batch), dtype=np.int32)
            m2 = plan.contract(self.mode, lanes, np.stack([j[2] for j in batch]), np.stack([j[3] for j in batch]),
                               np.stack([j[4] for j in batch]))
            launches += 1
            for lane, (b, coef, _, _, _, offset, freeflip) in enumerate(batch):
                ls, alive = logscale_from_maxima(m2[lane])
                if not alive:
                    continue
                plan.block_accumulate(b, lane, (offset >> self.stop) ^ freeflip, coef * self.bp.slice_scale(ls))
        for b in block_indices:
            plan.block_transform(b)
        self.ctx.synchronize()
        return [DeviceBlock(self, b) for b in block_indices], {
            'seconds': time.perf_counter() - start, 'slices_evaluated': counts, 'contractions': launches,
            'prefix_dtype': 'complex128' if self.use_fp64 else 'complex64', 'rmax': self.rmax}

    def device_bytes(self):
        """Device memory this engine holds (frontier(s), plan tables, blocks), as sized at alloc."""
        return self._plan.device_bytes() if self._plan is not None else 0

    def schedule(self):
        return self._plan.schedule(self.mode)

    def close(self):
        if self._plan is not None:
            self._plan.close()
            self._plan = None
        if getattr(self, 'ctx', None) is not None:
            if getattr(self, '_own_ctx', True):
                self.ctx.close()
            self.ctx = None


class InplaceConditionalBlockGPU(ConditionalBlockGPU):
    """Reference-compatible in-place engine (conditional_block_inplace_gpu.py:17-38)."""

    _c64_storage = _native.STORAGE_C64   # frontier storage behind dtype=np.complex64

    def alloc(self, dtype=np.complex64, device=0, fp64_kernel_path=None, lanes=1, schedule='compiled',
              fuse_max_dim=None, resident_bits=12, tile_log2=None, blocks=1, exact_rounding=False,
              pass_max_dim=None, lds_storage=True, group_policy=None, lds_buffers=None, graphs=False, context=None,
              factor=True):
        """Compile the in-place plan and allocate the device frontier (kernel_opt/inplace_gpu.py:85-105).

        dtype      np.complex64 (c64 storage, FP64 arithmetic) or np.complex128
        lanes      frontiers held at once; slices of one block are contracted `lanes` at a time
        schedule   'compiled' (resident runs + coset groups, one generated kernel per group, compiled
                   here with hipRTC), 'fused' (the same schedule on the generic coset kernel), 'steps'
                   (one launch per step, scale chain on the device) or 'steps_host' (one launch per
                   step, host reads every maximum)
        exact_rounding  fused schedules: round to storage precision after every step (mantissas equal
                   the per-step schedule) instead of once per register pass of up to three steps
        lds_storage  compiled schedule: hold coset tiles in LDS in storage precision (half the LDS of
                   FP64 tiles for complex64, so twice the workgroups per CU); values are then rounded
                   to storage once per register pass
        fuse_max_dim, tile_log2, pass_max_dim  largest coset dimension of a fused group, log2 entries
                   of its LDS tile (11, 12; 13 compiled only), span of one register pass (3, 4); in
                   the compiled schedule a group of dimension 9 keeps rows of 16 entries on a tile of
                   2^13 entries when tile_log2 is smaller.  The defaults are the fastest measured on
                   the 64-qubit plans: (9, 12, 4) for the f0-factored complex64 kernels (groups of
                   dimension 9 on 2^13 tiles, the others on 2^12; a group kernel that spills registers
                   with span-4 passes runs its span-3 version), (9, 13, 3) for complex128 in the
                   compiled schedule when neither of the first two is given, else (8, 12, 3)
        group_policy  0: greedy maximal coset groups; 1: the group boundaries with the fewest HBM
                   bytes (dynamic programming); 2: bytes and register-pass work summed; 3: the same two
                   costs overlapped, max plus a quarter of the smaller (both calibrated on complex64
                   tiles, the register passes priced lower for f0-factored kernels).  Default: 3 for the
                   f0-factored complex64 kernels, else 1 (the fastest measured on the 64-qubit plans)
        lds_buffers  compiled schedule: 1 one LDS tile per workgroup, the next tile waiting in
                   registers; 2 double-buffered tiles; 3 two tiles, the next one loaded straight from
                   HBM into LDS (global_load_lds) during the current tile's passes and the finished
                   tile stored in 16-byte pairs, where a group allows it (complex64 tiles in LDS, at
                   least two lanes), else as 1; the same arithmetic as 1, so the same blocks bit for
                   bit; 4 as 3 with the loads issued a whole tile ahead and
                   the last register pass storing to HBM from registers (the group's scale folded into
                   a step); 5 as 4 with the live inputs of the next two tiles in flight, in a ring of
                   compact LDS buffers, for groups whose input is a half or a quarter live (the same
                   blocks as 4 bit for bit).  Default: 3 (the fastest measured on the 64-qubit plans)
        factor     compiled schedule without per-step rounding (complex128, or tiles rounded once per
                   register pass): run every step whose f0 cannot be zero with the coefficients (1, f1 / f0) and
                   apply the pending f0s at each group's last step (BlockPlan.coefficient_classes says
                   whether f1 / f0 is real, imaginary or complex, which sets the FP64 work of the step);
                   a slice whose parameters do not fit runs the unfactored kernels; an array gives the
                   per-step classes instead of the plan's own
        graphs     replay each slice's launch sequence from a captured HIP graph (one per lane count)
        context    a _native.Context to share (engines contracted together by block_batch must share
                   one); the engine then leaves it open on close().  Default: a context of its own
        The legacy fp64_kernel_path argument is accepted and ignored, as in the reference."""
        dt = np.dtype(dtype)
        if dt not in (np.dtype(np.complex64), np.dtype(np.complex128)):
            raise ValueError(dtype)
        if schedule not in SCHEDULES:
            raise ValueError(f'unknown schedule {schedule!r}')
        bp = self.bp
        self.inplace_fp64 = self.use_fp64 = dt == np.dtype(np.complex128)
        self.inplace_steps, self.inplace_maxslots, self.inplace_rmax = bp.inplace_steps, bp.inplace_maxslots, bp.inplace_rmax
        self.terminal_basis, self.terminal_rank = bp.terminal_basis, bp.terminal_rank
        if bp.terminal_rank > bp.k:
            raise ValueError(f'terminal rank {bp.terminal_rank} exceeds the {bp.k} open vertices')
        self.mode = SCHEDULES[schedule]
        factored = (self.mode == _native.MODE_COMPILED and not self.use_fp64 and self._c64_storage == _native.STORAGE_C64
                    and lds_storage and not exact_rounding and not (factor is False or factor is None))
        if fuse_max_dim is None and tile_log2 is None and self.mode == _native.MODE_COMPILED and self.use_fp64:
            fuse_max_dim, tile_log2 = 9, 13
        fuse_max_dim = (9 if factored else 8) if fuse_max_dim is None else fuse_max_dim
        tile_log2 = 12 if tile_log2 is None else tile_log2
        pass_max_dim = (4 if factored else 3) if pass_max_dim is None else pass_max_dim
        if group_policy is None:
            group_policy = 3 if factored else 1
        if lds_buffers is None:
            lds_buffers = 3
        self.lanes = int(lanes)
        self._own_ctx = context is None
        self.ctx = _native.Context(device) if context is None else context
        storage = _native.STORAGE_C128 if self.use_fp64 else self._c64_storage
        self._plan = _native.Plan(self.ctx, bp.step_table(), bp.inplace_maxslots, storage, self.lanes, bp.k,
                                  bp.terminal_mask(), bp.syndrome_map(), bp.phase(), bp.open_unitaries(),
                                  blocks=blocks, fuse_max_dim=fuse_max_dim, resident_bits=resident_bits,
                                  tile_log2=tile_log2, exact_rounding=exact_rounding,
                                  pass_max_dim=pass_max_dim, lds_storage=lds_storage,
                                  group_policy=group_policy, lds_buffers=lds_buffers,
                                  coef_class=self._coef_classes(factor))
        self.schedule_options = {'schedule': schedule, 'fuse_max_dim': fuse_max_dim, 'tile_log2': tile_log2,
                                 'pass_max_dim': pass_max_dim, 'resident_bits': resident_bits,
                                 'group_policy': group_policy, 'lds_buffers': lds_buffers,
                                 'lds_storage': bool(lds_storage), 'exact_rounding': bool(exact_rounding),
                                 'factor': not (factor is False or factor is None)}
        self.compile_seconds = self._plan.compile() if self.mode == _native.MODE_COMPILED else 0.0
        if graphs:
            self._plan.use_graphs(True)
        self.schedule_options['graphs'] = bool(graphs)
        self.buf = True

    def _coef_classes(self, factor):
        if factor is False or factor is None:
            return None
        if factor is True:
            return self.bp.coefficient_classes()
        return np.asarray(factor, dtype=np.int32)

    def factoring(self):
        """(group kernels f0-factored, the last contraction ran them): see alloc(factor=...)."""
        return self._plan.factoring()

    # ---- reference mixin hooks -------------------------------------------------------------
    def prefix_inplace(self, f, yfix, lane=0):
        """(env or None, logscale, offset) of one slice (kernel_opt/inplace_gpu.py:107-124).
        env is the host copy of the terminal environment in storage precision."""
        coeffs, ov, ys, offset = self.bp.step_parameters(f, yfix)
        m2 = self._plan.contract(self.mode, [lane], coeffs, ov, ys)[0]
        ls, alive = logscale_from_maxima(m2)
        if not alive:
            return None, ls, offset
        return self._plan.frontier_terminal(lane), ls, offset

    def _prefix(self, f, yfix):
        env, ls, offset = self.prefix_inplace(f, yfix)
        return env, ls, offset >> self.stop

    def _sweep(self, f, yfix, renorm_every=None):
        env, ls, _ = self.prefix_inplace(f, yfix)
        return (0j if env is None else complex(env[0])), ls


class HalfConditionalBlockGPU(InplaceConditionalBlockGPU):
    """Binary16 storage of the in-place frontier, FP64 arithmetic (sampling/conditional_block_half_gpu.py:8-34,
    kernel_opt/inplace_half_gpu.py:14-78): each entry is two independently rounded IEEE float16 parts
    (4 bytes), stored through float with round to nearest; maxima and the scale chain use the stored
    values; the coherent slice sum stays complex128.  An approximate variant: whole-vector quality needs
    an FP64 check (the reference's own scope note)."""

    _c64_storage = _native.STORAGE_F16X2

    def alloc(self, dtype=np.complex64, device=0, fp64_kernel_path=None, group_policy=2, **kw):
        """As InplaceConditionalBlockGPU.alloc; group boundaries default to policy 2 (bytes and register
        passes), which measured fastest with 4-byte entries on the 64-qubit half_k1 plan."""
        if np.dtype(dtype) != np.dtype(np.complex64):
            raise ValueError('Half storage exports complex64; use baseline executor for complex128')
        super().alloc(dtype=np.complex64, device=device, fp64_kernel_path=fp64_kernel_path, group_policy=group_policy,
                      **kw)
        self.inplace_storage = 'float16-components'

    def block(self, *args, **kwargs):
        vector, meta = super().block(*args, **kwargs)
        meta.update({'intermediate_storage': 'float16 real/imaginary components',
                     'local_arithmetic': 'FP64 then float32-to-float16 rounded storage',
                     'coherent_accumulation': 'complex128',
                     'approximation_scope': 'Whole-vector quality requires independent FP64 validation'})
        return vector, meta


def block_batch(requests, w=None):
    """Blocks of prefixes of several circuits at once, the circuits' contractions sharing launches.

    requests  [(engine, x_bits, block_index)]: InplaceConditionalBlockGPU engines of different plans
              (different frontier widths allowed), allocated with schedule='fused' on one shared context
              (alloc(context=...)) with the default tile shapes; an engine may appear more than once,
              with distinct block indices.
    Each round contracts up to `lanes` pending slices of every engine in one dcs_contract_batch call, whose
    launches each run the same schedule segment of all engines together; every slice is then scattered
    into its request's block in the order block() uses, and each block gets its tail.  Every block equals
    engine.block(x_bits, block_index=b) alone, bit for bit.  Returns ([DeviceBlock], meta)."""
    if not requests:
        return [], {'seconds': 0.0, 'launches': 0, 'rounds': 0}
    engines = []
    for e, _, _ in requests:
        if e._plan is None:
            raise RuntimeError('call alloc() first')
        if e.mode != _native.MODE_FUSED:
            raise ValueError('block_batch runs engines allocated with schedule="fused"')
        if all(e is not f for f in engines):
            engines.append(e)
    ctx = engines[0].ctx
    if any(e.ctx is not ctx for e in engines):
        raise ValueError('engines contracted together must share one context: alloc(context=...)')
    if len(engines) > _native.BATCH_MAX:
        raise ValueError(f'at most {_native.BATCH_MAX} engines per batch')
    seen = set()
    for e, _, b in requests:
        if (id(e), b) in seen:
            raise ValueError(f'block index {b} requested twice for one engine')
        if not 0 <= b < e._plan.blocks:
            raise ValueError(f'block index {b} outside 0..{e._plan.blocks - 1}')
        seen.add((id(e), b))
    start = time.perf_counter()
    pending = {id(e): [] for e in engines}
    counts = []
    for e, x, b in requests:
        jobs, count = e._slice_jobs(x, w, 0, 1)
        pending[id(e)].extend((b,) + j for j in jobs)
        counts.append(count)
        e._plan.block_zero(b)
    launches = rounds = 0
    while any(pending.values()):
        taken, items = [], []
        for e in engines:
            jobs = pending[id(e)][:e.lanes]
            if not jobs:
                continue
            pending[id(e)] = pending[id(e)][e.lanes:]
            taken.append((e, jobs))
            items.append((e._plan, np.arange(len(jobs), dtype=np.int32), np.stack([j[2] for j in jobs]),
                          np.stack([j[3] for j in jobs]), np.stack([j[4] for j in jobs])))
        m2s, n = _native.contract_batch(items)
        launches += n
        rounds += 1
        for (e, jobs), m2 in zip(taken, m2s):
            for lane, (b, coef, _, _, _, offset, freeflip) in enumerate(jobs):
                ls, alive = logscale_from_maxima(m2[lane])
                if not alive:
                    continue
                e._plan.block_accumulate(b, lane, (offset >> e.stop) ^ freeflip, coef * e.bp.slice_scale(ls))
    for e, _, b in requests:
        e._plan.block_transform(b)
    ctx.synchronize()
    return [DeviceBlock(e, b) for e, _, b in requests], {
        'seconds': time.perf_counter() - start, 'slices_evaluated': counts, 'launches': launches,
        'rounds': rounds}
