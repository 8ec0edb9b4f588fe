"""ctypes binding of libdcs_amd.so (C ABI declared in include/dcs_amd.h).

The shared library is built in place next to this file by csrc/build.sh.  There is no fallback:
importing the engine without the library raises, so a missing build can never be mistaken for a
result."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / 'libdcs_amd.so'

DCS_OK, DCS_ERR_INVALID, DCS_ERR_RUNTIME, DCS_ERR_NOMEM, DCS_ERR_NONFINITE, DCS_ERR_ZERO, DCS_ERR_UNSUPPORTED = range(7)
DRAW_NOALLOC_MAX_K = 22   # DCS_DRAW_NOALLOC_MAX_K
STORAGE_C64, STORAGE_C128, STORAGE_F16X2 = 0, 1, 2
MODE_STEPS, MODE_STEPS_HOST, MODE_FUSED, MODE_COMPILED = 0, 1, 2, 3

EXPORTED = (
    'dcs_last_error', 'dcs_get_device_count', 'dcs_get_device_info', 'dcs_context_create',
    'dcs_context_destroy', 'dcs_synchronize', 'dcs_plan_create', 'dcs_gather_plan_create', 'dcs_plan_destroy',
    'dcs_plan_compile', 'dcs_plan_factoring', 'dcs_plan_kernel_source', 'dcs_plan_memory', 'dcs_plan_use_graphs',
    'dcs_plan_units',
    'dcs_plan_schedule', 'dcs_contract', 'dcs_block_zero', 'dcs_block_accumulate', 'dcs_block_transform',
    'dcs_block_read', 'dcs_block_device_pointer', 'dcs_draw', 'dcs_frontier_read', 'dcs_plan_profile',
    'dcs_stream_copy', 'dcs_draw_noalloc', 'dcs_contract_batch')
BATCH_MAX = 16   # DCS_BATCH_MAX


class DeviceInfo(C.Structure):
    _fields_ = [('name', C.c_char * 256), ('arch', C.c_char * 64), ('total_bytes', C.c_uint64),
                ('free_bytes', C.c_uint64), ('compute_units', C.c_int32), ('runtime_version', C.c_int32),
                ('driver_version', C.c_int32), ('device', C.c_int32)]


class PlanDesc(C.Structure):
    _fields_ = [('storage', C.c_int32), ('maxslots', C.c_int32), ('lanes', C.c_int32), ('nsteps', C.c_int32),
                ('kind', C.POINTER(C.c_int32)), ('K', C.POINTER(C.c_uint64)), ('Ka', C.POINTER(C.c_uint64)),
                ('pivot', C.POINTER(C.c_int32)), ('pivot_a', C.POINTER(C.c_int32)),
                ('beta', C.POINTER(C.c_uint64)), ('in_mask', C.POINTER(C.c_uint64)),
                ('out_mask', C.POINTER(C.c_uint64)), ('k', C.c_int32), ('terminal_mask', C.c_uint64),
                ('syndrome_map', C.POINTER(C.c_uint64)), ('phase', C.POINTER(C.c_double)),
                ('open_unitaries', C.POINTER(C.c_double)), ('blocks', C.c_int32),
                ('fuse_max_dim', C.c_int32), ('resident_bits', C.c_int32), ('tile_log2', C.c_int32),
                ('exact_rounding', C.c_int32), ('pass_max_dim', C.c_int32), ('lds_storage', C.c_int32),
                ('group_policy', C.c_int32), ('lds_buffers', C.c_int32), ('coef_class', C.POINTER(C.c_int32))]


class BatchItem(C.Structure):
    _fields_ = [('plan', C.c_void_p), ('nlanes', C.c_int), ('lanes', C.POINTER(C.c_int32)),
                ('coeffs', C.POINTER(C.c_double)), ('ov', C.POINTER(C.c_int32)), ('y', C.POINTER(C.c_int32)),
                ('unit_m2', C.POINTER(C.c_double))]


class GatherDesc(C.Structure):
    P = C.POINTER
    _fields_ = [('storage', C.c_int32), ('rmax', C.c_int32), ('nsteps', C.c_int32),
                ('r_in', P(C.c_int32)), ('r_out', P(C.c_int32)), ('is_free', P(C.c_int32)), ('nbytes', P(C.c_int32)),
                ('ncombos', P(C.c_int32)), ('tab_off', P(C.c_int64)), ('comb_off', P(C.c_int64)),
                ('tables', P(C.c_uint64)), ('combos', P(C.c_uint64)), ('beta', P(C.c_uint64)), ('k', C.c_int32),
                ('terminal_rank', C.c_int32), ('syndrome_map', P(C.c_uint64)), ('phase', P(C.c_double)),
                ('open_unitaries', P(C.c_double)), ('blocks', C.c_int32)]


class DrawResult(C.Structure):
    _fields_ = [('index', C.c_int64), ('weight', C.c_double), ('total', C.c_double),
                ('cdf_lower', C.c_double), ('cdf_upper', C.c_double), ('margin', C.c_double),
                ('mass_check', C.c_double), ('cdf_lower_check', C.c_double), ('cdf_upper_check', C.c_double)]


class ScheduleUnsupported(ValueError):
    """The requested schedule cannot be built on this device (DCS_ERR_UNSUPPORTED): the generated group
    kernels do not compile, or one of them cannot be resident.  A refusal, not a device failure."""


_lib = None


def load():
    """Load the library once; raise ImportError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f'{LIB_PATH} is missing: build it with csrc/build.sh (or __graft_entry__.build())')
    lib = C.CDLL(str(LIB_PATH))
    vp, i32, u64, dbl = C.c_void_p, C.c_int32, C.c_uint64, C.c_double
    P = C.POINTER
    sig = {
        'dcs_last_error': (C.c_char_p, []),
        'dcs_get_device_count': (C.c_int, [P(C.c_int)]),
        'dcs_get_device_info': (C.c_int, [C.c_int, P(DeviceInfo)]),
        'dcs_context_create': (C.c_int, [C.c_int, P(vp)]),
        'dcs_context_destroy': (None, [vp]),
        'dcs_synchronize': (C.c_int, [vp]),
        'dcs_plan_create': (C.c_int, [vp, P(PlanDesc), P(vp)]),
        'dcs_gather_plan_create': (C.c_int, [vp, P(GatherDesc), P(vp)]),
        'dcs_plan_destroy': (None, [vp]),
        'dcs_plan_compile': (C.c_int, [vp, P(dbl)]),
        'dcs_plan_factoring': (C.c_int, [vp, P(i32), P(i32)]),
        'dcs_plan_kernel_source': (C.c_int, [P(PlanDesc), C.c_char_p, u64, P(u64), P(i32), P(i32), P(i32), P(i32),
                                             C.c_int, P(C.c_int)]),
        'dcs_plan_memory': (C.c_int, [vp, P(u64)]),
        'dcs_plan_units': (C.c_int, [vp, C.c_int, P(C.c_int)]),
        'dcs_plan_schedule': (C.c_int, [vp, C.c_int, C.c_int, P(i32), P(i32), P(i32), P(i32), P(C.c_int)]),
        'dcs_contract': (C.c_int, [vp, C.c_int, C.c_int, P(i32), P(dbl), P(i32), P(i32), P(dbl)]),
        'dcs_block_zero': (C.c_int, [vp, C.c_int]),
        'dcs_block_accumulate': (C.c_int, [vp, C.c_int, C.c_int, u64, dbl, dbl]),
        'dcs_block_transform': (C.c_int, [vp, C.c_int]),
        'dcs_block_read': (C.c_int, [vp, C.c_int, P(dbl)]),
        'dcs_block_device_pointer': (C.c_int, [vp, C.c_int, P(vp)]),
        'dcs_draw': (C.c_int, [vp, C.c_int, dbl, P(DrawResult)]),
        'dcs_frontier_read': (C.c_int, [vp, C.c_int, vp]),
        'dcs_plan_profile': (C.c_int, [vp, C.c_int, P(dbl), P(C.c_int64)]),
        'dcs_plan_use_graphs': (C.c_int, [vp, C.c_int]),
        'dcs_stream_copy': (C.c_int, [vp, u64, C.c_int, P(dbl)]),
        'dcs_draw_noalloc': (C.c_int, [vp, vp, C.c_int, dbl, P(DrawResult)]),
        'dcs_contract_batch': (C.c_int, [C.c_int, P(BatchItem), P(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    """Translate a dcs_status into the reference's exception types."""
    if rc == DCS_OK:
        return
    msg = load().dcs_last_error().decode(errors='replace')
    if rc == DCS_ERR_INVALID:
        raise ValueError(msg)
    if rc == DCS_ERR_NOMEM:
        raise MemoryError(msg)
    if rc == DCS_ERR_UNSUPPORTED:
        raise ScheduleUnsupported(msg)
    if rc in (DCS_ERR_NONFINITE, DCS_ERR_ZERO):
        raise ArithmeticError(msg)
    raise RuntimeError(msg)


def draw_dict(r):
    """dcs_draw_result as the fields of draw_checked it covers: the primary CDF (numpy's cumsum order) and
    the independent interval, here summed with compensation on the device instead of math.fsum."""
    return {'A_index': int(r.index), 'selected_scaled_weight': r.weight, 'scaled_mass_cdf': r.total,
            'cdf_lower_numpy': r.cdf_lower, 'cdf_upper_numpy': r.cdf_upper, 'cdf_margin': r.margin,
            'scaled_mass_compensated': r.mass_check, 'cdf_lower_compensated': r.cdf_lower_check,
            'cdf_upper_compensated': r.cdf_upper_check, 'independent_cdf_check_passed': True}


def contract_batch(items):
    """dcs_contract_batch: items = [(plan, lanes, coeffs, ov, y)] of distinct plans on one context, laid out
    as for Plan.contract.  Runs their fused schedules together (each round's resident runs and coset
    groups in one launch each); returns ([float64[nlanes, nunits] unit maxima per item], launches)."""
    if not items:
        return [], 0
    arr = (BatchItem * len(items))()
    keep, outs = [], []
    for b, (plan, lanes, coeffs, ov, y) in zip(arr, items):
        lanes = np.ascontiguousarray(lanes, dtype=np.int32)
        nl = len(lanes)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(nl, plan.nsteps, 4)
        ov = np.ascontiguousarray(ov, dtype=np.int32).reshape(nl, plan.nsteps)
        y = np.ascontiguousarray(y, dtype=np.int32).reshape(nl, plan.nsteps)
        out = np.zeros((nl, plan.units(MODE_FUSED)), np.float64)
        keep.append((lanes, coeffs, ov, y))
        outs.append(out)
        b.plan, b.nlanes = plan.handle.value, nl
        b.lanes, b.coeffs = ptr(lanes, C.c_int32), ptr(coeffs, C.c_double)
        b.ov, b.y, b.unit_m2 = ptr(ov, C.c_int32), ptr(y, C.c_int32), ptr(out, C.c_double)
    launches = C.c_int(0)
    check(load().dcs_contract_batch(len(items), arr, C.byref(launches)))
    return outs, launches.value


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def device_count():
    n = C.c_int(0)
    check(load().dcs_get_device_count(C.byref(n)))
    return n.value


def device_info(device=0):
    info = DeviceInfo()
    check(load().dcs_get_device_info(device, C.byref(info)))
    return {'device_name': info.name.decode(), 'arch': info.arch.decode(), 'total_bytes': int(info.total_bytes),
            'free_bytes': int(info.free_bytes), 'compute_units': int(info.compute_units),
            'hip_runtime_version': int(info.runtime_version), 'hip_driver_version': int(info.driver_version),
            'device': int(info.device)}


class Context:
    """One device and one HIP stream (default priority)."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.dcs_context_create(int(device), C.byref(h)))
        self.handle = h
        self.device = int(device)

    def synchronize(self):
        check(self.lib.dcs_synchronize(self.handle))

    def stream_copy_ms(self, nbytes, reps=5):
        ms = C.c_double()
        check(self.lib.dcs_stream_copy(self.handle, C.c_uint64(nbytes), int(reps), C.byref(ms)))
        return ms.value

    def draw_noalloc(self, block_ptr, k, u):
        """The CDF draw of dcs_draw on a complex128 device block of 2^k entries at block_ptr, through the
        entry point that allocates nothing (one static scratch per device; calls on one device must not
        overlap)."""
        r = DrawResult()
        check(self.lib.dcs_draw_noalloc(self.handle, C.c_void_p(int(block_ptr)), int(k), float(u), C.byref(r)))
        return draw_dict(r)

    def close(self):
        if self.handle:
            self.lib.dcs_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plan_desc(table, maxslots, storage, lanes, k, terminal_mask, syndrome_map, phase, open_unitaries, blocks=1,
              fuse_max_dim=6, resident_bits=12, tile_log2=11, exact_rounding=False, pass_max_dim=3, lds_storage=False,
              group_policy=0, lds_buffers=2, coef_class=None):
    """dcs_plan_desc of an in-place step table and the arrays its pointers refer to (keep them alive).
    coef_class: per-step class of f1 / f0 for the f0-factored group kernels (None: off)."""
    n = len(table['kind'])
    keep = {}

    def arr(name, dt):
        a = np.ascontiguousarray(table[name], dtype=dt)
        keep[name] = a
        return a
    d = PlanDesc()
    d.storage, d.maxslots, d.lanes, d.nsteps = int(storage), int(maxslots), int(lanes), n
    d.kind = ptr(arr('kind', np.int32), C.c_int32)
    d.K = ptr(arr('K', np.uint64), C.c_uint64)
    d.Ka = ptr(arr('Ka', np.uint64), C.c_uint64)
    d.pivot = ptr(arr('pivot', np.int32), C.c_int32)
    d.pivot_a = ptr(arr('pivot_a', np.int32), C.c_int32)
    d.beta = ptr(arr('beta', np.uint64), C.c_uint64)
    d.in_mask = ptr(arr('in_mask', np.uint64), C.c_uint64)
    d.out_mask = ptr(arr('out_mask', np.uint64), C.c_uint64)
    keep['smap'] = np.ascontiguousarray(syndrome_map, dtype=np.uint64)
    keep['phase'] = np.ascontiguousarray(phase, dtype=np.float64)
    keep['U'] = np.ascontiguousarray(np.asarray(open_unitaries, dtype=np.complex128).reshape(-1, 4).view(np.float64))
    d.k, d.terminal_mask = int(k), C.c_uint64(int(terminal_mask))
    d.syndrome_map = ptr(keep['smap'], C.c_uint64)
    d.phase = ptr(keep['phase'], C.c_double)
    d.open_unitaries = ptr(keep['U'], C.c_double)
    d.blocks, d.fuse_max_dim, d.resident_bits, d.tile_log2 = int(blocks), int(fuse_max_dim), int(resident_bits), int(tile_log2)
    d.exact_rounding = int(bool(exact_rounding))
    d.pass_max_dim = int(pass_max_dim)
    d.lds_storage = int(bool(lds_storage))
    d.group_policy, d.lds_buffers = int(group_policy), int(lds_buffers)
    if coef_class is not None:
        keep['coef_class'] = np.ascontiguousarray(coef_class, dtype=np.int32)
        if keep['coef_class'].shape != (n,):
            raise ValueError(f'coef_class needs one entry per step ({n})')
        d.coef_class = ptr(keep['coef_class'], C.c_int32)
    return d, keep


def kernel_source(table, maxslots, storage, terminal_mask, k, **options):
    """Generated group-kernel source of a step table and its group segments, without a device:
    (source, [(kernel index, first step, step count, workgroup threads), ...]).  options: the
    schedule options of Plan (fuse_max_dim, resident_bits, tile_log2, exact_rounding, pass_max_dim,
    lds_storage, group_policy, lds_buffers)."""
    lib = load()
    kk = int(k)
    d, keep = plan_desc(table, maxslots, storage, 1, kk, terminal_mask, np.zeros(1, np.uint64), np.zeros(1),
                        np.tile(np.eye(2, dtype=np.complex128), (max(kk, 1), 1, 1)), **options)
    n, length = C.c_int(), C.c_uint64()
    check(lib.dcs_plan_kernel_source(C.byref(d), None, 0, C.byref(length), None, None, None, None, 0, C.byref(n)))
    buf = C.create_string_buffer(length.value + 1)
    seg = [np.zeros(max(1, n.value), np.int32) for _ in range(4)]
    check(lib.dcs_plan_kernel_source(C.byref(d), buf, length.value + 1, C.byref(length),
                                     *[ptr(a, C.c_int32) for a in seg], n.value, C.byref(n)))
    return buf.value.decode(), [tuple(int(a[i]) for a in seg) for i in range(n.value)]


class Plan:
    """Device-resident plan: frontier buffer(s), step tables, syndrome map, phase, block accumulators."""

    def __init__(self, ctx, table, maxslots, storage, lanes, k, terminal_mask, syndrome_map, phase,
                 open_unitaries, blocks=1, fuse_max_dim=6, resident_bits=12, tile_log2=11, exact_rounding=False,
                 pass_max_dim=3, lds_storage=False, group_policy=0, lds_buffers=2, coef_class=None):
        self.ctx, self.lib = ctx, ctx.lib
        n = len(table['kind'])
        d, self._keep = plan_desc(table, maxslots, storage, lanes, k, terminal_mask, syndrome_map, phase,
                                  open_unitaries, blocks, fuse_max_dim, resident_bits, tile_log2, exact_rounding,
                                  pass_max_dim, lds_storage, group_policy, lds_buffers, coef_class)
        h = C.c_void_p()
        check(self.lib.dcs_plan_create(ctx.handle, C.byref(d), C.byref(h)))
        self.handle = h
        self.nsteps, self.k, self.lanes, self.blocks = n, int(k), int(lanes), int(blocks)
        self.storage = int(storage)
        self.terminal_rank = bin(int(terminal_mask)).count('1')

    @classmethod
    def gather(cls, ctx, steps, rmax, storage, k, terminal_rank, syndrome_map, phase, open_unitaries, blocks=1):
        """Plan of the canonical gather executor.  steps: make_plan steps (r_in, r_out, free, beta,
        Lcols, kernel); their byte tables and kernel combinations are built here."""
        from .scan_plan import byte_tables, kernel_combos
        self = cls.__new__(cls)
        self.ctx, self.lib = ctx, ctx.lib
        m = len(steps)
        r_in = np.array([s.r_in for s in steps], np.int32)
        r_out = np.array([s.r_out for s in steps], np.int32)
        free = np.array([1 if s.free else 0 for s in steps], np.int32)
        beta = np.array([s.beta for s in steps], np.uint64)
        tabs = [byte_tables(s.Lcols, s.r_out) for s in steps]
        combs = [np.asarray(kernel_combos(s.kernel), np.uint64) for s in steps]
        nbytes = np.array([t.shape[0] for t in tabs], np.int32)
        ncomb = np.array([len(c) for c in combs], np.int32)
        tab_off = np.concatenate([[0], np.cumsum([t.size for t in tabs])[:-1]]).astype(np.int64)
        comb_off = np.concatenate([[0], np.cumsum(ncomb)[:-1]]).astype(np.int64)
        tables = np.ascontiguousarray(np.concatenate([t.reshape(-1) for t in tabs]), np.uint64)
        combos = np.ascontiguousarray(np.concatenate(combs), np.uint64)
        smap = np.ascontiguousarray(syndrome_map, dtype=np.uint64)
        ph = np.ascontiguousarray(phase, dtype=np.float64)
        U = np.ascontiguousarray(np.asarray(open_unitaries, dtype=np.complex128).reshape(-1, 4).view(np.float64))
        keep = (r_in, r_out, free, beta, nbytes, ncomb, tab_off, comb_off, tables, combos, smap, ph, U)
        d = GatherDesc()
        d.storage, d.rmax, d.nsteps = int(storage), int(rmax), m
        d.r_in, d.r_out, d.is_free = ptr(r_in, C.c_int32), ptr(r_out, C.c_int32), ptr(free, C.c_int32)
        d.nbytes, d.ncombos = ptr(nbytes, C.c_int32), ptr(ncomb, C.c_int32)
        d.tab_off, d.comb_off = ptr(tab_off, C.c_int64), ptr(comb_off, C.c_int64)
        d.tables, d.combos, d.beta = ptr(tables, C.c_uint64), ptr(combos, C.c_uint64), ptr(beta, C.c_uint64)
        d.k, d.terminal_rank = int(k), int(terminal_rank)
        d.syndrome_map, d.phase, d.open_unitaries = ptr(smap, C.c_uint64), ptr(ph, C.c_double), ptr(U, C.c_double)
        d.blocks = int(blocks)
        h = C.c_void_p()
        check(self.lib.dcs_gather_plan_create(ctx.handle, C.byref(d), C.byref(h)))
        del keep
        self.handle = h
        self.nsteps, self.k, self.lanes, self.blocks = m, int(k), 1, int(blocks)
        self.storage = int(storage)
        self.terminal_rank = int(terminal_rank)
        return self

    def compile(self):
        """Generate and compile the plan-specialised group kernels; returns the compile seconds
        (0.0 when the module was already cached in this process)."""
        sec = C.c_double(0.0)
        check(self.lib.dcs_plan_compile(self.handle, C.byref(sec)))
        return sec.value

    def factoring(self):
        """(module factored, last compiled contraction ran the factored kernels) as booleans."""
        a, b = C.c_int32(0), C.c_int32(0)
        check(self.lib.dcs_plan_factoring(self.handle, C.byref(a), C.byref(b)))
        return bool(a.value), bool(b.value)

    def use_graphs(self, enable=True):
        """Replay contract()'s launch sequence from captured HIP graphs (one per mode and lane count)."""
        check(self.lib.dcs_plan_use_graphs(self.handle, int(bool(enable))))

    def device_bytes(self):
        b = C.c_uint64(0)
        check(self.lib.dcs_plan_memory(self.handle, C.byref(b)))
        return int(b.value)

    def units(self, mode):
        u = C.c_int()
        check(self.lib.dcs_plan_units(self.handle, int(mode), C.byref(u)))
        return u.value

    def schedule(self, mode):
        cap = self.nsteps + 1
        kinds, first, count, dim = (np.zeros(cap, np.int32) for _ in range(4))
        nseg = C.c_int()
        check(self.lib.dcs_plan_schedule(self.handle, int(mode), cap, ptr(kinds, C.c_int32), ptr(first, C.c_int32),
                                         ptr(count, C.c_int32), ptr(dim, C.c_int32), C.byref(nseg)))
        m = nseg.value
        return [(int(kinds[i]), int(first[i]), int(count[i]), int(dim[i])) for i in range(m)]

    def contract(self, mode, lanes, coeffs, ov, y):
        """Run every step for the given lanes; returns float64[nlanes, nunits] of unit maxima."""
        lanes = np.ascontiguousarray(lanes, dtype=np.int32)
        nl = len(lanes)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(nl, self.nsteps, 4)
        ov = np.ascontiguousarray(ov, dtype=np.int32).reshape(nl, self.nsteps)
        y = np.ascontiguousarray(y, dtype=np.int32).reshape(nl, self.nsteps)
        out = np.zeros((nl, self.units(mode)), np.float64)
        check(self.lib.dcs_contract(self.handle, int(mode), nl, ptr(lanes, C.c_int32), ptr(coeffs, C.c_double),
                                    ptr(ov, C.c_int32), ptr(y, C.c_int32), ptr(out, C.c_double)))
        return out

    def block_zero(self, block=0):
        check(self.lib.dcs_block_zero(self.handle, int(block)))

    def block_accumulate(self, block, lane, xmask, coef):
        coef = complex(coef)
        check(self.lib.dcs_block_accumulate(self.handle, int(block), int(lane), C.c_uint64(int(xmask)),
                                            coef.real, coef.imag))

    def block_transform(self, block=0):
        check(self.lib.dcs_block_transform(self.handle, int(block)))

    def block_read(self, block=0):
        out = np.empty(1 << self.k, np.complex128)
        check(self.lib.dcs_block_read(self.handle, int(block), ptr(out.view(np.float64), C.c_double)))
        return out

    def block_pointer(self, block=0):
        p = C.c_void_p()
        check(self.lib.dcs_block_device_pointer(self.handle, int(block), C.byref(p)))
        return p.value

    def draw(self, block, u):
        r = DrawResult()
        check(self.lib.dcs_draw(self.handle, int(block), float(u), C.byref(r)))
        return draw_dict(r)

    def frontier_terminal(self, lane=0):
        if self.storage == STORAGE_F16X2:
            raw = np.empty((1 << self.terminal_rank, 2), np.uint16)
            check(self.lib.dcs_frontier_read(self.handle, int(lane), raw.ctypes.data_as(C.c_void_p)))
            f = raw.view(np.float16).astype(np.float32)
            return (f[:, 0] + 1j * f[:, 1]).astype(np.complex64)
        dt = np.complex64 if self.storage == STORAGE_C64 else np.complex128
        out = np.empty(1 << self.terminal_rank, dt)
        check(self.lib.dcs_frontier_read(self.handle, int(lane), out.ctypes.data_as(C.c_void_p)))
        return out

    def profile(self, enable):
        ms = np.zeros(5, np.float64)
        n = np.zeros(5, np.int64)
        check(self.lib.dcs_plan_profile(self.handle, int(bool(enable)), ptr(ms, C.c_double), ptr(n, C.c_int64)))
        return ms, n

    def close(self):
        if getattr(self, 'handle', None):
            self.lib.dcs_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
