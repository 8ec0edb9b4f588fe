"""HIP path against the oracle, through the C ABI, on seeded ring circuits.

Tolerances: complex128 per-step schedule == oracle bit for bit; fused schedules round to storage
precision after every step exactly like the per-step one, so they differ only through exp(ls) of the
slice coefficient (<= 1e-13 relative); complex64 storage vs the complex128 oracle <= 1e-5 relative
to the block's largest amplitude (BASELINE target)."""
import numpy as np
import pytest

from dcs_amd.block_plan import BlockPlan
from dcs_amd.conditional_block import InplaceConditionalBlockGPU
from dcs_amd.sampling_protocol import cdf_primary, draw_checked
from dcs_amd.stabilizer import synthetic_instance
from oracle import oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


@pytest.fixture(scope='module')
def c1():
    inst = synthetic_instance(16, 12, 20, seed=0)
    args = (inst['graph'], inst['unitaries'], 16, inst['natural_order'], [], list(range(16)))
    bp = BlockPlan(*args)
    x = np.zeros(16, np.uint8)
    return {'args': args, 'x': x, 'ref128': oracle.block(bp, x), 'ref64': oracle.block(bp, x, storage=0),
            'psi': oracle.statevector(inst['ops'], 16)}


def engine(args, **kw):
    e = InplaceConditionalBlockGPU(*args)
    e.alloc(**kw)
    return e


def test_c1_c128_steps_bit_exact(c1):
    print('case c1 c128 steps', flush=True)
    e = engine(c1['args'], dtype=np.complex128, schedule='steps')
    v, meta = e.block(c1['x'])
    assert np.array_equal(np.asarray(v), c1['ref128'])
    assert meta['prefix_dtype'] == 'complex128' and meta['slices_evaluated'] == 1
    e.close()


def test_c1_c128_host_chain_bit_exact(c1):
    print('case c1 c128 steps_host', flush=True)
    e = engine(c1['args'], dtype=np.complex128, schedule='steps_host')
    v, _ = e.block(c1['x'])
    assert np.array_equal(np.asarray(v), c1['ref128'])
    e.close()


def test_c1_c64_steps_bit_exact(c1):
    print('case c1 c64 steps', flush=True)
    e = engine(c1['args'], dtype=np.complex64, schedule='steps')
    v, _ = e.block(c1['x'])
    assert np.array_equal(np.asarray(v), c1['ref64'])
    assert rel(v, c1['ref128']) < 1e-5
    e.close()


@pytest.mark.parametrize('sched', ['fused', 'compiled'])
@pytest.mark.parametrize('rb,dim,tl,pj', [(12, 6, 12, 3), (6, 6, 12, 4), (6, 8, 11, 3), (0, 4, 11, 4), (6, 1, 12, 3)])
def test_c1_fused_matches_steps(c1, rb, dim, tl, pj, sched):
    """complex128 fused and complex64 fused with per-step rounding keep the per-step mantissas
    (differences only through exp(ls)); complex64 fused with per-pass rounding stays within 1e-5 of
    FP64 and no further from it than the per-step complex64 schedule (x2 margin)."""
    print(f'case c1 {sched} rb={rb} dim={dim} tl={tl} pj={pj}', flush=True)
    for dt, ref, tol, exact in ((np.complex128, c1['ref128'], 1e-13, False), (np.complex64, c1['ref64'], 1e-13, True)):
        e = engine(c1['args'], dtype=dt, schedule=sched, resident_bits=rb, fuse_max_dim=dim, tile_log2=tl,
                   exact_rounding=exact, pass_max_dim=pj)
        v, _ = e.block(c1['x'])
        assert rel(v, ref) < tol, (dt, rel(v, ref))
        e.close()
    e = engine(c1['args'], dtype=np.complex64, schedule=sched, resident_bits=rb, fuse_max_dim=dim, tile_log2=tl,
               pass_max_dim=pj)
    v, _ = e.block(c1['x'])
    err, err_steps = rel(v, c1['ref128']), rel(c1['ref64'], c1['ref128'])
    print(f'c64 fused error {err:.3e}, per-step c64 error {err_steps:.3e}', flush=True)
    assert err < 1e-5 and err <= 2 * err_steps
    e.close()


@pytest.fixture(scope='module')
def w20():
    """28 qubits, depth 32, 64 T: frontier width 20, so coset groups run on full 2^11 / 2^12 tiles."""
    inst = synthetic_instance(28, 32, 64, seed=0)
    args = (inst['graph'], inst['unitaries'], 28, inst['natural_order'], [], list(range(20, 28)))
    bp = BlockPlan(*args)
    assert bp.inplace_maxslots == 20
    x = np.random.default_rng(11).integers(0, 2, 28, dtype=np.uint8)
    return {'args': args, 'x': x, 'ref128': oracle.block(bp, x), 'ref64': oracle.block(bp, x, storage=0)}


@pytest.mark.parametrize('tl,pj,lds', [(11, 3, False), (11, 3, True), (12, 4, True), (12, 3, False)])
def test_w20_compiled_matches_oracle(w20, tl, pj, lds):
    """Generated group kernels: complex128 within 1e-13 of the oracle; with per-step rounding the
    complex64 block equals the generic coset kernel's bit for bit; the production complex64 setting
    stays within 1e-5 of FP64 and no further than the per-step complex64 schedule (x2 margin)."""
    print(f'case w20 compiled tl={tl} pj={pj} lds={lds}', flush=True)
    e = engine(w20['args'], dtype=np.complex128, schedule='compiled', tile_log2=tl, pass_max_dim=pj, lds_storage=lds)
    v = np.asarray(e.block(w20['x'])[0])
    e.close()
    assert rel(v, w20['ref128']) < 1e-13, rel(v, w20['ref128'])
    blocks = []
    for sched in ('fused', 'compiled'):
        e = engine(w20['args'], dtype=np.complex64, schedule=sched, tile_log2=tl, pass_max_dim=pj, exact_rounding=True,
                   lds_storage=lds)
        blocks.append(np.asarray(e.block(w20['x'])[0]))
        e.close()
    assert np.array_equal(blocks[0], blocks[1])
    assert rel(blocks[1], w20['ref64']) < 1e-13
    e = engine(w20['args'], dtype=np.complex64, schedule='compiled', tile_log2=tl, pass_max_dim=pj, lds_storage=lds)
    v = np.asarray(e.block(w20['x'])[0])
    e.close()
    err, err_steps = rel(v, w20['ref128']), rel(w20['ref64'], w20['ref128'])
    print(f'c64 compiled error {err:.3e}, per-step c64 error {err_steps:.3e}', flush=True)
    assert err < 1e-5 and err <= 2 * err_steps


@pytest.mark.parametrize('tl,dim,policy,buffers', [(11, 6, 1, 1), (12, 8, 1, 1), (12, 9, 1, 2), (13, 9, 1, 1),
                                                  (11, 12, 1, 2), (12, 12, 1, 1), (12, 8, 2, 1), (13, 9, 2, 1),
                                                  (12, 8, 1, 3), (11, 6, 0, 3), (13, 9, 2, 3), (12, 8, 3, 3)])
def test_w20_wide_groups_match_oracle(w20, tl, dim, policy, buffers):
    """Coset groups of up to 2^12 entries by dimension, on 2^11..2^13 tiles, with the fewest-bytes group
    boundaries (which put slots born or dying inside a group on coordinates above other pivots) and one
    or two LDS tiles per workgroup, or two filled by asynchronous global-to-LDS loads (buffers 3):
    complex128 within 1e-13 of the oracle, per-step-rounded complex64 within 1e-13 of the oracle's
    complex64 (and equal to the generic coset kernel where that kernel supports the tile), production
    complex64 no further from FP64 than the per-step schedule (x2); asynchronous loads change no
    arithmetic, so that block equals the one-tile pipeline's bit for bit."""
    print(f'case w20 wide tl={tl} dim={dim} policy={policy} buffers={buffers}', flush=True)
    kw = dict(schedule='compiled', tile_log2=tl, fuse_max_dim=dim, group_policy=policy, lds_buffers=buffers)
    e = engine(w20['args'], dtype=np.complex128, **kw)
    v = np.asarray(e.block(w20['x'])[0])
    e.close()
    assert rel(v, w20['ref128']) < 1e-13, rel(v, w20['ref128'])
    e = engine(w20['args'], dtype=np.complex64, exact_rounding=True, **kw)
    exact = np.asarray(e.block(w20['x'])[0])
    e.close()
    assert rel(exact, w20['ref64']) < 1e-13, rel(exact, w20['ref64'])
    if tl <= 12 and dim <= 8:
        e = engine(w20['args'], dtype=np.complex64, exact_rounding=True, schedule='fused', tile_log2=tl, fuse_max_dim=dim,
                   group_policy=policy, lds_storage=True)
        assert np.array_equal(np.asarray(e.block(w20['x'])[0]), exact)
        e.close()
    e = engine(w20['args'], dtype=np.complex64, **kw)
    v = np.asarray(e.block(w20['x'])[0])
    e.close()
    err, err_steps = rel(v, w20['ref128']), rel(w20['ref64'], w20['ref128'])
    assert err < 1e-5 and err <= 2 * err_steps, (err, err_steps)
    if buffers == 3:
        e = engine(w20['args'], dtype=np.complex64, **{**kw, 'lds_buffers': 1})
        assert np.array_equal(np.asarray(e.block(w20['x'])[0]), v)
        e.close()


def test_c1_matches_statevector_and_64_draws_seed0(c1):
    """C1: all 2^16 amplitudes agree with the dense state vector; 64 draws from seed 0 are identical
    between device draw, host draw_checked and the oracle."""
    print('case c1 draws', flush=True)
    e = engine(c1['args'], dtype=np.complex128, schedule='fused')
    vec, _ = e.block(c1['x'])
    v = np.asarray(vec)
    psi = c1['psi']
    i = int(np.argmax(np.abs(psi)))
    assert rel(v, (v[i] / psi[i]) * psi) < 1e-12
    assert abs(np.vdot(v, v).real - 1) < 1e-12
    rng = np.random.Generator(np.random.PCG64(0))
    for u in rng.random(64):
        d = vec.draw(u)
        h = draw_checked(v, u)
        o = oracle.draw(v, u)
        assert d['A_index'] == h['A_index'] == o['A_index']
        assert d['cdf_lower_numpy'] == h['cdf_lower_numpy'] and d['cdf_upper_numpy'] == h['cdf_upper_numpy']
        assert d['scaled_mass_cdf'] == h['scaled_mass_cdf']
    e.close()


def test_prefix_blocks_and_lanes():
    """12 qubits, 6 open vertices, several prefixes; lanes=3 batching gives the same blocks."""
    print('case prefix lanes', flush=True)
    inst = synthetic_instance(12, 10, 14, seed=1)
    opn = [9, 2, 7, 11, 4, 0]
    args = (inst['graph'], inst['unitaries'], 12, inst['natural_order'], [], opn)
    bp = BlockPlan(*args)
    rng = np.random.default_rng(5)
    xs = [rng.integers(0, 2, 12, dtype=np.uint8) for _ in range(4)]
    refs = [oracle.block(bp, x) for x in xs]
    e1 = engine(args, dtype=np.complex128, schedule='steps')
    for x, r in zip(xs, refs):
        assert np.array_equal(np.asarray(e1.block(x)[0]), r)
    e1.close()
    for sched in ('fused', 'compiled'):
        e3 = engine(args, dtype=np.complex64, schedule=sched, lanes=3, resident_bits=6)
        for x, r in zip(xs, refs):
            assert rel(e3.block(x)[0], r) < 1e-5
        e3.close()


def test_w20_batched_lanes_equal_alone(w20):
    """Width-20 plan with the default schedule (asynchronous tile loads): two prefixes batched into the
    same group-kernel launches (lanes=2, one lane each) give the blocks each prefix gives alone with
    lanes=1, bit for bit, and the same device draws."""
    print('case w20 batched lanes', flush=True)
    rng = np.random.default_rng(23)
    n = w20['args'][2]
    xs = [w20['x'], rng.integers(0, 2, n, dtype=np.uint8)]
    us = (0.013, 0.37, 0.91)
    e1 = engine(w20['args'], dtype=np.complex64)
    alone, draws = [], []
    for x in xs:
        v = e1.block(x)[0]
        alone.append(np.asarray(v))
        draws.append([v.draw(u)['A_index'] for u in us])
    e1.close()
    e2 = engine(w20['args'], dtype=np.complex64, lanes=2, blocks=2)
    vecs, meta = e2.block_many(xs)
    for v, a, d in zip(vecs, alone, draws):
        assert np.array_equal(np.asarray(v), a)
        assert [v.draw(u)['A_index'] for u in us] == d
    e2.close()
    assert rel(alone[0], w20['ref128']) < 1e-5


def test_parity_sliced_ring_circuit():
    """Parity slices on reference vertices (Y, X and Z operations) against the unsliced oracle."""
    print('case parity slices', flush=True)
    inst = synthetic_instance(10, 10, 12, seed=2)
    n = 10
    opn = [0, 3, 6, 8]
    refs = [v for v in range(n, inst['N']) if inst['graph'][v].any()]
    nb = [int(u) for u in np.flatnonzero(inst['graph'][refs[2]])]
    ops = [('Y', refs[0], None), ('Z', refs[1], None), ('X', refs[2], nb[0])]
    x = np.array([1, 0, 0, 1, 1, 0, 0, 1, 0, 1], np.uint8)
    full = oracle.block(BlockPlan(inst['graph'], inst['unitaries'], n, inst['natural_order'], [], opn), x)
    args = (inst['graph'], inst['unitaries'], n, inst['natural_order'], ops, opn)
    sl = oracle.block(BlockPlan(*args), x)
    # parity slicing changes local Cliffords by exact phases: compare up to one global phase
    i = int(np.argmax(np.abs(full)))
    assert rel(sl, (sl[i] / full[i]) * full) < 1e-12
    e = engine(args, dtype=np.complex128, schedule='steps', lanes=2)
    v, meta = e.block(x)
    assert np.array_equal(np.asarray(v), sl)
    assert meta['slices_evaluated'] == 8
    e.close()


@pytest.mark.parametrize('sched', ['steps', 'steps_host'])
def test_canonical_gather_engine_matches_oracle(c1, w20, sched):
    """The canonical engine (gather executor): complex128 and complex64 blocks equal the C oracle's
    gather blocks bit for bit on C1 (all 2^16 amplitudes) and on the width-20 ring circuit."""
    from dcs_amd.conditional_block import ConditionalBlockGPU
    for case in (c1, w20):
        bp = BlockPlan(*case['args'])
        for dt, st in ((np.complex128, 1), (np.complex64, 0)):
            print(f'case canonical {sched} {dt.__name__} N={bp.N} rmax={bp.rmax}', flush=True)
            e = ConditionalBlockGPU(*case['args'])
            e.alloc(dtype=dt, schedule=sched)
            v, meta = e.block(case['x'])
            want = oracle.gather_block(bp, case['x'], storage=st)
            assert np.array_equal(np.asarray(v), want)
            assert meta['rmax'] == bp.rmax and meta['terminal_rank'] == len(bp.gather_terminal_basis)
            e.close()
    psi_ok = rel(oracle.gather_block(BlockPlan(*c1['args']), c1['x']), c1['ref128'])
    assert psi_ok < 1e-14


def test_canonical_parity_slices_and_prefixes():
    """Canonical engine on parity slices (Y, Z, X) and on 12-qubit prefix blocks, complex128, bit-exact
    against the oracle's gather blocks."""
    from dcs_amd.conditional_block import ConditionalBlockGPU
    print('case canonical slices/prefixes', flush=True)
    inst = synthetic_instance(10, 10, 12, seed=2)
    refs = [v for v in range(10, inst['N']) if inst['graph'][v].any()]
    nb = [int(u) for u in np.flatnonzero(inst['graph'][refs[2]])]
    ops = [('Y', refs[0], None), ('Z', refs[1], None), ('X', refs[2], nb[0])]
    args = (inst['graph'], inst['unitaries'], 10, inst['natural_order'], ops, [0, 3, 6, 8])
    x = np.array([1, 0, 0, 1, 1, 0, 0, 1, 0, 1], np.uint8)
    e = ConditionalBlockGPU(*args)
    e.alloc(dtype=np.complex128)
    v, meta = e.block(x)
    assert np.array_equal(np.asarray(v), oracle.gather_block(BlockPlan(*args), x))
    assert meta['slices_evaluated'] == 8
    e.close()
    inst = synthetic_instance(12, 10, 14, seed=1)
    args = (inst['graph'], inst['unitaries'], 12, inst['natural_order'], [], [9, 2, 7, 11, 4, 0])
    bp = BlockPlan(*args)
    e = ConditionalBlockGPU(*args)
    e.alloc(dtype=np.complex128)
    for seed in range(3):
        x = np.random.default_rng(seed).integers(0, 2, 12, dtype=np.uint8)
        assert np.array_equal(np.asarray(e.block(x)[0]), oracle.gather_block(bp, x))
    e.close()


@pytest.mark.parametrize('sched', ['fused', 'compiled'])
def test_nonfinite_prefix_raises(c1, sched):
    print(f'case nonfinite {sched}', flush=True)
    e = engine(c1['args'], dtype=np.complex64, schedule=sched)
    w = {q: np.array([np.nan, 1.0]) for q in range(16, 36)}
    with pytest.raises(ArithmeticError):
        e.block(c1['x'], w=w)
    e.close()


def test_zero_block_draw_raises(c1):
    print('case zero block', flush=True)
    e = engine(c1['args'], dtype=np.complex128, schedule='steps')
    w = {q: np.array([0.0, 0.0]) for q in range(16, 36)}
    vec, _ = e.block(c1['x'], w=w)
    assert not np.any(np.asarray(vec))
    with pytest.raises(ArithmeticError):
        vec.draw(0.5)
    e.close()


def test_admission_refuses_oversized_frontier():
    print('case admission', flush=True)
    from dcs_amd import _native
    ctx = _native.Context(0)
    t = {'kind': [0], 'K': [0], 'Ka': [0], 'pivot': [39], 'pivot_a': [-1], 'beta': [0], 'in_mask': [0],
         'out_mask': [1 << 39]}
    with pytest.raises(MemoryError):
        _native.Plan(ctx, t, 40, 1, 2, 1, 1 << 39, np.array([0, 1], np.uint64), np.ones(2),
                     np.eye(2)[None].astype(complex))
    ctx.close()


def test_device_draw_random_vectors():
    """Device sequential-cumsum draw equals numpy's draw on random blocks (index and both bounds)."""
    print('case device draw random', flush=True)
    inst = synthetic_instance(14, 10, 16, seed=9)
    e = engine((inst['graph'], inst['unitaries'], 14, inst['natural_order'], [], list(range(14))),
               dtype=np.complex64, schedule='fused')
    vec, _ = e.block(np.zeros(14, np.uint8))
    v = np.asarray(vec)
    for u in np.random.default_rng(1).random(200):
        d = vec.draw(u)
        idx, w, cum, tot = cdf_primary(v, u)
        assert d['A_index'] == idx and d['scaled_mass_cdf'] == tot
        assert d['cdf_upper_numpy'] == cum[idx] and d['cdf_lower_numpy'] == (cum[idx - 1] if idx else 0.)
    e.close()


@pytest.mark.parametrize('sched', ['steps', 'steps_host'])
def test_c1_half_storage_per_step_bit_exact(c1, sched):
    """Binary16 storage (kernel_opt/inplace_half_gpu.py): the per-step schedules equal the oracle's
    half-storage block bit for bit (same conversions, same products, same scale chain)."""
    from dcs_amd.conditional_block import HalfConditionalBlockGPU
    print(f'case c1 half {sched}', flush=True)
    ref = oracle.block(BlockPlan(*c1['args']), c1['x'], storage=2)
    e = HalfConditionalBlockGPU(*c1['args'])
    e.alloc(schedule=sched)
    v, meta = e.block(c1['x'])
    assert np.array_equal(np.asarray(v), ref)
    assert meta['intermediate_storage'] == 'float16 real/imaginary components'
    assert e.inplace_storage == 'float16-components'
    e.close()


@pytest.mark.parametrize('sched', ['fused', 'compiled'])
def test_w20_half_storage_fused(w20, sched):
    """Binary16 storage on the width-20 plan: the per-step schedule equals the oracle bit for bit.  Fused
    schedules apply the power-of-two scale once per group; unlike complex64, binary16 rounding does not
    commute with that scale (its subnormal floor 2^-24 is absolute), so fused blocks are held to
    accuracy instead: with per-step or per-pass rounding, no further from FP64 than the per-step
    binary16 schedule (x2)."""
    from dcs_amd.conditional_block import HalfConditionalBlockGPU
    print(f'case w20 half {sched}', flush=True)
    bp = BlockPlan(*w20['args'])
    ref16 = oracle.block(bp, w20['x'], storage=2)
    e = HalfConditionalBlockGPU(*w20['args'])
    e.alloc(schedule='steps')
    assert np.array_equal(np.asarray(e.block(w20['x'])[0]), ref16)
    e.close()
    err_steps = rel(ref16, w20['ref128'])
    for exact in (True, False):
        e = HalfConditionalBlockGPU(*w20['args'])
        e.alloc(schedule=sched, exact_rounding=exact)
        v = np.asarray(e.block(w20['x'])[0])
        e.close()
        err = rel(v, w20['ref128'])
        print(f'half {sched} exact={exact} error {err:.3e}, per-step half error {err_steps:.3e}', flush=True)
        assert err <= 2 * err_steps, (exact, err, err_steps)


@pytest.mark.parametrize('sched', ['steps', 'compiled'])
def test_graph_replay_equals_direct_launches(w20, sched):
    """Captured-graph replay of the contraction (one graph per lane count) gives blocks bit-identical to
    direct launches, on every replay: prefixes change between replays (the parameter upload is part of
    the graph), direct work on another engine and host synchronisations run in between, and a
    two-lane batch uses its own graph."""
    print(f'case graph replay {sched}', flush=True)
    rng = np.random.default_rng(5)
    xs = [rng.integers(0, 2, 28, dtype=np.uint8) for _ in range(5)]
    direct = engine(w20['args'], dtype=np.complex64, schedule=sched, lanes=2, blocks=2)
    want = [np.asarray(direct.block(x)[0]) for x in xs]
    graphed = engine(w20['args'], dtype=np.complex64, schedule=sched, lanes=2, blocks=2, graphs=True)
    for rep in range(2):
        for x, w in zip(xs, want):
            got = np.asarray(graphed.block(x)[0])
            assert np.array_equal(got, w)
            assert np.array_equal(np.asarray(direct.block(xs[0])[0]), want[0])   # eager work in between
    vecs, _ = graphed.block_many(xs[:2])
    assert all(np.array_equal(np.asarray(v), w) for v, w in zip(vecs, want[:2]))
    vecs, _ = graphed.block_many(xs[2:4])
    assert all(np.array_equal(np.asarray(v), w) for v, w in zip(vecs, want[2:4]))
    graphed.close()
    direct.close()
