"""Plan compilers: frozen IBM checksums and field-exact agreement with the reference planners."""
import numpy as np
import pytest

from dcs_amd.block_plan import BlockPlan
from dcs_amd.inplace_plan import compile_inplace, orbit_descriptor, orbit_slots
from dcs_amd.scan_plan import byte_tables, kernel_combos, make_plan
from dcs_amd.stabilizer import synthetic_instance

# inputs/plans.json:54-105 (rmax); run_r1/worker0/anchor.json:13-15 and audits.jsonl (work per slice)
FROZEN = {'inplace_k2_fp32': (34, 504846186778, 2), 'inplace_k3_fp64': (33, 315122088218, 3),
          'canonical_k5_fp64': (32, None, 5), 'half_k1': (35, None, 1)}


@pytest.fixture(scope='module')
def ibm_plans(ibm64):
    return {name: BlockPlan(ibm64['graph'], ibm64['unitaries'], ibm64['n'], ibm64['order'],
                            ibm64['plans'][name]['operations'], ibm64['open_vertices']) for name in FROZEN}


@pytest.mark.parametrize('name', list(FROZEN))
def test_ibm_frozen_checksums(ibm_plans, ibm64, name):
    bp = ibm_plans[name]
    rmax, work, nslices = FROZEN[name]
    assert bp.rmax == rmax == ibm64['plans'][name]['rmax']
    assert bp.inplace_rmax == rmax and bp.inplace_maxslots == rmax
    assert bp.terminal_rank == 20 and bp.k == 20
    assert len(bp.slices) == nslices
    if work is not None:
        assert bp.work == work


def test_anchor_record(ibm_plans, records):
    a = records['anchor']
    bp = ibm_plans['inplace_k2_fp32']
    assert (bp.rmax, bp.terminal_rank, bp.work, 1 << len(bp.slices)) == (
        a['rmax'], a['terminal_rank'], a['work_per_slice'], a['slices_evaluated'])
    aud = next(iter(records['audits'].values()))
    assert (ibm_plans['inplace_k3_fp64'].work, ibm_plans['inplace_k3_fp64'].rmax) == (aud['work_per_slice'], aud['rmax'])


def _fixture_cases(ibm64):
    yield 'ibm64_inplace_k2_fp32', BlockPlan(ibm64['graph'], ibm64['unitaries'], 64, ibm64['order'],
                                             ibm64['plans']['inplace_k2_fp32']['operations'], ibm64['open_vertices'])
    yield 'ibm64_inplace_k3_fp64', BlockPlan(ibm64['graph'], ibm64['unitaries'], 64, ibm64['order'],
                                             ibm64['plans']['inplace_k3_fp64']['operations'], ibm64['open_vertices'])
    inst = synthetic_instance(16, 12, 20, seed=0)
    yield 'ring16_d12_t20_s0', BlockPlan(inst['graph'], inst['unitaries'], 16, inst['natural_order'], [], list(range(16)))
    inst = synthetic_instance(32, 40, 120, seed=0)
    yield 'ring32_d40_t120_s0', BlockPlan(inst['graph'], inst['unitaries'], 32, inst['natural_order'], [],
                                          list(range(16, 32)))


def test_plans_match_reference_planners(ibm64, plan_fixtures):
    """Every Step / InplaceStep field equals the reference planners' output on the same inputs."""
    for name, bp in _fixture_cases(ibm64):
        fx = plan_fixtures[name]
        assert fx['inputs']['order'] == list(map(int, bp.order)), name
        assert fx['inputs']['fixed'] == sorted(bp.fixed), name
        assert fx['inputs']['stop'] == bp.stop and fx['inputs']['n_edges'] == len(bp.edges)
        steps, profile = make_plan(bp.N, bp.edges, bp.order, bp.fixed)
        assert profile == fx['profile']
        for s, g in zip(steps, fx['make_plan']):
            assert (s.v, s.free, s.r_in, s.r_out, s.beta, s.Lcols, s.kernel) == (
                g['v'], g['free'], g['r_in'], g['r_out'], g['beta'], g['Lcols'], g['kernel']), name
        isteps, basis, maxslots = compile_inplace(bp.order, bp.fut, bp.fixed, bp.stop)
        assert maxslots == fx['maxslots']
        assert {str(k): v for k, v in sorted(basis.items())} == fx['terminal_basis']
        assert len(isteps) == len(fx['compile_inplace'])
        for s, g in zip(isteps, fx['compile_inplace']):
            assert (s.v, s.j, s.kind, list(s.active_in), list(s.active_out), s.beta, s.K, s.Ka, s.pivot,
                    s.pivot_a, s.free) == (g['v'], g['j'], g['kind'], g['active_in'], g['active_out'], g['beta'],
                                           g['K'], g['Ka'], g['pivot'], g['pivot_a'], g['free']), (name, s.j)


def test_inplace_orbits_are_disjoint_partitions():
    """Orbit representatives x -> {c, c^K, c^Ka, ...} cover the live slots exactly once (small plan)."""
    inst = synthetic_instance(12, 8, 10, seed=1)
    bp = BlockPlan(inst['graph'], inst['unitaries'], 12, inst['natural_order'], [], list(range(12)))
    for st in bp.inplace_steps:
        code, _ = orbit_descriptor(st)
        slots = orbit_slots(st)
        live = set()
        for x in range(1 << len(slots)):
            c = sum(((x >> i) & 1) << p for i, p in enumerate(slots))
            orb = {c}
            if code in (1, 6):
                orb |= {o ^ st.Ka for o in orb}
            if code in (3, 5, 6, 7):
                orb |= {o ^ st.K for o in orb}
            assert not (orb & live)
            live |= orb
        in_mask = sum(1 << p for p in st.active_in)
        assert live == {c for c in range(1 << (max(st.active_in) + 1 if st.active_in else 0)) if not c & ~in_mask} | ({0} if not st.active_in else set())


def test_byte_tables_and_combos():
    cols = [3, 5, 9, 17, 33, 65, 129, 257, 513]
    T = byte_tables(cols, len(cols))
    for c in (0, 1, 77, 300, 511):
        acc = 0
        for i in range(len(cols)):
            if c >> i & 1:
                acc ^= cols[i]
        assert int(T[0][c & 255] ^ T[1][c >> 8]) == acc
    assert kernel_combos([1, 2, 4]) == [0, 1, 2, 3, 4, 5, 6, 7]
