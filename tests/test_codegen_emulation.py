"""Generated coset-group kernels emulated on the CPU (no GPU).

dcs_plan_kernel_source returns, without a device, the exact source the library compiles with hipRTC
for a plan. This test compiles that source with the host compiler against tests/emu/shim.hpp (one
thread per work item, a barrier per workgroup, synchronous global-to-LDS loads) and runs every group
kernel of the width-20 synthetic plan. With per-step rounding (or complex128 storage, where rounding is
the identity) each kernel's live entries and recorded maximum must equal the per-step orbit formulas
bit for bit, and so must the whole chain of groups from |0...0> with the scale carried between them.

It pins what the device parity tests can only see through their results: the coset addressing, the
register passes, dead-entry handling, the LDS index permutation and deferred stores of the
asynchronous-load pipeline, and the one- and two-tile pipelines. It does not pin the asynchronous
timing, which the emulator runs synchronously; the ordering argument for that is in the generator's
comments.
"""
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dcs_amd import _native
from dcs_amd.block_plan import BlockPlan
from dcs_amd.stabilizer import synthetic_instance

EMU = Path(__file__).resolve().parent / 'emu'

# (storage, tile_log2, fuse_max_dim, group_policy, lds_buffers, exact_rounding, gridcap)
CASES = [
    (0, 12, 8, 1, 3, True, 1),    # complex64, asynchronous loads, one workgroup walks every tile
    (0, 12, 8, 1, 4, True, 1),    # complex64, asynchronous loads, last pass stores from registers
    (0, 12, 8, 3, 5, True, 2),    # complex64, ring of compact input buffers two tiles ahead where a group allows
    (0, 12, 8, 2, 4, True, 3),    # the same with three workgroups and bytes-and-passes groups
    (0, 12, 9, 3, 3, True, 2),    # groups of dimension 9 on 2^13 tiles beside 2^12-entry ones
    (0, 11, 6, 0, 2, True, 3),    # complex64, two register-fed LDS tiles on 2^11 tiles, greedy groups
    (1, 12, 8, 1, 1, False, 2),   # complex128, one LDS tile
    (2, 12, 8, 2, 3, True, 2),    # binary16 storage (one-tile pipeline), bytes-and-passes groups
]


def adapt(src):
    """The generated source as the host compiler takes it: no fixed-width typedefs (the shim's headers
    have them), no GPU attributes, the LDS-only barrier as a workgroup barrier, the non-temporal store
    as plain stores."""
    src = '\n'.join(line for line in src.split('\n') if not re.match(r'typedef .*int.*_t;', line))
    src = src.replace('__attribute__((amdgpu_waves_per_eu(4)))', '')
    src = re.sub(r'__attribute__\(\(address_space\(\d\)\)\)', '', src)
    src, n = re.subn(r'\{ asm volatile\(.*?s_barrier.*?\); \}', '{ __syncthreads(); }', src)
    assert n == 1, 'expected one LDS barrier definition in the prelude'
    src, n = re.subn(r'void dcs_store_nt\(float2 \*p, float2 r\) \{.*\}', 'void dcs_store_nt(float2 *p, float2 r) { *p = r; }', src)
    assert n == 1, 'expected one non-temporal store helper in the prelude'
    src, n = re.subn(r'void dcs_store_nt2\(float2 \*p, float2 a, float2 b\) \{.*\}',
                     'void dcs_store_nt2(float2 *p, float2 a, float2 b) { p[0] = a; p[1] = b; }', src)
    assert n == 1, 'expected one non-temporal pair store helper in the prelude'
    return src


@pytest.fixture(scope='module')
def w20():
    inst = synthetic_instance(28, 32, 64, seed=0)
    return BlockPlan(inst['graph'], inst['unitaries'], 28, inst['natural_order'], [], list(range(20, 28)))


@pytest.fixture(scope='module')
def w17():
    """A width-17 ring plan (eight groups of up to dimension 8, asynchronous loads, a ring group under
    group_policy 0): an eighth of w20's entries, so that its emulation takes seconds."""
    inst = synthetic_instance(24, 28, 48, seed=3)
    return BlockPlan(inst['graph'], inst['unitaries'], 24, inst['natural_order'], [], list(range(16, 24)))


def _emulate(w20, tmp_path, storage, tl, dim, policy, buffers, exact, gridcap, classes=None, tol=0.0, pj=3):
    """Generate the plan's group kernels (register passes of span pj), run them in the CPU emulator and
    return its report lines."""
    t = w20.step_table()
    extra = {} if classes is None else {'coef_class': classes}
    src, segs = _native.kernel_source(t, w20.inplace_maxslots, storage, w20.terminal_mask(), w20.k,
                                      fuse_max_dim=dim, resident_bits=0, tile_log2=tl, exact_rounding=exact,
                                      pass_max_dim=pj, lds_storage=True, group_policy=policy, lds_buffers=buffers,
                                      **extra)
    assert segs and len(set(s[0] for s in segs)) == len(segs)
    if buffers >= 3 and storage == 0:
        assert 'global_load_lds' in src        # the asynchronous pipeline is what runs here
    if buffers >= 4 and storage == 0:
        assert 'E[tb ^ gb0 ^ ' in src          # and its last pass stores from registers
    if buffers == 5 and storage == 0:
        assert 'ring_fill(F, ' in src          # some group takes the ring of compact input buffers
    if classes is not None:
        assert '(f0 factored)' in src and "(group's pending factors)" in src
    (tmp_path / 'kernels.inc').write_text(adapt(src))
    names = sorted({int(m) for m in re.findall(r'\bdcs_group_(\d+)\(St ', src)})
    assert names == list(range(len(names)))
    (tmp_path / 'table.inc').write_text(''.join(f'&dcs_group_{i},\n' for i in names))
    rows = [f"{len(t['kind'])} {w20.inplace_maxslots} {w20.terminal_mask()}"]
    for i in range(len(t['kind'])):
        rows.append(' '.join(str(int(t[k][i])) for k in ('kind', 'K', 'Ka', 'pivot', 'pivot_a', 'beta', 'in_mask',
                                                         'out_mask')))
    (tmp_path / 'plan.txt').write_text('\n'.join(rows) + '\n')
    (tmp_path / 'groups.txt').write_text(''.join(f'{g} {a} {c} {th}\n' for g, a, c, th in segs))
    cmd_extra = []
    if classes is not None:
        (tmp_path / 'classes.txt').write_text(' '.join(str(int(c)) for c in classes) + '\n')
        cmd_extra = [str(tmp_path / 'classes.txt'), repr(tol)]
    exe = tmp_path / 'emu'
    subprocess.run(['g++', '-std=c++20', '-O1', '-pthread', '-w', f'-I{EMU}', f'-I{tmp_path}', str(EMU / 'emu.cpp'),
                    '-o', str(exe)], check=True, timeout=600)
    r = subprocess.run([str(exe), str(tmp_path / 'plan.txt'), str(tmp_path / 'groups.txt'), str(gridcap)] + cmd_extra,
                       capture_output=True, text=True, timeout=900)
    print(r.stdout[-3000:])
    lines = r.stdout.splitlines()
    assert sum(line.startswith('group ') for line in lines) == len(segs)
    assert sum(line.startswith('chain unit ') for line in lines) == len(segs)  # resident_bits=0: groups cover all
    return r, lines


@pytest.mark.skipif(shutil.which('g++') is None, reason='needs the host C++ compiler')
@pytest.mark.parametrize('storage,tl,dim,policy,buffers,exact,gridcap', CASES)
def test_generated_kernels_match_steps_on_cpu(w20, tmp_path, storage, tl, dim, policy, buffers, exact, gridcap):
    r, lines = _emulate(w20, tmp_path, storage, tl, dim, policy, buffers, exact, gridcap)
    assert r.returncode == 0 and lines[-1] == 'RESULT ok (0 bad)', r.stdout[-2000:]


# f0-factored kernels: (storage, tile_log2, fuse_max_dim, group_policy, lds_buffers, gridcap, tolerance)
FACTORED = [(0, 12, 8, 1, 3, 1, 2e-6), (0, 12, 9, 3, 3, 2, 2e-6), (0, 12, 8, 3, 3, 2, 2e-6), (0, 12, 8, 3, 4, 3, 2e-6), (0, 12, 8, 0, 5, 1, 2e-6), (0, 12, 8, 2, 1, 2, 2e-6),
            (0, 11, 6, 0, 2, 3, 2e-6), (1, 12, 8, 1, 1, 2, 1e-12), (2, 12, 8, 2, 3, 2, 2e-2)]


@pytest.mark.skipif(shutil.which('g++') is None, reason='needs the host C++ compiler')
@pytest.mark.parametrize('exact', [True, False])
def test_span4_passes_match_steps_on_cpu(w17, tmp_path, exact):
    """Register passes over up to four basis vectors (16-entry orbits, the f0-factored default) with
    groups of dimension 9 on 2^13 tiles beside 2^12-entry ones, on the width-17 plan: per-step rounding
    bit for bit, and f0-factored within 2e-6 of the step formulas."""
    classes = None if exact else np.random.default_rng(5).integers(0, 4, len(w17.inplace_steps)).astype(np.int32)
    r, lines = _emulate(w17, tmp_path, 0, 12, 9, 3, 3, exact, 2, classes=classes, tol=0.0 if exact else 2e-6, pj=4)
    assert r.returncode == 0 and lines[-1] == 'RESULT ok (0 bad)', r.stdout[-2000:]


@pytest.mark.skipif(shutil.which('g++') is None, reason='needs the host C++ compiler')
@pytest.mark.parametrize('storage,tl,dim,policy,buffers,gridcap,tol', FACTORED)
def test_factored_kernels_match_steps_on_cpu(w17, tmp_path, storage, tl, dim, policy, buffers, gridcap, tol):
    """Group kernels with f0 factoring against the per-step formulas on the width-17 plan, within tol of
    the largest magnitude
    (complex64 and binary16 round once per register pass instead of once per step: 2e-6 and 2e-2;
    complex128 never rounds: 1e-12), on every group and the whole chain.  Classes are drawn per step
    (seeded) so that every class, and class-0 steps with f0 == 0, occur in the groups; the emulator
    draws step parameters that fit them.  A factored kernel that dropped the pending factors, used the
    wrong part of a ratio, or misplaced a lane sign would be off by O(1)."""
    classes = np.random.default_rng(5).integers(0, 4, len(w17.inplace_steps)).astype(np.int32)
    r, lines = _emulate(w17, tmp_path, storage, tl, dim, policy, buffers, False, gridcap, classes=classes, tol=tol)
    if storage == 2:
        # binary16: each group from O(1) random entries only.  The chain from |0...0> decays towards the
        # binary16 subnormal range (these coefficients have |f1| <= |f0|), where the per-step reference
        # rounds away what rounding once per pass keeps (the subnormal floor 2^-24 is absolute); the same
        # factored arithmetic with complex64 storage passes the chain at 2e-6 (the cases above)
        groups = [line for line in lines if line.startswith('group ')]
        assert groups and all(line.endswith('mismatches 0, max ok') for line in groups), r.stdout[-2000:]
    else:
        assert r.returncode == 0 and lines[-1] == 'RESULT ok (0 bad)', r.stdout[-2000:]


def test_kernel_source_modes_and_errors(w20):
    """Asynchronous global-to-LDS loads are generated for complex64 tiles only (complex128 and binary16
    groups take the one-tile pipeline under lds_buffers=3); a description the library rejects raises
    without touching a device."""
    t = w20.step_table()
    kw = dict(fuse_max_dim=8, resident_bits=12, tile_log2=12, exact_rounding=False, pass_max_dim=3,
              lds_storage=True, group_policy=1, lds_buffers=3)
    args = (t, w20.inplace_maxslots)
    c64, segs64 = _native.kernel_source(*args, 0, w20.terminal_mask(), w20.k, **kw)
    c128, segs128 = _native.kernel_source(*args, 1, w20.terminal_mask(), w20.k, **kw)
    h16, segs16 = _native.kernel_source(*args, 2, w20.terminal_mask(), w20.k, **kw)
    assert 'global_load_lds(' in c64 and 'global_load_lds(' not in c128 and 'global_load_lds(' not in h16
    assert segs64 == segs128 == segs16          # the grouping does not depend on the storage
    assert [s[1] for s in segs64] == sorted(s[1] for s in segs64)
    with pytest.raises(ValueError, match='group_policy'):
        _native.kernel_source(*args, 0, w20.terminal_mask(), w20.k, **{**kw, 'group_policy': 7})
    with pytest.raises(ValueError, match='tile_log2'):
        _native.kernel_source(*args, 0, w20.terminal_mask(), w20.k, **{**kw, 'tile_log2': 14})
    bad = dict(t)
    bad['in_mask'] = t['in_mask'].copy()
    bad['in_mask'][3] ^= 1 << 40                # a live slot beyond maxslots
    with pytest.raises(ValueError):
        _native.kernel_source(bad, w20.inplace_maxslots, 0, w20.terminal_mask(), w20.k, **kw)
