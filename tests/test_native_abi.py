"""The C ABI: the library loads and exports every entry point declared in include/dcs_amd.h."""
import ctypes
import re
from pathlib import Path

from dcs_amd import _native

HEADER = Path(__file__).resolve().parents[1] / 'include' / 'dcs_amd.h'


def declared():
    text = HEADER.read_text()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(dcs_[a-z_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    names = declared()
    assert len(names) >= 18
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_native.EXPORTED)


def test_binding_loads_and_reports_errors():
    lib = _native.load()
    assert isinstance(lib.dcs_last_error(), bytes)


def test_structure_layouts(tmp_path):
    """Every ctypes structure of the binding has the size and field offsets the C compiler gives the
    header's structure (gcc on a probe that includes include/dcs_amd.h)."""
    import subprocess
    pairs = {'dcs_draw_result': _native.DrawResult, 'dcs_device_info': _native.DeviceInfo,
             'dcs_plan_desc': _native.PlanDesc, 'dcs_gather_desc': _native.GatherDesc,
             'dcs_batch_item': _native.BatchItem}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {']
    for cname, cls in pairs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ['return 0;', '}']
    src = tmp_path / 'probe.c'
    src.write_text('\n'.join(lines) + '\n')
    exe = tmp_path / 'probe'
    subprocess.run(['gcc', '-o', str(exe), str(src)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, cls in pairs.items():
        assert got[(cname, 'size')] == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)
