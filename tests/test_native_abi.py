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


def test_structure_layouts():
    assert ctypes.sizeof(_native.DrawResult) == 48
    assert ctypes.sizeof(_native.DeviceInfo) == 256 + 64 + 16 + 16
