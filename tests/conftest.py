import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / 'dcs245-classical-sampling_amd', ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / 'tests' / 'golden'


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (gfx950) device')


@pytest.fixture(scope='session')
def ibm64():
    from dcs_amd.workloads import load_ibm64
    return load_ibm64()


@pytest.fixture(scope='session')
def records():
    import json
    return json.loads((GOLDEN / 'ibm64_records.json').read_text())


@pytest.fixture(scope='session')
def plan_fixtures():
    import json
    return json.loads((GOLDEN / 'plan_fixtures.json').read_text())
