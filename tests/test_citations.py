"""Every file:line citation in this repository points at lines that exist.

Citations of the reference are checked against tests/golden/reference_lines.json (the line count of
each of the reference's text files, written by tests/golden/make_reference_index.py); a cited path
may name the file from the reference's top level, from its source/ directory, or by a unique tail of
its path.  Citations of this repository's own files are checked against the files themselves."""
import json
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
LINES = json.loads((ROOT / 'tests' / 'golden' / 'reference_lines.json').read_text())

CITE = re.compile(r'(?<![\w/.-])((?:[\w.-]+/)*[\w.-]+\.(?:py|jsonl|json|md|qasm|txt|csv|hip|hpp|h|c)):'
                  r'(\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*)')
# files whose text is not this repository's (written by the survey, the judge or the driver)
NOT_OURS = {'SURVEY.md', 'VERDICT.md', 'ADVICE.md', 'BASELINE.md'}
SCAN_SUFFIXES = {'.py', '.md', '.hip', '.hpp', '.h', '.c', '.cpp', '.sh', '.txt'}


def scanned_files():
    for p in sorted(ROOT.rglob('*')):
        rel = p.relative_to(ROOT)
        if not p.is_file() or p.suffix not in SCAN_SUFFIXES or p.name in NOT_OURS:
            continue
        if rel.parts[0] in ('_work', 'gpurun_out', '.git') or '__pycache__' in rel.parts:
            continue
        yield p


def repo_files():
    out = {}
    for p in ROOT.rglob('*'):
        rel = p.relative_to(ROOT)
        if p.is_file() and rel.parts[0] not in ('_work', 'gpurun_out', '.git'):
            out.setdefault(p.name, []).append(p)
    return out


def resolve_reference(path):
    """Line count of the reference file `path` names, or None when it names none."""
    for cand in (path, 'source/' + path):
        if cand in LINES:
            return LINES[cand]
    tails = [k for k in LINES if k.endswith('/' + path)]
    if len(tails) == 1:
        return LINES[tails[0]]
    if len(tails) > 1:
        counts = {LINES[k] for k in tails}
        return max(counts)
    return None


def ranges(spec):
    for part in spec.replace(' ', '').split(','):
        a, _, b = part.partition('-')
        yield int(a), int(b or a)


def test_every_citation_is_in_range():
    own = repo_files()
    bad, checked = [], 0
    for f in scanned_files():
        text = f.read_text(errors='replace')
        for m in CITE.finditer(text):
            path, spec = m.group(1), m.group(2)
            name = path.rsplit('/', 1)[-1]
            n = None
            mine = [p for p in own.get(name, []) if str(p.relative_to(ROOT)).endswith(path)]
            if path.split('/')[0] in ('dcs_amd', 'tests', 'oracle', 'include', 'csrc', 'profiles') or \
                    (mine and resolve_reference(path) is None):
                if mine:
                    n = max(len(p.read_text(errors='replace').splitlines()) for p in mine)
            else:
                n = resolve_reference(path)
            if n is None:
                continue
            checked += 1
            for a, b in ranges(spec):
                if not (1 <= a <= b <= n):
                    line = text.count('\n', 0, m.start()) + 1
                    bad.append(f'{f.relative_to(ROOT)}:{line}: {path}:{spec} (file has {n} lines)')
    assert checked > 100
    assert not bad, '\n'.join(bad)
