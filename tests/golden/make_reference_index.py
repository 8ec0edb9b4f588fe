"""Write tests/golden/reference_lines.json: the line count of every text file of the reference
(paths relative to its top level), so that tests can check the file:line citations of this repository
without the reference present.  Run in the build container, where the reference is at /root/reference."""
import json
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
SUFFIXES = {'.py', '.json', '.jsonl', '.md', '.qasm', '.txt', '.csv'}


def main():
    out = {}
    for p in sorted(REF.rglob('*')):
        if p.is_file() and p.suffix in SUFFIXES:
            with open(p, 'rb') as fh:
                out[str(p.relative_to(REF))] = sum(1 for _ in fh)
    dst = Path(__file__).resolve().parent / 'reference_lines.json'
    dst.write_text(json.dumps(out, indent=0, sort_keys=True) + '\n')
    print(f'{len(out)} files -> {dst}')


if __name__ == '__main__':
    main()
