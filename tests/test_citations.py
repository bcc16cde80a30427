"""Every reference citation in the repository (File.ext:N or File.ext:N-M,
with same-file continuations ', N-M' or ', :N-M'; the evaluation tool's
files as tool/src/dir/File.ext) names lines that exist in the cited
reference file (line counts in tests/golden/
reference_line_counts.json)."""
import json
import os
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
COUNTS = json.loads((ROOT / "tests" / "golden" /
                     "reference_line_counts.json").read_text())
PAT = re.compile(r"((?:[A-Za-z0-9_]+/)*[A-Za-z0-9_&]+\.(?:cu|cpp|h)):(\d+)"
                 r"(?:-(\d+))?((?:,\s*:?\d+(?:-\d+)?)*)")
SKIP = {"SURVEY.md", "VERDICT.md", "ADVICE.md", "BASELINE.md"}


def _sources():
    for top in ("include", "markerfree_amd", "oracle", "tests"):
        for root, dirs, files in os.walk(ROOT / top):
            dirs[:] = [d for d in dirs if d not in ("lib", "__pycache__", "bin")]
            for f in files:
                if f.endswith((".cpp", ".hip", ".h", ".py", ".md")):
                    yield Path(root) / f
    for f in ("bench.py", "__graft_entry__.py", "DESIGN.md", "INTEGRATION.md",
              "README.md"):
        if (ROOT / f).exists() and f not in SKIP:
            yield ROOT / f


def test_cited_ranges_exist():
    bad, n = [], 0
    for p in _sources():
        for ln, line in enumerate(p.read_text(errors="replace").splitlines(), 1):
            for m in PAT.finditer(line):
                # The tool's files by path, the method's by file name.
                f = m.group(1)
                if not f.startswith("tool/src/"):
                    f = f.rsplit("/", 1)[-1]
                if f not in COUNTS:
                    continue
                ranges = [(int(m.group(2)), int(m.group(3) or m.group(2)))]
                for a, b in re.findall(r":?(\d+)(?:-(\d+))?", m.group(4) or ""):
                    ranges.append((int(a), int(b or a)))
                for lo, hi in ranges:
                    n += 1
                    if lo > hi or hi > COUNTS[f]:
                        bad.append(f"{p.relative_to(ROOT)}:{ln} {f}:{lo}-{hi}")
    assert n > 200
    assert not bad, bad
