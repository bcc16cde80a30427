"""Writes reference_line_counts.json: the number of lines of each source file
of the reference checkout that the repository cites (file name, or path for
the evaluation tool's files, -> lines), so
that tests/test_citations.py can check every cited range without the
reference present.

Usage: python tests/golden/make_reference_line_counts.py [/root/reference]
"""
import json
import os
import sys
from pathlib import Path

top = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
counts = {}
# The method's sources by file name; the evaluation tool's by their path
# from the checkout (cited as tool/src/...; some names repeat the method's).
for base, by_path in ((top / "Markerfree" / "src", False), (top / "tool" / "src", True)):
    for root, _, files in os.walk(base):
        for f in files:
            if f.endswith((".cu", ".cpp", ".h")):
                path = os.path.join(root, f)
                key = os.path.relpath(path, top) if by_path else f
                with open(path, "rb") as fh:
                    counts[key] = fh.read().count(b"\n") + 1
out = Path(__file__).with_name("reference_line_counts.json")
out.write_text(json.dumps(dict(sorted(counts.items())), indent=1) + "\n")
print(f"{len(counts)} files -> {out}")
