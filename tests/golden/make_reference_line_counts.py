"""Writes reference_line_counts.json: the number of lines of each source file
of the reference checkout that the repository cites (file name -> lines), so
that tests/test_citations.py can check every cited range without the
reference present.

Usage: python tests/golden/make_reference_line_counts.py [/root/reference]
"""
import json
import os
import sys
from pathlib import Path

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "Markerfree" / "src"
counts = {}
for root, _, files in os.walk(ref):
    for f in files:
        if f.endswith((".cu", ".cpp", ".h")):
            with open(os.path.join(root, f), "rb") as fh:
                counts[f] = fh.read().count(b"\n") + 1
out = Path(__file__).with_name("reference_line_counts.json")
out.write_text(json.dumps(dict(sorted(counts.items())), indent=1) + "\n")
print(f"{len(counts)} files -> {out}")
