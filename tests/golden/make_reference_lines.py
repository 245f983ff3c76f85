"""Writes tests/golden/reference_lines.json: the number of lines of every
Julia source file of the reference (path relative to its root -> lines), the
data tests/test_citations.py checks the repository's file:line citations
against.  Only line counts are recorded, no text.

Run from the repository root:  python tests/golden/make_reference_lines.py [REFERENCE_ROOT]
(default /root/reference).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    counts = {}
    for d, _, files in os.walk(ref):
        for f in files:
            if not f.endswith((".jl", ".csv")):
                continue
            p = os.path.join(d, f)
            txt = open(p, errors="replace").read()
            n = txt.count("\n") + (1 if txt and not txt.endswith("\n") else 0)
            counts[os.path.relpath(p, ref)] = n
    with open(os.path.join(HERE, "reference_lines.json"), "w") as fh:
        json.dump(dict(sorted(counts.items())), fh, indent=0)
        fh.write("\n")
    print(f"{len(counts)} files")


if __name__ == "__main__":
    main()
