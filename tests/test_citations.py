"""Every citation of a reference file (`name.jl:N`, `name.jl:N-M`, `name.jl:N,K-L`)
in the repository's sources, tests and documents names a file the reference
has and lines that file has (tests/golden/reference_lines.json holds the
reference's line counts).  A bare name matching several files (wang_landau.jl,
multicanonical.jl) must fit at least one of them."""
import glob
import json
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CITE = re.compile(r"([\w/.-]+\.(?:jl|csv))(:\d+(?:-\d+)?(?:,\d+(?:-\d+)?)*)")


def _top_docs():
    """The repository's own top-level documents: README.md and the documents it
    links; other top-level markdown files are inputs to the build, not its text."""
    readme = os.path.join(ROOT, "README.md")
    linked = set(re.findall(r"([\w-]+\.md)", open(readme).read()))
    return {"README.md"} | linked


def _files():
    docs = _top_docs()
    for p in glob.glob(os.path.join(ROOT, "**", "*"), recursive=True):
        rel = os.path.relpath(p, ROOT)
        if rel.startswith(("_work", ".git", "gpurun_out")) or not os.path.isfile(p):
            continue
        if rel.endswith(".md") and os.sep not in rel and rel not in docs:
            continue
        if p.endswith((".py", ".h", ".hip", ".cpp", ".c", ".md", ".txt")):
            yield rel, p


def test_every_reference_citation_exists():
    counts = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_lines.json")))
    assert len(counts) > 50
    bad, n = [], 0
    for rel, p in _files():
        for ln, line in enumerate(open(p, errors="replace"), 1):
            for m in CITE.finditer(line):
                name = m.group(1).lstrip("./")
                cands = [f for f in counts if f == name or f.endswith("/" + name)]
                lines = [int(x) for x in re.findall(r"\d+", m.group(2))]
                n += 1
                if not cands:
                    bad.append(f"{rel}:{ln}: {m.group(0)} (no such reference file)")
                elif max(lines) > max(counts[c] for c in cands) or min(lines) < 1:
                    bad.append(f"{rel}:{ln}: {m.group(0)} (file has "
                               f"{max(counts[c] for c in cands)} lines)")
    assert n > 200, n
    assert not bad, "\n".join(bad)
