"""Every reference citation `src/<file>:<lines>` or `include/<file>:<lines>`
in this repository's code, tests and design notes names a file of the
reference and lines inside it (file lengths in
tests/golden/ref_line_counts.json)."""
import glob
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
PAT = re.compile(r"\b((?:src|include)/[A-Za-z0-9_./]+?\.(?:cpp|cuh|cu|h|bak))(?![A-Za-z])"
                 r"(?::(\d+(?:-\d+)?(?:,\d+(?:-\d+)?)*))?")
OWN = {"include/mclics_hip.h"}  # this repository's own header


# what this repository ships: the package, the oracle, the header, the tests,
# the profiles, the bench and the three documents of its own
ROOTS = ("mclics-backend_amd", "oracle", "include", "tests", "profiles")
DOCS = ("DESIGN.md", "INTEGRATION.md", "README.md", "bench.py", "__graft_entry__.py")


def cited_files():
    for root in ROOTS:
        for f in glob.glob(os.path.join(REPO, root, "**", "*"), recursive=True):
            if os.path.isfile(f) and f.endswith((".c", ".h", ".cpp", ".hip", ".py", ".md")):
                yield os.path.relpath(f, REPO), f
    for name in DOCS:
        f = os.path.join(REPO, name)
        if os.path.isfile(f):
            yield name, f


def test_citations_point_inside_the_reference_files():
    counts = json.load(open(os.path.join(HERE, "golden", "ref_line_counts.json")))
    bad, n = [], 0
    for rel, f in cited_files():
        for ln, line in enumerate(open(f, errors="replace"), 1):
            for m in PAT.finditer(line):
                path = m.group(1)
                if path in OWN or "rocrand" in path:
                    continue
                if path not in counts:
                    bad.append(f"{rel}:{ln}: no such reference file {path}")
                    continue
                for part in (m.group(2) or "").split(",") if m.group(2) else []:
                    a, _, b = part.partition("-")
                    n += 1
                    if not (1 <= int(a) <= int(b or a) <= counts[path]):
                        bad.append(f"{rel}:{ln}: {path}:{part} outside 1..{counts[path]}")
    assert n > 150
    assert not bad, "\n".join(bad)
