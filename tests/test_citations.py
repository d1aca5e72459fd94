"""Every `file.jl:a-b` citation in the product, the oracle, the tests and the
documents points at lines that exist in the reference's file of that name.
Reads the reference sources as text; skipped where they are absent."""
import os
import re

import pytest

from conftest import REPO

REF = "/root/reference"
ROOTS = ["include", "oracle", "hdiscontinuousgalerkin.jl_amd", "tests", "bench.py", "__graft_entry__.py",
         "DESIGN.md", "INTEGRATION.md"]
EXTS = (".py", ".c", ".h", ".hip", ".cpp", ".jl", ".md")
CITE = re.compile(r"([\w/.]*?\w+\.jl):(\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*)")


def _reference_lengths():
    lens = {}
    for d, _, fs in os.walk(REF):
        for f in fs:
            if f.endswith(".jl"):
                p = os.path.join(d, f)
                with open(p, errors="ignore") as fh:
                    lens[os.path.relpath(p, REF)] = sum(1 for _ in fh)
    return lens


def _sources():
    for root in ROOTS:
        top = os.path.join(REPO, root)
        if os.path.isfile(top):
            yield top
            continue
        for d, dirs, fs in os.walk(top):
            dirs[:] = [x for x in dirs if x not in ("_build", "__pycache__")]
            for f in fs:
                if f.endswith(EXTS):
                    yield os.path.join(d, f)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
def test_reference_citations_in_range():
    lens = _reference_lengths()
    bad, seen = [], 0
    for p in _sources():
        with open(p, errors="ignore") as fh:
            lines = fh.readlines()
        for ln, line in enumerate(lines, 1):
            for m in CITE.finditer(line):
                name = m.group(1)
                base = os.path.basename(name)
                cands = [q for q in lens if os.path.basename(q) == base and (q.endswith(name) or name == base)]
                where = f"{os.path.relpath(p, REPO)}:{ln}: {m.group(0)}"
                if not cands:
                    bad.append(f"{where} (no such reference file)")
                    continue
                n = max(lens[q] for q in cands)
                for part in m.group(2).split(","):
                    a, _, b = part.strip().partition("-")
                    lo, hi = int(a), int(b or a)
                    seen += 1
                    if lo < 1 or hi < lo or hi > n:
                        bad.append(f"{where} (file has {n} lines)")
    assert seen > 100
    assert not bad, "\n".join(bad)
