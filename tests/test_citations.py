"""Every `file:line` citation of a reference file in this repository's sources and notes lies
inside the cited file. The reference tree is read as text; without it (for instance on a
machine that only runs the GPU tests) the test is skipped."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"

# `name.ext:12`, `name.ext:12-34`, `name.ext:12-34, 56-78`
CITE = re.compile(r"([A-Za-z0-9_./-]+\.(?:hpp|h|cu|cpp|py|toml|txt|cmake))"
                  r":(\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*)")
SCAN_DIRS = ("include", "oracle", "curadon_amd", "tests")
SCAN_FILES = ("bench.py", "__graft_entry__.py", "DESIGN.md", "INTEGRATION.md", "README.md")
SCAN_EXT = (".h", ".hpp", ".hip", ".cpp", ".c", ".py", ".md")


def _sources():
    for d in SCAN_DIRS:
        for dp, dn, fn in os.walk(os.path.join(ROOT, d)):
            dn[:] = [x for x in dn if not x.startswith((".", "_build", "__pycache__", "_lib"))]
            for f in fn:
                if f.endswith(SCAN_EXT) and f != os.path.basename(__file__):
                    yield os.path.join(dp, f)
    for f in SCAN_FILES:
        p = os.path.join(ROOT, f)
        if os.path.exists(p):
            yield p


def _reference_files():
    out = {}
    for dp, dn, fn in os.walk(REF):
        dn[:] = [x for x in dn if x != ".git"]
        for f in fn:
            rel = os.path.relpath(os.path.join(dp, f), REF)
            out[rel] = os.path.join(dp, f)
    return out


def _resolve(name, ref):
    """Reference files a cited name can mean: exact relative path, path suffix or basename."""
    if name in ref:
        return [name]
    if "/" in name:
        return [r for r in ref if r.endswith("/" + name)]
    return [r for r in ref if os.path.basename(r) == name]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_reference_citations_in_range():
    ref = _reference_files()
    nlines = {}
    bad, checked = [], 0
    for src in _sources():
        text = open(src, encoding="utf-8", errors="replace").read()
        for m in CITE.finditer(text):
            name, spec = m.group(1), m.group(2)
            cands = _resolve(name, ref)
            if not cands:
                continue  # not a reference file (a file of this repository)
            for rel in cands:
                if rel not in nlines:
                    with open(ref[rel], encoding="utf-8", errors="replace") as f:
                        nlines[rel] = sum(1 for _ in f)
            for part in spec.split(","):
                part = part.strip()
                lo, _, hi = part.partition("-")
                lo, hi = int(lo), int(hi or lo)
                checked += 1
                ok = lo >= 1 and lo <= hi and any(hi <= nlines[r] for r in cands)
                if not ok:
                    line = text.count("\n", 0, m.start()) + 1
                    bad.append(f"{os.path.relpath(src, ROOT)}:{line}: {name}:{part} "
                               f"(file has {max(nlines[r] for r in cands)} lines)")
    assert checked > 50, "citation scan found too few citations"
    assert not bad, "citations outside the cited file:\n" + "\n".join(bad)
