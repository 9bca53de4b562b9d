"""Every reference citation in this repository points inside the cited file.

Citations have the form `file.f:N` or `file.f:N-M` (also `.inc`, `.3d`,
`Makefile`); the line counts come from tests/golden/reference_lines.json,
which tests/golden/make_reference_lines.py writes from the reference's
cluster/ directory.  The entry points of the C ABI (include/fcmd.h) must cite
the line of the SUBROUTINE statement they replace.
"""
import json
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_lines.json")))
BY_BASE = {}
for _rel, _v in TABLE.items():
    BY_BASE.setdefault(os.path.basename(_rel), []).append(_v)

CITE = re.compile(r"(?<![\w/.-])((?:helpers/)?[A-Za-z0-9_][A-Za-z0-9_.-]*\.(?:f|inc|3d)|Makefile):(\d+)(?:-(\d+))?")
SKIP_TOP = {"SURVEY.md", "VERDICT.md", "ADVICE.md", "BASELINE.md", "BASELINE.json"}
SKIP_DIRS = {".git", "_work", "gpurun_out", "lib", "_ref", "__pycache__", ".pytest_cache", "golden"}
EXTS = (".py", ".md", ".h", ".hpp", ".hip", ".c", ".cpp", ".f", ".f90", ".txt")


def _files():
    for dp, dn, fn in os.walk(ROOT):
        dn[:] = [d for d in dn if d not in SKIP_DIRS]
        for f in fn:
            rel = os.path.relpath(os.path.join(dp, f), ROOT)
            if rel in SKIP_TOP or not f.endswith(EXTS) and f != "Makefile":
                continue
            yield rel


def _lookup(name):
    if name in TABLE:
        return [TABLE[name]]
    return BY_BASE.get(os.path.basename(name), [])


def test_every_citation_is_inside_its_file():
    bad, seen = [], 0
    for rel in _files():
        with open(os.path.join(ROOT, rel), errors="replace") as fh:
            for ln, line in enumerate(fh, 1):
                for m in CITE.finditer(line):
                    cands = _lookup(m.group(1))
                    if not cands:
                        continue  # not a reference file (e.g. this repository's own Makefile)
                    a = int(m.group(2))
                    b = int(m.group(3)) if m.group(3) else a
                    seen += 1
                    if not any(1 <= a <= b <= c["lines"] for c in cands):
                        bad.append(f"{rel}:{ln}: {m.group(0)} (file has {cands[0]['lines']} lines)")
    assert seen > 100
    assert not bad, "\n".join(bad)


ENTRY = {"AccAtom()": "accatom", "neighbor(step)": "neighbor", "integrator(step)": "integrator",
         "stoddard(step)": "stoddard"}


def test_abi_entry_points_cite_their_subroutine_statements():
    """include/fcmd.h's table of replaced subroutines cites, for each
    reference file, the line of its SUBROUTINE statement."""
    hdr = open(os.path.join(ROOT, "include", "fcmd.h")).read()
    table = hdr[hdr.index("Drop-in boundary"):hdr.index("The Fortran shims")]
    checked = 0
    unit = None
    for line in table.splitlines():
        for k, v in ENTRY.items():
            if k in line:
                unit = v
        for m in re.finditer(r"([A-Za-z0-9-]+\.f):(\d+)", line):
            if unit is None:
                continue
            want = TABLE[m.group(1)]["units"][unit]
            assert int(m.group(2)) == want, (m.group(0), want)
            checked += 1
    assert checked == 10


@pytest.mark.parametrize("name", ["for-a-lj-n-pbc.f", "for-a-lj-n-pbc-m.f", "neighbor-pbc.f",
                                  "neighbor-pbc-mav.f", "int-verlet.f", "int-verlet-mav.f",
                                  "int-verlet-nh-mav.f", "stoddard-pbc-n.f", "stoddard-pbc-n-m.f",
                                  "stoddard-pbc-n-m-fre.f"])
def test_table_holds_the_plugin_files(name):
    assert TABLE[name]["lines"] > 50
