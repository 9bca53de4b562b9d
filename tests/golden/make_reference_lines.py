"""Writes reference_lines.json: for every file of the reference's cluster/
directory, its line count and the line of each SUBROUTINE / PROGRAM
statement.  tests/test_citations.py checks every file:line citation in this
repository against it, so the check runs where the reference is absent.

    python tests/golden/make_reference_lines.py [/root/reference/cluster]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
UNIT = re.compile(r"^\s+(subroutine|program)\s+(\w+)", re.IGNORECASE)


def main(ref="/root/reference/cluster"):
    out = {}
    for dp, dn, fn in os.walk(ref):
        dn.sort()
        for f in sorted(fn):
            p = os.path.join(dp, f)
            rel = os.path.relpath(p, ref)
            if f.startswith("."):
                continue
            with open(p, "rb") as fh:
                lines = fh.read().split(b"\n")
            if lines and lines[-1] == b"":
                lines.pop()
            units = {}
            for k, raw in enumerate(lines, 1):
                m = UNIT.match(raw.decode("latin-1"))
                if m:
                    units.setdefault(m.group(2).lower(), k)
            out[rel] = {"lines": len(lines), "units": units}
    with open(os.path.join(HERE, "reference_lines.json"), "w") as fh:
        json.dump(out, fh, indent=0, sort_keys=True)
        fh.write("\n")


if __name__ == "__main__":
    main(*sys.argv[1:])
