"""Writes tests/golden/ref_tables.json from the reference-built table harness.

The harness (oracle/ref/ref_tables.cpp, built by oracle/ref/Makefile from the
hpGEM source tree) prints hpGEM's own quadrature rules, DG basis values and
gradients, reference-shape nodes, face maps and face-orientation maps. The
JSON it prints is committed verbatim as a fixture; this script only builds the
harness (when the reference tree is present) and captures its output.

Run from the repository root:  python tests/golden/make_ref_fixtures.py
"""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_tables")
OUT = os.path.join(ROOT, "tests", "golden", "ref_tables.json")


def main():
    if os.path.isdir("/root/reference/kernel"):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle", "ref"), "-j8"], check=True,
                       stdout=subprocess.DEVNULL)
    if not os.path.exists(HARNESS):
        sys.exit("reference table harness not built (needs the hpGEM source tree)")
    out = subprocess.run([HARNESS], check=True, capture_output=True, text=True).stdout
    data = json.loads(out)
    with open(OUT, "w") as f:
        json.dump(data, f, separators=(",", ":"))
        f.write("\n")
    print(f"wrote {OUT}: {len(data['tables'])} tables")


if __name__ == "__main__":
    main()
