"""Writes the reference fixtures under tests/golden/ from oracle/_ref/ref_models_dump.

The driver is compiled by oracle/Makefile against the reference's own
src/lgca_models.h; this script only runs it and stores its output as data:

  ref_models.json   per model: NUM_DIR, INV_DIR, MIR_DIR_X/Y, lattice-vector
                    FP32 bit patterns, and the collide (p = 0, 1), bounce_back
                    and bounce_forward_x/y tables over every input byte
  ref_offsets.json  the ten neighbour / wrap offset tables of each model's
                    ModelDescriptor constructor for a set of small lattices

Run from the repository root:  python tests/golden/make_ref_fixtures.py
"""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_models_dump")
OUT = os.path.dirname(os.path.abspath(__file__))

# (model, dim_x, dim_y): HPP accepts any size; the FHP models need an even dim_y
# (src/lattice.cpp:141).
OFFSET_CASES = [
    (0, 1, 1), (0, 1, 3), (0, 2, 2), (0, 3, 5), (0, 5, 4), (0, 8, 6), (0, 33, 7),
    (1, 1, 2), (1, 2, 2), (1, 3, 4), (1, 5, 6), (1, 33, 6),
    (2, 1, 2), (2, 2, 4), (2, 7, 2), (2, 8, 4), (2, 37, 10),
    (3, 2, 2), (3, 5, 6), (3, 64, 8),
]


def main() -> int:
    if not os.path.exists(DUMP):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    tables = json.loads(subprocess.run([DUMP, "tables"], check=True, capture_output=True,
                                       text=True).stdout)
    with open(os.path.join(OUT, "ref_models.json"), "w") as f:
        json.dump(tables, f, separators=(",", ":"))
        f.write("\n")
    offsets = []
    for m, dx, dy in OFFSET_CASES:
        rec = json.loads(subprocess.run([DUMP, "offsets", str(m), str(dx), str(dy)], check=True,
                                        capture_output=True, text=True).stdout)
        rec["model"] = m
        offsets.append(rec)
    with open(os.path.join(OUT, "ref_offsets.json"), "w") as f:
        json.dump(offsets, f, separators=(",", ":"))
        f.write("\n")
    print("wrote ref_models.json, ref_offsets.json")
    return 0


if __name__ == "__main__":
    sys.exit(main())
