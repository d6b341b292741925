"""Extract the mp_sqom map-file fixtures from the reference's BZO example.

Reads examples/single_crystal_BZO/bzo200d_sq_0001_0020_{07,11}.txt under
/root/reference (the reference's own output of mp_sqom 1.56: the total map
and the E = 0 map, 160 x 160) and writes tests/golden/bzo/bzo_sq_tables.json
with, for each table, its header records, its first and last three map
records, the number of map records and their lengths.  The maps' inputs are
not in the reference tree, so only the text layout is a fixture.
"""
import json
import os

SRC = "/root/reference/examples/single_crystal_BZO"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    out = {}
    for tag in ("07", "11"):
        lines = open(os.path.join(SRC, f"bzo200d_sq_0001_0020_{tag}.txt")).read().split("\n")
        if lines and lines[-1] == "":
            lines = lines[:-1]
        head = lines[:18]
        rows = lines[18:]
        out[tag] = {"header": head, "first_rows": rows[:3], "last_rows": rows[-3:], "n_rows": len(rows),
                    "row_lengths": sorted(set(len(r) for r in rows))}
    with open(os.path.join(HERE, "bzo", "bzo_sq_tables.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
