"""Extract the mp_dos table fixture from the reference's BZO example.

Reads examples/single_crystal_BZO/bzo200d_dos_0001_0020_01.txt under
/root/reference (the reference's own mp_dos output for 20 BZO snapshots)
and writes tests/golden/bzo/bzo_dos_table.json with its header records,
the frequency field of every row and the number of rows.  The snapshots
are not in the reference tree, so the DOS values themselves are not a
fixture; the file was written by an earlier version of the tool (MP_DOS55,
three decimals per value), whose rows differ from mp_dos56's f9.5 fields.
"""
import json
import os

SRC = "/root/reference/examples/single_crystal_BZO/bzo200d_dos_0001_0020_01.txt"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    lines = open(SRC).read().split("\n")
    if lines and lines[-1] == "":
        lines = lines[:-1]
    head, rows = lines[:11], lines[11:]
    out = {"header": head, "n_rows": len(rows), "freq_fields": [r[:7] for r in rows]}
    with open(os.path.join(HERE, "bzo", "bzo_dos_table.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
