"""Fixtures of the mp_sql golden comparison (run in a container that has the
reference tree; the outputs are committed, the reference is not needed later).

Inputs:  examples/amorphous_Al92U08/data/Al_U_large_n0001.dat (the snapshot,
         copied as is: it is the reference's own example data)
         code/ref/xray_ff.txt (Al, U form-factor rows)
Outputs: examples/amorphous_Al92U08/Al_U_large_sql_0001_40..43.txt, the tables
         the reference's mp_sql wrote for that snapshot (log:
         examples/amorphous_Al92U08/mp_sql_cmd_log.txt): Z(Q), I(Q), S(Q) and
         unscaled I(Q), X-ray weights, Q range 20 1/A, q_step 0.01, smoothing
         FWHM 5 steps.  Stored as float64 arrays in al_u_sql_golden.npz; the
         header records and the first and last three data records of each
         table as text in al_u_sql_text.json.
"""
import json
import os
import shutil

import numpy as np

REF = "/root/reference"
EX = os.path.join(REF, "examples", "amorphous_Al92U08")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "al_u")


def table(path):
    rows = []
    with open(path) as f:
        lines = f.read().splitlines()
    start = next(i for i, l in enumerate(lines) if l.strip().startswith("Q [A-1]")) + 1
    for l in lines[start:]:
        if l.strip():
            rows.append([float(v) for v in l.split()])
    return np.array(rows)


def main():
    os.makedirs(OUT, exist_ok=True)
    shutil.copyfile(os.path.join(EX, "data", "Al_U_large_n0001.dat"), os.path.join(OUT, "Al_U_large_n0001.dat"))
    ff = {}
    for l in open(os.path.join(REF, "code", "ref", "xray_ff.txt")):
        p = l.split()
        if p and p[0] in ("Al", "U"):
            ff[p[0]] = [float(v) for v in p[1:10]]
    t = {m: table(os.path.join(EX, f"Al_U_large_sql_0001_{n}.txt")) for m, n in
         (("Z", 40), ("I", 41), ("S", 42), ("Iu", 43))}
    np.savez_compressed(os.path.join(OUT, "al_u_sql_golden.npz"), ffpar=np.array([ff["Al"], ff["U"]]),
                        q_range=20.0, q_step=0.01, smooth_fwhm=5, **t)
    text = {}
    for m, n in (("Z", 40), ("I", 41), ("S", 42), ("Iu", 43)):
        lines = open(os.path.join(EX, f"Al_U_large_sql_0001_{n}.txt"), encoding="utf-8").read().splitlines()
        start = next(i for i, l in enumerate(lines) if l.strip().startswith("Q [A-1]")) + 1
        data = [l for l in lines[start:] if l.strip()]
        text[m] = {"header": lines[:start], "first_rows": data[:3], "last_rows": data[-3:], "n_rows": len(data)}
    with open(os.path.join(OUT, "al_u_sql_text.json"), "w", encoding="utf-8") as f:
        json.dump(text, f, indent=1, ensure_ascii=False)
    for k, v in t.items():
        print(k, v.shape, v[0])


if __name__ == "__main__":
    main()
