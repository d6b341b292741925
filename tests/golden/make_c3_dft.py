"""Writes tests/golden/c3_dft_10k.npz: the FP64 direct DFT (the definition
of dirft2d1, code/src/mp_sqom56.f90:2530-2602, in 3D) of the precision-sweep
workload -- one static frame of a 128^3 ABO3 perovskite (10,485,760 atoms,
sigma 0.05 A, seed 20261205; neutron-weighted species) -- at 10,000 seeded
random modes of the 512 x 512 x 64 box, as the product set of 25 k1 x 25 k2
x 16 k3 values.

The frame comes from numpy's seeded generator on the host
(mp_tools_amd.synth.static_frame), so it is the same wherever it is
regenerated; the file carries checksums of the positions so that a reader
can tell.  Run from the repository root:  python tests/golden/make_c3_dft.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mp_tools_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

NCELL, SEED, N = 128, 20261205, (512, 512, 64)
MODE_SEED = 31


def modes():
    rng = np.random.default_rng(MODE_SEED)
    k1s = np.sort(rng.choice(np.arange(-(N[0] // 2), (N[0] - 1) // 2 + 1), 25, replace=False))
    k2s = np.sort(rng.choice(np.arange(-(N[1] // 2), (N[1] - 1) // 2 + 1), 25, replace=False))
    k3s = np.sort(rng.choice(np.arange(-(N[2] // 2), (N[2] - 1) // 2 + 1), 16, replace=False))
    return k1s, k2s, k3s


def checksums(pos):
    p = pos[0, :, :3].astype(np.float64)
    return np.array([p.sum(), (p * p).sum(), p[::9973].sum()])


def staged(pos, counts, w):
    off = np.concatenate([[0], np.cumsum(counts)])
    parts = [O.stage(pos[0, off[j]:off[j + 1]], None, (0, 0, 0), np.eye(3), (2 * np.pi / NCELL,) * 3,
                     (0, 0, 0), 0, float(w[j])) for j in range(len(counts))]
    return [np.concatenate([p[k] for p in parts]) for k in range(4)]


def main():
    t = time.time()
    pos, counts = synth.static_frame(NCELL, sigma=0.05, seed=SEED)
    w = 0.1 * synth.B_NEUTRON_FM
    x, y, z, c = staged(pos, counts, w)
    k1s, k2s, k3s = modes()
    F = O.dirft_type1_prod(x, y, z, c, k1s, k2s, k3s)
    out = os.path.join(ROOT, "tests", "golden", "c3_dft_10k.npz")
    np.savez_compressed(out, k1s=k1s, k2s=k2s, k3s=k3s, F=F, checksums=checksums(pos), weights=w,
                        ncell=NCELL, seed=SEED, n_modes=np.array(N))
    print(f"wrote {out} ({F.size} modes) in {time.time() - t:.1f} s")


if __name__ == "__main__":
    main()
