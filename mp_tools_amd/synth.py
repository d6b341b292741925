"""Synthetic inputs of the S(Q,w) workloads (no datasets are available).

Cubic ABO3 perovskite supercells in the MP_TOOLS CELL layout: positions in
lattice units centred in [-n/2, n/2), species-major storage (A cells, B
cells, then the three O sites), each species in CELL record order with x
fastest (code/src/mp_lbin56.f90:1204).  Site set as in
examples/single_crystal_BZO/BZO200d.par:12-17; neutron coherent lengths Ba
5.07, Zr 7.16, O 5.803 fm (code/ref/neutron_xs.txt) are used for A/B/O.

Trajectories superpose plane-wave phonons commensurate with the supercell
(integer wave vectors, so the box stays periodic) plus Gaussian noise.
"""
from __future__ import annotations

import numpy as np

A_LATTICE = 4.19                       # Angstrom
SITES = np.array([[0.0, 0.0, 0.0],     # A
                  [0.5, 0.5, 0.5],     # B
                  [0.5, 0.5, 0.0],     # O1
                  [0.0, 0.5, 0.5],     # O2
                  [0.5, 0.0, 0.5]])    # O3
SPECIES = ("A", "B", "O")
B_NEUTRON_FM = np.array([5.07, 7.16, 5.80])


def perovskite_lattice(ncell: int):
    """Ideal sites R [n_tot, 3] (lattice units, float64) and species counts."""
    n = int(ncell)
    iz, iy, ix = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    blocks = [cells + SITES[s] - n / 2.0 for s in range(5)]
    R = np.concatenate(blocks, axis=0)
    counts = np.array([n ** 3, n ** 3, 3 * n ** 3], np.int64)
    return R, counts


def cell_geometry(ncell: int, dim: int = 3):
    """NUFFT staging for a cubic CELL box: unit axes, tpe = 2 pi / n, centre 0
    (mp_sqom56.f90:799-811 with e1=[100], e2=[010], e3=[001])."""
    return dict(axes=np.eye(3), tpe=(2 * np.pi / ncell,) * 3, centre=(0.0, 0.0, 0.0),
                q0=(0.0, 0.0, 0.0))


def static_frame(ncell: int, sigma: float = 0.05, seed: int = 0, a: float = A_LATTICE):
    """One frame: ideal sites + iid N(0, sigma) Angstrom displacements.
    Returns float32 [1, n_tot, 4] (x, y, z, 0) and species counts."""
    R, counts = perovskite_lattice(ncell)
    rng = np.random.default_rng(seed)
    pos = R + rng.standard_normal(R.shape) * (sigma / a)
    out = np.zeros((1, R.shape[0], 4), np.float32)
    out[0, :, :3] = pos
    return out, counts


def phonon_params(ncell: int, n_phonons: int = 64, rms: float = 0.06, seed: int = 0):
    rng = np.random.default_rng(seed)
    q = rng.integers(-(ncell // 2), ncell // 2, size=(n_phonons, 3)).astype(np.float64)
    pol = rng.standard_normal((n_phonons, 3))
    pol /= np.linalg.norm(pol, axis=1, keepdims=True)
    nu = rng.uniform(1.0, 20.0, n_phonons)          # THz
    phi = rng.uniform(0.0, 2 * np.pi, n_phonons)
    amp = rms * np.sqrt(2.0 / n_phonons)             # Angstrom, vector rms = rms
    return q, pol, nu, phi, amp


def phonon_trajectory(ncell: int, T: int, dt_ps: float = 0.01, n_phonons: int = 64,
                      rms: float = 0.06, noise: float = 0.01, seed: int = 0,
                      a: float = A_LATTICE, device="cpu"):
    """T frames at step dt_ps of the supercell.  Returns a float32 torch tensor
    [T, n_tot, 4] on `device` and the species counts.

    u_j(t) = sum_p amp * pol_p * cos(2 pi q_p . R_j / n - 2 pi nu_p t + phi_p)
             + N(0, noise) per component   (Angstrom; positions in lattice units)
    """
    import torch

    R, counts = perovskite_lattice(ncell)
    q, pol, nu, phi, amp = phonon_params(ncell, n_phonons, rms, seed)
    dev = torch.device(device)
    Rt = torch.tensor(R, dtype=torch.float64, device=dev)
    ph0 = (2 * np.pi / ncell) * (torch.tensor(q, dtype=torch.float64, device=dev) @ Rt.T)
    ph0 += torch.tensor(phi, dtype=torch.float64, device=dev)[:, None]      # [P, n_tot]
    polt = torch.tensor(pol * (amp / a), dtype=torch.float64, device=dev)     # [P, 3] lattice units
    om = torch.tensor(2 * np.pi * nu, dtype=torch.float64, device=dev)        # rad/ps
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed) + 1)
    out = torch.zeros((T, R.shape[0], 4), dtype=torch.float32, device=dev)
    for t in range(T):
        c = torch.cos(ph0 - om[:, None] * (t * dt_ps))                        # [P, n_tot]
        u = polt.T @ c                                                         # [3, n_tot]
        nz = torch.randn((R.shape[0], 3), generator=gen, dtype=torch.float64, device=dev)
        out[t, :, :3] = (Rt + u.T + nz * (noise / a)).to(torch.float32)
    return out, counts
