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


# ---------------------------------------------------------------------------
# amorphous binary boxes and the batched-jobs workload
# ---------------------------------------------------------------------------
B_AL_U_FM = np.array([3.449, 8.417])   # Al, U coherent lengths (code/ref/neutron_xs.txt)


def amorphous_box(side: float, density: float = 0.06, fractions=(0.92, 0.08), hard_core: float = 2.2,
                  seed: int = 0, device="cpu", max_rounds: int = 2000):
    """Random sequential addition of round(density * side^3) atoms, uniform in
    a periodic cube [-side/2, side/2)^3 (Angstrom), rejecting any candidate
    closer than `hard_core` to an accepted atom (minimum image).  Candidates
    are proposed in batches; inside a batch a candidate also yields to any
    lower-numbered candidate within the hard core.  The species are assigned
    by a seeded permutation in the given fractions and stored species-major.
    Returns (float64 torch tensor [n, 3], species counts)."""
    import torch

    dev = torch.device(device)
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    n = int(round(density * side ** 3))
    cell = hard_core / np.sqrt(3.0)          # at most one atom per grid cell
    G = max(1, int(np.floor(side / cell)))
    h = side / G
    reach = int(np.ceil(hard_core / h))
    grid = torch.full((G, G, G), -1, dtype=torch.int64, device=dev)
    pos = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    r = torch.arange(-reach, reach + 1, device=dev)
    off = torch.stack(torch.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)   # [K, 3]
    count = 0
    for _ in range(max_rounds):
        if count >= n:
            break
        K = min(max(8 * (n - count), 1 << 16), 1 << 17)
        cand = (torch.rand((K, 3), generator=g, dtype=torch.float64, device=dev) - 0.5) * side
        ci = torch.clamp(((cand + side / 2) / h).long(), 0, G - 1)
        nb = (ci[:, None, :] + off[None]) % G                                # [K, 27.., 3]
        idx = grid[nb[..., 0], nb[..., 1], nb[..., 2]]                          # accepted atoms nearby
        d = cand[:, None, :] - pos[idx.clamp(min=0)]
        d -= side * torch.round(d / side)
        close = ((d * d).sum(-1) < hard_core ** 2) & (idx >= 0)
        ok = ~close.any(1)
        cand, ci = cand[ok], ci[ok]
        if cand.shape[0] == 0:
            continue
        # conflicts among this batch's survivors: lowest index wins per cell,
        # then each survivor yields to a lower-numbered one within the core
        k = torch.arange(cand.shape[0], device=dev)
        tmp = torch.full((G * G * G,), cand.shape[0], dtype=torch.int64, device=dev)
        lin = (ci[:, 0] * G + ci[:, 1]) * G + ci[:, 2]
        tmp.scatter_reduce_(0, lin, k, reduce="amin")
        win = tmp[lin] == k
        cand, ci, k = cand[win], ci[win], k[win]
        tmp2 = torch.full((G, G, G), -1, dtype=torch.int64, device=dev)
        tmp2[ci[:, 0], ci[:, 1], ci[:, 2]] = torch.arange(cand.shape[0], device=dev)
        nb = (ci[:, None, :] + off[None]) % G
        j = tmp2[nb[..., 0], nb[..., 1], nb[..., 2]]
        d = cand[:, None, :] - cand[j.clamp(min=0)]
        d -= side * torch.round(d / side)
        me = torch.arange(cand.shape[0], device=dev)[:, None]
        clash = ((d * d).sum(-1) < hard_core ** 2) & (j >= 0) & (j < me)
        keep = ~clash.any(1)
        cand, ci = cand[keep], ci[keep]
        take = min(cand.shape[0], n - count)
        cand, ci = cand[:take], ci[:take]
        pos[count:count + take] = cand
        grid[ci[:, 0], ci[:, 1], ci[:, 2]] = torch.arange(count, count + take, device=dev)
        count += take
    pos = pos[:count]
    n0 = int(round(fractions[0] * count))
    perm = torch.randperm(count, generator=g, device=dev)
    counts = np.array([n0, count - n0], np.int64)
    return pos[perm], counts


def amorphous_trajectory(side: float, T: int, sigma: float = 0.05, seed: int = 0, device="cpu", **kw):
    """T frames of an amorphous box: the RSA configuration plus independent
    N(0, sigma) Angstrom displacements per frame.  Float32 [T, n, 4] (Angstrom)
    and the species counts."""
    import torch

    base, counts = amorphous_box(side, seed=seed, device=device, **kw)
    g = torch.Generator(device=base.device)
    g.manual_seed(int(seed) + 7)
    out = torch.zeros((T, base.shape[0], 4), dtype=torch.float32, device=base.device)
    for t in range(T):
        out[t, :, :3] = (base + sigma * torch.randn(base.shape, generator=g, dtype=torch.float64,
                                                    device=base.device)).to(torch.float32)
    return out, counts


BATCH_MODES = ((32, 32, 32), (64, 64, 32), (64, 64, 64), (128, 128, 32), (128, 128, 64), (256, 256, 32))


def batch_jobs(n_jobs: int = 64, seed: int = 20261206, device="cpu", eps: float = 1e-6,
               perovskite_sides=(8, 48), amorphous_sides=(30.0, 120.0), frames=(1, 64), dt_ps: float = 0.01,
               modes=BATCH_MODES):
    """The batched-jobs workload: half perovskite supercells (side drawn in
    perovskite_sides, phonon trajectories as in the S(Q,w) benchmark, neutron
    b of Ba/Zr/O), half amorphous binary boxes (side in amorphous_sides,
    density 0.06 / A^3, 92/8 % species with b of Al/U, 2.2 A hard core);
    frames and mode boxes drawn per job.  Returns a list of
    (engine kwargs, trajectory, reduction kwargs)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n_jobs):
        T = int(rng.integers(frames[0], frames[1] + 1))
        N = tuple(int(v) for v in modes[int(rng.integers(0, len(modes)))])
        if k % 2 == 0:
            side = int(rng.integers(perovskite_sides[0], perovskite_sides[1] + 1))
            pos, counts = phonon_trajectory(side, T, dt_ps=dt_ps, n_phonons=64, rms=0.06, noise=0.01,
                                            seed=seed + k, device=device)
            kw = dict(dim=3, n_modes=N, species_counts=counts, eps=eps, precision="fp32",
                      weights=0.1 * B_NEUTRON_FM, tpe=(2 * np.pi / side,) * 3)
        else:
            side = float(rng.uniform(amorphous_sides[0], amorphous_sides[1]))
            pos, counts = amorphous_trajectory(side, T, seed=seed + k, device=device)
            kw = dict(dim=3, n_modes=N, species_counts=counts, eps=eps, precision="fp32",
                      weights=0.1 * B_AL_U_FM, tpe=(2 * np.pi / side,) * 3)
        nat = float(np.sum(counts))
        red = dict(n_int=T, win_kind=0, dt=dt_ps, nat=nat, m_lo=0, n_f=max(1, T // 2))
        jobs.append((kw, pos, red))
    return jobs
