"""Synthetic workload generators (CPU): the amorphous boxes keep the 2.2 A
hard core under the minimum image and reach the requested density; the
species split and the batched-jobs list follow their definitions."""
import numpy as np


def test_amorphous_box_hard_core_and_density():
    import torch
    from mp_tools_amd import synth
    side = 24.0
    pos, counts = synth.amorphous_box(side, seed=3)
    n = pos.shape[0]
    assert n == round(0.06 * side ** 3)
    assert counts.sum() == n and counts[0] == round(0.92 * n)
    d = pos[:, None, :] - pos[None]
    d -= side * torch.round(d / side)
    r = (d * d).sum(-1).sqrt() + torch.eye(n, dtype=torch.float64) * 1e9
    assert r.min().item() >= 2.2
    assert pos.abs().max().item() <= side / 2


def test_batch_jobs_definition():
    from mp_tools_amd import synth
    jobs = synth.batch_jobs(n_jobs=4, seed=1, perovskite_sides=(2, 3), amorphous_sides=(12.0, 14.0),
                            frames=(1, 3), modes=((8, 8, 8),))
    assert len(jobs) == 4
    for k, (kw, pos, red) in enumerate(jobs):
        assert pos.shape[0] == red["n_int"] and 1 <= pos.shape[0] <= 3
        assert int(np.sum(kw["species_counts"])) == pos.shape[1]
        assert len(kw["species_counts"]) == (3 if k % 2 == 0 else 2)
