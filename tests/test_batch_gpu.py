"""Batched uneven S(Q,w) jobs (mpsq_batch_*): every job of a batch --
perovskite supercells and amorphous binary boxes of different sizes, frame
counts and mode boxes, all in one launch per phase -- must give exactly the
amplitudes and maps it gives when run alone (the spread accumulates in fixed
point, so the order in which tiles and atomics run does not matter), in one
round or in several, eagerly or replayed from a captured hipGraph; one job
is also checked against the FP64 direct DFT."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_jobs():
    from mp_tools_amd import synth
    return synth.batch_jobs(n_jobs=6, seed=7, device="cuda", perovskite_sides=(3, 5),
                            amorphous_sides=(14.0, 22.0), frames=(1, 6),
                            modes=((16, 16, 8), (24, 20, 12), (32, 32, 16)))


def _alone(kw, pos, red):
    from mp_tools_amd.sqw import SqwEngine
    kw = dict(kw)
    eng = SqwEngine(kw.pop("dim"), kw.pop("n_modes"), kw.pop("species_counts"), **kw)
    amp = eng.amplitudes(pos)
    S, St = eng.reduce(amp, **red)
    return amp, S, St


def test_batch_equals_alone(gpu_lib, small_jobs):
    import torch
    from mp_tools_amd.sqw import SqwBatch
    b = SqwBatch([kw for kw, _, _ in small_jobs])
    amps, S, St = b.run([p for _, p, _ in small_jobs], [r for _, _, r in small_jobs])
    torch.cuda.synchronize()
    for j, (kw, pos, red) in enumerate(small_jobs):
        a1, s1, t1 = _alone(kw, pos, red)
        torch.cuda.synchronize()
        assert torch.equal(amps[j], a1), f"job {j} amplitudes differ from the job run alone"
        assert torch.equal(S[j], s1) and torch.equal(St[j], t1), f"job {j} maps differ"


def test_batch_rounds_and_graph(gpu_lib, small_jobs):
    import torch
    from mp_tools_amd.sqw import SqwBatch
    pos = [p for _, p, _ in small_jobs]
    reds = [r for _, _, r in small_jobs]
    ref = SqwBatch([kw for kw, _, _ in small_jobs]).run(pos, reds)
    small = SqwBatch([kw for kw, _, _ in small_jobs], budget_bytes=1 << 22)   # several rounds
    got = small.run(pos, reds)
    g = SqwBatch([kw for kw, _, _ in small_jobs])
    out = g.run(pos, reds, use_graph=True)          # eager, allocates
    for _ in range(3):                               # capture, then replays
        for a in out[0]:
            a.zero_()
        g.run(pos, reds, out=out, use_graph=True)
        torch.cuda.synchronize()
        for j in range(len(pos)):
            assert torch.equal(out[0][j], ref[0][j]) and torch.equal(out[1][j], ref[1][j])
    torch.cuda.synchronize()
    for j in range(len(pos)):
        assert torch.equal(got[0][j], ref[0][j]) and torch.equal(got[1][j], ref[1][j])


def test_batch_job_vs_direct_dft(gpu_lib, small_jobs):
    import torch
    from mp_tools_amd.sqw import SqwBatch
    kw, pos, red = small_jobs[1]                     # an amorphous box
    amps, _, _ = SqwBatch([small_jobs[0][0], kw]).run([small_jobs[0][1], pos])
    torch.cuda.synchronize()
    N = kw["n_modes"]
    counts = kw["species_counts"]
    off = np.concatenate([[0], np.cumsum(counts)])
    pn = pos.cpu().numpy()
    ref = np.zeros((pn.shape[0], int(np.prod(N))), np.complex128)
    for t in range(pn.shape[0]):
        for s in range(len(counts)):
            x, y, z, c = O.stage(pn[t, off[s]:off[s + 1]], None, (0, 0, 0), np.eye(3), kw["tpe"],
                                 (0, 0, 0), 0, 1.0)
            ref[t] += kw["weights"][s] * O.dirft_type1(x, y, z, c, N).ravel()
    assert rel_l2(amps[1].cpu().numpy(), ref) < 2e-5


def test_autotune_tiles(gpu_lib, small_jobs):
    import torch
    from mp_tools_amd.sqw import SqwEngine, autotune_tiles
    kw, pos, red = small_jobs[0]
    kw = dict(kw)
    by, bz, ms = autotune_tiles(pos, kw["dim"], kw["n_modes"], kw["species_counts"],
                                **{k: v for k, v in kw.items() if k not in ("dim", "n_modes", "species_counts")})
    assert by in (1, 2, 4, 8) and bz in (1, 2, 4, 8) and ms > 0
    a0, _, _ = _alone(dict(kw, tile_by=1, tile_bz=1), pos, red)
    a1, _, _ = _alone(kw, pos, red)                  # uses the tuned shape
    torch.cuda.synchronize()
    assert rel_l2(a1.cpu().numpy(), a0.cpu().numpy()) < 1e-6
