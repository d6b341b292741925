"""Parity at the benchmark configurations that are not the bench line.

C1: one static frame of a 20^3 perovskite (40,000 atoms, 3 species, neutron
    weights), type-1 NUFFT to 128x128x16 modes at eps 1e-6 in FP64, elastic
    map |A(Q)|^2 / N^2 -- checked against the FP64 direct DFT of the staged
    atoms (the definition of dirft2d1, code/src/mp_sqom56.f90:2530-2602) at
    2,000 seeded modes.
C3: 128^3 perovskite (10,485,760 atoms), 512x512x64 modes, eps in
    {1e-3, 1e-4, 1e-6} (FP32 and FP64 forms) and 1e-9 (FP64), checked
    against the FP64 direct DFT at seeded random modes (120 here to keep the
    CPU oracle to seconds; bench-scale runs use more).

Tolerances (relative l2 over the sampled modes): FP64 form 10 * eps;
FP32 form max(10 * eps, 5e-5).
"""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _staged(pos, counts, ncell, w):
    """FP64 staged coordinates of every atom and the combined strengths."""
    off = np.concatenate([[0], np.cumsum(counts)])
    xs, ys, zs, cs = [], [], [], []
    for j in range(len(counts)):
        x, y, z, c = O.stage(pos[0, off[j]:off[j + 1]], None, (0, 0, 0), np.eye(3),
                             (2 * np.pi / ncell,) * 3, (0, 0, 0), 0, 1.0)
        xs.append(x)
        ys.append(y)
        zs.append(z)
        cs.append(c * w[j])
    return [np.concatenate(v) for v in (xs, ys, zs, cs)]


def _modes(rng, N, K):
    return [rng.integers(-(n // 2), (n - 1) // 2 + 1, K) for n in N]


def _pick(amp, N, k1, k2, k3):
    a = amp.reshape(N[2], N[1], N[0])
    return a[k3 + N[2] // 2, k2 + N[1] // 2, k1 + N[0] // 2]


def test_c1_static_elastic_map(gpu_lib):
    import torch
    from mp_tools_amd import synth
    from mp_tools_amd.sqw import SqwEngine
    ncell, N = 20, (128, 128, 16)
    pos, counts = synth.static_frame(ncell, sigma=0.05, seed=20261203)
    w = 0.1 * synth.B_NEUTRON_FM
    eng = SqwEngine(3, N, counts, eps=1e-6, precision="fp64", weights=w,
                    tpe=(2 * np.pi / ncell,) * 3)
    amp = eng.amplitudes(torch.tensor(pos, device="cuda"))
    nat = float(counts.sum())
    _, St = eng.reduce(amp, n_int=1, nat=nat, n_f=0)
    torch.cuda.synchronize()
    rng = np.random.default_rng(1)
    k1, k2, k3 = _modes(rng, N, 2000)
    x, y, z, c = _staged(pos, counts, ncell, w)
    ref = O.dirft_type1_sampled(x, y, z, c, k1, k2, k3)
    got = _pick(amp.cpu().numpy()[0], N, k1, k2, k3)
    assert rel_l2(got, ref) < 1e-5
    st = _pick(St.cpu().numpy(), N, k1, k2, k3)
    assert rel_l2(st, np.abs(ref) ** 2 / nat ** 2) < 2e-5


@pytest.mark.parametrize("precision,eps", [("fp64", 1e-9), ("fp64", 1e-6), ("fp64", 1e-4), ("fp64", 1e-3),
                                           ("fp32", 1e-6), ("fp32", 1e-4), ("fp32", 1e-3)])
def test_c3_precision_sweep(gpu_lib, precision, eps, c3_frame):
    import torch
    from mp_tools_amd.sqw import SqwEngine
    pos, counts, ncell, w, dpos, ref, modes = c3_frame
    N = (512, 512, 64)
    eng = SqwEngine(3, N, counts, eps=eps, precision=precision, weights=w,
                    tpe=(2 * np.pi / ncell,) * 3)
    amp = eng.amplitudes(dpos)
    torch.cuda.synchronize()
    got = _pick(amp.cpu().numpy()[0], N, *modes)
    err = rel_l2(got, ref)
    print(f"C3 {precision} eps={eps:g}: rel l2 vs direct DFT at {len(ref)} modes = {err:.3e}")
    tol = 10 * eps if precision == "fp64" else max(10 * eps, 5e-5)
    assert err < tol
    del amp, eng
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def c3_frame():
    import torch
    from mp_tools_amd import synth
    ncell = 128
    pos, counts = synth.static_frame(ncell, sigma=0.05, seed=20261205)
    w = 0.1 * synth.B_NEUTRON_FM
    rng = np.random.default_rng(3)
    modes = _modes(rng, (512, 512, 64), 120)
    x, y, z, c = _staged(pos, counts, ncell, w)
    ref = O.dirft_type1_sampled(x, y, z, c, *modes)
    dpos = torch.tensor(pos, device="cuda")
    return pos, counts, ncell, w, dpos, ref, modes
