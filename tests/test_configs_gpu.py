"""Parity at the benchmark configurations that are not the bench line.

C1: one static frame of a 20^3 perovskite (40,000 atoms, 3 species, neutron
    weights), type-1 NUFFT to 128x128x16 modes at eps 1e-6 in FP64, elastic
    map |A(Q)|^2 / N^2 -- checked against the FP64 direct DFT of the staged
    atoms (the definition of dirft2d1, code/src/mp_sqom56.f90:2530-2602) at
    2,000 seeded modes.
C3: 128^3 perovskite (10,485,760 atoms), 512x512x64 modes, eps in
    {1e-3, 1e-4, 1e-6, 1e-9}, FP32 and FP64 forms (FP32 floors eps at 1e-7),
    checked against the FP64 direct DFT at 10,000 seeded random modes (the
    product set 25 x 25 x 16 of tests/golden/c3_dft_10k.npz, written by
    tests/golden/make_c3_dft.py from the same seeded frame).

Tolerances (relative l2 over the sampled modes): FP64 form 10 * eps;
FP32 form max(10 * eps, 5e-5).  C1 is a nearly perfect lattice: its
amplitudes away from the Bragg points are ~1e-3 of sum |c|, so the relative
l2 error of the method itself (FINUFFT's, whose target is on the absolute
error, |f~ - f| ~ eps * sum_j |c_j|) is ~2.6e-5 at eps = 1e-6 there.  C1
is therefore checked against that bound, against the host restatement of
the same method at the same eps (the transform the reference's finufft2d1
call computes) to 1e-9, and at eps = 1e-9 also to 1e-7 relative l2 against
the direct DFT.
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


@pytest.mark.parametrize("eps", [1e-6, 1e-9])
def test_c1_static_elastic_map(gpu_lib, eps):
    import torch
    from mp_tools_amd import synth
    from mp_tools_amd.sqw import SqwEngine
    ncell, N = 20, (128, 128, 16)
    pos, counts = synth.static_frame(ncell, sigma=0.05, seed=20261203)
    w = 0.1 * synth.B_NEUTRON_FM
    eng = SqwEngine(3, N, counts, eps=eps, precision="fp64", weights=w,
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
    # the method's absolute error against the direct DFT: eps is FINUFFT's
    # target for max|f~ - f| / sum|c|, reached within 1.1-1.8x on this input
    assert np.abs(got - ref).max() <= 3 * eps * np.abs(c).sum()
    # the same method on the host at the same eps
    meth = _pick(O.cpu_nufft_type1(x, y, z, c, N, eps=eps), N, k1, k2, k3)
    err_m = rel_l2(got, meth)
    err_d = rel_l2(got, ref)
    print(f"C1 eps={eps:g}: rel l2 vs direct DFT {err_d:.3e}, vs host method {err_m:.3e}")
    assert err_m < 1e-9
    assert err_d < (5e-5 if eps >= 1e-6 else 1e-7)
    st = _pick(St.cpu().numpy(), N, k1, k2, k3)
    assert rel_l2(st, np.abs(meth) ** 2 / nat ** 2) < 2e-9
    assert rel_l2(st, np.abs(ref) ** 2 / nat ** 2) < (1e-4 if eps >= 1e-6 else 2e-7)


@pytest.mark.parametrize("precision,eps", [("fp64", 1e-9), ("fp64", 1e-6), ("fp64", 1e-4), ("fp64", 1e-3),
                                           ("fp32", 1e-9), ("fp32", 1e-6), ("fp32", 1e-4), ("fp32", 1e-3)])
def test_c3_precision_sweep(gpu_lib, precision, eps, c3_frame):
    import torch
    from mp_tools_amd.sqw import SqwEngine
    counts, ncell, w, dpos, F, sets = c3_frame
    N = (512, 512, 64)
    eng = SqwEngine(3, N, counts, eps=eps, precision=precision, weights=w,
                    tpe=(2 * np.pi / ncell,) * 3)
    amp = eng.amplitudes(dpos)
    torch.cuda.synchronize()
    idx = torch.tensor(_flat_index(N, *sets), device="cuda")
    got = amp[0, idx].cpu().numpy()
    err = rel_l2(got, F.ravel())
    print(f"C3 {precision} eps={eps:g}: rel l2 vs direct DFT at {F.size} modes = {err:.3e}")
    tol = 10 * eps if precision == "fp64" else max(10 * eps, 5e-5)
    assert err < tol
    del amp, eng
    torch.cuda.empty_cache()


def _flat_index(N, k1s, k2s, k3s):
    """CMCL flat indices of the product set, in its [k3][k2][k1] order."""
    K3, K2, K1 = np.meshgrid(k3s, k2s, k1s, indexing="ij")
    return ((K3 + N[2] // 2) * N[1] * N[0] + (K2 + N[1] // 2) * N[0] + (K1 + N[0] // 2)).ravel()


@pytest.fixture(scope="module")
def c3_frame():
    import os
    import torch
    from mp_tools_amd import synth
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "c3_dft_10k.npz"))
    ncell = int(g["ncell"])
    pos, counts = synth.static_frame(ncell, sigma=0.05, seed=int(g["seed"]))
    p = pos[0, :, :3].astype(np.float64)
    assert np.allclose([p.sum(), (p * p).sum(), p[::9973].sum()], g["checksums"], rtol=1e-12), \
        "the seeded frame differs from the one the fixture was made from"
    dpos = torch.tensor(pos, device="cuda")
    return counts, ncell, g["weights"], dpos, g["F"], (g["k1s"], g["k2s"], g["k3s"])
