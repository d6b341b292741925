"""Real strengths (q_center = 0: window and weights times exp(-i 2 pi q0 . r)
are real, mp_sqom56.f90:1115-1118) take the FP32 narrow spread kernel's real
form: one 32-bit fixed-point cell per row cell and two rows per complex x FFT,
split by Hermitian symmetry.  Checked against the FP64 oracle (staging and
direct DFT) on grids that take the wave-level and the block-wide FFT of that
kernel, and against the complex form on the same input: a q_center of 1e-30
keeps the complex kernel, whose fixed-point sums are then the same integers
(the imaginary contributions round to zero), so the two differ by FP32 FFT
rounding only."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _traj(ncell, T, seed):
    from mp_tools_amd import synth
    pos, counts = synth.phonon_trajectory(ncell, T, dt_ps=0.01, n_phonons=8, seed=seed, device="cpu")
    return pos.numpy(), counts


def _oracle(pos, counts, nm, ncell, window):
    dim = len(nm)
    T = pos.shape[0]
    off = np.concatenate([[0], np.cumsum(counts)])
    F = np.zeros((T, len(counts), int(np.prod(nm))), np.complex128)
    for t in range(T):
        for j in range(len(counts)):
            x, y, z, c = O.stage(pos[t, off[j]:off[j + 1]], None, (0, 0, 0), np.eye(3),
                                 (2 * np.pi / ncell,) * 3, (0.0, 0.0, 0.0), window, 1.0, dim=dim)
            F[t, j] = O.dirft_type1(x, y if dim >= 2 else None, z if dim >= 3 else None, c, nm).ravel()
    return F


# (200, 24, 8): n1 = 400 leaves each row pair's accumulators inside its FFT
# slot (the wave-level FFT); the smaller grids take the block-wide one.  The
# FP64 form has no real spread kernel but runs the half spectrum through its
# column passes as well.
@pytest.mark.parametrize("precision,eps,tol", [("fp32", 1e-6, 2e-5), ("fp64", 1e-10, 1e-8)])
@pytest.mark.parametrize("nm", [(200, 24, 8), (16, 12, 8), (15, 13, 9), (24, 20)])
@pytest.mark.parametrize("window", [False, True])
def test_real_strengths_vs_oracle(gpu_lib, precision, eps, tol, nm, window):
    import torch
    from mp_tools_amd.sqw import SqwEngine
    ncell, T = 5, 4
    pos, counts = _traj(ncell, T, seed=len(nm) + nm[0])
    w = np.array([0.507, 0.716, 0.580])
    F = _oracle(pos, counts, nm, ncell, window)
    dpos = torch.tensor(pos, device="cuda")
    eng = SqwEngine(len(nm), nm, counts, eps=eps, precision=precision, weights=w,
                    tpe=(2 * np.pi / ncell,) * 3, window=window)
    amp = eng.amplitudes(dpos)
    torch.cuda.synchronize()
    assert rel_l2(amp.cpu().numpy(), (F * w[None, :, None]).sum(axis=1)) < tol
    eng2 = SqwEngine(len(nm), nm, counts, eps=eps, precision=precision, mask=[1, 0, 1], combine=False,
                     tpe=(2 * np.pi / ncell,) * 3, window=window)
    a2 = eng2.amplitudes(dpos)
    torch.cuda.synchronize()
    assert rel_l2(a2.cpu().numpy(), F[:, [0, 2]]) < tol


@pytest.mark.parametrize("nm", [(256, 16, 8), (20, 16, 12)])
@pytest.mark.parametrize("iflag", [-1, 1])
def test_real_form_equals_complex_form(gpu_lib, nm, iflag):
    import torch
    from mp_tools_amd.sqw import SqwEngine
    ncell, T = 6, 3
    pos, counts = _traj(ncell, T, seed=7)
    dpos = torch.tensor(pos, device="cuda")
    w = np.array([0.507, 0.716, 0.580])
    kw = dict(eps=1e-6, precision="fp32", weights=w, tpe=(2 * np.pi / ncell,) * 3, iflag=iflag, window=True)
    real = SqwEngine(3, nm, counts, q0=(0.0, 0.0, 0.0), **kw).amplitudes(dpos)
    cplx = SqwEngine(3, nm, counts, q0=(1e-30, 0.0, 0.0), **kw).amplitudes(dpos)
    torch.cuda.synchronize()
    a, b = real.cpu().numpy(), cplx.cpu().numpy()
    assert np.all(np.isfinite(a))
    assert rel_l2(a, b) < 1e-6


def test_real_dense_tiles_take_the_wide_kernel(gpu_lib):
    """Points dense enough that tiles exceed the narrow kernel's covering
    bound: those tiles (32 rows with real strengths) are queued for the
    64-bit wide kernel in 8 passes of 4 rows; the map still matches the
    oracle."""
    import torch
    from mp_tools_amd.sqw import SqwEngine
    rng = np.random.default_rng(3)
    ncell, T, n = 4, 2, 12000
    counts = np.array([n // 2, n // 2])
    pos = np.zeros((T, n, 4), np.float32)
    pos[:, :, :3] = rng.uniform(0.0, ncell, (T, n, 3))
    nm = (64, 40, 24)
    w = np.array([0.6, -0.4])
    eng = SqwEngine(3, nm, counts, eps=1e-6, precision="fp32", weights=w, tpe=(2 * np.pi / ncell,) * 3)
    amp = eng.amplitudes(torch.tensor(pos, device="cuda"))
    torch.cuda.synchronize()
    F = _oracle(pos, counts, nm, ncell, False)
    assert rel_l2(amp.cpu().numpy(), (F * w[None, :, None]).sum(axis=1)) < 2e-5


@pytest.mark.parametrize("precision,eps,tol", [("fp32", 1e-6, 2e-5), ("fp64", 1e-10, 1e-8)])
def test_real_long_rows(gpu_lib, precision, eps, tol):
    """N1 = 1024 (fine rows of 2048 cells): a real-strength plan doubles the
    complex form's 4-row tiles to 8 rows of 4-byte cells, whose LDS (tile,
    2048 twiddles, tables) fits where the complex form's 8-byte cells would
    not."""
    import torch
    from mp_tools_amd.sqw import SqwEngine
    ncell, T = 4, 2
    pos, counts = _traj(ncell, T, seed=11)
    nm = (1024, 12, 8)
    w = np.array([0.507, 0.716, 0.580])
    eng = SqwEngine(3, nm, counts, eps=eps, precision=precision, weights=w, tpe=(2 * np.pi / ncell,) * 3)
    amp = eng.amplitudes(torch.tensor(pos, device="cuda"))
    torch.cuda.synchronize()
    F = _oracle(pos, counts, nm, ncell, False)
    assert rel_l2(amp.cpu().numpy(), (F * w[None, :, None]).sum(axis=1)) < tol
