"""mp_dos vibrational spectra (mp_dos56.f90:495-589): the oracle's direct DFT
pinned by numpy's FFT, a constant velocity and Parseval; the REAL(4)
frequency count and the plotted normalisation; on the GPU, the HIP path
(the reduction kernels on packed velocity columns) against the oracle."""
import numpy as np
import pytest

from mp_tools_amd import sqw
from oracle import oracle as O


def _vel(T, n_tot, nsp, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(T)[:, None]
    v = np.zeros((T, n_tot, 4), np.float32)
    freqs = rng.uniform(0.02, 0.45, (n_tot, 3))
    for c in range(3):
        v[:, :, c] = (np.cos(2 * np.pi * freqs[None, :, c] * t) + 0.3 * rng.standard_normal((T, n_tot)))
    ind = np.zeros((n_tot, 4), np.int32)
    ind[:, 3] = rng.integers(1, nsp + 1, n_tot)
    return v, ind


def _window(T):
    return 0.5 * (1.0 - np.cos(2 * np.pi * np.arange(1, T + 1) / (T + 1)))


def test_dos_matches_numpy_fft():
    T, n_tot, nsp = 48, 30, 3
    v, ind = _vel(T, n_tot, nsp, 1)
    n_freq = T // 2 + 1
    got = O.dos(v, ind, nsp, n_freq)
    X = np.fft.fft(v[:, :, :3].astype(np.float64) * _window(T)[:, None, None], axis=0)    # exp(-2 pi i f t / T)
    P = (np.abs(X) ** 2).sum(axis=2) / T                                                    # [T, n_tot]
    want = np.zeros((n_freq, nsp))
    for s in range(nsp):
        want[:, s] = P[:n_freq, ind[:, 3] == s + 1].sum(axis=1)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_dos_closed_forms():
    T = 40
    v = np.zeros((T, 2, 4), np.float32)
    v[:, 0, 0] = 1.5                                     # constant v_x on atom 0 (type 1)
    v[:, 1, 2] = np.cos(2 * np.pi * 5 * np.arange(T) / T)  # a tone on v_z of atom 1 (type 2)
    ind = np.array([[1, 1, 1, 1], [1, 1, 1, 2]], np.int32)
    got = O.dos(v, ind, 2, T // 2 + 1)
    w = _window(T)
    assert np.isclose(got[0, 0], (1.5 * w.sum()) ** 2 / T, rtol=1e-12)
    # Parseval for real signals over the half spectrum (T even)
    for a, s in ((0, 0), (1, 1)):
        tot = (v[:, a, :3].astype(np.float64) ** 2 * (w ** 2)[:, None]).sum()
        half = got[0, s] + 2 * got[1:T // 2, s].sum() + got[T // 2, s]
        assert np.isclose(half, tot, rtol=1e-12)


def test_reference_frequency_count_and_output_norm():
    # REAL(4) n_freq_plot = int(0.5 (nfile - 1) t_step / t_step) + 1
    assert sqw.dos_n_freq(512, 0.01) == 256
    assert sqw.dos_n_freq(101, 0.02) == 51
    assert sqw.dos_n_freq(1, 0.01) == 1
    cs = np.array([[1.0, 2.0], [3.0, 4.0]])
    out = sqw.dos_output(cs, weights=[1.0, 4.0], occup=[1.0, 0.5], nsuper=[10, 10], nfile=8, t_step=0.02)
    i = np.arange(1, 9)
    ws = np.sum(0.5 * (1 - np.cos(2 * np.pi * i / 9)))
    aws = (1.0 * 1 * 10 + 0.5 * 4 * 10) * ws ** 2 * 2.0
    assert np.allclose(out[:, 0], cs[:, 0] * 8 / aws) and np.allclose(out[:, 1], cs[:, 1] * 4 * 8 / aws)
    assert np.allclose(out[:, 2], out[:, 0] + out[:, 1])


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
@pytest.mark.parametrize("T", [64, 50, 513])
def test_dos_gpu_vs_oracle(gpu_lib, precision, tol, T):
    """One window over the frames: T = 64 and 50 take the LDS time FFT;
    513 = 27 x 19 has a factor the radix plan lacks and takes the sliding
    DFT kernel.  Types from at_ind."""
    import torch
    from conftest import rel_l2
    n_tot, nsp = 37, 3
    v, ind = _vel(T, n_tot, nsp, T)
    n_freq = sqw.dos_n_freq(T, 0.01)
    ref = O.dos(v, ind, nsp, n_freq)
    got = sqw.dos_spectra(torch.tensor(v, device="cuda"), torch.tensor(ind, device="cuda"), nsp, n_freq,
                          precision=precision)
    torch.cuda.synchronize()
    assert got.shape == (n_freq, nsp)
    assert rel_l2(got.cpu().numpy(), ref) < tol


@pytest.mark.gpu
def test_dos_from_dat_velocities(gpu_lib, tmp_path):
    """Snapshots with velocity records (N_TRAJ = 1) written like mp_lbin,
    loaded to the device by their velocity block, through mpsq_dos, against
    the oracle on the same arrays."""
    import os
    import torch
    from conftest import rel_l2
    from mp_tools_amd import mpdat
    here = os.path.dirname(os.path.abspath(__file__))
    s = mpdat.read_dat(os.path.join(here, "golden", "al_u", "Al_U_large_n0001.dat"))
    n, T = 300, 24
    d1 = dict(s.groups[0][1])
    d1.update(N_TOT=n, N_REC=-(-4 * n // mpdat.L_REC), N_TRAJ=1, J_SHELL_OUT=0)
    d2 = dict(s.groups[1][1])
    d2.update(NSUPER_R=[250, 50])
    v, _ = _vel(T, n, 2, 4)
    ind = s.at_ind[:n].copy()
    ind[:, 3] = np.repeat([1, 2], [250, 50])
    paths = []
    for t in range(T):
        snap = mpdat.MPSnapshot(version=s.version, n_head=s.n_head, time_stamp="2026/01/01 00:00:00",
                                groups=[("DATA_HEADER_1", d1), ("DATA_HEADER_2", d2), s.groups[2]],
                                at_ind=ind, at_pos=np.zeros((n, 4), np.float32), at_vel=v[t])
        p = str(tmp_path / f"v{t}.dat")
        mpdat.write_dat(p, snap)
        paths.append(p)
    vel, first = mpdat.load_trajectory(paths, device="cuda", block="at_vel")
    nf = sqw.dos_n_freq(T, 0.01)
    got = sqw.dos_spectra(vel, torch.tensor(first.at_ind, device="cuda"), 2, nf)
    torch.cuda.synchronize()
    assert rel_l2(got.cpu().numpy(), O.dos(v, ind, 2, nf)) < 2e-6
