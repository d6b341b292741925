"""GPU parity of the S(Q,w) engine against the FP64 oracle.

Amplitudes: the oracle stages each species (a3, mp_sqom56.f90:1115-1118),
runs the FP64 direct DFT (dirft2d1's definition) and combines the species
with their weights (mp_sqom56.f90:1338-1356).
Reduction: the oracle's direct windowed DFT (mp_sqom56.f90:1437-1478; the
E(Q) branch :1480-1517 gives the same |X|) with the normalisation of
:1527 / :1576.

Tolerances (relative l2): FP64 form at eps 1e-10: 1e-8; FP32 form at
eps 1e-6: 1e-5 on S(Q,w) (the accuracy bar of the FP32 form), 2e-5 on amplitudes.
"""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _traj(ncell, T, seed=0):
    from mp_tools_amd import synth
    pos, counts = synth.phonon_trajectory(ncell, T, dt_ps=0.01, n_phonons=8, seed=seed,
                                          device="cpu")
    return pos.numpy(), counts


def _oracle_amps(pos, counts, dim, nm, q0, weights, axes=None, tpe=None, window=False,
                 sel_fn=None):
    """F_j(t,Q) for every species by staging + direct DFT (FP64)."""
    T = pos.shape[0]
    ncell = round((counts[0]) ** (1 / 3))
    axes = np.eye(3) if axes is None else axes
    tpe = (2 * np.pi / ncell,) * 3 if tpe is None else tpe
    off = np.concatenate([[0], np.cumsum(counts)])
    F = np.zeros((T, len(counts), int(np.prod(nm))), np.complex128)
    for t in range(T):
        for j in range(len(counts)):
            block = pos[t, off[j]:off[j + 1]]
            sel = sel_fn(block) if sel_fn else None
            x, y, z, c = O.stage(block, sel, (0, 0, 0), axes, tpe, q0, window, 1.0, dim=dim)
            F[t, j] = O.dirft_type1(x, y if dim >= 2 else None, z if dim >= 3 else None, c,
                                    nm).ravel()
    return F


@pytest.mark.parametrize("precision,eps,tol", [("fp64", 1e-10, 1e-8), ("fp32", 1e-6, 2e-5)])
@pytest.mark.parametrize("dim,nm", [(3, (16, 12, 8)), (2, (24, 20))])
def test_amplitudes_vs_oracle(gpu_lib, precision, eps, tol, dim, nm):
    import torch
    from mp_tools_amd.sqw import SqwEngine
    ncell, T = 5, 6
    pos, counts = _traj(ncell, T, seed=1)
    w = np.array([0.507, 0.716, 0.580])
    q0 = (0.5, 0.0, 0.25)
    eng = SqwEngine(dim, nm, counts, eps=eps, precision=precision, weights=w,
                    tpe=(2 * np.pi / ncell,) * 3, q0=q0, batch_frames=4)
    dpos = torch.tensor(pos, device="cuda")
    amp = eng.amplitudes(dpos)
    torch.cuda.synchronize()
    F = _oracle_amps(pos, counts, dim, nm, q0, w)
    ref = (F * w[None, :, None]).sum(axis=1)
    assert rel_l2(amp.cpu().numpy(), ref) < tol
    # per-species output (no combine) with one species masked off
    eng2 = SqwEngine(dim, nm, counts, eps=eps, precision=precision, mask=[1, 0, 1],
                     combine=False, tpe=(2 * np.pi / ncell,) * 3, q0=q0)
    a2 = eng2.amplitudes(dpos)
    torch.cuda.synchronize()
    assert a2.shape == (T, 2, int(np.prod(nm)))
    assert rel_l2(a2.cpu().numpy(), F[:, [0, 2]]) < tol


def test_selection_window_and_table_weights(gpu_lib):
    """BULK-box selection (mp_sqom56.f90:1121-1132), vacancies (:1106),
    spatial Hann window (:1118, :1167) and per-Q weight tables (:1351)."""
    import torch
    from mp_tools_amd.sqw import SqwEngine
    ncell, T, nm = 5, 3, (20, 16)
    pos, counts = _traj(ncell, T, seed=2)
    pos[:, 7, :3] = 0.0                       # a vacancy
    pmin = np.array([-2.0, -2.4, -2.5], np.float32)
    pmax = np.array([1.9, 2.0, 2.4], np.float32)
    margin = 0.1
    NQ = int(np.prod(nm))
    rng = np.random.default_rng(4)
    wtab = rng.uniform(0.5, 2.0, (3, NQ)).astype(np.float32)
    tpe = (2 * np.pi / ncell,) * 3
    eng = SqwEngine(2, nm, counts, eps=1e-10, precision="fp64", weight_table=wtab, tpe=tpe,
                    window=True, select_mode=2, pmin=pmin, pmax=pmax, margin=margin)
    amp = eng.amplitudes(torch.tensor(pos, device="cuda"))
    torch.cuda.synchronize()

    def sel_fn(block):
        s = O.select_bulk(block, pmin, pmax, margin)
        nz = np.any(block[s, :3] != 0, axis=1)
        return s[nz]

    F = _oracle_amps(pos, counts, 2, nm, (0, 0, 0), None, tpe=tpe, window=True, sel_fn=sel_fn)
    ref = (F * wtab[None].astype(np.float64)).sum(axis=1)
    assert rel_l2(amp.cpu().numpy(), ref) < 1e-8


RED_CASES = [
    # T, n_int, win_kind, m_lo, n_f, subtract_mean, qfac
    (32, 32, 0, 0, 17, True, False),     # one window over all frames: LDS FFT path
    (32, 32, 1, -8, 16, True, True),
    (20, 9, 0, 0, 5, True, False),       # sliding windows (reference default shape, odd n_int)
    (21, 10, 1, -3, 7, False, True),
    (12, 1, 0, 0, 1, True, False),       # n_int = 1: window 1
    (53, 53, 0, 0, 27, True, True),      # one window, T without a 2,3,5 radix plan: DFT recurrence
    (37, 37, 1, -5, 12, False, False),
    (512, 512, 0, 0, 257, True, False),   # one window over 512 frames (the C2 shape)
    (256, 256, 1, -100, 200, False, True),  # 256 frames, negative m_lo, per-Q factor
    # long single windows: the LDS time FFT at its largest sizes (FP64 2048 / 4096,
    # FP32 up to 8192 frames; FP64 8192 takes the sliding path)
    (2048, 2048, 0, 0, 64, True, False),
    (4096, 4096, 1, -5, 40, True, True),
    (8192, 8192, 0, 0, 24, True, False),
]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("T,n_int,kind,m_lo,n_f,sub,useq", RED_CASES)
def test_reduce_vs_oracle(gpu_lib, precision, T, n_int, kind, m_lo, n_f, sub, useq):
    import torch
    from mp_tools_amd.sqw import SqwEngine
    nm = (16, 12) if T < 256 else (15, 13)  # 195 Q: the last group of 16 columns is ragged
    NQ = int(np.prod(nm))
    rng = np.random.default_rng(T * 100 + n_int)
    # amplitudes with a strong time-constant part (Bragg-like) plus fluctuations
    A = rng.standard_normal((T, NQ)) + 1j * rng.standard_normal((T, NQ))
    A += 5.0 * (rng.standard_normal(NQ) + 1j * rng.standard_normal(NQ))[None]
    qf = rng.uniform(0.2, 1.5, NQ).astype(np.float32) if useq else None
    eng = SqwEngine(2, nm, [10, 10], precision=precision)
    cdt = torch.complex128 if precision == "fp64" else torch.complex64
    if precision == "fp32":
        A = A.astype(np.complex64).astype(np.complex128)
    damp = torch.tensor(A, dtype=cdt, device="cuda")
    S, St = eng.reduce(damp, n_int=n_int, win_kind=kind, dt=0.01, nat=20.0, m_lo=m_lo, n_f=n_f,
                       subtract_mean=sub,
                       qfac=None if qf is None else torch.tensor(qf, device="cuda"))
    torch.cuda.synchronize()
    Sr, Str = O.sqw_direct(A[:, None, :], n_int=n_int, win_kind=kind, dt=0.01, nat=20.0,
                           m_lo=m_lo, n_f=n_f, subtract_mean=sub,
                           qfac=None if qf is None else qf.astype(np.float64))
    tol = 1e-11 if precision == "fp64" else 2e-6
    assert rel_l2(St.cpu().numpy(), Str) < tol
    assert rel_l2(S.cpu().numpy(), Sr) < tol


@pytest.mark.parametrize("precision,eps,tol", [("fp64", 1e-10, 1e-8), ("fp32", 1e-6, 1e-5)])
def test_sqw_end_to_end(gpu_lib, precision, eps, tol):
    """Trajectory -> S(Q,w): relative l2 against the FP64 oracle pipeline."""
    import torch
    from mp_tools_amd.sqw import SqwEngine
    ncell, T, nm = 4, 16, (16, 16, 8)
    pos, counts = _traj(ncell, T, seed=3)
    w = np.array([0.507, 0.716, 0.580])
    eng = SqwEngine(3, nm, counts, eps=eps, precision=precision, weights=w,
                    tpe=(2 * np.pi / ncell,) * 3, batch_frames=5)
    amp = eng.amplitudes(torch.tensor(pos, device="cuda"))
    nat = float(counts.sum())
    S, St = eng.reduce(amp, n_int=T, dt=0.01, nat=nat, n_f=T // 2 + 1)
    torch.cuda.synchronize()
    F = _oracle_amps(pos, counts, 3, nm, (0, 0, 0), w)
    Sr, Str = O.sqw_direct(F, wsc=w, n_int=T, dt=0.01, nat=nat, n_f=T // 2 + 1)
    assert rel_l2(St.cpu().numpy(), Str) < tol
    assert rel_l2(S.cpu().numpy(), Sr) < tol


def test_map_file_end_to_end(gpu_lib, tmp_path):
    """The reference's inputs -> Q-plane set-up -> engine total map -> map file
    (mp_sqom56.f90:800-857, 1381-1387, 1574, 1994-2045): the written map parses
    back to the oracle's S_tot to the file's three digits, rows along e1."""
    import torch
    from mp_tools_amd import mptxt
    from mp_tools_amd.sqw import SqwEngine, qplane
    ncell, T = 6, 4
    pos, counts = _traj(ncell, T, seed=7)
    g = qplane((ncell,) * 3, 2, (1, 0, 0), (0, 1, 0), q_center=(1, 1, 0))
    nm = g["n_modes"]
    assert nm == (12, 12)
    w = np.array([0.507, 0.716, 0.580])
    nat = float(np.sum(counts))
    eng = SqwEngine(2, nm, counts, eps=1e-10, precision="fp64", weights=w, **g["engine_kw"])
    amp = eng.amplitudes(torch.tensor(pos, device="cuda"))
    S, St = eng.reduce(amp, n_int=T, dt=0.01, nat=nat, n_f=T // 2 + 1)
    torch.cuda.synchronize()
    F = _oracle_amps(pos, counts, 2, nm, g["engine_kw"]["q0"], w, axes=g["engine_kw"]["axes"],
                     tpe=g["engine_kw"]["tpe"])
    C = (F * w[None, :, None]).sum(axis=1)
    st_ref = (np.abs(C) ** 2).sum(axis=0) / (T * nat * nat)
    assert rel_l2(St.cpu().numpy(), st_ref) < 1e-8
    cs = mptxt.cs_plot_total(St.cpu().numpy(), *nm)
    path = str(tmp_path / "map.txt")
    mptxt.write_sqom_table(path, cs, n_row=(ncell,) * 3, at_names=["A", "B", "O"], at_occup=[1, 1, 3],
                           b_coh=w, q_center=(1, 1, 0), qx_range=g["qx_range"], qy_range=g["qy_range"])
    lines = open(path).read().split("\n")
    assert lines[15].split()[-3:] == ["12", "0.00000000", "1.83333337"]  # 12 pixels from Q = 0, 1/6 apart
    back = np.array([[float(v) for v in r.split()] for r in lines[18:18 + nm[0]]])
    ref = mptxt.cs_plot_total(st_ref, *nm)
    np.testing.assert_allclose(back, ref, rtol=5e-3, atol=1e-6 * ref.max())
