"""mp_sql on the GPU: the |Q| histogram kernel against numpy, and the whole
path against the reference's own golden tables.

Golden: examples/amorphous_Al92U08/Al_U_large_sql_0001_40..43.txt, computed
by the reference's mp_sql (FINUFFT 2D slices at eps 1e-6 stacked over k3, FP32
amplitudes and histograms) from the snapshot in tests/golden/al_u, with X-ray
weights, Q range 20 1/A, q_step 0.01, smoothing FWHM 5 steps.  Both sides
carry the NUFFT's eps-level error and the reference its FP32 sums, so the
comparison is to a relative l2 tolerance per column (stated below), not bit
for bit.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_sql_hist_kernel_vs_numpy(gpu_lib):
    import ctypes as C
    import torch
    from mp_tools_amd import _lib
    rng = np.random.default_rng(4)
    N = (15, 13, 11)
    nsp = 3
    nq = int(np.prod(N))
    a = (rng.standard_normal((nsp, nq)) + 1j * rng.standard_normal((nsp, nq))).astype(np.complex128)
    dq = np.array([0.31, 0.27, 0.4], np.float32)
    ainv = np.array([[1.0, 0.1, 0.0], [0.0, 0.9, 0.2], [0.05, 0.0, 1.1]], np.float32)
    q_range, q_step, n_hist = np.float32(2.5), np.float32(0.05), 51
    hist = torch.zeros(n_hist * 6, dtype=torch.float64, device="cuda")
    da = torch.tensor(a, device="cuda")
    rc = gpu_lib.mpsq_sql_hist(C.c_void_p(da.data_ptr()), nsp, 1, (C.c_int64 * 3)(*N), (C.c_float * 3)(*dq),
                               (C.c_float * 9)(*ainv.reshape(-1)), float(q_range), float(q_step), n_hist,
                               C.c_void_p(hist.data_ptr()), None)
    assert rc == 0
    torch.cuda.synchronize()
    got = hist.cpu().numpy().reshape(n_hist, 6)
    # numpy: the same FP32 |Q| and bin, FP64 products
    i3, i2, i1 = np.meshgrid(*[np.arange(n) for n in N[::-1]], indexing="ij")
    u = [((i - n // 2).astype(np.float32) * dq[d]).ravel() for d, (i, n) in enumerate(zip((i1, i2, i3), N))]
    v = [(ainv[r, 0] * u[0] + ainv[r, 1] * u[1]).astype(np.float32) + ainv[r, 2] * u[2] for r in range(3)]
    q = np.sqrt((v[0] * v[0] + v[1] * v[1]).astype(np.float32) + v[2] * v[2]).astype(np.float32)
    ok = q <= q_range
    b = np.floor(q / q_step + np.float32(0.5)).astype(np.int64)
    ref = np.zeros((n_hist, 6))
    p = 0
    for s in range(nsp):
        for t in range(s + 1):
            np.add.at(ref[:, p], b[ok], (a[s] * np.conj(a[t])).real[ok])
            p += 1
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()
    # 64 KB of bins x pairs is the limit
    rc = gpu_lib.mpsq_sql_hist(C.c_void_p(da.data_ptr()), nsp, 1, (C.c_int64 * 3)(*N), (C.c_float * 3)(*dq),
                               (C.c_float * 9)(*ainv.reshape(-1)), 100.0, 0.001, 100001,
                               C.c_void_p(hist.data_ptr()), None)
    assert rc == _lib.ERR_SIZE


def _al_u(precision):
    from mp_tools_amd import mpdat, sql
    s = mpdat.read_dat(os.path.join(HERE, "golden", "al_u", "Al_U_large_n0001.dat"))
    g = sql.geometry(s.a_par, s.a_cell_inv, s.header["N_ROW"], s.n_tot, 20.0, 0.01)
    tm = {}
    h = sql.histogram([s.at_pos], s.species_counts, g, eps=1e-6, precision=precision, timing=tm)
    print(f"mp_sql Al/U 509^3 {precision}: {tm['ms']:.1f} ms per snapshot (2 species)")
    return s, g, h


@pytest.fixture(scope="module")
def al_u_hist():
    return _al_u("fp64")


@pytest.fixture(scope="module")
def al_u_hist32():
    """The FP32 form: FP32 NUFFT (the reference's own amplitudes are REAL(4)
    complex), FP64 histogram sums."""
    return _al_u("fp32")


# relative l2 (over the written rows) of each column against the reference
GOLDEN_TOL = 2e-3


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("key,mode", [("Z", 5), ("I", 6), ("S", 4), ("Iu", 7)])
def test_al_u_against_reference_tables(gpu_lib, request, precision, key, mode):
    from mp_tools_amd import sql
    s, g, h = request.getfixturevalue("al_u_hist" if precision == "fp64" else "al_u_hist32")
    gold = np.load(os.path.join(HERE, "golden", "al_u", "al_u_sql_golden.npz"))
    ref = gold[key]
    q, tot, parts = sql.output(h, g, mode, s.occupancies, "xray", ffpar=gold["ffpar"],
                               smooth_fwhm=int(gold["smooth_fwhm"]))
    got = parts if tot is None else np.column_stack([tot, parts])
    assert got.shape == ref[:, 1:].shape
    np.testing.assert_allclose(q, ref[:, 0], rtol=1e-6)
    for c in range(got.shape[1]):
        err = np.linalg.norm(got[:, c] - ref[:, c + 1]) / np.linalg.norm(ref[:, c + 1])
        print(f"{key} column {c}: rel l2 vs reference table {err:.2e}")
        assert err < GOLDEN_TOL
    # the whole text file from these columns: the reference's header records
    # and as many data records, each parsing back to the REAL(4) values
    import json
    text = json.load(open(os.path.join(HERE, "golden", "al_u", "al_u_sql_text.json"), encoding="utf-8"))[key]
    hdr = text["header"]
    amp = sql.header_amplitudes("xray", s.occupancies, ffpar=gold["ffpar"])
    out = sql.table_text(q, tot, parts, mode, subst_name="Al_U", time_stamp=" ".join(hdr[0].split()[-2:]),
                         file_dat_t0="./data/Al_U_large_n0001.dat", file_dat="./data/Al_U_large_n0001.dat",
                         proc_num=2, n_row=s.header["N_ROW"], a_par=s.a_par, at_names=s.species_names,
                         occ=s.occupancies, amplitudes=amp, q_step=0.01,
                         smooth_fwhm=int(gold["smooth_fwhm"])).split("\n")
    assert out[0].split() == hdr[0].split() and out[1:len(hdr)] == hdr[1:]
    rows = out[len(hdr):-1]
    assert len(rows) == text["n_rows"]
    back = np.array([[float(v) for v in r.split()] for r in rows])
    np.testing.assert_array_equal(back[:, 1:].astype(np.float32), got.astype(np.float32))


def test_al_u_monte_carlo_against_gauss(gpu_lib, al_u_hist):
    """The Monte Carlo integration (mp_sql56.f90:1146-1215, normalised at
    :1260-1265) of the same frame estimates the same shell averages as the
    Gauss sum: times (n_int - skips) / n_int (the reference leaves the
    skipped samples out of its count) and h dq / q_range (the sampled
    radius), the diagonal partials agree within 1% over 1/A windows and the
    cross partial within 1% of sqrt(S_11 S_22).  16 x 10^6 samples put about
    8 x 10^5 in each window: a sampling error near 0.1%."""
    from mp_tools_amd import sql
    s, g, h = al_u_hist
    n_h = 16
    assert n_h <= sql.n_h_max(g)
    tm = {}
    m = sql.histogram([s.at_pos], s.species_counts, g, eps=1e-6, precision="fp64", integration="mc", n_h=n_h,
                      seed=2024, timing=tm)
    n_int = n_h * sql.N_MC
    sk = tm["skips"]
    print(f"mp_sql MC Al/U: {tm['ms']:.1f} ms per snapshot, skips {sk}")
    assert 0 < sk[0] + sk[1] < 0.05 * n_int
    hd = np.float32(g.n_qq[0] // 2) * g.dq
    q_max = [float(np.linalg.norm(g.a_cell_inv.astype(np.float64) @ (np.eye(3)[d] * hd[d]))) for d in range(3)]
    assert max(q_max) / min(q_max) < 1.01
    m = m * (n_int - sk[0] - sk[1]) / n_int * (q_max[0] / float(g.q_range))
    for lo in range(200, 1800, 100):
        w = slice(lo, lo + 100)
        a00, a11 = h[0, 0, w].mean(), h[1, 1, w].mean()
        assert abs(m[0, 0, w].mean() / a00 - 1) < 0.01
        assert abs(m[1, 1, w].mean() / a11 - 1) < 0.01
        assert abs(m[1, 0, w].mean() - h[1, 0, w].mean()) < 0.01 * np.sqrt(a00 * a11)


@pytest.mark.gpu
def test_load_trajectory_to_device_and_engine(gpu_lib, tmp_path):
    """Snapshots -> one device tensor [T, n_tot, 4] -> the S(Q,w) engine's
    amplitudes, against the same frames passed from host arrays."""
    import torch
    from mp_tools_amd import mpdat
    from mp_tools_amd.sqw import SqwEngine
    src = os.path.join(HERE, "golden", "al_u", "Al_U_large_n0001.dat")
    s = mpdat.read_dat(src)
    rng = np.random.default_rng(5)
    paths, frames = [], []
    for k in range(3):
        pos = s.at_pos.copy()
        pos[:, :3] += rng.normal(0, 0.05, (s.n_tot, 3)).astype(np.float32)
        snap = mpdat.MPSnapshot(version=s.version, n_head=s.n_head, header=s.header, at_ind=s.at_ind,
                                at_pos=pos, time_stamp=s.time_stamp, groups=s.groups)
        p = str(tmp_path / f"t{k}.dat")
        mpdat.write_dat(p, snap)
        paths.append(p)
        frames.append(pos)
    traj, first = mpdat.load_trajectory(paths, device="cuda")
    assert traj.is_cuda and tuple(traj.shape) == (3, s.n_tot, 4)
    ref = torch.from_numpy(np.stack(frames)).cuda()
    assert torch.equal(traj, ref)
    L = float(s.a_par[0])
    eng = SqwEngine(3, (32, 32, 8), s.species_counts, eps=1e-6, precision="fp32",
                    weights=[1.0, 1.0], tpe=(2 * np.pi / L,) * 3)
    a = eng.amplitudes(traj)
    b = eng.amplitudes(ref)
    torch.cuda.synchronize()
    assert torch.equal(a, b) and torch.isfinite(torch.view_as_real(a)).all()
    eng.destroy()
