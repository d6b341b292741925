"""mp_sql's Monte Carlo integration (j_acc = 2, mp_sql56.f90:1146-1215,
normalised at :1260-1265).

The reference draws from Fortran's random_number, so its samples cannot be
replayed; the port draws from Philox4x32-10 under (seed, tag, chain, sample)
and the oracle restates the same draw.  Pins of the oracle (CPU):
  * the generator: the Random123 suite's published known-answer vectors;
  * constant amplitudes c_s: every kept sample adds exactly Re(c_i c_j*), so
    the bins hold integer counts that, with the skips, add up to n_h n_mc,
    and the fraction at d = 0 matches its closed form over uniform
    directions, P = E[min(1, 1 / (2 h max_d |u_d|))];
  * the radius law: counts are flat in |Q| (uniform radius fraction; a
    uniform-in-ball law would grow as Q^2);
  * isotropy: the sphere's direction moments <cos^2> = 1/3, <cos^4> = 1/5
    (points projected from the cube without the ball test give 0.180);
  * the normalised estimate of a smooth |A(Q)|^2 = 1 + Q^2 against its value
    and against the Gauss sum over the same mode box (numpy).
On the GPU: skip counts equal the oracle's, histograms to 1e-12 relative l2
(FP64 sums in another order), integer counts bit for bit; the Al/U frame's
MC S(Q) against its Gauss S(Q) within the sampling error."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O
from mp_tools_amd import sql


def test_philox_known_answers():
    """Random123 kat_vectors, philox4x32 with 10 rounds."""
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        assert tuple(int(v) for v in O.philox4x32_10(ctr, key)) == want


def _grid(n):
    k = np.arange(n) - n // 2
    K3, K2, K1 = np.meshgrid(k, k, k, indexing="ij")
    return K1, K2, K3


def _p_zero(h, n=2_000_000):
    u = np.random.default_rng(0).standard_normal((n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    return float(np.minimum(1.0, 0.5 / (h * np.abs(u).max(axis=1))).mean())


def test_constant_amplitudes_count_every_sample():
    N, h = (31, 31, 31), 15
    amp = np.zeros((2, 31 ** 3), complex)
    amp[0], amp[1] = 1 + 2j, 0.5 - 1j
    total = 4 * 250_000
    hist, (s1, s2) = O.sql_hist_mc(amp, N, (0.1,) * 3, np.eye(3), 2.0, 0.05, 41, 4, 250_000, seed=7)
    n = hist[:, 0] / 5.0
    assert np.array_equal(n, np.rint(n))
    assert np.array_equal(hist[:, 1], -1.5 * n) and np.array_equal(hist[:, 2], 1.25 * n)
    assert s2 == 0 and n.sum() + s1 + s2 == total
    assert abs(s1 / total - _p_zero(h)) < 0.03 * _p_zero(h)
    # a smaller q_range: samples beyond it are skip 2 and leave their bins empty
    hist2, (t1, t2) = O.sql_hist_mc(amp, N, (0.1,) * 3, np.eye(3), 1.0, 0.05, 41, 4, 250_000, seed=7)
    n2 = hist2[:, 0] / 5.0
    assert t1 == s1 and t2 > 0 and n2.sum() + t1 + t2 == total
    assert np.all(n2[21:] == 0) and np.array_equal(n2[:20], n[:20])


def test_radius_law_flat_in_q():
    amp = np.ones((1, 31 ** 3), complex)
    hist, (s1, _) = O.sql_hist_mc(amp, (31,) * 3, (0.1,) * 3, np.eye(3), 2.0, 0.05, 41, 4, 250_000, seed=5)
    # radius h r, r uniform in [0, 1): 30 bins of 0.05 up to |Q| = 1.5
    w = hist[2:30, 0].reshape(4, 7).sum(axis=1) / (1e6 * 7 / 30)
    assert np.all(np.abs(w - 1.0) < 0.1), w


def test_isotropy_direction_moments():
    K1, K2, K3 = _grid(41)
    r = np.sqrt(K1 ** 2 + K2 ** 2 + K3 ** 2).astype(float)
    r[r == 0] = 1.0
    amp = np.zeros((2, 41 ** 3), complex)
    amp[0] = 1.0
    amp[1] = ((K1 / r) ** 2).reshape(-1)                  # A_1 = cos^2 of the angle to the first axis
    hist, _ = O.sql_hist_mc(amp, (41,) * 3, (0.1,) * 3, np.eye(3), 2.0, 0.05, 41, 4, 250_000, seed=3)
    for lo, hi in [(6, 14), (14, 22), (22, 30), (30, 38)]:
        cnt = hist[lo:hi, 0].sum()
        assert abs(hist[lo:hi, 1].sum() / cnt - 1 / 3) < 0.015 / 3
        assert abs(hist[lo:hi, 2].sum() / cnt - 1 / 5) < 0.015 / 5


def test_normalised_estimate_of_smooth_intensity():
    """|A(Q)|^2 = 1 + Q^2 on a cubic box with h dq = q_range: the Gauss sum
    over the modes of each shell and the MC estimate both give 1 + Q^2, the
    MC one times n_int / (n_int - skips) -- the reference's normalisation
    leaves the skipped samples out of the count although they were drawn
    from the same radial law."""
    K1, K2, K3 = _grid(41)
    q = 0.1 * np.sqrt(K1 ** 2 + K2 ** 2 + K3 ** 2)
    g = 1.0 + q ** 2
    geo = sql.SqlGeometry(np.float32([0.1] * 3), np.eye(3, dtype=np.float32), (41,) * 3, np.float32(2.0),
                          np.float32(0.05), 41, (2 * np.pi / 0.1) ** 3)
    n_int = 4 * 250_000
    hist, sk = O.sql_hist_mc(np.sqrt(g).reshape(1, -1), geo.n_qq, geo.dq, geo.a_cell_inv, geo.q_range, geo.q_step,
                             41, 4, 250_000, seed=11)
    mc = sql.mc_normalise(hist.T.reshape(1, 1, 41), 1, geo, 1, n_int, sk)[0, 0]
    mc *= (n_int - sk[0] - sk[1]) / n_int
    b = np.rint(q / 0.05).astype(int)
    sel = q <= 2.0
    gauss = np.bincount(b[sel], weights=g[sel], minlength=41)
    qb = np.arange(41) * 0.05
    gauss[1:] *= 2 * np.pi ** 2 / (0.05 * qb[1:] ** 2 * geo.at_volume)
    # windows of 16 bins: the lattice's shells make single bins uneven
    for lo, hi in [(4, 20), (20, 36)]:
        want = (1 + qb[lo:hi] ** 2).mean()
        assert abs(mc[lo:hi].mean() / want - 1) < 0.03
        assert abs(mc[lo:hi].mean() / gauss[lo:hi].mean() - 1) < 0.03


def test_draw_is_a_function_of_seed_tag_and_index():
    rng = np.random.default_rng(2)
    amp = rng.standard_normal((2, 15 ** 3)) + 1j * rng.standard_normal((2, 15 ** 3))
    args = ((15,) * 3, (0.2,) * 3, np.eye(3), 1.2, 0.1, 13, 7, 3001)
    a, sa = O.sql_hist_mc(amp, *args, seed=5, tag=0)
    nthr = O.lib().orc_num_threads()
    O.lib().orc_set_num_threads(1)
    try:
        b, sb = O.sql_hist_mc(amp, *args, seed=5, tag=0)
    finally:
        O.lib().orc_set_num_threads(nthr)
    assert sa == sb and rel_l2(a, b) < 1e-14
    c, sc = O.sql_hist_mc(amp, *args, seed=5, tag=1)
    d, sd = O.sql_hist_mc(amp, *args, seed=6, tag=0)
    assert not np.array_equal(a, c) and not np.array_equal(a, d)


def test_chain_limit_and_normalisation():
    # int(.5 prod(n_qq) / 1e6) in REAL(4) (mp_sql56.f90:1063)
    geo = sql.SqlGeometry(np.float32([0.1] * 3), np.eye(3, dtype=np.float32), (509, 509, 509), np.float32(20.0),
                          np.float32(0.01), 2001, 1.0)
    assert sql.n_h_max(geo) == 65
    geo.n_qq = (101, 101, 101)
    assert sql.n_h_max(geo) == 0
    h = np.ones((2, 2, 5))
    geo = sql.SqlGeometry(np.float32([0.1] * 3), np.eye(3, dtype=np.float32), (9,) * 3, np.float32(0.4),
                          np.float32(0.1), 5, 1.0)
    out = sql.mc_normalise(h, 10, geo, 2, 1000, (100, 150))
    assert np.allclose(out[:, :, 0], 0.1)
    assert np.allclose(out[:, :, 1:], 0.1 / (np.float32(0.1) * 2) * np.float32(0.4) / 750)


def test_histogram_argument_checks():
    """Before any device work: the integration name, n_h >= 1 and the
    reference's chain limit for 10^6-sample chains (mp_sql56.f90:1069)."""
    geo = sql.SqlGeometry(np.float32([0.1] * 3), np.eye(3, dtype=np.float32), (509, 509, 509), np.float32(20.0),
                          np.float32(0.01), 2001, 1.0)
    with pytest.raises(ValueError):
        sql.histogram([], [1], geo, integration="gaus")
    with pytest.raises(ValueError):
        sql.histogram([], [1], geo, integration="mc", n_h=0)
    with pytest.raises(ValueError):
        sql.histogram([], [1], geo, integration="mc", n_h=66)


# ---------------------------------------------------------------------------
# GPU: the HIP kernel through the C ABI against the oracle
# ---------------------------------------------------------------------------
def _mc_gpu(L, amp_t, nsp, prec, N, dq, ainv, q_range, q_step, n_hist, n_h, n_mc, seed, tag, hist=None, skip=None):
    import torch
    npairs = nsp * (nsp + 1) // 2
    if hist is None:
        hist = torch.zeros(n_hist * npairs, dtype=torch.float64, device="cuda")
        skip = torch.zeros(2, dtype=torch.int64, device="cuda")
    rc = L.mpsq_sql_hist_mc(C.c_void_p(amp_t.data_ptr()), nsp, prec, (C.c_int64 * 3)(*N), (C.c_float * 3)(*dq),
                            (C.c_float * 9)(*np.asarray(ainv, np.float32).reshape(-1)), float(q_range),
                            float(q_step), n_hist, n_h, n_mc, seed, tag, C.c_void_p(hist.data_ptr()),
                            C.c_void_p(skip.data_ptr()), None)
    return rc, hist, skip


HEX = np.array([[1.0, 0.57735, 0.0], [0.0, 1.1547, 0.0], [0.0, 0.0, 0.8]], np.float32)
MC_CASES = [
    # N, dq, a_cell_inv, q_range, q_step, n_hist, nsp, precision, n_h, n_mc
    ((31, 31, 31), (0.1, 0.1, 0.1), np.eye(3, dtype=np.float32), 2.0, 0.05, 41, 2, "fp64", 3, 200_000),
    ((25, 19, 33), (0.13, 0.17, 0.11), HEX, 1.6, 0.02, 81, 3, "fp32", 2, 150_000),
    ((24, 20, 16), (0.1, 0.12, 0.15), np.eye(3, dtype=np.float32), 1.5, 0.05, 31, 1, "fp64", 1, 300_000),
    ((21, 21, 21), (0.1, 0.1, 0.1), np.eye(3, dtype=np.float32), 0.7, 0.05, 10, 2, "fp32", 400, 997),
    ((1, 17, 17), (0.1, 0.1, 0.1), np.eye(3, dtype=np.float32), 1.0, 0.1, 11, 1, "fp64", 5, 10_000),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(MC_CASES)))
def test_mc_gpu_vs_oracle(gpu_lib, case):
    """Skip counts exact, histogram to 1e-12 relative l2; cases: cubic,
    non-orthogonal with three species in FP32, even box sides, many short
    chains (n_mc 997, 400 chains: the lane stride wraps chains), a flat box
    (N_1 = 1), a q_range short of the box (skip 2) and n_hist short of
    q_range / q_step (bins dropped)."""
    import torch
    N, dq, ainv, q_range, q_step, n_hist, nsp, prec, n_h, n_mc = MC_CASES[case]
    rng = np.random.default_rng(case)
    nm = int(np.prod(N))
    amp = rng.standard_normal((nsp, nm)) + 1j * rng.standard_normal((nsp, nm))
    if prec == "fp32":
        amp = amp.astype(np.complex64).astype(np.complex128)
    cdt = torch.complex128 if prec == "fp64" else torch.complex64
    seed, tag = 0x1234_5678_9ABC + case, case
    print(f"mc case {case}: N={N} nsp={nsp} {prec} n_h={n_h} n_mc={n_mc}", flush=True)
    rc, hist, skip = _mc_gpu(gpu_lib, torch.tensor(amp, dtype=cdt, device="cuda"), nsp, 1 if prec == "fp64" else 0,
                             N, dq, ainv, q_range, q_step, n_hist, n_h, n_mc, seed, tag)
    assert rc == 0
    torch.cuda.synchronize()
    ref, sk = O.sql_hist_mc(amp.reshape(nsp, *N[::-1]), N, dq, ainv, q_range, q_step, n_hist, n_h, n_mc, seed, tag)
    assert tuple(skip.cpu().tolist()) == sk
    assert rel_l2(hist.cpu().numpy().reshape(n_hist, -1), ref) < 1e-12
    # constant amplitudes: integer counts, bit for bit
    one = torch.ones((nsp, nm), dtype=cdt, device="cuda")
    rc, h1, s1 = _mc_gpu(gpu_lib, one, nsp, 1 if prec == "fp64" else 0, N, dq, ainv, q_range, q_step, n_hist,
                         n_h, n_mc, seed, tag)
    torch.cuda.synchronize()
    r1, k1 = O.sql_hist_mc(np.ones((nsp, nm)), N, dq, ainv, q_range, q_step, n_hist, n_h, n_mc, seed, tag)
    assert rc == 0 and tuple(s1.cpu().tolist()) == k1
    assert np.array_equal(h1.cpu().numpy().reshape(n_hist, -1), r1)


@pytest.mark.gpu
def test_mc_gpu_accumulates_and_rejects_bad_arguments(gpu_lib):
    import torch
    from mp_tools_amd import _lib
    N = (15, 15, 15)
    amp = torch.ones((1, 15 ** 3), dtype=torch.complex128, device="cuda")
    args = (N, (0.1,) * 3, np.eye(3), 1.0, 0.1, 11)
    rc, h, s = _mc_gpu(gpu_lib, amp, 1, 1, *args, 2, 5000, 9, 0)
    rc2, h, s = _mc_gpu(gpu_lib, amp, 1, 1, *args, 2, 5000, 9, 1, hist=h, skip=s)
    torch.cuda.synchronize()
    r0, k0 = O.sql_hist_mc(np.ones((1, 15 ** 3)), *args, 2, 5000, 9, 0)
    r1, k1 = O.sql_hist_mc(np.ones((1, 15 ** 3)), *args, 2, 5000, 9, 1)
    assert rc == rc2 == 0
    assert np.array_equal(h.cpu().numpy().reshape(11, 1), r0 + r1)
    assert tuple(s.cpu().tolist()) == (k0[0] + k1[0], k0[1] + k1[1])
    for n_h, n_mc in [(0, 10), (1, 0), (1, 2 ** 31), (2 ** 32, 1)]:
        rc, _, _ = _mc_gpu(gpu_lib, amp, 1, 1, *args, n_h, n_mc, 9, 0)
        assert rc == _lib.ERR_BAD_ARG
    hist = torch.zeros(11, dtype=torch.float64, device="cuda")
    rc = gpu_lib.mpsq_sql_hist_mc(C.c_void_p(amp.data_ptr()), 1, 1, (C.c_int64 * 3)(*N), (C.c_float * 3)(0.1, 0.1, 0.1),
                                  (C.c_float * 9)(*np.eye(3, dtype=np.float32).reshape(-1)), 1.0, 0.1, 11, 1, 10, 9, 0,
                                  C.c_void_p(hist.data_ptr()), None, None)
    assert rc == _lib.ERR_BAD_ARG
