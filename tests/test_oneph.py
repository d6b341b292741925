"""The one-phonon approximation (j_oneph = 1, mp_sqom56.f90:1037-1098): the
oracle's restatement pinned by closed forms on small lattices -- ideal
positions, one displaced atom, a plane-wave displacement, vacancies and the
(hhl) column map -- and, on the GPU, the HIP path against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O


def lattice(n, bases, u=None, seed=None):
    """CELL data: nsp blocks of n^3 atoms with 1-based cell indices (a
    fastest) at pos = ind - n/2 - 1 + base + u (lattice units), the
    convention of the reference's CELL window (mp_sqom56.f90:1110)."""
    nsp = len(bases)
    c = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1)
    cells = c.transpose(2, 1, 0, 3).reshape(-1, 3) + 1          # a fastest
    ind = np.zeros((nsp * n ** 3, 4), np.int32)
    pos = np.zeros((nsp * n ** 3, 4), np.float32)
    for j, b in enumerate(bases):
        sl = slice(j * n ** 3, (j + 1) * n ** 3)
        ind[sl, :3] = cells
        ind[sl, 3] = j + 1
        pos[sl, :3] = cells - n // 2 - 1 + np.asarray(b, np.float64)
    if u is not None:
        pos[:, :3] = (pos[:, :3].astype(np.float64) + u).astype(np.float32)
    elif seed is not None:
        rng = np.random.default_rng(seed)
        pos[:, :3] = (pos[:, :3].astype(np.float64) + 0.05 * rng.standard_normal((pos.shape[0], 3))).astype(np.float32)
    return pos, ind


# no ideal site at (0, 0, 0), which the reference reads as a vacancy
BASES = [(0.25, 0.0, 0.0), (0.5, 0.5, 0.5), (0.5, 0.5, 0.0)]


def pixel_q(n, n_bz, q_min):
    nq = n_bz * n
    X = np.arange(nq)
    qx = (np.float32(q_min[0]) + (X // n).astype(np.float32)) + (X % n).astype(np.float32) / np.float32(n)
    return qx.astype(np.float64)


def test_ideal_lattice_bragg_only():
    """Ideal positions: every (hk0) displacement is pos - base - ind + n/2 =
    -1 (the reference's 1-based convention), so U is n_row(2) n_row(3)
    (-1) at the zone centre only and A(Q) = exp(-2 pi i Q.base) (q_x + q_y)
    (-n^2) at the Bragg pixels, zero elsewhere."""
    n, n_bz = 4, 2
    pos, ind = lattice(n, BASES)
    A = O.oneph(pos[None], ind, BASES, (n, n, n), n_bz, 1, (0.0, 0.0))
    q = pixel_q(n, n_bz, (0.0, 0.0))
    QX, QY = np.meshgrid(q, q, indexing="xy")                       # [Y, X]
    bragg = (np.abs(QX - np.round(QX)) < 1e-6) & (np.abs(QY - np.round(QY)) < 1e-6)
    for j, b in enumerate(BASES):
        a = A[0, j].reshape(n_bz * n, n_bz * n)
        ph = np.exp(-2j * np.pi * (QX * b[0] + QY * b[1]))
        want = np.where(bragg, ph * (QX + QY) * (-(n * n * n) / n), 0.0)
        assert np.allclose(a, want, atol=1e-4)


def test_single_displaced_atom_closed_form():
    """One atom of type 1 at cell (k0, l0, m0) moved by (dx, dy): the
    amplitudes change by exp(-2 pi i Q.base) (q_x dx + q_y dy)
    exp(-2 pi i ((k0-1) kk / n + (l0-1) ll / n)) / n at every pixel."""
    n, n_bz, q_min = 6, 2, (-0.5, 0.25)
    pos0, ind = lattice(n, BASES)
    k0, l0, m0 = 2, 5, 3
    a = n ** 3 + (k0 - 1) + n * (l0 - 1) + n * n * (m0 - 1)          # type 1 (0-based block 1)
    assert tuple(ind[a, :3]) == (k0, l0, m0)
    d = np.array([0.0625, -0.125], np.float32)
    pos1 = pos0.copy()
    pos1[a, :2] += d
    A0 = O.oneph(pos0[None], ind, BASES, (n, n, n), n_bz, 1, q_min)
    A1 = O.oneph(pos1[None], ind, BASES, (n, n, n), n_bz, 1, q_min)
    nq = n_bz * n
    X = np.arange(nq)
    qx = pixel_q(n, n_bz, q_min)
    qy = (np.float32(q_min[1]) + (X // n).astype(np.float32)) + (X % n).astype(np.float32) / np.float32(n)
    QX, QY = np.meshgrid(qx, qy.astype(np.float64), indexing="xy")
    nqx_min = int(np.float32(n) * (np.float32(q_min[0]) - np.floor(np.float32(q_min[0]))))
    nqy_min = int(np.float32(n) * (np.float32(q_min[1]) - np.floor(np.float32(q_min[1]))))
    KK = (nqx_min + (X % n)) % n
    LL = (nqy_min + (X % n)) % n
    KKg, LLg = np.meshgrid(KK, LL, indexing="xy")
    b = BASES[1]
    want = (np.exp(-2j * np.pi * (QX * b[0] + QY * b[1])) * (QX * d[0] + QY * d[1]) *
            np.exp(-2j * np.pi * ((k0 - 1) * KKg / n + (l0 - 1) * LLg / n)) / n)
    diff = (A1[0, 1] - A0[0, 1]).reshape(nq, nq)
    assert np.allclose(diff, want, atol=2e-6)
    assert np.array_equal(A1[0, 0], A0[0, 0]) and np.array_equal(A1[0, 2], A0[0, 2])


def test_plane_wave_displacement_peaks():
    """u_x = 0.1 cos(2 pi h (a - 1) / n) on every atom of type 0: U_x is
    0.1 n^3 / (2 n) at kk = h and n - h besides the zone-centre term, so the
    change of A at q_x = h/n (mod 1) is exp(-2 pi i q_x base_x) q_x 0.05 n^2
    and zero at the other off-centre pixels."""
    n, h = 8, 3
    pos0, ind = lattice(n, BASES[:1])
    u = np.zeros((n ** 3, 3))
    u[:, 0] = 0.1 * np.cos(2 * np.pi * h * (ind[:, 0] - 1) / n)
    pos1, _ = lattice(n, BASES[:1], u=u)
    A0 = O.oneph(pos0[None], ind, BASES[:1], (n, n, n), 1, 1)
    A1 = O.oneph(pos1[None], ind, BASES[:1], (n, n, n), 1, 1)
    d = (A1 - A0)[0, 0].reshape(n, n)                                 # [ll, kk], q_min = 0
    qx = np.arange(n) / n
    for kk in range(n):
        ph = np.exp(-2j * np.pi * qx[kk] * BASES[0][0])
        want = ph * qx[kk] * 0.05 * n * n if kk in (h, n - h) else 0.0
        assert np.allclose(d[0, kk], want, atol=2e-5)
    assert np.allclose(d[1:, :], 0.0, atol=2e-5)


def test_vacancy_and_linearity():
    """An atom at (0, 0, 0) is a vacancy: the amplitudes equal those with
    that atom's displacement column contribution removed, i.e. the
    difference to the full lattice is minus the single-atom term."""
    n = 4
    pos, ind = lattice(n, BASES, seed=3)
    a = 2 * n ** 3 + 5
    A = O.oneph(pos[None], ind, BASES, (n, n, n), 1, 1)
    pv = pos.copy()
    pv[a, :3] = 0.0
    Av = O.oneph(pv[None], ind, BASES, (n, n, n), 1, 1)
    # the lone atom, every other atom of its type at its ideal position minus itself
    lone = lattice(n, BASES)[0]
    lone_only = lone.copy()
    lone_only[a] = pos[a]
    A1 = O.oneph(lone_only[None], ind, BASES, (n, n, n), 1, 1)
    lone_v = lone.copy()
    lone_v[a, :3] = 0.0
    A2 = O.oneph(lone_v[None], ind, BASES, (n, n, n), 1, 1)
    assert np.allclose(A - Av, A1 - A2, atol=1e-9)
    assert np.array_equal(A[0, :2], Av[0, :2])


def test_hhl_column_map():
    """(hhl): the atom at cells (a, b, c) lands in column ((a + b - 2) mod n
    + 1, c) with d_x = (d_a + d_b) / sqrt 2 (+2 / sqrt 2 for the reference's
    offsets) and d_y = d_c + 1: moving one atom by (e, e, 0) changes U_x
    by 2 e / sqrt 2 at that column only: A changes by exp(-2 pi i q_x
    (base_a + base_b)) q_x (2 e / sqrt 2) exp(-2 pi i ((k-1) kk + (l-1) ll) / n) / n."""
    n = 4
    pos0, ind = lattice(n, BASES[:1])
    a = 1 + n * 2 + n * n * 3                                         # cell (2, 3, 4)
    e = np.float32(0.125)
    pos1 = pos0.copy()
    pos1[a, 0] += e
    pos1[a, 1] += e
    A0 = O.oneph(pos0[None], ind, BASES[:1], (n, n, n), 1, 2, (0.0, 0.0))
    A1 = O.oneph(pos1[None], ind, BASES[:1], (n, n, n), 1, 2, (0.0, 0.0))
    k = ((2 - 1 + 3 - 1) % n) + 1
    l = 4
    nq = n
    d = (A1 - A0)[0, 0].reshape(nq, nq)
    q = np.arange(n) / n
    for ll in range(n):
        for kk in range(n):
            ph = np.exp(-2j * np.pi * q[kk] * (BASES[0][0] + BASES[0][1]))
            want = ph * q[kk] * (2 * e / np.float32(np.sqrt(2.0))) * np.exp(
                -2j * np.pi * ((k - 1) * kk / n + (l - 1) * ll / n)) / n
            assert abs(d[ll, kk] - want) < 2e-6


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
@pytest.mark.parametrize("j_plane,q_min,n_bz", [(1, (0.0, 0.0), 2), (1, (-1.25, 0.5), 3), (2, (0.5, -0.75), 2)])
def test_oneph_gpu_vs_oracle(gpu_lib, precision, tol, j_plane, q_min, n_bz):
    import torch
    from conftest import rel_l2
    from mp_tools_amd import sqw
    n, T = 6, 3
    frames, ind = [], None
    for t in range(T):
        p, ind = lattice(n, BASES, seed=10 + t)
        frames.append(p)
    pos = np.stack(frames)
    pos[1, 7, :3] = 0.0                                                # a vacancy
    ref = O.oneph(pos, ind, BASES, (n, n, n), n_bz, j_plane, q_min)
    dpos = torch.tensor(pos, device="cuda")
    dind = torch.tensor(ind, device="cuda")
    A = sqw.oneph_amplitudes(dpos, dind, BASES, (n, n, n), n_bz, j_plane, q_min, precision=precision)
    torch.cuda.synchronize()
    assert A.shape == ref.shape
    assert rel_l2(A.cpu().numpy(), ref) < tol
    # species combined with scalar weights, then with a per-Q table
    w = np.array([0.7, -0.3, 1.1])
    Aw = sqw.oneph_amplitudes(dpos, dind, BASES, (n, n, n), n_bz, j_plane, q_min, precision=precision, weights=w)
    assert rel_l2(Aw.cpu().numpy(), np.einsum("j,tjq->tq", w, ref)) < tol
    NQ = ref.shape[2]
    wt = torch.rand((3, NQ), device="cuda")
    At = sqw.oneph_amplitudes(dpos, dind, BASES, (n, n, n), n_bz, j_plane, q_min, precision=precision, w_table=wt)
    torch.cuda.synchronize()
    want = (ref * wt.cpu().numpy().astype(np.float64)[None]).sum(axis=1)
    assert rel_l2(At.cpu().numpy(), want) < tol


@pytest.mark.gpu
def test_oneph_gpu_argument_errors(gpu_lib):
    import torch
    from mp_tools_amd import _lib, sqw
    n = 4
    pos, ind = lattice(n, BASES)
    dpos = torch.tensor(pos[None], device="cuda")
    bad = ind.copy()
    bad[3, 0] = n + 1                                                   # a cell index past the columns
    with pytest.raises(_lib.MpsqError):
        sqw.oneph_amplitudes(dpos, torch.tensor(bad, device="cuda"), BASES, (n, n, n), 1)
    with pytest.raises(_lib.MpsqError):                                 # non-square plane
        sqw.oneph_amplitudes(dpos, torch.tensor(ind, device="cuda"), BASES, (n, n + 2, n), 1)
