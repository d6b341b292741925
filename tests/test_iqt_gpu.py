"""GPU parity of the reduction's other outputs against the FP64 oracle:
the intermediate function I(Q,t) (mp_sqom56.f90:1467-1471 map form,
:1504-1510 cut form, normalised by nat^2 at :1574-1575; direct sums below
16 lags, FFT correlation from 16 on), the E(Q) /
I(Q,t) cuts through a mode index list (:1480-1510: the same per-Q
arithmetic as the map, at the pixels of the cut) and the reference's
window-length rule (:683-686).

Tolerances: I(Q,t) is summed in FP64 from the stored amplitudes (FP32 or
FP64), so both forms are held to 1e-12 relative l2; a cut is bit-exact
against the full map's columns."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NAT = 20.0


def _amps(T, NQ, seed, precision):
    import torch
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((T, NQ)) + 1j * rng.standard_normal((T, NQ))
    A += 5.0 * (rng.standard_normal(NQ) + 1j * rng.standard_normal(NQ))[None]
    if precision == "fp32":
        A = A.astype(np.complex64).astype(np.complex128)
    cdt = torch.complex128 if precision == "fp64" else torch.complex64
    return A, torch.tensor(A, dtype=cdt, device="cuda")


IQT_CASES = [
    # T, n_int, kind, m_lo (first lag), n_f (lags), subtract_mean, qfac
    (40, 20, 1, 0, 21, True, False),       # map form, every lag
    (40, 20, 2, 0, 21, True, True),        # cut form (unscaled), per-Q factor
    (64, 32, 1, 5, 10, False, True),       # a lag window, no mean subtraction
    (33, 1, 2, 0, 33, True, False),        # n_int = 1
    (512, 256, 1, 0, 257, True, False),    # the reference's default n_int = nfile / 2 at 512 frames
    # 16 lags or more take the FFT correlation (T with a radix plan)
    (48, 7, 2, 3, 30, True, True),         # a lag offset, per-Q factor
    (100, 50, 1, 0, 51, False, False),     # T = 4 x 5 x 5, no mean subtraction
    (2048, 1024, 1, 0, 17, True, False),   # one column per block
    (3000, 1500, 2, 1480, 20, True, True), # T = 8 x 3 x 5^3, the last lags
]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("T,n_int,kind,m_lo,n_f,sub,useq", IQT_CASES)
def test_iqt_vs_oracle(gpu_lib, precision, T, n_int, kind, m_lo, n_f, sub, useq):
    import torch
    from mp_tools_amd.sqw import SqwEngine
    nm = (15, 13)
    NQ = int(np.prod(nm))
    A, damp = _amps(T, NQ, T * 7 + n_int + kind, precision)
    rng = np.random.default_rng(T)
    qf = rng.uniform(0.2, 1.5, NQ).astype(np.float32) if useq else None
    eng = SqwEngine(2, nm, [10, 10], precision=precision)
    I, St = eng.reduce(damp, n_int=n_int, nat=NAT, m_lo=m_lo, n_f=n_f, subtract_mean=sub, kind=kind,
                       qfac=None if qf is None else torch.tensor(qf, device="cuda"))
    torch.cuda.synchronize()
    Ir = O.iqt(A[:, None, :], n_int=n_int, lag_lo=m_lo, n_lag=n_f, subtract_mean=sub, scaled=(kind == 1),
               nat=NAT, qfac=None if qf is None else qf.astype(np.float64))
    _, Str = O.sqw_direct(A[:, None, :], n_int=n_int, nat=NAT, n_f=1, subtract_mean=sub)
    assert I.shape == (n_f, NQ)
    assert rel_l2(I.cpu().numpy(), Ir) < 1e-12
    assert rel_l2(St.cpu().numpy(), Str) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fft_correlation_equals_direct_sums(gpu_lib, precision):
    """16 lags take the FFT correlation, 15 the direct sums: on the same
    amplitudes the shared lags agree to 1e-13 (FP64 either way)."""
    import torch
    from mp_tools_amd.sqw import SqwEngine
    T, n_int, nm = 128, 64, (16, 12)
    _, damp = _amps(T, 192, 17, precision)
    qf = torch.rand(192, device="cuda") + 0.5
    eng = SqwEngine(2, nm, [10, 10], precision=precision)
    for kind in (1, 2):
        a, ta = eng.reduce(damp, n_int=n_int, nat=NAT, m_lo=7, n_f=16, kind=kind, qfac=qf)
        b, tb = eng.reduce(damp, n_int=n_int, nat=NAT, m_lo=7, n_f=15, kind=kind, qfac=qf)
        torch.cuda.synchronize()
        assert rel_l2(a[:15].cpu().numpy(), b.cpu().numpy()) < 1e-13
        assert rel_l2(ta.cpu().numpy(), tb.cpu().numpy()) < 1e-14


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("T,n_int,kind", [(48, 48, 0), (48, 17, 0), (48, 24, 1), (48, 24, 2)])
def test_cut_equals_map_columns(gpu_lib, precision, T, n_int, kind):
    """A cut from sqw.qcut through the mode index list gives the full map's
    values at its pixels, bit for bit (S(Q,w) over one window -- the LDS
    time FFT -- and sliding windows; I(Q,t) map and cut forms)."""
    import torch
    from mp_tools_amd import sqw
    g = sqw.qplane((24, 24, 24), 2, (1, 0, 0), (0, 1, 0), q_center=(1, 1, 0))
    nm = g["n_modes"]
    NQ = int(np.prod(nm))
    _, damp = _amps(T, NQ, 99, precision)
    eng = sqw.SqwEngine(2, nm, [10, 10], precision=precision)
    n_f = 13 if kind == 0 else T - n_int + 1
    m_lo = -4 if kind == 0 else 0
    qf = torch.rand(NQ, device="cuda") + 0.5
    S, St = eng.reduce(damp, n_int=n_int, nat=NAT, m_lo=m_lo, n_f=n_f, kind=kind, qfac=qf)
    for q1, q2 in [((0.2, 0.4, 0), (1.8, 1.5, 0)), ((1.5, 1.9, 0), (0.6, 0.1, 0))]:
        c = sqw.qcut(g, q1, q2, (24, 24, 24))
        Sc, Stc = eng.reduce(damp, n_int=n_int, nat=NAT, m_lo=m_lo, n_f=n_f, kind=kind, qfac=qf, qidx=c["qidx"])
        torch.cuda.synchronize()
        idx = torch.tensor(c["qidx"], device="cuda", dtype=torch.long)
        assert Sc.shape == (n_f, c["n_qdisp"])
        assert torch.equal(Sc, S[:, idx]) and torch.equal(Stc, St[idx])
    # an arbitrary index list, repeats included
    qi = np.random.default_rng(1).integers(0, NQ, 300)
    Sc, Stc = eng.reduce(damp, n_int=n_int, nat=NAT, m_lo=m_lo, n_f=n_f, kind=kind, qfac=qf, qidx=qi)
    torch.cuda.synchronize()
    idx = torch.tensor(qi, device="cuda", dtype=torch.long)
    assert torch.equal(Sc, S[:, idx]) and torch.equal(Stc, St[idx])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_reference_n_int_rule(gpu_lib, precision):
    """odd_n_int: n_int 0 is nfile / 2, grown to odd for S(Q,w); the default
    output is then the reference's n_freq_max frequencies or lags."""
    import torch
    from mp_tools_amd.sqw import SqwEngine, reference_n_int
    T, nm = 40, (8, 6)
    _, damp = _amps(T, 48, 5, precision)
    eng = SqwEngine(2, nm, [10, 10], precision=precision)
    for kind, n_in in [(0, 0), (0, 10), (0, 9), (1, 0), (2, 0), (1, 7)]:
        n_eff, n_all = reference_n_int(T, n_in, sqw=(kind == 0))
        S, St = eng.reduce(damp, n_int=n_in, nat=NAT, kind=kind, odd_n_int=True)
        Sr, Str = eng.reduce(damp, n_int=n_eff, nat=NAT, kind=kind, n_f=n_all)
        torch.cuda.synchronize()
        assert S.shape == (n_all, 48)
        assert torch.equal(S, Sr) and torch.equal(St, Str)


def test_iqt_argument_errors(gpu_lib):
    from mp_tools_amd.sqw import SqwEngine
    _, damp = _amps(20, 48, 6, "fp64")
    eng = SqwEngine(2, (8, 6), [10, 10], precision="fp64")
    with pytest.raises(RuntimeError):
        eng.reduce(damp, n_int=10, kind=1, m_lo=5, n_f=7)      # last lag 11 > T - n_int = 10
    with pytest.raises(RuntimeError):
        eng.reduce(damp, n_int=10, kind=3, n_f=1)
    with pytest.raises(RuntimeError):
        eng.reduce(damp, n_int=10, kind=1, m_lo=-1, n_f=2)


@pytest.mark.parametrize("n_int", [32, 9])
def test_out_of_range_index_gives_nan(gpu_lib, n_int):
    """Through the C ABI (the Python front end rejects such lists): index
    list entries outside [0, NQ) give NaN columns and read nothing; the
    other columns are the map's (one window and sliding windows)."""
    import ctypes as C
    import torch
    from mp_tools_amd import _lib
    from mp_tools_amd.sqw import SqwEngine
    T, NQ = 32, 48
    _, damp = _amps(T, NQ, 8, "fp64")
    eng = SqwEngine(2, (8, 6), [10, 10], precision="fp64")
    S, St = eng.reduce(damp, n_int=n_int, nat=NAT, n_f=5)
    qi = torch.tensor([3, -1, 47, NQ, 10, NQ + 1000, 0], dtype=torch.int32, device="cuda")
    out = torch.zeros((5, qi.numel()), dtype=torch.float64, device="cuda")
    tot = torch.zeros(qi.numel(), dtype=torch.float64, device="cuda")
    r = _lib.ReduceSpec()
    r.n_int, r.win_kind, r.dt, r.nat, r.m_lo, r.n_f, r.subtract_mean = n_int, 0, 1.0, NAT, 0, 5, 1
    r.d_qidx, r.n_qidx = qi.data_ptr(), qi.numel()
    _lib.check("sqw_reduce", eng._L.mpsq_sqw_reduce(eng._h, damp.data_ptr(), T, C.byref(r), out.data_ptr(),
                                                    tot.data_ptr()))
    torch.cuda.synchronize()
    bad = torch.tensor([1, 3, 5], device="cuda")
    good, src = torch.tensor([0, 2, 4, 6], device="cuda"), torch.tensor([3, 47, 10, 0], device="cuda")
    assert torch.isnan(out[:, bad]).all() and torch.isnan(tot[bad]).all()
    assert torch.equal(out[:, good], S[:, src]) and torch.equal(tot[good], St[src])
    with pytest.raises(ValueError):
        eng.reduce(damp, n_int=n_int, nat=NAT, n_f=5, qidx=[0, NQ])
