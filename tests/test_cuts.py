"""Host logic of the E(Q) / I(Q,t) cuts, the E = const stacks and the
reference's window-length rule (mp_sqom56.f90:683-701, 1266-1284,
1310-1321, 1389-1432), and the oracle's intermediate function I(Q,t)
(:1467-1471, :1504-1510) against closed forms and numpy."""
import numpy as np
import pytest

from mp_tools_amd import sqw
from oracle import oracle as O


def test_reference_n_int_rule():
    # n_int = 0 -> nfile / 2; S(Q,w) takes an odd n_int and (n_int - 1) / 2 + 1 frequencies
    assert sqw.reference_n_int(512) == (257, 129)
    assert sqw.reference_n_int(513) == (257, 129)
    assert sqw.reference_n_int(100, 40) == (41, 21)
    assert sqw.reference_n_int(100, 41) == (41, 21)
    # I(Q,t): n_int as given, nfile - n_int + 1 lags
    assert sqw.reference_n_int(512, 0, sqw=False) == (256, 257)
    assert sqw.reference_n_int(100, 40, sqw=False) == (40, 61)


def test_econst_stack_indices():
    # E_min 0.5, E_step 0.25 at 0.1 THz steps: i_step = nint(2.5) = 3 (half away from zero)
    m, lab = sqw.econst_stack(0.5, 0.25, 6, 0.1)
    assert m == [5, 8]                    # i_plot 1, 4: j_freq 6, 9 -> m 5, 8
    assert np.allclose(lab, [0.5, 0.8])
    assert sqw.econst_stack(0.5, 0.25, 7, 0.1)[0] == [5, 8, 11]     # i_plot 7 <= n_plot
    # a step below freq_step becomes one frequency step
    m, _ = sqw.econst_stack(0.5, 0.01, 3, 0.1)
    assert m == [5, 6, 7]
    # one map (E = 0): the zero frequency
    assert sqw.econst_stack(0.0, 0.1, 1, 0.1)[0] == [0]
    # I(Q,t): lag j_freq, one more than the frequency index
    assert sqw.econst_stack(0.0, 0.02, 3, 0.01, iqt=True)[0] == [1, 3]


def _plane():
    return sqw.qplane((40, 40, 40), 2, (1, 0, 0), (0, 1, 0), q_center=(1, 1, 0))


def test_qcut_axis_cuts():
    g = _plane()
    n_qx = g["n_qx"]
    assert (n_qx, g["n_qy"]) == (80, 80)
    c = sqw.qcut(g, (0.5, 1, 0), (1.5, 1, 0), (40, 40, 40))
    # 40 pixels per r.l.u.: Q_x = 0.5 is pixel 21, Q_y = 1 is pixel 41
    assert c["n_qdisp"] == 40 and c["label"] == "Q_x"
    assert np.array_equal(c["pixels"][:, 0], np.arange(21, 61)) and np.all(c["pixels"][:, 1] == 41)
    assert np.array_equal(c["qidx"], (c["pixels"][:, 0] - 1) + n_qx * 40)
    # the reversed cut walks down from Q1's pixel
    r = sqw.qcut(g, (1.5, 1, 0), (0.5, 1, 0), (40, 40, 40))
    assert np.array_equal(r["pixels"][:, 0], np.arange(61, 21, -1))
    # along Q_y
    c = sqw.qcut(g, (1, 0.25, 0), (1, 1.75, 0), (40, 40, 40))
    assert c["label"] == "Q_y" and c["n_qdisp"] == 60
    assert np.all(c["pixels"][:, 0] == 41) and np.array_equal(c["pixels"][:, 1], np.arange(11, 71))


@pytest.mark.parametrize("q1,q2", [((0.1, 0.3, 0), (1.9, 1.2, 0)), ((1.7, 0.05, 0), (0.2, 1.95, 0)),
                                   ((0.3, 0.3, 0), (1.3, 1.3, 0)), ((0.8, 1.9, 0), (1.05, 0.1, 0))])
def test_qcut_pixels_follow_the_line(q1, q2):
    """Every pixel of an oblique cut is one step along the longer axis from
    the previous one and lies within half a pixel (plus FP32 rounding) of
    the straight line between the end pixels."""
    g = _plane()
    c = sqw.qcut(g, q1, q2, (40, 40, 40))
    px, py = c["pixels"][:, 0].astype(float), c["pixels"][:, 1].astype(float)
    a = np.array([px[0], py[0]])
    e1r = 40.0
    b = np.array([np.rint((q2[0] - 0.0) * e1r) + 1, np.rint((q2[1] - 0.0) * e1r) + 1])
    long_ax = 0 if abs(b[0] - a[0]) >= abs(b[1] - a[1]) else 1
    steps = np.diff(c["pixels"][:, long_ax])
    assert np.all(np.abs(steps) == 1) and len(set(np.sign(steps))) == 1
    assert len(px) == abs(b[long_ax] - a[long_ax])
    d = b - a
    t = (np.stack([px, py], 1) - a) @ d / (d @ d)
    off = np.stack([px, py], 1) - (a + t[:, None] * d)
    assert np.max(np.abs(off[:, 1 - long_ax])) <= 0.5 + 1e-4
    # the pixels' mode indices are those of the plane's CMCL order
    assert np.array_equal(c["qidx"], (c["pixels"][:, 0] - 1) + g["n_qx"] * (c["pixels"][:, 1] - 1))


def test_qcut_series_shift_and_errors():
    g = _plane()
    q1, q2 = (0.5, 0.5, 0), (1.5, 0.5, 0)
    dq = sqw.cut_step(g, q1, q2, 0.25)
    # in plane, perpendicular to Q2 - Q1, of length q_step, pointing -(dq x ev) = +e2 here
    assert np.allclose(dq, [0, 0.25, 0])
    c0 = sqw.qcut(g, q1, q2, (40, 40, 40), i_plot=2, n_plot=3, dq_p=dq)   # the middle plot is not shifted
    c1 = sqw.qcut(g, q1, q2, (40, 40, 40), i_plot=1, n_plot=3, dq_p=dq)   # shifted by -dq
    c2 = sqw.qcut(g, q1, q2, (40, 40, 40), i_plot=3, n_plot=3, dq_p=dq)
    assert np.all(c0["pixels"][:, 1] == 21)
    assert np.all(c1["pixels"][:, 1] == 11) and np.all(c2["pixels"][:, 1] == 31)
    with pytest.raises(ValueError, match="not in the Q plane"):
        sqw.qcut(g, (0.5, 0.5, 0.1), q2, (40, 40, 40))
    with pytest.raises(ValueError, match="out of the map range"):
        sqw.qcut(g, (0.5, 0.5, 0), (2.5, 0.5, 0), (40, 40, 40))
    with pytest.raises(ValueError, match="one pixel"):
        sqw.qcut(g, (0.5, 0.5, 0), (0.51, 0.5, 0), (40, 40, 40))


# ---------------------------------------------------------------------------
# oracle I(Q,t)
# ---------------------------------------------------------------------------

def test_iqt_closed_forms():
    T, NQ = 48, 5
    t = np.arange(T)
    # constant amplitudes: zero after mean subtraction
    F = np.ones((T, 1, NQ), np.complex128) * (0.3 - 0.7j)
    assert np.abs(O.iqt(F, n_int=20, lag_lo=0, n_lag=T - 19)).max() < 1e-14
    # a pure tone a e^{-i w t} with w = 2 pi k / T (zero mean over the T frames):
    # I(lag) = |a|^2 cos(w lag) scaled, n_int |a|^2 cos(w lag) unscaled
    k = np.array([1, 3, 5, 7, 11])
    a = np.array([1.0, 0.5 + 0.5j, -2.0, 0.25j, 3.0])
    w = 2 * np.pi * k / T
    F = (a[None, :] * np.exp(-1j * np.outer(t, w)))[:, None, :]
    n_int, lags = 20, np.arange(T - 20 + 1)
    I_s = O.iqt(F, n_int=n_int, lag_lo=0, n_lag=len(lags))
    I_u = O.iqt(F, n_int=n_int, lag_lo=0, n_lag=len(lags), scaled=False)
    want = np.abs(a)[None, :] ** 2 * np.cos(np.outer(lags, w))
    assert np.allclose(I_s, want, atol=1e-12) and np.allclose(I_u, n_int * want, atol=1e-11)
    # a lag window starting past zero is the same rows
    assert np.array_equal(O.iqt(F, n_int=n_int, lag_lo=7, n_lag=5), I_s[7:12])


def test_iqt_is_numpy_correlation():
    """Re sum_i C'(i + lag) conj C'(i), C' = sum_j w_j F_j - mean over t, with
    scalar weights, then a per-Q weight table in their place (the X-ray
    form factors), one species masked off."""
    rng = np.random.default_rng(5)
    T, nsp, NQ = 40, 3, 7
    F = rng.standard_normal((T, nsp, NQ)) + 1j * rng.standard_normal((T, nsp, NQ))
    wsc = np.array([0.5, -1.2, 2.0])
    wtab = rng.random((nsp, NQ))
    mask = np.array([1, 0, 1], np.int32)
    n_int, lag_lo, n_lag = 15, 3, 20
    for tab in (None, wtab):
        got = O.iqt(F, wsc=wsc, wtab=tab, mask=mask, n_int=n_int, lag_lo=lag_lo, n_lag=n_lag, scaled=False)
        C = np.zeros((T, NQ), np.complex128)
        for j in (0, 2):
            C += (wsc[j] if tab is None else tab[j][None, :]) * F[:, j]
        C -= C.mean(axis=0)
        for q in range(NQ):
            full = np.correlate(C[:, q], C[:n_int, q], mode="valid")    # sum_i C(i + lag) conj(C(i))
            assert np.allclose(got[:, q], full[lag_lo:lag_lo + n_lag].real, rtol=1e-12, atol=1e-12)
        # lag 0 of the scaled form is the mean |C'|^2 over the first n_int frames
        I0 = O.iqt(F, wsc=wsc, wtab=tab, mask=mask, n_int=n_int, lag_lo=0, n_lag=1)
        assert np.allclose(I0[0], (np.abs(C[:n_int]) ** 2).mean(axis=0), rtol=1e-12)


def test_iqt_argument_checks():
    F = np.zeros((10, 1, 2), np.complex128)
    with pytest.raises(RuntimeError):
        O.iqt(F, n_int=5, lag_lo=4, n_lag=3)       # last lag 6 > T - n_int = 5
    with pytest.raises(RuntimeError):
        O.iqt(F, n_int=11)


# ---------------------------------------------------------------------------
# speckle filter
# ---------------------------------------------------------------------------

def test_speckle_pass_known_answers():
    # a lone spike takes its neighbours' max, a lone dip their min
    v = O.speckle_pass([1.0, 1.0, 9.0, 2.0, 1.0], 2.0)
    assert np.array_equal(v, [1.0, 1.0, 2.0, 2.0, 1.0])
    v = O.speckle_pass([5.0, 6.0, 0.5, 4.0, 5.0], 2.0)
    assert np.array_equal(v, [5.0, 6.0, 4.0, 4.0, 5.0])
    # the walk is sequential: element 2 (30) sees element 1 already lowered to 50, so
    # it is no dip (30 >= min(50, 200) / 2); a filter over the original values would
    # raise it to 200.  Element 3 (200) is a spike against (30, 50).
    v = O.speckle_pass([50.0, 1000.0, 30.0, 200.0, 50.0], 2.0)
    assert np.array_equal(v, [50.0, 50.0, 30.0, 50.0, 50.0])
    # cyclic: the first element's left neighbour is the last one
    v = O.speckle_pass([100.0, 1.0, 2.0, 5.0], 3.0)
    assert np.array_equal(v, [5.0, 1.0, 2.0, 5.0])
    # s_trig 1 pins every value between its neighbours' min and max
    rng = np.random.default_rng(3)
    x = rng.random(200)
    v = O.speckle_pass(x, 1.0)
    for k in range(1, 199):
        assert v[k] >= min(v[k - 1], x[k + 1]) - 1e-15 and v[k] <= max(v[k - 1], x[k + 1]) + 1e-15


def test_speckle_filter_matches_oracle_walks():
    """The library's filter (host code) against the oracle's pass in the
    reference's walk orders: a map x fastest then y fastest, a cut frequency
    fastest; s_trig 0 is no filter."""
    rng = np.random.default_rng(11)
    m = rng.lognormal(0.0, 1.5, (24, 31))
    got = sqw.speckle_filter(m, 3.0)
    want = O.speckle_pass(m.ravel(), 3.0).reshape(m.shape)           # x fastest
    want = O.speckle_pass(want.T.ravel(), 3.0).reshape(m.T.shape).T  # y fastest
    assert np.array_equal(got, want)
    assert not np.array_equal(got, m)
    c = rng.lognormal(0.0, 1.5, (17, 40))                                # [n_f, n_qdisp]
    got = sqw.speckle_filter(c, 2.5, cut=True)
    assert np.array_equal(got, O.speckle_pass(c.T.ravel(), 2.5).reshape(c.T.shape).T)
    assert np.array_equal(sqw.speckle_filter(m, 0.0), m)
