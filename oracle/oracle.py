"""ctypes front of oracle/liboracle.so -- the CPU FP64 restatement of the
S(Q,w) hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the cpu_baseline leg of ``bench.py``.  The product package
(``mp_tools_amd``) never imports this module.

Citations are to the jkulda/MP_tools tree; see the per-function notes in
oracle.c for the reference lines each routine restates.
"""
from __future__ import annotations

import ctypes as C
import functools
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with gcc (OpenMP, no FP contraction)."""
    src = os.path.join(_HERE, "oracle.c")
    if force or not os.path.exists(_LIB_PATH) or (
        os.path.getmtime(src) > os.path.getmtime(_LIB_PATH)
    ):
        cmd = ["make", "-C", _HERE, "-s", "liboracle.so"]
        subprocess.check_call(cmd)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        i64, i32, dbl = C.c_int64, C.c_int, C.c_double
        vp = C.c_void_p
        L.orc_dirft_type1.argtypes = [i32, i64, vp, vp, vp, vp, i32, i64, i64, i64, vp]
        L.orc_dirft_type1_sampled.argtypes = [i32, i64, vp, vp, vp, vp, i32, i64, vp, vp, vp, vp]
        L.orc_dirft_type1_prod.argtypes = [i32, i64, vp, vp, vp, vp, i32, i64, vp, i64, vp, i64, vp, vp]
        L.orc_qtables.argtypes = [i32, i32, C.c_float, C.c_float, C.c_float, C.c_float, vp, vp, vp, vp, vp, vp,
                                  i32, vp, vp, vp]
        L.orc_qtables.restype = None
        L.orc_fold.argtypes = [i32, i64, vp, vp, vp, vp, i32, i32, i32, vp, vp, vp]
        L.orc_counting_sort.argtypes = [i64, vp, i32, vp, vp]
        L.orc_select_cell.argtypes = [i64, vp, vp, vp, vp, vp, vp]
        L.orc_select_cell.restype = i64
        L.orc_select_bulk.argtypes = [i64, vp, vp, vp, C.c_float, vp]
        L.orc_select_bulk.restype = i64
        L.orc_stage.argtypes = [i32, i64, vp, vp, vp, vp, vp, vp, i32, dbl, vp, vp, vp, vp]
        L.orc_stage.restype = None
        L.orc_sqw_direct.argtypes = [i64, i64, i32, vp, vp, vp, vp, vp, i64, i32, dbl, dbl,
                                     i64, i64, i32, vp, vp]
        L.orc_iqt.argtypes = [i64, i64, i32, vp, vp, vp, vp, vp, i64, i64, i64, i32, i32, dbl, vp]
        L.orc_speckle_pass.argtypes = [i64, vp, dbl]
        L.orc_dos.argtypes = [i64, i64, vp, vp, i32, i64, vp]
        L.orc_philox4x32_10.argtypes = [vp, vp, vp]
        L.orc_sql_hist_mc.argtypes = [i32, vp, vp, vp, C.c_float, C.c_float, i32, i64, i64, C.c_uint64, C.c_uint32,
                                      vp, vp, vp]
        L.orc_oneph.argtypes = [i64, i32, vp, i32, i32, C.c_float, C.c_float, vp, vp, vp, vp]
        L.orc_speckle_pass.restype = None
        L.orc_es_spread.argtypes = [i32, i64, vp, vp, vp, vp, vp, i32, dbl, vp]
        L.orc_num_threads.restype = i32
        L.orc_set_num_threads.argtypes = [i32]
        L.orc_set_num_threads.restype = None
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def set_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def _coords(dim, x, y, z):
    x = _f64(x)
    y = _f64(y) if (dim >= 2 and y is not None) else np.zeros(1)
    z = _f64(z) if (dim >= 3 and z is not None) else np.zeros(1)
    return x, y, z


def dirft_type1(x, y, z, c, n_modes, iflag=-1):
    """Dense FP64 direct type-1 DFT; returns complex128 array of shape
    (N3, N2, N1) (CMCL: k1 fastest) -- squeezed to the used dims."""
    n_modes = tuple(int(v) for v in n_modes)
    dim = len(n_modes)
    N = list(n_modes) + [1] * (3 - dim)
    x, y, z = _coords(dim, x, y, z)
    c = _c128(c)
    f = np.zeros(N[0] * N[1] * N[2], np.complex128)
    rc = lib().orc_dirft_type1(dim, x.size, _p(x), _p(y), _p(z), _p(c), iflag,
                               N[0], N[1], N[2], _p(f))
    if rc:
        raise RuntimeError(f"orc_dirft_type1 rc={rc}")
    return f.reshape(tuple(reversed(n_modes)))


def dirft_type1_sampled(x, y, z, c, k1, k2=None, k3=None, iflag=-1):
    """FP64 direct sum at explicit integer modes; returns complex128[K]."""
    dim = 1 + (k2 is not None) + (k3 is not None)
    x, y, z = _coords(dim, x, y, z)
    c = _c128(c)
    k1 = np.ascontiguousarray(k1, np.int64)
    K = k1.size
    k2 = np.ascontiguousarray(k2 if k2 is not None else np.zeros(K), np.int64)
    k3 = np.ascontiguousarray(k3 if k3 is not None else np.zeros(K), np.int64)
    f = np.zeros(K, np.complex128)
    rc = lib().orc_dirft_type1_sampled(dim, x.size, _p(x), _p(y), _p(z), _p(c), iflag, K,
                                       _p(k1), _p(k2), _p(k3), _p(f))
    if rc:
        raise RuntimeError(f"orc_dirft_type1_sampled rc={rc}")
    return f


def dirft_type1_prod(x, y, z, c, k1s, k2s=None, k3s=None, iflag=-1):
    """FP64 direct sum on the product set k1s x k2s x k3s (each at most 4096
    values); returns complex128[len(k3s), len(k2s), len(k1s)]."""
    dim = 1 + (k2s is not None) + (k3s is not None)
    x, y, z = _coords(dim, x, y, z)
    c = _c128(c)
    k1s = np.ascontiguousarray(k1s, np.int64)
    k2s = np.ascontiguousarray(k2s if k2s is not None else [0], np.int64)
    k3s = np.ascontiguousarray(k3s if k3s is not None else [0], np.int64)
    if max(k2s.size, k3s.size) > 4096:
        raise ValueError("at most 4096 values of k2 and k3")
    f = np.zeros((k3s.size, k2s.size, k1s.size), np.complex128)
    rc = lib().orc_dirft_type1_prod(dim, x.size, _p(x), _p(y), _p(z), _p(c), iflag, k1s.size, _p(k1s),
                                    k2s.size, _p(k2s), k3s.size, _p(k3s), _p(f))
    if rc:
        raise RuntimeError(f"orc_dirft_type1_prod rc={rc}")
    return f


def qtables(plane, x_ffpar=None):
    """a8: (q2_norm [n_qx*n_qy], x_ffq [nsp, n_qx*n_qy] or None) of a Q plane
    (the dict of mp_tools_amd.sqw.qplane), REAL(4) as mp_sqom56.f90:909-936."""
    f32 = lambda v: np.ascontiguousarray(v, np.float32)
    n_qx, n_qy = int(plane["n_qx"]), int(plane["n_qy"])
    q2 = np.zeros(n_qx * n_qy, np.float32)
    ff = None if x_ffpar is None else f32(x_ffpar).reshape(-1, 9)
    nsp = 0 if ff is None else ff.shape[0]
    xq = None if ff is None else np.zeros((nsp, n_qx * n_qy), np.float32)
    vec = [f32(plane[k]) for k in ("e1", "e2", "q_v", "e1p", "e2p", "q_vp")]
    lib().orc_qtables(n_qx, n_qy, float(plane["qx_range"][0]), float(plane["qy_range"][0]), float(plane["bz_nx"]),
                      float(plane["bz_ny"]), *[_p(v) for v in vec], nsp, _p(ff), _p(q2), _p(xq))
    return q2, xq


def fold(x, y, z, n_fine, w, by=1, bz=1):
    """FP64 fold: returns (l int32[M,3], d float64[M,3], key int32[M], rc)."""
    dim = len(n_fine)
    n = np.array(list(n_fine) + [1] * (3 - dim), np.int64)
    x, y, z = _coords(dim, x, y, z)
    M = x.size
    l = np.zeros((M, 3), np.int32)
    d = np.zeros((M, 3), np.float64)
    key = np.zeros(M, np.int32)
    rc = lib().orc_fold(dim, M, _p(x), _p(y), _p(z), _p(n), int(w), int(by), int(bz),
                        _p(l), _p(d), _p(key))
    return l, d, key, rc


def counting_sort(key, nkeys):
    key = np.ascontiguousarray(key, np.int32)
    perm = np.zeros(key.size, np.int32)
    off = np.zeros(int(nkeys) + 1, np.int64)
    rc = lib().orc_counting_sort(key.size, _p(key), int(nkeys), _p(perm), _p(off))
    if rc:
        raise RuntimeError(f"orc_counting_sort rc={rc}")
    return perm, off


def select_cell(pos, ind, n_row, pmin, pmax):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
    ind = np.ascontiguousarray(ind, np.int32).reshape(-1, 4)
    n_row = np.ascontiguousarray(n_row, np.int32)
    pmin = np.ascontiguousarray(pmin, np.float32)
    pmax = np.ascontiguousarray(pmax, np.float32)
    sel = np.zeros(pos.shape[0], np.int32)
    k = lib().orc_select_cell(pos.shape[0], _p(pos), _p(ind), _p(n_row), _p(pmin), _p(pmax),
                              _p(sel))
    return sel[:k]


def select_bulk(pos, pmin, pmax, margin):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
    pmin = np.ascontiguousarray(pmin, np.float32)
    pmax = np.ascontiguousarray(pmax, np.float32)
    sel = np.zeros(pos.shape[0], np.int32)
    k = lib().orc_select_bulk(pos.shape[0], _p(pos), _p(pmin), _p(pmax), float(margin),
                              _p(sel))
    return sel[:k]


def stage(pos, sel, centre, axes, tpe, q0, window, weight, dim=3):
    """a3 staging: returns (x, y, z, c) FP64 / complex128."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
    sel = None if sel is None else np.ascontiguousarray(sel, np.int32)
    M = pos.shape[0] if sel is None else sel.size
    centre = _f64(centre)
    axes = _f64(axes).reshape(3, 3)
    tpe = _f64(tpe)
    q0 = _f64(q0)
    x = np.zeros(M)
    y = np.zeros(M)
    z = np.zeros(M)
    c = np.zeros(M, np.complex128)
    lib().orc_stage(dim, M, _p(pos), _p(sel), _p(centre), _p(axes), _p(tpe), _p(q0),
                    int(window), float(weight), _p(x), _p(y), _p(z), _p(c))
    return x, y, z, c


def sqw_direct(F, wsc=None, wtab=None, mask=None, qfac=None, n_int=None, win_kind=0,
               dt=1.0, nat=1.0, m_lo=0, n_f=None, subtract_mean=True):
    """FP64 reduction.  F: complex [T, nsp, NQ].  Returns (S[n_f, NQ], Stot[NQ])."""
    F = _c128(F)
    T, nsp, NQ = F.shape
    if n_int is None:
        n_int = T
    if n_f is None:
        n_f = n_int
    wsc_a = _f64(wsc if wsc is not None else np.ones(nsp))
    wtab_a = None if wtab is None else _f64(wtab).reshape(nsp, NQ)
    mask_a = np.ascontiguousarray(mask if mask is not None else np.ones(nsp), np.int32)
    qfac_a = None if qfac is None else _f64(qfac).reshape(NQ)
    S = np.zeros((n_f, NQ))
    St = np.zeros(NQ)
    rc = lib().orc_sqw_direct(T, NQ, nsp, _p(F), _p(wsc_a), _p(wtab_a), _p(mask_a),
                              _p(qfac_a), int(n_int), int(win_kind), float(dt), float(nat),
                              int(m_lo), int(n_f), int(bool(subtract_mean)), _p(S), _p(St))
    if rc:
        raise RuntimeError(f"orc_sqw_direct rc={rc}")
    return S, St


# ---------------------------------------------------------------------------
# FINUFFT-method restatement on the CPU (the "host FINUFFT-method
# restatement" timed beside the GPU; FINUFFT itself is not available).
# ---------------------------------------------------------------------------

def iqt(F, wsc=None, wtab=None, mask=None, qfac=None, n_int=None, lag_lo=0, n_lag=1, subtract_mean=True,
        scaled=True, nat=1.0):
    """FP64 intermediate function I(Q,t): [n_lag, NQ] (see orc_iqt)."""
    F = _c128(F)
    T, nsp, NQ = F.shape
    n_int = T // 2 if n_int is None else int(n_int)
    wsc_a = _f64(wsc if wsc is not None else np.ones(nsp))
    wtab_a = None if wtab is None else _f64(wtab).reshape(nsp, NQ)
    mask_a = np.ascontiguousarray(mask if mask is not None else np.ones(nsp), np.int32)
    qfac_a = None if qfac is None else _f64(qfac).reshape(NQ)
    out = np.zeros((n_lag, NQ))
    rc = lib().orc_iqt(T, NQ, nsp, _p(F), _p(wsc_a), _p(wtab_a), _p(mask_a), _p(qfac_a), int(n_int), int(lag_lo),
                       int(n_lag), int(bool(subtract_mean)), int(bool(scaled)), float(nat), _p(out))
    if rc:
        raise RuntimeError(f"orc_iqt rc={rc}")
    return out


def speckle_pass(v, s_trig):
    """One sequential pass of the speckle filter over a flat array (copy)."""
    v = _f64(np.array(v, np.float64, copy=True).ravel())
    lib().orc_speckle_pass(v.size, _p(v), float(s_trig))
    return v


def oneph(pos, at_ind, base, n_row, n_bz, j_plane=1, q_min=(0.0, 0.0)):
    """FP64 one-phonon amplitudes [T, nsp, NQ] (see orc_oneph)."""
    pos = np.ascontiguousarray(pos, np.float32)
    T = pos.shape[0]
    base = np.ascontiguousarray(base, np.float32).reshape(-1, 3)
    nsp = base.shape[0]
    nr = np.ascontiguousarray(n_row, np.int32)
    ind = np.ascontiguousarray(at_ind, np.int32)
    NQ = n_bz * int(nr[0]) * n_bz * int(nr[1])
    out = np.zeros((T, nsp, NQ), np.complex128)
    rc = lib().orc_oneph(T, nsp, _p(nr), int(n_bz), int(j_plane), float(q_min[0]), float(q_min[1]),
                         _p(pos), _p(ind), _p(base), _p(out))
    if rc:
        raise RuntimeError(f"orc_oneph rc={rc}")
    return out


def dos(vel, at_ind, nsp, n_freq):
    """FP64 mp_dos spectra [n_freq, nsp] (see orc_dos)."""
    vel = np.ascontiguousarray(vel, np.float32)
    T, n_tot = vel.shape[0], vel.shape[1]
    ind = np.ascontiguousarray(at_ind, np.int32)
    out = np.zeros((int(n_freq), int(nsp)))
    rc = lib().orc_dos(T, n_tot, _p(vel), _p(ind), int(nsp), int(n_freq), _p(out))
    if rc:
        raise RuntimeError(f"orc_dos rc={rc}")
    return out


def philox4x32_10(ctr, key):
    """Philox4x32-10 of one 128-bit counter under a 64-bit key (see orc_philox4x32_10)."""
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def sql_hist_mc(amp, n_modes, dq, a_cell_inv, q_range, q_step, n_hist, n_h, n_mc, seed, tag=0):
    """mp_sql's Monte Carlo histogram of one frame (see orc_sql_hist_mc):
    amp complex [nsp, N3, N2, N1] (or [nsp, nmodes]); returns (hist
    [n_hist, npairs], (skip1, skip2))."""
    amp = np.ascontiguousarray(amp, np.complex128)
    nsp = amp.shape[0]
    N = np.ascontiguousarray(n_modes, np.int64)
    dq = np.ascontiguousarray(dq, np.float32)
    ai = np.ascontiguousarray(a_cell_inv, np.float32).reshape(-1)
    hist = np.zeros((int(n_hist), nsp * (nsp + 1) // 2))
    skip = np.zeros(2, np.int64)
    rc = lib().orc_sql_hist_mc(nsp, _p(N), _p(dq), _p(ai), float(q_range), float(q_step), int(n_hist), int(n_h),
                               int(n_mc), int(seed), int(tag), _p(amp), _p(hist), _p(skip))
    if rc:
        raise RuntimeError(f"orc_sql_hist_mc rc={rc}")
    return hist, (int(skip[0]), int(skip[1]))


def es_params(eps: float, single: bool = False):
    """Kernel width and shape for upsampling factor 2, after the published
    FINUFFT rule (Barnett, Magland & af Klinteberg, SISC 2019): w = ceil(-log10(eps/10)),
    2 <= w <= 16, beta = gamma*w with gamma = 2.20/2.26/2.38 for w = 2/3/4 and 2.30 above."""
    w = int(np.ceil(-np.log10(eps / 10.0)))
    w = max(2, min(16, w))
    if single:
        w = min(w, 8)
    gamma = {2: 2.20, 3: 2.26, 4: 2.38}.get(w, 2.30)
    return w, gamma * w


def next_smooth_even(n: int) -> int:
    """Smallest even integer >= n whose only prime factors are 2, 3, 5."""
    n = max(2, int(n))
    if n % 2:
        n += 1
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 2


def fine_size(N: int, w: int) -> int:
    return next_smooth_even(max(2 * N, 2 * w))


def es_phihat(k, n, w, beta, nodes: int = 200):
    """phi_hat(k/n) = 2 * int_0^{w/2} phi(t) cos(2 pi k t / n) dt by
    Gauss-Legendre quadrature in FP64."""
    t, wt = np.polynomial.legendre.leggauss(nodes)
    hw = 0.5 * w
    t = 0.5 * hw * (t + 1.0)
    wt = 0.5 * hw * wt
    phi = np.exp(beta * (np.sqrt(np.maximum(0.0, 1.0 - (t / hw) ** 2)) - 1.0))
    k = np.asarray(k, np.float64)
    return 2.0 * (np.cos(2.0 * np.pi * np.outer(k, t) / n) * (phi * wt)).sum(axis=1)


@functools.lru_cache(maxsize=64)
def _deconv(N, n, w, beta):
    """1 / phi_hat on the kept modes k = -N/2 .. (N-1)/2 (cached per plan)."""
    k = np.arange(-(N // 2), (N - 1) // 2 + 1)
    return 1.0 / es_phihat(k, n, w, beta)


def cpu_nufft_type1(x, y, z, c, n_modes, eps=1e-6, iflag=-1, workers=None):
    """Type-1 NUFFT by the FINUFFT method on the host: FP64 ES spreading
    (C, OpenMP), scipy.fft (pocketfft, multi-threaded), deconvolution.
    Returns complex128 of shape reversed(n_modes)."""
    import scipy.fft as sfft

    n_modes = tuple(int(v) for v in n_modes)
    dim = len(n_modes)
    w, beta = es_params(eps)
    nf = [fine_size(N, w) for N in n_modes] + [1] * (3 - dim)
    x, y, z = _coords(dim, x, y, z)
    c = _c128(c)
    grid = np.zeros(nf[0] * nf[1] * nf[2], np.complex128)
    nfa = np.array(nf, np.int64)
    rc = lib().orc_es_spread(dim, x.size, _p(x), _p(y), _p(z), _p(c), _p(nfa), w, beta,
                             _p(grid))
    if rc:
        raise RuntimeError(f"orc_es_spread rc={rc}")
    g = grid.reshape(tuple(reversed(nf[:dim])))
    if workers is None:
        workers = num_threads()
    if iflag >= 0:
        G = sfft.ifftn(g, workers=workers, norm="forward")
    else:
        G = sfft.fftn(g, workers=workers)
    out = G
    for ax_rev, (N, n) in enumerate(zip(reversed(n_modes), reversed(nf[:dim]))):
        k = np.arange(-(N // 2), (N - 1) // 2 + 1)
        idx = np.mod(k, n)
        out = np.take(out, idx, axis=ax_rev)
        shape = [1] * dim
        shape[ax_rev] = N
        out = out * _deconv(N, n, w, beta).reshape(shape)
    return out
