"""GPU parity of the type-1 NUFFT (libmpsq.so) against the FP64 oracle.

The oracle is the direct DFT with the definition of the reference's
dirft2d1 (code/src/mp_sqom56.f90:2530-2602); FINUFFT itself is not in the
reference tree, so parity is to the transform's definition at the plan's
tolerance.  The bin-sort permutation is compared bit for bit with the CPU
fold + stable counting sort.

Tolerances (relative l2 over all modes):
    FP64 plans: 10 * eps (the ES kernel's error at w = ceil(-log10(eps/10)))
    FP32 plans: max(10 * eps, 2e-5)  (FP32 roundoff of spread + FFT)
"""
import ctypes as C
import zlib

import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _pts(rng, M, dim, spread=1.0):
    return [rng.uniform(-np.pi * spread, np.pi * spread, M) for _ in range(dim)]


def _tol(dtype, eps):
    return 10 * eps if dtype == "complex128" else max(10 * eps, 2e-5)


CASES = [(1, (64,)), (2, (48, 40)), (3, (32, 24, 16)), (3, (20, 18, 10))]


@pytest.mark.parametrize("dim,nm", CASES)
@pytest.mark.parametrize("dtype,eps", [("complex128", 1e-6), ("complex128", 1e-10),
                                       ("complex64", 1e-6), ("complex64", 1e-3)])
@pytest.mark.parametrize("isign", [-1, 1])
def test_type1_vs_direct_dft(gpu_lib, dim, nm, dtype, eps, isign):
    from mp_tools_amd import nufft
    rng = np.random.default_rng(zlib.crc32(repr((dim, nm, dtype, eps, isign)).encode()))
    M = 3000
    co = _pts(rng, M, dim, spread=1.5)   # points beyond [-pi, pi) fold periodically
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    ref = O.dirft_type1(*(co + [None] * (3 - dim)), c, nm, iflag=isign)
    p = nufft.Plan(1, nm, eps=eps, isign=isign, dtype=dtype)
    p.setpts(*co)
    f = p.execute(c)
    assert f.shape == ref.shape
    err = rel_l2(f, ref)
    assert err < _tol(dtype, eps), f"rel l2 {err:.3e}"


@pytest.mark.parametrize("case", ["dense2d", "clustered3d", "sparse3d"])
def test_fp32_accumulation_modes(gpu_lib, case):
    """FP32 plans spread each tile in 32-bit fixed point when the tile's
    covering bound allows it, else in 64-bit fixed point (a second launch).
    dense2d: every tile takes the 64-bit path; clustered3d: 20000 of 22000
    points in a small ball, so the tiles through the ball (~13%) take the
    64-bit path and the others the 32-bit one in the same call; sparse3d:
    every tile 32-bit.
    Each is checked against the direct DFT and against the FP64 plan."""
    from mp_tools_amd import nufft
    rng = np.random.default_rng(zlib.crc32(case.encode()))
    if case == "dense2d":
        nm, M = (32, 28), 120000
        co = _pts(rng, M, 2)
    elif case == "clustered3d":
        nm, M = (40, 36, 24), 22000
        co = _pts(rng, M, 3)
        k = 20000
        for d in range(3):
            co[d][:k] = rng.normal(0.5, 0.15, k)
    else:
        nm, M = (40, 36, 24), 2000
        co = _pts(rng, M, 3)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    ref = O.dirft_type1(*(co + [None] * (3 - len(nm))), c, nm)
    p32 = nufft.Plan(1, nm, eps=1e-6, dtype="complex64")
    p32.setpts(*co)
    f32 = p32.execute(c)
    p64 = nufft.Plan(1, nm, eps=1e-6, dtype="complex128")
    p64.setpts(*co)
    f64 = p64.execute(c)
    e32, e64 = rel_l2(f32, ref), rel_l2(f64, ref)
    print(f"{case}: FP32 {e32:.3e}, FP64 {e64:.3e} vs direct DFT")
    assert e64 < 1e-5
    assert e32 < 2e-5
    assert rel_l2(f32, f64) < 2e-5


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("dim,nm", [(2, (40, 36)), (3, (24, 20, 12))])
def test_sort_permutation_bit_exact(gpu_lib, dtype, dim, nm):
    """The bin-sort permutation equals the CPU stable counting sort of the
    FP64 fold (keys from l_y / BY, l_z / BZ)."""
    from mp_tools_amd import nufft
    rng = np.random.default_rng(7 + dim)
    M = 20000
    co = _pts(rng, M, dim, spread=2.0)
    p = nufft.Plan(1, nm, eps=1e-6, dtype=dtype)
    if dtype == "complex64":
        co = [v.astype(np.float32) for v in co]
    p.setpts(*co)
    info = p.info()
    w, beta = O.es_params(1e-6, single=(dtype == "complex64"))
    assert info["w"] == w
    assert info["n_fine"][:dim] == tuple(O.fine_size(N, w) for N in nm)
    by, bz = info["sort_bin"]
    nf = info["n_fine"][:dim]
    l, d, key, rc = O.fold(*(list(co) + [None] * (3 - dim)), nf, w, by, bz)
    assert rc == 0
    nkeys = (nf[1] // by) * (nf[2] // bz if dim == 3 else 1)
    perm_ref, _ = O.counting_sort(key, nkeys)
    perm = p.sort_perm()
    assert np.array_equal(perm, perm_ref)


def test_ntrans_and_empty_and_errors(gpu_lib):
    from mp_tools_amd import nufft, _lib
    rng = np.random.default_rng(11)
    M, nm = 2500, (30, 26)
    x, y = _pts(rng, M, 2)
    c = rng.standard_normal((3, M)) + 1j * rng.standard_normal((3, M))
    p = nufft.Plan(1, nm, n_trans=3, eps=1e-9)
    p.setpts(x, y)
    f = p.execute(c)
    assert f.shape == (3, 26, 30)
    for t in range(3):
        assert rel_l2(f[t], O.dirft_type1(x, y, None, c[t], nm)) < 1e-8
    # no points -> all-zero modes
    p1 = nufft.Plan(1, nm, eps=1e-6)
    p1.setpts(np.zeros(0), np.zeros(0))
    assert np.all(p1.execute(np.zeros(0, np.complex128)) == 0)
    # a non-finite coordinate is an error, and the plan forgets its points
    x2 = x.copy()
    x2[5] = np.nan
    with pytest.raises(_lib.MpsqError) as ei:
        p1.setpts(x2, y)
    assert ei.value.code == _lib.ERR_NONFINITE_POINT
    # eps below the FP32 floor -> warning code, result still computed
    pf = nufft.Plan(1, nm, eps=1e-9, dtype="complex64")
    assert pf.warning == _lib.WARN_EPS_TOO_SMALL


def test_device_api_torch(gpu_lib):
    import torch
    from mp_tools_amd import nufft
    rng = np.random.default_rng(5)
    M, nm = 4000, (32, 16, 8)
    co = _pts(rng, M, 3)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    ref = O.dirft_type1(*co, c, nm)
    p = nufft.Plan(1, nm, eps=1e-6, dtype="complex64", stream="torch")
    p.setpts(*[torch.tensor(v, dtype=torch.float32, device="cuda") for v in co])
    f = p.execute(torch.tensor(c, dtype=torch.complex64, device="cuda"))
    torch.cuda.synchronize()
    assert rel_l2(f.cpu().numpy(), ref) < 2e-5


def test_simple_interface_and_fortran_shim(gpu_lib):
    """finufft2d1_ with the reference's by-reference argument list
    (mp_sqom56.f90:1148): int64 nj, double x/y, complex16 c, int iflag,
    double eps, int64 ms/mt, complex16 fk, null opts, int ier."""
    from mp_tools_amd import nufft
    rng = np.random.default_rng(3)
    nj, ms, mt = 5000, 64, 48
    x, y = _pts(rng, nj, 2)
    c = rng.standard_normal(nj) + 1j * rng.standard_normal(nj)
    ref = O.dirft_type1(x, y, None, c, (ms, mt))
    f1 = nufft.nufft2d1(x, y, c, (ms, mt), eps=1e-6)
    assert rel_l2(f1, ref) < 1e-5
    fk = np.zeros((mt, ms), np.complex128)
    args = [C.c_int64(nj), x, y, c, C.c_int(-1), C.c_double(1e-6), C.c_int64(ms), C.c_int64(mt)]
    ier = C.c_int(-7)
    ptr = lambda a: a.ctypes.data if isinstance(a, np.ndarray) else C.addressof(a)
    gpu_lib.finufft2d1_(*[ptr(a) for a in args], fk.ctypes.data, None, C.addressof(ier))
    assert ier.value == 0
    assert rel_l2(fk, ref) < 1e-5
    # 3D shim on the same points (z = 0 plane) equals the 2D transform for N3 = 1
    fk3 = np.zeros((1, mt, ms), np.complex128)
    z = np.zeros(nj)
    args3 = [C.c_int64(nj), x, y, z, c, C.c_int(-1), C.c_double(1e-6), C.c_int64(ms),
             C.c_int64(mt), C.c_int64(1)]
    gpu_lib.finufft3d1_(*[ptr(a) for a in args3], fk3.ctypes.data, None, C.addressof(ier))
    assert ier.value == 0
    assert rel_l2(fk3[0], ref) < 1e-5
