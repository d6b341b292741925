"""The allocation-free |Q| histogram (mpsq_sql_hist_noalloc, partial slices
in a static device array) against the allocating entry point (mpsq_sql_hist,
partial slices in memory it allocates), through the C ABI: two calls in a
row, the second with a smaller input (fewer modes, fewer bins, one species
fewer), in FP64 and FP32, each to 1e-12 relative (the sums are FP64; the
order of a workgroup's LDS atomics is free); and the size limit as an error
code."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(L, fn, amp, nsp, fp64, nm, dq, ainv, q_range, q_step, n_hist, hist):
    return getattr(L, fn)(C.c_void_p(amp.data_ptr()), nsp, 1 if fp64 else 0, (C.c_int64 * 3)(*nm),
                          (C.c_float * 3)(*dq), (C.c_float * 9)(*ainv.reshape(-1)), C.c_float(q_range),
                          C.c_float(q_step), n_hist, C.c_void_p(hist.data_ptr()), None)


@pytest.mark.parametrize("fp64", [True, False])
def test_noalloc_hist_equals_allocating(gpu_lib, fp64):
    import torch
    L = gpu_lib
    rng = np.random.default_rng(17 if fp64 else 18)
    ainv = np.array([[1.0, 0.1, 0.0], [0.0, 0.95, 0.05], [0.02, 0.0, 1.05]], np.float32)
    cdt = torch.complex128 if fp64 else torch.complex64
    cases = [((48, 40, 36), 3, (0.21, 0.19, 0.23), 9.0, 0.01),    # 901 bins x 6 pairs
             ((20, 18, 16), 2, (0.3, 0.3, 0.3), 3.0, 0.02)]        # the second call: smaller
    for nm, nsp, dq, q_range, q_step in cases:
        n_hist = int(round(q_range / q_step)) + 1
        npairs = nsp * (nsp + 1) // 2
        nmodes = int(np.prod(nm))
        a = rng.standard_normal((nsp, nmodes)) + 1j * rng.standard_normal((nsp, nmodes))
        amp = torch.tensor(a, dtype=cdt, device="cuda")
        h_ref = torch.zeros((n_hist, npairs), dtype=torch.float64, device="cuda")
        h_new = torch.zeros_like(h_ref)
        assert _call(L, "mpsq_sql_hist", amp, nsp, fp64, nm, dq, ainv, q_range, q_step, n_hist, h_ref) == 0
        assert _call(L, "mpsq_sql_hist_noalloc", amp, nsp, fp64, nm, dq, ainv, q_range, q_step, n_hist, h_new) == 0
        torch.cuda.synchronize()
        ref, new = h_ref.cpu().numpy(), h_new.cpu().numpy()
        assert np.abs(ref).sum() > 0
        assert np.linalg.norm(new - ref) <= 1e-12 * np.linalg.norm(ref)
        # accumulates like the allocating entry point
        assert _call(L, "mpsq_sql_hist_noalloc", amp, nsp, fp64, nm, dq, ainv, q_range, q_step, n_hist, h_new) == 0
        torch.cuda.synchronize()
        assert np.linalg.norm(h_new.cpu().numpy() - 2 * ref) <= 1e-12 * np.linalg.norm(ref)


def test_noalloc_hist_size_limit(gpu_lib):
    import torch
    from mp_tools_amd import _lib
    amp = torch.zeros((3, 8 * 8 * 8), dtype=torch.complex128, device="cuda")
    ainv = np.eye(3, dtype=np.float32)
    n_hist = 8192 // 6 + 1                     # 6 pairs x n_hist > 8192 doubles of LDS
    hist = torch.zeros((n_hist, 6), dtype=torch.float64, device="cuda")
    rc = _call(gpu_lib, "mpsq_sql_hist_noalloc", amp, 3, True, (8, 8, 8), (0.1, 0.1, 0.1), ainv, 1.0, 0.001,
               n_hist, hist)
    assert rc == _lib.ERR_SIZE
