"""C ABI of libmpsq.so on the CPU: the library loads, exports every symbol
include/mpsq.h declares, and follows the FINUFFT error convention when no
HIP device is present (error code, zero-filled outputs, no abort).  No
compute call is made here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "mpsq.h")


@pytest.fixture(scope="module")
def lib():
    from mp_tools_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "mp_tools_amd", "csrc"), "-s", "-j4"])
    return _lib.load()


def declared_functions():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|int64_t|char)\s*\*?\s*(\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_lists_the_abi():
    names = declared_functions()
    for must in ("mpsq_makeplan", "mpsqf_execute", "finufft2d1_", "finufft3d1_", "mpsq_sqw_reduce"):
        assert must in names
    from mp_tools_amd import _lib
    assert sorted(_lib.EXPORTS) == names


def test_every_declared_symbol_is_exported(lib):
    out = subprocess.check_output(["nm", "-D", "--defined-only", os.path.join(ROOT, "mp_tools_amd", "libmpsq.so")],
                                  text=True)
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"not exported: {missing}"
    assert lib.mpsq_version().decode().startswith("mpsq")


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_no_device_error_convention(lib):
    from mp_tools_amd import _lib
    nm = (C.c_int64 * 3)(16, 16, 1)
    h = C.c_void_p(123)
    rc = lib.mpsq_makeplan(1, 2, nm, -1, 1, 1e-6, C.byref(h), None)
    assert rc == _lib.ERR_DEVICE and not h.value
    # the Fortran shim zero-fills fk and reports through ier
    nj, ms, mt = 10, 8, 6
    x = np.zeros(nj)
    c = np.ones(nj, np.complex128)
    fk = np.full((mt, ms), 7.0 + 7.0j)
    ier = C.c_int(-1)
    args = [C.c_int64(nj), C.c_int(-1), C.c_double(1e-6), C.c_int64(ms), C.c_int64(mt)]
    lib.finufft2d1_(C.addressof(args[0]), x.ctypes.data, x.ctypes.data, c.ctypes.data,
                    C.addressof(args[1]), C.addressof(args[2]), C.addressof(args[3]),
                    C.addressof(args[4]), fk.ctypes.data, None, C.addressof(ier))
    assert ier.value == _lib.ERR_DEVICE
    assert np.all(fk == 0)
    spec = _lib.SqwSpec()
    lib.mpsq_sqw_default_spec(C.byref(spec))
    counts = np.array([5, 5], np.int64)
    spec.dim, spec.nsp, spec.n_tot, spec.sp_count = 2, 2, 10, counts.ctypes.data
    spec.n_modes[0], spec.n_modes[1] = 8, 8
    e = C.c_void_p(5)
    assert lib.mpsq_sqw_create(C.byref(spec), C.byref(e)) == _lib.ERR_DEVICE and not e.value


def test_bad_arguments_are_rejected_before_any_device_call(lib):
    from mp_tools_amd import _lib
    h = C.c_void_p()
    nm = (C.c_int64 * 3)(16, 16, 16)
    assert lib.mpsq_makeplan(2, 3, nm, -1, 1, 1e-6, C.byref(h), None) == _lib.ERR_BAD_ARG  # type 2
    assert lib.mpsq_makeplan(1, 3, nm, -1, 0, 1e-6, C.byref(h), None) == _lib.ERR_BAD_ARG  # ntrans 0
    assert lib.mpsq_sqw_create(None, C.byref(h)) == _lib.ERR_BAD_ARG
    assert lib.mpsq_execute(None, None, None) == _lib.ERR_BAD_ARG
    # the entry points of the rows around the S(Q,w) map
    dq = (C.c_float * 3)(0.1, 0.1, 0.1)
    ai = (C.c_float * 9)(1, 0, 0, 0, 1, 0, 0, 0, 1)
    buf = (C.c_double * 64)()
    skip = (C.c_int64 * 2)()
    assert lib.mpsq_sql_hist(None, 1, 1, nm, dq, ai, 1.0, 0.1, 11, buf, None) == _lib.ERR_BAD_ARG
    assert lib.mpsq_sql_hist_mc(None, 1, 1, nm, dq, ai, 1.0, 0.1, 11, 1, 10, 1, 0, buf, skip, None) == _lib.ERR_BAD_ARG
    assert lib.mpsq_sql_hist_mc(buf, 1, 1, nm, dq, ai, 1.0, 0.1, 11, 0, 10, 1, 0, buf, skip, None) == _lib.ERR_BAD_ARG
    assert lib.mpsq_sql_hist_mc(buf, 1, 1, nm, dq, ai, 1.0, 0.1, 11, 1, 10, 1, 0, buf, None, None) == _lib.ERR_BAD_ARG
    assert lib.mpsq_dos(None, 8, 4, buf, 1, 5, 0, buf, None) == _lib.ERR_BAD_ARG
    assert lib.mpsq_dos(buf, 8, 4, buf, 1, 6, 0, buf, None) == _lib.ERR_BAD_ARG          # n_freq > T / 2 + 1
    assert lib.mpsq_dos(buf, 8, 4, buf, 1, 5, 7, buf, None) == _lib.ERR_BAD_ARG          # precision
    assert lib.mpsq_oneph_amplitudes(None, buf, buf, 1, None, None, buf, None) == _lib.ERR_BAD_ARG
    assert lib.mpsq_speckle_filter(None, 4, 4, 0, 2.0) == _lib.ERR_BAD_ARG
    assert lib.mpsq_speckle_filter(buf, 0, 4, 0, 2.0) == _lib.ERR_BAD_ARG
