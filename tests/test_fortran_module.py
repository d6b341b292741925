"""The Fortran module's interfaces to the host-side and argument-checking
paths, without a device: an amdflang program built against
mp_tools_amd/fortran/mpsq_mod.f90 runs the speckle filter
(mp_sqom56.f90:1530-1570) on a Fortran array and gets the oracle's pass,
and the device entry points (mp_sql's integration loops, mp_dos) reject
null device pointers with MPSQ_ERR_BAD_ARG through the same interfaces."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
LIB = os.path.join(ROOT, "mp_tools_amd", "libmpsq.so")

PROGRAM = """
program t
  use mpsq
  use, intrinsic :: iso_c_binding
  implicit none
  real(c_double) :: m(3, 4)
  integer(c_int) :: rc
  integer(c_int64_t) :: nm(3)
  real(c_float) :: dq(3), ai(9)
  m = reshape((/ 5d0, 60d0, 4d0, 3d0, 0.1d0, 2d0, 1d0, 1d0, 9d0, 2d0, 2d0, 40d0 /), (/ 3, 4 /))
  rc = mpsq_speckle_filter(m, 1_c_int64_t, 12_c_int64_t, 0_c_int, 3d0)
  write (*, '(i3, 12(1x, es24.16))') rc, m
  nm = 9
  dq = 0.1
  ai = (/ 1., 0., 0., 0., 1., 0., 0., 0., 1. /)
  rc = mpsq_sql_hist(c_null_ptr, 1_c_int, 1_c_int, nm, dq, ai, 1.0, 0.1, 11_c_int, c_null_ptr, c_null_ptr)
  write (*, '(i3)') rc
  rc = mpsq_sql_hist_mc(c_null_ptr, 1_c_int, 1_c_int, nm, dq, ai, 1.0, 0.1, 11_c_int, 1_c_int64_t, &
                        10_c_int64_t, 7_c_int64_t, 0_c_int32_t, c_null_ptr, c_null_ptr, c_null_ptr)
  write (*, '(i3)') rc
  rc = mpsq_dos(c_null_ptr, 8_c_int64_t, 4_c_int64_t, c_null_ptr, 1_c_int, 5_c_int64_t, 0_c_int, &
                c_null_ptr, c_null_ptr)
  write (*, '(i3)') rc
end program t
"""


@pytest.mark.skipif(FC is None or not os.path.exists(LIB), reason="amdflang or libmpsq.so missing")
def test_fortran_module_speckle_and_argument_errors(tmp_path):
    src = os.path.join(ROOT, "mp_tools_amd", "fortran", "mpsq_mod.f90")
    subprocess.check_call([FC, "-O2", "-c", src, "-o", str(tmp_path / "mpsq_mod.o"), "-J", str(tmp_path)])
    (tmp_path / "t.f90").write_text(PROGRAM)
    exe = str(tmp_path / "t")
    subprocess.check_call([FC, "-O2", "-I", str(tmp_path), str(tmp_path / "t.f90"), str(tmp_path / "mpsq_mod.o"),
                           "-L", os.path.dirname(LIB), "-lmpsq", f"-Wl,-rpath,{os.path.dirname(LIB)}", "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    first = out[0].split()
    assert int(first[0]) == 0
    got = np.array([float(v) for v in first[1:]])
    # a column-major (3, 4) array walked first index fastest: its memory order
    want = O.speckle_pass([5.0, 60.0, 4.0, 3.0, 0.1, 2.0, 1.0, 1.0, 9.0, 2.0, 2.0, 40.0], 3.0)
    assert np.array_equal(got, want)
    assert [int(v) for v in out[1:4]] == [2, 2, 2]
