"""The Fortran drop-in: an amdflang program calls libmpsq exactly as the
reference calls FINUFFT,
    call finufft2d1(n_fft,xf,yf,cf,iflag,eps_fft,n_qx8,n_qy8,ampl_tot,nul_opt,ier)
(code/src/mp_sqom56.f90:1148, argument kinds :94-106), and through the
ISO_C_BINDING plan interface; both results are checked against the FP64
direct DFT oracle (tolerance 10 * eps relative l2)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mp_tools_amd", "fortran", "dropin_sqom")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (amdflang missing)")
def test_fortran_finufft2d1_call_shape(tmp_path):
    rng = np.random.default_rng(12)
    nj, ms, mt, eps = 7000, 160, 160, 1e-6    # the BZO example's 160 x 160 maps
    x = rng.uniform(-np.pi, np.pi, nj)
    y = rng.uniform(-np.pi, np.pi, nj)
    c = np.exp(-1j * rng.uniform(0, 2 * np.pi, nj))
    fin = tmp_path / "in.bin"
    fout = tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        np.array([nj, ms, mt], np.int64).tofile(fh)
        np.array([eps], np.float64).tofile(fh)
        x.tofile(fh)
        y.tofile(fh)
        c.astype(np.complex128).tofile(fh)
    print(f"running {DRIVER} nj={nj} ms={ms} mt={mt}", flush=True)
    subprocess.run([DRIVER, str(fin), str(fout)], check=True, timeout=120)
    raw = open(fout, "rb").read()
    ier, rc = np.frombuffer(raw[:8], np.int32)
    f = np.frombuffer(raw[8:], np.complex128).reshape(3, mt, ms)
    assert ier == 0 and rc == 0
    ref = O.dirft_type1(x, y, None, c, (ms, mt))
    assert rel_l2(f[0], ref) < 10 * eps
    assert rel_l2(f[1], ref) < 10 * eps
    assert rel_l2(f[2], 2 * ref) < 10 * eps
