"""mp_sqom's per-Q tables of a Q plane (mp_sqom56.f90:909-936): q2_norm =
1/|Q| (r.l.u., 1 at Q = 0) and the X-ray form-factor weights x_ffq = c +
sum_k a_k exp(-0.25 b_k q_sq') with q_sq' = |Q'|^2 / 4 (1/A).

CPU: the oracle's restatement (orc_qtables) against closed forms -- 1/|Q| on
an axis-aligned plane computed in FP64, the Q = 0 pixel, a vertical
component of q_center, the form factors at Q = 0 against the atomic numbers
of the reference's coefficient table (code/ref/xray_ff.txt: O, Zr, Ba), and
the decay of a single Gaussian term.  GPU: mpsq_qtables against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O

# a1, b1, a2, b2, a3, b3, a4, b4, c rows of code/ref/xray_ff.txt (O, Zr, Ba)
FF = np.array([[3.0485, 13.2771, 2.2868, 5.7011, 1.5463, 0.3239, 0.867, 32.9089, 0.2508],
               [17.8765, 1.27618, 10.948, 11.916, 5.41732, 0.117622, 3.65721, 87.6627, 2.06929],
               [20.3361, 3.216, 19.297, 0.2756, 10.888, 20.2073, 2.6959, 167.202, 2.7731]], np.float32)
Z = np.array([8, 40, 56])
A = 4.19


def _plane(q_center=(0.0, 0.0, 0.0), e1=(1, 0, 0), e2=(0, 1, 0)):
    from mp_tools_amd.sqw import qplane
    return qplane((32, 32, 32), 5.0, e1, e2, q_center=q_center, a_cell_inv=np.eye(3) / A)


def _qcoords(g):
    i = np.arange(g["n_qx"], dtype=np.float64)
    j = np.arange(g["n_qy"], dtype=np.float64)
    ax = g["qx_range"][0] + i * (float(g["bz_nx"]) / g["n_qx"])
    ay = g["qy_range"][0] + j * (float(g["bz_ny"]) / g["n_qy"])
    AY, AX = np.meshgrid(ay, ax, indexing="ij")
    return AX.ravel(), AY.ravel()


def test_q2_norm_closed_forms():
    g = _plane()
    assert g["n_qx"] == 160 and g["n_qy"] == 160
    q2, _ = O.qtables(g)
    ax, ay = _qcoords(g)
    qq = np.sqrt(ax ** 2 + ay ** 2)
    zero = qq < 1e-6
    assert zero.sum() == 1 and q2[zero][0] == 1.0           # Q = 0: 1 (mp_sqom56.f90:914-917)
    np.testing.assert_allclose(q2[~zero], 1.0 / qq[~zero], rtol=2e-6)
    # a vertical component of q_center adds |q_v|^2 (mp_sqom56.f90:911-913)
    g = _plane(q_center=(0.5, 0.25, 1.5))
    assert np.allclose(g["q_v"], [0, 0, 1.5])
    q2, _ = O.qtables(g)
    ax, ay = _qcoords(g)
    np.testing.assert_allclose(q2, 1.0 / np.sqrt(ax ** 2 + ay ** 2 + 2.25), rtol=2e-6)


def test_form_factors_closed_forms():
    g = _plane()
    _, xq = O.qtables(g, FF)
    ax, ay = _qcoords(g)
    zero = np.hypot(ax, ay) < 1e-6
    # at Q = 0 the form factor is the electron count of the (neutral) atom
    assert np.all(np.abs(xq[:, zero][:, 0] - Z) < 0.05)   # the table's own rounding: 7.9994, 39.968, 55.990
    # one Gaussian term: c + a exp(-0.25 b q_sq'), q_sq' = |Q'|^2 / 4 with Q' in 1/A
    one = np.array([[2.0, 10.0, 0, 0, 0, 0, 0, 0, 0.5]], np.float32)
    _, x1 = O.qtables(g, one)
    qa2 = (2 * np.pi / A) ** 2 * (ax ** 2 + ay ** 2)
    np.testing.assert_allclose(x1[0], 0.5 + 2.0 * np.exp(-0.25 * 10.0 * 0.25 * qa2), rtol=2e-6, atol=1e-7)
    # large |Q|: the form factor decays towards c
    assert np.all(xq[:, ~zero] < Z[:, None])


@pytest.mark.gpu
@pytest.mark.parametrize("qc,e1,e2", [((2, 2, 0), (1, 0, 0), (0, 1, 0)), ((1, 0, 0.5), (1, 1, 0), (0, 0, 1)),
                                      ((0.3, -0.7, 1.1), (1, 2, 0), (0, 1, 1))])
def test_device_tables_vs_oracle(gpu_lib, qc, e1, e2):
    import torch
    from mp_tools_amd.sqw import qtables
    g = _plane(q_center=qc, e1=e1, e2=e2)
    q2, xq = qtables(g, FF)
    torch.cuda.synchronize()
    r2, rx = O.qtables(g, FF)
    np.testing.assert_allclose(q2.cpu().numpy(), r2, rtol=1e-6)
    np.testing.assert_allclose(xq.cpu().numpy(), rx, rtol=2e-6, atol=1e-6)
