"""Spherically averaged S(Q) of MD / RMC snapshots: the mp_sql path
(code/src/mp_sql56.f90) on the device.

  1. geometry (mp_sql56.f90:956-1021): d_q = 2 pi / box, a mode box
     n_qq = 2 nint(q_range a_d / d_q) + 1 per axis around Q = 0;
  2. per snapshot and species, the 3D type-1 NUFFT of the atoms at
     x = r * d_q with unit strengths, iflag -1, eps 1e-6 (:1096-1135; the
     reference stacks 2D transforms over k3, which is the same sum);
  3. the Gauss integration (:1224-1254): Re(A_i A_j*) of every mode with
     |Q| <= q_range into bin nint(|Q| / q_step) -- ``mpsq_sql_hist`` --
     and its normalisation (:1266-1273); or the Monte Carlo integration
     (:1146-1215, j_acc = 2) over n_h x 10^6 radially sampled modes --
     ``mpsq_sql_hist_mc`` -- and its normalisation (:1260-1265);
  4. the output functions (:1366-1500): Gaussian smoothing, X-ray /
     neutron / unit weights, F(Q), S(Q), Z(Q) (Faber-Ziman), I(Q),
     unscaled I(Q), as the columns the reference writes (:1812-1829).

Steps 1, 3 (binning) and the transforms run on the GPU through libmpsq;
the per-bin arithmetic of step 4 (a few thousand bins) is host code, as
in the reference.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib

MODES = {3: "F(Q)", 4: "S(Q)", 5: "Z(Q)", 6: "I(Q)", 7: "I(Q)_unscaled"}


@dataclass
class SqlGeometry:
    dq: np.ndarray          # float32 [3], Q step per mode index (1/A)
    a_cell_inv: np.ndarray  # float32 [3, 3]
    n_qq: tuple             # modes per axis (odd)
    q_range: np.float32
    q_step: np.float32
    n_hist: int
    at_volume: float


def geometry(a_par, a_cell_inv, n_row, n_tot, q_range: float, q_step: float = 0.01) -> SqlGeometry:
    """mp_sql56.f90:956-1021 in the reference's FP32 arithmetic."""
    f32 = np.float32
    a_par = np.asarray(a_par, f32)
    n_row = np.asarray(n_row)
    nsuper = int(np.prod(n_row))
    span = a_par if nsuper == 1 else n_row.astype(f32)          # at_pos_max - at_pos_min
    dq = (2.0 * math.pi / span.astype(np.float64)).astype(f32)
    ainv = np.asarray(a_cell_inv, f32)
    q_range, q_step = f32(q_range), f32(q_step)
    bz_n = f32(2.0) * q_range
    e = [ainv[:, k].astype(f32) for k in range(3)]
    ev = [np.cross(e[1], e[2]).astype(f32), np.cross(e[2], e[0]).astype(f32), np.cross(e[0], e[1]).astype(f32)]
    n_qq = []
    for d in range(3):
        height = f32(np.dot(e[d], ev[d]) / np.linalg.norm(ev[d]).astype(f32))
        bz = f32(bz_n / height)
        n_qq.append(2 * int(np.rint(f32(f32(0.5) * bz) / dq[d])) + 1)
    n_hist = 1 + int(f32(q_range / q_step))
    at_volume = float(nsuper) * float(np.prod(a_par.astype(np.float64))) / float(n_tot)
    return SqlGeometry(dq, ainv, tuple(n_qq), q_range, q_step, n_hist, at_volume)


N_MC = 1_000_000     # samples per Monte Carlo chain (mp_sql56.f90:40)


def n_h_max(geo: SqlGeometry) -> int:
    """Most Monte Carlo chains per frame, int(.5 prod(n_qq) / n_mc) in REAL(4)
    (mp_sql56.f90:1063)."""
    prod = np.float32(1.0)
    for n in geo.n_qq:
        prod = np.float32(prod * np.float32(n))
    return int(np.float32(np.float32(0.5) * prod) / np.float32(N_MC))


def mc_normalise(h: np.ndarray, n_sup_total: int, geo: SqlGeometry, nfile: int, n_int: int, skips) -> np.ndarray:
    """sq_hist from the summed Monte Carlo histogram h [nsp, nsp, n_hist]:
    amplitudes / sqrt(sum n_sup) (mp_sql56.f90:1153), then bins 2.. over
    q_step nfile and times q_range / (n_int - n_skip1 - n_skip2)
    (:1260-1265).  The reference resets the skip counts per frame
    (:1150-1151) and normalises after the frame loop, so ``skips`` are the
    last frame's."""
    sq = np.array(h, np.float64) / float(n_sup_total)
    sq[:, :, 1:] /= float(geo.q_step) * nfile
    sq[:, :, 1:] *= float(geo.q_range) / float(n_int - skips[0] - skips[1])
    return sq


def histogram(frames, counts, geo: SqlGeometry, eps: float = 1e-6, precision: str = "fp64",
              device: int = 0, timing: dict | None = None, integration: str = "gauss", n_h: int = 1,
              seed: int = 1, vary_points: bool = False, n_mc: int = N_MC) -> np.ndarray:
    """Normalised partial histograms sq_hist[i, j, bin] of the snapshots
    ``frames`` (each float [n_tot, >=3] positions, species-major with
    ``counts`` atoms per species): Gauss integration (mp_sql56.f90:1224-1273)
    or, with integration="mc", the Monte Carlo integration over n_h chains of
    n_mc samples per frame (:1146-1215, :1260-1265).  The reference's default
    seeding (j_rand = 1, :1160-1166) restarts every chain from the same seed
    in each frame, so every frame samples the same modes; vary_points=True
    draws new ones per frame (the frame index is the generator's tag).
    ``timing`` receives the device time of the loop ("ms") and, for "mc",
    the last frame's skip counts ("skips")."""
    import torch

    if integration not in ("gauss", "mc"):
        raise ValueError("integration must be 'gauss' or 'mc'")
    mc = integration == "mc"
    if mc and not 1 <= n_h:
        raise ValueError("n_h must be >= 1")
    if mc and n_mc == N_MC and n_h > n_h_max(geo):
        raise ValueError(f"n_h exceeds the {n_h_max(geo)} chains of 10^6 samples this Q box allows (mp_sql56.f90:1069)")

    from .nufft import Plan
    L = _lib.load()
    nsp = len(counts)
    npairs = nsp * (nsp + 1) // 2
    dev = torch.device("cuda", device)
    dtype = "complex128" if precision == "fp64" else "complex64"
    cdt = torch.complex128 if precision == "fp64" else torch.complex64
    N1, N2, N3 = geo.n_qq
    nq = N1 * N2 * N3
    hist = torch.zeros(geo.n_hist * npairs, dtype=torch.float64, device=dev)
    amp = torch.empty((nsp, nq), dtype=cdt, device=dev)
    off = np.concatenate([[0], np.cumsum(counts)])
    stream = torch.cuda.current_stream(dev)
    plans = [Plan(1, geo.n_qq, eps=eps, isign=-1, dtype=dtype, device=device, stream="torch") for _ in range(nsp)]
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    skip = torch.zeros(2, dtype=torch.int64, device=dev)
    nfile = 0
    for ifile, pos in enumerate(frames):
        nfile += 1
        pos = np.asarray(pos)
        for s in range(nsp):
            r = pos[off[s]:off[s + 1], :3].astype(np.float32)
            # x = at_pos * d_q in FP32 (mp_sql56.f90:1115-1118), handed on as FP64
            xyz = [torch.tensor((r[:, d] * geo.dq[d]).astype(np.float64), device=dev) for d in range(3)]
            if precision != "fp64":
                xyz = [v.float() for v in xyz]
            plans[s].setpts(*xyz)
            c = torch.ones(len(r), dtype=cdt, device=dev)
            amp[s] = plans[s].execute(c).reshape(-1)
        nm = (C.c_int64 * 3)(N1, N2, N3)
        dq = (C.c_float * 3)(*[float(v) for v in geo.dq])
        ai = (C.c_float * 9)(*[float(v) for v in geo.a_cell_inv.reshape(-1)])
        prec = 1 if precision == "fp64" else 0
        if mc:
            skip.zero_()
            rc = L.mpsq_sql_hist_mc(C.c_void_p(amp.data_ptr()), nsp, prec, nm, dq, ai, float(geo.q_range),
                                    float(geo.q_step), geo.n_hist, int(n_h), int(n_mc), int(seed),
                                    ifile if vary_points else 0, C.c_void_p(hist.data_ptr()),
                                    C.c_void_p(skip.data_ptr()), C.c_void_p(stream.cuda_stream))
            _lib.check("sql_hist_mc", rc)
        else:
            rc = L.mpsq_sql_hist(C.c_void_p(amp.data_ptr()), nsp, prec, nm, dq, ai, float(geo.q_range),
                                 float(geo.q_step), geo.n_hist, C.c_void_p(hist.data_ptr()),
                                 C.c_void_p(stream.cuda_stream))
            _lib.check("sql_hist", rc)
    ev1.record(stream)
    torch.cuda.synchronize(dev)
    skips = tuple(int(v) for v in skip.cpu().tolist())
    if timing is not None:
        timing["ms"] = ev0.elapsed_time(ev1)
        timing["frames"] = nfile
        if mc:
            timing["skips"] = skips
    h = hist.cpu().numpy().reshape(geo.n_hist, npairs)
    sq = np.zeros((nsp, nsp, geo.n_hist))
    p = 0
    for i in range(nsp):
        for j in range(i + 1):
            sq[i, j] = sq[j, i] = h[:, p]
            p += 1
    if mc:
        return mc_normalise(sq, int(np.sum(counts)), geo, nfile, int(n_h) * int(n_mc), skips)
    # amplitudes / sum(n_sup) (:1231), then the Gauss normalisation (:1266-1272)
    sq /= float(np.sum(counts)) ** 2
    qh = np.arange(geo.n_hist, dtype=np.float64) * float(geo.q_step)
    sq[:, :, 1:] /= float(geo.q_step) * qh[1:] ** 2
    sq[:, :, 1:] *= 2.0 * math.pi ** 2 / (nfile * geo.at_volume)
    return sq


def xray_ff(ffpar: np.ndarray, q: np.ndarray) -> np.ndarray:
    """f(Q) = c + sum_k a_k exp(-b_k (Q / 4 pi)^2) per species; ffpar rows
    (a1, b1, a2, b2, a3, b3, a4, b4, c) (mp_sql56.f90:1051-1059)."""
    ffpar = np.atleast_2d(np.asarray(ffpar, np.float64))
    q_sq = 0.25 * (np.asarray(q, np.float64) / (2.0 * math.pi)) ** 2
    f = np.repeat(ffpar[:, 8:9], len(q_sq), axis=1)
    for k in range(4):
        f += ffpar[:, 2 * k:2 * k + 1] * np.exp(-ffpar[:, 2 * k + 1:2 * k + 2] * q_sq[None, :])
    return f


def output(sq_hist: np.ndarray, geo: SqlGeometry, mode: int, occ, weighting: str = "xray", ffpar=None,
           b_coh=None, smooth_fwhm: int = 1, masks=None, scf=None, mixed: bool = True):
    """The reference's output columns (mp_sql56.f90:1366-1500, written at
    :1812-1829): returns (q, total, partials) over the written rows, with
    partials[:, p] for (1,1), (2,1), (2,2), (3,1), ... and total None for Z(Q).
    mixed: all species share the sites (amorphous / Faber-Ziman type)."""
    if mode not in MODES:
        raise ValueError(f"mode must be one of {sorted(MODES)}")
    nsp, n_hist = sq_hist.shape[0], sq_hist.shape[2]
    occ = np.asarray(occ, np.float64)
    occ = occ / occ.sum()
    masks = np.ones(nsp) if masks is None else np.asarray(masks, np.float64)
    scf = np.ones(nsp) if scf is None else np.asarray(scf, np.float64)
    q_step = float(geo.q_step)
    qh = np.arange(n_hist, dtype=np.float64) * q_step
    i0 = int(np.float32(0.5) / geo.q_step) - 1          # 0-based first processed bin (Fortran int(.5/q_step))
    # Gaussian smoothing (:1370-1404)
    n_smooth = 1 if smooth_fwhm <= 1 else 5 * smooth_fwhm // 2 + 1
    fs = np.array([2.0 ** (-(((j - n_smooth // 2 - 1) / (0.5 * smooth_fwhm)) ** 2)) for j in range(1, n_smooth + 1)])
    fs /= fs.sum()
    sp = np.zeros_like(sq_hist)
    for i in range(i0, n_hist):                           # 0-based i <-> Fortran i + 1
        if n_smooth == 1:
            sp[:, :, i] = sq_hist[:, :, i]
            continue
        for j in range(1, n_smooth + 1):
            idx = (i + 1) - n_smooth // 2 + j - 1         # Fortran index
            if idx <= 2:
                continue
            src = (n_hist - j) if idx >= n_hist else idx
            sp[:, :, i] += sq_hist[:, :, src - 1] * fs[j - 1]
    sp[:, :, :i0 + 1] = 0.0
    if weighting == "xray":
        f = xray_ff(ffpar, qh)                            # [nsp, n_hist]
        if mode != 5:
            for i in range(i0, n_hist):
                fi = f[:, i]
                av_sq = np.sum(occ * scf * fi) ** 2
                sq_av = np.sum(occ * (scf * fi ** 2))
                wv = masks * scf * fi
                sp[:, :, i] *= np.outer(wv, wv)
                if mode == 6:
                    sp[:, :, i] /= sq_av
                if mixed and mode < 5:
                    for a in range(nsp):
                        sp[a, a, i] -= occ[a] * masks[a] * (scf[a] * fi[a]) ** 2
                    if mode == 4:
                        ow = occ * masks * scf * fi
                        sp[:, :, i] += np.outer(ow, ow)
                    sp[:, :, i] /= av_sq
        elif mixed:
            for a in range(nsp):
                sp[a, a, :] -= occ[a]
            sp /= np.outer(occ, occ)[:, :, None]
    else:
        w = np.ones(nsp) if weighting == "unit" else np.asarray(b_coh, np.float64)
        # average weights for mixed occupancies, updated in place as the
        # reference's loop does (:1349-1352)
        avm = np.ones((nsp, nsp)) if mixed else np.eye(nsp)
        for j in range(nsp):
            w[j] = np.sum(w * occ * avm[j]) / np.sum(occ * avm[j])
        sq_av = np.sum(occ * (scf * w) ** 2)
        av_sq = np.sum(occ * scf * w) ** 2
        if mode != 5:
            wm = np.outer(w, w) * np.outer(masks * scf, masks * scf)
            sp[:, :, i0:] *= wm[:, :, None]
            if mode == 6:
                sp /= sq_av
            if mixed and mode < 5:
                for a in range(nsp):
                    sp[a, a, 1:] -= occ[a] * masks[a] * (scf[a] * w[a]) ** 2
                if mode == 4:
                    sp += (np.outer(occ, occ) * wm)[:, :, None]
                sp /= av_sq
        elif mixed:
            for a in range(nsp):
                sp[a, a, :] -= occ[a]
            sp /= np.outer(occ, occ)[:, :, None]
    total = None
    if mode != 5:
        total = sp.sum(axis=(0, 1))
        for a in range(nsp):
            for b in range(a):
                sp[a, b] = sp[a, b] + sp[b, a]
                sp[b, a] = sp[a, b]
    cols = [sp[a, b] for a in range(nsp) for b in range(a + 1)]
    i_start = i0 + n_smooth // 2                            # 0-based first written row
    rows = slice(i_start, n_hist)
    q = (np.arange(n_hist, dtype=np.float32) * geo.q_step)[rows]
    return q, (None if total is None else total[rows]), np.stack([c[rows] for c in cols], axis=1)


# names the text file gives the output modes (pdf_out, mp_sql56.f90:171)
TEXT_NAMES = {3: "F(Q)", 4: "S(Q)", 5: "Z(Q)", 6: "I(Q)", 7: "I(Q)"}


def header_amplitudes(weighting: str, occ, ffpar=None, b_coh=None, scf=None, mixed: bool = True) -> np.ndarray:
    """at_scf * at_weight as the text header prints them (mp_sql56.f90:1309-1352):
    the species weight (1, b_coh, or the X-ray form factor at the first Q
    point, Q = 0), averaged in place over mixed occupancies, REAL(4)."""
    occ = np.asarray(occ, np.float32)
    occ = occ / occ.sum()
    nsp = len(occ)
    scf = np.ones(nsp, np.float32) if scf is None else np.asarray(scf, np.float32)
    if weighting == "xray":
        w = xray_ff(ffpar, np.zeros(1))[:, 0].astype(np.float32)
    elif weighting == "neutron":
        w = np.asarray(b_coh, np.float32).copy()
    else:
        w = np.ones(nsp, np.float32)
    avm = np.ones((nsp, nsp), np.float32) if mixed else np.eye(nsp, dtype=np.float32)
    for j in range(nsp):
        w[j] = np.float32(np.sum(w * occ * avm[j]) / np.sum(occ * avm[j]))
    return (scf * w).astype(np.float32)


def table_text(q, total, partials, mode: int, *, at_names, occ, amplitudes, q_step: float, smooth_fwhm: int = 1,
               subst_name: str = "", shell: str = "", mp_tool: str = "MP_SQL 1.56", time_stamp: str = "",
               file_dat_t0: str = "", file_dat: str = "", nfile_step: int = 1, proc_num: int = 1,
               n_row=(1, 1, 1), a_par=(1.0, 1.0, 1.0), scf=None, masks=None) -> str:
    """The text of mp_sql's output file (mp_sql56.f90:1800-1829) for the columns
    ``output`` returns: the header records, then one list-directed record per
    Q row (Q, [total,] the partials (1,1), (2,1), (2,2), ...), REAL(4)."""
    from .mptxt import ld_record, _R4

    nsp = len(at_names)
    scf = np.ones(nsp) if scf is None else np.asarray(scf)
    masks = np.ones(nsp, dtype=int) if masks is None else np.asarray(masks, dtype=int)
    occ = np.asarray(occ, np.float32)
    occ = occ / occ.sum()
    rec = [
        ld_record("Substance:   ", subst_name.strip(), "  ", shell.ljust(10)[:10], "       ",
                  (mp_tool.strip() + " " + time_stamp.strip())),
        ld_record("Input files: ", file_dat_t0.strip(), " to ", file_dat.strip(), " step", int(nfile_step)),
        ld_record("OMP processes  = ", int(proc_num)),
        ld_record("Supercell size:", [int(v) for v in n_row]),
        ld_record("Unit cell parameter:", [_R4(v) for v in a_par]),
        ld_record("Atoms/unit_cell:", int(nsp)),
        ld_record("Atoms       :", *["    " + str(n).ljust(4)[:4] for n in at_names]),
        " Atom no.    :" + "".join("%8d" % (i + 1) for i in range(nsp)),
        " Occupancies: " + "".join("%8.4f" % v for v in occ),
        " Amplitudes  : " + "".join("%8.4f" % np.float32(v) for v in amplitudes),
        " Scale factors:" + "".join(" %8.4f" % np.float32(v) for v in scf),
        " Masks\t      : " + "".join("%8d" % int(v) for v in masks),
        ld_record("Smoothing FWHM [Å]", _R4(np.float32(smooth_fwhm) * np.float32(q_step))),
        ld_record(),
        ld_record("Output: ", TEXT_NAMES[mode]),
    ]
    pairs = [(a + 1, b + 1) for a in range(nsp) for b in range(a + 1)]
    title = "    Q [A-1]          " + ("" if mode == 5 else "total          ")
    # '(a,50(2i3,11x))': the 11x after the last pair writes nothing
    rec.append(title + (" " * 11).join("%3d%3d" % p for p in pairs))
    part = np.asarray(partials, np.float32)
    for i in range(len(q)):
        items = [_R4(q[i])] + ([] if mode == 5 else [_R4(total[i])]) + [_R4(v) for v in part[i]]
        rec.append(ld_record(*items))
    return "".join(r + "\n" for r in rec)
