"""S(Q) and S(Q,w) maps from device-resident trajectories.

Python face of the ``mpsq_sqw_*`` C ABI (``include/mpsq.h``).  The reference
computes, per frame and species, a type-1 NUFFT of staged positions
(code/src/mp_sqom56.f90:1101-1156), combines the species with weights
(:1338-1377), and reduces the amplitudes to the total map (:1381-1387) or to
S(Q,w) by a Hann-windowed time transform (:1437-1517) with the normalisation
of :1527 and :1572-1577.  ``SqwEngine.amplitudes`` runs the first two steps,
``SqwEngine.reduce`` the last, both on the GPU.

Device memory and streams come from torch (plumbing only); all arithmetic
runs in libmpsq's HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def make_spec(dim, n_modes, species_counts, eps=1e-6, precision="fp32", iflag=-1,
              weights=None, weight_table=None, mask=None, combine=True,
              centre=(0.0, 0.0, 0.0), axes=None, tpe=(1.0, 1.0, 1.0), q0=(0.0, 0.0, 0.0),
              window=False, vacancy_check=True, select_mode=0, n_row=(1, 1, 1),
              pmin=(0.0, 0.0, 0.0), pmax=(0.0, 0.0, 0.0), margin=0.1, at_ind=None,
              batch_frames=0, device=0, stream="torch", tile_by=0, tile_bz=0):
    """An mpsq_sqw_spec from the reference's inputs (see SqwEngine); returns
    (spec, host arrays the spec points into, derived sizes)."""
    import torch

    L = _lib.load()
    dim = int(dim)
    n_modes = tuple(int(v) for v in n_modes)
    if len(n_modes) != dim:
        raise ValueError("n_modes must have dim entries")
    NQ = int(np.prod(n_modes))
    fp64 = precision == "fp64"
    counts = np.ascontiguousarray(species_counts, np.int64)
    nsp = counts.size
    n_tot = int(counts.sum())
    s = _lib.SqwSpec()
    L.mpsq_sqw_default_spec(C.byref(s))
    s.dim = dim
    for d in range(3):
        s.n_modes[d] = n_modes[d] if d < dim else 1
    s.eps = float(eps)
    s.precision = _lib.FP64 if fp64 else _lib.FP32
    s.iflag = int(iflag)
    s.n_tot = n_tot
    s.nsp = nsp
    keep = [counts]
    s.sp_count = counts.ctypes.data
    if mask is not None:
        m = np.ascontiguousarray(mask, np.int32)
        keep.append(m)
        s.mask = m.ctypes.data
    if weights is not None:
        w = np.ascontiguousarray(weights, np.float64)
        keep.append(w)
        s.w_scalar = w.ctypes.data
    if weight_table is not None:
        wt = np.ascontiguousarray(weight_table, np.float32).reshape(nsp, NQ)
        keep.append(wt)
        s.w_table = wt.ctypes.data
    s.combine = 1 if combine else 0
    ax = np.eye(3) if axes is None else np.asarray(axes, np.float64).reshape(3, 3)
    for d in range(3):
        s.centre[d] = float(centre[d])
        s.tpe[d] = float(tpe[d])
        s.q0[d] = float(q0[d])
        s.n_row[d] = int(n_row[d])
        s.pmin[d] = float(pmin[d])
        s.pmax[d] = float(pmax[d])
    for k in range(9):
        s.axes[k] = float(ax.ravel()[k])
    s.window = int(bool(window))
    s.vacancy_check = int(bool(vacancy_check))
    s.select_mode = int(select_mode)
    s.margin = float(margin)
    if at_ind is not None:
        ai = np.ascontiguousarray(at_ind, np.int32).reshape(n_tot, 4)
        keep.append(ai)
        s.at_ind = ai.ctypes.data
    s.batch_frames = int(batch_frames)
    s.opts.device = int(device)
    if stream == "torch":
        torch.cuda.set_device(int(device))
        stream = torch.cuda.current_stream().cuda_stream
    s.opts.stream = stream
    s.opts.tile_by, s.opts.tile_bz = int(tile_by), int(tile_bz)
    info = dict(dim=dim, n_modes=n_modes, NQ=NQ, precision=precision, fp64=fp64,
                cdtype=torch.complex128 if fp64 else torch.complex64, nsp=nsp, n_tot=n_tot,
                combine=bool(combine),
                Mact=int(counts.sum() if mask is None else counts[np.asarray(mask) != 0].sum()))
    return s, keep, info


class SqwEngine:
    """Per-species NUFFT + species combine + time reduction on one GPU.

    Parameters mirror the reference's inputs: ``species_counts`` are the
    nsuper_r of each species in species-major storage order, ``weights``
    the per-species scalars (1 or 0.1*b_coh), ``weight_table`` per-Q weights
    [nsp, NQ] (X-ray form factors), ``mask`` the at_mask species set.
    Geometry: ``axes`` rows are the unit plane vectors, ``tpe`` the phase
    scale 2 pi / (e_d . (pmax - pmin)), ``centre`` the real-space centre,
    ``q0`` the Q-centre of the strengths' phase exp(-i 2 pi q0 . r).
    """

    def __init__(self, dim, n_modes, species_counts, **kw):
        self._L = _lib.load()
        s, keep, info = make_spec(dim, n_modes, species_counts, **kw)
        self._keep = keep
        self.__dict__.update(info)
        h = C.c_void_p()
        rc = self._L.mpsq_sqw_create(C.byref(s), C.byref(h))
        _lib.check("sqw_create", rc)
        self.warning = rc
        self._h = h
        self.nact = int(self._L.mpsq_sqw_nact(h))

    def amplitudes(self, pos, out=None):
        """pos: CUDA float32 tensor [T, n_tot, 4] (x, y, z, charge).
        Returns C(t,Q) [T, NQ] (combine) or F_j(t,Q) [T, nact, NQ]."""
        import torch

        if pos.dtype != torch.float32 or not pos.is_cuda:
            raise ValueError("pos must be a CUDA float32 tensor [T, n_tot, 4]")
        pos = pos.contiguous()
        T = pos.shape[0]
        if pos.numel() != T * self.n_tot * 4:
            raise ValueError("pos must have shape [T, n_tot, 4]")
        shape = (T, self.NQ) if self.combine else (T, self.nact, self.NQ)
        if out is None:
            out = torch.empty(shape, dtype=self.cdtype, device=pos.device)
        rc = self._L.mpsq_sqw_amplitudes(self._h, pos.data_ptr(), T, out.data_ptr())
        _lib.check("sqw_amplitudes", rc)
        return out

    def reduce(self, amp, n_int=None, win_kind=0, dt=1.0, nat=1.0, m_lo=0, n_f=None,
               subtract_mean=True, qfac=None, out=None, out_tot=None, kind=0, odd_n_int=False, qidx=None):
        """S(Q,w) [n_f, NQ] and S_tot(Q) [NQ] (float64 CUDA tensors).

        kind 0: S(Q,w) at frequencies m_lo .. m_lo+n_f-1; kind 1 / 2: the
        intermediate function I(Q,t) at lags m_lo .. (map form, 1/n_int;
        cut form, unscaled).  odd_n_int: the reference's window length rule
        (n_int 0 = T/2; S(Q,w) takes an odd n_int).  qidx: mode indices to
        reduce (an E(Q) cut); outputs are then [n_f, len(qidx)]."""
        import torch

        T = amp.shape[0]
        if odd_n_int:
            n_int = 0 if n_int is None else int(n_int)
            _, n_all = reference_n_int(T, n_int, sqw=(int(kind) == 0))
        else:
            n_int = T if n_int is None else int(n_int)
            n_all = n_int if int(kind) == 0 else T - n_int + 1
        n_f = n_all if n_f is None else int(n_f)
        r = _lib.ReduceSpec()
        r.n_int, r.win_kind, r.dt, r.nat = n_int, int(win_kind), float(dt), float(nat)
        r.m_lo, r.n_f, r.subtract_mean = int(m_lo), n_f, int(bool(subtract_mean))
        r.kind, r.odd_n_int = int(kind), int(bool(odd_n_int))
        nq = self.NQ
        if qfac is not None:
            qf = qfac.to(torch.float32).contiguous()
            r.d_qfac = qf.data_ptr()
        if qidx is not None:
            qi = torch.as_tensor(qidx, dtype=torch.int32, device=amp.device).contiguous()
            if qi.numel() < 1 or int(qi.min()) < 0 or int(qi.max()) >= self.NQ:
                raise ValueError(f"qidx entries must lie in [0, {self.NQ})")
            r.d_qidx, r.n_qidx = qi.data_ptr(), qi.numel()
            nq = qi.numel()
        if out is None:
            out = torch.empty((n_f, nq), dtype=torch.float64, device=amp.device)
        if out_tot is None:
            out_tot = torch.empty(nq, dtype=torch.float64, device=amp.device)
        rc = self._L.mpsq_sqw_reduce(self._h, amp.data_ptr(), T, C.byref(r), out.data_ptr(),
                                     out_tot.data_ptr())
        _lib.check("sqw_reduce", rc)
        return out, out_tot

    def debug_sorted(self, pos, frame=0):
        """Staged, bin-sorted points of one frame as the spread reads them:
        (atom index in the frame, first kernel cell [n, 3], selection flag),
        int32 numpy arrays in sorted order."""
        import numpy as np

        if not pos.is_cuda:
            raise ValueError("pos must be a CUDA float32 tensor [T, n_tot, 4]")
        pos = pos.contiguous()
        n = int(self.Mact)
        atom = np.zeros(n, np.int32)
        cell = np.zeros((n, 3), np.int32)
        keep = np.zeros(n, np.int32)
        rc = self._L.mpsq_sqw_debug_sorted(self._h, pos.data_ptr(), int(frame), atom.ctypes.data,
                                           cell.ctypes.data, keep.ctypes.data, n)
        _lib.check("sqw_debug_sorted", rc)
        return atom, cell, keep

    def info(self):
        """Fine grid, kernel width, spread tile and sort-bin shape."""
        out = (C.c_int64 * 8)()
        _lib.check("sqw_info", self._L.mpsq_sqw_info(self._h, out))
        return dict(zip(("n1", "n2", "n3", "w", "by", "bz", "BY", "BZ"), [int(v) for v in out]))

    def set_timing(self, on=True):
        """Record HIP events around every launch group (per-stage times)."""
        _lib.check("sqw_set_timing", self._L.mpsq_sqw_set_timing(self._h, int(bool(on))))

    def timing(self):
        """{stage: (summed ms, launches)} since set_timing(True)."""
        ms = (C.c_double * 6)()
        n = (C.c_int64 * 6)()
        _lib.check("sqw_get_timing", self._L.mpsq_sqw_get_timing(self._h, ms, n, 6))
        return {k: (ms[i], n[i]) for i, k in enumerate(_lib.STAGES)}

    def destroy(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.mpsq_sqw_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _reduce_spec(n_int, win_kind=0, dt=1.0, nat=1.0, m_lo=0, n_f=None, subtract_mean=True, qfac=None, kind=0,
                 odd_n_int=False):
    r = _lib.ReduceSpec()
    r.n_int, r.win_kind, r.dt, r.nat = int(n_int), int(win_kind), float(dt), float(nat)
    r.m_lo, r.n_f, r.subtract_mean = int(m_lo), int(n_int if n_f is None else n_f), int(bool(subtract_mean))
    r.kind, r.odd_n_int = int(kind), int(bool(odd_n_int))
    if qfac is not None:
        r.d_qfac = qfac.data_ptr()
    return r


class SqwBatch:
    """Many S(Q,w) jobs through one launch per phase (``mpsq_batch_*``).

    ``jobs`` is a list of dicts of SqwEngine arguments (dim, n_modes,
    species_counts, eps, weights, tpe, ...); all FP32 or all FP64, combined.
    ``run`` takes one device trajectory per job and, optionally, one
    reduction (dict of SqwEngine.reduce arguments) per job; with
    ``use_graph`` the second call with the same tensors captures the whole
    sequence in a hipGraph and later calls replay it.
    """

    def __init__(self, jobs, budget_bytes=None, stream="torch", device=0):
        self._L = _lib.load()
        specs = (_lib.SqwSpec * len(jobs))()
        self._keep, self.info = [], []
        for j, kw in enumerate(jobs):
            kw = dict(kw)
            kw.setdefault("stream", stream)
            kw.setdefault("device", device)
            s, keep, info = make_spec(kw.pop("dim"), kw.pop("n_modes"), kw.pop("species_counts"), **kw)
            specs[j] = s
            self._keep.append(keep)
            self.info.append(info)
        h = C.c_void_p()
        rc = self._L.mpsq_batch_create(len(jobs), specs, C.byref(h))
        _lib.check("batch_create", rc)
        self.warning = rc
        self._h = h
        if budget_bytes:
            _lib.check("batch_set_budget", self._L.mpsq_batch_set_budget(h, int(budget_bytes)))

    def run(self, positions, reductions=None, out=None, use_graph=False):
        """positions: list of CUDA float32 [T_j, n_tot_j, 4].  Returns the
        amplitudes [T_j, NQ_j] and, with reductions, (S [n_f, NQ], S_tot [NQ])
        per job.  Outputs are allocated on the first call and may be passed
        back through ``out`` (required for graph replay)."""
        import torch

        nj = len(self.info)
        if len(positions) != nj:
            raise ValueError("one trajectory per job")
        if out is None:
            amps = [torch.empty((p.shape[0], self.info[j]["NQ"]), dtype=self.info[j]["cdtype"],
                                device=p.device) for j, p in enumerate(positions)]
            S = St = None
            if reductions is not None:
                S, St = [], []
                for j, rd in enumerate(reductions):
                    T = positions[j].shape[0]
                    n_f = rd.get("n_f", rd.get("n_int", T))
                    S.append(torch.empty((n_f, self.info[j]["NQ"]), dtype=torch.float64, device=positions[j].device))
                    St.append(torch.empty(self.info[j]["NQ"], dtype=torch.float64, device=positions[j].device))
            out = (amps, S, St)
        amps, S, St = out
        P = (C.c_void_p * nj)(*[p.data_ptr() for p in positions])
        T = (C.c_int64 * nj)(*[p.shape[0] for p in positions])
        A = (C.c_void_p * nj)(*[a.data_ptr() for a in amps])
        R = dS = dSt = None
        if reductions is not None:
            R = (_lib.ReduceSpec * nj)()
            for j, rd in enumerate(reductions):
                rd = dict(rd)
                rd.setdefault("n_int", positions[j].shape[0])
                R[j] = _reduce_spec(**rd)
            dS = (C.c_void_p * nj)(*[x.data_ptr() for x in S])
            dSt = (C.c_void_p * nj)(*[x.data_ptr() for x in St])
        rc = self._L.mpsq_batch_run(self._h, P, T, A, R, dS, dSt, int(bool(use_graph)))
        _lib.check("batch_run", rc)
        return out

    def set_timing(self, on=True):
        _lib.check("batch_set_timing", self._L.mpsq_batch_set_timing(self._h, int(bool(on))))

    def timing(self):
        ms = (C.c_double * 6)()
        n = (C.c_int64 * 6)()
        _lib.check("batch_get_timing", self._L.mpsq_batch_get_timing(self._h, ms, n, 6))
        return {k: (ms[i], n[i]) for i, k in enumerate(_lib.STAGES)}

    def destroy(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.mpsq_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def autotune_tiles(pos, dim, n_modes, species_counts, **kw):
    """Time the spread tile shapes for this geometry on the first frames of
    ``pos`` and cache the fastest for engines created afterwards.  Returns
    (tile_by, tile_bz, ms per frame)."""
    L = _lib.load()
    s, keep, info = make_spec(dim, n_modes, species_counts, **kw)
    by, bz, ms = C.c_int(), C.c_int(), C.c_double()
    _lib.check("autotune_tiles", L.mpsq_autotune_tiles(C.byref(s), pos.data_ptr(), pos.shape[0],
                                                      C.byref(by), C.byref(bz), C.byref(ms)))
    return by.value, bz.value, ms.value


def qplane(n_row, bz_n: float, e1=(1, 0, 0), e2=(0, 1, 0), q_center=(0.0, 0.0, 0.0), a_cell_inv=None,
           a_par=None, nsuper: int = 2) -> dict:
    """The reference's Q-plane set-up (mp_sqom56.f90:800-857, 884-907), in its
    REAL(4) arithmetic: e1, e2 (normalised) and their normal ev, the NUFFT
    phase scales tpe = 2 pi / (e . (pos_max - pos_min)) with the box
    [-n_row/2, n_row/2) (lattice data) or [-a_par/2, a_par/2) (one bulk cell,
    nsuper == 1), the Brillouin-zone counts bz_nx, bz_ny (widened for
    oblique reciprocal axes), the pixel counts n_qx, n_qy (truncated, then
    made even) and the plot ranges qx/qy_min, max.  ``engine_kw`` holds the
    matching SqwEngine arguments (axes, tpe, q0) and ``n_modes``."""
    f = np.float32
    twopi = 2.0 * np.pi  # REAL(8) in the reference (mp_sqom56.f90:64); the rest is REAL(4)
    e1 = np.asarray(e1, f)
    e2 = np.asarray(e2, f)
    ev = np.array([e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]], f)
    n1, n2, nv = (f(np.sqrt(np.dot(v, v))) for v in (e1, e2, ev))
    e1, e2, ev = e1 / n1, e2 / n2, ev / nv
    if nsuper == 1:
        pmin, pmax = -np.asarray(a_par, f) / f(2), np.asarray(a_par, f) / f(2)
    else:
        pmin, pmax = -np.asarray(n_row, f) / f(2), np.asarray(n_row, f) / f(2)
    span = (pmax - pmin).astype(f)
    with np.errstate(divide="ignore"):  # a plane normal across a cubic box gives tpev = +Inf, as in Fortran
        tpe1, tpe2, tpev = (f(np.float64(twopi) / np.float64(f(np.dot(v, span)))) for v in (e1, e2, ev))
    ainv = np.eye(3, dtype=f) if a_cell_inv is None else np.asarray(a_cell_inv, f)
    bz_n = f(bz_n)
    # plane vectors in 1/A (mp_sqom56.f90:812-820): 2 pi a_cell_inv e for lattices
    if nsuper > 1:
        e1p, e2p, evp = ((twopi * (ainv @ v).astype(np.float64)).astype(f) for v in (e1, e2, ev))
    else:
        e1p, e2p, evp = ((ainv @ v).astype(f) for v in (e1, e2, ev))
    if nsuper > 1:
        e1pn, e2pn = f(np.sqrt(np.dot(e1p, e1p))), f(np.sqrt(np.dot(e2p, e2p)))
        cosa = f(np.dot(e1p, e2p) / (e1pn * e1pn))
        bz_ny = max(bz_n, f(bz_n * (e1pn / (e2pn * f(np.sqrt(f(1) - cosa * cosa))))))
        bz_nx = f(bz_n + bz_ny * abs(cosa)) if abs(cosa) > 0 else bz_n
        n_qx, n_qy = int(twopi * float(bz_nx) / float(tpe1)), int(twopi * float(bz_ny) / float(tpe2))
    else:
        bz_nx = bz_ny = bz_n
        n_qx, n_qy = int(f(bz_nx / tpe1)), int(f(bz_ny / tpe2))  # REAL(4) / REAL(4)
    if n_qx > 1 and n_qx % 2:
        n_qx += 1
    if n_qy > 1 and n_qy % 2:
        n_qy += 1
    qc = np.asarray(q_center, f)
    qx0, qy0 = f(np.dot(e1, qc) / n1), f(np.dot(e2, qc) / n2)
    qx_min = qx_max = qx0
    qy_min = qy_max = qy0
    if n_qx > 1:
        qx_min, qx_max = f(qx0 - f(0.5) * bz_nx), f(qx0 + f(0.5) * bz_nx)
    if n_qy > 1:
        qy_min, qy_max = f(qy0 - f(0.5) * bz_ny), f(qy0 + f(0.5) * bz_ny)
    # vertical components of q_center (:911, :920), in r.l.u. and in 1/A
    q_v = (f(np.dot(ev, qc) / nv) * ev).astype(f)
    evpn = f(np.sqrt(np.dot(evp, evp)))
    q_vp = (f(np.dot(evp, qc) / evpn) * evp).astype(f) if evpn > 0 else np.zeros(3, f)
    return dict(e1=e1, e2=e2, ev=ev, tpe=(tpe1, tpe2, tpev), bz_nx=bz_nx, bz_ny=bz_ny, n_qx=n_qx, n_qy=n_qy,
                qx_range=(qx_min, qx_max), qy_range=(qy_min, qy_max), n_modes=(n_qx, n_qy),
                e1p=e1p, e2p=e2p, evp=evp, q_v=q_v, q_vp=q_vp, e_norms=(n1, n2, nv),
                engine_kw=dict(axes=np.stack([e1, e2, ev]).astype(np.float64),
                               tpe=tuple(float(t) for t in (tpe1, tpe2, tpev)), q0=tuple(float(v) for v in qc)))


def qtables(plane, x_ffpar=None, stream=None, device=0):
    """The reference's per-Q tables of a Q plane from qplane(), computed on
    the device (mpsq_qtables; mp_sqom56.f90:909-936): q2_norm [n_qx * n_qy]
    (1/|Q| in r.l.u., the factor on C' of the S(Q)/Q^2 scaling) and, when
    x_ffpar ([nsp, 9]: a1, b1, a2, b2, a3, b3, a4, b4, c per species, as the
    reference reads them) is given, the X-ray weights x_ffq [nsp, n_qx *
    n_qy]; CUDA float32 tensors in the CMCL order of the plane's amplitudes."""
    import torch

    L = _lib.load()
    q = _lib.QPlane()
    q.n_qx, q.n_qy = int(plane["n_qx"]), int(plane["n_qy"])
    q.qx_min, q.qy_min = float(plane["qx_range"][0]), float(plane["qy_range"][0])
    q.bz_nx, q.bz_ny = float(plane["bz_nx"]), float(plane["bz_ny"])
    for d in range(3):
        q.e1[d], q.e2[d], q.q_v[d] = float(plane["e1"][d]), float(plane["e2"][d]), float(plane["q_v"][d])
        q.e1p[d], q.e2p[d], q.q_vp[d] = float(plane["e1p"][d]), float(plane["e2p"][d]), float(plane["q_vp"][d])
    npix = q.n_qx * q.n_qy
    dev = torch.device("cuda", int(device))
    q2 = torch.empty(npix, dtype=torch.float32, device=dev)
    ff, xq, nsp = None, None, 0
    if x_ffpar is not None:
        ff = np.ascontiguousarray(x_ffpar, np.float32).reshape(-1, 9)
        nsp = ff.shape[0]
        xq = torch.empty((nsp, npix), dtype=torch.float32, device=dev)
    if stream is None:
        stream = torch.cuda.current_stream(dev).cuda_stream
    rc = L.mpsq_qtables(C.byref(q), nsp, None if ff is None else ff.ctypes.data, q2.data_ptr(),
                        None if xq is None else xq.data_ptr(), stream)
    _lib.check("qtables", rc)
    return q2, xq


def _nint(x):
    """Fortran NINT: nearest integer, halves away from zero."""
    x = float(x)
    return int(np.floor(x + 0.5)) if x >= 0 else -int(np.floor(-x + 0.5))


def reference_n_int(nfile, n_int=0, sqw=True):
    """The reference's window length (mp_sqom56.f90:683-701): n_int = 0 means
    nfile / 2; for S(Q,w) an even n_int grows by one.  Returns (n_int,
    n_freq_max): (n_int - 1) / 2 + 1 frequencies of S(Q,w), or the
    nfile - n_int + 1 lags of I(Q,t)."""
    n_int = int(n_int) if n_int else int(nfile) // 2
    if sqw:
        if n_int % 2 == 0:
            n_int += 1
        return n_int, (n_int - 1) // 2 + 1
    return n_int, int(nfile) - n_int + 1


def econst_stack(f_map_min, f_map_step, n_plot, freq_step, iqt=False):
    """Indices of the maps of an E = const (or t = const) stack
    (mp_sqom56.f90:1266-1284, 1437-1438): plots i_plot = 1, 1 + i_step, ...
    <= n_plot with n_freq_min = nint(f_map_min / freq_step) + 1, i_step =
    nint(f_map_step / freq_step) (a smaller step is raised to freq_step, as
    the reference's dialogue does) and j_freq = n_freq_min + i_plot - 1.
    S(Q,w): frequency m = j_freq - 1 of exp(-2 pi i m t / n_int) (:1443);
    I(Q,t) (freq_step = the frame step, :701): the map correlates frames
    j_freq + 1 .. j_freq + n_int with 1 .. n_int (:1468-1469), i.e. lag
    j_freq.  Returns (indices for SqwEngine.reduce, plot labels
    f_map_min + (i_plot - 1) freq_step)."""
    f = np.float32
    freq_step = f(freq_step)
    n_freq_min = _nint(f(f_map_min) / freq_step) + 1
    i_step = _nint(f(f_map_step) / freq_step)
    if i_step < 1:
        i_step = _nint(freq_step / freq_step)
    plots = list(range(1, int(n_plot) + 1, i_step))
    j_freq = [n_freq_min + i - 1 for i in plots]
    idx = j_freq if iqt else [j - 1 for j in j_freq]
    labels = [float(f(f_map_min) + f(i - 1) * freq_step) for i in plots]
    return idx, labels


def qcut(plane, q1_3d, q2_3d, n_row, i_plot=1, n_plot=1, dq_p=(0.0, 0.0, 0.0)):
    """Pixels of an E(Q) / I(Q,t) cut (mp_sqom56.f90:1290-1301, 1389-1432,
    1489-1498) in the reference's REAL(4) arithmetic: Q1, Q2 (r.l.u.) minus
    the plane's vertical component, the in-plane test, the plot i_plot of a
    series of n_plot shifted by dq_p, the end pixels n_q1x = nint(|q1_x -
    qx_min| (e1 . n_row)) + 1 (and y), n_qdisp steps along the longer axis,
    and pixel ii = (n_q1x + nint((ii-1) q_nx), n_q1y + nint((ii-1) q_ny)) for
    ii = 1 .. n_qdisp.  Returns a dict with the CMCL mode index of every
    pixel (``qidx``, for SqwEngine.reduce) and the axis (q_min, q_max, label).
    Raises ValueError where the reference prints and asks again."""
    f = np.float32
    e1, e2, ev = (np.asarray(plane[k], f) for k in ("e1", "e2", "ev"))
    q_v = np.asarray(plane["q_v"], f)
    q1 = (np.asarray(q1_3d, f) - q_v).astype(f)
    q2 = (np.asarray(q2_3d, f) - q_v).astype(f)
    if f(np.dot(q1, ev)) != 0 or f(np.dot(q2, ev)) != 0:
        raise ValueError("Q1 or Q2 is not in the Q plane")
    e1n, e2n = f(plane["e_norms"][0]), f(plane["e_norms"][1])
    sh = (f(i_plot - int(n_plot) // 2 - 1) * np.asarray(dq_p, f)).astype(f)
    q1_x, q1_y = f(np.dot(e1, (q1 + sh).astype(f)) / e1n), f(np.dot(e2, (q1 + sh).astype(f)) / e2n)
    q2_x, q2_y = f(np.dot(e1, (q2 + sh).astype(f)) / e1n), f(np.dot(e2, (q2 + sh).astype(f)) / e2n)
    (qx_min, qx_max), (qy_min, qy_max) = plane["qx_range"], plane["qy_range"]
    for qx, qy in ((q1_x, q1_y), (q2_x, q2_y)):
        if qx < qx_min or qx > qx_max or qy < qy_min or qy > qy_max:
            raise ValueError("Q1 or Q2 is out of the map range")
    nr = np.asarray(n_row, f)
    e1r, e2r = f(np.dot(e1, nr)), f(np.dot(e2, nr))
    n_q1x, n_q2x = _nint(f(abs(f(q1_x - qx_min)) * e1r)) + 1, _nint(f(abs(f(q2_x - qx_min)) * e1r)) + 1
    n_q1y, n_q2y = _nint(f(abs(f(q1_y - qy_min)) * e2r)) + 1, _nint(f(abs(f(q2_y - qy_min)) * e2r)) + 1
    if abs(n_q2x - n_q1x) >= abs(n_q2y - n_q1y):
        n_qdisp, q_min, q_max, label = abs(n_q2x - n_q1x), q1_x, q2_x, "Q_x"
    else:
        n_qdisp, q_min, q_max, label = abs(n_q2y - n_q1y), q1_y, q2_y, "Q_y"
    if n_qdisp == 0:
        raise ValueError("Q1 and Q2 fall on one pixel")
    q_nx = f((n_q2x - n_q1x) / f(n_qdisp))
    q_ny = f((n_q2y - n_q1y) / f(n_qdisp))
    ii = np.arange(n_qdisp)
    px = np.array([n_q1x + _nint(f(f(k) * q_nx)) for k in ii])
    py = np.array([n_q1y + _nint(f(f(k) * q_ny)) for k in ii])
    n_qx = int(plane["n_qx"])
    if px.min() < 1 or py.min() < 1 or px.max() > n_qx or py.max() > int(plane["n_qy"]):
        raise ValueError("the cut leaves the map's pixels")
    return dict(qidx=(px - 1) + n_qx * (py - 1), pixels=np.stack([px, py], 1), n_qdisp=int(n_qdisp),
                q_range=(float(q_min), float(q_max)), label=label, q_nx=float(q_nx), q_ny=float(q_ny))


def cut_step(plane, q1_3d, q2_3d, q_step):
    """dq_p of an E(Q) stack (mp_sqom56.f90:1300-1305): q_step times the unit
    in-plane normal of Q2 - Q1, with the reference's sign."""
    f = np.float32
    q_v = np.asarray(plane["q_v"], f)
    dq = ((np.asarray(q2_3d, f) - q_v) - (np.asarray(q1_3d, f) - q_v)).astype(f)
    ev = np.asarray(plane["ev"], f)
    p = np.array([dq[1] * ev[2] - dq[2] * ev[1], dq[2] * ev[0] - dq[0] * ev[2], dq[0] * ev[1] - dq[1] * ev[0]], f)
    p = (f(-1.0) * p / f(np.sqrt(np.dot(p, p)))).astype(f)
    return (f(q_step) * p).astype(f)


def speckle_filter(S, s_trig, cut=False):
    """The reference's speckle filter (mp_sqom56.f90:1530-1570) on a copy of
    a host map: an E = const map [n_qy, n_qx] is filtered x fastest, then y
    fastest; an E(Q) cut [n_f, n_qdisp] (all n_freq_max frequencies, as the
    reference filters them) once, frequency fastest.  s_trig 0 is no filter."""
    a = np.array(S, np.float64, order="C", copy=True)
    if a.ndim != 2:
        raise ValueError("speckle_filter takes a 2-D map")
    L = _lib.load()
    p = a.ctypes.data_as(C.c_void_p)
    for walk in ((1,) if cut else (0, 1)):
        _lib.check("speckle_filter", L.mpsq_speckle_filter(p, a.shape[0], a.shape[1], walk, float(s_trig)))
    return a


def oneph_amplitudes(pos, at_ind, at_base, n_row, n_bz, j_plane=1, q_min=(0.0, 0.0), precision="fp32",
                     weights=None, w_table=None, stream=None):
    """One-phonon amplitudes (mp_sqom56.f90:1037-1098) of lattice (CELL)
    data on the GPU: ``pos`` float32 CUDA [T, n_tot, 4] (nsp blocks of
    n_row(1) n_row(2) n_row(3) atoms), ``at_ind`` int32 CUDA [n_tot, 4]
    (1-based cells), ``at_base`` [nsp, 3].  Returns [T, nsp, NQ] or, with
    scalar ``weights`` [nsp] or a per-Q ``w_table`` [nsp, NQ] (CUDA float32),
    the combined [T, NQ]; NQ = (n_bz n_row(1)) x (n_bz n_row(2)), x fastest."""
    import torch

    if not (pos.is_cuda and at_ind.is_cuda):
        raise ValueError("pos and at_ind must be CUDA tensors")
    base = np.ascontiguousarray(at_base, np.float32).reshape(-1, 3)
    nsp = base.shape[0]
    T = pos.shape[0]
    n_row = [int(v) for v in n_row]
    NQ = (n_bz * n_row[0]) * (n_bz * n_row[1])
    s = _lib.OnephSpec()
    s.n_row[:] = n_row
    s.n_bz, s.j_plane, s.nsp = int(n_bz), int(j_plane), nsp
    s.qx_min, s.qy_min = float(q_min[0]), float(q_min[1])
    s.at_base = base.ctypes.data
    s.precision = _lib.FP64 if precision == "fp64" else _lib.FP32
    cdt = torch.complex128 if precision == "fp64" else torch.complex64
    combine = weights is not None or w_table is not None
    out = torch.empty((T, NQ) if combine else (T, nsp, NQ), dtype=cdt, device=pos.device)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    wt = None if w_table is None else w_table.to(torch.float32).contiguous()
    pos = pos.contiguous()
    ind = at_ind.to(torch.int32).contiguous()
    rc = _lib.load().mpsq_oneph_amplitudes(C.byref(s), pos.data_ptr(), ind.data_ptr(), T,
                                           None if w is None else w.ctypes.data,
                                           None if wt is None else wt.data_ptr(), out.data_ptr(), stream)
    _lib.check("oneph_amplitudes", rc)
    return out


def dos_n_freq(nfile, t_step):
    """mp_dos's frequency count n_freq_plot (mp_dos56.f90:362, 476-477):
    int(0.5 t_tot / t_step) + 1 with t_tot = (nfile - 1) t_step, REAL(4)."""
    f = np.float32
    t_tot = f(nfile - 1) * f(t_step)
    return int(f(f(0.5) * t_tot) / f(t_step)) + 1


def dos_spectra(vel, at_ind, nsp, n_freq, precision="fp32", stream=None):
    """Per-type vibrational spectra (mp_dos56.f90:495-527) on the GPU:
    ``vel`` float32 CUDA [T, n_tot, 4] velocities, ``at_ind`` int32 CUDA
    [n_tot, 4] (type in column 3, 1-based).  Returns float64 CUDA
    [n_freq, nsp]: sum over the type's atoms and components of
    |FFT(v w)(f)|^2 / T, w the Hann window 0.5 (1 - cos(2 pi (t+1) / (T+1)))."""
    import torch

    if not (vel.is_cuda and at_ind.is_cuda):
        raise ValueError("vel and at_ind must be CUDA tensors")
    vel = vel.to(torch.float32).contiguous()
    ind = at_ind.to(torch.int32).contiguous()
    T, n_tot = vel.shape[0], vel.shape[1]
    out = torch.empty((int(n_freq), int(nsp)), dtype=torch.float64, device=vel.device)
    rc = _lib.load().mpsq_dos(vel.data_ptr(), T, n_tot, ind.data_ptr(), int(nsp), int(n_freq),
                              _lib.FP64 if precision == "fp64" else _lib.FP32, out.data_ptr(), stream)
    _lib.check("dos", rc)
    return out


def dos_output(cs_plot, weights, occup, nsuper, nfile, t_step, mask=None):
    """mp_dos's plotted spectra (mp_dos56.f90:571-589): per-type columns
    cs_plot[f, s] * w_s and their total, times nfile / at_weight_sum with
    at_weight_sum = sum(occup w nsuper) (sum of the Hann window)^2
    (t_step / 0.01).  weights: 1 (uniform) or b_coh^2 per type; mask 0/1."""
    cs = np.asarray(cs_plot, np.float64)
    nsp = cs.shape[1]
    w = np.asarray(weights, np.float64) * (np.ones(nsp) if mask is None else np.asarray(mask, np.float64))
    i = np.arange(1, int(nfile) + 1, dtype=np.float64)
    wind_sum = np.sum(0.5 * (1.0 - np.cos(2.0 * np.pi * i / (nfile + 1))))
    aws = np.sum(np.asarray(occup, np.float64) * w * np.asarray(nsuper, np.float64)) * wind_sum ** 2 * (t_step / 0.01)
    out = np.empty((cs.shape[0], nsp + 1))
    out[:, :nsp] = cs * w[None, :]
    out[:, nsp] = out[:, :nsp].sum(axis=1)
    return out * nfile / aws
