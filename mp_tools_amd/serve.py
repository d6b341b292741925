"""Map server: device-resident trajectories, HBM-budget admission and
continuous batching (``mpsq_server_*``), plus the Poisson-arrival load
generator of the serving benchmark.

The server is native (a worker thread in libmpsq): requests are admitted in
arrival order while their device data fit the HBM budget, every step runs
the next frames of every admitted request through one launch per phase, and
a request leaves once its S(Q,w) reduction has run.  This module only
submits, polls and collects.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib

QUEUED, ADMITTED, DONE = 0, 1, 2


class MapServer:
    def __init__(self, hbm_budget=64 << 30, step_budget=8 << 30, max_active=64, device=0):
        self._L = _lib.load()
        o = _lib.ServerOpts(int(device), int(hbm_budget), int(step_budget), int(max_active))
        h = C.c_void_p()
        _lib.check("server_create", self._L.mpsq_server_create(C.byref(o), C.byref(h)))
        self._h = h
        self._trajs = []     # keeps the tensors alive
        self._req_keep = {}
        self.device = device

    def add_trajectory(self, pos, species_counts, tpe, qscale=None):
        """pos: CUDA float32 [T, n_tot, 4], species-major; tpe: 2 pi / box
        length (position units) per axis; qscale: 2 pi / box length in
        Angstrom per axis (X-ray form factors)."""
        counts = np.ascontiguousarray(species_counts, np.int64)
        tp = np.ascontiguousarray(tpe, np.float64)
        qs = np.ascontiguousarray(qscale if qscale is not None else tpe, np.float64)
        tid = C.c_int()
        _lib.check("server_add_trajectory", self._L.mpsq_server_add_trajectory(
            self._h, pos.data_ptr(), pos.shape[0], counts.size, counts.ctypes.data, tp.ctypes.data,
            qs.ctypes.data, C.byref(tid)))
        self._trajs.append((pos, counts))
        return tid.value

    def submit(self, traj, n_modes, f0, T, eps=1e-6, weights=None, xray=None, mask=None, dt=0.01, n_f=None):
        """weights: per-species scalars (neutron 0.1 b); xray: [nsp, 9] form
        factor coefficients (a1..a4, b1..b4, c).  Returns a ticket."""
        nsp = self._trajs[traj][1].size
        r = _lib.Request()
        r.traj = int(traj)
        for d in range(3):
            r.n_modes[d] = int(n_modes[d])
        r.f0, r.T, r.eps, r.dt = int(f0), int(T), float(eps), float(dt)
        r.n_f = int(max(1, T // 2) if n_f is None else n_f)
        keep = []
        if mask is not None:
            m = np.ascontiguousarray(mask, np.int32)
            keep.append(m)
            r.mask = m.ctypes.data
        if xray is not None:
            x = np.ascontiguousarray(xray, np.float64).reshape(nsp, 9)
            keep.append(x)
            r.weight_kind, r.xff = 1, x.ctypes.data
        else:
            w = np.ascontiguousarray(weights if weights is not None else np.ones(nsp), np.float64)
            keep.append(w)
            r.weight_kind, r.w_scalar = 0, w.ctypes.data
        t = C.c_int64()
        _lib.check("server_submit", self._L.mpsq_server_submit(self._h, C.byref(r), C.byref(t)))
        self._req_keep[t.value] = (r.n_f, int(np.prod(n_modes)))
        return t.value

    def poll(self, ticket):
        st, q, l = C.c_int(), C.c_double(), C.c_double()
        _lib.check("server_poll", self._L.mpsq_server_poll(self._h, int(ticket), C.byref(st), C.byref(q),
                                                          C.byref(l)))
        if st.value < 0:
            raise _lib.MpsqError("server request", -st.value)
        return st.value, q.value, l.value

    def result(self, ticket):
        """(S [n_f, NQ], S_tot [NQ]) of a done request, copied to torch tensors."""
        import torch
        n_f, nq = self._req_keep[ticket]
        S = torch.empty((n_f, nq), dtype=torch.float64, device=f"cuda:{self.device}")
        St = torch.empty(nq, dtype=torch.float64, device=f"cuda:{self.device}")
        _lib.check("server_copy_result", self._L.mpsq_server_copy_result(self._h, int(ticket), S.data_ptr(),
                                                                        St.data_ptr()))
        return S, St

    def release(self, ticket):
        _lib.check("server_release", self._L.mpsq_server_release(self._h, int(ticket)))
        self._req_keep.pop(ticket, None)

    def wait_idle(self):
        _lib.check("server_wait_idle", self._L.mpsq_server_wait_idle(self._h))

    def stats(self):
        s, d, b = C.c_int64(), C.c_int64(), C.c_double()
        _lib.check("server_stats", self._L.mpsq_server_stats(self._h, C.byref(s), C.byref(d), C.byref(b)))
        return {"steps": s.value, "done": d.value, "host_step_ms": b.value}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.mpsq_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def random_request(rng, trajs, box_modes=((64, 64, 8), (128, 128, 16), (256, 256, 32), (64, 128, 16),
                                         (128, 256, 32)), max_frames=256):
    """One map request of the serving workload: a trajectory, a species set,
    a Q box, 1..256 frames, and neutron or X-ray weights (X-ray coefficients
    a_k in [0.5, 10], b_k in [0.5, 60], c in [0, 1], seeded)."""
    k = int(rng.integers(0, len(trajs)))
    tr = trajs[k]
    nsp, T_avail = tr["nsp"], tr["T"]
    T = int(rng.integers(1, min(max_frames, T_avail) + 1))
    f0 = int(rng.integers(0, T_avail - T + 1))
    mask = rng.integers(0, 2, nsp)
    if mask.sum() == 0:
        mask[int(rng.integers(0, nsp))] = 1
    N = box_modes[int(rng.integers(0, len(box_modes)))]
    req = dict(traj=k, n_modes=N, f0=f0, T=T, mask=mask, dt=tr.get("dt", 0.01), n_f=max(1, T // 2))
    if rng.random() < 0.5:
        req["weights"] = tr["b"]
    else:
        a = rng.uniform(0.5, 10.0, (nsp, 4))
        b = rng.uniform(0.5, 60.0, (nsp, 4))
        c = rng.uniform(0.0, 1.0, (nsp, 1))
        req["xray"] = np.concatenate([a, b, c], axis=1)
    return req


def poisson_load(server, trajs, rate, duration_s, seed=0, eps=1e-6, max_inflight=None):
    """Submit requests with exponential inter-arrival times (rate per second)
    for duration_s, collect every completion, and return the per-request
    latencies (ms), queue waits (ms), and the wall time from the first
    submission to the last completion."""
    rng = np.random.default_rng(seed)
    t_start = time.perf_counter()
    next_t = t_start
    pending, lat, qwait, nreq = {}, [], [], 0
    while True:
        now = time.perf_counter()
        if now >= next_t and now - t_start < duration_s:
            req = random_request(rng, trajs)
            pending[server.submit(eps=eps, **req)] = now
            nreq += 1
            next_t += rng.exponential(1.0 / rate)
            continue
        for tk in list(pending):
            st, qms, lms = server.poll(tk)
            if st == DONE:
                lat.append(lms)
                qwait.append(qms)
                server.release(tk)
                del pending[tk]
        if now - t_start >= duration_s and not pending:
            break
        time.sleep(0.0005)
    wall = time.perf_counter() - t_start
    return np.array(lat), np.array(qwait), wall, nreq
