"""Type-1 NUFFT on the GPU behind FINUFFT's Python interface shape.

The reference reaches FINUFFT through ``call finufft2d1(...)``
(code/src/mp_sqom56.f90:1148, code/src/mp_sql56.f90:1129).  This module is
the Python face of the same C ABI (``include/mpsq.h``): a ``Plan`` with
``setpts`` / ``execute`` (FINUFFT's guru interface) and the simple
``nufft1d1`` / ``nufft2d1`` / ``nufft3d1`` calls.

Conventions (those of FINUFFT's type 1):
    f[k] = sum_j c_j exp(isign * i * (k1 x_j + k2 y_j + k3 z_j)),
    k_d = -N_d/2 .. (N_d-1)/2.
Outputs are returned in the C ABI's memory order, k1 fastest, i.e. as arrays
of shape ``reversed(n_modes)`` (``(N3, N2, N1)`` in 3D), with a leading
``n_trans`` axis when ``n_trans > 1``.

Inputs may be numpy arrays (host path: copied to the device, result copied
back) or CUDA torch tensors (device path: no host copies, work queued on the
torch stream that was current when the plan was made).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


class Plan:
    """GPU type-1 NUFFT plan (FP64 for complex128, FP32 for complex64)."""

    def __init__(self, nufft_type, n_modes, n_trans=1, eps=1e-6, isign=-1,
                 dtype="complex128", device=0, stream=None, tile_by=0, tile_bz=0,
                 sort_chunk=0):
        if nufft_type != 1:
            raise ValueError("only type-1 transforms are on the S(Q,w) path")
        self._L = _lib.load()
        self.n_modes = tuple(int(v) for v in np.atleast_1d(n_modes))
        self.dim = len(self.n_modes)
        if not 1 <= self.dim <= 3:
            raise ValueError("n_modes must have 1..3 entries")
        self.n_trans = int(n_trans)
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.complex64, np.complex128):
            raise ValueError("dtype must be complex64 or complex128")
        self.single = self.dtype == np.complex64
        self.rdtype = np.float32 if self.single else np.float64
        self._pre = "mpsqf" if self.single else "mpsq"
        self.isign = int(isign)
        opts = _lib.Opts()
        self._L.mpsq_default_opts(C.byref(opts))
        opts.device = int(device)
        if stream == "torch":
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        opts.stream = stream
        opts.tile_by, opts.tile_bz, opts.sort_chunk = int(tile_by), int(tile_bz), int(sort_chunk)
        nm = (C.c_int64 * 3)(*(list(self.n_modes) + [1] * (3 - self.dim)))
        h = C.c_void_p()
        rc = getattr(self._L, f"{self._pre}_makeplan")(1, self.dim, nm, self.isign,
                                                      self.n_trans, float(eps), C.byref(h), C.byref(opts))
        _lib.check("makeplan", rc)
        self.warning = rc
        self._h = h
        self.M = None
        self._keep = None

    # -- points -------------------------------------------------------------
    def setpts(self, x, y=None, z=None):
        co = [x, y, z][: self.dim]
        if any(v is None for v in co):
            raise ValueError(f"{self.dim}D plan needs {self.dim} coordinate arrays")
        if _is_torch(x):
            import torch
            tdt = torch.float32 if self.single else torch.float64
            co = [v.to(tdt).contiguous() for v in co]
            M = co[0].numel()
            ptrs = [v.data_ptr() for v in co] + [None] * (3 - self.dim)
            rc = getattr(self._L, f"{self._pre}_setpts_dev")(self._h, M, *ptrs)
        else:
            co = [np.ascontiguousarray(v, dtype=self.rdtype).ravel() for v in co]
            M = co[0].size
            ptrs = [v.ctypes.data for v in co] + [None] * (3 - self.dim)
            rc = getattr(self._L, f"{self._pre}_setpts")(self._h, M, *ptrs)
        _lib.check("setpts", rc)
        self._keep = co
        self.M = M
        return rc

    # -- transform ----------------------------------------------------------
    def _out_shape(self):
        sh = tuple(reversed(self.n_modes))
        return ((self.n_trans,) + sh) if self.n_trans > 1 else sh

    def execute(self, data, out=None):
        if self.M is None:
            raise RuntimeError("execute before setpts")
        if _is_torch(data):
            import torch
            tdt = torch.complex64 if self.single else torch.complex128
            c = data.to(tdt).contiguous()
            if c.numel() != self.M * self.n_trans:
                raise ValueError("strength count does not match M * n_trans")
            if out is None:
                out = torch.empty(self._out_shape(), dtype=tdt, device=c.device)
            rc = getattr(self._L, f"{self._pre}_execute_dev")(self._h, c.data_ptr(), out.data_ptr())
        else:
            c = np.ascontiguousarray(data, dtype=self.dtype).ravel()
            if c.size != self.M * self.n_trans:
                raise ValueError("strength count does not match M * n_trans")
            if out is None:
                out = np.empty(self._out_shape(), dtype=self.dtype)
            rc = getattr(self._L, f"{self._pre}_execute")(self._h, c.ctypes.data, out.ctypes.data)
        _lib.check("execute", rc)
        return out

    def sort_perm(self) -> np.ndarray:
        """Bin-sort permutation of the last setpts (perm[k] = input index)."""
        p = np.empty(self.M, np.int32)
        _lib.check("get_sort_perm",
                   getattr(self._L, f"{self._pre}_get_sort_perm")(self._h, p.ctypes.data))
        return p

    def info(self) -> dict:
        o = (C.c_int64 * 8)()
        _lib.check("plan_info", getattr(self._L, f"{self._pre}_plan_info")(self._h, o))
        return dict(n_fine=(o[0], o[1], o[2]), w=o[3], tile=(o[4], o[5]), sort_bin=(o[6], o[7]))

    def destroy(self):
        if getattr(self, "_h", None) and self._h.value:
            getattr(self._L, f"{self._pre}_destroy")(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _simple(dim, coords, c, n_modes, eps, isign):
    L = _lib.load()
    co = [np.ascontiguousarray(v, np.float64).ravel() for v in coords]
    c = np.ascontiguousarray(c, np.complex128).ravel()
    n_modes = tuple(int(v) for v in np.atleast_1d(n_modes))
    if len(n_modes) != dim:
        raise ValueError(f"n_modes must have {dim} entries")
    f = np.empty(tuple(reversed(n_modes)), np.complex128)
    M = co[0].size
    if dim == 1:
        rc = L.mpsq_nufft1d1(M, co[0].ctypes.data, c.ctypes.data, int(isign), float(eps),
                             n_modes[0], f.ctypes.data, None)
    elif dim == 2:
        rc = L.mpsq_nufft2d1(M, co[0].ctypes.data, co[1].ctypes.data, c.ctypes.data, int(isign),
                             float(eps), n_modes[0], n_modes[1], f.ctypes.data, None)
    else:
        rc = L.mpsq_nufft3d1(M, co[0].ctypes.data, co[1].ctypes.data, co[2].ctypes.data,
                             c.ctypes.data, int(isign), float(eps), n_modes[0], n_modes[1],
                             n_modes[2], f.ctypes.data, None)
    _lib.check(f"nufft{dim}d1", rc)
    return f


def nufft1d1(x, c, n_modes, eps=1e-6, isign=-1):
    return _simple(1, [x], c, n_modes, eps, isign)


def nufft2d1(x, y, c, n_modes, eps=1e-6, isign=-1):
    return _simple(2, [x, y], c, n_modes, eps, isign)


def nufft3d1(x, y, z, c, n_modes, eps=1e-6, isign=-1):
    return _simple(3, [x, y, z], c, n_modes, eps, isign)
