"""Loader for libmpsq.so (the HIP/gfx950 S(Q,w) engine).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
mp_tools_amd/csrc``).  There is no CPU fallback: importing a compute entry
point without the library raises, and every compute call on a machine
without a HIP device returns an error code.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpsq.so")

# status codes (include/mpsq.h)
OK = 0
WARN_EPS_TOO_SMALL = 1
ERR_BAD_ARG = 2
ERR_NONFINITE_POINT = 3
ERR_SIZE = 4
ERR_DEVICE = 5
ERR_NOT_SET = 6
FP32 = 0
FP64 = 1

# every symbol include/mpsq.h declares
EXPORTS = [
    "mpsq_default_opts", "mpsq_makeplan", "mpsq_setpts", "mpsq_execute", "mpsq_setpts_dev",
    "mpsq_execute_dev", "mpsq_destroy", "mpsq_get_sort_perm", "mpsq_plan_info",
    "mpsqf_makeplan", "mpsqf_setpts", "mpsqf_execute", "mpsqf_setpts_dev", "mpsqf_execute_dev",
    "mpsqf_destroy", "mpsqf_get_sort_perm", "mpsqf_plan_info",
    "mpsq_nufft1d1", "mpsq_nufft2d1", "mpsq_nufft3d1",
    "finufft1d1_", "finufft2d1_", "finufft3d1_",
    "mpsq_makeplan_", "mpsq_setpts_", "mpsq_execute_", "mpsq_destroy_",
    "mpsq_sqw_default_spec", "mpsq_sqw_create", "mpsq_sqw_amplitudes", "mpsq_sqw_reduce",
    "mpsq_sqw_destroy", "mpsq_sqw_sync", "mpsq_sqw_nq", "mpsq_sqw_nact", "mpsq_version",
    "mpsq_sqw_set_timing", "mpsq_sqw_get_timing", "mpsq_sqw_debug_sorted", "mpsq_sqw_info", "mpsq_qtables",
    "mpsq_speckle_filter", "mpsq_oneph_amplitudes", "mpsq_dos",
    "mpsq_batch_create", "mpsq_batch_run", "mpsq_batch_set_budget", "mpsq_batch_set_timing",
    "mpsq_batch_get_timing", "mpsq_batch_destroy", "mpsq_autotune_tiles",
    "mpsq_server_create", "mpsq_server_add_trajectory", "mpsq_server_submit", "mpsq_server_poll",
    "mpsq_server_result", "mpsq_server_copy_result", "mpsq_server_release", "mpsq_server_wait_idle",
    "mpsq_server_stats", "mpsq_server_destroy", "mpsq_sql_hist", "mpsq_sql_hist_noalloc",
    "mpsq_sql_hist_mc",
]
STAGES = ("stage", "sort", "spread_xfft", "colfft", "final_colfft", "reduce")


class MpsqError(RuntimeError):
    def __init__(self, where: str, code: int):
        names = {2: "bad argument", 3: "non-finite point", 4: "size beyond LDS tile",
                 5: "HIP device/runtime failure", 6: "execute before setpts"}
        super().__init__(f"{where}: mpsq error {code} ({names.get(code, 'unknown')})")
        self.code = code


class QPlane(C.Structure):
    _fields_ = [("n_qx", C.c_int32), ("n_qy", C.c_int32), ("qx_min", C.c_float), ("qy_min", C.c_float),
                ("bz_nx", C.c_float), ("bz_ny", C.c_float), ("e1", C.c_float * 3), ("e2", C.c_float * 3),
                ("q_v", C.c_float * 3), ("e1p", C.c_float * 3), ("e2p", C.c_float * 3), ("q_vp", C.c_float * 3)]


class Opts(C.Structure):
    _fields_ = [("device", C.c_int), ("stream", C.c_void_p), ("tile_by", C.c_int),
                ("tile_bz", C.c_int), ("sort_chunk", C.c_int), ("debug", C.c_int)]


class SqwSpec(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("n_modes", C.c_int64 * 3), ("eps", C.c_double),
        ("precision", C.c_int), ("iflag", C.c_int), ("n_tot", C.c_int64), ("nsp", C.c_int),
        ("sp_count", C.c_void_p), ("mask", C.c_void_p), ("w_scalar", C.c_void_p),
        ("w_table", C.c_void_p), ("combine", C.c_int), ("centre", C.c_double * 3),
        ("axes", C.c_double * 9), ("tpe", C.c_double * 3), ("q0", C.c_double * 3),
        ("window", C.c_int), ("vacancy_check", C.c_int), ("select_mode", C.c_int),
        ("n_row", C.c_int32 * 3), ("pmin", C.c_float * 3), ("pmax", C.c_float * 3),
        ("margin", C.c_float), ("at_ind", C.c_void_p), ("batch_frames", C.c_int),
        ("opts", Opts),
    ]


class ServerOpts(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget", C.c_int64), ("step_budget", C.c_int64),
                ("max_active", C.c_int)]


class Request(C.Structure):
    _fields_ = [("traj", C.c_int), ("mask", C.c_void_p), ("n_modes", C.c_int64 * 3), ("f0", C.c_int64),
                ("T", C.c_int64), ("eps", C.c_double), ("weight_kind", C.c_int), ("w_scalar", C.c_void_p),
                ("xff", C.c_void_p), ("dt", C.c_double), ("n_f", C.c_int64)]


class OnephSpec(C.Structure):
    _fields_ = [("n_row", C.c_int * 3), ("n_bz", C.c_int), ("j_plane", C.c_int), ("nsp", C.c_int),
                ("qx_min", C.c_float), ("qy_min", C.c_float), ("at_base", C.c_void_p), ("precision", C.c_int)]


class ReduceSpec(C.Structure):
    _fields_ = [("n_int", C.c_int64), ("win_kind", C.c_int), ("dt", C.c_double),
                ("nat", C.c_double), ("m_lo", C.c_int64), ("n_f", C.c_int64),
                ("subtract_mean", C.c_int), ("d_qfac", C.c_void_p), ("kind", C.c_int), ("odd_n_int", C.c_int),
                ("d_qidx", C.c_void_p), ("n_qidx", C.c_int64)]


_lib = None


def load():
    """Load libmpsq.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    # torch ships its own libamdhip64.so.7; two HIP runtimes in one process
    # do not share the device.  Loading torch first makes the dynamic linker
    # resolve libmpsq's libamdhip64.so.7 to torch's copy (same SONAME), so
    # torch tensors and libmpsq kernels live in one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    L.mpsq_default_opts.argtypes = [vp]
    L.mpsq_default_opts.restype = None
    for pre in ("mpsq", "mpsqf"):
        getattr(L, f"{pre}_makeplan").argtypes = [i32, i32, vp, i32, i32, dbl, vp, vp]
        getattr(L, f"{pre}_setpts").argtypes = [vp, i64, vp, vp, vp]
        getattr(L, f"{pre}_setpts_dev").argtypes = [vp, i64, vp, vp, vp]
        getattr(L, f"{pre}_execute").argtypes = [vp, vp, vp]
        getattr(L, f"{pre}_execute_dev").argtypes = [vp, vp, vp]
        getattr(L, f"{pre}_destroy").argtypes = [vp]
        getattr(L, f"{pre}_get_sort_perm").argtypes = [vp, vp]
        getattr(L, f"{pre}_plan_info").argtypes = [vp, vp]
    L.mpsq_nufft1d1.argtypes = [i64, vp, vp, i32, dbl, i64, vp, vp]
    L.mpsq_nufft2d1.argtypes = [i64, vp, vp, vp, i32, dbl, i64, i64, vp, vp]
    L.mpsq_nufft3d1.argtypes = [i64, vp, vp, vp, vp, i32, dbl, i64, i64, i64, vp, vp]
    for f in ("finufft1d1_", "finufft2d1_", "finufft3d1_", "mpsq_makeplan_", "mpsq_setpts_",
              "mpsq_execute_", "mpsq_destroy_"):
        getattr(L, f).restype = None
    L.finufft2d1_.argtypes = [vp] * 11
    L.finufft3d1_.argtypes = [vp] * 13
    L.finufft1d1_.argtypes = [vp] * 9
    L.mpsq_sqw_default_spec.argtypes = [vp]
    L.mpsq_sqw_default_spec.restype = None
    L.mpsq_sqw_create.argtypes = [vp, vp]
    L.mpsq_sqw_amplitudes.argtypes = [vp, vp, i64, vp]
    L.mpsq_sqw_reduce.argtypes = [vp, vp, i64, vp, vp, vp]
    L.mpsq_sqw_destroy.argtypes = [vp]
    L.mpsq_sqw_sync.argtypes = [vp]
    L.mpsq_sqw_nq.argtypes = [vp]
    L.mpsq_sqw_nq.restype = i64
    L.mpsq_sqw_nact.argtypes = [vp]
    L.mpsq_version.restype = C.c_char_p
    L.mpsq_sqw_set_timing.argtypes = [vp, i32]
    L.mpsq_sqw_debug_sorted.argtypes = [vp, vp, i64, vp, vp, vp, i64]
    L.mpsq_qtables.argtypes = [vp, i32, vp, vp, vp, vp]
    L.mpsq_speckle_filter.argtypes = [vp, i64, i64, i32, C.c_double]
    L.mpsq_oneph_amplitudes.argtypes = [vp, vp, vp, i64, vp, vp, vp, vp]
    L.mpsq_dos.argtypes = [vp, i64, i64, vp, i32, i64, i32, vp, vp]
    L.mpsq_sqw_info.argtypes = [vp, vp]
    L.mpsq_sqw_get_timing.argtypes = [vp, vp, vp, i32]
    L.mpsq_batch_create.argtypes = [i32, vp, vp]
    L.mpsq_batch_run.argtypes = [vp, vp, vp, vp, vp, vp, vp, i32]
    L.mpsq_batch_set_budget.argtypes = [vp, i64]
    L.mpsq_batch_set_timing.argtypes = [vp, i32]
    L.mpsq_batch_get_timing.argtypes = [vp, vp, vp, i32]
    L.mpsq_batch_destroy.argtypes = [vp]
    L.mpsq_autotune_tiles.argtypes = [vp, vp, i64, vp, vp, vp]
    L.mpsq_server_create.argtypes = [vp, vp]
    L.mpsq_server_add_trajectory.argtypes = [vp, vp, i64, i32, vp, vp, vp, vp]
    L.mpsq_server_submit.argtypes = [vp, vp, vp]
    L.mpsq_server_poll.argtypes = [vp, i64, vp, vp, vp]
    L.mpsq_server_result.argtypes = [vp, i64, vp, vp, vp]
    L.mpsq_server_copy_result.argtypes = [vp, i64, vp, vp]
    L.mpsq_server_release.argtypes = [vp, i64]
    L.mpsq_server_wait_idle.argtypes = [vp]
    L.mpsq_server_stats.argtypes = [vp, vp, vp, vp]
    L.mpsq_server_destroy.argtypes = [vp]
    flt = C.c_float
    L.mpsq_sql_hist.argtypes = [vp, i32, i32, vp, vp, vp, flt, flt, i32, vp, vp]
    L.mpsq_sql_hist_noalloc.argtypes = [vp, i32, i32, vp, vp, vp, flt, flt, i32, vp, vp]
    L.mpsq_sql_hist_mc.argtypes = [vp, i32, i32, vp, vp, vp, flt, flt, i32, i64, i64, C.c_uint64, C.c_uint32,
                                   vp, vp, vp]
    _lib = L
    return L


def check(where: str, rc: int) -> int:
    if rc >= 2:
        raise MpsqError(where, rc)
    return rc
