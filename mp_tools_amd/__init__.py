"""mp_tools_amd -- MI355X-native S(Q,w) hot path of jkulda/MP_tools.

    libmpsq.so (HIP, gfx950)   type-1 NUFFT (fold, bin-sort, spread + FFT,
                               deconvolution) and the S(Q,w) reduction,
                               behind a FINUFFT-shaped C ABI (include/mpsq.h)
    nufft                      Python Plan / nufft{1,2,3}d1 over that ABI
    sqw                        SqwEngine: trajectory -> C(t,Q) -> S(Q), S(Q,w)
    synth                      synthetic perovskite trajectories
"""
from . import _lib  # noqa: F401

__all__ = ["nufft", "sqw", "synth"]
