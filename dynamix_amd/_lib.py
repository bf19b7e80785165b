"""ctypes binding of libdxg2.so (the C ABI declared in include/dxg2.h).

The library is built in-tree by dynamix_amd.build; loading fails loudly if it is
missing -- there is no CPU fallback for any correlator.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdxg2.so")
HEADER = os.path.join(HERE, "..", "include", "dxg2.h")

DXG2_OK = 0
DXG2_EINVAL = 1
DXG2_ENOMEM = 2
DXG2_EHIP = 3
DXG2_EUNSUPPORTED = 4
DXG2_ECAPACITY = 5

DXG2_FLAG_TILE128 = 1
DXG2_FLAG_NO_QUEUE = 2

DXG2_STATIC_SCRATCH_WORDS = 65536   # static workspace of dxg2_events_window_noalloc

DXG2_U8 = 1
DXG2_I8 = 2
DXG2_U16 = 3

DTYPE_CODES = {np.dtype(np.uint8): DXG2_U8, np.dtype(np.int8): DXG2_I8,
               np.dtype(np.uint16): DXG2_U16}

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_double = ctypes.c_double
c_size_t = ctypes.c_size_t


class PlanDesc(ctypes.Structure):
    _fields_ = [("height", c_int), ("width", c_int), ("nframes", c_int), ("dtype", c_int),
                ("qmask", c_void_p), ("scale_factors", c_void_p),
                ("n_scale_factors", c_int), ("flags", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    _fields_ = [("tile", c_int), ("band", c_int), ("group", c_int), ("queue", c_int),
                ("kchunk", c_int)]


ALLOC_FN = ctypes.CFUNCTYPE(c_void_p, c_size_t, c_void_p, c_void_p)
FREE_FN = ctypes.CFUNCTYPE(None, c_void_p, c_void_p, c_void_p)

_SIGNATURES = {
    "dxg2_set_allocator": (c_int, [ALLOC_FN, FREE_FN, c_void_p]),
    "dxg2_plan_create": (c_int, [ctypes.POINTER(PlanDesc), ctypes.POINTER(c_void_p)]),
    "dxg2_measure_lds_atomics": (c_int, [ctypes.c_int32, ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_double), c_void_p]),
    "dxg2_plan_dry_run": (c_int, [ctypes.POINTER(PlanDesc), ctypes.POINTER(ctypes.c_int64),
                                  ctypes.c_int32]),
    "dxg2_plan_destroy": (c_int, [c_void_p]),
    "dxg2_plan_destroy_async": (c_int, [c_void_p, c_void_p]),
    "dxg2_plan_query": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "dxg2_plan_footprint": (c_int, [c_void_p, c_int, ctypes.POINTER(c_size_t)]),
    "dxg2_correlate_dense": (c_int, [c_void_p] * 7),
    "dxg2_correlate_dense_std": (c_int, [c_void_p] * 8),
    "dxg2_correlate_dense_graph": (c_int, [c_void_p] * 7),
    "dxg2_correlate_dense_ttcf": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "dxg2_plan_get_tuning": (c_int, [c_void_p, ctypes.POINTER(Tuning)]),
    "dxg2_plan_set_tuning": (c_int, [c_void_p, ctypes.POINTER(Tuning)]),
    "dxg2_correlate_dense_plain": (c_int, [c_void_p] * 4),
    "dxg2_batch_create": (c_int, [ctypes.POINTER(c_void_p)]),
    "dxg2_batch_destroy": (c_int, [c_void_p]),
    "dxg2_batch_correlate_dense": (c_int, [c_void_p, c_int] + [c_void_p] * 7),
    "dxg2_batch_correlate_events": (c_int, [c_void_p, c_int] + [c_void_p] * 9),
    "dxg2_correlate_fft": (c_int, [c_void_p] * 7),
    "dxg2_correlate_events_csr": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_i64,
                                          c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dxg2_correlate_events_coo": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_i64,
                                          c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dxg2_plan_event_status": (c_int, [c_void_p, ctypes.POINTER(c_int), c_void_p]),
    "dxg2_plan_set_profiling": (c_int, [c_void_p, c_int]),
    "dxg2_plan_stage_times": (c_int, [c_void_p, c_void_p, c_int]),
    "dxg2_synth_speckle": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_double,
                                   c_u64, c_void_p]),
    "dxg2_synth_events": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_u64, c_int,
                                  c_void_p, c_void_p, c_void_p, c_i64, ctypes.POINTER(c_i64),
                                  c_void_p]),
    "dxg2_events_window": (c_int, [c_void_p, c_void_p, c_void_p, c_i64, c_i64, c_int, c_int,
                                   c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_i64,
                                   c_void_p]),
    "dxg2_events_window_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_i64, c_i64, c_int,
                                           c_int, c_int, c_void_p, c_void_p, c_void_p, c_i64,
                                           c_void_p]),
    "dxg2_events_to_columns": (c_int, [c_void_p, c_void_p, c_void_p, c_i64, c_int, c_int,
                                       c_void_p, c_i64, c_void_p, c_void_p, c_void_p]),
    "dxg2_events_to_coo": (c_int, [c_void_p, c_void_p, c_void_p, c_i64, c_int, c_int, c_int,
                                   c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dxg2_compress_scratch_bytes": (c_int, [c_int, c_int, c_int, ctypes.POINTER(c_size_t)]),
    "dxg2_compress_count": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                    c_void_p]),
    "dxg2_compress_fill": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p]),
    "dxg2_compress_push": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_void_p]),
    "dxg2_compress_compact": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int,
                                      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dxg2_last_error": (ctypes.c_char_p, []),
    "dxg2_version": (ctypes.c_char_p, []),
}

_lib = None
_lock = threading.Lock()


def load():
    """Load libdxg2.so (raises RuntimeError if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libdxg2.so not built: run `python -m dynamix_amd.build` "
                               "(the HIP library is required; there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc, what=""):
    """Turn a DXG2 status into the exception class the reference raises for it."""
    if rc == DXG2_OK:
        return
    msg = load().dxg2_last_error().decode(errors="replace")
    if what:
        msg = "%s: %s" % (what, msg)
    if rc == DXG2_EINVAL:
        raise ValueError(msg)
    if rc == DXG2_ENOMEM:
        raise MemoryError(msg)
    if rc == DXG2_EUNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == DXG2_ECAPACITY:
        raise ValueError(msg)
    raise RuntimeError(msg)


def header_symbols():
    """Function names declared in include/dxg2.h."""
    import re
    with open(HEADER) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(dxg2_[a-z0-9_]+)\s*\(", txt)))


def ptr(t):
    """Device address of a torch tensor (or None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())
