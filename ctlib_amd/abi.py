"""ctypes binding of the C ABI in libctlib_amd.so (include/ctlib_abi.h).

This is the binding a non-torch caller would use (a ctypes stub over the C ABI); the
package's torch ops go through the C++ extension `ctlib_amd._C` instead.  It
is also what the host-side tests use to check that the library loads, exports
every declared symbol, and builds the same geometry plan on any machine.
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libctlib_amd.so")

STATUS = {0: "OK", 1: "INVALID_ARG", 2: "BAD_SCAN_TYPE", 3: "SHAPE", 4: "HIP", 5: "UNSUPPORTED_DTYPE",
          6: "WORKSPACE", 7: "GEOMETRY"}

OPS = {"projection": 0, "projection_t": 1, "backprojection": 2, "backprojection_t": 3,
       "backprojection_t_true": 4, "fbp": 5, "fbp_t": 6}

# Every function declared in include/ctlib_abi.h.
EXPORTS = ("ctlib_parse_options", "ctlib_plan_create", "ctlib_plan_destroy", "ctlib_workspace_bytes", "ctlib_run",
           "ctlib_run_residual", "ctlib_run_residual_loss", "ctlib_residual_workspace_bytes", "ctlib_plan_geom", "ctlib_plan_view_info",
           "ctlib_plan_export", "ctlib_timing_enable", "ctlib_timing_collect", "ctlib_last_error",
           "ctlib_status_string", "ctlib_batch_create", "ctlib_batch_destroy", "ctlib_batch_workspace_bytes",
           "ctlib_batch_set_schedule", "ctlib_batch_bind", "ctlib_batch_run", "ctlib_batch_info",
           "ctlib_batch_last_error", "ctlib_plan_set_tuning", "ctlib_plan_set_chunks",
           "ctlib_plan_set_line_tables", "ctlib_batch_check",
           "ctlib_laplacian", "ctlib_laplacian_workspace_bytes", "ctlib_run_noalloc")


class Geom(ctypes.Structure):
    _fields_ = [("views", ctypes.c_int32), ("dets", ctypes.c_int32), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("scan_type", ctypes.c_int32), ("dImg", ctypes.c_double),
                ("dDet", ctypes.c_double), ("Ang0", ctypes.c_double), ("dAng", ctypes.c_double),
                ("s2r", ctypes.c_double), ("d2r", ctypes.c_double), ("binshift", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


VIEW_DTYPE = np.dtype({"names": ["sx", "sy", "lk", "ea_sc", "ea_off", "xmode", "rev", "nlines", "npix"],
                       "formats": ["<f8", "<f8", ("<f4", 4), "<f4", "<f4", "<i4", "<i4", "<i4", "<i4"],
                       "offsets": [0, 8, 16, 32, 36, 40, 44, 48, 52], "itemsize": 64})
LINE_DTYPE = np.dtype({"names": ["a", "s", "A", "B", "w0", "w1", "w2", "inv_s"],
                       "formats": ["<f8"] * 8, "offsets": [8 * i for i in range(8)], "itemsize": 64})

TABLES = {"views": (0, VIEW_DTYPE), "edges": (1, np.float64), "post_prj": (2, np.float64),
          "post_bprj_t": (3, np.float64), "lines": (4, LINE_DTYPE), "xviews": (5, np.int32),
          "yviews": (6, np.int32), "taps": (7, np.float64), "preweight": (8, np.float64),
          "ea_atan": (9, np.float32), "adj_lk": (10, np.float32),
          "adj_wk": (11, np.float32)}

_lib = None


class CtlibError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m ctlib_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.ctlib_parse_options.argtypes = [dp, ctypes.c_int, ctypes.POINTER(Geom)]
        lib.ctlib_plan_export.argtypes = [ctypes.POINTER(Geom), ctypes.c_int, ctypes.c_void_p,
                                          ctypes.POINTER(ctypes.c_size_t)]
        lib.ctlib_last_error.restype = ctypes.c_char_p
        lib.ctlib_status_string.restype = ctypes.c_char_p
        lib.ctlib_status_string.argtypes = [ctypes.c_int]
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        raise CtlibError(rc, load().ctlib_last_error().decode())


def parse_options(options):
    o = np.ascontiguousarray(np.asarray(options, dtype=np.float64).ravel())
    g = Geom()
    _check(load().ctlib_parse_options(o.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), o.size, ctypes.byref(g)))
    return g


def plan_export(options, name):
    """Host copy of one plan table (no GPU needed)."""
    g = parse_options(options)
    idx, dt = TABLES[name]
    n = ctypes.c_size_t(0)
    _check(load().ctlib_plan_export(ctypes.byref(g), idx, None, ctypes.byref(n)))
    buf = np.zeros(n.value, dtype=np.uint8)
    _check(load().ctlib_plan_export(ctypes.byref(g), idx, buf.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)))
    return buf.view(dt)


def exported_symbols():
    lib = load()
    return [s for s in EXPORTS if hasattr(lib, s)]
