"""ctypes front end of the FP64 CPU oracle (oracle/ct_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the ctlib_amd package.

Every function takes the options vector as the reference does (12 entries for
fan beam, >= 9 for parallel beam, scan_type last; ctlib.cpp:18), converted to
float64 from whatever dtype the caller used, and works on float64 numpy arrays
shaped like the reference's tensors without the channel axis:
image [B, H, W], sinogram [B, V, D], per-view images (backprojection_sv)
[B, V, H, W].
"""
import ctypes
import os
import subprocess

import numpy as np

OP_PRJ, OP_PRJ_T, OP_BPRJ, OP_BPRJ_T, OP_BPRJ_T_TRUE, OP_FBP, OP_BPRJ_SV = range(7)
OP_NAMES = {
    "projection": OP_PRJ,
    "projection_t": OP_PRJ_T,
    "backprojection": OP_BPRJ,
    "backprojection_t": OP_BPRJ_T,
    "backprojection_t_true": OP_BPRJ_T_TRUE,
    "fbp": OP_FBP,
    "backprojection_sv": OP_BPRJ_SV,
}

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libct_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    """Compile the oracle library with its Makefile (g++ -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        lib.ctor_parse.argtypes = [_dp, ctypes.c_int, _ip, _dp]
        lib.ctor_view.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp]
        lib.ctor_apply.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int]
        lib.ctor_ramp.argtypes = [_dp, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int]
        lib.ctor_rl_taps.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        lib.ctor_fbp_preweight.argtypes = [_dp, ctypes.c_int, _dp, _dp]
        lib.ctor_brute.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int]
        lib.ctor_max_threads.argtypes = []
        _lib = lib
    return _lib


def _opt(options):
    o = np.ascontiguousarray(np.asarray(options, dtype=np.float64).ravel())
    return o, o.ctypes.data_as(_dp), o.size


def _p(a):
    return a.ctypes.data_as(_dp)


class OracleError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise OracleError(f"{what}: oracle status {rc}")


def parse(options):
    o, op, n = _opt(options)
    ints = (ctypes.c_int * 5)()
    dbl = (ctypes.c_double * 7)()
    _check(_load().ctor_parse(op, n, ints, dbl), "parse")
    keys_i = ("views", "dets", "width", "height", "scan_type")
    keys_d = ("dImg", "dDet", "Ang0", "dAng", "s2r", "d2r", "binshift")
    g = {k: ints[i] for i, k in enumerate(keys_i)}
    g.update({k: dbl[i] for i, k in enumerate(keys_d)})
    return g


def view(options, v):
    """(xmode, rev, internal edges [D+1], bin cosines [D]) of view v."""
    g = parse(options)
    o, op, n = _opt(options)
    mr = (ctypes.c_int * 2)()
    e = np.zeros(g["dets"] + 1)
    c = np.zeros(g["dets"])
    _check(_load().ctor_view(op, n, v, mr, _p(e), _p(c)), "view")
    return bool(mr[0]), bool(mr[1]), e, c


def _shape_in(op, g):
    if op in (OP_PRJ, OP_BPRJ_T, OP_BPRJ_T_TRUE):
        return (g["height"], g["width"]), (g["views"], g["dets"])
    if op == OP_BPRJ_SV:
        return (g["views"], g["dets"]), (g["views"], g["height"], g["width"])
    return (g["views"], g["dets"]), (g["height"], g["width"])


def apply(op, options, x, chunk=0, threads=0):
    """Apply operator `op` (code or name) to float64 array x of shape [B, ...]."""
    if isinstance(op, str):
        op = OP_NAMES[op]
    g = parse(options)
    sin, sout = _shape_in(op, g)
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    if x.ndim == 2:
        x = x[None]
    if x.shape[1:] != sin:
        raise ValueError(f"oracle input shape {x.shape} does not match {sin}")
    B = x.shape[0]
    out = np.zeros((B,) + sout)
    o, op_, n = _opt(options)
    _check(_load().ctor_apply(op, op_, n, _p(x), _p(out), B, chunk, threads), "apply")
    return out


def brute(op, options, x):
    """Independent all-pairs interval-intersection implementation (small sizes only)."""
    if isinstance(op, str):
        op = OP_NAMES[op]
    g = parse(options)
    sin, sout = _shape_in(op, g)
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    if x.ndim == 2:
        x = x[None]
    B = x.shape[0]
    out = np.zeros((B,) + sout)
    o, op_, n = _opt(options)
    _check(_load().ctor_brute(op, op_, n, _p(x), _p(out), B), "brute")
    return out


def ramp(options, y, threads=0):
    """Pre-weighted RL-filtered sinogram rows (the conv2d stage of fbp)."""
    g = parse(options)
    y = np.ascontiguousarray(np.asarray(y, dtype=np.float64))
    rows = y.size // g["dets"]
    q = np.zeros_like(y)
    o, op_, n = _opt(options)
    _check(_load().ctor_ramp(op_, n, _p(y), _p(q), rows, threads), "ramp")
    return q


def rl_taps(D, delta):
    h = np.zeros(2 * D - 1)
    _check(_load().ctor_rl_taps(D, delta, _p(h)), "rl_taps")
    return h


def fbp_preweight(options):
    g = parse(options)
    pw = np.zeros(g["dets"])
    delta = ctypes.c_double()
    o, op_, n = _opt(options)
    _check(_load().ctor_fbp_preweight(op_, n, _p(pw), ctypes.byref(delta)), "preweight")
    return pw, delta.value


def max_threads():
    return _load().ctor_max_threads()
