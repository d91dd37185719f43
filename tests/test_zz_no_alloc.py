"""ctlib_run_noalloc (the allocation-free variant of ctlib_run) through the C ABI: its
output equals ctlib_run's on the same inputs for every operator, on two calls in a row
with the second on a smaller batch, and a batch whose scratch exceeds the fixed array
is refused before anything is launched."""
import ctypes

import numpy as np
import pytest
import torch

from ctlib_amd import abi, phantom

pytestmark = pytest.mark.gpu

OPS = {"projection": (0, "img"), "projection_t": (1, "sino"), "backprojection": (2, "sino"),
       "backprojection_t": (3, "img"), "backprojection_t_true": (4, "img"), "fbp": (5, "sino"),
       "fbp_t": (6, "img")}
SCRATCH_BYTES = 64 << 20


def _lib():
    lib = abi.load()
    lib.ctlib_plan_create.argtypes = [ctypes.POINTER(abi.Geom), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.ctlib_plan_destroy.argtypes = [ctypes.c_void_p]
    lib.ctlib_workspace_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int]
    lib.ctlib_workspace_bytes.restype = ctypes.c_size_t
    lib.ctlib_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_int64, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.ctlib_run_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    return lib


class Plan:
    def __init__(self, options, device):
        self.lib = _lib()
        self.g = abi.parse_options(options)
        self.p = ctypes.c_void_p()
        rc = self.lib.ctlib_plan_create(ctypes.byref(self.g), device, ctypes.byref(self.p))
        assert rc == 0, self.lib.ctlib_last_error()

    def close(self):
        self.lib.ctlib_plan_destroy(self.p)

    def shapes(self, kind, B):
        g = self.g
        img, sino = (B, 1, g.height, g.width), (B, 1, g.views, g.dets)
        return (img, sino) if kind == "img" else (sino, img)

    def run(self, op, x, out):
        """ctlib_run with a workspace from torch."""
        dt = 0 if x.dtype == torch.float32 else 1
        B = x.shape[0]
        wsb = self.lib.ctlib_workspace_bytes(self.p, op, B, dt)
        ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=x.device)
        s = torch.cuda.current_stream(x.device).cuda_stream
        rc = self.lib.ctlib_run(self.p, op, dt, x.data_ptr(), out.data_ptr(), B, ws.data_ptr(), wsb, s)
        assert rc == 0, self.lib.ctlib_last_error()

    def run_noalloc(self, op, x, out):
        dt = 0 if x.dtype == torch.float32 else 1
        s = torch.cuda.current_stream(x.device).cuda_stream
        return self.lib.ctlib_run_noalloc(self.p, op, dt, x.data_ptr(), out.data_ptr(), x.shape[0], s)


def rel(a, b):
    return float(torch.linalg.vector_norm((a - b).double()) / torch.linalg.vector_norm(b.double()))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("cfg", [0, 1, 2])
def test_noalloc_matches_run(dev, cfg, dtype):
    opts = phantom.CONFIGS[cfg]["options"]
    plan = Plan(opts, dev.index or 0)
    try:
        gen = torch.Generator().manual_seed(100 + cfg)
        for name, (op, kind) in OPS.items():
            for B in (2, 1):  # second call on a smaller input
                sin, sout = plan.shapes(kind, B)
                x = torch.rand(sin, generator=gen, dtype=torch.float64).to(dev, dtype)
                ref = torch.empty(sout, dtype=dtype, device=dev)
                got = torch.full(sout, float("nan"), dtype=dtype, device=dev)
                print(f"noalloc cfg {cfg} {dtype} {name} B={B}", flush=True)
                plan.run(op, x, ref)
                rc = plan.run_noalloc(op, x, got)
                assert rc == 0, plan.lib.ctlib_last_error()
                torch.cuda.synchronize()
                tol = 1e-5 if dtype == torch.float32 else 1e-7
                err = rel(got, ref)
                assert err <= tol, (name, B, err)
    finally:
        plan.close()


def test_noalloc_builds_line_tables_per_block_when_they_do_not_fit(dev):
    """projection at fan-ED 512^2, 736 x 736, B=8: its prebuilt line tables alone (8 x 1024 lines x
    514 entries x 16 bytes = 67 MB) exceed the fixed array, so ctlib_run_noalloc builds the tables in
    each block (and splits the lines into fewer chunks) - still ctlib_run's result, which comes from
    the prebuilt tables, to fp32 summation order."""
    opts = phantom.CONFIGS[1]["options"]
    plan = Plan(opts, dev.index or 0)
    try:
        g = plan.g
        B = 8
        assert B * (g.width + g.height) * 514 * 16 > SCRATCH_BYTES
        assert plan.lib.ctlib_workspace_bytes(plan.p, 0, B, 0) > SCRATCH_BYTES
        gen = torch.Generator().manual_seed(7)
        for Bc in (B, 3):  # the second call needs no fallback
            x = torch.rand(Bc, 1, g.height, g.width, generator=gen).to(dev)
            ref = torch.empty(Bc, 1, g.views, g.dets, device=dev)
            got = torch.full((Bc, 1, g.views, g.dets), float("nan"), device=dev)
            print(f"noalloc projection per-block tables B={Bc}", flush=True)
            plan.run(0, x, ref)
            rc = plan.run_noalloc(0, x, got)
            assert rc == 0, plan.lib.ctlib_last_error()
            torch.cuda.synchronize()
            assert rel(got, ref) <= 1e-6, (Bc, rel(got, ref))
    finally:
        plan.close()


def test_noalloc_refuses_oversized_scratch(dev):
    """fbp at fan-ED 512^2, 736 x 736, B=16 needs the ramp rows and bin tables (> 64 MiB even
    unsplit): refused with the workspace status, the output untouched."""
    opts = phantom.CONFIGS[1]["options"]
    plan = Plan(opts, dev.index or 0)
    try:
        g = plan.g
        B = 16
        wsb = plan.lib.ctlib_workspace_bytes(plan.p, 5, B, 0)
        assert B * g.views * (g.dets + 2) * 16 > SCRATCH_BYTES and wsb > SCRATCH_BYTES
        x = torch.rand(B, 1, g.views, g.dets, device=dev)
        out = torch.full((B, 1, g.height, g.width), 7.0, device=dev)
        print("noalloc oversized fbp B=16", flush=True)
        rc = plan.run_noalloc(5, x, out)
        assert rc == 6, rc  # CTLIB_ERR_WORKSPACE
        assert b"scratch" in plan.lib.ctlib_last_error()
        torch.cuda.synchronize()
        assert bool((out == 7.0).all())
    finally:
        plan.close()
