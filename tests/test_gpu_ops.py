"""torch.library operators on the GPU: torch.library.opcheck (schema, fake tensor,
autograd registration, AOT dispatch), equality with the plain API, torch.compile
with fullgraph=True, gradients against the autograd module, and replay of a
captured HIP graph."""
import math

import numpy as np
import pytest
import torch

import ctlib_amd
from ctlib_amd import autograd, ops, phantom

pytestmark = pytest.mark.gpu


def small(st):
    if st == 2:
        return phantom.para_options(24, 17, 12, 10, 0.1, 0.13, Ang0=0.1, dAng=2 * math.pi / 24, binshift=0.02)
    return phantom.fan_options(24, 17, 12, 10, 0.1, 0.03 if st == 1 else 0.13, 3.0, 2.0, st, Ang0=0.1, binshift=0.02)


IMG_IN = ("projection", "backprojection_t", "backprojection_t_true", "fbp_t")
NAMES = ("projection", "projection_t", "backprojection", "backprojection_t", "backprojection_t_true", "fbp", "fbp_t")


def inputs(name, o, dev, dtype=torch.float64, B=2, seed=0):
    g = ctlib_amd.parse_options(o)
    gen = torch.Generator().manual_seed(seed)
    shape = (B, 1, g["height"], g["width"]) if name in IMG_IN else (B, 1, g["views"], g["dets"])
    return torch.rand(shape, generator=gen, dtype=dtype).to(dev)


@pytest.mark.parametrize("st", [0, 1, 2])
@pytest.mark.parametrize("name", NAMES)
def test_opcheck_and_matches_plain_api(dev, st, name):
    o = small(st)
    ol = ops.as_list(o)
    x = inputs(name, o, dev)
    print(f"[case] opcheck {name} st={st}", flush=True)
    op = getattr(torch.ops.ctlib_amd, name)
    torch.library.opcheck(op, (x, ol))
    assert torch.equal(op(x, ol), getattr(ctlib_amd, name)(x, torch.from_numpy(o)))


@pytest.mark.parametrize("st", [0, 1, 2])
@pytest.mark.parametrize("name", ["projection", "backprojection", "backprojection_t", "fbp"])
def test_op_gradients_match_autograd_module(dev, st, name):
    o = small(st)
    ol = ops.as_list(o)
    x = inputs(name, o, dev, seed=1).requires_grad_(True)
    w = torch.rand_like(getattr(ctlib_amd, name)(x.detach(), torch.from_numpy(o)))
    print(f"[case] op grad {name} st={st}", flush=True)
    (gx_op,) = torch.autograd.grad((getattr(torch.ops.ctlib_amd, name)(x, ol) * w).sum(), x)
    (gx_ag,) = torch.autograd.grad((getattr(autograd, name)(x, torch.from_numpy(o)) * w).sum(), x)
    assert torch.allclose(gx_op, gx_ag, rtol=1e-12, atol=0)


def test_residual_step_and_sv_ops(dev):
    o = small(0)
    ol = ops.as_list(o)
    x = inputs("projection", o, dev, dtype=torch.float32)
    m = inputs("projection_t", o, dev, dtype=torch.float32)
    print("[case] residual_step / backprojection_sv ops", flush=True)
    y, g = torch.ops.ctlib_amd.residual_step(x, m, ol)
    y2, g2 = ctlib_amd.residual_step(x, m, torch.from_numpy(o))
    assert torch.equal(y, y2) and torch.equal(g, g2)
    sv = torch.ops.ctlib_amd.backprojection_sv(m, ol)
    assert torch.equal(sv, ctlib_amd.backprojection_sv(m, torch.from_numpy(o)))
    torch.library.opcheck(torch.ops.ctlib_amd.residual_step, (x, m, ol),
                          test_utils=("test_schema", "test_faketensor"))


def test_compile_fullgraph_step(dev):
    o = phantom.CONFIGS[1]["options"]
    ol = ops.as_list(o)
    x = torch.from_numpy(phantom.phantom_batch(2, 512, 512, config=1, scale=0.02)).to(dev)
    m = ctlib_amd.projection(x * 1.05, torch.from_numpy(o))

    def step(x, m, alpha):
        r = torch.ops.ctlib_amd.projection(x, ol) - m
        return torch.relu(x - alpha * torch.ops.ctlib_amd.projection_t(r, ol)), 0.5 * (r * r).sum()

    print("[case] torch.compile fullgraph", flush=True)
    eager = step(x, m, 1e-3)
    compiled = torch.compile(step, fullgraph=True)
    for _ in range(2):
        got = compiled(x, m, 1e-3)
    assert torch.allclose(got[0], eager[0], rtol=1e-6, atol=1e-9)
    assert torch.allclose(got[1], eager[1], rtol=1e-6)


def test_graph_capture_replays(dev):
    o = phantom.CONFIGS[1]["options"]
    ol = ops.as_list(o)
    x = torch.from_numpy(phantom.phantom_batch(2, 512, 512, config=1, scale=0.02)).to(dev)
    ref = torch.ops.ctlib_amd.projection_t(torch.ops.ctlib_amd.projection(x, ol), ol)
    print("[case] graph capture of the ops", flush=True)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        torch.ops.ctlib_amd.projection_t(torch.ops.ctlib_amd.projection(x, ol), ol)  # warm-up (plan, workspace)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        out = torch.ops.ctlib_amd.projection_t(torch.ops.ctlib_amd.projection(x, ol), ol)
    for rep in range(3):
        out.zero_()
        junk = torch.rand(1 << 20, device=dev).sum()  # eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
        del junk


def test_captured_plan_survives_cache_clear_and_eviction(dev):
    """A graph captured through the public API keeps its plan's device tables alive: after
    clear_plan_cache() and after cycling more geometries than the cache holds (least recently
    used evicted), replays still equal the eager result."""
    from ctlib_amd import _C

    o = torch.from_numpy(phantom.fan_options(120, 97, 64, 64, 0.02, 0.03, 5.95, 4.906, 0))
    x = torch.from_numpy(phantom.phantom_batch(2, 64, 64, config=1, scale=0.02)).to(dev)
    ref = ctlib_amd.projection_t(ctlib_amd.projection(x, o), o)
    print("[case] captured plan survives clear / eviction", flush=True)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        ctlib_amd.projection_t(ctlib_amd.projection(x, o), o)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    pinned0 = _C.plan_cache_info()["pinned"]
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        out = ctlib_amd.projection_t(ctlib_amd.projection(x, o), o)
    info = _C.plan_cache_info()
    assert info["pinned"] == pinned0 + 1  # one geometry, pinned once
    _C.clear_plan_cache()
    assert _C.plan_cache_info()["entries"] == 0
    xs = torch.rand(1, 1, 16, 16, device=dev)
    cap = info["capacity"]
    for k in range(cap + 6):  # more geometries than the cache holds
        ok = torch.from_numpy(phantom.fan_options(8 + k, 11, 16, 16, 0.1, 0.13, 3.0, 2.0, 0))
        ctlib_amd.projection(xs, ok)
    info2 = _C.plan_cache_info()
    assert info2["entries"] == cap
    for rep in range(3):
        out.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
    # least recently used first: the newest geometries are still cached (no rebuild)
    created = _C.plan_cache_info()["created"]
    for k in range(cap + 6 - 8, cap + 6):
        ok = torch.from_numpy(phantom.fan_options(8 + k, 11, 16, 16, 0.1, 0.13, 3.0, 2.0, 0))
        ctlib_amd.projection(xs, ok)
    assert _C.plan_cache_info()["created"] == created


def test_reference_module_name_runs_unchanged(dev):
    """A script written against the reference (`import ctlib`, device options of the data dtype as
    ctlib.cpp:17 requires) runs unchanged and gives exactly the library's results."""
    import ctlib

    o = torch.from_numpy(phantom.fan_options(90, 80, 64, 64, 0.02, 0.03, 5.95, 4.906, 0)).float().to(dev)
    x = torch.from_numpy(phantom.phantom_batch(2, 64, 64, config=1, scale=0.02)).to(dev)
    print("[case] import ctlib", flush=True)
    sino = ctlib.projection(x, o)
    outs = {"projection": sino, "projection_t": ctlib.projection_t(sino, o),
            "backprojection": ctlib.backprojection(sino, o), "backprojection_t": ctlib.backprojection_t(x, o),
            "fbp": ctlib.fbp(sino, o), "backprojection_sv": ctlib.backprojection_sv(sino, o)}
    for name, got in outs.items():
        inp = x if name in ("projection", "backprojection_t") else sino
        assert torch.equal(got, getattr(ctlib_amd, name)(inp, o)), name
    pts = torch.rand(300, 5, device=dev)
    W = ctlib.laplacian(pts, 6)
    W2 = ctlib_amd.laplacian(pts, 6)
    assert torch.equal(W._indices(), W2._indices()) and torch.equal(W._values(), W2._values())
