"""torch.library registration of the projector family, checked without a GPU:
every operator exists under torch.ops.ctlib_amd, its fake implementation gives
the shapes the options imply, and a step built from them traces to one graph
(make_fx in fake mode) whose nodes are the registered operators."""
import math

import numpy as np
import pytest
import torch
from torch.fx.experimental.proxy_tensor import make_fx

import ctlib_amd  # noqa: F401
from ctlib_amd import ops, phantom
from torch._subclasses.fake_tensor import FakeTensorMode

CASES = {
    "fan_ed": phantom.CONFIGS[1]["options"],
    "fan_ea": phantom.CONFIGS[2]["options"],
    "para9": np.array([30, 21, 16, 12, 0.1, 0.1, 0.0, math.pi / 30, 2], dtype=np.float32),
}


@pytest.mark.parametrize("case", list(CASES))
def test_fake_shapes(case):
    o = ops.as_list(CASES[case])
    g = ctlib_amd.parse_options(CASES[case])
    with FakeTensorMode():
        x = torch.empty(3, 1, g["height"], g["width"], device="meta")
        y = torch.empty(3, 1, g["views"], g["dets"], device="meta")
        img, sino = (3, 1, g["height"], g["width"]), (3, 1, g["views"], g["dets"])
        assert torch.ops.ctlib_amd.projection(x, o).shape == sino
        assert torch.ops.ctlib_amd.backprojection_t(x, o).shape == sino
        assert torch.ops.ctlib_amd.backprojection_t_true(x, o).shape == sino
        assert torch.ops.ctlib_amd.fbp_t(x, o).shape == sino
        assert torch.ops.ctlib_amd.projection_t(y, o).shape == img
        assert torch.ops.ctlib_amd.backprojection(y, o).shape == img
        assert torch.ops.ctlib_amd.fbp(y, o).shape == img
        a, b = torch.ops.ctlib_amd.residual_step(x, y, o)
        assert a.shape == sino and b.shape == img
        if g["scan_type"] == 0:
            assert torch.ops.ctlib_amd.backprojection_sv(y, o).shape == (3, g["views"], g["height"], g["width"])
        else:
            with pytest.raises(ValueError):
                torch.ops.ctlib_amd.backprojection_sv(y, o)


def test_step_traces_to_registered_ops():
    o = ops.as_list(phantom.CONFIGS[1]["options"])

    def step(x, m):
        y = torch.ops.ctlib_amd.projection(x, o)
        g = torch.ops.ctlib_amd.projection_t(y - m, o)
        return x - 0.1 * g

    gm = make_fx(step, tracing_mode="fake")(torch.zeros(2, 1, 512, 512), torch.zeros(2, 1, 736, 736))
    targets = [str(n.target) for n in gm.graph.nodes if n.op == "call_function"]
    assert "ctlib_amd.projection.default" in targets and "ctlib_amd.projection_t.default" in targets
    out = [n for n in gm.graph.nodes if n.op == "output"][0]
    assert tuple(out.args[0][0].meta["val"].shape) == (2, 1, 512, 512) if isinstance(out.args[0], (list, tuple)) \
        else tuple(out.args[0].meta["val"].shape) == (2, 1, 512, 512)


def test_autograd_registered():
    for name in ("projection", "projection_t", "backprojection", "backprojection_t", "backprojection_t_true",
                 "fbp", "fbp_t"):
        op = getattr(torch.ops.ctlib_amd, name).default
        assert torch._C._dispatch_has_kernel_for_dispatch_key(op.name(), "Autograd"), name
