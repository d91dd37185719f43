"""backprojection_sv (per-view backprojection, ctlib.cpp:45-54, fan_ed_kernel.cu:945-1161)
against the FP64 oracle, and against the GPU backprojection it decomposes.

Bars: relative L2 <= 1e-5 (fp32) and <= 1e-7 (fp64) against the oracle over the
whole [B, V, H, W] output; the view sum equals backprojection to 1e-6.
"""
import math

import numpy as np
import pytest
import torch

import ctlib_amd
from ctlib_amd import _C, phantom
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def small(st, V=24, D=17, W=12, H=10, Ang0=0.1, bs=0.02, dAng=None, dImg=0.1):
    dAng = 2 * math.pi / V if dAng is None else dAng
    if st == 2:
        return phantom.para_options(V, D, W, H, dImg, 0.13, Ang0=Ang0, dAng=dAng, binshift=bs)
    return phantom.fan_options(V, D, W, H, dImg, 0.03 if st == 1 else 0.13, 3.0, 2.0, st, Ang0=Ang0, dAng=dAng,
                               binshift=bs)


CASES = {
    "ed": (0, {}),
    "ed_odd": (0, dict(V=36, D=16, W=9, H=14, Ang0=-0.4, bs=-0.07)),
    "ed_ties": (0, dict(V=8, D=21, W=10, H=10, Ang0=0.0, dAng=math.pi / 4)),
    "ed_wide_rows": (0, dict(V=30, D=70, W=130, H=66, dImg=0.02)),   # three 64-column groups per row
    "ea": (1, {}),
    "para": (2, {}),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sv_small_against_oracle(dev, name, dtype):
    st, kw = CASES[name]
    o = small(st, **kw)
    g = O.parse(o)
    rng = np.random.default_rng(31)
    y = rng.random((2, 1, g["views"], g["dets"])).astype(dtype)
    print(f"[case] sv {name} {np.dtype(dtype).name}", flush=True)
    got = _C.backprojection_sv(torch.from_numpy(y).to(dev), torch.from_numpy(o), True).cpu().numpy()
    ref = O.apply("backprojection_sv", o.astype(np.float64), y[:, 0].astype(np.float64))
    assert got.dtype == dtype and got.shape == ref.shape == (2, g["views"], g["height"], g["width"])
    tol = 1e-5 if dtype == np.float32 else 1e-7
    assert rel_l2(got, ref) <= tol, (name, rel_l2(got, ref))


def test_sv_public_api_is_fan_ed_only(dev):
    o = small(0)
    y = torch.rand(1, 1, 24, 17, device=dev)
    print("[case] sv public api", flush=True)
    out = ctlib_amd.backprojection_sv(y, torch.from_numpy(o))
    assert out.shape == (1, 24, 10, 12)
    for st in (1, 2):
        with pytest.raises(ValueError):
            ctlib_amd.backprojection_sv(y, torch.from_numpy(small(st)))
    assert ctlib_amd.backprojection_sv(torch.zeros(0, 1, 24, 17, device=dev), torch.from_numpy(o)).shape == \
        (0, 24, 10, 12)


def test_sv_config1_slice(dev):
    """Fan-ED 512^2, 736 views x 736 bins, one slice (the reference's only sv geometry)."""
    o = phantom.CONFIGS[1]["options"]
    x = phantom.phantom_batch(1, 512, 512, config=1, scale=0.02)
    y = O.apply("projection", o, x[:, 0].astype(np.float64))[:, None].astype(np.float32)
    yt = torch.from_numpy(y).to(dev)
    ot = torch.from_numpy(o)
    print("[case] sv config1", flush=True)
    sv = ctlib_amd.backprojection_sv(yt, ot)
    full = ctlib_amd.backprojection(yt, ot)
    s = sv.double().sum(dim=1, keepdim=True)
    assert rel_l2(s.cpu().numpy(), full.double().cpu().numpy()) <= 1e-6
    got = sv.cpu().numpy()
    del sv, s
    ref = O.apply("backprojection_sv", o.astype(np.float64), y[:, 0].astype(np.float64))
    err = rel_l2(got, ref)
    print(f"  rel_l2={err:.3e}")
    assert err <= 1e-5


def test_sv_deterministic_and_batch_independent(dev):
    o = small(0, V=64, D=41, W=32, H=24)
    rng = np.random.default_rng(8)
    y = torch.from_numpy(rng.random((3, 1, 64, 41)).astype(np.float32)).to(dev)
    ot = torch.from_numpy(o)
    print("[case] sv determinism", flush=True)
    a = ctlib_amd.backprojection_sv(y, ot)
    b = ctlib_amd.backprojection_sv(y, ot)
    assert torch.equal(a, b)
    assert torch.equal(ctlib_amd.backprojection_sv(y[1:2].contiguous(), ot), a[1:2])
