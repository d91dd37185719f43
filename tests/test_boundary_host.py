"""The drop-in boundary without a GPU: every public operator refuses host tensors the way the
reference's CHECK_INPUT does (RuntimeError, ctlib.cpp:7-9), options errors raise ValueError
instead of exit(0) (ctlib.cpp:26), and nothing falls back to a CPU computation."""
import math

import numpy as np
import pytest
import torch

import ctlib_amd
from ctlib_amd import autograd, phantom

O_FAN = torch.from_numpy(phantom.fan_options(24, 17, 12, 10, 0.1, 0.13, 3.0, 2.0, 0))
O_PARA = torch.from_numpy(phantom.para_options(24, 17, 12, 10, 0.1, 0.13, dAng=math.pi / 24))

IMG_IN = ("projection", "backprojection_t", "backprojection_t_true", "fbp_t")
SINO_IN = ("projection_t", "backprojection", "fbp", "backprojection_sv")


@pytest.mark.parametrize("name", IMG_IN + SINO_IN)
def test_host_tensor_is_refused(name):
    x = torch.zeros(1, 1, 10, 12) if name in IMG_IN else torch.zeros(1, 1, 24, 17)
    with pytest.raises(RuntimeError, match="GPU"):
        getattr(ctlib_amd, name)(x, O_FAN)


def test_other_entry_points_refuse_host_tensors():
    with pytest.raises(RuntimeError):
        ctlib_amd.residual_step(torch.zeros(1, 1, 10, 12), torch.zeros(1, 1, 24, 17), O_FAN)
    with pytest.raises(RuntimeError):
        ctlib_amd.laplacian(torch.rand(10, 3), 2)
    with pytest.raises(RuntimeError):
        autograd.projection(torch.zeros(1, 1, 10, 12, requires_grad=True), O_FAN)


@pytest.mark.parametrize("bad, err", [
    ([24, 17, 12, 10, 0.1, 0.13, 0.0, 0.26, 3.0, 2.0, 0.0, 5], ValueError),   # unknown scan_type
    ([24, 17, 12, 10, 0.1, 0.13, 0.0, 0.26, 3.0, 0], ValueError),             # fan vector too short
    ([24, 17, 12, 10, 0.1, 0.13, 2], ValueError),                              # parallel vector too short
    ([0, 17, 12, 10, 0.1, 0.13, 0.0, 0.26, 0.0, 2], ValueError),               # zero views
    ([24, 17.5, 12, 10, 0.1, 0.13, 0.0, 0.26, 0.0, 2], ValueError),            # fractional size
])
def test_options_errors_raise(bad, err):
    with pytest.raises(err):
        ctlib_amd.parse_options(torch.tensor(bad, dtype=torch.float64))


def test_nine_entry_parallel_vector_reads_binshift_from_entry_8():
    o9 = torch.tensor([24, 17, 12, 10, 0.1, 0.13, 0.0, math.pi / 24, 2], dtype=torch.float32)
    g = ctlib_amd.parse_options(o9)
    assert g["scan_type"] == 2 and g["binshift"] == 2.0
    g10 = ctlib_amd.parse_options(O_PARA)
    assert g10["binshift"] == 0.0 and g10["views"] == 24 and g10["dets"] == 17
    assert ctlib_amd.parse_options(O_FAN)["s2r"] == pytest.approx(3.0)


def test_public_surface_matches_the_reference_module():
    # ctlib.cpp:106-114 registers these seven names
    for name in ("projection", "projection_t", "backprojection", "backprojection_t", "backprojection_sv", "fbp",
                 "laplacian"):
        assert callable(getattr(ctlib_amd, name)), name


def test_reference_module_name_imports():
    """`import ctlib` (the reference's module name, setup.py:9) exposes exactly the seven functions of
    ctlib.cpp:106-114, each the library's own entry point, and refuses host tensors like the reference."""
    import ctlib

    names = ("projection", "projection_t", "backprojection_t", "backprojection", "backprojection_sv", "fbp",
             "laplacian")
    assert sorted(ctlib.__all__) == sorted(names)
    for name in names:
        assert getattr(ctlib, name) is getattr(ctlib_amd, name), name
    with pytest.raises(RuntimeError, match="GPU"):
        ctlib.projection(torch.zeros(1, 1, 10, 12), O_FAN)
