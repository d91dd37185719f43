"""Seeded ellipse phantoms, Poisson-noised sinograms and reference geometries.

The phantom: 10 ellipses with centres U[-0.7, 0.7]^2, semi-axes U[0.05, 0.4],
rotation U[0, pi] and additive intensity U[0, 1] * scale, drawn per ellipse in
that order from numpy's PCG64(seed), on coordinates normalised to the image
half-extent (X = (c - W/2 + 1/2)/(W/2), Y = (H/2 - r - 1/2)/(H/2)); a pixel
holds the sum of the intensities of the ellipses containing its centre.
"""
import math

import numpy as np


def ellipse_phantom(height, width, seed, n=10, scale=1.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    c = (np.arange(width) - width / 2 + 0.5) / (width / 2)
    r = (height / 2 - np.arange(height) - 0.5) / (height / 2)
    X, Y = np.meshgrid(c, r)
    img = np.zeros((height, width))
    for _ in range(n):
        cx, cy = rng.uniform(-0.7, 0.7, 2)
        a, b = rng.uniform(0.05, 0.4, 2)
        phi = rng.uniform(0.0, math.pi)
        rho = rng.uniform(0.0, 1.0) * scale
        dx, dy = X - cx, Y - cy
        u = dx * math.cos(phi) + dy * math.sin(phi)
        v = -dx * math.sin(phi) + dy * math.cos(phi)
        img[(u / a) ** 2 + (v / b) ** 2 <= 1.0] += rho
    return img.astype(np.float32)


def phantom_batch(batch, height, width, config, scale=1.0):
    """[batch, 1, H, W] float32, slice s seeded with 1000 * config + s."""
    return np.stack([ellipse_phantom(height, width, 1000 * config + s, scale=scale)[None]
                     for s in range(batch)])


def poisson_sinogram(clean, I0, seed):
    """y = -ln(max(N, 1) / I0), N ~ Poisson(I0 * exp(-clean))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = rng.poisson(I0 * np.exp(-np.asarray(clean, dtype=np.float64)))
    return (-np.log(np.maximum(counts, 1) / I0)).astype(np.float32)


def fan_options(views, dets, width, height, dImg, dDet, s2r, d2r, scan_type, Ang0=0.0, dAng=None, binshift=0.0):
    if dAng is None:
        dAng = 2 * math.pi / views
    return np.array([views, dets, width, height, dImg, dDet, Ang0, dAng, s2r, d2r, binshift, scan_type],
                    dtype=np.float32)


def para_options(views, dets, width, height, dImg, dDet, Ang0=0.0, dAng=None, binshift=0.0):
    if dAng is None:
        dAng = math.pi / views
    return np.array([views, dets, width, height, dImg, dDet, Ang0, dAng, binshift, 2], dtype=np.float32)


# The geometries the projector family is measured on.
CONFIGS = {
    0: dict(options=para_options(180, 367, 256, 256, 1 / 256, 1 / 256), batch=1, scale=1.0),
    1: dict(options=fan_options(736, 736, 512, 512, 0.0066, 0.0072, 5.95, 4.906, 0), batch=8, scale=0.02),
    2: dict(options=fan_options(1024, 888, 512, 512, 0.0066, 0.0011, 5.95, 4.906, 1), batch=16, scale=0.02),
    4: dict(options=fan_options(1440, 1536, 1024, 1024, 0.0033, 0.0036, 5.95, 4.906, 0), batch=4, scale=0.02),
}


def hetero_slices(n=64, seed=3999):
    """Heterogeneous batch: per slice an image size from {256, 384, 512}^2, views U[96, 1152],
    bins from {512, 736, 1024} and scan_type from {0, 1, 2}.  Fan slices use the fan-ED / fan-EA
    geometries above (source 5.95, detector 4.906, same field of view and detector extent), parallel
    slices the parallel one; pixel and bin spacings scale with the counts.  Returns
    (options list, phantom list) with phantoms seeded 3000 + slice."""
    rng = np.random.default_rng(seed)
    opts, imgs = [], []
    for s in range(n):
        size = int(rng.choice([256, 384, 512]))
        views = int(rng.integers(96, 1153))
        dets = int(rng.choice([512, 736, 1024]))
        st = int(rng.choice([0, 1, 2]))
        if st == 0:
            o = fan_options(views, dets, size, size, 0.0066 * 512 / size, 0.0072 * 736 / dets, 5.95, 4.906, 0)
            scale = 0.02
        elif st == 1:
            o = fan_options(views, dets, size, size, 0.0066 * 512 / size, 0.0011 * 888 / dets, 5.95, 4.906, 1)
            scale = 0.02
        else:
            o = para_options(views, dets, size, size, 1.0 / size, (367 / 256) / dets)
            scale = 1.0
        opts.append(o)
        imgs.append(ellipse_phantom(size, size, 3000 + s, scale=scale))
    return opts, imgs
