"""Numpy model of the GPU kernels' arithmetic (cumulative-integral tables).

Runs the exact algorithm of ctlib_amd/csrc/dd_kernels.hip on the host, on the
plan tables the library exports (ctlib_plan_export), so that the geometry plan
and the table formulation can be checked against the FP64 oracle without a GPU.
Float32 steps of the kernels (table index from tau, the Moebius / atan bin
lookup, the pixel-weight reciprocal) are reproduced in float32.
"""
import numpy as np

from ctlib_amd import abi


class Plan:
    def __init__(self, options):
        self.g = abi.parse_options(options).as_dict()
        self.vp = abi.plan_export(options, "views")
        V, D = self.g["views"], self.g["dets"]
        self.edges = abi.plan_export(options, "edges").reshape(V, D + 1)
        self.post_prj = abi.plan_export(options, "post_prj").reshape(V, D)
        self.post_bt = abi.plan_export(options, "post_bprj_t").reshape(V, D)
        self.max_lines = max(self.g["height"], self.g["width"])
        self.lines = abi.plan_export(options, "lines").reshape(V, self.max_lines)
        self.taps = abi.plan_export(options, "taps")
        self.preweight = abi.plan_export(options, "preweight")


def _line_values(x, xmode):
    """Lines of a batch of images: X rows left->right; Y columns bottom->top."""
    return x if xmode else np.ascontiguousarray(np.flip(x, axis=1).transpose(0, 2, 1))


def _put_lines(img, vals, xmode):
    if xmode:
        img += vals
    else:
        img += np.flip(vals.transpose(0, 2, 1), axis=1)


def forward(plan, x, weighted=False, post=None):
    """Line -> bin operator: projection (post = post_prj) or backprojection_t (weighted)."""
    g = plan.g
    V, D = g["views"], g["dets"]
    B = x.shape[0]
    y = np.zeros((B, V, D))
    for v in range(V):
        vp = plan.vp[v]
        xmode = bool(vp["xmode"])
        L = _line_values(x.astype(np.float64), xmode)   # [B, nl, n]
        nl, n = L.shape[1], L.shape[2]
        e = plan.edges[v]
        S = np.zeros((B, D + 1))
        for l in range(nl):
            lc = plan.lines[v, l]
            vals = L[:, l, :]
            if weighted:
                p = np.arange(n, dtype=np.float64)
                u = lc["w1"] * p + lc["w0"]
                d = u * u + lc["w2"]
                vals = vals * (lc["inv_s"] * (np.float32(1.0) / d.astype(np.float32)).astype(np.float64))
            P = np.concatenate([np.zeros((B, 1)), np.cumsum(vals, axis=1)], axis=1)  # [B, n+1]
            j = np.arange(1, n + 1)
            C = np.zeros((B, n + 2))
            sl = np.zeros((B, n + 2))
            C[:, 1:n + 1] = P[:, :n] - j * vals
            sl[:, 1:n + 1] = vals
            C[:, n + 1] = P[:, n]
            tau = lc["A"] * e + lc["B"]
            jj = np.clip(tau.astype(np.float32), 0, np.float32(n + 1)).astype(np.int64)
            F = C[:, jj] + tau * sl[:, jj]
            S += lc["s"] * F
        val = S[:, 1:] - S[:, :-1]
        if post is not None:
            val = val * post[v]
        o = np.arange(D)[::-1] if vp["rev"] else np.arange(D)
        y[:, v, o] = val
    return y


def bin_tables(plan, y, pre=None, scale=1.0):
    g = plan.g
    V, D = g["views"], g["dets"]
    B = y.shape[0]
    tabs = np.zeros((B, V, D + 2, 2))
    for v in range(V):
        o = np.arange(D)[::-1] if plan.vp[v]["rev"] else np.arange(D)
        gk = y[:, v, o].astype(np.float64)
        if pre is not None:
            gk = gk * pre[v]
        gk = gk * scale
        e = plan.edges[v]
        w = e[1:] - e[:-1]
        Q = np.concatenate([np.zeros((B, 1)), np.cumsum(w * gk, axis=1)], axis=1)
        tabs[:, v, 1:D + 1, 0] = Q[:, :D] - e[:D] * gk
        tabs[:, v, 1:D + 1, 1] = gk
        tabs[:, v, D + 1, 0] = Q[:, D]
    return tabs


def _lookup(vp, pi, D, scan_type):
    pf = pi.astype(np.float32)
    lk = vp["lk"].astype(np.float32)
    if scan_type == 2:
        q = lk[0] * pf + lk[1]
    else:
        num = lk[0] * pf + lk[1]
        den = lk[2] * pf + lk[3]
        q = (num * (np.float32(1) / den)).astype(np.float32)
        if scan_type == 1:
            q = (np.arctan(q).astype(np.float32) * np.float32(vp["ea_sc"]) + np.float32(vp["ea_off"]))
    return np.clip(q, 0, np.float32(D + 1)).astype(np.int64)


def adjoint(plan, y, weighted=False, pre=None, wscale=1.0):
    """Bin -> pixel operator: projection_t (pre = post_prj) or backprojection (weighted)."""
    g = plan.g
    V, D, H, W = g["views"], g["dets"], g["height"], g["width"]
    B = y.shape[0]
    tabs = bin_tables(plan, y, pre)
    img = np.zeros((B, H, W))
    for v in range(V):
        vp = plan.vp[v]
        xmode = bool(vp["xmode"])
        nl, n = (H, W) if xmode else (W, H)
        out = np.zeros((B, nl, n))
        for l in range(nl):
            lc = plan.lines[v, l]
            p = np.arange(n + 1, dtype=np.float64)
            pi = lc["s"] * p + lc["a"]
            j = _lookup(vp, pi, D, g["scan_type"])
            G = tabs[:, v, j, 0] + pi * tabs[:, v, j, 1]
            diff = G[:, 1:] - G[:, :-1]
            if weighted:
                pp = np.arange(n, dtype=np.float64)
                u = lc["w1"] * pp + lc["w0"]
                d = u * u + lc["w2"]
                diff = diff * (lc["inv_s"] * wscale * (np.float32(1.0) / d.astype(np.float32)).astype(np.float64))
            out[:, l, :] = diff
        _put_lines(img, out, xmode)
    return img


def ramp(plan, y, pre=None, post=None, scale=1.0):
    D = plan.g["dets"]
    h = plan.taps
    T = h[(np.arange(D)[:, None] - np.arange(D)[None, :]) + D - 1]  # T[m][j] = h(m - j)
    a = y.reshape(-1, D).astype(np.float64)
    if pre is not None:
        a = a * pre
    out = a @ T
    if post is not None:
        out = out * post
    return (out * scale).reshape(y.shape)


def apply(plan, op, x):
    """All public ops plus the exact adjoints, as the library composes them."""
    st = plan.g["scan_type"]
    if op == "projection":
        return forward(plan, x, False, plan.post_prj)
    if op == "backprojection_t":
        return forward(plan, x, True, plan.post_bt if st == 1 else None)
    if op == "backprojection_t_true":
        return forward(plan, x, True, None)
    if op == "projection_t":
        return adjoint(plan, x, False, plan.post_prj)
    if op == "backprojection":
        return adjoint(plan, x, True, None)
    if op == "fbp":
        q = ramp(plan, x, pre=plan.preweight)
        return adjoint(plan, q, True, None, wscale=plan.g["dAng"] / 2)
    if op == "fbp_t":
        z = forward(plan, x, True, None)
        return ramp(plan, z, post=plan.preweight, scale=plan.g["dAng"] / 2)
    raise ValueError(op)
