"""FP64 oracle of the motion-proposal network and of the refinement rule.

TEST INFRASTRUCTURE ONLY (the checker of umi_amd/csrc/proposer.hip).  The
network: 3x3 convolutions with dilation d and zero padding d, + bias, ReLU
on all but the last layer; Huber loss (delta 1) against block labels / 4,
mean over all outputs.  The backward pass is written out analytically and
pinned by finite differences in tests/test_oracle_proposer.py.
"""
from __future__ import annotations

import numpy as np


def shift(x, dy, dx):
    """y[..., i, j] = x[..., i + dy, j + dx], zero outside."""
    out = np.zeros_like(x)
    H, W = x.shape[-2:]
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    out[..., yd, xd] = x[..., ys, xs]
    return out


def taps(d):
    return [((t // 3 - 1) * d, (t % 3 - 1) * d) for t in range(9)]


def conv(x, w, b, d):
    N, Ci, H, W = x.shape
    Co = w.shape[0]
    out = np.broadcast_to(np.asarray(b, np.float64)[None, :, None, None], (N, Co, H, W)).copy()
    for t, (dy, dx) in enumerate(taps(d)):
        xs = shift(x, dy, dx)
        out += np.einsum("oc,nchw->nohw", w[:, :, t // 3, t % 3], xs)
    return out


def forward(x, params, layers):
    acts = []
    a = np.asarray(x, np.float64)
    for (w, b), (_, _, d, relu) in zip(params, layers):
        a = conv(a, np.asarray(w, np.float64), b, d)
        if relu:
            a = np.maximum(a, 0.0)
        acts.append(a)
    return acts


def huber(pred, labels):
    """(loss, dloss/dpred); labels [N][H4/2][W4/2][3] int, target = label / 4."""
    N, _, H4, W4 = pred.shape
    tgt = np.repeat(np.repeat(labels[..., :2].transpose(0, 3, 1, 2).astype(np.float64) / 4.0, 2, axis=2), 2, axis=3)
    e = pred - tgt
    ae = np.abs(e)
    n = pred.size
    loss = np.where(ae <= 1, 0.5 * e * e, ae - 0.5).sum() / n
    return loss, np.clip(e, -1, 1) / n


def backward(x, params, layers, acts, g_out):
    """Analytic gradients [(dw, db)] of every layer given dloss/d(last output)."""
    grads = [None] * len(layers)
    g = g_out
    for li in range(len(layers) - 1, -1, -1):
        w = np.asarray(params[li][0], np.float64)
        d, relu = layers[li][2], layers[li][3]
        if relu:
            g = g * (acts[li] > 0)
        inp = np.asarray(x, np.float64) if li == 0 else acts[li - 1]
        dw = np.zeros_like(w)
        for t, (dy, dx) in enumerate(taps(d)):
            dw[:, :, t // 3, t % 3] = np.einsum("nohw,nchw->oc", g, shift(inp, dy, dx))
        db = g.sum(axis=(0, 2, 3))
        grads[li] = (dw, db)
        if li > 0:
            gi = np.zeros_like(inp)
            for t, (dy, dx) in enumerate(taps(d)):
                gi += shift(np.einsum("oc,nohw->nchw", w[:, :, t // 3, t % 3], g), -dy, -dx)
            g = gi
    return grads


def block_sad(cur, ref, by, bx, dy, dx):
    H, W = cur.shape
    ys = np.clip(np.arange(by * 8, by * 8 + 8) + dy, 0, H - 1)
    xs = np.clip(np.arange(bx * 8, bx * 8 + 8) + dx, 0, W - 1)
    return int(np.abs(cur[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8].astype(np.int64) - ref[np.ix_(ys, xs)]).sum())


def refine(cur, ref, prop, R, slack):
    """(mv_sad [nblk][3], fallback set) of the +-2 refinement rule."""
    H, W = cur.shape
    nby, nbx = H // 8, W // 8
    out = np.zeros((nby * nbx, 3), np.int64)
    fb = set()
    for by in range(nby):
        for bx in range(nbx):
            b = by * nbx + bx
            py, px = prop[b]
            best, bi = None, 0
            for i in range(25):
                dy = int(np.clip(py + i // 5 - 2, -R, R))
                dx = int(np.clip(px + i % 5 - 2, -R, R))
                s = block_sad(cur, ref, by, bx, dy, dx)
                if best is None or s < best:
                    best, bi, bmv = s, i, (dy, dx)
            nb = None
            for o in (0, 1, 2, 3, 5, 6, 7, 8):
                ny, nx = by + o // 3 - 1, bx + o % 3 - 1
                if 0 <= ny < nby and 0 <= nx < nbx:
                    s = block_sad(cur, ref, by, bx, *prop[ny * nbx + nx])
                    nb = s if nb is None else min(nb, s)
            out[b] = (bmv[0], bmv[1], best)
            border = bi // 5 in (0, 4) or bi % 5 in (0, 4)
            worse = nb is not None and best > slack * min(nb, best)
            if border or worse:
                fb.add(b)
    return out, fb
