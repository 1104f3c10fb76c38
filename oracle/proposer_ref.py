"""FP64 oracle of the motion-proposal network and of the refinement rule.

TEST INFRASTRUCTURE ONLY (the checker of umi_amd/csrc/proposer.hip).  The
network: 3x3 convolutions with dilation d and zero padding d, + bias, ReLU
on all but the last layer; Huber loss (delta 1) against block labels / 4,
mean over all outputs.  The backward pass is written out analytically and
pinned by finite differences in tests/test_oracle_proposer.py.

The refinement (seed, iterated +-2 descent, in-tile and global jump-flood
propagation rounds, fallback check) has no counterpart in the reference; it is pinned in
tests/test_oracle_proposer.py by the C oracle's SAD volume, by exact-shift
frames (SAD 0 at the true vector), by the local-minimum property of every
converged block and by a scalar recomputation of the fallback test.
"""
from __future__ import annotations

import numpy as np


def shift(x, dy, dx):
    """y[..., i, j] = x[..., i + dy, j + dx], zero outside."""
    out = np.zeros_like(x)
    H, W = x.shape[-2:]
    if abs(dy) >= H or abs(dx) >= W:
        return out
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


def bf16(x):
    """Round to bfloat16 (via float32, round to nearest even), as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def forward_bf16(x, params, layers):
    """Inference form of the network: inputs, weights and the activations
    between layers rounded to bf16, sums in float64; returns the last output."""
    a = bf16(x)
    for li, ((w, b), (_, _, d, relu)) in enumerate(zip(params, layers)):
        a = conv(a, bf16(w), np.asarray(b, np.float64), d)
        if relu:
            a = np.maximum(a, 0.0)
        if li < len(layers) - 1:
            a = bf16(a)
    return a


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


# distances of in-tile round p: 1 and LOCAL_JUMPS[p % 3]; of global round p:
# 1 and the two of GLOBAL_JUMPS[p % 4]
LOCAL_JUMPS = ((4,), (2,), ())
GLOBAL_JUMPS = ((8, 32), (4, 16), (2, 8), (4, 32))
TILE = (4, 8)            # tile of blocks (rows, columns) for the in-tile rounds, aligned to the block grid
RING = (0, 1, 2, 3, 5, 6, 7, 8)            # 3x3 neighbourhood in raster order without the centre (4)


class _Blocks:
    """All 8x8 blocks of a frame pair, with vectorised clamped SADs."""

    def __init__(self, cur, ref):
        self.H, self.W = cur.shape
        self.nby, self.nbx = self.H // 8, self.W // 8
        self.n = self.nby * self.nbx
        self.by, self.bx = np.divmod(np.arange(self.n), self.nbx)
        self.cur = (cur[:self.nby * 8, :self.nbx * 8].reshape(self.nby, 8, self.nbx, 8)
                    .transpose(0, 2, 1, 3).reshape(self.n, 8, 8).astype(np.int64))
        self.ref = np.asarray(ref)

    def sad(self, mv):
        """SAD of every block at its vector mv [n][2] (ref coordinates clamped)."""
        ys = np.clip(self.by[:, None] * 8 + np.arange(8)[None] + mv[:, :1], 0, self.H - 1)
        xs = np.clip(self.bx[:, None] * 8 + np.arange(8)[None] + mv[:, 1:2], 0, self.W - 1)
        return np.abs(self.cur - self.ref[ys[:, :, None], xs[:, None, :]]).sum((1, 2))

    def neighbour(self, o, d, tile=False):
        """(mask, index) of neighbour o (0..8, raster, 4 = self) at distance d:
        inside the frame and, when tile is set, inside the block's tile."""
        ny, nx = self.by + (o // 3 - 1) * d, self.bx + (o % 3 - 1) * d
        ok = (ny >= 0) & (ny < self.nby) & (nx >= 0) & (nx < self.nbx)
        if tile:
            ok &= (ny // TILE[0] == self.by // TILE[0]) & (nx // TILE[1] == self.bx // TILE[1])
        return ok, np.where(ok, ny * self.nbx + nx, 0)

    def tile_id(self):
        return (self.by // TILE[0]) * (-(-self.nbx // TILE[1])) + self.bx // TILE[1]

    def first_min(self, cands):
        """First minimum (strict <, in list order) over [(mask, mv [n][2])]:
        (sad, mv); sad = 2**28 - 1 where no candidate is valid."""
        best = np.full(self.n, 2 ** 28 - 1, np.int64)
        bmv = np.zeros((self.n, 2), np.int64)
        for ok, mv in cands:
            s = np.where(ok, self.sad(mv), 2 ** 28 - 1)
            m = s < best
            best, bmv = np.where(m, s, best), np.where(m[:, None], mv, bmv)
        return best, bmv


def _descend(B, c, active, R, iters):
    """Iterated +-2 descent of the active blocks from c: (mv, sad, unconv)."""
    c = c.copy()
    sad = np.zeros(B.n, np.int64)
    unconv = np.ones(B.n, bool)
    act = active.copy()
    offs = [(i // 5 - 2, i % 5 - 2) for i in range(25)]
    for _ in range(iters):
        if not act.any():
            break
        best, nmv = B.first_min([(np.ones(B.n, bool), np.clip(c + np.array(o), -R, R)) for o in offs])
        moved = (np.abs(nmv - c) == 2).any(axis=1)
        c = np.where(act[:, None], nmv, c)
        sad = np.where(act, best, sad)
        unconv = np.where(act & ~moved, False, unconv)
        act = act & moved
    return c, sad, unconv


def _rounds(B, mv, sad, unconv, R, iters, n, jumps, tile, all_rounds=False):
    """Up to n propagation rounds (Jacobi: every block reads the states from
    before the round).  Round p tries the 8 neighbours at distance 1, then at
    each distance of jumps[p % len] (inside the tile when tile is set); a
    block descends from the first minimum when it beats its own SAD.  With
    tile set each tile stops after its first round without a change, else
    all stop together; all_rounds runs all n (umi_me_refine_fused).
    Returns (mv, sad, unconv, changed flags)."""
    tid = B.tile_id() if tile else np.zeros(B.n, np.int64)
    active = np.ones(tid.max() + 1, bool)
    changed = []
    for p in range(n):
        if all_rounds:
            active[:] = True
        if not active.any():
            break
        cands = []
        for d in (1,) + tuple(jumps[p % len(jumps)]):
            for o in RING:
                ok, nb = B.neighbour(o, d, tile)
                cands.append((ok, mv[nb]))
        best, bmv = B.first_min(cands)
        imp = (best < sad) & active[tid]
        m2, s2, u2 = _descend(B, bmv, imp, R, iters)
        mv = np.where(imp[:, None], m2, mv)
        sad = np.where(imp, s2, sad)
        unconv = np.where(imp, u2, unconv)
        hit = np.zeros_like(active)
        hit[tid[imp]] = True
        active &= hit
        changed.append(bool(imp.any()))
    return mv, sad, unconv, changed


def full_search_block(cur, ref, by, bx, R):
    """(dy, dx, SAD) of block (by, bx) over [-R, R]^2, first minimum in raster
    order, ref coordinates clamped (oracle_me_full_search's rule)."""
    H, W = cur.shape
    ys = np.clip(np.arange(by * 8 - R, by * 8 + 8 + R), 0, H - 1)
    xs = np.clip(np.arange(bx * 8 - R, bx * 8 + 8 + R), 0, W - 1)
    win = ref[np.ix_(ys, xs)].astype(np.int64)
    c = cur[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8].astype(np.int64)
    S = 2 * R + 1
    v = np.lib.stride_tricks.sliding_window_view(win, (8, 8))[:S, :S]
    sads = np.abs(v - c).sum((2, 3))
    i = int(np.argmin(sads))  # first minimum in raster order
    return i // S - R, i % S - R, int(sads.flat[i])


def refine(cur, ref, prop, R, iters, local_rounds, rounds, slack, all_rounds=False):
    """Refinement of block proposals restating umi_me_refine: seed, descent,
    in-tile rounds, global rounds, fallback check, full search of the
    fallback blocks.  all_rounds: every global round runs, as in
    umi_me_refine_fused.  Returns (mv_sad [n][3], fallback set, changed flags
    of the global rounds)."""
    B = _Blocks(cur, ref)
    prop = np.clip(np.asarray(prop, np.int64).reshape(B.n, 2), -R, R)
    seeds = [(np.ones(B.n, bool), prop)]
    for o in RING:
        ok, nb = B.neighbour(o, 1)
        seeds.append((ok, prop[nb]))
    _, c = B.first_min(seeds)
    mv, sad, unconv = _descend(B, c, np.ones(B.n, bool), R, iters)
    mv, sad, unconv, _ = _rounds(B, mv, sad, unconv, R, iters, local_rounds, LOCAL_JUMPS, True)
    mv, sad, unconv, changed = _rounds(B, mv, sad, unconv, R, iters, rounds, GLOBAL_JUMPS, False, all_rounds)
    changed += [False] * (rounds - len(changed))
    nbmin = np.full(B.n, 2 ** 32 - 1, np.int64)
    for o in RING:
        ok, nb = B.neighbour(o, 1)
        nbmin = np.minimum(nbmin, np.where(ok, B.sad(mv[nb]), 2 ** 32 - 1))
    worse = sad.astype(np.float32) > np.float32(slack) * np.minimum(sad, nbmin).astype(np.float32)
    fb = set(np.nonzero(unconv | worse)[0].tolist())
    out = np.concatenate([mv, sad[:, None]], axis=1)
    for b in fb:
        out[b] = full_search_block(cur, ref, b // B.nbx, b % B.nbx, R)
    return out, fb, changed
