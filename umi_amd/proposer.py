"""Learned motion proposals for full-search block motion estimation.

A 4-layer 3x3 convolution network (2 -> 32 -> 32 -> 32 -> 2 channels,
dilations 1, 2, 4, 8, ReLU between layers; 19,682 parameters) reads the frame
pair at quarter resolution and predicts the motion field.  The mean of each
8x8 block's 2x2 quarter-resolution outputs, times 4, is the block's proposal.
Exact SAD refinement (umi_me_refine) starts from the best of the block's
own and its 8 neighbours' proposals, descends through +-2 windows, then
runs jump-flood propagation rounds, first inside 4 x 8-block tiles, then
over the frame, in which blocks adopt better neighbour vectors.  A block falls back to the full +-R search when its descent never
settles, or when its SAD exceeds `slack` (1.1) x the smallest SAD it has at
its 8 neighbours' final vectors.

Training runs entirely on the GPU with this project's kernels (fp32 forward,
backward and Adam).  The labels are the plain full search's motion vectors.
Inference runs all four layers on bf16 MFMA (weights, inputs and activations
rounded to bf16, fp32 accumulation).
"""
from __future__ import annotations

import ctypes as C
import struct
from pathlib import Path

import numpy as np

from ._lib import UMI_ERR_UNSUPPORTED, UmiError, call
from .device import DeviceBuffer, Stream, Timer
from . import ops

LAYERS = [(2, 32, 1, True), (32, 32, 2, True), (32, 32, 4, True), (32, 2, 8, False)]  # (cin, cout, dilation, relu)
MAGIC = b"UMIPROP1"
CHECKPOINT = Path(__file__).resolve().parent / "data" / "proposer_1080p_r32.bin"


def _s(st):
    return None if st is None else st.handle


def n_params() -> int:
    return sum(co * ci * 9 + co for ci, co, _, _ in LAYERS)


# ---------------------------------------------------------------- synthesis
def bump_params(seed: int, H: int, W: int) -> np.ndarray:
    """4 Gaussian-weighted affine motion bumps: (cy, cx, 1/(2 s^2), ty, tx, a00, a01, a10, a11)."""
    g = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros((4, 9), np.float32)
    for k in range(4):
        s = g.uniform(200.0, 700.0)
        p[k] = [g.uniform(0, H), g.uniform(0, W), 1.0 / (2 * s * s), g.uniform(-24, 24), g.uniform(-24, 24),
                *g.uniform(-0.03, 0.03, size=4)]
    return p


class PairMaker:
    """Synthetic 1920x1080 pairs on the GPU: ref = smoothed noise, cur = ref
    warped by a smooth field (|mv| <= 32) plus N(0, sigma^2) noise."""

    def __init__(self, H=1080, W=1920, sigma=2.0):
        self.H, self.W, self.sigma = H, W, sigma
        self.bumps = DeviceBuffer(4 * 9 * 4)

    def make(self, seed: int, ref: DeviceBuffer, cur: DeviceBuffer, stream=None):
        call("umi_synth_frame0", ref.ptr, self.H, self.W, self.W, C.c_uint64(2 * seed + 1), _s(stream))
        self.bumps.upload(bump_params(seed, self.H, self.W))
        call("umi_synth_warp", ref.ptr, cur.ptr, self.H, self.W, self.W, self.bumps.ptr, C.c_float(self.sigma),
             C.c_uint64(2 * seed + 2), _s(stream))


# ---------------------------------------------------------------- network
class Network:
    """fp32 master parameters (device), Adam state, forward/backward buffers."""

    def __init__(self, seed: int = 0):
        g = np.random.Generator(np.random.PCG64(seed))
        self.host = []
        for ci, co, _, _ in LAYERS:
            w = (g.standard_normal((co, ci, 3, 3)) * np.sqrt(2.0 / (ci * 9))).astype(np.float32)
            self.host.append((w, np.zeros(co, np.float32)))
        self._upload()
        self.step = 0

    def _upload(self):
        self.w = [DeviceBuffer.from_array(w) for w, _ in self.host]
        self.b = [DeviceBuffer.from_array(b) for _, b in self.host]
        self.m = [(DeviceBuffer.zeros(w.nbytes), DeviceBuffer.zeros(b.nbytes)) for w, b in self.host]
        self.v = [(DeviceBuffer.zeros(w.nbytes), DeviceBuffer.zeros(b.nbytes)) for w, b in self.host]
        # gradient buffers: weights then bias, contiguous (the backward kernel's output layout)
        self.gw = [DeviceBuffer(w.nbytes + b.nbytes) for w, b in self.host]

    def params(self):
        return [(w.download(hw.shape, np.float32), b.download(hb.shape, np.float32))
                for (w, b), (hw, hb) in zip(zip(self.w, self.b), self.host)]

    # -- checkpoint: magic, version, step, per layer (cin, cout, dil, relu) + w + b
    def save(self, path):
        with open(path, "wb") as f:
            f.write(MAGIC)
            f.write(struct.pack("<iii", 1, self.step, len(LAYERS)))
            for (ci, co, d, r), (w, b) in zip(LAYERS, self.params()):
                f.write(struct.pack("<iiii", ci, co, d, int(r)))
                f.write(w.astype("<f4").tobytes())
                f.write(b.astype("<f4").tobytes())

    @classmethod
    def load(cls, path):
        net = cls.__new__(cls)
        data = Path(path).read_bytes()
        if data[:8] != MAGIC:
            raise ValueError(f"{path}: not a proposer checkpoint")
        ver, step, nl = struct.unpack_from("<iii", data, 8)
        if ver != 1 or nl != len(LAYERS):
            raise ValueError(f"{path}: version {ver}, {nl} layers")
        off = 20
        net.host = []
        for ci, co, d, r in LAYERS:
            hdr = struct.unpack_from("<iiii", data, off)
            if hdr != (ci, co, d, int(r)):
                raise ValueError(f"{path}: layer {hdr} != {(ci, co, d, r)}")
            off += 16
            w = np.frombuffer(data, "<f4", co * ci * 9, off).reshape(co, ci, 3, 3).astype(np.float32)
            off += w.nbytes
            b = np.frombuffer(data, "<f4", co, off).astype(np.float32)
            off += b.nbytes
            net.host.append((w, b))
        net._upload()
        net.step = step
        return net


class Trainer:
    """Forward + backward + Adam on batches of N quarter-resolution pairs."""

    def __init__(self, net: Network, N: int, H4: int, W4: int):
        self.net, self.N, self.H4, self.W4 = net, N, H4, W4
        px = N * H4 * W4
        self.act = [DeviceBuffer(px * co * 4) for _, co, _, _ in LAYERS]
        self.grad = [DeviceBuffer(px * co * 4) for _, co, _, _ in LAYERS]
        self.loss_parts = DeviceBuffer(((px + 255) // 256) * 4)
        ws = 0
        for ci, co, _, _ in LAYERS:
            n = C.c_size_t()
            call("umi_cnn_conv_bwd_weight_workspace", N, ci, co, H4, W4, C.byref(n))
            ws = max(ws, n.value)
        self.ws = DeviceBuffer(ws)

    def forward(self, x: DeviceBuffer, stream=None):
        src = x
        for (ci, co, d, relu), w, b, a in zip(LAYERS, self.net.w, self.net.b, self.act):
            call("umi_cnn_conv_fwd", src.ptr, self.N, ci, co, self.H4, self.W4, d, w.ptr, b.ptr, int(relu), a.ptr,
                 _s(stream))
            src = a
        return self.act[-1]

    def backward(self, x: DeviceBuffer, labels: DeviceBuffer, stream=None) -> None:
        """Writes the loss partial sums and every layer's (dW, db) into net.gw."""
        nparts = C.c_int()
        call("umi_cnn_huber", self.act[-1].ptr, labels.ptr, self.N, self.H4, self.W4, self.grad[-1].ptr,
             self.loss_parts.ptr, C.byref(nparts), _s(stream))
        self.nparts = nparts.value
        for li in range(len(LAYERS) - 1, -1, -1):
            ci, co, d, relu = LAYERS[li]
            inp = x if li == 0 else self.act[li - 1]
            call("umi_cnn_conv_bwd_weight", self.grad[li].ptr, self.act[li].ptr, inp.ptr, self.N, ci, co, self.H4,
                 self.W4, d, int(relu), self.net.gw[li].ptr, self.ws.ptr, self.ws.nbytes, _s(stream))
            if li > 0:
                call("umi_cnn_conv_bwd_data", self.grad[li].ptr, self.act[li].ptr, self.N, ci, co, self.H4, self.W4,
                     d, self.net.w[li].ptr, int(relu), self.grad[li - 1].ptr, _s(stream))

    def adam(self, lr=1e-3, stream=None):
        self.net.step += 1
        for li, (ci, co, _, _) in enumerate(LAYERS):
            nw = co * ci * 9
            gw = self.net.gw[li]
            call("umi_adam", self.net.w[li].ptr, gw.ptr, self.net.m[li][0].ptr, self.net.v[li][0].ptr, nw,
                 C.c_float(lr), C.c_float(0.9), C.c_float(0.999), C.c_float(1e-8), self.net.step, _s(stream))
            call("umi_adam", self.net.b[li].ptr, gw.ptr + 4 * nw, self.net.m[li][1].ptr, self.net.v[li][1].ptr, co,
                 C.c_float(lr), C.c_float(0.9), C.c_float(0.999), C.c_float(1e-8), self.net.step, _s(stream))

    def loss(self) -> float:
        parts = self.loss_parts.download((self.nparts,), np.float32)
        return float(parts.astype(np.float64).sum() / (2 * self.N * self.H4 * self.W4))


class Dataset:
    """Quarter-resolution inputs and full-search labels of `n` seeded pairs, on the GPU."""

    def __init__(self, n: int, seed0: int, H=1080, W=1920, R=32, stream=None):
        self.n, self.H, self.W, self.R = n, H, W, R
        self.H4, self.W4 = H // 4, W // 4
        self.nby, self.nbx = H // 8, W // 8
        self.x = DeviceBuffer(n * 2 * self.H4 * self.W4 * 4)
        self.labels = DeviceBuffer(n * self.nby * self.nbx * 3 * 4)
        pm = PairMaker(H, W)
        ref, cur = DeviceBuffer(H * W), DeviceBuffer(H * W)
        plane4 = 2 * self.H4 * self.W4 * 4
        lab = self.nby * self.nbx * 3 * 4
        for i in range(n):
            pm.make(seed0 + i, ref, cur, stream)
            call("umi_cnn_downsample_pair", cur.ptr, ref.ptr, H, W, W, self.x.ptr + i * plane4, 0, _s(stream))
            call("umi_me_full_search_u8", cur.ptr, ref.ptr, H, W, W, 8, R, self.labels.ptr + i * lab, None, _s(stream))

    def batch(self, idx):
        """Device pointers of a contiguous batch [idx, idx + N)."""
        idx = int(idx)
        return (self.x.ptr + idx * 2 * self.H4 * self.W4 * 4, self.labels.ptr + idx * self.nby * self.nbx * 12)


class _View:
    """A device pointer with the DeviceBuffer interface the kernels' wrappers use."""

    def __init__(self, ptr, nbytes):
        self.ptr, self.nbytes = int(ptr), int(nbytes)


def learning_rate(lr: float, ep: int, epochs: int, schedule: str) -> float:
    """'step': halve every epochs / 3; 'cosine': lr/2 (1 + cos(pi ep / epochs))."""
    if schedule == "cosine":
        return lr * 0.5 * (1.0 + np.cos(np.pi * ep / epochs))
    return lr * (0.5 ** (ep // max(1, epochs // 3)))


def train(net: Network, data: Dataset, epochs: int, batch: int = 8, lr: float = 2e-3, log=print, stream=None,
          schedule: str = "step", on_epoch=None):
    """Adam over shuffled batches; on_epoch(ep, net) runs after each epoch."""
    tr = Trainer(net, batch, data.H4, data.W4)
    g = np.random.Generator(np.random.PCG64(1234))
    history = []
    nb = data.n // batch
    for ep in range(epochs):
        order = g.permutation(nb)
        tot = 0.0
        for k, bi in enumerate(order):
            xp, lp = data.batch(bi * batch)
            x = _View(xp, 0)
            tr.forward(x, stream)
            tr.backward(x, _View(lp, 0), stream)
            tr.adam(learning_rate(lr, ep, epochs, schedule), stream)
            tot += tr.loss()
        history.append(tot / nb)
        log(f"[proposer] epoch {ep}: mean Huber loss {history[-1]:.4f}")
        if on_epoch is not None:
            on_epoch(ep, net)
    return history


# ---------------------------------------------------------------- inference
def layer_kind(li: int) -> int:
    """umi_cnn_infer_layer kind of layer li: 0 first, 2 last, 1 between."""
    return 0 if li == 0 else (2 if li == len(LAYERS) - 1 else 1)


def pack_layer(net: Network, li: int, stream=None) -> DeviceBuffer:
    """Layer li's weights rounded to bf16 in the MFMA operand order."""
    n = C.c_size_t()
    call("umi_cnn_packed_bytes", layer_kind(li), C.byref(n))
    buf = DeviceBuffer(n.value)
    call("umi_cnn_pack_layer", layer_kind(li), net.w[li].ptr, buf.ptr, _s(stream))
    return buf


class Proposer:
    """Proposal + refinement + fallback for 1920x1080 pairs with a trained network."""

    def __init__(self, net: Network, H=1080, W=1920, R=32, slack=1.1, iters=4, local_rounds=4, rounds=4,
                 fused=None):
        """fused: refine in one launch (umi_me_refine_fused, every global round
        run) -- the default where it applies (at most 1024 tiles of 4 x 8
        blocks and 32 rounds).  It needs the GPU's workgroup slots to itself,
        so pipelines that run concurrently on several streams take fused=False
        (umi_me_refine: one launch per stage, stopping after an idle round)."""
        self.net, self.H, self.W, self.R, self.slack = net, H, W, R, slack
        self.iters, self.local_rounds, self.rounds = iters, local_rounds, rounds
        ntiles = -(-(W // 8) // 8) * -(-(H // 8) // 4)
        self._auto = fused is None  # an automatic choice falls back if the device refuses the one-launch form
        self.fused = (ntiles <= 1024 and rounds <= 32) if fused is None else fused
        self.H4, self.W4 = H // 4, W // 4
        px = self.H4 * self.W4
        self.nblk = (H // 8) * (W // 8)
        self.x = DeviceBuffer(2 * px * 4)
        self.a = [DeviceBuffer(px * 32 * 2) for _ in range(3)]
        self.flow = DeviceBuffer(2 * px * 4)
        self.prop = DeviceBuffer(self.nblk * 8)
        self.fb_list = DeviceBuffer(self.nblk * 4)
        self.fb_count = DeviceBuffer(4)
        self.packed = [pack_layer(net, li) for li in range(len(LAYERS))]
        n = C.c_size_t()
        if self.fused:
            call("umi_me_refine_fused_workspace", H, W, rounds, C.byref(n))
            self.ws = DeviceBuffer.zeros(n.value)  # zeroed once; every call leaves it consistent
        else:
            call("umi_me_refine_workspace", H, W, rounds, C.byref(n))
            self.ws = DeviceBuffer(n.value)

    def propose(self, cur: DeviceBuffer, ref: DeviceBuffer, stream=None):
        """Network and block proposals: the first two layers read the frames
        in one launch, the last one writes the proposals (3 launches)."""
        s = _s(stream)
        call("umi_cnn_first_two_layers_frames", cur.ptr, ref.ptr, self.H, self.W, self.W, self.packed[0].ptr,
             self.net.b[0].ptr, LAYERS[1][2], self.packed[1].ptr, self.net.b[1].ptr, self.a[1].ptr, s)
        call("umi_cnn_infer_layer", 1, self.a[1].ptr, 1, self.H4, self.W4, LAYERS[2][2], self.packed[2].ptr,
             self.net.b[2].ptr, self.a[2].ptr, s)
        call("umi_cnn_last_layer_proposals", self.a[2].ptr, self.H4, self.W4, LAYERS[3][2], self.packed[3].ptr,
             self.net.b[3].ptr, self.R, self.flow.ptr, self.prop.ptr, s)

    def refine_only(self, cur, ref, mv_sad: DeviceBuffer, stream=None):
        args = (cur.ptr, ref.ptr, self.H, self.W, self.W, self.R, self.prop.ptr, self.iters, self.local_rounds,
                self.rounds, C.c_float(self.slack), mv_sad.ptr, self.fb_list.ptr, self.fb_count.ptr)
        if self.fused:
            try:
                call("umi_me_refine_fused", *args, self.ws.ptr, self.ws.nbytes, _s(stream))
                return
            except UmiError as e:
                # the device cannot hold the whole grid at once: one launch per stage instead
                if not (self._auto and e.status == UMI_ERR_UNSUPPORTED):
                    raise
                self.fused = False
                n = C.c_size_t()
                call("umi_me_refine_workspace", self.H, self.W, self.rounds, C.byref(n))
                self.ws = DeviceBuffer(n.value)
        call("umi_me_refine", *args, self.ws.ptr, self.ws.nbytes, _s(stream))

    def refine(self, cur, ref, mv_sad: DeviceBuffer, stream=None):
        """Refinement; the fallback blocks come back with the full search's result."""
        self.refine_only(cur, ref, mv_sad, stream)

    def run(self, cur, ref, mv_sad, stream=None, proposals=True):
        """proposals=False: the same refinement and fallback around zero motion
        (the method without the network)."""
        if proposals:
            self.propose(cur, ref, stream)
        else:
            self.prop.fill(0, stream)
        self.refine(cur, ref, mv_sad, stream)

    def fallback_count(self) -> int:
        return int(self.fb_count.download((1,), np.int32)[0])

    def status(self):
        """(change flag of each global round of the last refinement, whether a
        wait of the fused refinement ran out); synchronises the device."""
        if not self.fused:
            nb = self.nblk * 16
            off = 2 * ((nb + 255) // 256 * 256)
            raw = self.ws.download((self.ws.nbytes,), np.uint8)[off:off + 4 * self.rounds].tobytes()
            return [bool(v) for v in np.frombuffer(raw, np.int32)], False
        changed = (C.c_int32 * max(self.rounds, 1))()
        late = C.c_int()
        call("umi_me_refine_fused_status", self.ws.ptr, self.rounds, C.cast(changed, C.c_void_p), C.byref(late))
        return [bool(changed[p]) for p in range(self.rounds)], bool(late.value)

    def rounds_run(self) -> int:
        """Propagation rounds of the last refinement that changed some block."""
        return sum(self.status()[0])


def evaluate(prop: Proposer, pairs: int, seed0: int, proposals=True, reps: int = 10, graphs: bool = False):
    """Fraction of blocks at the full search's minimum SAD, fallback fraction,
    and mean times (ms) of the proposal pipeline (with its stages) and of the
    plain full search, launched eagerly; with `graphs` also both as one graph
    replay per pair (`ms_pipeline_graph`, `ms_full_search_graph`,
    `speedup_graph`)."""
    H, W, R = prop.H, prop.W, prop.R
    pm = PairMaker(H, W)
    ref, cur = DeviceBuffer(H * W), DeviceBuffer(H * W)
    mv, judge = DeviceBuffer(prop.nblk * 12), DeviceBuffer(prop.nblk * 12)
    st = Stream()
    at_min, fb, rounds = [], [], []
    times = {"pipeline": [], "network": [], "refine": [], "full_search": [], "pipeline_graph": [],
             "full_search_graph": []}

    def timed(fn):
        fn()
        st.synchronize()
        t = Timer()
        t.start(st)
        for _ in range(reps):
            fn()
        t.stop(st)
        return t.ms() / reps

    for i in range(pairs):
        pm.make(seed0 + i, ref, cur, st)
        ops.me_full_search(cur, ref, H, W, W, 8, R, judge, stream=st)
        prop.run(cur, ref, mv, st, proposals)
        st.synchronize()
        a = mv.download((prop.nblk, 3), np.int32)
        j = judge.download((prop.nblk, 3), np.int32)
        at_min.append(float(np.mean(a[:, 2] == j[:, 2])))
        fb.append(prop.fallback_count() / prop.nblk)
        flags, late = prop.status()
        if late:
            raise RuntimeError("proposer: a wait of the fused refinement ran out")
        rounds.append(sum(flags))
        times["pipeline"].append(timed(lambda: prop.run(cur, ref, mv, st, proposals)))
        if proposals:
            times["network"].append(timed(lambda: prop.propose(cur, ref, st)))
        times["refine"].append(timed(lambda: prop.refine_only(cur, ref, mv, st)))
        times["full_search"].append(timed(lambda: ops.me_full_search(cur, ref, H, W, W, 8, R, judge, stream=st)))
        if graphs:
            # the same two, each captured once as a graph and replayed: three
            # replays (with eager work between) must give the eager vectors
            from .device import Graph
            gp, gf = Graph(st), Graph(st)
            gp.begin()
            prop.run(cur, ref, mv, st, proposals)
            gp.end()
            gf.begin()
            ops.me_full_search(cur, ref, H, W, W, 8, R, judge, stream=st)
            gf.end()
            for _ in range(3):
                mv.fill(0, st)
                gp.launch()
                st.synchronize()
                if not np.array_equal(mv.download((prop.nblk, 3), np.int32), a):
                    raise RuntimeError("proposer: graph replay differs from the eager pipeline")
            times["pipeline_graph"].append(timed(gp.launch))
            times["full_search_graph"].append(timed(gf.launch))
    out = {"refinement": "one launch (umi_me_refine_fused)" if prop.fused else "one launch per stage (umi_me_refine)",
           "at_min_sad": float(np.mean(at_min)), "min_at_min_sad": float(np.min(at_min)),
           "fallback_frac": float(np.mean(fb)), "rounds_changed": float(np.mean(rounds))}
    out.update({f"ms_{k}": float(np.mean(v)) for k, v in times.items() if v})
    out["speedup"] = out["ms_full_search"] / out["ms_pipeline"]
    if graphs:
        out["speedup_graph"] = out["ms_full_search_graph"] / out["ms_pipeline_graph"]
    return out


def throughput(net: "Network", pairs: int, seed0: int, reps: int = 20, R: int = 32, gate: int = 80) -> dict:
    """Blocks per second with `pairs` independent pairs in flight at once (as
    consecutive pairs of a video would be), one stream and one graph-captured
    pipeline per pair, against the plain full search given the same streams
    and graphs.  The first replay of every pipeline graph must reproduce its
    eager result bit for bit.

    GPU-timed: stream 0 first runs `gate` full searches while the host
    enqueues `reps` rounds of one graph launch per stream behind a fork from
    stream 0; an event pair on stream 0 spans the fork to the join.  The host
    enqueue time is reported with the gate's length: the measurement holds
    while the host finishes first (`host_ahead`)."""
    import time
    from .device import Graph, synchronize

    H, W = 1080, 1920
    pm = PairMaker(H, W)
    refs = [DeviceBuffer(H * W) for _ in range(pairs)]
    curs = [DeviceBuffer(H * W) for _ in range(pairs)]
    streams = [Stream() for _ in range(pairs)]
    props = [Proposer(net, H, W, R=R, fused=False) for _ in range(pairs)]  # pipelines run concurrently
    nblk = props[0].nblk
    mvs = [DeviceBuffer(nblk * 12) for _ in range(pairs)]
    judges = [DeviceBuffer(nblk * 12) for _ in range(pairs)]
    gate_out = DeviceBuffer(nblk * 12)
    at_min, gp, gf = [], [], []
    for i in range(pairs):
        pm.make(seed0 + i, refs[i], curs[i], streams[i])
        ops.me_full_search(curs[i], refs[i], H, W, W, 8, R, judges[i], stream=streams[i])
        props[i].run(curs[i], refs[i], mvs[i], streams[i])
        streams[i].synchronize()
        eager = mvs[i].download((nblk, 3), np.int32)
        at_min.append(float(np.mean(eager[:, 2] == judges[i].download((nblk, 3), np.int32)[:, 2])))
        g = Graph(streams[i])
        g.begin()
        props[i].run(curs[i], refs[i], mvs[i], streams[i])
        g.end()
        mvs[i].fill(0, streams[i])
        g.launch()
        streams[i].synchronize()
        if not np.array_equal(mvs[i].download((nblk, 3), np.int32), eager):
            raise RuntimeError("proposer: graph replay differs from the eager pipeline")
        gp.append(g)
        g = Graph(streams[i])
        g.begin()
        ops.me_full_search(curs[i], refs[i], H, W, W, 8, R, judges[i], stream=streams[i])
        g.end()
        gf.append(g)
    s0 = streams[0]

    def run_gate():
        for _ in range(gate):
            ops.me_full_search(curs[0], refs[0], H, W, W, 8, R, gate_out, stream=s0)

    t = Timer()
    t.start(s0)
    run_gate()
    t.stop(s0)
    gate_ms = t.ms()

    def timed(graphs):
        for _ in range(2):
            for g in graphs:
                g.launch()
        synchronize()
        run_gate()
        tm = Timer()
        tm.start(s0)
        for st in streams[1:]:
            st.wait(s0)
        h0 = time.perf_counter()
        for _ in range(reps):
            for g in graphs:
                g.launch()
        host_ms = (time.perf_counter() - h0) * 1e3
        for st in streams[1:]:
            s0.wait(st)
        tm.stop(s0)
        return tm.ms() / (reps * len(graphs)), host_ms

    (tp, hp), (tf, hf) = timed(gp), timed(gf)
    return {"streams": pairs, "first_seed": seed0, "at_min_sad": float(np.mean(at_min)),
            "min_at_min_sad": float(np.min(at_min)), "us_per_pair_pipeline": tp * 1e3,
            "us_per_pair_full_search": tf * 1e3, "blocks_per_s_pipeline": nblk / (tp * 1e-3),
            "blocks_per_s_full_search": nblk / (tf * 1e-3), "speedup": tf / tp, "gate_ms": gate_ms,
            "host_enqueue_ms": max(hp, hf), "host_ahead": max(hp, hf) < gate_ms}


def main(argv=None):
    """python -m umi_amd.proposer train --pairs 2000 --epochs 12 --out ckpt.bin
       python -m umi_amd.proposer eval --ckpt ckpt.bin --pairs 4"""
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(prog="python -m umi_amd.proposer")
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--schedule", choices=["step", "cosine"], default="step")
    ap.add_argument("--seed0", type=int, default=0, help="first training pair seed")
    ap.add_argument("--eval-seed0", type=int, default=1_000_000, help="first held-out pair seed")
    ap.add_argument("--ckpt", default=str(CHECKPOINT))
    ap.add_argument("--rounds", type=int, default=4, help="global propagation rounds of the refinement")
    ap.add_argument("--local-rounds", type=int, default=4, help="in-tile propagation rounds of the refinement")
    ap.add_argument("--iters", type=int, default=4, help="descent steps per block of the refinement")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        t0 = time.time()
        data = Dataset(a.pairs, a.seed0)
        print(f"[proposer] {a.pairs} pairs synthesised and labelled in {time.time() - t0:.1f} s", flush=True)
        net = Network(seed=1)
        t0 = time.time()
        out = a.out or a.ckpt
        Path(out).parent.mkdir(parents=True, exist_ok=True)

        def checkpoint(ep, n):  # every 10 epochs, so that an interrupted run keeps its progress
            if (ep + 1) % 10 == 0:
                n.save(out)

        hist = train(net, data, a.epochs, a.batch, a.lr, log=lambda m: print(m, flush=True), schedule=a.schedule,
                     on_epoch=checkpoint)
        print(f"[proposer] trained {a.epochs} epochs in {time.time() - t0:.1f} s", flush=True)
        net.save(out)
        print(json.dumps({"loss_history": hist, "checkpoint": out}), flush=True)
    else:
        net = Network.load(a.ckpt)
        prop = Proposer(net, local_rounds=a.local_rounds, rounds=a.rounds, iters=a.iters)
        with_net = evaluate(prop, a.pairs, a.eval_seed0, proposals=True)
        without = evaluate(prop, a.pairs, a.eval_seed0, proposals=False)
        long = Proposer(net, rounds=32)  # the same refinement with a larger budget of global rounds
        without_32 = evaluate(long, a.pairs, a.eval_seed0, proposals=False)
        print(json.dumps({"with_network": with_net, "without_network": without,
                          "without_network_32_rounds": without_32}), flush=True)


if __name__ == "__main__":
    main()
