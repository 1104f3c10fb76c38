"""Batches of images of uneven sizes: one persistent launch per operator.

``ImageBatch`` packs a list of images into one device buffer (each image at
its own element offset, pitch = width) and creates the library's batch
object (the per-image work tables live in device memory), so every batched
operator is a single launch that allocates nothing and can be captured in a
graph.  ``autotune`` picks the fastest tile configuration of an operator by
timing graph replays, and caches the choice on disk.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
from pathlib import Path

import numpy as np

from ._lib import call, lib
from .device import DeviceBuffer, Graph, Stream, Timer, device_name

OP_FILTER, OP_BILATERAL, OP_IIMG = 0, 1, 2


def _stream(s):
    return None if s is None else s.handle


def n_configs(op: int) -> int:
    n = C.c_int(0)
    call("umi_batch_configs", op, C.byref(n))
    return n.value


class ImageBatch:
    """n images (2-D numpy arrays of one dtype) resident in one device buffer."""

    def __init__(self, images, out_dtype=None):
        self.shapes = [tuple(im.shape) for im in images]
        self.dtype = np.dtype(images[0].dtype)
        self.out_dtype = np.dtype(out_dtype or self.dtype)
        sizes = [h * w for h, w in self.shapes]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.total = int(sum(sizes))
        flat = np.concatenate([np.ascontiguousarray(im, self.dtype).ravel() for im in images])
        self.inp = DeviceBuffer.from_array(flat)
        self.out = DeviceBuffer(self.total * self.out_dtype.itemsize)
        n = len(images)
        H = np.array([s[0] for s in self.shapes], np.int32)
        W = np.array([s[1] for s in self.shapes], np.int32)
        h = C.c_void_p()
        call("umi_batch_create", n, H.ctypes.data, W.ctypes.data, self.offsets.ctypes.data, W.ctypes.data,
             self.offsets.ctypes.data, W.ctypes.data, C.byref(h))
        self.handle = h.value
        self.signature = hashlib.sha1(H.tobytes() + W.tobytes()).hexdigest()[:16]

    def __del__(self):
        try:
            if self.handle:
                lib().umi_batch_destroy(self.handle)
        except Exception:
            pass

    def outputs(self, dtype=None):
        dt = np.dtype(dtype or self.out_dtype)
        flat = self.out.download((self.total,), dt)
        return [flat[o:o + h * w].reshape(h, w) for o, (h, w) in zip(self.offsets, self.shapes)]

    # ---- operators (out written to self.out) ----
    def wfilter3x3(self, weights: DeviceBuffer, init=0.0, pad=0.0, mode=0, config=-1, stream=None, out=None):
        call("umi_batch_wfilter3x3_f32", self.handle, self.inp.ptr, weights.ptr, init, pad, mode,
             (out or self.out).ptr, config, _stream(stream))

    def bilateral(self, radius: int, spatial: DeviceBuffer, norm: float, pad=0.0, mode=0, config=-1, stream=None,
                  out=None):
        call("umi_batch_bilateral_f32", self.handle, self.inp.ptr, radius, spatial.ptr, norm, pad, mode,
             (out or self.out).ptr, config, _stream(stream))

    def iimg_workspace(self) -> DeviceBuffer:
        """The integral image's workspace, zero-filled (as it must be before its first use)."""
        n = C.c_size_t(0)
        kind = 0 if self.dtype == np.uint8 else 2
        call("umi_batch_iimg_workspace_size", self.handle, kind, C.byref(n))
        return DeviceBuffer.zeros(max(n.value, 16))

    def iimg(self, workspace: DeviceBuffer, stream=None, out=None, config=-1):
        name = "umi_batch_iimg_u8_u32" if self.dtype == np.uint8 else "umi_batch_iimg_f32"
        call(name, self.handle, self.inp.ptr, (out or self.out).ptr, workspace.ptr, workspace.nbytes, config,
             _stream(stream))

    def box_sum(self, ry: int, rx: int, out: DeviceBuffer, stream=None, iimg: DeviceBuffer | None = None):
        """Box sums of every image from its uint32 integral image (`iimg`,
        default self.out, as `iimg` writes it) into `out` (same layout)."""
        call("umi_batch_box_sum_u32", self.handle, (iimg or self.out).ptr, ry, rx, out.ptr, _stream(stream))


def _cache_path() -> Path:
    p = os.environ.get("UMI_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "umi_amd" / "tune.json"


def time_graph(run, stream: Stream, replays: int = 20) -> float:
    """Capture run(stream) in a graph and return the mean ms of `replays` replays."""
    run(stream)  # eager warm-up
    g = Graph(stream)
    g.begin()
    run(stream)
    g.end()
    g.launch()
    stream.synchronize()
    t = Timer()
    t.start(stream)
    for _ in range(replays):
        g.launch()
    t.stop(stream)
    stream.synchronize()
    return t.ms() / replays


def autotune(op: int, batch: ImageBatch, run_config, replays: int = 20, use_cache: bool = True):
    """Returns (best config, {config: ms}).  run_config(config, stream) issues the op."""
    key = f"{device_name()}|op{op}|{batch.signature}"
    path = _cache_path()
    cache = {}
    if use_cache and path.exists():
        try:
            cache = json.loads(path.read_text())
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return cache[key]["best"], {int(k): v for k, v in cache[key]["ms"].items()}
    st = Stream()
    times = {c: time_graph(lambda s, c=c: run_config(c, s), st, replays) for c in range(n_configs(op))}
    best = min(times, key=times.get)
    if use_cache:
        cache[key] = {"best": best, "ms": times}
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            path.write_text(json.dumps(cache, indent=1))
        except OSError:
            pass
    return best, times
