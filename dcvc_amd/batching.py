"""Many concurrent video streams on one GPU (the batching form).

Each ``StreamContext`` is one stream at its own resolution: an HT-S encoder
and decoder proxy sharing one weight copy in HBM, with the stream's temporal
context resident on the device between chunks.  ``code_concurrently`` codes
one chunk of every stream at the same time: one host worker thread per
stream drives that stream's proxies (the C ABI releases the GIL), each on its
own HIP streams, so the kernels of different streams overlap on the GPU and
the CPU entropy coding of one stream overlaps the others' GPU work.
Results are bit-identical to coding each stream alone: every kernel is
batch- and tile-invariant and a stream's state is its own.

``BatchGroup`` goes further for streams of one resolution: one encoder and
one decoder proxy code the whole group as a batch (``DMCHTSBatchProxy``), so
every layer is one launch over the group's pictures, each stream keeping its
own QP, temporal context and bitstream.  ``code_groups_concurrently`` runs
the groups of a mix side by side, one worker thread per group.
"""
from __future__ import annotations

import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np
import torch

from . import codec, synthetic
from .inference_extensions_cuda import DMCHTSBatchProxy


@dataclass
class ChunkResult:
    stream: int
    qp: int
    bits: bytes
    frames: list               # 8 decoded frames (copies), [1, 3, Hp, Wp] fp16
    encode_s: float
    decode_s: float


class StreamContext:
    """One stream: encoder + decoder contexts sharing `shared`'s weights."""

    def __init__(self, sid: int, height: int, width: int, shared: codec.VideoCodecHTS, precision="fp16"):
        self.sid, self.H, self.W = sid, height, width
        self.enc = codec.VideoCodecHTS(share_from=shared)
        self.dec = codec.VideoCodecHTS(share_from=shared)
        self.enc.proxy.set_precision(precision)
        self.dec.proxy.set_precision(precision)
        self.pad_r, self.pad_b = synthetic.get_padding_size(height, width, 16)
        self.lock = threading.Lock()   # chunks of one stream are coded in order
        self.chunks_coded = 0

    def start(self, intra_enc: codec.IntraCodec, intra_dec: codec.IntraCodec, frame, qp=32):
        """Code the stream's I frame and seed both temporal contexts."""
        e = intra_enc.compress(frame, qp, self.pad_b, self.pad_r)
        self.enc.add_ref_feature_from_frame(e["x_hat"])
        d = intra_dec.decompress(e["bit_stream"], {"height": self.H, "width": self.W}, qp, e["ec_parallel"])
        self.dec.add_ref_feature_from_frame(d["x_hat"], apply_feature_adaptor=False)
        torch.cuda.synchronize()
        return len(e["bit_stream"])

    def code_chunk(self, x, qp) -> ChunkResult:
        """Encode then decode one 8-frame chunk [1, 24, H, W] on the calling
        thread's current torch stream."""
        with self.lock:
            t0 = time.perf_counter()
            bits, ec = self.enc.proxy.compress(x, qp, False, self.pad_b, self.pad_r)
            t1 = time.perf_counter()
            outs = self.dec.proxy.decompress(bits, qp, self.H, self.W, ec, False)
            frames = [o.clone() for o in outs]
            torch.cuda.current_stream().synchronize()
            t2 = time.perf_counter()
            self.chunks_coded += 1
            return ChunkResult(self.sid, qp, bits.tobytes(), frames, t1 - t0, t2 - t1)

    def device_bytes(self) -> int:
        """HBM held by this stream's two contexts (activation buffers + outputs)."""
        n = 0
        for p in (self.enc.proxy, self.dec.proxy):
            n += int(p.device_bytes())
        return n


_tls = threading.local()


def _worker_stream():
    """A private torch stream per worker thread, so that the proxies' entry /
    exit ordering never couples different host threads."""
    s = getattr(_tls, "stream", None)
    if s is None:
        s = torch.cuda.Stream()
        _tls.stream = s
    return s


def code_one(ctx: StreamContext, x, qp) -> ChunkResult:
    with torch.cuda.stream(_worker_stream()):
        return ctx.code_chunk(x, qp)


def code_sequentially(contexts, inputs, qps):
    """Reference schedule: one stream at a time."""
    return [code_one(c, x, q) for c, x, q in zip(contexts, inputs, qps)]


def code_concurrently(contexts, inputs, qps, workers=None):
    """All streams at once, one worker thread per stream (or `workers`)."""
    with ThreadPoolExecutor(max_workers=workers or len(contexts)) as pool:
        futs = [pool.submit(code_one, c, x, q) for c, x, q in zip(contexts, inputs, qps)]
        return [f.result() for f in futs]


MIX_16 = [(480, 854)] * 4 + [(720, 1280)] * 6 + [(1080, 1920)] * 4 + [(2160, 3840)] * 2


@dataclass
class BatchingSetup:
    contexts: list
    chunks: list               # per stream: list of chunk inputs on the GPU
    qps: list
    intra_bytes: list = field(default_factory=list)


def build_streams(sdi, sdp, resolutions=MIX_16, n_chunks=2, seed=0, precision="fp16", skip_thres=0.15,
                  device="cuda"):
    """Streams of the given resolutions with synthetic content (texture +
    integer motion + occlusion), one random qp in [0, 63] per stream."""
    rng = np.random.Generator(np.random.PCG64(seed))
    intra_e = codec.IntraCodec(sdi, skip_thres)
    intra_d = codec.IntraCodec(sdi, skip_thres)
    shared = codec.VideoCodecHTS(sdp, skip_thres)
    contexts, chunks, qps, ib = [], [], [], []
    for sid, (H, W) in enumerate(resolutions):
        frames = synthetic.motion_sequence_yuv420(H, W, 1 + 4 * n_chunks, seed=seed + 17 * sid)
        ctx = StreamContext(sid, H, W, shared, precision)
        ib.append(ctx.start(intra_e, intra_d, codec.to_input(synthetic.frame_to_input(frames[0]), device)))
        contexts.append(ctx)
        chunks.append([codec.to_input(synthetic.chunk_to_input(frames[1 + 4 * c:5 + 4 * c], 8), device)
                       for c in range(n_chunks)])
        qps.append(int(rng.integers(0, 64)))
    return BatchingSetup(contexts, chunks, qps, ib), shared


class BatchGroup:
    """Streams of one resolution coded as one batch: an encoder and a decoder
    DMCHTSBatchProxy sharing `shared`'s weights.  The streams advance together,
    one chunk each per call."""

    def __init__(self, sids, height: int, width: int, shared: codec.VideoCodecHTS, precision="fp16"):
        self.sids = list(sids)
        self.H, self.W = height, width
        n = len(self.sids)
        self.enc = DMCHTSBatchProxy(n)
        self.dec = DMCHTSBatchProxy(n)
        self.enc.share_weights_from(shared.proxy)
        self.dec.share_weights_from(shared.proxy)
        self.enc.set_precision(precision)
        self.dec.set_precision(precision)
        self.pad_r, self.pad_b = synthetic.get_padding_size(height, width, 16)
        self.lock = threading.Lock()

    def start(self, intra_enc: codec.IntraCodec, intra_dec: codec.IntraCodec, frames, qp=32):
        """Code each stream's I frame (one intra proxy pair) and seed the
        group's temporal contexts; `frames` holds one [1, 3, H, W] input per
        stream.  Returns the I-frame sizes."""
        enc_hat, dec_hat, sizes = [], [], []
        for f in frames:
            e = intra_enc.compress(f, qp, self.pad_b, self.pad_r)
            d = intra_dec.decompress(e["bit_stream"], {"height": self.H, "width": self.W}, qp, e["ec_parallel"])
            enc_hat.append(e["x_hat"].clone())
            dec_hat.append(d["x_hat"].clone())
            sizes.append(len(e["bit_stream"]))
        self.enc.add_ref_feature_from_frame(_stack(enc_hat))
        self.dec.add_ref_feature_from_frame(_stack(dec_hat), False)
        torch.cuda.synchronize()
        return sizes

    def code_chunk(self, xs, qps):
        """Encode then decode one chunk of every stream (`xs`: one [1, 24, H, W]
        input per stream); returns one ChunkResult per stream, with the
        group's encode / decode time."""
        with self.lock:
            x = _stack(xs)
            t0 = time.perf_counter()
            bits, ecs = self.enc.compress_batch(x, qps, False, self.pad_b, self.pad_r)
            t1 = time.perf_counter()
            outs = self.dec.decompress_batch(bits, qps, self.H, self.W, ecs, False)
            frames = [[o[i:i + 1].clone() for o in outs] for i in range(len(self.sids))]
            torch.cuda.current_stream().synchronize()
            t2 = time.perf_counter()
            return [ChunkResult(sid, int(q), b.tobytes(), f, t1 - t0, t2 - t1)
                    for sid, q, b, f in zip(self.sids, qps, bits, frames)]


def _stack(ts):
    return torch.cat(list(ts), 0).contiguous(memory_format=torch.channels_last)


def code_group(group: BatchGroup, xs, qps):
    with torch.cuda.stream(_worker_stream()):
        return group.code_chunk(xs, qps)


def code_groups_concurrently(groups, inputs, qps):
    """Every group at once, one worker thread per group; `inputs` / `qps` hold
    one list per group.  Returns the ChunkResults of all streams, by group."""
    with ThreadPoolExecutor(max_workers=len(groups)) as pool:
        futs = [pool.submit(code_group, g, x, q) for g, x, q in zip(groups, inputs, qps)]
        return [r for f in futs for r in f.result()]


@dataclass
class GroupSetup:
    groups: list
    chunks: list               # per group: per chunk index, one input per stream
    qps: list                  # per group: one qp per stream
    intra_bytes: list = field(default_factory=list)


def build_groups(sdi, sdp, resolutions=MIX_16, n_chunks=2, seed=0, precision="fp16", skip_thres=0.15,
                 device="cuda", max_batch=DMCHTSBatchProxy.MAX_STREAMS):
    """The streams of build_streams (same content and qps, stream for
    stream), grouped by resolution into BatchGroups of at most max_batch
    streams."""
    rng = np.random.Generator(np.random.PCG64(seed))
    qps = [int(rng.integers(0, 64)) for _ in resolutions]
    intra_e = codec.IntraCodec(sdi, skip_thres)
    intra_d = codec.IntraCodec(sdi, skip_thres)
    shared = codec.VideoCodecHTS(sdp, skip_thres)
    by_res = {}
    for sid, hw in enumerate(resolutions):
        by_res.setdefault(hw, []).append(sid)
    parts = []
    for hw, sids in by_res.items():
        for i in range(0, len(sids), max_batch):
            parts.append((hw, sids[i:i + max_batch]))
    groups, chunks, gqps, ib = [], [], [], []
    for (H, W), sids in parts:
        g = BatchGroup(sids, H, W, shared, precision)
        firsts, per_chunk = [], [[] for _ in range(n_chunks)]
        for sid in sids:
            frames = synthetic.motion_sequence_yuv420(H, W, 1 + 4 * n_chunks, seed=seed + 17 * sid)
            firsts.append(codec.to_input(synthetic.frame_to_input(frames[0]), device))
            for c in range(n_chunks):
                per_chunk[c].append(codec.to_input(synthetic.chunk_to_input(frames[1 + 4 * c:5 + 4 * c], 8), device))
        ib.extend(g.start(intra_e, intra_d, firsts))
        groups.append(g)
        chunks.append(per_chunk)
        gqps.append([qps[sid] for sid in sids])
    return GroupSetup(groups, chunks, gqps, ib), shared
