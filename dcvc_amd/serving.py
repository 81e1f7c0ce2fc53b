"""Serving form: a continuous-batching request loop for many persistent
streams on one GPU.

Requests (encode or decode of one 4-frame chunk of one stream) arrive over
time.  ``Server`` admits a request when
  * a slot is free (at most ``slots`` requests run on the device at once),
  * its stream has no request in flight (a stream's chunks are coded in
    order: each one updates the stream's temporal context), and
  * the stream's contexts are resident, or fit in the HBM byte budget.
    Contexts stay resident once admitted (the temporal context lives on the
    device), so the budget bounds how many distinct streams are served.
A request that arrives while others run starts as soon as these hold: its
work overlaps theirs on the device (continuous batching, not a queue in
front of one stream at a time).  The scheduling policy is first-come
first-served among admissible requests.

The executor is an interface (``footprint(stream)``, ``run(request)``), so
the scheduler is tested on the CPU with a mock and driven on the GPU by
``GpuExecutor`` (HT-S proxies sharing one weight copy).
"""
from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    t_arrive: float          # seconds from the start of the trace
    stream: int
    kind: str                # "encode" | "decode"
    qp: int
    t_start: float = -1.0
    t_done: float = -1.0
    result: object = None
    error: str = ""

    @property
    def latency(self):
        return self.t_done - self.t_arrive


def poisson_trace(rate, duration, n_streams, seed=0, decode_fraction=0.5):
    """Poisson arrivals at `rate` per second over `duration` seconds; each
    request picks a stream uniformly, encode or decode, and qp in [0, 63]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t, reqs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        kind = "decode" if rng.random() < decode_fraction else "encode"
        reqs.append(Request(len(reqs), t, int(rng.integers(0, n_streams)), kind, int(rng.integers(0, 64))))
    return reqs


class Server:
    def __init__(self, executor, slots: int, hbm_budget: int):
        self.ex = executor
        self.slots = slots
        self.budget = hbm_budget
        self.resident = {}              # stream -> bytes
        self.resident_bytes = 0
        self.busy_streams = set()
        self.pending = deque()
        self.running = 0
        self.cv = threading.Condition()
        self.max_running = 0
        self.max_resident_bytes = 0
        self.rejected = []

    # ---------------------------------------------------------------- policy
    def _admissible(self, r: Request) -> bool:
        if self.running >= self.slots or r.stream in self.busy_streams:
            return False
        if r.stream in self.resident:
            return True
        return self.resident_bytes + self.ex.footprint(r.stream) <= self.budget

    def _refuse_unfittable(self):
        """Resident contexts are never evicted, so a stream that does not fit
        now never will: refuse its pending requests (status 'budget')."""
        keep = deque()
        refused = 0
        for r in self.pending:
            if r.stream not in self.resident and self.resident_bytes + self.ex.footprint(r.stream) > self.budget:
                r.error = "budget: the stream's contexts do not fit in the HBM budget"
                self.rejected.append(r)
                refused += 1
            else:
                keep.append(r)
        self.pending = keep
        return refused

    def _take(self):
        for i, r in enumerate(self.pending):
            if self._admissible(r):
                del self.pending[i]
                if r.stream not in self.resident:
                    need = self.ex.footprint(r.stream)
                    self.resident[r.stream] = need
                    self.resident_bytes += need
                    self.max_resident_bytes = max(self.max_resident_bytes, self.resident_bytes)
                self.busy_streams.add(r.stream)
                self.running += 1
                self.max_running = max(self.max_running, self.running)
                return r
        return None

    # ---------------------------------------------------------------- loop
    def run(self, requests, realtime=True):
        """Feed `requests` (sorted by arrival) and serve them; returns them
        with start / done times filled.  realtime=False submits everything at
        t = 0 (a closed-loop batch)."""
        for r in requests:
            if self.ex.footprint(r.stream) > self.budget:
                r.error = "stream footprint exceeds the HBM budget"
                self.rejected.append(r)
        todo = [r for r in requests if not r.error]
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        done_count = [0]
        pre_refused = len(requests) - len(todo)

        def worker(r: Request):
            try:
                r.result = self.ex.run(r)
            except Exception as e:  # surfaced in the request, the loop keeps serving
                r.error = f"{type(e).__name__}: {e}"
            with self.cv:
                r.t_done = now()
                self.running -= 1
                self.busy_streams.discard(r.stream)
                done_count[0] += 1
                self.cv.notify_all()

        def feeder():
            for r in todo:
                if realtime:
                    dt = r.t_arrive - now()
                    if dt > 0:
                        time.sleep(dt)
                with self.cv:
                    if not realtime:
                        r.t_arrive = now()
                    self.pending.append(r)
                    self.cv.notify_all()

        threads = []
        ft = threading.Thread(target=feeder, daemon=True)
        ft.start()
        with self.cv:
            while done_count[0] + len(self.rejected) - pre_refused < len(todo):
                r = self._take()
                if r is None:
                    if self._refuse_unfittable():
                        continue
                    self.cv.wait(timeout=0.05)
                    continue
                r.t_start = now()
                t = threading.Thread(target=worker, args=(r,), daemon=True)
                t.start()
                threads.append(t)
        ft.join()
        for t in threads:
            t.join()
        return requests

    @staticmethod
    def stats(requests, frames_per_request=4):
        """Latency percentiles over all served requests, and over the warm ones:
        each (stream, kind)'s first request also allocates that stream's
        context and captures its graphs, a one-time cost of a new stream."""
        ok = [r for r in requests if not r.error and r.t_done >= 0]
        lat = np.array([r.latency for r in ok]) * 1e3
        span = max(r.t_done for r in ok) - min(r.t_arrive for r in ok) if ok else 0.0
        seen, warm, first = set(), [], []
        for r in sorted(ok, key=lambda r: r.t_start):
            key = (r.stream, r.kind)
            (warm if key in seen else first).append(r.latency * 1e3)
            seen.add(key)
        return {"requests": len(requests), "served": len(ok), "errors": len(requests) - len(ok),
                "p50_ms": round(float(np.percentile(lat, 50)), 3) if ok else None,
                "p99_ms": round(float(np.percentile(lat, 99)), 3) if ok else None,
                "mean_ms": round(float(lat.mean()), 3) if ok else None,
                "warm_p50_ms": round(float(np.percentile(warm, 50)), 3) if warm else None,
                "warm_p99_ms": round(float(np.percentile(warm, 99)), 3) if warm else None,
                "first_requests": len(first),
                "first_request_mean_ms": round(float(np.mean(first)), 3) if first else None,
                "frames_per_s": round(frames_per_request * len(ok) / span, 2) if span > 0 else None}


# ---------------------------------------------------------------------------
# GPU executor over HT-S proxies
# ---------------------------------------------------------------------------

class GpuExecutor:
    """Streams at the given resolutions; content is shared per resolution
    (synthetic), contexts are per stream.  Decode requests decode the stream's
    next bitstream of a sequence pre-encoded at setup from the same I frame;
    encode requests code the stream's next chunk with its own encoder."""

    def __init__(self, sdi, sdp, resolutions, chunks_per_stream=4, skip_thres=0.15, seed=0, device="cuda",
                 precision="fp16", keep_results=True):
        import torch
        from . import codec, synthetic
        self.torch = torch
        self.codec = codec
        self.res = list(resolutions)
        self.device = device
        # keep_results=False returns sizes instead of the bitstream and copies
        # of the decoded frames, so a long trace does not hold every result
        self.keep_results = keep_results
        self.precision = precision
        self.shared = codec.VideoCodecHTS(sdp, skip_thres)
        intra_e, intra_d = codec.IntraCodec(sdi, skip_thres), codec.IntraCodec(sdi, skip_thres)
        self.content = {}
        self._fp = {}
        for (H, W) in sorted(set(self.res)):
            frames = synthetic.motion_sequence_yuv420(H, W, 1 + 4 * chunks_per_stream, seed=seed + H)
            pad_r, pad_b = synthetic.get_padding_size(H, W, 16)
            x0 = codec.to_input(synthetic.frame_to_input(frames[0]), device)
            e = intra_e.compress(x0, 32, pad_b, pad_r)
            d = intra_d.decompress(e["bit_stream"], {"height": H, "width": W}, 32, e["ec_parallel"])
            chunks = [codec.to_input(synthetic.chunk_to_input(frames[1 + 4 * c:5 + 4 * c], 8), device)
                      for c in range(chunks_per_stream)]
            # bitstreams the decode requests consume, one per chunk and qp 32
            pre = codec.VideoCodecHTS(share_from=self.shared)
            pre.proxy.set_precision(precision)
            pre.add_ref_feature_from_frame(e["x_hat"])
            bits = []
            for c in chunks:
                b, ec = pre.proxy.compress(c, 32, False, pad_b, pad_r)
                bits.append((b.copy(), ec))
            self.content[(H, W)] = {"x_hat_e": e["x_hat"].clone(), "x_hat_d": d["x_hat"].clone(), "chunks": chunks,
                                    "bits": bits, "pad": (pad_b, pad_r)}
            # footprint of one stream (encoder + decoder contexts) at this size
            self._fp[(H, W)] = 2 * pre.proxy.device_bytes() + 2 * 8 * 3 * (H + pad_b) * (W + pad_r) * 2
            del pre
        torch.cuda.synchronize()
        self.ctx = {}
        self.lock = threading.Lock()

    def footprint(self, stream):
        return self._fp[self.res[stream]]

    def _context(self, stream):
        with self.lock:
            c = self.ctx.get(stream)
            if c is None:
                H, W = self.res[stream]
                cont = self.content[(H, W)]
                enc = self.codec.VideoCodecHTS(share_from=self.shared)
                dec = self.codec.VideoCodecHTS(share_from=self.shared)
                enc.proxy.set_precision(self.precision)
                dec.proxy.set_precision(self.precision)
                c = {"enc": enc, "dec": dec, "ne": 0, "nd": 0, "init": False}
                self.ctx[stream] = c
            return c

    def prewarm(self):
        """Allocate every stream's contexts and capture their graphs before
        serving (two encodes and two decodes each, so both the first-chunk
        and the steady-state graph variants exist), then put every stream
        back at its I frame.  Without it a stream's first request pays for
        the allocation and the captures."""
        torch = self.torch
        s = getattr(_tls, "stream", None)
        if s is None:
            s = _tls.stream = torch.cuda.Stream()
        for stream in range(len(self.res)):
            H, W = self.res[stream]
            cont = self.content[(H, W)]
            c = self._context(stream)
            pad_b, pad_r = cont["pad"]
            with torch.cuda.stream(s):
                c["enc"].add_ref_feature_from_frame(cont["x_hat_e"])
                c["dec"].add_ref_feature_from_frame(cont["x_hat_d"], apply_feature_adaptor=False)
                for i in range(2):
                    c["enc"].proxy.compress(cont["chunks"][i], 32, False, pad_b, pad_r)
                    bits, ec = cont["bits"][i]
                    c["dec"].proxy.decompress(bits, 32, H, W, ec, False)
                c["enc"].add_ref_feature_from_frame(cont["x_hat_e"])
                c["dec"].add_ref_feature_from_frame(cont["x_hat_d"], apply_feature_adaptor=False)
                torch.cuda.current_stream().synchronize()
            c.update(ne=0, nd=0, init=True)

    def run(self, r: Request):
        torch = self.torch
        H, W = self.res[r.stream]
        cont = self.content[(H, W)]
        c = self._context(r.stream)
        pad_b, pad_r = cont["pad"]
        s = getattr(_tls, "stream", None)
        if s is None:
            s = _tls.stream = torch.cuda.Stream()
        with torch.cuda.stream(s):
            if not c["init"]:
                c["enc"].add_ref_feature_from_frame(cont["x_hat_e"])
                c["dec"].add_ref_feature_from_frame(cont["x_hat_d"], apply_feature_adaptor=False)
                c["init"] = True
            if r.kind == "encode":
                x = cont["chunks"][c["ne"] % len(cont["chunks"])]
                c["ne"] += 1
                bits, ec = c["enc"].proxy.compress(x, r.qp, False, pad_b, pad_r)
                out = bits.tobytes() if self.keep_results else len(bits)
            else:
                n = len(cont["bits"])
                if c["nd"] and c["nd"] % n == 0:
                    # the pre-encoded sequence starts over from the I frame
                    c["dec"].add_ref_feature_from_frame(cont["x_hat_d"], apply_feature_adaptor=False)
                bits, ec = cont["bits"][c["nd"] % n]
                c["nd"] += 1
                frames = c["dec"].proxy.decompress(bits, 32, H, W, ec, False)
                out = [f.clone() for f in frames] if self.keep_results else len(frames)
            torch.cuda.current_stream().synchronize()
        return out


_tls = threading.local()

MIX_64 = "480p 30 %, 720p 40 %, 1080p 25 %, 2160p 5 %"


def resolution_mix(n_streams, seed=0):
    """Stream resolutions drawn from {480p: .3, 720p: .4, 1080p: .25, 2160p: .05}."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = [(480, 854), (720, 1280), (1080, 1920), (2160, 3840)]
    idx = rng.choice(4, size=n_streams, p=[0.3, 0.4, 0.25, 0.05])
    return [sizes[i] for i in idx]
