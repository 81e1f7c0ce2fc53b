"""Batching and serving forms on the GPU: every stream's result is the same
whether it is coded alone or in the concurrent mix (bit for bit)."""
import numpy as np
import pytest
import torch

from dcvc_amd import batching, codec, serving

pytestmark = pytest.mark.gpu

SIZES = [(128, 192), (144, 256), (96, 128), (256, 256), (128, 192)]


@pytest.fixture(scope="module")
def models(gpu):
    return codec.make_state_dicts(seed=0)


def _same(a, b):
    if isinstance(a, (bytes, bytearray)):
        return a == b
    return all(torch.equal(x.view(torch.int16), y.view(torch.int16)) for x, y in zip(a, b))


def test_concurrent_streams_equal_alone(gpu, models):
    sdi, sdp = models
    alone, _ = batching.build_streams(sdi, sdp, SIZES, n_chunks=2, seed=3)
    mixed, _ = batching.build_streams(sdi, sdp, SIZES, n_chunks=2, seed=3)
    for c in range(2):
        xs_a = [ch[c] for ch in alone.chunks]
        xs_m = [ch[c] for ch in mixed.chunks]
        ra = batching.code_sequentially(alone.contexts, xs_a, alone.qps)
        rm = batching.code_concurrently(mixed.contexts, xs_m, mixed.qps)
        for a, m in zip(ra, rm):
            assert a.bits == m.bits, f"stream {a.stream} chunk {c}: bitstream differs"
            assert _same(a.frames, m.frames), f"stream {a.stream} chunk {c}: reconstruction differs"


def test_serving_mix_equals_serial(gpu, models):
    sdi, sdp = models
    res = [SIZES[i % len(SIZES)] for i in range(6)]
    reqs = serving.poisson_trace(rate=200, duration=0.15, n_streams=6, seed=5)
    assert len(reqs) >= 12
    out = {}
    for slots in (1, 4):
        ex = serving.GpuExecutor(sdi, sdp, res, chunks_per_stream=2, seed=1)
        trace = [serving.Request(r.rid, r.t_arrive, r.stream, r.kind, r.qp) for r in reqs]
        srv = serving.Server(ex, slots=slots, hbm_budget=1 << 40)
        srv.run(trace, realtime=(slots > 1))
        assert all(not r.error for r in trace), [r.error for r in trace if r.error]
        out[slots] = trace
        st = serving.Server.stats(trace)
        print(f"slots {slots}: {st}, max running {srv.max_running}")
    for a, b in zip(out[1], out[4]):
        assert _same(a.result, b.result), f"request {a.rid} ({a.kind}, stream {a.stream}) differs in the mix"


def test_mix16_at_size_alone_equals_concurrent_bf16(gpu, models):
    """The batching form's workload at its real sizes and precision: the 16
    streams of 4x 854x480, 6x 1280x720, 4x 1920x1080 and 2x 3840x2160, one
    4-frame chunk each at a random qp, in BF16.  Every stream's bitstream and
    reconstruction are the same coded alone (one stream at a time) and with
    all 16 streams in flight at once."""
    import gc
    sdi, sdp = models
    print("MIX_16 alone", flush=True)
    alone, _ = batching.build_streams(sdi, sdp, batching.MIX_16, n_chunks=1, seed=3, precision="bf16")
    ra = batching.code_sequentially(alone.contexts, [c[0] for c in alone.chunks], alone.qps)
    ra = [(r.stream, r.bits, [f.cpu() for f in r.frames]) for r in ra]
    del alone
    gc.collect()
    torch.cuda.empty_cache()
    print("MIX_16 concurrent", flush=True)
    mixed, _ = batching.build_streams(sdi, sdp, batching.MIX_16, n_chunks=1, seed=3, precision="bf16")
    rm = batching.code_concurrently(mixed.contexts, [c[0] for c in mixed.chunks], mixed.qps)
    assert len(rm) == len(ra) == 16
    for (sid, bits, frames), m in zip(ra, rm):
        assert sid == m.stream
        assert bits == m.bits, f"stream {sid}: bitstream differs"
        assert _same(frames, [f.cpu() for f in m.frames]), f"stream {sid}: reconstruction differs"
    del mixed
    gc.collect()
    torch.cuda.empty_cache()


def test_serving_at_the_request_mix_sizes_equals_serial(gpu, models):
    """The serving form over streams drawn from its resolution mix (480p 30 %,
    720p 40 %, 1080p 25 %, 2160p 5 %): with contexts prewarmed, each request's
    result is the same served one at a time and with 4 requests in flight."""
    import gc
    sdi, sdp = models
    res = serving.resolution_mix(8, seed=1)
    print("serving mix", res, flush=True)
    reqs = serving.poisson_trace(rate=60, duration=0.4, n_streams=8, seed=7)
    assert len(reqs) >= 10
    out = {}
    for slots in (1, 4):
        ex = serving.GpuExecutor(sdi, sdp, res, chunks_per_stream=2, seed=1)
        ex.prewarm()
        trace = [serving.Request(r.rid, r.t_arrive, r.stream, r.kind, r.qp) for r in reqs]
        srv = serving.Server(ex, slots=slots, hbm_budget=1 << 40)
        srv.run(trace, realtime=(slots > 1))
        assert all(not r.error for r in trace), [r.error for r in trace if r.error]
        out[slots] = [(r.rid, r.result if isinstance(r.result, bytes) else [f.cpu() for f in r.result])
                      for r in trace]
        print(f"slots {slots}: {serving.Server.stats(trace)}, max running {srv.max_running}")
        del ex, srv
        gc.collect()
        torch.cuda.empty_cache()
    for (ra, a), (rb, b) in zip(out[1], out[4]):
        assert ra == rb
        assert _same(a, b), f"request {ra} differs in the mix"
