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
