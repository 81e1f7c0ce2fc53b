"""Batches of HT-S streams: one proxy codes n streams of one resolution with
one launch per layer over the n pictures.  Each stream's bitstream and
reconstruction must equal those of a proxy of its own, with per-stream QPs,
resets, padding, and in each product precision."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _streams(n, H, W, n_chunks, seed, gpu):
    from dcvc_amd import codec, synthetic
    out = []
    for s in range(n):
        fr = synthetic.motion_sequence_yuv420(H, W, 1 + 4 * n_chunks, seed=seed + 17 * s)
        x0 = codec.to_input(synthetic.frame_to_input(fr[0]), gpu)
        chunks = [codec.to_input(synthetic.chunk_to_input(fr[1 + 4 * c:5 + 4 * c], 8), gpu) for c in range(n_chunks)]
        out.append((x0, chunks))
    return out


def _cat(ts):
    return torch.cat([t.float() for t in ts], 0).half().contiguous(memory_format=torch.channels_last)


@pytest.mark.parametrize("H,W,n,prec", [(96, 160, 3, "fp16"), (100, 176, 2, "fp16"), (64, 128, 4, "fp8")])
def test_batch_equals_separate_proxies(gpu, H, W, n, prec):
    from dcvc_amd import codec, synthetic
    from dcvc_amd.inference_extensions_cuda import DMCHTSBatchProxy, DMCHTSProxy
    sdi, sdp = codec.make_state_dicts(seed=3)
    pad_r, pad_b = synthetic.get_padding_size(H, W, 16)
    streams = _streams(n, H, W, 3, seed=5, gpu=gpu)
    intra = codec.IntraCodec(sdi, 0.15)
    x_hat = [intra.compress(x0, 32, pad_b, pad_r)["x_hat"] for x0, _ in streams]
    # per-chunk QPs and resets of every stream
    rng = np.random.Generator(np.random.PCG64(n))
    qps = [[int(v) for v in rng.integers(0, 64, n)] for _ in range(3)]
    resets = [False, True, False]

    singles = []
    for s in range(n):
        e, d = DMCHTSProxy(), DMCHTSProxy()
        e.set_param(sdp, 0.15)
        d.share_weights_from(e)
        e.set_precision(prec)
        d.set_precision(prec)
        e.add_ref_feature_from_frame(x_hat[s], True)
        d.add_ref_feature_from_frame(x_hat[s], False)
        res = []
        for c in range(3):
            bits, ec = e.compress(streams[s][1][c], qps[c][s], resets[c], pad_b, pad_r)
            frames = d.decompress(bits, qps[c][s], H, W, ec, resets[c])
            res.append((bits.tobytes(), ec, torch.cat([f.float() for f in frames], 1).cpu()))
        singles.append(res)
        del e, d

    be, bd = DMCHTSBatchProxy(n), DMCHTSBatchProxy(n)
    be.set_param(sdp, 0.15)
    bd.share_weights_from(be)
    be.set_precision(prec)
    bd.set_precision(prec)
    be.add_ref_feature_from_frame(_cat(x_hat), True)
    bd.add_ref_feature_from_frame(_cat(x_hat), False)
    for c in range(3):
        bits, ecs = be.compress_batch(_cat([st[1][c] for st in streams]), qps[c], resets[c], pad_b, pad_r)
        frames = bd.decompress_batch(bits, qps[c], H, W, ecs, resets[c])
        rec = torch.cat([f.float() for f in frames], 1).cpu()
        for s in range(n):
            want_bits, want_ec, want_rec = singles[s][c]
            assert bits[s].tobytes() == want_bits, (c, s)
            assert ecs[s] == want_ec
            assert torch.equal(rec[s:s + 1], want_rec), (c, s)
            assert be.last_symbol_count(s) >= 0


def test_batch_proxy_refuses_single_stream_calls(gpu):
    from dcvc_amd import codec
    from dcvc_amd.inference_extensions_cuda import DMCHTSBatchProxy
    _, sdp = codec.make_state_dicts(seed=3)
    p = DMCHTSBatchProxy(2)
    p.set_param(sdp, 0.15)
    with pytest.raises(TypeError):
        p.compress(None, 32, False, 0, 0)
    with pytest.raises(ValueError):
        DMCHTSBatchProxy(17)
    x = torch.zeros((2, 3, 64, 64), dtype=torch.float16, device=gpu).contiguous(memory_format=torch.channels_last)
    p.add_ref_feature_from_frame(x, True)
    chunk = torch.zeros((2, 24, 64, 64), dtype=torch.float16, device=gpu).contiguous(memory_format=torch.channels_last)
    with pytest.raises(ValueError):
        p.compress_batch(chunk, [32], False, 0, 0)  # one qp per stream
    from dcvc_amd import _lib
    with pytest.raises(_lib.DcvcError):
        p.compress_batch(chunk, [32, 64], False, 0, 0)  # qp out of range
    torch.cuda.synchronize()


def test_resolution_groups_equal_separate_streams(gpu):
    """batching.build_groups (one batch proxy pair per resolution, groups on
    concurrent host threads) codes every stream of a mix exactly as
    build_streams' one-proxy-per-stream contexts do."""
    from dcvc_amd import batching, codec
    sdi, sdp = codec.make_state_dicts(seed=0)
    mix = [(96, 128)] * 3 + [(128, 192)] * 2 + [(64, 64)]
    setup, _ = batching.build_streams(sdi, sdp, mix, n_chunks=2, seed=1)
    ref = {}
    for c in range(2):
        for r in batching.code_sequentially(setup.contexts, [ch[c] for ch in setup.chunks], setup.qps):
            ref[(c, r.stream)] = r
    del setup
    gs, _ = batching.build_groups(sdi, sdp, mix, n_chunks=2, seed=1)
    assert sorted(len(g.sids) for g in gs.groups) == [1, 2, 3]
    for c in range(2):
        res = batching.code_groups_concurrently(gs.groups, [ch[c] for ch in gs.chunks], gs.qps)
        assert sorted(r.stream for r in res) == list(range(len(mix)))
        for r in res:
            want = ref[(c, r.stream)]
            assert r.qp == want.qp and r.bits == want.bits, (c, r.stream)
            assert all(torch.equal(a, b) for a, b in zip(r.frames, want.frames))
