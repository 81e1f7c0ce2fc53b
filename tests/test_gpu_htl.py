"""End-to-end GPU parity of the HT-L chunk codec (DMCHTLProxy).

* The decoder rebuilds, from the bitstream alone, the encoder's y_hat,
  temporal context and every per-step index stream bit for bit, over chunks
  with and without feature-memory reset, several QPs and a picture size that
  needs the 1/64-scale padding.
* Against the FP32 CPU oracle fed the same fp16-rounded weights and the same
  reference frame: per-step symbol / index match rates, bitstream size and
  reconstruction PSNR.
"""
import numpy as np
import pytest
import torch

from dcvc_amd import entropy, synthetic, weights
from oracle import dcvc_oracle as O

pytestmark = pytest.mark.gpu


def _state(model, seed):
    sd = weights.perturb_biases(weights.init_state_dict(model, seed), seed + 100)
    sd = weights.round_to_half(sd)
    return entropy.add_cdf_to_state_dict(weights.to_torch(sd))


@pytest.fixture(scope="module")
def models(gpu):
    return _state("dmci", 0), _state("htl", 2)


def _to_gpu_nchw(arr, gpu):
    return torch.from_numpy(arr)[None].half().to(gpu).contiguous(memory_format=torch.channels_last)


def psnr(a, b):
    mse = float(((a.double() - b.double()) ** 2).mean())
    return 10 * np.log10(1.0 / max(mse, 1e-12))


def _intra(sdi, frame, gpu, H, W, pad_b, pad_r):
    from dcvc_amd.inference_extensions_cuda import DMCIProxy
    e = DMCIProxy()
    e.set_param(sdi, 0.0)
    bits, x_hat, ec = e.compress(_to_gpu_nchw(frame, gpu), 32, pad_b, pad_r)
    d = DMCIProxy()
    d.set_param(sdi, 0.0)
    return x_hat.clone(), d.decompress(bits, 32, H, W, ec).clone()


@pytest.mark.parametrize("H,W", [(128, 192), (136, 200), (240, 416), (1080, 1920)])
def test_htl_chunks_encoder_decoder_bitexact(gpu, models, H, W):
    from dcvc_amd.inference_extensions_cuda import DMCHTLProxy
    sdi, sdp = models
    pad_r, pad_b = synthetic.get_padding_size(H, W, 16)
    Hp, Wp = H + pad_b, W + pad_r
    frames = synthetic.motion_sequence_yuv420(H, W, 17, seed=5)
    x_hat_i, x_hat_id = _intra(sdi, synthetic.frame_to_input(frames[0]), gpu, H, W, pad_b, pad_r)
    enc, dec = DMCHTLProxy(), DMCHTLProxy()
    enc.set_param(sdp, 0.15)
    dec.set_param(sdp, 0.15)
    enc.add_ref_feature_from_frame(x_hat_i, True)
    dec.add_ref_feature_from_frame(x_hat_id, False)
    P16, P8 = (Hp // 16) * (Wp // 16), (Hp // 8) * (Wp // 8)
    for chunk, (reset, qp) in enumerate([(False, 32), (True, 5), (False, 63)]):
        src = frames[1 + 8 * chunk: 9 + 8 * chunk] if chunk < 2 else frames[9:17]
        bits, ec = enc.compress(_to_gpu_nchw(synthetic.chunk_to_input(src), gpu), qp, reset, pad_b, pad_r)
        yhat_e = enc.debug_copy(0, P16 * 256 * 2, np.float16).copy()
        feat_e = enc.debug_copy(3, P8 * 1024 * 2, np.float16).reshape(P8, 1024)[:, 512:].copy()
        counts_e = [enc.step_symbol_count(k) for k in range(4)]
        outs = dec.decompress(bits, qp, H, W, ec, reset)
        torch.cuda.synchronize()
        yhat_d = dec.debug_copy(0, P16 * 256 * 2, np.float16)
        feat_d = dec.debug_copy(3, P8 * 1024 * 2, np.float16).reshape(P8, 1024)[:, 512:]
        assert [dec.step_symbol_count(k) for k in range(4)] == counts_e, f"chunk {chunk}: step counts"
        assert sum(counts_e) > 0
        assert np.array_equal(yhat_e.view(np.uint16), yhat_d.view(np.uint16)), f"chunk {chunk}: y_hat differs"
        assert np.array_equal(feat_e.view(np.uint16), feat_d.view(np.uint16)), f"chunk {chunk}: feature differs"
        for f in outs:
            assert f.shape == (1, 3, Hp, Wp)
            assert torch.isfinite(f.float()).all() and float(f.abs().max()) <= 0.5


def test_htl_symbols_and_psnr_vs_oracle(gpu, models):
    """End to end against the FP32 oracle: y within fp16 storage error; the
    per-step symbols, bitstream size and PSNR close.  Conditioned on the
    GPU's own y and z_hat (z rounding is discontinuous: a z that lands on the
    other side of .5 changes the whole prior around it), the oracle's common
    params and per-step symbols must agree much more tightly."""
    from dcvc_amd.inference_extensions_cuda import DMCHTLProxy
    sdi, sdp = models
    H, W = 128, 192
    frames = synthetic.motion_sequence_yuv420(H, W, 9, seed=6)
    x_hat_i, _ = _intra(sdi, synthetic.frame_to_input(frames[0]), gpu, H, W, 0, 0)
    ref_frame = x_hat_i.float().cpu().contiguous()
    enc, dec = DMCHTLProxy(), DMCHTLProxy()
    enc.set_param(sdp, 0.0)
    dec.set_param(sdp, 0.0)
    enc.add_ref_feature_from_frame(x_hat_i, True)
    dec.add_ref_feature_from_frame(x_hat_i, False)
    chunk = synthetic.chunk_to_input(frames[1:9])
    bits, ec = enc.compress(_to_gpu_nchw(chunk, gpu), 32, False, 0, 0)
    H16, W16 = H // 16, W // 16
    P16, nq = H16 * W16, H16 * W16 * 64
    sym_g = [enc.debug_copy(5 + k, nq * 2, np.int16).astype(np.int64) for k in range(4)]
    y_g = enc.debug_copy(1, P16 * 256 * 2, np.float16).astype(np.float32)
    com_g = enc.debug_copy(2, P16 * 768 * 2, np.float16).astype(np.float32).reshape(P16, 768)
    z_g = enc.debug_copy(13, (H16 // 4) * (W16 // 4) * 128, np.int8)
    outs = [o.float().cpu() for o in dec.decompress(bits, 32, H, W, ec, False)]
    sdo = {k: v.float() for k, v in sdp.items()}
    st = O.HTSState()
    O.htl_add_ref_feature_from_frame(sdo, st, ref_frame, True)
    memory = st.memory.clone()
    r = O.htl_compress(sdo, st, torch.from_numpy(chunk)[None].half().float(), 32, False, recon=True)
    nchw = lambda a, C: torch.from_numpy(a.reshape(1, H16, W16, C)).permute(0, 3, 1, 2).contiguous()

    def match(res):
        sm, im = [], []
        for k in range(4):
            o = res.steps_dense[k].astype(np.int64)
            sm.append(float(np.mean((sym_g[k] >> 8) == (o >> 8))))
            im.append(float(np.mean((sym_g[k] & 0xFF) == (o & 0xFF))))
        return sm, im

    sym_m, idx_m = match(r)
    y_o = (r.y * torch.clamp_min(r.common_params[:, :256], 0.5)).permute(0, 2, 3, 1).reshape(-1).numpy()
    y_rel = float(np.linalg.norm(y_g - y_o) / np.linalg.norm(y_o))
    z_match = float(np.mean(z_g == r.z_int8))
    # conditioned: the oracle's prior from the GPU's z_hat, its steps from the GPU's y
    z_hat = torch.from_numpy(z_g.astype(np.float32)).reshape(1, H16 // 4, W16 // 4, 128).permute(0, 3, 1, 2)
    params_c = O._htl_prior(sdo, z_hat.contiguous(), memory, 32)
    com_o = params_c.permute(0, 2, 3, 1).reshape(P16, 768).numpy()
    com_rel = float(np.linalg.norm(com_g - com_o) / np.linalg.norm(com_o))
    rc = O.htl_code_y(sdo, nchw(y_g, 256), params_c)
    sym_c, idx_c = match(rc)
    enc_o = entropy.RansEncoder()
    enc_o.set_cdf(sdp["bit_estimator_z.quantized_cdf"].numpy(), sdp["bit_estimator_z.cdf_length"].numpy(), 0)
    enc_o.set_cdf(sdp["gaussian_encoder.quantized_cdf"].numpy(), sdp["gaussian_encoder.cdf_length"].numpy(), 1)
    enc_o.reset()
    enc_o.set_entropy_coder_parallel(ec)
    for k in (3, 2, 1, 0):
        enc_o.encode_y(r.steps_sym16[k])
    enc_o.encode_z(r.z_int8, 32 * 128, 128)
    enc_o.flush()
    n_o = len(enc_o.get_encoded_stream())
    pg = [psnr(outs[i], torch.from_numpy(chunk[3 * i:3 * i + 3])[None]) for i in range(8)]
    po = [psnr(r.x_hat[i], torch.from_numpy(chunk[3 * i:3 * i + 3])[None]) for i in range(8)]
    print(f"HT-L vs FP32 oracle: y rel {y_rel:.2e}, z match {z_match:.4f}; end to end per step symbol {sym_m} "
          f"index {idx_m}; bytes gpu {len(bits)} oracle {n_o}; PSNR gpu {np.round(pg, 4)} oracle {np.round(po, 4)}")
    print(f"conditioned on the GPU's y and z_hat: common params rel {com_rel:.2e}, per step symbol {sym_c} "
          f"index {idx_c}")
    assert y_rel < 4e-3 and z_match > 0.99
    assert com_rel < 4e-3
    # conditioned (measured on MI355X: symbols 99.9 / 99.8 / 99.6 / 99.1 %,
    # indexes 98.3 / 96.5 / 94.9 / 92.8 % over the steps; a relative scale
    # error e moves the index with probability ~ e / 0.039, the log-bin width,
    # and every step's spatial prior sees the earlier steps' differences)
    assert min(sym_c) > 0.985 and min(idx_c) > 0.90
    # end to end, symbols / indexes and bytes are checked against the fp16
    # storage floor in tests/test_gpu_parity.py
    assert max(abs(a - b) for a, b in zip(pg, po)) <= 0.02


def test_htl_shared_weights_and_precision_forms(gpu, models):
    """Contexts sharing one weight copy code identically to a context with its
    own copy; BF16 / FP8 keep encoder and decoder bit-identical."""
    from dcvc_amd.inference_extensions_cuda import DMCHTLProxy
    sdi, sdp = models
    H, W = 64, 128
    frames = synthetic.motion_sequence_yuv420(H, W, 9, seed=7)
    x_hat_i, _ = _intra(sdi, synthetic.frame_to_input(frames[0]), gpu, H, W, 0, 0)
    x = _to_gpu_nchw(synthetic.chunk_to_input(frames[1:9]), gpu)
    own = DMCHTLProxy()
    own.set_param(sdp, 0.15)
    shared = DMCHTLProxy()
    shared.share_weights_from(own)
    res = []
    for p in (own, shared):
        p.add_ref_feature_from_frame(x_hat_i, True)
        b, _ = p.compress(x, 40, False, 0, 0)
        res.append(b.tobytes())
    assert res[0] == res[1]
    P16 = (H // 16) * (W // 16)
    for prec in ("bf16", "fp8"):
        enc, dec = DMCHTLProxy(), DMCHTLProxy()
        enc.share_weights_from(own)
        dec.share_weights_from(own)
        enc.set_precision(prec)
        dec.set_precision(prec)
        enc.add_ref_feature_from_frame(x_hat_i, True)
        dec.add_ref_feature_from_frame(x_hat_i, False)
        bits, ec = enc.compress(x, 40, False, 0, 0)
        ye = enc.debug_copy(0, P16 * 256 * 2, np.float16).copy()
        outs = dec.decompress(bits, 40, H, W, ec, False)
        torch.cuda.synchronize()
        yd = dec.debug_copy(0, P16 * 256 * 2, np.float16)
        assert np.array_equal(ye.view(np.uint16), yd.view(np.uint16)), prec
        assert all(torch.isfinite(o.float()).all() for o in outs)


def test_htl_decoder_recovers_after_a_corrupt_bitstream(gpu, models):
    """As for HT-S and LD: a truncated chunk bitstream raises in the host z
    decode with the temporal-context work already queued on the side stream;
    after both sides re-anchor on the intra frame, the next chunk decodes bit
    for bit."""
    from dcvc_amd.inference_extensions_cuda import DMCHTLProxy
    sdi, sdp = models
    H, W = 128, 192
    frames = synthetic.motion_sequence_yuv420(H, W, 17, seed=6)
    x_hat_i, x_hat_id = _intra(sdi, synthetic.frame_to_input(frames[0]), gpu, H, W, 0, 0)
    enc, dec = DMCHTLProxy(), DMCHTLProxy()
    enc.set_param(sdp, 0.15)
    dec.set_param(sdp, 0.15)
    enc.add_ref_feature_from_frame(x_hat_i, True)
    dec.add_ref_feature_from_frame(x_hat_id, False)
    bits, ec = enc.compress(_to_gpu_nchw(synthetic.chunk_to_input(frames[1:9]), gpu), 32, False, 0, 0)
    with pytest.raises(Exception):
        dec.decompress(np.ascontiguousarray(bits[:3]), 32, H, W, ec, False)
    enc.add_ref_feature_from_frame(x_hat_i, True)
    dec.add_ref_feature_from_frame(x_hat_id, False)
    P16 = (H // 16) * (W // 16)
    bits, ec = enc.compress(_to_gpu_nchw(synthetic.chunk_to_input(frames[9:17]), gpu), 32, False, 0, 0)
    yhat_e = enc.debug_copy(0, P16 * 256 * 2, np.float16).copy()
    dec.decompress(bits, 32, H, W, ec, False)
    torch.cuda.synchronize()
    yhat_d = dec.debug_copy(0, P16 * 256 * 2, np.float16)
    assert np.array_equal(yhat_e.view(np.uint16), yhat_d.view(np.uint16))
