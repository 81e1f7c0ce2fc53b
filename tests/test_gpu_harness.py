"""The sequence harness (codec.encode_sequence / decode_sequence, the
counterpart of test_video.py:166-399) on the GPU, and the decoders' call
order check.

* The container written by encode_sequence carries, frame for frame, the
  bitstreams that direct proxy calls produce, with the qp, ec_parallel and
  reset flags the reference loop derives (I frame, intra_period, frame
  delay 8 for HT-S with a short last chunk padded by repetition, 1 for LD).
* decode_sequence parses it back and reconstructs bit for bit what direct
  decoder calls do; bpp counts the container bytes.
"""
import io

import numpy as np
import pytest
import torch

from dcvc_amd import codec, entropy, stream_helper, synthetic, weights
from dcvc_amd.inference_extensions_cuda import DMCHTLProxy, DMCHTSProxy, DMCIProxy, DMCLDProxy

pytestmark = pytest.mark.gpu

H, W = 72, 120  # padded to 80 x 128 by the proxies


def _state(model, seed):
    sd = weights.perturb_biases(weights.init_state_dict(model, seed), seed + 100)
    return entropy.add_cdf_to_state_dict(weights.to_torch(weights.round_to_half(sd)))


def _direct(frames, sdi, sdp, proxy_cls, delay, qp_i, qp_p, reset_interval, intra_period, gpu):
    """The reference loop written out with the proxies alone: returns the
    per-NAL (is_intra, qp, ec, reset, bits) and the decoded frames."""
    pad_r, pad_b = synthetic.get_padding_size(H, W, 16)
    ienc, idec = DMCIProxy(), DMCIProxy()
    ienc.set_param(sdi, 0.0)
    idec.set_param(sdi, 0.0)
    penc, pdec = proxy_cls(), proxy_cls()
    penc.set_param(sdp, 0.0)
    pdec.set_param(sdp, 0.0)
    nals, recon = [], []
    idx = 0
    while idx < len(frames):
        intra = idx == 0 or (intra_period > 1 and idx != 1 and idx % intra_period == 1)
        if intra:
            x = codec.to_input(synthetic.frame_to_input(frames[idx]), gpu)
            bits, x_hat, ec = ienc.compress(x, qp_i, pad_b, pad_r)
            penc.add_ref_feature_from_frame(x_hat, True)
            bits = bits.copy()
            xd = idec.decompress(bits, qp_i, H, W, ec)
            pdec.add_ref_feature_from_frame(xd, False)
            recon.append(xd[:, :, :H, :W].float().cpu())
            nals.append((True, qp_i, ec, 0, bits.tobytes()))
            idx += 1
            continue
        n = min(delay, len(frames) - idx)
        reset = int(reset_interval > 0 and (idx + delay) % reset_interval == 1)
        if delay == 1:
            x = codec.to_input(synthetic.frame_to_input(frames[idx]), gpu)
        else:
            x = codec.to_input(synthetic.chunk_to_input(frames[idx:idx + n], delay), gpu)
        bits, ec = penc.compress(x, qp_p, reset, pad_b, pad_r)
        bits = bits.copy()
        out = pdec.decompress(bits, qp_p, H, W, ec, reset)
        out = out if isinstance(out, list) else [out]
        recon += [o[:, :, :H, :W].float().cpu() for o in out[:n]]
        nals.append((False, qp_p, ec, reset, bits.tobytes()))
        idx += n
    return nals, recon


@pytest.mark.parametrize("structure,n_frames,reset_interval,intra_period",
                         [("hts", 13, 16, -1), ("ld", 7, 3, 4)])
def test_sequence_harness_matches_direct_proxy_calls(gpu, structure, n_frames, reset_interval, intra_period):
    print(f"harness {structure}: {n_frames} frames {H}x{W}", flush=True)
    sdi, sdp = _state("dmci", 0), _state(structure, 1)
    frames = synthetic.motion_sequence_yuv420(H, W, n_frames, seed=5)
    cls = codec.video_codec(structure)
    qp_i, qp_p = 30, 37

    i_enc, p_enc = codec.IntraCodec(sdi), cls(sdp)
    res = codec.encode_sequence(frames, i_enc, p_enc, qp_i, qp_p, reset_interval=reset_interval,
                                intra_period=intra_period, device=gpu)
    i_dec, p_dec = codec.IntraCodec(sdi), cls(sdp)
    res = codec.decode_sequence(res.bitstream, frames, i_dec, p_dec, res, keep_recon=True)

    proxy_cls = {"hts": DMCHTSProxy, "htl": DMCHTLProxy, "ld": DMCLDProxy}[structure]
    nals, recon = _direct(frames, sdi, sdp, proxy_cls, cls.FRAME_DELAY, qp_i, qp_p, reset_interval, intra_period,
                          gpu)

    # the container holds the direct calls' bitstreams and header fields
    f = io.BytesIO(res.bitstream)
    got = []
    while f.tell() < len(res.bitstream):
        h = stream_helper.read_header(f)
        if h["nal_type"] == stream_helper.NalType.NAL_SPS:
            assert stream_helper.read_sps_remaining(f, h["sps_id"]) == {"sps_id": 0, "height": H, "width": W}
            continue
        qp, ec, reset, payload = stream_helper.read_ip_remaining(f)
        got.append((h["nal_type"] == stream_helper.NalType.NAL_I, qp, ec, reset, payload))
    assert got == nals
    assert sum(res.bits) == 8 * len(res.bitstream)
    assert len(res.frame_types) == n_frames and res.frame_types.count(0) == sum(n[0] for n in nals)
    assert any(n[3] for n in nals) and not all(n[3] for n in nals if not n[0])
    # decode_sequence reconstructs what the direct decoder calls do, frame for frame
    assert len(res.recon) == len(recon) == n_frames == len(res.psnr)
    for i, (a, b) in enumerate(zip(res.recon, recon)):
        assert torch.equal(a, b), f"frame {i}"
    assert all(np.isfinite(p) for p in res.psnr)
    print(structure, "bpp", res.bpp(), "psnr", np.mean(res.psnr))


@pytest.mark.parametrize("cls", [DMCHTSProxy, DMCHTLProxy, DMCLDProxy])
def test_decompress_before_reference_frame_is_a_call_order_error(gpu, cls):
    from dcvc_amd._lib import DcvcError
    model = {DMCHTSProxy: "hts", DMCHTLProxy: "htl", DMCLDProxy: "ld"}[cls]
    p = cls()
    p.set_param(_state(model, 1), 0.0)
    p._prepare(80, 128)
    with pytest.raises(DcvcError) as e:
        p.decompress(np.zeros(16, np.uint8), 32, H, W, 1, False)
    assert e.value.code == 5
