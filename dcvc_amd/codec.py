"""Host-side codec objects and the sequence loop around the proxies.

``IntraCodec`` / ``VideoCodecHTS`` / ``VideoCodecHTL`` mirror the inference interface of the
reference model classes (DMCI.compress/decompress, image_model.py:194-217;
DMC.add_ref_feature_from_frame/compress/decompress, video_model_ht.py:
413-450) over a state dict with the reference's keys.  ``encode_sequence``
and ``decode_sequence`` follow the reference harness
(test_video.py:166-399): an I frame, then chunks of up to the codec's frame
delay (8 for HT-S / HT-L, video_model_ht.py:16; 1 for LD,
video_model_ld.py:16) padded by repetition, an I frame again every
``intra_period`` frames, feature-memory reset every ``reset_interval``
frames, the stream_helper container, PSNR on YUV420 with 6:1:1 weighting.
"""
from __future__ import annotations

import io
import time
from dataclasses import dataclass, field

import numpy as np
import torch
import torch.nn.functional as F

from . import entropy, stream_helper, synthetic, weights
from .inference_extensions_cuda import DMCHTLProxy, DMCHTSProxy, DMCIProxy, DMCLDProxy


def make_state_dicts(seed=0, perturb_bias=False, video="hts"):
    """(intra, video) state dicts: seeded init, fp16-rounded, plus CDF tables.
    ``video`` is the inter model structure, 'hts', 'htl' or 'ld'."""
    out = []
    for i, model in enumerate(("dmci", video)):
        sd = weights.init_state_dict(model, seed + i)
        if perturb_bias:
            sd = weights.perturb_biases(sd, seed + 100 + i)
        sd = weights.to_torch(weights.round_to_half(sd))
        out.append(entropy.add_cdf_to_state_dict(sd))
    return tuple(out)


def to_input(frame_or_chunk, device):
    """numpy [C, H, W] in [-0.5, 0.5] -> fp16 [1, C, H, W] channels_last on the GPU."""
    t = torch.from_numpy(np.ascontiguousarray(frame_or_chunk))[None].to(device=device, dtype=torch.float16)
    return t.contiguous(memory_format=torch.channels_last)


class IntraCodec:
    def __init__(self, state_dict, skip_thres=0.0):
        self.proxy = DMCIProxy()
        self.proxy.set_param(state_dict, skip_thres)

    def compress(self, x, qp, padding_b, padding_r):
        bits, x_hat, ec = self.proxy.compress(x, qp, padding_b, padding_r)
        return {"bit_stream": bits.tobytes(), "x_hat": x_hat, "ec_parallel": ec}

    def decompress(self, bit_stream, sps, qp, ec_part):
        x_hat = self.proxy.decompress(np.frombuffer(bit_stream, dtype=np.uint8), qp, sps["height"], sps["width"],
                                      ec_part)
        return {"x_hat": x_hat}


class VideoCodecHTS:
    PROXY = DMCHTSProxy
    FRAME_DELAY = 8  # g_frame_delay, video_model_ht.py:16

    def __init__(self, state_dict=None, skip_thres=0.0, share_from: "VideoCodecHTS | None" = None):
        self.proxy = self.PROXY()
        if share_from is not None:
            self.proxy.share_weights_from(share_from.proxy)
        else:
            self.proxy.set_param(state_dict, skip_thres)

    def add_ref_feature_from_frame(self, frame, apply_feature_adaptor=True):
        self.proxy.add_ref_feature_from_frame(frame, apply_feature_adaptor)

    def clear_dpb(self):
        """State is reset by the next add_ref_feature_from_frame (video_model_ht.py:364-367)."""

    def compress(self, x, qp, reset_feature_memory, padding_b, padding_r):
        bits, ec = self.proxy.compress(x, qp, reset_feature_memory, padding_b, padding_r)
        return {"bit_stream": bits.tobytes(), "ec_parallel": ec}

    def decompress(self, bit_stream, sps, qp, ec_part, reset_feature_memory):
        x_hat = self.proxy.decompress(np.frombuffer(bit_stream, dtype=np.uint8), qp, sps["height"], sps["width"],
                                      ec_part, reset_feature_memory)
        return {"x_hat": x_hat}


class VideoCodecHTL(VideoCodecHTS):
    """DMC with ModelStructure.HTL (video_model_ht.py:413-450 select DMCHTLProxy)."""
    PROXY = DMCHTLProxy


class VideoCodecLD(VideoCodecHTS):
    """The low-delay DMC (video_model_ld.py:273-308): one frame per call."""
    PROXY = DMCLDProxy
    FRAME_DELAY = 1  # g_frame_delay, video_model_ld.py:16


def video_codec(structure: str):
    """'hts', 'htl' or 'ld' -> the codec class (video_model_ht.py:413-424, video_model_ld.py:273-289)."""
    return {"hts": VideoCodecHTS, "htl": VideoCodecHTL, "ld": VideoCodecLD}[structure]


def calc_psnr(a, b, data_range=255.0):
    """src/utils/metrics.py:10-24."""
    mse = float(np.mean(np.square(a.astype(np.float64) - b.astype(np.float64))))
    if np.isnan(mse) or np.isinf(mse):
        return -999.9
    p = 10 * np.log10(data_range * data_range / mse) if mse > 1e-10 else 999.9
    return min(p, 99.9)


def yuv420_psnr(x_hat, yuv):
    """test_video.py:30-47: 444 recon -> 420 (avg-pooled chroma), PSNR 6:1:1."""
    x = x_hat.float() + 0.5
    y_rec = torch.clamp(x[:, :1] * 255, 0, 255)[0, 0].cpu().numpy()
    uv = torch.clamp(F.avg_pool2d(x[:, 1:], 2, 2) * 255, 0, 255)[0].cpu().numpy()
    y, u, v = yuv
    py, pu, pv = calc_psnr(y, y_rec), calc_psnr(u, uv[0]), calc_psnr(v, uv[1])
    return (6 * py + pu + pv) / 8, py, pu, pv


@dataclass
class SequenceResult:
    bitstream: bytes = b""
    bits: list = field(default_factory=list)
    psnr: list = field(default_factory=list)
    enc_time: list = field(default_factory=list)
    dec_time: list = field(default_factory=list)
    frame_types: list = field(default_factory=list)
    recon: list = field(default_factory=list)
    frame_num: int = 0
    height: int = 0
    width: int = 0

    def bpp(self):
        return sum(self.bits) / (self.frame_num * self.height * self.width)


def _is_intra(frame_idx, intra_period, frame_delay):
    """test_video.py:205-212: frame 0, every frame at intra_period 1, and
    frames at intra_period * k + 1 (never frame 1) otherwise."""
    if frame_idx == 0 or intra_period == 1:
        return True
    if intra_period > 1 and frame_idx != 1:
        if intra_period % frame_delay != 0:
            raise ValueError("intra_period must be a multiple of the frame delay")
        return frame_idx % intra_period == 1
    return False


def _input_of(frames, frame_delay, device):
    """get_src_frame (test_video.py:75-122): one frame on 3 channels, or a
    chunk of frame_delay frames on 3 * frame_delay channels."""
    if frame_delay == 1 and len(frames) == 1:
        return to_input(synthetic.frame_to_input(frames[0]), device)
    return to_input(synthetic.chunk_to_input(frames, frame_delay), device)


def encode_sequence(frames, i_codec, p_codec, qp_i, qp_p, reset_interval=32, intra_period=-1, device="cuda"):
    """frames: list of uint8 YUV420 tuples.  Returns SequenceResult with the
    bitstream, per-frame bits (the chunk's bits on its first frame) and frame
    types (0 I, 1 P), as test_video.py:243-250 accounts them."""
    h, w = frames[0][0].shape
    pad_r, pad_b = synthetic.get_padding_size(h, w, 16)
    delay = p_codec.FRAME_DELAY
    res = SequenceResult(frame_num=len(frames), height=h, width=w)
    out = io.BytesIO()
    sps_helper = stream_helper.SPSHelper()
    idx = 0
    while idx < len(frames):
        is_intra = _is_intra(idx, intra_period, delay)
        n = 1 if is_intra else min(delay, len(frames) - idx)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if is_intra:
            x = to_input(synthetic.frame_to_input(frames[idx]), device)
            enc = i_codec.compress(x, qp_i, pad_b, pad_r)
            p_codec.clear_dpb()
            p_codec.add_ref_feature_from_frame(enc["x_hat"])
            reset, qp = 0, qp_i
        else:
            reset = 1 if (reset_interval > 0 and (idx + delay) % reset_interval == 1) else 0
            x = _input_of(frames[idx:idx + n], delay, device)
            enc = p_codec.compress(x, qp_p, reset, pad_b, pad_r)
            qp = qp_p
        sps = {"sps_id": -1, "height": h, "width": w}
        sps_id, new = sps_helper.get_sps_id(sps)
        sps["sps_id"] = sps_id
        nb = stream_helper.write_sps(out, sps) if new else 0
        nb += stream_helper.write_ip(out, is_intra, sps_id, qp, enc["ec_parallel"], reset, enc["bit_stream"])
        torch.cuda.synchronize()
        res.enc_time.append(time.perf_counter() - t0)
        res.bits += [nb * 8] + [0] * (n - 1)
        res.frame_types += [0] if is_intra else [1] * n
        idx += n
    res.bitstream = out.getvalue()
    return res


def decode_sequence(bitstream, frames, i_codec, p_codec, res: SequenceResult | None = None, keep_recon=False):
    """Decodes the container; PSNR against the source frames.  Returns res;
    with keep_recon, res.recon holds a host copy of every decoded frame."""
    h, w = frames[0][0].shape
    res = res or SequenceResult(frame_num=len(frames), height=h, width=w)
    delay = p_codec.FRAME_DELAY
    f = io.BytesIO(bitstream)
    sps_helper = stream_helper.SPSHelper()
    done = 0
    while done < len(frames):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        header = stream_helper.read_header(f)
        while header["nal_type"] == stream_helper.NalType.NAL_SPS:
            sps_helper.add_sps_by_id(stream_helper.read_sps_remaining(f, header["sps_id"]))
            header = stream_helper.read_header(f)
        sps = sps_helper.get_sps_by_id(header["sps_id"])
        qp, ec_part, reset, payload = stream_helper.read_ip_remaining(f)
        if header["nal_type"] == stream_helper.NalType.NAL_I:
            x_hat = i_codec.decompress(payload, sps, qp, ec_part)["x_hat"]
            p_codec.clear_dpb()
            p_codec.add_ref_feature_from_frame(x_hat, apply_feature_adaptor=False)
            n = 1
        else:
            x_hat = p_codec.decompress(payload, sps, qp, ec_part, reset)["x_hat"]
            n = min(delay, len(frames) - done)
        recon = x_hat if isinstance(x_hat, (list, tuple)) else [x_hat]
        torch.cuda.synchronize()
        res.dec_time.append(time.perf_counter() - t0)
        for i in range(n):
            r = recon[i][:, :, :h, :w]
            res.psnr.append(yuv420_psnr(r, frames[done + i])[0])
            if keep_recon:
                res.recon.append(r.float().cpu())
        done += n
    return res
