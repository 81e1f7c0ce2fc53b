"""State-dict schema and seeded random initialisation for DCVC-UF models.

The key names and tensor shapes are the reference model's ``state_dict()``
interface, which ``set_param`` consumes:

* ``DepthConvBlock``           src/layers/layers.py:128-159
* ``ResidualBlockUpsample``    src/layers/layers.py:162-173 (``SubpelConv2x`` :92-103)
* ``ResidualBlockWithStride2`` src/layers/layers.py:176-188
* ``DMCI``                     src/models/image_model.py:21-148
* ``DMC`` (HT-S / HT-L)        src/models/video_model_ht.py:26-351
* ``DMC`` (low delay)          src/models/video_model_ld.py:24-216
* ``BitEstimator``             src/models/entropy_models.py:78-89

Initialisation follows the reference's scheme (src/models/common_model.py:49-54):
xavier-normal (gain 1) conv weights, zero conv biases, ones for the per-QP
quantisation scales, N(0, 0.01) for the bit-estimator parameters.  The random
stream is numpy PCG64, not torch's, so the values are this project's own seeded
weights shared by every path (GPU kernels, CPU oracle), not a bit-copy of a
torch-initialised reference model.
"""
from __future__ import annotations

import math
import re
from collections import OrderedDict

import numpy as np

QP_NUM = 64

# --------------------------------------------------------------------------
# schema builders: each returns a list of (key, shape, kind)
#   kind: "conv_w", "conv_b", "ones", "normal001"
# --------------------------------------------------------------------------


def _conv(prefix, cout, cin, k=1, groups=1, bias=True):
    out = [(prefix + "weight", (cout, cin // groups, k, k), "conv_w")]
    if bias:
        out.append((prefix + "bias", (cout,), "conv_b"))
    return out


def dcb_schema(prefix, in_ch, out_ch, dcb2=False, force_adaptor=False):
    """DepthConvBlock parameters (layers.py:128-150)."""
    r = 2 if dcb2 else 1
    s = []
    if in_ch != out_ch or force_adaptor:
        s += _conv(prefix + "adaptor.", out_ch, in_ch)
    s += _conv(prefix + "dc.0.", out_ch // r, out_ch)
    s += _conv(prefix + "dc.2.", out_ch // r, out_ch // r, k=3, groups=out_ch // r)
    s += _conv(prefix + "dc.3.", out_ch, out_ch // r)
    s += _conv(prefix + "ffn.0.", out_ch * 4 // r, out_ch)
    s += _conv(prefix + "ffn.2.", out_ch, out_ch // r)
    return s


def rbu_schema(prefix, in_ch, out_ch, force_bias=False, dcb2=False):
    """ResidualBlockUpsample: SubpelConv2x(k=1) + DepthConvBlock (layers.py:92-173)."""
    s = _conv(prefix + "up.conv.0.", out_ch * 4, in_ch, bias=force_bias)
    s += dcb_schema(prefix + "conv.", out_ch, out_ch, dcb2=dcb2)
    return s


def rbs2_schema(prefix, in_ch, out_ch, dcb2=False):
    """ResidualBlockWithStride2: 1x1 over pixel_unshuffle(x, 2) + DCB (layers.py:176-188)."""
    s = _conv(prefix + "down.", out_ch, in_ch * 4)
    s += dcb_schema(prefix + "conv.", out_ch, out_ch, dcb2=dcb2)
    return s


def bit_estimator_schema(prefix, channel):
    return [
        (prefix + "h", (QP_NUM, channel, 4), "normal001"),
        (prefix + "b", (QP_NUM, channel, 4), "normal001"),
        (prefix + "a", (QP_NUM, channel, 3), "normal001"),
    ]


def dmci_schema():
    """DMCI (image_model.py:126-148): g_ch_src=192, enc/dec=384, y=256, z=128."""
    src, ed, y, z = 192, 384, 256, 128
    s = []
    s += bit_estimator_schema("bit_estimator_z.", z)
    s += dcb_schema("enc.enc_1.", src, ed)
    for i in range(6):
        s += dcb_schema(f"enc.enc_2.{i}.", ed, ed)
    s += _conv("enc.enc_2.6.", y, ed, k=3)
    s += dcb_schema("hyper_enc.conv.0.", y, z)
    s += rbs2_schema("hyper_enc.conv.1.", z, z)
    s += rbs2_schema("hyper_enc.conv.2.", z, z)
    s += rbu_schema("hyper_dec.conv.0.", z, z)
    s += rbu_schema("hyper_dec.conv.1.", z, z)
    s += dcb_schema("hyper_dec.conv.2.", z, y)
    s += dcb_schema("y_prior_fusion.conv.0.", y, y * 2)
    s += dcb_schema("y_prior_fusion.conv.1.", y * 2, y * 2)
    s += dcb_schema("y_prior_fusion.conv.2.", y * 2, y * 2)
    s += _conv("y_prior_fusion.conv.3.", y * 2, y * 2)
    s += _conv("y_spatial_prior_reduction.", y, y * 2)
    for i in (1, 2, 3):
        s += dcb_schema(f"y_spatial_prior_adaptor_{i}.", y * 2, y * 2, force_adaptor=True)
    for i in range(3):
        s += dcb_schema(f"y_spatial_prior.conv.{i}.", y * 2, y * 2)
    s += _conv("y_spatial_prior.conv.3.", y * 2, y * 2)
    s += rbu_schema("dec.dec_1.0.", y, ed)
    for i in range(1, 13):
        s += dcb_schema(f"dec.dec_1.{i}.", ed, ed)
    s += dcb_schema("dec.dec_2.", ed, src)
    s += [("q_scale_enc", (QP_NUM, ed), "ones"), ("q_scale_dec", (QP_NUM, ed), "ones"),
          ("q_scale_y_enc", (QP_NUM, y), "ones"), ("q_scale_y_dec", (QP_NUM, y), "ones")]
    return s


def dmc_hts_schema():
    """DMC with ModelStructure.HTS (video_model_ht.py:26-351)."""
    fd, src_i, src, y, z, d, m, rc = 8, 192, 192 * 8, 256, 128, 512, 512, 256
    s = []
    s += bit_estimator_schema("bit_estimator_z.", z)
    s += dcb_schema("feature_adaptor_i.conv.0.", src_i, m, dcb2=True)
    for i in range(1, 4):
        s += dcb_schema(f"feature_adaptor_i.conv.{i}.", m, m, dcb2=True)
    s += dcb_schema("feature_adaptor_m.conv.0.", m + d, m, dcb2=True)
    for i in range(1, 6):
        s += dcb_schema(f"feature_adaptor_m.conv.{i}.", m, m, dcb2=True)
    s += dcb_schema("feature_extractor.conv.0.", m, d, dcb2=True)
    for i in range(1, 5):
        s += dcb_schema(f"feature_extractor.conv.{i}.", d, d, dcb2=True)
    s += dcb_schema("encoder.conv1.0.", src + d, d, dcb2=True)
    for i in range(1, 6):
        s += dcb_schema(f"encoder.conv1.{i}.", d, d, dcb2=True)
    s += _conv("encoder.down.", y, d, k=3)
    s += dcb_schema("hyper_encoder.conv.0.", y, y)
    s += rbs2_schema("hyper_encoder.conv.1.", y, y)
    s += rbs2_schema("hyper_encoder.conv.2.", y, z)
    s += rbu_schema("hyper_decoder.conv.0.", z, y)
    s += rbu_schema("hyper_decoder.conv.1.", y, y)
    s += dcb_schema("hyper_decoder.conv.2.", y, y)
    s += rbs2_schema("temporal_prior_encoder.conv.", d, y * 2)
    for i in range(3):
        s += dcb_schema(f"y_prior_fusion.conv.{i}.", y * 3, y * 3)
    s += _conv("y_prior_fusion.conv.3.", y * 3, y * 3)
    s += _conv("y_spatial_prior_reduction.", y, y * 3)
    for i in (1, 2, 3):
        s += dcb_schema(f"y_spatial_prior_adaptor_{i}.", y * 2, y * 2, force_adaptor=True)
    for i in range(3):
        s += dcb_schema(f"y_spatial_prior.conv.{i}.", y * 2, y * 2)
    s += _conv("y_spatial_prior.conv.3.", y, y * 2)
    s += _conv("decoder.up.conv.0.", d * 4, y, bias=False)
    s += dcb_schema("decoder.conv1.0.", d * 2, d, dcb2=True)
    for i in range(1, 7):
        s += dcb_schema(f"decoder.conv1.{i}.", d, d, dcb2=True)
    for i in range(fd // 2):
        s += dcb_schema(f"recon_head.conv1.{i}.0.", d, d)
    for i in range(fd):
        s += dcb_schema(f"recon_head.conv2.{i}.0.", d, rc)
        s += dcb_schema(f"recon_head.conv2.{i}.1.", rc, rc)
        s += dcb_schema(f"recon_head.conv2.{i}.2.", rc, rc)
        s += _conv(f"recon_head.conv2.{i}.3.", src_i, rc)
    s += [("q_encoder", (QP_NUM, d), "ones"), ("q_decoder", (QP_NUM, d), "ones"),
          ("q_feature", (QP_NUM, d), "ones")]
    return s


def dmc_htl_schema():
    """DMC with ModelStructure.HTL (video_model_ht.py:26-351, the ``not is_hts``
    branches): ratio-1 DepthConvBlocks, a 3x3 biased sub-pixel decoder up
    conv, force_bias hyper-decoder up convs, a scales+means spatial prior and
    a 5-block recon head per frame."""
    fd, src_i, src, y, z, d, m, rc = 8, 192, 192 * 8, 256, 128, 512, 512, 256
    s = []
    s += bit_estimator_schema("bit_estimator_z.", z)
    s += dcb_schema("feature_adaptor_i.conv.0.", src_i, m)
    for i in range(1, 3):
        s += dcb_schema(f"feature_adaptor_i.conv.{i}.", m, m)
    s += dcb_schema("feature_adaptor_m.conv.0.", m + d, m)
    for i in range(1, 10):
        s += dcb_schema(f"feature_adaptor_m.conv.{i}.", m, m)
    s += dcb_schema("feature_extractor.conv.0.", m, d)
    s += dcb_schema("feature_extractor.conv.1.", d, d)
    s += dcb_schema("encoder.conv1.0.", src + d, d)
    for i in range(1, 7):
        s += dcb_schema(f"encoder.conv1.{i}.", d, d)
    s += _conv("encoder.down.", y, d, k=3)
    s += dcb_schema("hyper_encoder.conv.0.", y, y)
    s += rbs2_schema("hyper_encoder.conv.1.", y, y)
    s += rbs2_schema("hyper_encoder.conv.2.", y, z)
    s += rbu_schema("hyper_decoder.conv.0.", z, y, force_bias=True)
    s += rbu_schema("hyper_decoder.conv.1.", y, y, force_bias=True)
    s += dcb_schema("hyper_decoder.conv.2.", y, y)
    s += rbs2_schema("temporal_prior_encoder.conv.", d, y * 2)
    for i in range(3):
        s += dcb_schema(f"y_prior_fusion.conv.{i}.", y * 3, y * 3)
    s += _conv("y_prior_fusion.conv.3.", y * 3, y * 3)
    s += _conv("y_spatial_prior_reduction.", y, y * 3)
    for i in (1, 2, 3):
        s += dcb_schema(f"y_spatial_prior_adaptor_{i}.", y * 2, y * 2, force_adaptor=True)
    for i in range(3):
        s += dcb_schema(f"y_spatial_prior.conv.{i}.", y * 2, y * 2)
    s += _conv("y_spatial_prior.conv.3.", y * 2, y * 2)
    s += _conv("decoder.up.conv.0.", d * 4, y, k=3)
    s += dcb_schema("decoder.conv1.0.", d * 2, d)
    for i in range(1, 11):
        s += dcb_schema(f"decoder.conv1.{i}.", d, d)
    for i in range(fd):
        s += dcb_schema(f"recon_head.conv.{i}.0.", d, rc)
        for j in range(1, 5):
            s += dcb_schema(f"recon_head.conv.{i}.{j}.", rc, rc)
        s += _conv(f"recon_head.conv.{i}.5.", src_i, rc)
    s += [("q_encoder", (QP_NUM, d), "ones"), ("q_decoder", (QP_NUM, d), "ones"),
          ("q_feature", (QP_NUM, d), "ones")]
    return s


def dmc_ld_schema():
    """DMC low-delay (video_model_ld.py:24-216): one frame per call, dcb2
    blocks throughout, y = 128, d = m = 256 channels, a 2x checkerboard."""
    src, y, z, d, m = 192, 128, 128, 256, 256
    s = []
    s += bit_estimator_schema("bit_estimator_z.", z)
    s += dcb_schema("feature_adaptor_i.conv.0.", src, m, dcb2=True)
    for i in range(1, 4):
        s += dcb_schema(f"feature_adaptor_i.conv.{i}.", m, m, dcb2=True)
    s += dcb_schema("feature_adaptor_m.conv.0.", m + d, m, dcb2=True)
    for i in range(1, 4):
        s += dcb_schema(f"feature_adaptor_m.conv.{i}.", m, m, dcb2=True)
    for i in range(5):
        s += dcb_schema(f"feature_extractor.conv.{i}.", m, m, dcb2=True)
    s += dcb_schema("encoder.conv1.0.", src + m, d, dcb2=True)
    s += dcb_schema("encoder.conv1.1.", d, d, dcb2=True)
    s += dcb_schema("encoder.conv2.", d, d, dcb2=True)
    s += _conv("encoder.down.", y, d, k=3)
    s += dcb_schema("hyper_encoder.conv.0.", y, z, dcb2=True)
    s += rbs2_schema("hyper_encoder.conv.1.", z, z, dcb2=True)
    s += rbs2_schema("hyper_encoder.conv.2.", z, z, dcb2=True)
    s += rbu_schema("hyper_decoder.conv.0.", z, z, dcb2=True)
    s += rbu_schema("hyper_decoder.conv.1.", z, z, dcb2=True)
    s += dcb_schema("hyper_decoder.conv.2.", z, y, dcb2=True)
    s += rbs2_schema("temporal_prior_encoder.conv.", m, y * 2, dcb2=True)
    for i in range(3):
        s += dcb_schema(f"y_prior_fusion.conv.{i}.", y * 3, y * 3, dcb2=True)
    s += _conv("y_prior_fusion.conv.3.", y * 3, y * 3)
    s += dcb_schema("y_spatial_prior.conv.0.", y * 4, y * 2, dcb2=True)
    s += dcb_schema("y_spatial_prior.conv.1.", y * 2, y * 2, dcb2=True)
    s += _conv("y_spatial_prior.conv.2.", y, y * 2)
    s += _conv("decoder.up.conv.0.", d * 4, y, bias=False)
    s += dcb_schema("decoder.conv1.0.", d + m, d, dcb2=True)
    s += dcb_schema("decoder.conv1.1.", d, d, dcb2=True)
    s += dcb_schema("decoder.conv1.2.", d, d, dcb2=True)
    s += _conv("decoder.conv2.", d, d)
    for i in range(3):
        s += dcb_schema(f"recon_head.conv.{i}.", d, d, dcb2=True)
    s += _conv("recon_head.head.", src, d)
    s += [("q_encoder", (QP_NUM, d), "ones"), ("q_decoder", (QP_NUM, d), "ones"),
          ("q_feature", (QP_NUM, y * 2), "ones")]
    return s


SCHEMAS = {"dmci": dmci_schema, "hts": dmc_hts_schema, "htl": dmc_htl_schema, "ld": dmc_ld_schema}


def _xavier_std(shape, gain=1.0):
    # torch.nn.init._calculate_fan_in_and_fan_out for conv weights
    receptive = int(np.prod(shape[2:])) if len(shape) > 2 else 1
    fan_in = shape[1] * receptive
    fan_out = shape[0] * receptive
    return gain * math.sqrt(2.0 / float(fan_in + fan_out))


RESIDUAL_BRANCH_GAIN = 0.25
RECON_PROJECTION_GAIN = 0.1


def _is_residual_branch_out(key):
    return key.endswith("dc.3.weight") or key.endswith("ffn.2.weight")


def _is_recon_projection(key):
    if key == "recon_head.head.weight":
        return True
    if key.startswith("dec.dec_2.") and key.endswith(".weight"):
        return any(s in key for s in ("adaptor.", "dc.3.", "ffn.2."))
    if key.startswith("recon_head.conv2.") and key.endswith(".3.weight"):
        return True
    return re.match(r"recon_head\.conv\.\d+\.5\.weight$", key) is not None


def init_state_dict(model: str, seed: int = 0, q_jitter: float = 0.0, stable: bool = True):
    """Return an OrderedDict of float32 numpy arrays for ``model`` in {'dmci','hts','htl','ld'}.

    ``stable`` (default) scales the output conv of every residual branch
    (``dc.3``, ``ffn.2``) by 0.25 and the reconstruction projections by 0.1.
    With the reference's plain gain-1 init each DepthConvBlock multiplies the
    activation std by ~1.6, so a 13-block decoder reaches ~1e6 and FP16
    overflows; the scaled init keeps activations O(1) like a trained model.
    ``stable=False`` gives the plain reference scheme.

    ``q_jitter`` > 0 replaces the all-ones per-QP scales with 1 + U(-j, j) so
    that tests can see the per-QP vectors being selected; 0 keeps the
    reference initialisation.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    sd = OrderedDict()
    for key, shape, kind in SCHEMAS[model]():
        if kind == "conv_w":
            v = rng.standard_normal(shape, dtype=np.float64) * _xavier_std(shape)
        elif kind == "conv_b":
            v = np.zeros(shape, dtype=np.float64)
        elif kind == "ones":
            v = np.ones(shape, dtype=np.float64)
            if q_jitter > 0:
                v = v + rng.uniform(-q_jitter, q_jitter, size=shape)
        elif kind == "normal001":
            v = rng.standard_normal(shape, dtype=np.float64) * 0.01
        else:  # pragma: no cover
            raise ValueError(kind)
        if stable and kind == "conv_w":
            if _is_residual_branch_out(key):
                v = v * RESIDUAL_BRANCH_GAIN
            if _is_recon_projection(key):
                v = v * RECON_PROJECTION_GAIN
        sd[key] = v.astype(np.float32)
    return sd


QP_ENCODER_SIDE = ("q_encoder", "q_feature", "q_scale_enc", "q_scale_y_enc")
QP_DECODER_SIDE = ("q_decoder", "q_scale_dec", "q_scale_y_dec")


def qp_ladder(sd, lo: float = 0.5, hi: float = 2.0):
    """Per-QP scale rows that differ by qp, in place of the reference's
    all-ones initialisation (video_model_ht.py:348-350, image_model.py):
    encoder-side tables (features and latents multiplied before
    quantisation) rise geometrically from `lo` at qp 0 to `hi` at qp 63, and
    decoder-side tables are their reciprocals, the shape a trained model's
    quantisation ladder has.  Used where a qp sweep must exercise the
    per-QP selection (the precision form's qp rows); the default weights
    keep the reference's ones."""
    out = OrderedDict(sd)
    g = np.exp(np.linspace(np.log(lo), np.log(hi), QP_NUM))
    for k, v in sd.items():
        base = k.split(".")[-1]
        if base in QP_ENCODER_SIDE:
            out[k] = np.repeat(g[:, None], v.shape[1], axis=1).astype(np.float32)
        elif base in QP_DECODER_SIDE:
            out[k] = np.repeat(1.0 / g[:, None], v.shape[1], axis=1).astype(np.float32)
    return out


def perturb_biases(sd, seed: int = 1, scale: float = 0.02):
    """Give every conv bias a small seeded random value.

    The reference initialisation zeroes all biases, which leaves every bias
    add of a kernel untested; tests call this to make bias handling visible.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    for k in sd:
        if k.endswith(".bias"):
            sd[k] = (rng.standard_normal(sd[k].shape) * scale).astype(np.float32)
    return sd


def to_torch(sd, dtype=None, device="cpu"):
    import torch
    out = OrderedDict()
    for k, v in sd.items():
        t = torch.from_numpy(np.ascontiguousarray(v))
        if dtype is not None and k not in ():
            t = t.to(dtype)
        out[k] = t.to(device)
    return out


def round_to_half(sd):
    """fp16-round every array (the GPU path stores weights in fp16)."""
    return OrderedDict((k, v.astype(np.float16).astype(np.float32)) for k, v in sd.items())


def param_count(model: str) -> int:
    return int(sum(np.prod(shape) for _, shape, _ in SCHEMAS[model]()))
