"""CPU oracle for the DCVC-UF inference hot path (test infrastructure only).

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it.

It restates, in float32 (or float64) torch on the CPU, what the reference's
CUDA proxies compute.  The network structure follows the reference Python
forward (``src/layers/layers.py``, ``src/models/*.py``); the integer
boundary follows the proxy kernels, whose semantics differ from the Python
forward (``elementwise/stream.cu``):

* ``y_q`` is zeroed where the predicted scale is <= ``skip_thres`` and clamped
  to [-128, 127]                                    (stream.cu:581-592)
* rounding is C ``roundf`` (half away from zero), the function the proxy
  kernels call on ``at::Half`` values                (stream.cu:587, :873)
* ``z_hat`` is clamped to [-64, 63]                   (stream.cu:873-878)
* scale -> CDF index uses the truncated constants of def_const.h:6-11.  It is
  evaluated in FP32 semantics through a 126-entry threshold table computed in
  float64 (``scale_index_thresholds``), so host and device agree bit-for-bit.

What pins each function (the check in tests/ that fails on a plausible
mistake; none of them retypes the oracle's formula):

* scale_index_thresholds / scale_to_index: float64 formula on every finite
  fp16 scale, random float32 and edges, the entropy model's scale table, and
  the C++ table (test_oracle_boundary.py); special cases incl. the top-end
  126 of the truncated constants (test_oracle_pins.py).
* round_half_away, round_z, quantize_masked, build_index: hand-derived
  special cases of the stream.cu semantics (ties, clamps, skip threshold,
  NHWC order, symbol packing) (test_oracle_boundary.py, test_oracle_pins.py).
* compact / recover: stability and inverse invariants incl. empty and full
  inputs (test_oracle_pins.py).  compute_ec_parallel: closed form.
* mask_4x: the reference's get_mask_4x output (golden vectors,
  test_oracle_golden.py) and the partition / permutation pattern.
* conv / wsilu / dcb / rbs2 / subpel / rbu and every DMCI / HT-S submodule
  (encoder, hyper codec, priors, spatial priors, decoder, recon head,
  feature adaptors / extractor): the reference modules' own forward on the
  same seeded weights (golden vectors, test_oracle_golden.py).
* dmci_compress / dmci_decompress / hts_compress / hts_decompress (step
  order, y_hat accumulation, context update): composed of the pinned parts;
  each decoder, fed the encoder's compacted symbols, asks for exactly the
  encoder's CDF indexes (DMCI: in each of its four steps, quarter view) and
  rebuilds the encoder's y_hat and reconstruction exactly
  (test_oracle_pins.py).
* storage_mode(fp16=True): the same functions with fp16 rounding at the GPU
  library's storage points.  With the rounding switched off it is the FP32
  restatement bit for bit (test_oracle_pins.py); where each rounding sits is
  pinned on the GPU, block by block, by bit-exact agreement of the GPU
  DepthConvBlock with it on nearly every element (tests/test_gpu_parity.py).
* hts/intra/htl *_f32 step restatements: agree with the torch masked steps
  on symbols, conditions and y_hat (test_oracle_pins.py, HT-S and HT-L
  forms; the HT-L form covers the DMCI quarter steps it is built on).
* ld_* submodules and mask_2x: the reference low-delay modules and
  get_mask_2x on the same seeded weights (golden vectors,
  test_oracle_golden.py).  ld_compress / ld_decompress: the decoder rebuilds
  the encoder's indexes, y_hat, reconstruction and context exactly
  (test_oracle_pins.py).  ld_*_f32 step restatements: against the torch
  masked halves (test_oracle_pins.py).
* htl_* submodules: the reference HT-L modules on the same seeded weights
  (golden vectors, test_oracle_golden.py).  htl_compress / htl_decompress:
  composed of the pinned parts; the decoder rebuilds the encoder's indexes
  in each of the four steps, y_hat, reconstruction and context exactly
  (test_oracle_pins.py).
* ref_gemm_epilogue / ref_dw3x3: torch float64 conv2d / sigmoid forms
  (test_oracle_pins.py).
"""
from __future__ import annotations

import contextlib
import math
from dataclasses import dataclass, field

import numpy as np
import torch
import torch.nn.functional as F

# def_const.h:6-11
SCALE_MIN = np.float32(0.11)
SCALE_MAX = np.float32(16.0)
LOG_SCALE_MIN = np.float32(-2.2073)
LOG_SCALE_MAX = np.float32(2.7726)
SCALE_LEVEL = 128
LOG_SCALE_STEP = np.float32((LOG_SCALE_MAX - LOG_SCALE_MIN) / np.float32(SCALE_LEVEL - 1))
LOG_SCALE_STEP_RECIP = np.float32(np.float32(1.0) / LOG_SCALE_STEP)
MIN_SYMBOLS_PER_STREAM = 32768   # def_const.h:18
MAX_EC_PARALLEL = 8              # py_rans.h:15
QP_NUM = 64


# ---------------------------------------------------------------------------
# integer boundary
# ---------------------------------------------------------------------------

def scale_index_exact(s):
    """floor((ln(clamp(s,.11,16)) - LOG_SCALE_MIN) * STEP_RECIP) in float64.

    The real-number value of stream.cu:77-87 with the float32 constants.
    """
    s = np.asarray(s, dtype=np.float64)
    s = np.clip(s, np.float64(SCALE_MIN), np.float64(SCALE_MAX))
    v = (np.log(s) - np.float64(LOG_SCALE_MIN)) * np.float64(LOG_SCALE_STEP_RECIP)
    return np.floor(v).astype(np.int64)


def scale_index_thresholds():
    """float32 thresholds T[k], k=1..127: index(s) = #{k : s >= T[k]}.

    T[k] is the smallest float32 s whose float64 index is >= k; computed by a
    float32 bisection on ``scale_index_exact``.  Entries beyond the clamp
    (SCALE_MAX) are +inf.  Pinned by tests/test_oracle_boundary.py against
    the float64 formula on every float16 value and on random float32 values.
    """
    th = np.full(SCALE_LEVEL, np.inf, dtype=np.float32)  # th[0] unused
    lo_bits = np.float32(SCALE_MIN).view(np.int32)
    for k in range(1, SCALE_LEVEL):
        approx = math.exp(k / float(LOG_SCALE_STEP_RECIP) + float(LOG_SCALE_MIN))
        if approx > float(SCALE_MAX) * 1.01:
            continue
        if scale_index_exact(np.float32(SCALE_MAX)) < k:
            continue
        # bisection over positive float32 bit patterns
        lo = lo_bits
        hi = np.float32(SCALE_MAX).view(np.int32)
        while lo < hi:
            mid = (int(lo) + int(hi)) // 2
            v = np.int32(mid).view(np.float32)
            if scale_index_exact(v) >= k:
                hi = mid
            else:
                lo = mid + 1
        th[k] = np.int32(lo).view(np.float32)
    return th


_THRESHOLDS = None


def scale_to_index(s):
    """Index table lookup (FP32 semantics of stream.cu:77-87, 104-110)."""
    global _THRESHOLDS
    if _THRESHOLDS is None:
        _THRESHOLDS = scale_index_thresholds()
    s = np.asarray(s, dtype=np.float32)
    idx = np.zeros(s.shape, dtype=np.int64)
    for k in range(1, SCALE_LEVEL):
        if np.isinf(_THRESHOLDS[k]):
            break
        idx += (s >= _THRESHOLDS[k])
    return idx


def round_half_away(x):
    """C roundf semantics (stream.cu:587 calls round() on a float).

    Computed exactly as trunc(x) + sign(x) * [|x - trunc(x)| >= 0.5]; the
    frequent floor(|x| + 0.5) shortcut is wrong for 0.49999997f.
    """
    if isinstance(x, torch.Tensor):
        t = torch.trunc(x)
        return t + torch.sign(x) * ((x - t).abs() >= 0.5).to(x.dtype)
    x = np.asarray(x)
    t = np.trunc(x)
    return t + np.sign(x) * (np.abs(x - t) >= 0.5).astype(x.dtype)


def round_z(z):
    """round_z_kernel (stream.cu:862-884): clamp(round(z), -64, 63)."""
    return torch.clamp(round_half_away(z), -64, 63)


def quantize_masked(y, means, scales, mask, skip_thres):
    """One masked quantisation step (stream.cu:581-594).

    Returns (y_q, y_hat) where masked-out positions are 0.  ``scales`` is the
    unmasked scale map; ``y_q`` is forced to 0 where scale <= skip_thres.
    """
    m = mask.to(y.dtype)
    means_hat = means * m
    y_res = (y - means_hat) * m
    y_q = round_half_away(y_res)
    y_q = y_q * ((scales * m) > skip_thres).to(y.dtype)
    y_q = torch.clamp(y_q, -128, 127)
    return y_q, y_q + means_hat


def build_index(y_q, scales, skip_thres):
    """build_index_enc/dec (stream.cu:89-173) over NCHW inputs.

    Returns (sym16, idx, cond) flattened in NHWC-linear order: sym16 =
    (y_q << 8) + idx as int16; cond = scale > skip_thres.
    """
    s = scales.permute(0, 2, 3, 1).reshape(-1).numpy().astype(np.float32)
    idx = scale_to_index(s)
    cond = s > np.float32(skip_thres)
    sym16 = None
    if y_q is not None:
        q = y_q.permute(0, 2, 3, 1).reshape(-1).numpy().astype(np.int64)
        sym16 = ((q << 8) + idx).astype(np.int16)
    return sym16, idx.astype(np.uint8), cond


def compact(values, cond):
    """Stable stream compaction by cond (stream.cu:175-340)."""
    return np.asarray(values)[np.asarray(cond, dtype=bool)]


def recover(decoded, cond, n):
    """conditional_recover_with_type_conversion (stream.cu:359-383)."""
    out = np.zeros(n, dtype=np.int8)
    out[np.asarray(cond, dtype=bool)] = np.asarray(decoded, dtype=np.int8)[: int(np.sum(cond))]
    return out


def compute_ec_parallel(symbol_count):
    """dmc_common.cpp:31-35."""
    n = symbol_count // MIN_SYMBOLS_PER_STREAM
    return max(1, min(MAX_EC_PARALLEL, n))


# ---------------------------------------------------------------------------
# masks (common_model.py:96-100, 174-195; dmci_proxy.cpp:678-699)
# ---------------------------------------------------------------------------

def _one_mask(micro, H, W):
    m = torch.tensor(micro, dtype=torch.bool)
    m = m.repeat((H + 1) // 2, (W + 1) // 2)
    return m[None, None, :H, :W]


def mask_4x(C, H, W):
    m0 = _one_mask(((1, 0), (0, 0)), H, W)
    m1 = _one_mask(((0, 1), (0, 0)), H, W)
    m2 = _one_mask(((0, 0), (1, 0)), H, W)
    m3 = _one_mask(((0, 0), (0, 1)), H, W)
    q = C // 4

    def cat(a, b, c, d):
        return torch.cat([a.expand(1, q, H, W), b.expand(1, q, H, W),
                          c.expand(1, q, H, W), d.expand(1, q, H, W)], dim=1)
    return [cat(m0, m1, m2, m3), cat(m3, m2, m1, m0), cat(m2, m3, m0, m1), cat(m1, m0, m3, m2)]


# ---------------------------------------------------------------------------
# layers (layers.py), FP32 on NCHW tensors
# ---------------------------------------------------------------------------

def _nchw(x):
    """x with standard NCHW strides.  Tensors with size-1 dims count as
    contiguous under several stride patterns, and the CPU convolution picks
    its summation order from the strides; normalising them makes the result
    depend on the values only (the encoder and decoder sides then agree bit
    for bit)."""
    std, acc = [], 1
    for d in reversed(x.shape):
        std.append(acc)
        acc *= d
    if tuple(reversed(std)) == tuple(x.stride()):
        return x
    return x.clone(memory_format=torch.contiguous_format).as_strided(x.shape, tuple(reversed(std)))


# ---------------------------------------------------------------------------
# storage modes
#
# The default is the FP32 restatement.  Under ``storage_mode(fp16=True)`` the
# oracle rounds to fp16 wherever the GPU library stores a kernel's output
# (``st16``): every GEMM / conv / depthwise output after its fused epilogue,
# each quantiser's y_hat, and the per-channel multiplies run as their own
# kernels.  It also folds the depthwise bias into dc.3 in float64, as the
# library does at load.  Arithmetic inside each kernel stays float32, so this
# mode isolates what fp16 activation storage alone changes.  Under
# ``acc64=True`` every convolution accumulates in float64 and rounds once to
# float32: a second summation order, used to measure how far two
# implementations that differ only in fp32 summation order drift apart.
# ---------------------------------------------------------------------------

_MODE = {"fp16": False, "acc64": False}


@contextlib.contextmanager
def storage_mode(fp16=False, acc64=False):
    old = dict(_MODE)
    _MODE.update(fp16=bool(fp16), acc64=bool(acc64))
    try:
        yield
    finally:
        _MODE.update(old)


def st16(x):
    """A stored activation: fp16-rounded under storage_mode(fp16=True)."""
    return x.half().float() if _MODE["fp16"] else x


def _conv2d(x, w, b, stride=1, padding=0, groups=1):
    if _MODE["acc64"]:
        return F.conv2d(x.double(), w.double(), None if b is None else b.double(),
                        stride=stride, padding=padding, groups=groups).to(x.dtype)
    return F.conv2d(x, w.to(x.dtype), None if b is None else b.to(x.dtype),
                    stride=stride, padding=padding, groups=groups)


def conv(sd, prefix, x, stride=1, padding=0, groups=1):
    x = _nchw(x)
    return _conv2d(x, sd[prefix + "weight"], sd.get(prefix + "bias"), stride=stride, padding=padding, groups=groups)


def conv16(sd, prefix, x, stride=1, padding=0, groups=1):
    """A convolution launched on its own: its output is stored."""
    return st16(conv(sd, prefix, x, stride=stride, padding=padding, groups=groups))


def wsilu(x):
    """layers.py:106-111."""
    return torch.sigmoid(4.0 * x) * x


def dcb(sd, p, x, shortcut=False, quant=None):
    """DepthConvBlock.forward (layers.py:152-159) with the proxy's quant tail.

    ``quant`` multiplies the block output per channel, the fusion the proxy
    applies in conv1x1_bias_shortcut_with_quant (layers_proxy.cpp:92-94); in
    the Python model it is the ``feature * quant_step`` right after the block.
    """
    if _MODE["fp16"]:
        return _dcb_fp16(sd, p, x, shortcut, quant)
    if (p + "adaptor.weight") in sd:
        x = conv(sd, p + "adaptor.", x)
    c = sd[p + "dc.2.weight"].shape[0]
    t = wsilu(conv(sd, p + "dc.0.", x))
    t = conv(sd, p + "dc.2.", t, padding=1, groups=c)
    out = conv(sd, p + "dc.3.", t) + x
    t = wsilu(conv(sd, p + "ffn.0.", out))
    t = t[:, 0::4] + t[:, 1::4] + t[:, 2::4] + t[:, 3::4]
    out = conv(sd, p + "ffn.2.", t) + out
    if shortcut:
        out = out + x
    if quant is not None:
        out = out * quant.to(out.dtype).reshape(1, -1, 1, 1)
    return out


def _dcb_fp16(sd, p, x, shortcut, quant):
    """DepthConvBlock with the GPU library's storage points: the adaptor,
    dc.0 + WSiLU, the bias-free depthwise conv, dc.3 with the folded bias and
    the shortcut, the quad GEMM's WSiLU chunk sum, and ffn.2 with its tail
    each store fp16 (layers_proxy.cpp:71-100, 160-206: the same launches)."""
    if (p + "adaptor.weight") in sd:
        x = conv16(sd, p + "adaptor.", x)
    w_dw = sd[p + "dc.2.weight"]
    t = st16(wsilu(conv(sd, p + "dc.0.", x)))
    t = st16(_conv2d(_nchw(t), w_dw, None, padding=1, groups=w_dw.shape[0]))
    w3 = sd[p + "dc.3.weight"]
    b3, b_dw = sd.get(p + "dc.3.bias"), sd.get(p + "dc.2.bias")
    fold = torch.zeros(w3.shape[0], dtype=torch.float64) if b3 is None else b3.double().clone()
    if b_dw is not None:
        fold = fold + w3.double()[:, :, 0, 0] @ b_dw.double()
    out = st16(_conv2d(_nchw(t), w3, fold.float()) + x)
    h = wsilu(conv(sd, p + "ffn.0.", out))
    t = st16(h[:, 0::4] + h[:, 1::4] + h[:, 2::4] + h[:, 3::4])
    o = conv(sd, p + "ffn.2.", t) + out
    if shortcut:
        o = o + x
    if quant is not None:
        o = o * quant.to(o.dtype).reshape(1, -1, 1, 1)
    return st16(o)


def rbs2(sd, p, x, shortcut):
    """ResidualBlockWithStride2 (layers.py:176-188)."""
    out = conv16(sd, p + "down.", F.pixel_unshuffle(x, 2))
    return dcb(sd, p + "conv.", out, shortcut=shortcut)


def subpel(sd, p, x):
    """SubpelConv2x with kernel 1 (layers.py:92-103)."""
    return F.pixel_shuffle(conv16(sd, p + "conv.0.", x), 2)


def rbu(sd, p, x, shortcut):
    """ResidualBlockUpsample (layers.py:162-173)."""
    return dcb(sd, p + "conv.", subpel(sd, p + "up.", x), shortcut=shortcut)


def replicate_pad(x, pad_b, pad_r):
    if pad_b == 0 and pad_r == 0:
        return x
    return F.pad(x, (0, pad_r, 0, pad_b), mode="replicate")


def pad_to(x, p):
    H, W = x.shape[2], x.shape[3]
    pb = (H + p - 1) // p * p - H
    pr = (W + p - 1) // p * p - W
    return replicate_pad(x, pb, pr)


def qsel(sd, key, qp):
    return sd[key][qp]


# ---------------------------------------------------------------------------
# DMCI (intra) — image_model.py:21-217 with dmci_proxy.cpp:296-602 semantics
# ---------------------------------------------------------------------------

def dmci_encoder(sd, x_unsh, q):
    out = dcb(sd, "enc.enc_1.", x_unsh, quant=q)  # * q_scale_enc in the block tail (image_model.py:66-68)
    for i in range(6):
        out = dcb(sd, f"enc.enc_2.{i}.", out)
    return conv16(sd, "enc.enc_2.6.", out, stride=2, padding=1)


def dmci_hyper_enc(sd, y):
    out = dcb(sd, "hyper_enc.conv.0.", y)
    out = rbs2(sd, "hyper_enc.conv.1.", out, True)
    return rbs2(sd, "hyper_enc.conv.2.", out, True)


def dmci_hyper_dec(sd, z_hat):
    out = rbu(sd, "hyper_dec.conv.0.", z_hat, True)
    out = rbu(sd, "hyper_dec.conv.1.", out, True)
    return dcb(sd, "hyper_dec.conv.2.", out)


def dmci_prior_fusion(sd, x):
    for i in range(3):
        x = dcb(sd, f"y_prior_fusion.conv.{i}.", x)
    return conv16(sd, "y_prior_fusion.conv.3.", x)


def dmci_spatial_prior(sd, i, x):
    x = dcb(sd, f"y_spatial_prior_adaptor_{i}.", x)
    for j in range(3):
        x = dcb(sd, f"y_spatial_prior.conv.{j}.", x)
    return conv16(sd, "y_spatial_prior.conv.3.", x)


def dmci_decoder(sd, y_hat, q):
    out = rbu(sd, "dec.dec_1.0.", y_hat, True)
    for i in range(1, 13):  # * q_scale_dec in the last block's tail (image_model.py:42-44)
        out = dcb(sd, f"dec.dec_1.{i}.", out, quant=q if i == 12 else None)
    out = dcb(sd, "dec.dec_2.", out)
    return torch.clamp(F.pixel_shuffle(out, 8), -0.5, 0.5)


@dataclass
class IntraResult:
    steps_sym16: list = field(default_factory=list)   # per step: compacted int16
    z_int8: np.ndarray = None                          # NHWC-linear int8
    y_hat: torch.Tensor = None
    x_hat: torch.Tensor = None
    y: torch.Tensor = None
    scales_steps: list = field(default_factory=list)
    steps_dense: list = field(default_factory=list)
    steps_cond: list = field(default_factory=list)


def dmci_compress(sd, x, qp, skip_thres=0.0):
    """DMCIProxy::compress semantics (dmci_proxy.cpp:296-421).

    ``x``: [1,3,H,W] float tensor in [-0.5, 0.5] (unpadded). Returns IntraResult.
    """
    H, W = x.shape[2], x.shape[3]
    xp = pad_to(x, 16)
    y = dmci_encoder(sd, F.pixel_unshuffle(xp, 8), qsel(sd, "q_scale_enc", qp))
    y_pad = pad_to(y, 4)
    z = dmci_hyper_enc(sd, y_pad)
    z_hat = round_z(z)
    params = dmci_prior_fusion(sd, dmci_hyper_dec(sd, z_hat))
    yH, yW = y.shape[2], y.shape[3]
    params = params[:, :, :yH, :yW]
    scales, means = params.chunk(2, 1)
    y = st16(y * qsel(sd, "q_scale_y_enc", qp).reshape(1, -1, 1, 1))
    reduced = conv16(sd, "y_spatial_prior_reduction.", params)
    C = y.shape[1]
    masks = mask_4x(C, yH, yW)
    res = IntraResult(y=y)
    y_hat_so_far = None
    for step in range(4):
        if step > 0:
            sp = dmci_spatial_prior(sd, step, torch.cat([y_hat_so_far, reduced], 1))
            scales, means = sp.chunk(2, 1)
        y_q, y_hat = quantize_masked(y, means, scales, masks[step], skip_thres)
        y_hat = st16(y_hat)
        m = masks[step].to(y.dtype)
        s_hat = scales * m
        q = C // 4
        y_q_w = y_q[:, :q] + y_q[:, q:2 * q] + y_q[:, 2 * q:3 * q] + y_q[:, 3 * q:]
        s_w = s_hat[:, :q] + s_hat[:, q:2 * q] + s_hat[:, 2 * q:3 * q] + s_hat[:, 3 * q:]
        sym16, _, cond = build_index(y_q_w, s_w, skip_thres)
        res.steps_sym16.append(compact(sym16, cond))
        res.steps_dense.append(sym16)
        res.steps_cond.append(cond)
        res.scales_steps.append(s_w)
        y_hat_so_far = y_hat if y_hat_so_far is None else y_hat_so_far + y_hat
    y_hat = st16(y_hat_so_far * qsel(sd, "q_scale_y_dec", qp).reshape(1, -1, 1, 1))
    res.y_hat = y_hat
    res.z_int8 = z_hat.permute(0, 2, 3, 1).reshape(-1).numpy().astype(np.int8)
    res.x_hat = dmci_decoder(sd, y_hat, qsel(sd, "q_scale_dec", qp))
    return res


def dmci_decompress(sd, z_int8, decode_y, qp, height, width, skip_thres=0.0):
    """DMCIProxy::decompress semantics (dmci_proxy.cpp:423-602).

    ``decode_y(step, indexes_uint8) -> int8 quarter symbols`` is the entropy
    decoder callback; the encoder wrote steps 3..0 after z, so the decoder
    reads z, then steps 0..3.  Step k: CDF indexes of the active channel
    group's scales (single_part_for_reading_4x + build_index_dec, :667-676),
    the entropy round trip, recover + restore_y_4x(_and_add_inplace)
    (stream.cu:756-860) into y_hat_so_far, then the spatial prior for the
    next step.  y_hat = y_hat_so_far * q_scale_y_dec, then the decoder.
    Returns (x_hat, y_hat, indexes per step).
    """
    H16, W16 = (height + 15) // 16, (width + 15) // 16
    H64, W64 = (H16 + 3) // 4, (W16 + 3) // 4
    z_hat = torch.from_numpy(z_int8.astype(np.float32)).reshape(1, H64, W64, -1).permute(0, 3, 1, 2).contiguous()
    params = dmci_prior_fusion(sd, dmci_hyper_dec(sd, z_hat))
    params = params[:, :, :H16, :W16]
    scales, means = params.chunk(2, 1)
    reduced = conv16(sd, "y_spatial_prior_reduction.", params)
    C = scales.shape[1]
    q = C // 4
    masks = mask_4x(C, H16, W16)
    y_hat_so_far = None
    seen = []
    for step in range(4):
        if step > 0:
            sp = dmci_spatial_prior(sd, step, torch.cat([y_hat_so_far, reduced], 1))
            scales, means = sp.chunk(2, 1)
        m = masks[step].to(scales.dtype)
        s_hat = scales * m
        s_w = s_hat[:, :q] + s_hat[:, q:2 * q] + s_hat[:, 2 * q:3 * q] + s_hat[:, 3 * q:]
        _, idx, cond = build_index(None, s_w, skip_thres)
        seen.append(compact(idx, cond))
        decoded = decode_y(step, seen[-1])
        yq = recover(decoded, cond, idx.size)
        yq = torch.from_numpy(yq.astype(np.float32)).reshape(1, H16, W16, q).permute(0, 3, 1, 2)
        yq = yq.to(scales.dtype).contiguous().repeat(1, 4, 1, 1)
        y_hat = st16((yq + means) * m)
        y_hat_so_far = y_hat if y_hat_so_far is None else y_hat_so_far + y_hat
    y_hat = st16(y_hat_so_far * qsel(sd, "q_scale_y_dec", qp).reshape(1, -1, 1, 1))
    return dmci_decoder(sd, y_hat, qsel(sd, "q_scale_dec", qp)), y_hat, seen


# ---------------------------------------------------------------------------
# DMC HT-S — video_model_ht.py with dmc_hts_proxy.cpp:492-710 semantics
# ---------------------------------------------------------------------------

def seq(sd, prefix, n, x, quant_last=None, start=0):
    for i in range(start, n):
        x = dcb(sd, f"{prefix}{i}.", x, quant=quant_last if i == n - 1 else None)
    return x


def hts_feature_adaptor_i(sd, f):
    return seq(sd, "feature_adaptor_i.conv.", 4, f)


def hts_feature_adaptor_m(sd, memory, feature):
    return seq(sd, "feature_adaptor_m.conv.", 6, torch.cat([memory, feature], 1))


def hts_feature_extractor(sd, memory):
    return seq(sd, "feature_extractor.conv.", 5, memory)


def hts_recon_head(sd, feature, only_reset=False):
    """ReconHead.internal_forward for HT-S (video_model_ht.py:252-265)."""
    outs = []
    heads = []
    common = None
    frames = [7] if only_reset else range(8)
    for i in frames:
        if only_reset or i % 2 == 0:
            common = dcb(sd, f"recon_head.conv1.{i // 2}.0.", feature)
        out = common
        for j in range(3):
            out = dcb(sd, f"recon_head.conv2.{i}.{j}.", out)
        out = conv16(sd, f"recon_head.conv2.{i}.3.", out)
        heads.append(out)
        outs.append(torch.clamp(F.pixel_shuffle(out, 8), -0.5, 0.5))
    return outs, heads[-1]


@dataclass
class HTSState:
    """Per-stream temporal context held by the proxy (dmc_hts_proxy.h:320-323)."""
    feature_i: torch.Tensor = None     # [1,192,H8,W8] from an I frame or recon-head reset
    feature_p: torch.Tensor = None     # [1,512,H8,W8] decoder output of the last chunk
    memory: torch.Tensor = None
    ctx: torch.Tensor = None
    memory_has_value: bool = False


@dataclass
class HTSResult:
    sym16: np.ndarray = None
    z_int8: np.ndarray = None
    sym16_dense: np.ndarray = None
    cond_dense: np.ndarray = None
    y: torch.Tensor = None
    y_q: torch.Tensor = None
    y_hat: torch.Tensor = None
    scales: torch.Tensor = None
    common_params: torch.Tensor = None
    x_hat: list = None


def hts_add_ref_feature_from_frame(sd, st: HTSState, frame, apply_adaptor=True):
    """DMCHTSProxy::add_ref_feature_from_frame (dmc_hts_proxy.cpp:492-502)."""
    st.feature_i = F.pixel_unshuffle(frame, 8)
    if apply_adaptor:
        st.memory = hts_feature_adaptor_i(sd, st.feature_i)
        st.ctx = hts_feature_extractor(sd, st.memory)
    st.memory_has_value = apply_adaptor
    return st


def hts_encoder(sd, x_padded, ctx, qp):
    """Encoder (video_model_ht.py:63-92) with the quant tail fused in the last block."""
    feat = torch.cat([F.pixel_unshuffle(x_padded, 8), ctx], 1)
    feat = seq(sd, "encoder.conv1.", 6, feat, quant_last=qsel(sd, "q_encoder", qp))
    return conv16(sd, "encoder.down.", feat, stride=2, padding=1)


def hts_hyper_encoder(sd, y_pad):
    """HyperEncoder, HT-S (shortcut=False; video_model_ht.py:184-198)."""
    z = dcb(sd, "hyper_encoder.conv.0.", y_pad)
    z = rbs2(sd, "hyper_encoder.conv.1.", z, False)
    return rbs2(sd, "hyper_encoder.conv.2.", z, False)


def hts_hyper_decoder(sd, z_hat):
    """HyperDecoder, HT-S (shortcut=False; video_model_ht.py:167-181)."""
    hyper = dcb(sd, "hyper_decoder.conv.0.conv.", subpel(sd, "hyper_decoder.conv.0.up.", z_hat))
    hyper = dcb(sd, "hyper_decoder.conv.1.conv.", subpel(sd, "hyper_decoder.conv.1.up.", hyper))
    return dcb(sd, "hyper_decoder.conv.2.", hyper)


def hts_temporal_prior(sd, memory, qp):
    """TemporalPriorEncoder (video_model_ht.py:307-317)."""
    return rbs2(sd, "temporal_prior_encoder.conv.", st16(memory * qsel(sd, "q_feature", qp).reshape(1, -1, 1, 1)), False)


def _hts_prior(sd, y, z_hat, memory, qp):
    temporal = hts_temporal_prior(sd, memory, qp)
    hyper = hts_hyper_decoder(sd, z_hat)
    hyper = hyper[:, :, :temporal.shape[2], :temporal.shape[3]]
    x = torch.cat([hyper, temporal], 1)
    for i in range(3):
        x = dcb(sd, f"y_prior_fusion.conv.{i}.", x)
    return conv16(sd, "y_prior_fusion.conv.3.", x)


def _hts_spatial_prior(sd, i, y_hat, reduced):
    x = dcb(sd, f"y_spatial_prior_adaptor_{i}.", torch.cat([y_hat, reduced], 1))
    for j in range(3):
        x = dcb(sd, f"y_spatial_prior.conv.{j}.", x)
    return conv16(sd, "y_spatial_prior.conv.3.", x)


def _hts_decoder(sd, y_hat, ctx, qp):
    up = F.pixel_shuffle(conv16(sd, "decoder.up.conv.0.", y_hat), 2)
    return seq(sd, "decoder.conv1.", 7, torch.cat([up, ctx], 1), quant_last=qsel(sd, "q_decoder", qp))


def _hts_apply_adaptor(sd, st: HTSState):
    if st.memory_has_value:
        st.memory = hts_feature_adaptor_m(sd, st.memory, st.feature_p)
    else:
        st.memory = hts_feature_adaptor_i(sd, st.feature_i)
    st.ctx = hts_feature_extractor(sd, st.memory)


def hts_compress(sd, st: HTSState, x, qp, reset_feature_memory, skip_thres=0.0, recon=False):
    """DMCHTSProxy::compress (dmc_hts_proxy.cpp:504-585).

    ``x``: [1,24,H,W] (8 frames on channels).  The state must hold memory/ctx
    (adaptor applied).  Updates the state for the next chunk as the proxy does.
    """
    y = hts_encoder(sd, pad_to(x, 16), st.ctx, qp)
    z = hts_hyper_encoder(sd, pad_to(y, 4))
    z_hat = round_z(z)
    params = _hts_prior(sd, y, z_hat, st.memory, qp)
    q_dec, scales, means = params.chunk(3, 1)
    q_dec = torch.clamp_min(q_dec, 0.5)
    y = y * (1.0 / q_dec)
    reduced = conv16(sd, "y_spatial_prior_reduction.", params)
    C, yH, yW = y.shape[1], y.shape[2], y.shape[3]
    masks = mask_4x(C, yH, yW)
    y_q_tot = torch.zeros_like(y)
    y_hat = torch.zeros_like(y)
    for step in range(4):
        if step > 0:
            means = _hts_spatial_prior(sd, step, y_hat, reduced)
        y_q, yh = quantize_masked(y, means, scales, masks[step], skip_thres)
        y_q_tot = y_q_tot + y_q
        y_hat = y_hat + st16(yh)
    y_hat = st16(y_hat * q_dec)
    sym16, _, cond = build_index(y_q_tot, scales, skip_thres)
    res = HTSResult(sym16=compact(sym16, cond), y=y, y_q=y_q_tot, y_hat=y_hat, scales=scales,
                    common_params=params,
                    z_int8=z_hat.permute(0, 2, 3, 1).reshape(-1).numpy().astype(np.int8))
    res.sym16_dense = sym16
    res.cond_dense = cond
    # lambda_enc_1: decoder + context update (dmc_hts_proxy.cpp:571-574, 832-842)
    st.feature_p = _hts_decoder(sd, y_hat, st.ctx, qp)
    if recon:
        res.x_hat, _ = hts_recon_head(sd, st.feature_p)
    if reset_feature_memory:
        _, head = hts_recon_head(sd, st.feature_p, only_reset=True)
        st.memory = hts_feature_adaptor_i(sd, head)
    else:
        st.memory = hts_feature_adaptor_m(sd, st.memory, st.feature_p)
    st.ctx = hts_feature_extractor(sd, st.memory)
    st.memory_has_value = True
    return res


def hts_decompress(sd, st: HTSState, z_int8, decode_y, qp, height, width, reset_feature_memory,
                   skip_thres=0.0):
    """DMCHTSProxy::decompress (dmc_hts_proxy.cpp:587-710).

    ``decode_y(indexes_uint8) -> int8 symbols`` is the entropy-decoder callback;
    ``z_int8`` the decoded hyper latent in NHWC-linear order.
    """
    _hts_apply_adaptor(sd, st)
    H16, W16 = (height + 15) // 16, (width + 15) // 16
    H64, W64 = (H16 + 3) // 4, (W16 + 3) // 4
    # contiguous NCHW, as on the encoder side: a channels-last view would send
    # the CPU convolutions down another summation order
    z_hat = torch.from_numpy(z_int8.astype(np.float32)).reshape(1, H64, W64, -1).permute(0, 3, 1, 2)
    z_hat = z_hat.to(st.memory.dtype).contiguous()
    params = _hts_prior(sd, None, z_hat, st.memory, qp)
    q_dec, scales, means = params.chunk(3, 1)
    q_dec = torch.clamp_min(q_dec, 0.5)
    _, idx, cond = build_index(None, scales, skip_thres)
    decoded = decode_y(compact(idx, cond))
    C = scales.shape[1]
    y_q = recover(decoded, cond, idx.size)
    y_q = torch.from_numpy(y_q.astype(np.float32)).reshape(1, H16, W16, C).permute(0, 3, 1, 2)
    y_q = y_q.to(params.dtype).contiguous()
    reduced = conv16(sd, "y_spatial_prior_reduction.", params)
    masks = mask_4x(C, H16, W16)
    y_hat = torch.zeros_like(y_q)
    for step in range(4):
        if step > 0:
            means = _hts_spatial_prior(sd, step, y_hat, reduced)
        y_hat = y_hat + st16((y_q + means) * masks[step].to(y_q.dtype))
    y_hat = st16(y_hat * q_dec)
    st.feature_p = _hts_decoder(sd, y_hat, st.ctx, qp)
    x_hat, head = hts_recon_head(sd, st.feature_p)
    st.feature_i = head
    st.memory_has_value = not reset_feature_memory
    return x_hat, y_hat


# ---------------------------------------------------------------------------
# DMC HT-L — the ``not is_hts`` branches of video_model_ht.py with
# dmc_htl_proxy.cpp:476-1178 semantics
# ---------------------------------------------------------------------------

def htl_feature_adaptor_i(sd, f):
    """FeatureAdaptorI, HT-L: 3 ratio-1 blocks (video_model_ht.py:95-113)."""
    return seq(sd, "feature_adaptor_i.conv.", 3, f)


def htl_feature_adaptor_m(sd, memory, feature):
    """FeatureAdaptorM, HT-L: 10 blocks over cat(memory, feature) (:116-143)."""
    return seq(sd, "feature_adaptor_m.conv.", 10, torch.cat([memory, feature], 1))


def htl_feature_extractor(sd, memory):
    """FeatureExtractor, HT-L: 2 blocks (:146-164)."""
    return seq(sd, "feature_extractor.conv.", 2, memory)


def htl_recon_head(sd, feature, only_reset=False):
    """ReconHead, HT-L (video_model_ht.py:236-248, 266-276): per frame 5
    blocks and a 1x1 to 192 channels, then pixel_shuffle(8) and the proxy's
    clamp to [-0.5, 0.5]; the reset branch runs frame 7's head only
    (dmc_htl_proxy.cpp:347-359)."""
    outs, heads = [], []
    for i in ([7] if only_reset else range(8)):
        out = seq(sd, f"recon_head.conv.{i}.", 5, feature)
        out = conv16(sd, f"recon_head.conv.{i}.5.", out)
        heads.append(out)
        outs.append(torch.clamp(F.pixel_shuffle(out, 8), -0.5, 0.5))
    return outs, heads[-1]


def htl_add_ref_feature_from_frame(sd, st: HTSState, frame, apply_adaptor=True):
    """DMCHTLProxy::add_ref_feature_from_frame (dmc_htl_proxy.cpp:583-593)."""
    st.feature_i = F.pixel_unshuffle(frame, 8)
    if apply_adaptor:
        st.memory = htl_feature_adaptor_i(sd, st.feature_i)
        st.ctx = htl_feature_extractor(sd, st.memory)
    st.memory_has_value = apply_adaptor
    return st


def htl_encoder(sd, x_padded, ctx, qp):
    """Encoder, HT-L (video_model_ht.py:63-92): 7 blocks, * q_encoder as a
    separate multiply (dmc_htl_proxy.cpp:93-106), 3x3 stride-2 down conv."""
    feat = torch.cat([F.pixel_unshuffle(x_padded, 8), ctx], 1)
    feat = seq(sd, "encoder.conv1.", 7, feat)
    feat = st16(feat * qsel(sd, "q_encoder", qp).reshape(1, -1, 1, 1))
    return conv16(sd, "encoder.down.", feat, stride=2, padding=1)


def htl_hyper_encoder(sd, y_pad):
    """HyperEncoder, HT-L: the stride-2 blocks keep their shortcut (:184-198)."""
    z = dcb(sd, "hyper_encoder.conv.0.", y_pad)
    z = rbs2(sd, "hyper_encoder.conv.1.", z, True)
    return rbs2(sd, "hyper_encoder.conv.2.", z, True)


def htl_hyper_decoder(sd, z_hat):
    """HyperDecoder, HT-L: force_bias up convs, shortcut blocks (:167-181)."""
    hyper = rbu(sd, "hyper_decoder.conv.0.", z_hat, True)
    hyper = rbu(sd, "hyper_decoder.conv.1.", hyper, True)
    return dcb(sd, "hyper_decoder.conv.2.", hyper)


def htl_temporal_prior(sd, memory, qp):
    """TemporalPriorEncoder, HT-L: shortcut stride-2 block (:307-317)."""
    return rbs2(sd, "temporal_prior_encoder.conv.", st16(memory * qsel(sd, "q_feature", qp).reshape(1, -1, 1, 1)), True)


def _htl_prior(sd, z_hat, memory, qp):
    temporal = htl_temporal_prior(sd, memory, qp)
    hyper = htl_hyper_decoder(sd, z_hat)
    hyper = hyper[:, :, :temporal.shape[2], :temporal.shape[3]]
    x = torch.cat([hyper, temporal], 1)
    for i in range(3):
        x = dcb(sd, f"y_prior_fusion.conv.{i}.", x)
    return conv16(sd, "y_prior_fusion.conv.3.", x)


def _htl_spatial_prior(sd, i, y_hat_so_far, reduced):
    """SpatialPriorAdaptor_i on cat(y_hat_so_far, reduced) -> SpatialPrior ->
    (scales, means) (common_model.py:237-282, spatial_prior_has_scales)."""
    x = dcb(sd, f"y_spatial_prior_adaptor_{i}.", torch.cat([y_hat_so_far, reduced], 1))
    for j in range(3):
        x = dcb(sd, f"y_spatial_prior.conv.{j}.", x)
    return conv16(sd, "y_spatial_prior.conv.3.", x).chunk(2, 1)


def _htl_decoder(sd, y_hat, ctx, qp):
    """Decoder, HT-L: 3x3 biased sub-pixel up conv, 11 blocks, * q_decoder
    as a separate multiply (dmc_htl_proxy.cpp:22-42)."""
    up = F.pixel_shuffle(conv16(sd, "decoder.up.conv.0.", y_hat, padding=1), 2)
    feat = seq(sd, "decoder.conv1.", 11, torch.cat([up, ctx], 1))
    return st16(feat * qsel(sd, "q_decoder", qp).reshape(1, -1, 1, 1))


def _quarter(x, mask):
    """single_part_for_writing/reading_4x (stream.cu:896-962): the sum of the
    four channel groups of x * mask, i.e. the active group's value."""
    q = x.shape[1] // 4
    xm = x * mask.to(x.dtype)
    return xm[:, :q] + xm[:, q:2 * q] + xm[:, 2 * q:3 * q] + xm[:, 3 * q:]


@dataclass
class HTLResult:
    steps_sym16: list = field(default_factory=list)   # per step: compacted int16 (quarter order)
    steps_dense: list = field(default_factory=list)
    steps_cond: list = field(default_factory=list)
    z_int8: np.ndarray = None
    y: torch.Tensor = None                             # y * (1 / q_dec)
    y_hat: torch.Tensor = None
    common_params: torch.Tensor = None
    x_hat: list = None


def htl_compress(sd, st: HTSState, x, qp, reset_feature_memory, skip_thres=0.0, recon=False):
    """DMCHTLProxy::compress (dmc_htl_proxy.cpp:595-716): four checkerboard
    steps, each with its own scales and means (step 0 from the common
    params, steps 1-3 from the spatial prior over the y_hat accumulated so
    far) and its own quarter-tensor symbol stream; then the decoder and the
    context update."""
    y = htl_encoder(sd, pad_to(x, 16), st.ctx, qp)
    z = htl_hyper_encoder(sd, pad_to(y, 4))
    z_hat = round_z(z)
    params = _htl_prior(sd, z_hat, st.memory, qp)
    res = htl_code_y(sd, y, params, skip_thres)
    res.z_int8 = z_hat.permute(0, 2, 3, 1).reshape(-1).numpy().astype(np.int8)
    st.feature_p = _htl_decoder(sd, res.y_hat, st.ctx, qp)
    if recon:
        res.x_hat, _ = htl_recon_head(sd, st.feature_p)
    if reset_feature_memory:
        _, head = htl_recon_head(sd, st.feature_p, only_reset=True)
        st.memory = htl_feature_adaptor_i(sd, head)
    else:
        st.memory = htl_feature_adaptor_m(sd, st.memory, st.feature_p)
    st.ctx = htl_feature_extractor(sd, st.memory)
    st.memory_has_value = True
    return res


def htl_code_y(sd, y, params, skip_thres=0.0):
    """The four checkerboard steps of DMCHTLProxy::compress
    (dmc_htl_proxy.cpp:619-688) from the encoder output y and the common
    params: y / max(q_dec, .5), step 0 on the common scales / means, steps
    1-3 on the spatial prior over the y_hat accumulated so far; returns an
    HTLResult with the per-step symbol streams and the final y_hat."""
    q_dec, scales, means = params.chunk(3, 1)
    q_dec = torch.clamp_min(q_dec, 0.5)
    y = y * (1.0 / q_dec)
    reduced = conv16(sd, "y_spatial_prior_reduction.", params)
    C, yH, yW = y.shape[1], y.shape[2], y.shape[3]
    masks = mask_4x(C, yH, yW)
    res = HTLResult(y=y, common_params=params)
    so_far = torch.zeros_like(y)
    for step in range(4):
        if step > 0:
            scales, means = _htl_spatial_prior(sd, step, so_far, reduced)
        y_q, yh = quantize_masked(y, means, scales, masks[step], skip_thres)
        yh = st16(yh)
        sym16, _, cond = build_index(_quarter(y_q, masks[step]), _quarter(scales, masks[step]), skip_thres)
        res.steps_sym16.append(compact(sym16, cond))
        res.steps_dense.append(sym16)
        res.steps_cond.append(cond)
        if step < 3:
            so_far = so_far + yh
        else:
            res.y_hat = st16((yh + so_far) * q_dec)
    return res


def htl_decompress(sd, st: HTSState, z_int8, decode_y, qp, height, width, reset_feature_memory,
                   skip_thres=0.0):
    """DMCHTLProxy::decompress (dmc_htl_proxy.cpp:718-915): four entropy
    round trips.  ``decode_y(step, indexes_uint8) -> int8 symbols``."""
    if st.memory_has_value:
        st.memory = htl_feature_adaptor_m(sd, st.memory, st.feature_p)
    else:
        st.memory = htl_feature_adaptor_i(sd, st.feature_i)
    H16, W16 = (height + 15) // 16, (width + 15) // 16
    H64, W64 = (H16 + 3) // 4, (W16 + 3) // 4
    z_hat = torch.from_numpy(z_int8.astype(np.float32)).reshape(1, H64, W64, -1).permute(0, 3, 1, 2)
    z_hat = z_hat.to(st.memory.dtype).contiguous()
    params = _htl_prior(sd, z_hat, st.memory, qp)
    st.ctx = htl_feature_extractor(sd, st.memory)
    q_dec, scales, means = params.chunk(3, 1)
    q_dec = torch.clamp_min(q_dec, 0.5)
    reduced = conv16(sd, "y_spatial_prior_reduction.", params)
    C = scales.shape[1]
    masks = mask_4x(C, H16, W16)
    so_far = torch.zeros((1, C, H16, W16), dtype=params.dtype)
    y_hat = None
    for step in range(4):
        if step > 0:
            scales, means = _htl_spatial_prior(sd, step, so_far, reduced)
        _, idx, cond = build_index(None, _quarter(scales, masks[step]), skip_thres)
        decoded = decode_y(step, compact(idx, cond))
        yq = recover(decoded, cond, idx.size)
        yq = torch.from_numpy(yq.astype(np.float32)).reshape(1, H16, W16, C // 4).permute(0, 3, 1, 2)
        yq = yq.to(params.dtype).contiguous().repeat(1, 4, 1, 1)
        yh = st16((yq + means) * masks[step].to(params.dtype))
        if step < 3:
            so_far = so_far + yh
        else:
            y_hat = st16((yh + so_far) * q_dec)
    st.feature_p = _htl_decoder(sd, y_hat, st.ctx, qp)
    x_hat, head = htl_recon_head(sd, st.feature_p)
    st.feature_i = head
    st.memory_has_value = not reset_feature_memory
    return x_hat, y_hat


# ---------------------------------------------------------------------------
# DMC low delay — video_model_ld.py with dmc_ld_proxy.cpp:380-862 semantics
# ---------------------------------------------------------------------------

def mask_2x(C, H, W):
    """get_mask_2x (common_model.py:157-172, dmc_ld_proxy.cpp:678-691)."""
    m0 = _one_mask(((1, 0), (0, 1)), H, W)
    m1 = _one_mask(((0, 1), (1, 0)), H, W)
    q = C // 2

    def cat(a, b):
        return torch.cat([a.expand(1, q, H, W), b.expand(1, q, H, W)], dim=1)
    return [cat(m0, m1), cat(m1, m0)]


@dataclass
class LDState:
    """Per-stream context of DMCLDProxy (dmc_ld_proxy.h): as HT, plus the
    temporal params, which the proxy computes once the memory is known."""
    feature_i: torch.Tensor = None     # [1,192,H8,W8]
    feature_p: torch.Tensor = None     # [1,256,H8,W8]
    memory: torch.Tensor = None
    ctx: torch.Tensor = None
    temporal: torch.Tensor = None      # [1,256,H16,W16], before the q_feature multiply
    memory_has_value: bool = False


@dataclass
class LDResult:
    sym16: np.ndarray = None
    sym16_dense: np.ndarray = None
    cond_dense: np.ndarray = None
    z_int8: np.ndarray = None
    y: torch.Tensor = None             # y * (1 / q_dec)
    y_hat: torch.Tensor = None
    common_params: torch.Tensor = None
    x_hat: torch.Tensor = None


def ld_feature_adaptor_i(sd, f):
    """FeatureAdaptorI, LD: 4 dcb2 blocks (video_model_ld.py:62-75)."""
    return seq(sd, "feature_adaptor_i.conv.", 4, f)


def ld_feature_adaptor_m(sd, memory, feature):
    """FeatureAdaptorM, LD: 4 blocks over cat(memory, feature) (:78-89)."""
    return seq(sd, "feature_adaptor_m.conv.", 4, torch.cat([memory, feature], 1))


def ld_feature_extractor(sd, memory):
    """FeatureExtractor, LD: 5 blocks (:92-106)."""
    return seq(sd, "feature_extractor.conv.", 5, memory)


def ld_recon_head(sd, feature):
    """ReconHead, LD (:151-166): 3 blocks, 1x1 to 192, pixel_shuffle(8) with
    the proxy's clamp; also returns the head output (the reset feature)."""
    out = seq(sd, "recon_head.conv.", 3, feature)
    head = conv16(sd, "recon_head.head.", out)
    return torch.clamp(F.pixel_shuffle(head, 8), -0.5, 0.5), head


def ld_add_ref_feature_from_frame(sd, st: LDState, frame, apply_adaptor=True):
    """DMCLDProxy::add_ref_feature_from_frame (dmc_ld_proxy.cpp:407-418)."""
    st.feature_i = F.pixel_unshuffle(frame, 8)
    if apply_adaptor:
        st.memory = ld_feature_adaptor_i(sd, st.feature_i)
        st.ctx = ld_feature_extractor(sd, st.memory)
        st.temporal = ld_temporal_prior(sd, st.memory)
    st.memory_has_value = apply_adaptor
    return st


def ld_encoder(sd, x_padded, ctx, qp):
    """Encoder, LD (video_model_ld.py:43-59): 2 blocks, the third with the
    q_encoder multiply fused in its tail (dmc_ld_proxy.cpp:63-71), 3x3
    stride-2 down conv."""
    feat = torch.cat([F.pixel_unshuffle(x_padded, 8), ctx], 1)
    feat = seq(sd, "encoder.conv1.", 2, feat)
    feat = dcb(sd, "encoder.conv2.", feat, quant=qsel(sd, "q_encoder", qp))
    return conv16(sd, "encoder.down.", feat, stride=2, padding=1)


def ld_hyper_encoder(sd, y_pad):
    """HyperEncoder, LD (:124-134)."""
    z = dcb(sd, "hyper_encoder.conv.0.", y_pad)
    z = rbs2(sd, "hyper_encoder.conv.1.", z, False)
    return rbs2(sd, "hyper_encoder.conv.2.", z, False)


def ld_hyper_decoder(sd, z_hat):
    """HyperDecoder, LD (:109-121)."""
    hyper = rbu(sd, "hyper_decoder.conv.0.", z_hat, False)
    hyper = rbu(sd, "hyper_decoder.conv.1.", hyper, False)
    return dcb(sd, "hyper_decoder.conv.2.", hyper)


def ld_temporal_prior(sd, memory):
    """TemporalPriorEncoder, LD (:182-188): no quant on its input."""
    return rbs2(sd, "temporal_prior_encoder.conv.", memory, False)


def _ld_prior(sd, z_hat, temporal, qp):
    """PriorFusion over cat(hyper, temporal * q_feature) (:137-148)."""
    hyper = ld_hyper_decoder(sd, z_hat)
    hyper = hyper[:, :, :temporal.shape[2], :temporal.shape[3]]
    x = torch.cat([hyper, st16(temporal * qsel(sd, "q_feature", qp).reshape(1, -1, 1, 1))], 1)
    for i in range(3):
        x = dcb(sd, f"y_prior_fusion.conv.{i}.", x)
    return conv16(sd, "y_prior_fusion.conv.3.", x)


def ld_spatial_prior(sd, y_hat0, params):
    """SpatialPrior, LD (:169-179): means for the second half from
    cat(y_hat of step 0, common params)."""
    x = torch.cat([y_hat0, params], 1)
    x = dcb(sd, "y_spatial_prior.conv.0.", x)
    x = dcb(sd, "y_spatial_prior.conv.1.", x)
    return conv16(sd, "y_spatial_prior.conv.2.", x)


def _ld_decoder(sd, y_hat, ctx, qp):
    """Decoder, LD (:24-40): 1x1 sub-pixel up conv, 3 blocks, 1x1 conv with
    the q_decoder multiply fused (conv1x1_bias_with_quant, dmc_ld_proxy.cpp:27-35)."""
    up = subpel(sd, "decoder.up.", y_hat)
    feat = seq(sd, "decoder.conv1.", 3, torch.cat([up, ctx], 1))
    return st16(conv(sd, "decoder.conv2.", feat) * qsel(sd, "q_decoder", qp).reshape(1, -1, 1, 1))


def ld_code_y(sd, y, params, skip_thres=0.0):
    """The two checkerboard halves of DMCLDProxy::compress
    (dmc_ld_proxy.cpp:432-455, forward_prior_2x common_model.py:212-232):
    y / max(q_dec, .5); half 0 on the common scales / means; half 1 on the
    same scales and the spatial prior's means; one symbol stream over the
    whole of y; y_hat = (y_hat_0 + y_hat_1) * max(q_dec, .5)."""
    q_dec, scales, means = params.chunk(3, 1)
    q_dec = torch.clamp_min(q_dec, 0.5)
    y = y * (1.0 / q_dec)
    C, yH, yW = y.shape[1], y.shape[2], y.shape[3]
    masks = mask_2x(C, yH, yW)
    y_q0, yh0 = quantize_masked(y, means, scales, masks[0], skip_thres)
    yh0 = st16(yh0)
    means1 = ld_spatial_prior(sd, yh0, params)
    y_q1, yh1 = quantize_masked(y, means1, scales, masks[1], skip_thres)
    yh1 = st16(yh1)
    sym16, _, cond = build_index(y_q0 + y_q1, scales, skip_thres)
    res = LDResult(sym16=compact(sym16, cond), sym16_dense=sym16, cond_dense=cond, y=y, common_params=params)
    res.y_hat = st16((yh0 + yh1) * q_dec)
    return res


def ld_compress(sd, st: LDState, x, qp, reset_feature_memory, skip_thres=0.0, recon=False):
    """DMCLDProxy::compress (dmc_ld_proxy.cpp:420-485); x: [1,3,H,W] one frame."""
    y = ld_encoder(sd, pad_to(x, 16), st.ctx, qp)
    z = ld_hyper_encoder(sd, pad_to(y, 4))
    z_hat = round_z(z)
    params = _ld_prior(sd, z_hat, st.temporal, qp)
    res = ld_code_y(sd, y, params, skip_thres)
    res.z_int8 = z_hat.permute(0, 2, 3, 1).reshape(-1).numpy().astype(np.int8)
    # lambda_enc_1 (:466-470, 639-649)
    st.feature_p = _ld_decoder(sd, res.y_hat, st.ctx, qp)
    if recon:
        res.x_hat, _ = ld_recon_head(sd, st.feature_p)
    if reset_feature_memory:
        _, head = ld_recon_head(sd, st.feature_p)
        st.memory = ld_feature_adaptor_i(sd, head)
    else:
        st.memory = ld_feature_adaptor_m(sd, st.memory, st.feature_p)
    st.ctx = ld_feature_extractor(sd, st.memory)
    st.temporal = ld_temporal_prior(sd, st.memory)
    st.memory_has_value = True
    return res


def ld_decompress(sd, st: LDState, z_int8, decode_y, qp, height, width, reset_feature_memory, skip_thres=0.0):
    """DMCLDProxy::decompress (dmc_ld_proxy.cpp:487-593): one entropy round
    trip.  ``decode_y(indexes_uint8) -> int8 symbols``."""
    if st.memory_has_value:
        st.memory = ld_feature_adaptor_m(sd, st.memory, st.feature_p)
    else:
        st.memory = ld_feature_adaptor_i(sd, st.feature_i)
    st.temporal = ld_temporal_prior(sd, st.memory)
    H16, W16 = (height + 15) // 16, (width + 15) // 16
    H64, W64 = (H16 + 3) // 4, (W16 + 3) // 4
    z_hat = torch.from_numpy(z_int8.astype(np.float32)).reshape(1, H64, W64, -1).permute(0, 3, 1, 2)
    z_hat = z_hat.to(st.memory.dtype).contiguous()
    params = _ld_prior(sd, z_hat, st.temporal, qp)
    st.ctx = ld_feature_extractor(sd, st.memory)
    q_dec, scales, means = params.chunk(3, 1)
    q_dec = torch.clamp_min(q_dec, 0.5)
    _, idx, cond = build_index(None, scales, skip_thres)
    decoded = decode_y(compact(idx, cond))
    C = scales.shape[1]
    y_q = recover(decoded, cond, idx.size)
    y_q = torch.from_numpy(y_q.astype(np.float32)).reshape(1, H16, W16, C).permute(0, 3, 1, 2)
    y_q = y_q.to(params.dtype).contiguous()
    masks = mask_2x(C, H16, W16)
    yh0 = st16((y_q + means) * masks[0].to(y_q.dtype))
    means1 = ld_spatial_prior(sd, yh0, params)
    y_hat = st16((yh0 + st16((y_q + means1) * masks[1].to(y_q.dtype))) * q_dec)
    st.feature_p = _ld_decoder(sd, y_hat, st.ctx, qp)
    x_hat, head = ld_recon_head(sd, st.feature_p)
    st.feature_i = head
    st.memory_has_value = not reset_feature_memory
    return x_hat, y_hat


# ---------------------------------------------------------------------------
# per-op references used by the GPU kernel parity tests (float64)
# ---------------------------------------------------------------------------

def ref_gemm_epilogue(a, w, bias, epi, s1=None, s2=None, q=None):
    """1x1-conv GEMM + epilogue in float64 (cutlass_epilogue.h:79-131, 175-224).

    a [M,K], w [N,K], bias [N]; epi in {bias, wsilu, shortcut, shortcut2,
    shortcut_quant, quant, quad}.  For 'quad', w has 4N rows in the
    reference ffn.0 order and the result is sum_j wsilu(a w[j::4]^T + b[j::4]).
    """
    a = np.asarray(a, np.float64)
    w = np.asarray(w, np.float64)
    b = np.asarray(bias, np.float64)
    x = a @ w.T + b
    if epi == "quad":
        x = x / (1.0 + np.exp(-4.0 * x))
        return x[:, 0::4] + x[:, 1::4] + x[:, 2::4] + x[:, 3::4]
    if epi == "wsilu":
        return x / (1.0 + np.exp(-4.0 * x))
    if epi in ("shortcut", "shortcut2", "shortcut_quant"):
        x = x + np.asarray(s1, np.float64)
    if epi == "shortcut2":
        x = x + np.asarray(s2, np.float64)
    if epi in ("shortcut_quant", "quant"):
        x = x * np.asarray(q, np.float64)[None, :]
    return x


def ref_dw3x3(x_hwc, w_c33):
    """Depthwise 3x3, pad 1, no bias, float64. x [H,W,C], w [C,3,3]."""
    x = np.asarray(x_hwc, np.float64)
    w = np.asarray(w_c33, np.float64)
    H, W, C = x.shape
    xp = np.zeros((H + 2, W + 2, C))
    xp[1:-1, 1:-1] = x
    out = np.zeros((H, W, C))
    for ky in range(3):
        for kx in range(3):
            out += xp[ky:ky + H, kx:kx + W] * w[None, None, :, ky, kx]
    return out


# ---------------------------------------------------------------------------
# float32 step-exact references for the fused quantisation kernels.  Inputs
# are the same fp16 arrays the kernels read; every operation is one IEEE
# float32 op in the order the proxy kernels apply them (stream.cu:422-443,
# 548-630, 685-860), so results must match the GPU bit-for-bit.
# ---------------------------------------------------------------------------

f32 = np.float32


def _rha32(d):
    return round_half_away(np.asarray(d, np.float32))


def _quant32(y, mean, scale, thr):
    r = _rha32((y - mean).astype(np.float32)).astype(np.float32)
    r = np.where(scale > f32(thr), r, f32(0.0))
    return np.clip(r, -128, 127).astype(np.int64)


def mask4_active(step, C, H, W, n_img=1):
    """[n_img*H*W, C] bool: active positions of checkerboard step (mask_4x)."""
    m = mask_4x(C, H, W)[step][0].permute(1, 2, 0).reshape(H * W, C).numpy()
    return np.tile(m, (n_img, 1))


def hts_quant_step_f32(y, scales, means, qdec, yhat_in, step, thr, C, H, W, n_img=1):
    """One HT-S encoder step on [P, C] fp16 arrays -> (yhat fp16, sym int16, cond u8, active)."""
    act = mask4_active(step, C, H, W, n_img)
    q = np.maximum(qdec.astype(np.float32), f32(0.5))
    ys = (y.astype(np.float32) * (f32(1.0) / q)).astype(np.float32)
    mean = means.astype(np.float32)
    sc = scales.astype(np.float32)
    yq = _quant32(ys, mean, sc, thr)
    yh = (yq.astype(np.float32) + mean).astype(np.float16)
    idx = scale_to_index(sc)
    sym = ((yq << 8) + idx).astype(np.int16)
    cond = (sc > f32(thr)).astype(np.uint8)
    old = np.zeros_like(yhat_in) if step == 0 else yhat_in
    out = np.where(act, yh, old)
    if step == 3:
        out = (out.astype(np.float32) * q).astype(np.float16)
    return out.astype(np.float16), sym, cond, act


def hts_restore_step_f32(yq, means, qdec, yhat_in, step, C, H, W, n_img=1):
    act = mask4_active(step, C, H, W, n_img)
    yh = (yq.astype(np.float32) + means.astype(np.float32)).astype(np.float16)
    old = np.zeros_like(yhat_in) if step == 0 else yhat_in
    out = np.where(act, yh, old)
    if step == 3:
        out = (out.astype(np.float32) * np.maximum(qdec.astype(np.float32), f32(0.5))).astype(np.float16)
    return out.astype(np.float16)


def intra_quant_step_f32(y, scales, means, ysf_in, q_row, step, thr, C, H, W, n_img=1):
    """DMCI encoder step -> (ysf fp16, yout fp16 or None, sym quarter, cond quarter)."""
    act = mask4_active(step, C, H, W, n_img)
    mean = means.astype(np.float32)
    sc = scales.astype(np.float32)
    yq = _quant32(y.astype(np.float32), mean, sc, thr)
    yh = (yq.astype(np.float32) + mean).astype(np.float16)
    C4 = C // 4
    # quarter view: the active group's value at each (pixel, c')
    sel = lambda a: (a.reshape(-1, 4, C4) * act.reshape(-1, 4, C4)).sum(1)
    sym = ((sel(yq) << 8) + scale_to_index(sel(sc * act).astype(np.float32))).astype(np.int16).reshape(-1)
    cond = (sel(sc * act).astype(np.float32) > f32(thr)).astype(np.uint8).reshape(-1)
    old = np.zeros_like(ysf_in) if step == 0 else ysf_in
    ysf = np.where(act, yh, old).astype(np.float16)
    yout = None
    if step == 3:
        yout = (ysf.astype(np.float32) * _qrows(q_row)).astype(np.float16)
        ysf = ysf_in
    return ysf, yout, sym, cond


def intra_restore_step_f32(yq_quarter, means, ysf_in, q_row, step, C, H, W, n_img=1):
    act = mask4_active(step, C, H, W, n_img)
    C4 = C // 4
    yq = np.repeat(yq_quarter.reshape(-1, 1, C4), 4, axis=1).reshape(-1, C)
    yh = (yq.astype(np.float32) + means.astype(np.float32)).astype(np.float16)
    old = np.zeros_like(ysf_in) if step == 0 else ysf_in
    ysf = np.where(act, yh, old).astype(np.float16)
    if step == 3:
        return ysf_in, (ysf.astype(np.float32) * _qrows(q_row)).astype(np.float16)
    return ysf, None


def _qrows(q):
    """A per-channel [C] scale as a row, or an elementwise [P, C] scale as is."""
    q = np.asarray(q).astype(np.float32)
    return q[None, :] if q.ndim == 1 else q


def htl_quant_step_f32(y, scales, means, qdec, ysf_in, step, thr, C, H, W, n_img=1):
    """HT-L encoder step: the DMCI quarter step on y / max(q_dec, .5), whose
    step-3 output is scaled by max(q_dec, .5) elementwise
    (divide_with_clamp_min, add_and_multiply_with_clamp_min; stream.cu:40-66,
    422-443).  -> (ysf, yout or None, sym quarter, cond quarter)"""
    q = np.maximum(qdec.astype(np.float32), f32(0.5))
    ys = (y.astype(np.float32) * (f32(1.0) / q)).astype(np.float32)
    return intra_quant_step_f32(ys, scales, means, ysf_in, q, step, thr, C, H, W, n_img)


def htl_restore_step_f32(yq_quarter, means, qdec, ysf_in, step, C, H, W, n_img=1):
    q = np.maximum(qdec.astype(np.float32), f32(0.5))
    return intra_restore_step_f32(yq_quarter, means, ysf_in, q, step, C, H, W, n_img)


def mask2_active(step, C, H, W, n_img=1):
    """[n_img*H*W, C] bool: active positions of 2x checkerboard half `step`."""
    m = mask_2x(C, H, W)[step][0].permute(1, 2, 0).reshape(H * W, C).numpy()
    return np.tile(m, (n_img, 1))


def ld_quant_step_f32(y, scales, means, qdec, yhat_in, step, thr, C, H, W, n_img=1):
    """One LD encoder half (process_with_mask_no_scale(_add_and_multiply_inplace),
    stream.cu:548-645) on [P, C] fp16 arrays -> (yhat fp16, sym, cond, active)."""
    act = mask2_active(step, C, H, W, n_img)
    q = np.maximum(qdec.astype(np.float32), f32(0.5))
    ys = (y.astype(np.float32) * (f32(1.0) / q)).astype(np.float32)
    mean = means.astype(np.float32)
    sc = scales.astype(np.float32)
    yq = _quant32(ys, mean, sc, thr)
    yh = (yq.astype(np.float32) + mean).astype(np.float16)
    sym = ((yq << 8) + scale_to_index(sc)).astype(np.int16)
    cond = (sc > f32(thr)).astype(np.uint8)
    old = np.zeros_like(yhat_in) if step == 0 else yhat_in
    out = np.where(act, yh, old)
    if step == 1:
        out = (out.astype(np.float32) * q).astype(np.float16)
    return out.astype(np.float16), sym, cond, act


def ld_restore_step_f32(yq, means, qdec, yhat_in, step, C, H, W, n_img=1):
    """restore_y / restore_y_and_add_multiply_inplace (stream.cu:685-754)."""
    act = mask2_active(step, C, H, W, n_img)
    yh = (yq.astype(np.float32) + means.astype(np.float32)).astype(np.float16)
    old = np.zeros_like(yhat_in) if step == 0 else yhat_in
    out = np.where(act, yh, old)
    if step == 1:
        out = (out.astype(np.float32) * np.maximum(qdec.astype(np.float32), f32(0.5))).astype(np.float16)
    return out.astype(np.float16)
