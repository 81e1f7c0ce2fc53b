"""Torch-facing wrappers of the kernel-level C ABI (include/dcvc_ops.h).

Each function mirrors one operator of the reference's def_cutlass.h /
def_elementwise.h and takes CUDA (HIP) tensors.  Activations are 2-D
[pixels, C] fp16 views whose row stride is the NHWC pixel stride, so channel
slices of concatenation buffers work unchanged.  Everything runs on the
current torch stream; nothing falls back to torch compute.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

_lib_call = _lib.call


def _p(t):
    return None if t is None else t.data_ptr()


def _ld(t):
    assert t.dim() == 2 and t.stride(1) == 1, "expected a [pixels, C] row-major view"
    return t.stride(0)


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _check(t, dtype=torch.float16):
    if not t.is_cuda:
        raise ValueError("tensor must live on the GPU")
    if t.dtype != dtype:
        raise ValueError(f"expected {dtype}, got {t.dtype}")


_TH = {}


def thresholds(device="cuda"):
    """Device copy of the 128-entry scale->index threshold table."""
    key = str(device)
    if key not in _TH:
        th = np.zeros(128, dtype=np.float32)
        _lib_call("dcvc_scale_index_thresholds", th.ctypes.data)
        _TH[key] = torch.from_numpy(th).to(device)
    return _TH[key]


def host_thresholds():
    th = np.zeros(128, dtype=np.float32)
    _lib_call("dcvc_scale_index_thresholds", th.ctypes.data)
    return th


# ----------------------------- GEMMs ---------------------------------------

EPI_CODES = {"bias": 0, "wsilu": 1, "shortcut": 2, "shortcut2": 3, "shortcut_quant": 4, "quant": 5, "quad": 6}
PRECISIONS = {"fp16": 0, "bf16": 1, "fp8": 2}


def weights_to_lowp(w, prec):
    """fp16 weights [rows, K] -> (bf16 [rows, K], None) or (fp8 e4m3 bytes
    [rows, K] as uint8, fp32 per-row scales) on the device."""
    _check(w)
    w = w.contiguous()
    rows, K = w.shape
    p = PRECISIONS[prec]
    if p == 1:
        out = torch.empty((rows, K), dtype=torch.bfloat16, device=w.device)
        _lib_call("dcvc_weights_to_lowp", _p(w), rows, K, p, _p(out), None, _stream())
        return out, None
    if p == 2:
        out = torch.empty((rows, K), dtype=torch.uint8, device=w.device)
        scale = torch.empty(rows, dtype=torch.float32, device=w.device)
        _lib_call("dcvc_weights_to_lowp", _p(w), rows, K, p, _p(out), _p(scale), _stream())
        return out, scale
    raise ValueError(prec)


def gemm_lowp(a, wq, scale, prec, bias, epi="bias", s1=None, s2=None, q=None, qp=None, out=None):
    """conv1x1 with the product in `prec` ('fp16', 'bf16', 'fp8'): wq from
    weights_to_lowp (or fp16 weights for 'fp16')."""
    _check(a)
    M, K = a.shape
    N = wq.shape[0] // 4 if epi == "quad" else wq.shape[0]
    if out is None:
        out = torch.empty((M, N), dtype=torch.float16, device=a.device)
    bias = bias.float().contiguous()
    qs = 0 if q is None else (q.stride(0) if q.dim() == 2 else q.numel())
    _lib_call("dcvc_gemm_lowp", _p(a), _ld(a), M, K, _p(wq.contiguous()), N, _p(scale), PRECISIONS[prec],
              EPI_CODES[epi], _p(bias), _p(s1), _ld(s1) if s1 is not None else 0, _p(s2),
              _ld(s2) if s2 is not None else 0, _p(q), qs, _p(qp), _p(out), _ld(out), _stream())
    return out

def conv1x1(a, w, bias, epi="bias", s1=None, s2=None, q=None, qp=None, out=None):
    """a [M,K] fp16, w [N,K] fp16 (or [4N,K] for 'quad'), bias [N] fp32."""
    _check(a)
    _check(w)
    M, K = a.shape
    N = w.shape[0] // 4 if epi == "quad" else w.shape[0]
    if out is None:
        out = torch.empty((M, N), dtype=torch.float16, device=a.device)
    bias = bias.float().contiguous()
    w = w.contiguous()
    base = (_p(a), _ld(a), M, K, _p(w), N, _p(bias))
    st = _stream()
    qs = 0 if q is None else (q.stride(0) if q.dim() == 2 else q.numel())
    if epi == "bias":
        _lib_call("dcvc_conv1x1_bias", *base, _p(out), _ld(out), st)
    elif epi == "wsilu":
        _lib_call("dcvc_conv1x1_bias_wsilu", *base, _p(out), _ld(out), st)
    elif epi == "shortcut":
        _lib_call("dcvc_conv1x1_bias_shortcut", *base, _p(s1), _ld(s1), _p(out), _ld(out), st)
    elif epi == "shortcut2":
        _lib_call("dcvc_conv1x1_bias_shortcut2", *base, _p(s1), _ld(s1), _p(s2), _ld(s2), _p(out), _ld(out), st)
    elif epi == "shortcut_quant":
        _lib_call("dcvc_conv1x1_bias_shortcut_with_quant", *base, _p(s1), _ld(s1), _p(q), qs, _p(qp),
                  _p(out), _ld(out), st)
    elif epi == "shortcut_then_quant":
        _lib_call("dcvc_conv1x1_bias_shortcut_then_quant", *base, _p(s1), _ld(s1), _p(q), qs, _p(qp),
                  _p(out), _ld(out), st)
    elif epi == "quant":
        _lib_call("dcvc_conv1x1_bias_with_quant", *base, _p(q), qs, _p(qp), _p(out), _ld(out), st)
    elif epi == "quad":
        _lib_call("dcvc_conv1x1_bias_wsilu_chunk_add", *base, _p(out), _ld(out), st)
    else:
        raise ValueError(epi)
    return out


def gemm_fp8_act(a, wq, scale, bias, epi="bias", a_fp8=True, s1=None, s2=None, q=None, qp=None, out=None,
                 out8=None, fp16_out=True, prec="fp8"):
    """gemm_lowp with e4m3 activations: `a` uint8 e4m3 bytes [M, K] when
    a_fp8 (prec 'fp8' only), else fp16; out8 (uint8 [M, N]) receives an e4m3
    copy of the output, which with fp16_out=False is the only output.
    Returns (out or None, out8 or None)."""
    _check(a, torch.uint8 if a_fp8 else torch.float16)
    M, K = a.shape
    N = wq.shape[0] // 4 if epi == "quad" else wq.shape[0]
    if fp16_out and out is None:
        out = torch.empty((M, N), dtype=torch.float16, device=a.device)
    o = out if fp16_out else None
    bias = bias.float().contiguous()
    qs = 0 if q is None else (q.stride(0) if q.dim() == 2 else q.numel())
    _lib_call("dcvc_gemm_fp8_act", _p(a), _ld(a), 1 if a_fp8 else 0, M, K, _p(wq.contiguous()), N, _p(scale),
              PRECISIONS[prec], EPI_CODES[epi], _p(bias), _p(s1), _ld(s1) if s1 is not None else 0, _p(s2),
              _ld(s2) if s2 is not None else 0, _p(q), qs, _p(qp), _p(o), _ld(o) if o is not None else 0,
              _p(out8), _ld(out8) if out8 is not None else 0, _stream())
    return o, out8


def set_fp8_activations(on):
    """Codec proxies in precision 'fp8': store e4m3 activations between
    products (True, default) or convert fp16 inside each product (False).
    Same bits either way; takes effect at the next graph capture."""
    _lib_call("dcvc_set_fp8_activations", 1 if on else 0)


def fp8_activations():
    return bool(_lib.load().dcvc_get_fp8_activations())


GROUPED_EPIS = {"bias": 0, "wsilu": 1, "shortcut": 2, "shortcut2": 3, "quad": 6}


def conv1x1_grouped(a, w, bias, epi="bias", s1=None, s2=None, a_div=1, out=None):
    """batch problems in one launch: a [batch // a_div, M, K] (problem b reads
    a[b // a_div]), w [batch, rows, K], bias [batch, rows] (rows = 4N for
    'quad'), s1 / s2 / out [batch, M, N]."""
    nb = w.shape[0]
    M, K = a.shape[1], a.shape[2]
    rows = w.shape[1]
    N = rows // 4 if epi == "quad" else rows
    if a.shape[0] * a_div != nb:
        raise ValueError("a must hold batch / a_div problems")
    if out is None:
        out = torch.empty((nb, M, N), dtype=torch.float16, device=a.device)
    w, bias = w.contiguous(), bias.float().contiguous()
    for t in (a, out) + tuple(x for x in (s1, s2) if x is not None):
        assert t.is_contiguous(), "expected contiguous [batch, M, C] tensors"
    _lib_call("dcvc_conv1x1_grouped", _p(a), K, M * K, a_div, M, K, _p(w), rows * K, N, _p(bias), rows,
              GROUPED_EPIS[epi], _p(s1), N, M * N if s1 is not None else 0, _p(s2), N,
              M * N if s2 is not None else 0, _p(out), N, M * N, nb, _stream())
    return out


FFN_TAILS = {"shortcut": 2, "shortcut2": 3, "shortcut_quant": 4}


def depth_conv_ffn(a, w0, b0, w2, b2, tail="shortcut", s2=None, q=None, qp=None, q_round=False, out=None):
    """The ffn half of DepthConvBlock in one launch: a [M,C] fp16, w0 [4CM,C]
    (ffn.0, reference channel order), w2 [C,CM] (ffn.2), fp32 biases.
    out = tail(quad(a w0^T + b0) w2^T + b2 + a); out may be a or s2."""
    _check(a)
    M, C = a.shape
    CM = w2.shape[1]
    if out is None:
        out = torch.empty((M, C), dtype=torch.float16, device=a.device)
    w0, w2 = w0.contiguous(), w2.contiguous()
    b0, b2 = b0.float().contiguous(), b2.float().contiguous()
    qs = 0 if q is None else (q.stride(0) if q.dim() == 2 else q.numel())
    _lib_call("dcvc_depth_conv_ffn", _p(a), _ld(a), M, C, CM, _p(w0), _p(b0), _p(w2), _p(b2), FFN_TAILS[tail],
              _p(s2), _ld(s2) if s2 is not None else 0, _p(q), qs, _p(qp), 1 if q_round else 0, _p(out),
              _ld(out), _stream())
    return out


def conv_bias(x, n_img, H, W, wt_ohwi, bias, kernel, stride, out=None):
    """x [n_img*H*W, Cin]; wt_ohwi [N, k*k*Cin] fp16 tap-major."""
    Cin = wt_ohwi.shape[1] // (kernel * kernel)
    N = wt_ohwi.shape[0]
    Ho = (H + stride - 1) // stride if kernel == 3 else H // 2
    Wo = (W + stride - 1) // stride if kernel == 3 else W // 2
    if kernel == 3 and stride == 1:
        Ho, Wo = H, W
    if out is None:
        out = torch.empty((n_img * Ho * Wo, N), dtype=torch.float16, device=x.device)
    _lib_call("dcvc_conv_bias", _p(x), _ld(x), n_img, H, W, Cin, _p(wt_ohwi.contiguous()), N,
              _p(bias.float().contiguous()), kernel, stride, _p(out), _ld(out), _stream())
    return out


def transposed_conv(x, n_img, H, W, wt_sub, cout, bias=None, out=None):
    """wt_sub [4*Cout, Cin] sub-pixel-major rows; output [n_img*2H*2W, Cout]."""
    Cin = wt_sub.shape[1]
    if bias is None:
        bias = torch.zeros(4 * cout, dtype=torch.float32, device=x.device)
    if out is None:
        out = torch.empty((n_img * 4 * H * W, cout), dtype=torch.float16, device=x.device)
    _lib_call("dcvc_transposed_conv", _p(x), _ld(x), n_img, H, W, Cin, _p(wt_sub.contiguous()), cout,
              _p(bias.float().contiguous()), _p(out), _ld(out), _stream())
    return out


def subpel_conv3x3(x, n_img, H, W, wt_sub, cout, bias, out=None):
    """wt_sub [4*Cout, 9*Cin] sub-pixel-major rows, tap-major K; output [n_img*2H*2W, Cout]."""
    Cin = wt_sub.shape[1] // 9
    if out is None:
        out = torch.empty((n_img * 4 * H * W, cout), dtype=torch.float16, device=x.device)
    _lib_call("dcvc_subpel_conv3x3", _p(x), _ld(x), n_img, H, W, Cin, _p(wt_sub.contiguous()), cout,
              _p(bias.float().contiguous()), _p(out), _ld(out), _stream())
    return out


def ld_quant_step(y, n_img, H, W, scales, means, qdec, yhat, sym, cond, thr, step):
    C = y.shape[1]
    _lib_call("dcvc_ld_quant_step", _p(y), _ld(y), n_img, H, W, C, _p(scales), _ld(scales), _p(means),
              _ld(means), _p(qdec), _ld(qdec), _p(yhat), _ld(yhat), _p(sym), _p(cond),
              _p(thresholds(y.device)), float(thr), step, _stream())


def ld_restore_step(yq, n_img, H, W, C, means, qdec, yhat, step):
    _lib_call("dcvc_ld_restore_step", _p(yq), n_img, H, W, C, _p(means), _ld(means), _p(qdec), _ld(qdec),
              _p(yhat), _ld(yhat), step, _stream())


def htl_quant_step(y, n_img, H, W, scales, means, qdec, ysf, yout, sym, cond, thr, step):
    C = y.shape[1]
    _lib_call("dcvc_htl_quant_step", _p(y), _ld(y), n_img, H, W, C, _p(scales), _ld(scales), _p(means),
              _ld(means), _p(qdec), _ld(qdec), _p(ysf), _ld(ysf), _p(yout), 0 if yout is None else _ld(yout),
              _p(sym), _p(cond), _p(thresholds(y.device)), float(thr), step, _stream())


def htl_restore_step(yq, n_img, H, W, C, means, qdec, ysf, yout, step):
    _lib_call("dcvc_htl_restore_step", _p(yq), n_img, H, W, C, _p(means), _ld(means), _p(qdec), _ld(qdec), _p(ysf),
              _ld(ysf), _p(yout), 0 if yout is None else _ld(yout), step, _stream())


def d3x3_fp8_out(x, n_img, H, W, wt9c, out=None, out8=None, fp16_out=True):
    """d3x3 that also (or only, fp16_out=False) stores an e4m3 copy of its
    output into out8 (uint8 [n_img*H*W, C]); returns (out or None, out8)."""
    C = wt9c.shape[1]
    M = n_img * H * W
    if fp16_out and out is None:
        out = torch.empty((M, C), dtype=torch.float16, device=x.device)
    if out8 is None:
        out8 = torch.empty((M, C), dtype=torch.uint8, device=x.device)
    o = out if fp16_out else None
    _lib_call("dcvc_d3x3_fp8_out", _p(x), _ld(x), n_img, H, W, C, _p(wt9c.contiguous()), _p(o),
              _ld(o) if o is not None else 0, _p(out8), _ld(out8), _stream())
    return o, out8


def d3x3(x, n_img, H, W, wt9c, out=None):
    C = wt9c.shape[1]
    if out is None:
        out = torch.empty((n_img * H * W, C), dtype=torch.float16, device=x.device)
    _lib_call("dcvc_d3x3", _p(x), _ld(x), n_img, H, W, C, _p(wt9c.contiguous()), _p(out), _ld(out), _stream())
    return out


# ----------------------------- layout ---------------------------------------

def pad_and_unshuffle(x, n_img, H, W, cin, f, pad_b, pad_r, out=None):
    Ho, Wo = (H + pad_b) // f, (W + pad_r) // f
    if out is None:
        out = torch.empty((n_img * Ho * Wo, cin * f * f), dtype=torch.float16, device=x.device)
    _lib_call("dcvc_pad_and_unshuffle", _p(x), _ld(x), n_img, H, W, cin, f, pad_b, pad_r, _p(out), _ld(out),
              _stream())
    return out


def pixel_shuffle_8(x, n_img, H8, W8, C, out=None):
    if out is None:
        out = torch.empty((n_img * H8 * 8 * W8 * 8, C), dtype=torch.float16, device=x.device)
    _lib_call("dcvc_pixel_shuffle_8", _p(x), _ld(x), n_img, H8, W8, C, _p(out), _ld(out), _stream())
    return out


def multiply_with_broadcast(x, q, qp=None, out=None):
    M, C = x.shape
    if out is None:
        out = torch.empty_like(x)
    qs = q.stride(0) if q.dim() == 2 else q.numel()
    _lib_call("dcvc_multiply_with_broadcast", _p(x), _ld(x), M, C, _p(q), qs, _p(qp), _p(out), _ld(out), _stream())
    return out


def round_z(z):
    M, C = z.shape
    zhat = torch.empty_like(z)
    zi8 = torch.empty(M * C, dtype=torch.int8, device=z.device)
    _lib_call("dcvc_round_z", _p(z), _ld(z), M, C, _p(zhat), _ld(zhat), _p(zi8), _stream())
    return zhat, zi8


def int8_to_half(zi8, M, C):
    out = torch.empty((M, C), dtype=torch.float16, device=zi8.device)
    _lib_call("dcvc_int8_to_half", _p(zi8), M, C, _p(out), _ld(out), _stream())
    return out


def replicate_pad(x, n_img, H, W, Ho, Wo):
    C = x.shape[1]
    out = torch.empty((n_img * Ho * Wo, C), dtype=torch.float16, device=x.device)
    _lib_call("dcvc_replicate_pad", _p(x), _ld(x), n_img, H, W, C, Ho, Wo, _p(out), _ld(out), _stream())
    return out


# ----------------------------- quantisation ---------------------------------

def hts_quant_step(y, n_img, H, W, scales, means, qdec, yhat, sym, cond, thr, step):
    C = y.shape[1]
    _lib_call("dcvc_hts_quant_step", _p(y), _ld(y), n_img, H, W, C, _p(scales), _ld(scales), _p(means),
              _ld(means), _p(qdec), _ld(qdec), _p(yhat), _ld(yhat), _p(sym), _p(cond),
              _p(thresholds(y.device)), float(thr), step, _stream())


def hts_build_index_dec(scales, n_img, H, W, thr):
    M, C = scales.shape
    idx = torch.empty(M * C, dtype=torch.uint8, device=scales.device)
    cond = torch.empty(M * C, dtype=torch.uint8, device=scales.device)
    _lib_call("dcvc_hts_build_index_dec", _p(scales), _ld(scales), n_img, H, W, C, _p(thresholds(scales.device)),
              float(thr), _p(idx), _p(cond), _stream())
    return idx, cond


def hts_restore_step(yq, n_img, H, W, C, means, qdec, yhat, step):
    _lib_call("dcvc_hts_restore_step", _p(yq), n_img, H, W, C, _p(means), _ld(means), _p(qdec), _ld(qdec),
              _p(yhat), _ld(yhat), step, _stream())


def intra_quant_step(y, n_img, H, W, scales, means, ysf, q, qp, yout, sym, cond, thr, step):
    C = y.shape[1]
    qs = 0 if q is None else (q.stride(0) if q.dim() == 2 else q.numel())
    _lib_call("dcvc_intra_quant_step", _p(y), _ld(y), n_img, H, W, C, _p(scales), _ld(scales), _p(means),
              _ld(means), _p(ysf), _ld(ysf), _p(q), qs, _p(qp), _p(yout), 0 if yout is None else _ld(yout),
              _p(sym), _p(cond), _p(thresholds(y.device)), float(thr), step, _stream())


def intra_build_index_dec(scales, n_img, H, W, thr, step):
    M, C = scales.shape
    n = M * C // 4
    idx = torch.empty(n, dtype=torch.uint8, device=scales.device)
    cond = torch.empty(n, dtype=torch.uint8, device=scales.device)
    _lib_call("dcvc_intra_build_index_dec", _p(scales), _ld(scales), n_img, H, W, C, _p(thresholds(scales.device)),
              float(thr), step, _p(idx), _p(cond), _stream())
    return idx, cond


def intra_restore_step(yq, n_img, H, W, C, means, ysf, q, qp, yout, step):
    qs = 0 if q is None else (q.stride(0) if q.dim() == 2 else q.numel())
    _lib_call("dcvc_intra_restore_step", _p(yq), n_img, H, W, C, _p(means), _ld(means), _p(ysf), _ld(ysf), _p(q),
              qs, _p(qp), _p(yout), 0 if yout is None else _ld(yout), step, _stream())


def conditional_index(vals, cond):
    """Returns (compacted tensor view, count, block_offsets)."""
    N = vals.numel()
    esize = vals.element_size()
    out = torch.empty_like(vals)
    nb = _lib.load().dcvc_compact_blocks(N)
    offs = torch.empty(max(nb, 1), dtype=torch.int32, device=vals.device)
    count = torch.empty(1, dtype=torch.int32, device=vals.device)
    _lib_call("dcvc_conditional_index", _p(vals), esize, _p(cond), N, _p(out), _p(offs), _p(count), _stream())
    n = int(count.item())
    return out[:n], n, offs


def conditional_index_noalloc(vals, cond, out, count):
    """dcvc_conditional_index_noalloc into caller-owned `out` (same size and
    dtype as vals) and `count` (one int32): the entry point allocates
    nothing.  Returns the compacted view and the count."""
    _lib_call("dcvc_conditional_index_noalloc", _p(vals), vals.element_size(), _p(cond), vals.numel(), _p(out),
              _p(count), _stream())
    n = int(count.item())
    return out[:n], n


def conditional_recover(dec, cond, offs):
    N = cond.numel()
    out = torch.empty(N, dtype=torch.int8, device=cond.device)
    _lib_call("dcvc_conditional_recover", _p(dec), _p(cond), N, _p(offs), _p(out), _stream())
    return out
