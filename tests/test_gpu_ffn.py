"""GPU parity of the fused DepthConvBlock FFN (dcvc_depth_conv_ffn).

* Bit-exact against the two launches it replaces, conv1x1_bias_wsilu_chunk_add
  then conv1x1_bias_shortcut / _shortcut2 / _shortcut_with_quant (the
  reference proxy's sequence, layers_proxy.cpp:84-97), for every tail, both
  supported widths, ragged pixel counts and in-place outputs.
* Against a float64 restatement of layers.py:114-159 (ffn.0, WSiLU chunk add,
  ffn.2, shortcut) with the hidden activation rounded to fp16 as the proxy
  stores it.
* Codec level: HT-S and HT-L bitstreams and reconstructions are identical with
  the fused path forced on every DepthConvBlock and with it off.
"""
import numpy as np
import pytest
import torch

from dcvc_amd import _lib, entropy, synthetic, weights

pytestmark = pytest.mark.gpu


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def f16(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a).astype(np.float16)).to(dev)


@pytest.fixture
def fusion_mode():
    """Set the DepthConvBlock fusion mode for one test, restoring the default (off)."""
    def set_mode(m):
        _lib.call("dcvc_set_ffn_fusion", m)
    yield set_mode
    _lib.call("dcvc_set_ffn_fusion", 0)


def _ffn_operands(gpu, M, C, CM, seed):
    r = rng(seed)
    a = f16(r.standard_normal((M, C)), gpu)
    w0 = f16(r.standard_normal((4 * CM, C)) / np.sqrt(C), gpu)
    b0 = torch.from_numpy((r.standard_normal(4 * CM) * 0.1).astype(np.float32)).to(gpu)
    w2 = f16(r.standard_normal((C, CM)) / np.sqrt(CM), gpu)
    b2 = torch.from_numpy((r.standard_normal(C) * 0.1).astype(np.float32)).to(gpu)
    s2 = f16(r.standard_normal((M, C)), gpu)
    qtab = f16(0.5 + r.random((64, C)), gpu)
    qp = torch.tensor([11], dtype=torch.int32, device=gpu)
    return a, w0, b0, w2, b2, s2, qtab, qp


def _two_launches(kernels, a, w0, b0, w2, b2, tail, s2, qtab, qp, q_round):
    t1 = kernels.conv1x1(a, w0, b0, epi="quad")
    if tail == "shortcut":
        return kernels.conv1x1(t1, w2, b2, epi="shortcut", s1=a)
    if tail == "shortcut2":
        return kernels.conv1x1(t1, w2, b2, epi="shortcut2", s1=a, s2=s2)
    epi = "shortcut_then_quant" if q_round else "shortcut_quant"
    return kernels.conv1x1(t1, w2, b2, epi=epi, s1=a, q=qtab, qp=qp)


CASES = [(512, 256, 32640), (512, 256, 1000), (256, 256, 32640), (256, 256, 129), (512, 256, 1), (256, 256, 128)]


@pytest.mark.parametrize("C,CM,M", CASES)
@pytest.mark.parametrize("tail,q_round", [("shortcut", False), ("shortcut2", False), ("shortcut_quant", False),
                                          ("shortcut_quant", True)])
def test_depth_conv_ffn_equals_two_launches(gpu, C, CM, M, tail, q_round):
    from dcvc_amd import kernels
    print(f"ffn C={C} CM={CM} M={M} tail={tail} q_round={q_round}", flush=True)
    a, w0, b0, w2, b2, s2, qtab, qp = _ffn_operands(gpu, M, C, CM, C + M)
    ref = _two_launches(kernels, a, w0, b0, w2, b2, tail, s2, qtab, qp, q_round)
    got = kernels.depth_conv_ffn(a, w0, b0, w2, b2, tail=tail, s2=s2, q=qtab, qp=qp, q_round=q_round)
    torch.cuda.synchronize()
    assert torch.isfinite(got.float()).all()
    diff = int((got.view(torch.int16) != ref.view(torch.int16)).sum())
    assert diff == 0, f"{diff} of {got.numel()} elements differ from the two-launch path"


@pytest.mark.parametrize("C,inplace", [(512, "a"), (512, "s2"), (256, "a")])
def test_depth_conv_ffn_in_place(gpu, C, inplace):
    """The DepthConvBlock call patterns: out is the block input o (no block
    shortcut) or the block input x (block shortcut, out = x)."""
    from dcvc_amd import kernels
    M, CM = 32640, 256
    a, w0, b0, w2, b2, s2, qtab, qp = _ffn_operands(gpu, M, C, CM, 7 * C)
    tail = "shortcut2" if inplace == "s2" else "shortcut"
    ref = _two_launches(kernels, a, w0, b0, w2, b2, tail, s2, qtab, qp, False)
    dst = (a if inplace == "a" else s2).clone()
    if inplace == "a":
        got = kernels.depth_conv_ffn(dst, w0, b0, w2, b2, tail=tail, out=dst)
    else:
        got = kernels.depth_conv_ffn(a, w0, b0, w2, b2, tail=tail, s2=dst, out=dst)
    torch.cuda.synchronize()
    assert got.data_ptr() == dst.data_ptr()
    assert torch.equal(dst.view(torch.int16), ref.view(torch.int16))


def test_depth_conv_ffn_vs_float64(gpu):
    """layers.py:114-159 in float64: t1 = sum_j wsilu(o W0[4n+j] + b0), rounded
    to fp16 as the proxy stores it; out = t1 W2^T + b2 + o (+ S2).  Bound: the
    fp16 output ulp, the propagated fp16 rounding of t1 (|W2| . ulp(t1)) and
    fp32 accumulation error over both products."""
    from dcvc_amd import kernels
    M, C, CM = 777, 512, 256
    a, w0, b0, w2, b2, s2, qtab, qp = _ffn_operands(gpu, M, C, CM, 99)
    got = kernels.depth_conv_ffn(a, w0, b0, w2, b2, tail="shortcut2", s2=s2).double().cpu().numpy()
    A, W0, B0 = a.double().cpu().numpy(), w0.double().cpu().numpy(), b0.double().cpu().numpy()
    W2, B2, S2 = w2.double().cpu().numpy(), b2.double().cpu().numpy(), s2.double().cpu().numpy()
    pre = A @ W0.T + B0
    act = pre / (1.0 + np.exp(-4.0 * pre))
    t1 = act[:, 0::4] + act[:, 1::4] + act[:, 2::4] + act[:, 3::4]
    # the channel order of layers.py:119-125 is x[:, j::4]: row 4n + j feeds hidden channel n
    t1h = t1.astype(np.float16).astype(np.float64)
    ref = t1h @ W2.T + B2 + A + S2

    def ulp16(x):
        e = np.floor(np.log2(np.maximum(np.abs(x), 2.0 ** -14)))
        return 2.0 ** (e - 10)

    acc1 = np.abs(A) @ np.abs(W0).T
    t1_err = ulp16(t1) + 4e-6 * (np.abs(acc1[:, 0::4]) + np.abs(acc1[:, 1::4]) + np.abs(acc1[:, 2::4])
                                 + np.abs(acc1[:, 3::4])) + 1e-6 * np.abs(t1)
    bound = ulp16(ref) + t1_err @ np.abs(W2).T + 2e-6 * (np.abs(t1h) @ np.abs(W2).T) + 1e-6 * np.abs(ref)
    err = np.abs(got - ref)
    assert (err <= bound).all(), f"worst err {err.max():.3g}, {int((err > bound).sum())} outside the bound"


def test_depth_conv_ffn_rejects_bad_arguments(gpu):
    from dcvc_amd import kernels
    a, w0, b0, w2, b2, s2, qtab, qp = _ffn_operands(gpu, 256, 512, 256, 1)
    lib = _lib.load()
    # unsupported hidden width, unsupported tail, missing second shortcut
    assert lib.dcvc_depth_conv_ffn(a.data_ptr(), 512, 256, 512, 128, w0.data_ptr(), b0.data_ptr(), w2.data_ptr(),
                                   b2.data_ptr(), 2, None, 0, None, 0, None, 0, a.data_ptr(), 512, None) == 1
    assert lib.dcvc_depth_conv_ffn(a.data_ptr(), 512, 256, 512, 256, w0.data_ptr(), b0.data_ptr(), w2.data_ptr(),
                                   b2.data_ptr(), 6, None, 0, None, 0, None, 0, a.data_ptr(), 512, None) == 1
    assert lib.dcvc_depth_conv_ffn(a.data_ptr(), 512, 256, 512, 256, w0.data_ptr(), b0.data_ptr(), w2.data_ptr(),
                                   b2.data_ptr(), 3, None, 0, None, 0, None, 0, a.data_ptr(), 512, None) == 1
    assert lib.dcvc_set_ffn_fusion(3) == 1
    del kernels


# ------------------------------------------------------------------ codec level

H, W = 128, 192


def _state(model, seed):
    sd = weights.perturb_biases(weights.init_state_dict(model, seed), seed + 100)
    sd = weights.round_to_half(sd)
    return entropy.add_cdf_to_state_dict(weights.to_torch(sd))


def _nchw(arr, gpu):
    return torch.from_numpy(arr)[None].half().to(gpu).contiguous(memory_format=torch.channels_last)


def _chunk_run(gpu, proxy_cls, sdp, x_ref, chunks, mode, set_mode):
    """Encode the chunks with fusion `mode`, decode them with the other mode;
    returns the bitstreams and the decoder's reconstructions."""
    set_mode(mode)
    enc = proxy_cls()
    enc.set_param(sdp, 0.0)
    enc.add_ref_feature_from_frame(x_ref, True)
    streams = []
    for x in chunks:
        bits, ec = enc.compress(x, 32, False, 0, 0)
        streams.append((bits.copy(), ec))
    torch.cuda.synchronize()
    set_mode(2 - mode)
    dec = proxy_cls()
    dec.set_param(sdp, 0.0)
    dec.add_ref_feature_from_frame(x_ref, False)
    recon = []
    for bits, ec in streams:
        outs = dec.decompress(bits, 32, H, W, ec, False)
        recon.append(torch.stack([f.clone() for f in outs]))
    torch.cuda.synchronize()
    return [s for s, _ in streams], recon


@pytest.mark.parametrize("model", ["hts", "htl"])
def test_codec_identical_with_and_without_fusion(gpu, fusion_mode, model):
    from dcvc_amd.inference_extensions_cuda import DMCHTLProxy, DMCHTSProxy, DMCIProxy
    print(f"codec fusion parity {model}", flush=True)
    proxy_cls = DMCHTSProxy if model == "hts" else DMCHTLProxy
    sdi, sdp = _state("dmci", 0), _state(model, 1)
    frames = synthetic.motion_sequence_yuv420(H, W, 9, seed=5)
    fusion_mode(0)
    intra = DMCIProxy()
    intra.set_param(sdi, 0.0)
    _, x_ref, _ = intra.compress(_nchw(synthetic.frame_to_input(frames[0]), gpu), 32, 0, 0)
    x_ref = x_ref.clone()
    chunks = [_nchw(synthetic.chunk_to_input(frames[1:9]), gpu), _nchw(synthetic.chunk_to_input(frames[1:9][::-1]), gpu)]
    bits0, rec0 = _chunk_run(gpu, proxy_cls, sdp, x_ref, chunks, 0, fusion_mode)
    bits2, rec2 = _chunk_run(gpu, proxy_cls, sdp, x_ref, chunks, 2, fusion_mode)
    for c in range(len(chunks)):
        assert np.array_equal(bits0[c], bits2[c]), f"chunk {c}: bitstream differs with the fused FFN"
        assert torch.equal(rec0[c].view(torch.int16), rec2[c].view(torch.int16)), f"chunk {c}: reconstruction differs"
