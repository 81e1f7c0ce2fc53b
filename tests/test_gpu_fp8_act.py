"""FP8 activations in memory (precision form): producers store e4m3 copies
of their outputs, and products read them instead of converting fp16
operands as fragments leave LDS.  The e4m3 bytes follow round to nearest
even at scale 1 with saturation at +-448, and every product and codec gives
the same bits either way."""
import numpy as np
import pytest
import torch

from e4m3_model import e4m3_rne_bytes

pytestmark = pytest.mark.gpu


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def f16(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a).astype(np.float16)).to(dev)


def test_d3x3_fp8_copy_is_rne_of_its_fp16_output(gpu):
    from dcvc_amd import kernels
    n_img, H, W, C = 2, 37, 45, 96
    r = rng(1)
    x = f16(r.standard_normal((n_img * H * W, C)) * np.exp(r.uniform(-9, 4, (n_img * H * W, 1))), gpu)
    wt = f16(r.standard_normal((9, C)), gpu)
    ref = kernels.d3x3(x, n_img, H, W, wt)
    y, y8 = kernels.d3x3_fp8_out(x, n_img, H, W, wt)
    _, y8_only = kernels.d3x3_fp8_out(x, n_img, H, W, wt, fp16_out=False)
    torch.cuda.synchronize()
    assert torch.equal(y.view(torch.int16), ref.view(torch.int16))
    assert torch.equal(y8, y8_only)
    assert np.array_equal(y8.cpu().numpy(), e4m3_rne_bytes(ref.cpu().numpy()))


CASES = [(2999, 256, 192, "bias"), (4096, 512, 256, "wsilu"), (3000, 768, 128, "shortcut"),
         (1000, 256, 256, "shortcut2"), (2048, 512, 128, "quad"), (777, 1024, 64, "quant")]


@pytest.mark.parametrize("M,K,N,epi", CASES)
def test_fp8_activation_product_equals_in_loop_conversion(gpu, M, K, N, epi):
    """A product on e4m3 activations made by a producer equals the product on
    the producer's fp16 output converted in the K loop, bit for bit; its e4m3
    output copy is the RNE of its own fp16 output."""
    from dcvc_amd import kernels
    r = rng(M + K + N)
    nb = 4 * N if epi == "quad" else N
    # the activations come out of a depthwise conv (fp16 + e4m3 copies)
    x = f16(r.standard_normal((M, K)) * np.exp(r.uniform(-8, 3, (M, 1))), gpu)
    wt = f16(r.standard_normal((9, K)) * 0.5, gpu)
    a16, a8 = kernels.d3x3_fp8_out(x, 1, 1, M, wt)
    w = f16(r.standard_normal((nb, K)) / np.sqrt(K), gpu)
    b = torch.from_numpy((r.standard_normal(nb) * 0.1).astype(np.float32)).to(gpu)
    s1 = f16(r.standard_normal((M, N)), gpu)
    s2 = f16(r.standard_normal((M, N)), gpu)
    q = f16(r.uniform(0.5, 2.0, (64, N)), gpu)
    qp = torch.tensor([5], dtype=torch.int32, device=gpu)
    wq, scale = kernels.weights_to_lowp(w, "fp8")
    kw = dict(epi=epi, s1=s1 if epi.startswith("shortcut") else None, s2=s2 if epi == "shortcut2" else None,
              q=q if epi == "quant" else None, qp=qp if epi == "quant" else None)
    ref = kernels.gemm_lowp(a16, wq, scale, "fp8", b, **kw)
    got, got8 = kernels.gemm_fp8_act(a8, wq, scale, b, a_fp8=True,
                                     out8=torch.empty((M, N), dtype=torch.uint8, device=gpu), **kw)
    _, only8 = kernels.gemm_fp8_act(a8, wq, scale, b, a_fp8=True, fp16_out=False,
                                    out8=torch.empty((M, N), dtype=torch.uint8, device=gpu), **kw)
    # fp16 activations in, e4m3 copy out
    got_b, got8_b = kernels.gemm_fp8_act(a16, wq, scale, b, a_fp8=False,
                                         out8=torch.empty((M, N), dtype=torch.uint8, device=gpu), **kw)
    torch.cuda.synchronize()
    assert torch.equal(got.view(torch.int16), ref.view(torch.int16))
    assert torch.equal(got_b.view(torch.int16), ref.view(torch.int16))
    assert torch.equal(got8, only8) and torch.equal(got8, got8_b)
    assert np.array_equal(got8.cpu().numpy(), e4m3_rne_bytes(ref.cpu().numpy()))
    assert np.array_equal(a8.cpu().numpy(), e4m3_rne_bytes(a16.cpu().numpy()))


@pytest.mark.parametrize("epi", ["wsilu", "shortcut", "quad"])
def test_fp8_activation_tile_configs_bit_identical(gpu, epi):
    from dcvc_amd import _lib, kernels
    lib = _lib.load()
    r = rng(3)
    M, K, N = 2999, 512, 200
    nb = 4 * N if epi == "quad" else N
    # e4m3 copy of a16 from a depthwise conv with only its centre tap (= 1)
    a16 = f16(r.standard_normal((M, K)), gpu)
    ident = torch.zeros((9, K), dtype=torch.float16, device=gpu)
    ident[4] = 1
    same, a8 = kernels.d3x3_fp8_out(a16, 1, 1, M, ident)
    assert torch.equal(same, a16)
    w = f16(r.standard_normal((nb, K)) / np.sqrt(K), gpu)
    b = torch.from_numpy((r.standard_normal(nb) * 0.1).astype(np.float32)).to(gpu)
    s1 = f16(r.standard_normal((M, N)), gpu)
    wq, scale = kernels.weights_to_lowp(w, "fp8")
    outs = []
    try:
        for policy in range(0, 15):
            assert lib.dcvc_gemm_set_tile_policy(policy) == 0
            o, _ = kernels.gemm_fp8_act(a8, wq, scale, b, epi=epi, s1=s1 if epi == "shortcut" else None)
            outs.append(o)
    finally:
        lib.dcvc_gemm_set_tile_policy(0)
    ref = kernels.gemm_lowp(a16, wq, scale, "fp8", b, epi=epi, s1=s1 if epi == "shortcut" else None)
    torch.cuda.synchronize()
    for o in outs:
        assert torch.equal(o.view(torch.int16), ref.view(torch.int16))


def test_fp8_activation_arguments_rejected(gpu):
    from dcvc_amd import _lib
    lib = _lib.load()
    a = torch.zeros((64, 128), dtype=torch.uint8, device=gpu)  # room for 64 x 64 fp16 activations too
    w = torch.zeros((64, 128), dtype=torch.uint8, device=gpu)  # room for 64 x 64 fp16 weights too
    s = torch.ones(64, dtype=torch.float32, device=gpu)
    d = torch.zeros((64, 64), dtype=torch.float16, device=gpu)
    d8 = torch.zeros((64, 64), dtype=torch.uint8, device=gpu)
    P = lambda t: None if t is None else t.data_ptr()

    def call(a_fp8=1, prec=2, lda=64, D=d, D8=d8, ldd8=64):
        return lib.dcvc_gemm_fp8_act(P(a), lda, a_fp8, 64, 64, P(w), 64, P(s), prec, 0, P(s), None, 0, None, 0,
                                     None, 0, None, P(D), 64, P(D8), ldd8, None)
    assert call() == 0
    torch.cuda.synchronize()
    assert call(prec=1) == 1          # e4m3 activations need e4m3 weights
    assert call(lda=72) == 1          # e4m3 rows must stay 16-byte aligned
    assert call(D=None, D8=None) == 1  # no output at all
    assert call(ldd8=60) == 1
    assert call(a_fp8=2) == 1
    assert call(a_fp8=0, prec=0, lda=64) == 1  # e4m3 output copies come from fp8 products only
    assert call(a_fp8=0, prec=0, lda=64, D8=None) == 0


@pytest.mark.parametrize("model", ["hts", "htl", "ld"])
def test_codec_fp8_activations_same_bits(gpu, model):
    """Each codec in precision fp8 codes and reconstructs the same bits with
    e4m3 activations stored between products as with fp16 activations
    converted inside every product."""
    from dcvc_amd import codec, kernels, synthetic
    H, W = 96, 160
    sdi, sdp = codec.make_state_dicts(seed=0, video=model)
    cls = codec.video_codec(model)
    pad_r, pad_b = synthetic.get_padding_size(H, W, 16)
    frames = synthetic.motion_sequence_yuv420(H, W, 9, seed=2)
    x0 = codec.to_input(synthetic.frame_to_input(frames[0]), gpu)
    if cls.FRAME_DELAY == 1:
        chunks = [codec.to_input(synthetic.frame_to_input(frames[i]), gpu) for i in (1, 2)]
    else:
        chunks = [codec.to_input(synthetic.chunk_to_input(frames[1 + 4 * c:5 + 4 * c], cls.FRAME_DELAY), gpu)
                  for c in range(2)]
    runs = []
    try:
        for on in (False, True):
            kernels.set_fp8_activations(on)
            assert kernels.fp8_activations() == on
            ie, idc = codec.IntraCodec(sdi, 0.15), codec.IntraCodec(sdi, 0.15)
            pe = cls(sdp, 0.15)
            pd = cls(share_from=pe)
            for o in (ie, idc, pe, pd):
                o.proxy.set_precision("fp8")
            e = ie.compress(x0, 32, pad_b, pad_r)
            d = idc.decompress(e["bit_stream"], {"height": H, "width": W}, 32, e["ec_parallel"])
            pe.add_ref_feature_from_frame(e["x_hat"])
            pd.add_ref_feature_from_frame(d["x_hat"], apply_feature_adaptor=False)
            res = [e["bit_stream"], d["x_hat"].cpu().numpy().tobytes()]
            for i, x in enumerate(chunks):
                c = pe.compress(x, 21 + 10 * i, 0, pad_b, pad_r)
                out = pd.decompress(c["bit_stream"], {"height": H, "width": W}, 21 + 10 * i, c["ec_parallel"], 0)
                xs = out["x_hat"] if isinstance(out["x_hat"], (list, tuple)) else [out["x_hat"]]
                res.append(c["bit_stream"])
                res.extend(t.cpu().numpy().tobytes() for t in xs)
            torch.cuda.synchronize()
            runs.append(res)
            del ie, idc, pe, pd
    finally:
        kernels.set_fp8_activations(True)
    assert len(runs[0]) == len(runs[1])
    for a, b in zip(runs[0], runs[1]):
        assert a == b
