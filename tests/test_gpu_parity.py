"""Parity of the GPU codecs with the FP32 oracle, measured against the floor
that fp16 activation storage sets.

The GPU path stores every kernel output in fp16 (as the reference's CUDA
path does) and computes inside each kernel in fp32.  The oracle runs the
same codec three ways on the same inputs (oracle.storage_mode):

* fp32        the FP32 restatement of the reference;
* fp16        fp16 rounding at exactly the GPU library's storage points;
* fp16_acc64  the same, with every convolution summed in float64 first: an
              equally valid fp32 summation order.

Agreement of fp16 with fp32 (and of fp16_acc64 with fp32) is how close an
exact fp16-storage implementation gets to the FP32 path; agreement of fp16
with fp16_acc64 is how far two such implementations drift apart through
summation order alone (rounding ties, z at .5, scales at CDF-index bin
edges).  Each test asserts that the GPU is within twice those floors, that
is no further from the FP32 path than fp16 storage itself puts any
implementation, and that PSNR and bitstream size stay within 0.02 dB and
0.5 % of the FP32 path.  A block-level test pins where the oracle's fp16
roundings sit: the GPU's DepthConvBlocks reproduce the fp16 mode bit for bit
on nearly every element.
"""
import numpy as np
import pytest
import torch

from dcvc_amd import entropy, synthetic, weights
from oracle import dcvc_oracle as O

pytestmark = pytest.mark.gpu

H, W = 128, 192
MODES = {"fp32": {}, "fp16": {"fp16": True}, "fp16_acc64": {"fp16": True, "acc64": True}}


def _state(model, seed):
    sd = weights.perturb_biases(weights.init_state_dict(model, seed), seed + 100)
    return entropy.add_cdf_to_state_dict(weights.to_torch(weights.round_to_half(sd)))


def _f32(sd):
    return {k: v.float() for k, v in sd.items()}


def _gpu_nchw(arr, gpu):
    return torch.from_numpy(arr)[None].half().to(gpu).contiguous(memory_format=torch.channels_last)


def _psnr(a, b):
    mse = float(((a.double() - b.double()) ** 2).mean())
    return 10 * np.log10(1.0 / max(mse, 1e-12))


def _agree(a, b):
    """(symbol, index) agreement of two lists of dense int16 (y_q << 8) + idx arrays."""
    a = np.concatenate([x.reshape(-1) for x in a]).astype(np.int64)
    b = np.concatenate([x.reshape(-1) for x in b]).astype(np.int64)
    return float(np.mean((a >> 8) == (b >> 8))), float(np.mean((a & 0xFF) == (b & 0xFF)))


def _coded_bytes(sdp, sym_streams, z_int8, qp, ec):
    enc = entropy.RansEncoder()
    enc.set_cdf(sdp["bit_estimator_z.quantized_cdf"].numpy(), sdp["bit_estimator_z.cdf_length"].numpy(), 0)
    enc.set_cdf(sdp["gaussian_encoder.quantized_cdf"].numpy(), sdp["gaussian_encoder.cdf_length"].numpy(), 1)
    enc.reset()
    enc.set_entropy_coder_parallel(ec)
    for s in sym_streams:
        enc.encode_y(s)
    enc.encode_z(z_int8, qp * 128, 128)
    enc.flush()
    return len(enc.get_encoded_stream())


def _check_floor(name, gpu, oracle):
    """gpu: list of dense symbol arrays; oracle: mode -> the same from the oracle."""
    a = _agree(gpu, oracle["fp32"])
    b = _agree(gpu, oracle["fp16"])
    f1 = _agree(oracle["fp16"], oracle["fp32"])
    f2 = _agree(oracle["fp16"], oracle["fp16_acc64"])
    f3 = _agree(oracle["fp16_acc64"], oracle["fp32"])
    n = sum(x.size for x in gpu)
    print(f"{name} (N={n}) [symbols, indexes]: GPU~fp32 {a}  GPU~fp16 {b} | floors: fp16~fp32 {f1}  "
          f"fp16~fp16_acc64 {f2}  fp16_acc64~fp32 {f3}", flush=True)
    slack = 5.0 / n
    for i, what in enumerate(("symbols", "indexes")):
        floor = min(f1[i], f3[i])
        assert 1 - a[i] <= 2 * (1 - floor) + slack, f"{name} {what}: GPU~fp32 {a[i]} below the storage floor {floor}"
        assert 1 - b[i] <= 2 * (1 - f2[i]) + slack, f"{name} {what}: GPU~fp16 {b[i]} below the order floor {f2[i]}"
    return a, b, f1, f2, f3


@pytest.fixture(scope="module")
def intra(gpu):
    """One DMCI frame on the GPU (encoder and decoder) and the oracle's three runs."""
    from dcvc_amd.inference_extensions_cuda import DMCIProxy
    sdi = _state("dmci", 0)
    frame = synthetic.frame_to_input(synthetic.texture_yuv420(H, W, 0))
    enc, dec = DMCIProxy(), DMCIProxy()
    enc.set_param(sdi, 0.0)
    dec.set_param(sdi, 0.0)
    bits, x_hat_e, ec = enc.compress(_gpu_nchw(frame, gpu), 32, 0, 0)
    x_hat_e = x_hat_e.clone()
    x_hat_d = dec.decompress(bits, 32, H, W, ec).clone()
    nq = (H // 16) * (W // 16) * 64
    sym = [enc.debug_copy(k, nq * 2, np.int16).copy() for k in range(4)]
    torch.cuda.synchronize()
    sdo = _f32(sdi)
    xin = torch.from_numpy(frame)[None].half().float()
    runs = {}
    for m, kw in MODES.items():
        with torch.inference_mode(), O.storage_mode(**kw):
            runs[m] = O.dmci_compress(sdo, xin, 32)
    return dict(sdi=sdi, frame=frame, bits=bits, ec=ec, x_hat_e=x_hat_e, x_hat_d=x_hat_d, sym=sym, runs=runs)


def test_dmci_parity_within_the_fp16_storage_floor(intra):
    r = intra["runs"]
    assert torch.equal(intra["x_hat_e"], intra["x_hat_d"])
    _check_floor("DMCI", intra["sym"], {m: r[m].steps_dense for m in MODES})
    src = torch.from_numpy(intra["frame"])[None]
    pg, po = _psnr(intra["x_hat_e"].float().cpu(), src), _psnr(r["fp32"].x_hat, src)
    n_o = _coded_bytes(intra["sdi"], r["fp32"].steps_sym16[::-1], r["fp32"].z_int8, 32, intra["ec"])
    print(f"DMCI PSNR gpu {pg:.4f} fp32 {po:.4f}; bytes gpu {len(intra['bits'])} fp32 {n_o}")
    assert abs(pg - po) <= 0.02
    assert abs(len(intra["bits"]) - n_o) <= 0.005 * n_o


def _p_codec(gpu, model, seed, x_hat_i, chunk, qp, thr=0.0):
    from dcvc_amd.inference_extensions_cuda import DMCHTLProxy, DMCHTSProxy, DMCLDProxy
    cls = {"hts": DMCHTSProxy, "htl": DMCHTLProxy, "ld": DMCLDProxy}[model]
    sdp = _state(model, seed)
    enc, dec = cls(), cls()
    enc.set_param(sdp, thr)
    dec.set_param(sdp, thr)
    enc.add_ref_feature_from_frame(x_hat_i, True)
    dec.add_ref_feature_from_frame(x_hat_i, False)
    bits, ec = enc.compress(_gpu_nchw(chunk, gpu), qp, False, 0, 0)
    return sdp, enc, dec, bits, ec


def test_hts_parity_within_the_fp16_storage_floor(gpu, intra):
    frames = synthetic.motion_sequence_yuv420(H, W, 9, seed=4)
    chunk = synthetic.chunk_to_input(frames[1:9])
    x_hat_i = intra["x_hat_e"]
    sdp, enc, dec, bits, ec = _p_codec(gpu, "hts", 1, x_hat_i, chunk, 32)
    ny = (H // 16) * (W // 16) * 256
    sym = [enc.debug_copy(5, ny * 2, np.int16).copy()]
    outs = [o.float().cpu() for o in dec.decompress(bits, 32, H, W, ec, False)]
    sdo = _f32(sdp)
    ref = x_hat_i.float().cpu().contiguous()
    xin = torch.from_numpy(chunk)[None].half().float()
    runs = {}
    for m, kw in MODES.items():
        with torch.inference_mode(), O.storage_mode(**kw):
            st = O.HTSState()
            O.hts_add_ref_feature_from_frame(sdo, st, ref, True)
            runs[m] = O.hts_compress(sdo, st, xin, 32, False, recon=True)
    _check_floor("HT-S", sym, {m: [runs[m].sym16_dense] for m in MODES})
    n_o = _coded_bytes(sdp, [runs["fp32"].sym16], runs["fp32"].z_int8, 32, ec)
    dp = [abs(_psnr(outs[i], torch.from_numpy(chunk[3 * i:3 * i + 3])[None])
              - _psnr(runs["fp32"].x_hat[i], torch.from_numpy(chunk[3 * i:3 * i + 3])[None])) for i in range(8)]
    print(f"HT-S bytes gpu {len(bits)} fp32 {n_o}; max PSNR delta {max(dp):.4f} dB")
    assert max(dp) <= 0.02
    assert abs(len(bits) - n_o) <= 0.005 * n_o


def test_htl_parity_within_the_fp16_storage_floor(gpu, intra):
    frames = synthetic.motion_sequence_yuv420(H, W, 9, seed=6)
    chunk = synthetic.chunk_to_input(frames[1:9])
    x_hat_i = intra["x_hat_e"]
    sdp, enc, dec, bits, ec = _p_codec(gpu, "htl", 2, x_hat_i, chunk, 32)
    nq = (H // 16) * (W // 16) * 64
    sym = [enc.debug_copy(5 + k, nq * 2, np.int16).copy() for k in range(4)]
    outs = [o.float().cpu() for o in dec.decompress(bits, 32, H, W, ec, False)]
    sdo = _f32(sdp)
    ref = x_hat_i.float().cpu().contiguous()
    xin = torch.from_numpy(chunk)[None].half().float()
    runs = {}
    for m, kw in MODES.items():
        with torch.inference_mode(), O.storage_mode(**kw):
            st = O.HTSState()
            O.htl_add_ref_feature_from_frame(sdo, st, ref, True)
            runs[m] = O.htl_compress(sdo, st, xin, 32, False, recon=True)
    _check_floor("HT-L", sym, {m: runs[m].steps_dense for m in MODES})
    nb = {m: _coded_bytes(sdp, runs[m].steps_sym16[::-1], runs[m].z_int8, 32, ec) for m in MODES}
    dp = [abs(_psnr(outs[i], torch.from_numpy(chunk[3 * i:3 * i + 3])[None])
              - _psnr(runs["fp32"].x_hat[i], torch.from_numpy(chunk[3 * i:3 * i + 3])[None])) for i in range(8)]
    print(f"HT-L bytes gpu {len(bits)} {nb}; max PSNR delta {max(dp):.4f} dB")
    assert max(dp) <= 0.02
    # bitstream size: within 0.5 %, or within twice the fp16 implementations' own spread
    spread = max(abs(nb["fp16"] - nb["fp32"]), abs(nb["fp16_acc64"] - nb["fp32"]))
    assert abs(len(bits) - nb["fp32"]) <= max(0.005 * nb["fp32"], 2 * spread)


def test_ld_parity_within_the_fp16_storage_floor(gpu, intra):
    frames = synthetic.motion_sequence_yuv420(H, W, 2, seed=10)
    frame = synthetic.frame_to_input(frames[1])
    x_hat_i = intra["x_hat_e"]
    sdp, enc, dec, bits, ec = _p_codec(gpu, "ld", 3, x_hat_i, frame, 32)
    ny = (H // 16) * (W // 16) * 128
    sym = [enc.debug_copy(5, ny * 2, np.int16).copy()]
    x_hat = dec.decompress(bits, 32, H, W, ec, False).float().cpu()
    sdo = _f32(sdp)
    ref = x_hat_i.float().cpu().contiguous()
    xin = torch.from_numpy(frame)[None].half().float()
    runs = {}
    for m, kw in MODES.items():
        with torch.inference_mode(), O.storage_mode(**kw):
            st = O.LDState()
            O.ld_add_ref_feature_from_frame(sdo, st, ref, True)
            runs[m] = O.ld_compress(sdo, st, xin, 32, False, recon=True)
    _check_floor("LD", sym, {m: [runs[m].sym16_dense] for m in MODES})
    n_o = _coded_bytes(sdp, [runs["fp32"].sym16], runs["fp32"].z_int8, 32, ec)
    src = torch.from_numpy(frame)[None]
    dp = abs(_psnr(x_hat, src) - _psnr(runs["fp32"].x_hat, src))
    print(f"LD bytes gpu {len(bits)} fp32 {n_o}; PSNR delta {dp:.4f} dB")
    assert dp <= 0.02
    assert abs(len(bits) - n_o) <= 0.005 * n_o


def _gpu_dcb(sd, p, x, H8, W8, gpu):
    """One DepthConvBlock as the library's DCB::forward launches it: dc.0 +
    WSiLU, depthwise (bias folded into dc.3 in float64), dc.3 + shortcut,
    quad, ffn.2 + shortcut."""
    from dcvc_amd import kernels
    f16 = lambda t: t.half().to(gpu).contiguous()
    w0 = f16(sd[p + "dc.0.weight"][:, :, 0, 0])
    wdw = sd[p + "dc.2.weight"]
    C = wdw.shape[0]
    w3 = sd[p + "dc.3.weight"][:, :, 0, 0]
    fold = (sd[p + "dc.3.bias"].double() + w3.double() @ sd[p + "dc.2.bias"].double()).float()
    t1 = kernels.conv1x1(x, w0, sd[p + "dc.0.bias"].to(gpu), epi="wsilu")
    t2 = kernels.d3x3(t1, 1, H8, W8, f16(wdw.reshape(C, 9).t()))
    o = kernels.conv1x1(t2, f16(w3), fold.to(gpu), epi="shortcut", s1=x)
    t3 = kernels.conv1x1(o, f16(sd[p + "ffn.0.weight"][:, :, 0, 0]), sd[p + "ffn.0.bias"].to(gpu), epi="quad")
    return kernels.conv1x1(t3, f16(sd[p + "ffn.2.weight"][:, :, 0, 0]), sd[p + "ffn.2.bias"].to(gpu),
                           epi="shortcut", s1=o)


def test_blocks_reproduce_the_fp16_storage_oracle(gpu, intra):
    """Where the fp16 mode rounds is where the GPU stores.  One
    DepthConvBlock launched kernel by kernel, and the HT-S feature extractor
    (5 blocks) run by the proxy, on the proxy's own memory tensor: the GPU
    output equals the fp16-mode oracle bit for bit far more often than the
    FP32 oracle rounded once at the end.  (The rest are 1-ulp flips where
    the two fp32 summation orders straddle an fp16 rounding point; a flip
    spreads to the next 1x1 conv's outputs, so the chain keeps fewer.)"""
    from dcvc_amd.inference_extensions_cuda import DMCHTSProxy
    sdp = _state("hts", 1)
    p = DMCHTSProxy()
    p.set_param(sdp, 0.0)
    p.add_ref_feature_from_frame(intra["x_hat_e"], True)
    H8, W8 = H // 8, W // 8
    P8 = H8 * W8
    mem = p.debug_copy(3, P8 * 1024 * 2, np.float16).reshape(P8, 1024)[:, :512].copy()
    ctx = p.debug_copy(4, P8 * 2048 * 2, np.float16).reshape(P8, 2048)[:, 1536:].copy()
    one = _gpu_dcb(_f32(sdp), "feature_extractor.conv.0.", torch.from_numpy(mem).to(gpu), H8, W8, gpu)
    one = one.cpu().numpy()
    torch.cuda.synchronize()
    m = torch.from_numpy(mem.astype(np.float32).reshape(1, H8, W8, 512)).permute(0, 3, 1, 2).contiguous()
    sdo = _f32(sdp)
    exact = {}
    for name, kw in (("fp16", {"fp16": True}), ("fp32", {})):
        with torch.inference_mode(), O.storage_mode(**kw):
            b = O.dcb(sdo, "feature_extractor.conv.0.", m)
            c = O.hts_feature_extractor(sdo, m)
        nhwc = lambda t: t.half().permute(0, 2, 3, 1).reshape(P8, 512).numpy().view(np.uint16)
        exact[name] = (float(np.mean(nhwc(b) == one.view(np.uint16))), float(np.mean(nhwc(c) == ctx.view(np.uint16))))
    print(f"bit-exact elements [one block, 5-block chain]: fp16 mode {exact['fp16']}, "
          f"fp32 rounded once {exact['fp32']}")
    for i in range(2):
        assert 1 - exact["fp16"][i] < 0.6 * (1 - exact["fp32"][i])
    assert exact["fp16"][0] > 0.9


def test_hts_1080p_chunks_bitexact_and_within_the_floor(gpu):
    """The headline workload at its size: 1920x1080, an intra frame, then two
    chunks of 4 real frames padded to the native 8 by repeating the last
    (test_video.py:104-110), qp 32.  Encoder and decoder agree bit for bit
    (y_hat, the next chunk's context, symbol counts, reconstructions), and
    the first chunk's symbols sit within the fp16 storage floor of the FP32
    oracle."""
    from dcvc_amd.inference_extensions_cuda import DMCHTSProxy, DMCIProxy
    Hh, Wh = 1080, 1920
    pad_r, pad_b = synthetic.get_padding_size(Hh, Wh, 16)
    Hp, Wp = Hh + pad_b, Wh + pad_r
    print(f"HT-S 1080p: {Hh}x{Wh} padded to {Hp}x{Wp}", flush=True)
    sdi, sdp = _state("dmci", 0), _state("hts", 1)
    frames = synthetic.motion_sequence_yuv420(Hh, Wh, 9, seed=7)
    ie, idec = DMCIProxy(), DMCIProxy()
    ie.set_param(sdi, 0.0)
    idec.set_param(sdi, 0.0)
    bits_i, x_hat_e, ec_i = ie.compress(_gpu_nchw(synthetic.frame_to_input(frames[0]), gpu), 32, pad_b, pad_r)
    x_hat_e = x_hat_e.clone()
    x_hat_d = idec.decompress(bits_i, 32, Hh, Wh, ec_i).clone()
    assert torch.equal(x_hat_e, x_hat_d)
    enc, dec = DMCHTSProxy(), DMCHTSProxy()
    enc.set_param(sdp, 0.0)
    dec.set_param(sdp, 0.0)
    enc.add_ref_feature_from_frame(x_hat_e, True)
    dec.add_ref_feature_from_frame(x_hat_d, False)
    P16, P8 = (Hp // 16) * (Wp // 16), (Hp // 8) * (Wp // 8)
    first = None
    for c, src in enumerate((frames[1:5], frames[5:9])):
        chunk = synthetic.chunk_to_input(src, 8)
        bits, ec = enc.compress(_gpu_nchw(chunk, gpu), 32, False, pad_b, pad_r)
        yhat_e = enc.debug_copy(0, P16 * 512 * 2, np.float16).reshape(P16, 512)[:, :256].copy()
        feat_e = enc.debug_copy(3, P8 * 1024 * 2, np.float16).reshape(P8, 1024)[:, 512:].copy()
        sym_e = enc.debug_copy(5, P16 * 256 * 2, np.int16).copy()
        n_e = enc.last_symbol_count()
        outs = dec.decompress(bits, 32, Hh, Wh, ec, False)
        torch.cuda.synchronize()
        yhat_d = dec.debug_copy(0, P16 * 512 * 2, np.float16).reshape(P16, 512)[:, :256]
        feat_d = dec.debug_copy(3, P8 * 1024 * 2, np.float16).reshape(P8, 1024)[:, 512:]
        assert dec.last_symbol_count() == n_e
        assert np.array_equal(yhat_e.view(np.uint16), yhat_d.view(np.uint16)), f"chunk {c}: y_hat"
        assert np.array_equal(feat_e.view(np.uint16), feat_d.view(np.uint16)), f"chunk {c}: feature"
        for f in outs:
            assert f.shape == (1, 3, Hp, Wp) and torch.isfinite(f.float()).all()
        # frames 4-7 repeat frame 3, so their reconstructions come from the same chunk input
        print(f"chunk {c}: {len(bits)} bytes, {n_e} y symbols, ec {ec}", flush=True)
        if c == 0:
            first = (chunk, sym_e, bits, ec)
    chunk, sym_e, bits, ec = first
    sdo = _f32(sdp)
    ref = x_hat_e.float().cpu().contiguous()
    xin = torch.from_numpy(chunk)[None].half().float()
    runs = {}
    for m in ("fp32", "fp16"):
        with torch.inference_mode(), O.storage_mode(**MODES[m]):
            st = O.HTSState()
            O.hts_add_ref_feature_from_frame(sdo, st, ref, True)
            runs[m] = O.hts_compress(sdo, st, xin, 32, False)
    a = _agree([sym_e], [runs["fp32"].sym16_dense])
    f1 = _agree([runs["fp16"].sym16_dense], [runs["fp32"].sym16_dense])
    print(f"HT-S 1080p (N={sym_e.size}) [symbols, indexes]: GPU~fp32 {a}, floor fp16~fp32 {f1}", flush=True)
    for i in range(2):
        assert 1 - a[i] <= 2 * (1 - f1[i]) + 5.0 / sym_e.size
    n_o = _coded_bytes(sdp, [runs["fp32"].sym16], runs["fp32"].z_int8, 32, ec)
    print(f"HT-S 1080p bytes gpu {len(bits)} fp32 {n_o}")
    assert abs(len(bits) - n_o) <= 0.005 * n_o


# ------------------------------------------------------------- 1920x1080
H1080, W1080 = 1080, 1920


def _check_floor_2(name, gpu, o32, o16):
    """The 1080p form of _check_floor: FP32 and fp16-storage oracle runs only
    (the float64-summation run would double the oracle's time at this size)."""
    a = _agree(gpu, o32)
    f1 = _agree(o16, o32)
    n = sum(x.size for x in gpu)
    print(f"{name} 1080p (N={n}) [symbols, indexes]: GPU~fp32 {a} | floor fp16~fp32 {f1}", flush=True)
    for i, what in enumerate(("symbols", "indexes")):
        assert 1 - a[i] <= 2 * (1 - f1[i]) + 5.0 / n, f"{name} {what}: GPU~fp32 {a[i]} below the floor {f1[i]}"


def _psnr_cropped(x_hat, src):
    return _psnr(x_hat[..., :src.shape[-2], :src.shape[-1]], src)


@pytest.fixture(scope="module")
def intra_1080p(gpu):
    """A 1920x1080 DMCI frame (padded to 1088 rows) on the GPU, encoder and
    decoder, and the oracle's FP32 and fp16-storage runs of it."""
    from dcvc_amd.inference_extensions_cuda import DMCIProxy
    pad_r, pad_b = synthetic.get_padding_size(H1080, W1080, 16)
    print(f"1080p intra: padding bottom {pad_b} right {pad_r}", flush=True)
    sdi = _state("dmci", 0)
    frames = synthetic.motion_sequence_yuv420(H1080, W1080, 9, seed=11)
    frame = synthetic.frame_to_input(frames[0])
    enc, dec = DMCIProxy(), DMCIProxy()
    enc.set_param(sdi, 0.0)
    dec.set_param(sdi, 0.0)
    bits, x_hat_e, ec = enc.compress(_gpu_nchw(frame, gpu), 32, pad_b, pad_r)
    x_hat_e = x_hat_e.clone()
    x_hat_d = dec.decompress(bits, 32, H1080, W1080, ec).clone()
    Hp, Wp = H1080 + pad_b, W1080 + pad_r
    nq = (Hp // 16) * (Wp // 16) * 64
    sym = [enc.debug_copy(k, nq * 2, np.int16).copy() for k in range(4)]
    torch.cuda.synchronize()
    sdo = _f32(sdi)
    xin = torch.from_numpy(frame)[None].half().float()
    runs = {}
    for m in ("fp32", "fp16"):
        with torch.inference_mode(), O.storage_mode(**MODES[m]):
            runs[m] = O.dmci_compress(sdo, xin, 32)
    return dict(sdi=sdi, frames=frames, frame=frame, bits=bits, ec=ec, x_hat_e=x_hat_e, x_hat_d=x_hat_d, sym=sym,
                runs=runs, pad_b=pad_b, pad_r=pad_r, Hp=Hp, Wp=Wp)


def test_dmci_1080p_within_the_fp16_storage_floor(intra_1080p):
    """The intra codec at 1920x1080: encoder and decoder agree bit for bit,
    the four quarter-step symbol arrays sit within the fp16 storage floor of
    the FP32 oracle, and PSNR / bytes stay within 0.02 dB / 0.5 %."""
    t = intra_1080p
    r = t["runs"]
    assert torch.equal(t["x_hat_e"], t["x_hat_d"])
    _check_floor_2("DMCI", t["sym"], r["fp32"].steps_dense, r["fp16"].steps_dense)
    src = torch.from_numpy(t["frame"])[None]
    pg, po = _psnr_cropped(t["x_hat_e"].float().cpu(), src), _psnr_cropped(r["fp32"].x_hat, src)
    n_o = _coded_bytes(t["sdi"], r["fp32"].steps_sym16[::-1], r["fp32"].z_int8, 32, t["ec"])
    print(f"DMCI 1080p PSNR gpu {pg:.4f} fp32 {po:.4f}; bytes gpu {len(t['bits'])} fp32 {n_o}")
    assert abs(pg - po) <= 0.02
    assert abs(len(t["bits"]) - n_o) <= 0.005 * n_o


@pytest.mark.parametrize("model", ["htl", "ld"])
def test_htl_ld_1080p_within_the_fp16_storage_floor(gpu, intra_1080p, model):
    """HT-L (one chunk of 4 real frames padded to 8) and LD (one frame) at
    1920x1080 from the intra frame above: encoder and decoder agree bit for
    bit, the symbols sit within the fp16 storage floor of the FP32 oracle,
    and PSNR / bytes stay within 0.02 dB / 0.5 % (HT-L bytes: or within twice
    the fp16 run's own distance, as at 128x192)."""
    from dcvc_amd.inference_extensions_cuda import DMCHTLProxy, DMCLDProxy
    t = intra_1080p
    pad_b, pad_r, Hp, Wp = t["pad_b"], t["pad_r"], t["Hp"], t["Wp"]
    x_hat_i = t["x_hat_e"]
    sdp = _state(model, 2 if model == "htl" else 3)
    enc, dec = (DMCHTLProxy(), DMCHTLProxy()) if model == "htl" else (DMCLDProxy(), DMCLDProxy())
    enc.set_param(sdp, 0.0)
    dec.set_param(sdp, 0.0)
    enc.add_ref_feature_from_frame(x_hat_i, True)
    dec.add_ref_feature_from_frame(t["x_hat_d"], False)
    P16 = (Hp // 16) * (Wp // 16)
    if model == "htl":
        x = synthetic.chunk_to_input(t["frames"][1:5], 8)
    else:
        x = synthetic.frame_to_input(t["frames"][1])
    print(f"{model} 1080p: input {x.shape}", flush=True)
    bits, ec = enc.compress(_gpu_nchw(x, gpu), 32, False, pad_b, pad_r)
    if model == "htl":
        sym = [enc.debug_copy(5 + k, P16 * 64 * 2, np.int16).copy() for k in range(4)]
        yhat_e = enc.debug_copy(0, P16 * 256 * 2, np.float16).copy()
        outs = [o.float().cpu() for o in dec.decompress(bits, 32, H1080, W1080, ec, False)]
        yhat_d = dec.debug_copy(0, P16 * 256 * 2, np.float16)
    else:
        sym = [enc.debug_copy(5, P16 * 128 * 2, np.int16).copy()]
        yhat_e = enc.debug_copy(0, P16 * 512 * 2, np.float16).reshape(P16, 512)[:, :128].copy()
        outs = [dec.decompress(bits, 32, H1080, W1080, ec, False).float().cpu()]
        yhat_d = dec.debug_copy(0, P16 * 512 * 2, np.float16).reshape(P16, 512)[:, :128]
    torch.cuda.synchronize()
    assert np.array_equal(yhat_e.view(np.uint16), yhat_d.view(np.uint16)), "decoder y_hat differs"
    sdo = _f32(sdp)
    ref = x_hat_i.float().cpu().contiguous()
    xin = torch.from_numpy(x)[None].half().float()
    runs = {}
    for m in ("fp32", "fp16"):
        with torch.inference_mode(), O.storage_mode(**MODES[m]):
            if model == "htl":
                st = O.HTSState()
                O.htl_add_ref_feature_from_frame(sdo, st, ref, True)
                runs[m] = O.htl_compress(sdo, st, xin, 32, False, recon=True)
            else:
                st = O.LDState()
                O.ld_add_ref_feature_from_frame(sdo, st, ref, True)
                runs[m] = O.ld_compress(sdo, st, xin, 32, False, recon=True)
    if model == "htl":
        _check_floor_2("HT-L", sym, runs["fp32"].steps_dense, runs["fp16"].steps_dense)
        nb = {m: _coded_bytes(sdp, runs[m].steps_sym16[::-1], runs[m].z_int8, 32, ec) for m in runs}
        srcs = [torch.from_numpy(x[3 * i:3 * i + 3])[None] for i in range(8)]
        dp = max(abs(_psnr_cropped(outs[i], srcs[i]) - _psnr_cropped(runs["fp32"].x_hat[i], srcs[i]))
                 for i in range(8))
        tol = max(0.005 * nb["fp32"], 2 * abs(nb["fp16"] - nb["fp32"]))
    else:
        _check_floor_2("LD", sym, [runs["fp32"].sym16_dense], [runs["fp16"].sym16_dense])
        nb = {m: _coded_bytes(sdp, [runs[m].sym16], runs[m].z_int8, 32, ec) for m in runs}
        src = torch.from_numpy(x)[None]
        dp = abs(_psnr_cropped(outs[0], src) - _psnr_cropped(runs["fp32"].x_hat, src))
        tol = 0.005 * nb["fp32"]
    print(f"{model} 1080p bytes gpu {len(bits)} {nb}; max PSNR delta {dp:.4f} dB", flush=True)
    assert dp <= 0.02
    assert abs(len(bits) - nb["fp32"]) <= tol


# ------------------------------------------------------------- 3840x2160
H2160, W2160 = 2160, 3840


@pytest.fixture(scope="module")
def intra_2160p(gpu):
    """A 3840x2160 DMCI frame: encoder and decoder reconstructions."""
    from dcvc_amd.inference_extensions_cuda import DMCIProxy
    pad_r, pad_b = synthetic.get_padding_size(H2160, W2160, 16)
    print(f"2160p intra: padding bottom {pad_b} right {pad_r}", flush=True)
    frames = synthetic.motion_sequence_yuv420(H2160, W2160, 5, seed=13)
    e, d = DMCIProxy(), DMCIProxy()
    e.set_param(_state("dmci", 0), 0.0)
    d.set_param(_state("dmci", 0), 0.0)
    bits, x_hat_e, ec = e.compress(_gpu_nchw(synthetic.frame_to_input(frames[0]), gpu), 32, pad_b, pad_r)
    x_hat_e = x_hat_e.clone()
    x_hat_d = d.decompress(bits, 32, H2160, W2160, ec).clone()
    torch.cuda.synchronize()
    return dict(frames=frames, x_hat_e=x_hat_e, x_hat_d=x_hat_d, pad_b=pad_b, pad_r=pad_r,
                Hp=H2160 + pad_b, Wp=W2160 + pad_r, nbytes=len(bits))


def test_dmci_2160p_encoder_decoder_bitexact(intra_2160p):
    t = intra_2160p
    assert t["nbytes"] > 0
    assert torch.equal(t["x_hat_e"], t["x_hat_d"]), "decoder reconstruction differs from the encoder's"


@pytest.mark.parametrize("model", ["hts", "htl", "ld"])
def test_p_codecs_2160p_encoder_decoder_bitexact(gpu, intra_2160p, model):
    """The largest picture size the reference tunes for (3840x2160): from
    the intra frame above, one chunk of 4 real frames padded to 8 (HT-S,
    HT-L) or one frame (LD); the decoder rebuilds the encoder's y_hat bit
    for bit and every reconstruction is finite and in range."""
    from dcvc_amd.inference_extensions_cuda import DMCHTLProxy, DMCHTSProxy, DMCLDProxy
    t = intra_2160p
    pad_b, pad_r, Hp, Wp = t["pad_b"], t["pad_r"], t["Hp"], t["Wp"]
    cls = {"hts": DMCHTSProxy, "htl": DMCHTLProxy, "ld": DMCLDProxy}[model]
    sdp = _state(model, {"hts": 1, "htl": 2, "ld": 3}[model])
    enc, dec = cls(), cls()
    enc.set_param(sdp, 0.0)
    dec.set_param(sdp, 0.0)
    enc.add_ref_feature_from_frame(t["x_hat_e"], True)
    dec.add_ref_feature_from_frame(t["x_hat_d"], False)
    P16 = (Hp // 16) * (Wp // 16)
    x = synthetic.frame_to_input(t["frames"][1]) if model == "ld" else synthetic.chunk_to_input(t["frames"][1:5], 8)
    print(f"{model} 2160p: input {x.shape}", flush=True)
    bits, ec = enc.compress(_gpu_nchw(x, gpu), 32, False, pad_b, pad_r)
    if model == "htl":
        yhat_e = enc.debug_copy(0, P16 * 256 * 2, np.float16).copy()
    else:
        cy = 256 if model == "hts" else 128
        yhat_e = enc.debug_copy(0, P16 * 512 * 2, np.float16).reshape(P16, 512)[:, :cy].copy()
    outs = dec.decompress(bits, 32, H2160, W2160, ec, False)
    outs = [outs] if model == "ld" else outs
    torch.cuda.synchronize()
    if model == "htl":
        yhat_d = dec.debug_copy(0, P16 * 256 * 2, np.float16)
    else:
        yhat_d = dec.debug_copy(0, P16 * 512 * 2, np.float16).reshape(P16, 512)[:, :cy]
    print(f"{model} 2160p: {len(bits)} bytes, ec {ec}", flush=True)
    assert len(bits) > 0
    assert np.array_equal(yhat_e.view(np.uint16), yhat_d.view(np.uint16)), "decoder y_hat differs"
    for f in outs:
        assert f.shape == (1, 3, Hp, Wp)
        assert torch.isfinite(f.float()).all() and float(f.abs().max()) <= 0.5


@pytest.mark.parametrize("H,W", [(240, 416), (480, 832), (480, 854), (720, 1280)])
def test_hts_reference_resolutions_encoder_decoder_bitexact(gpu, H, W):
    """HT-S at the other resolutions the reference tunes for (416x240 to
    1280x720; 854 columns need padding, 240 and 480 rows the 1/64-scale
    padding inside the hyperprior): an intra frame, then one chunk of 4 real
    frames padded to 8; the decoder rebuilds the encoder's y_hat bit for bit."""
    from dcvc_amd.inference_extensions_cuda import DMCHTSProxy, DMCIProxy
    pad_r, pad_b = synthetic.get_padding_size(H, W, 16)
    Hp, Wp = H + pad_b, W + pad_r
    print(f"HT-S {H}x{W} padded to {Hp}x{Wp}", flush=True)
    frames = synthetic.motion_sequence_yuv420(H, W, 5, seed=17)
    ie, idec = DMCIProxy(), DMCIProxy()
    ie.set_param(_state("dmci", 0), 0.0)
    idec.set_param(_state("dmci", 0), 0.0)
    bits_i, x_hat_e, ec_i = ie.compress(_gpu_nchw(synthetic.frame_to_input(frames[0]), gpu), 32, pad_b, pad_r)
    x_hat_e = x_hat_e.clone()
    x_hat_d = idec.decompress(bits_i, 32, H, W, ec_i).clone()
    assert torch.equal(x_hat_e, x_hat_d)
    sdp = _state("hts", 1)
    enc, dec = DMCHTSProxy(), DMCHTSProxy()
    enc.set_param(sdp, 0.0)
    dec.set_param(sdp, 0.0)
    enc.add_ref_feature_from_frame(x_hat_e, True)
    dec.add_ref_feature_from_frame(x_hat_d, False)
    P16 = (Hp // 16) * (Wp // 16)
    bits, ec = enc.compress(_gpu_nchw(synthetic.chunk_to_input(frames[1:5], 8), gpu), 32, False, pad_b, pad_r)
    yhat_e = enc.debug_copy(0, P16 * 512 * 2, np.float16).reshape(P16, 512)[:, :256].copy()
    n_e = enc.last_symbol_count()
    outs = dec.decompress(bits, 32, H, W, ec, False)
    torch.cuda.synchronize()
    yhat_d = dec.debug_copy(0, P16 * 512 * 2, np.float16).reshape(P16, 512)[:, :256]
    assert dec.last_symbol_count() == n_e
    assert np.array_equal(yhat_e.view(np.uint16), yhat_d.view(np.uint16))
    for f in outs:
        assert f.shape == (1, 3, Hp, Wp) and torch.isfinite(f.float()).all()
