"""Pins of the oracle functions that the reference modules' golden vectors do
not cover (CPU only).  Each check is chosen so that a plausible mistake in
the oracle (a wrong clamp bound, rounding mode, mask / step order, NHWC
flattening order, transposed operand) fails it: hand-derived special cases
from the reference kernels' documented semantics, invariants, and agreement
between independently written restatements (torch float64 vs numpy)."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from dcvc_amd import synthetic, weights
from oracle import dcvc_oracle as O


def test_round_z_special_cases():
    """round_z_kernel (stream.cu:862-884): C roundf (half away from zero), then
    clamp to [-64, 63]."""
    z = torch.tensor([63.4, 63.5, 64.7, -63.5, -64.5, -65.2, 0.5, -0.5, 2.5, -2.5, 0.49999997, 1e9, -1e9],
                     dtype=torch.float32)
    want = [63, 63, 63, -64, -64, -64, 1, -1, 3, -3, 0, 63, -64]
    assert O.round_z(z).tolist() == want


def test_quantize_masked_special_cases():
    """stream.cu:581-594: y_q = clamp(roundf((y - mean) * m) * [scale*m > thr],
    -128, 127); y_hat = y_q + mean * m."""
    y = torch.tensor([2.6, 2.6, 200.0, -300.0, 1.0, -2.6]).reshape(1, 6, 1, 1)
    means = torch.tensor([0.1, 0.1, 0.0, 0.0, 0.4, -0.1]).reshape(1, 6, 1, 1)
    scales = torch.tensor([1.0, 0.1, 1.0, 1.0, 1.0, 0.15]).reshape(1, 6, 1, 1)
    mask = torch.tensor([1, 1, 1, 1, 0, 1], dtype=torch.bool).reshape(1, 6, 1, 1)
    y_q, y_hat = O.quantize_masked(y, means, scales, mask, 0.15)
    assert y_q.flatten().tolist() == [3, 0, 127, -128, 0, 0]   # 2.5 -> 3; scale 0.1, 0.15 <= thr -> 0
    np.testing.assert_allclose(y_hat.flatten().numpy(), [3.1, 0.1, 127, -128, 0, -0.1], rtol=0, atol=1e-6)


def test_build_index_packing_and_nhwc_order():
    """(y_q << 8) + idx as int16 over NHWC-linear order (stream.cu:89-173)."""
    # NCHW [1, 2, 1, 2]: channel-major values a b | c d -> NHWC order a c b d
    y_q = torch.tensor([[-3.0, 0.0], [127.0, -128.0]]).reshape(1, 2, 1, 2)
    scales = torch.tensor([[1.0, 0.05], [16.0, 0.11]]).reshape(1, 2, 1, 2)
    sym, idx, cond = O.build_index(y_q, scales, 0.1)
    assert (sym.astype(np.int64) >> 8).tolist() == [-3, 127, 0, -128]
    assert idx.tolist() == [int(O.scale_to_index(np.float32(v))) for v in (1.0, 16.0, 0.05, 0.11)]
    # clamp ends: ln(16) = 2.772589 < log_scale_max 2.7726 (def_const.h:9), so
    # the top scale lands at 126.9997 -> 126; scales <= 0.11 -> 0
    assert idx[1] == 126 and idx[2] == 0 and idx[3] == 0
    assert (sym.astype(np.int64) & 0xFF).tolist() == idx.tolist()
    assert cond.tolist() == [True, True, False, True]


@pytest.mark.parametrize("n,p", [(0, 0.5), (1, 1.0), (1000, 0.0), (5000, 0.3), (4097, 1.0)])
def test_compact_recover_invariants(n, p):
    """Stable compaction and its inverse (stream.cu:175-383)."""
    r = np.random.Generator(np.random.PCG64(n))
    v = r.integers(-128, 128, n).astype(np.int16)
    cond = r.random(n) < p
    c = O.compact(v, cond)
    assert len(c) == int(cond.sum())
    assert c.tolist() == [int(x) for x, k in zip(v, cond) if k]          # order kept
    back = O.recover(c.astype(np.int8), cond, n)
    assert back.tolist() == [int(x) if k else 0 for x, k in zip(v.astype(np.int8), cond)]


def test_mask_steps_partition_each_channel_group():
    m = torch.stack([mk[0] for mk in O.mask_4x(16, 5, 6)])
    assert (m.sum(0) == 1).all()                                         # each element active once
    # step k applies position g ^ {0,3,2,1}[k] to group g (dmci_proxy.cpp:678-699)
    for k, x in enumerate((0, 3, 2, 1)):
        for g in range(4):
            pos = g ^ x
            assert bool(m[k, g * 4, pos // 2, pos % 2])


def test_fp32_step_restatement_agrees_with_torch_steps():
    """The float32 per-step restatement (NHWC, active groups) and the torch
    NCHW masked steps give the same symbols, indexes and y_hat."""
    r = np.random.Generator(np.random.PCG64(5))
    C, H, W = 8, 4, 6
    P = H * W
    y = r.standard_normal((P, C)).astype(np.float16) * np.float16(3)
    sc = np.exp(r.uniform(np.log(0.05), np.log(20), (P, C))).astype(np.float16)
    mn = r.standard_normal((P, C)).astype(np.float16)
    qd = np.ones((P, C), np.float16)
    thr = 0.15
    to_nchw = lambda a: torch.from_numpy(a.astype(np.float32).reshape(1, H, W, C)).permute(0, 3, 1, 2)
    masks = O.mask_4x(C, H, W)
    yh = np.zeros((P, C), np.float16)
    y_hat_t = torch.zeros(1, C, H, W)
    for step in range(4):
        yh, sym, cond, act = O.hts_quant_step_f32(y, sc, mn, qd, yh, step, thr, C, H, W)
        y_q, yh_t = O.quantize_masked(to_nchw(y), to_nchw(mn), to_nchw(sc), masks[step], thr)
        y_hat_t = y_hat_t + yh_t
        sym_t, idx_t, cond_t = O.build_index(y_q, to_nchw(sc), thr)
        a = act.reshape(-1)
        assert np.array_equal(sym.reshape(-1)[a], sym_t[a])
        assert np.array_equal(cond.reshape(-1).astype(bool), cond_t)
    ref = y_hat_t.permute(0, 2, 3, 1).reshape(P, C).numpy()
    np.testing.assert_allclose(yh.astype(np.float32), ref, rtol=2 ** -10, atol=2 ** -14)


def test_gemm_reference_against_torch():
    """ref_gemm_epilogue (numpy) against torch float64 conv1x1 / sigmoid forms."""
    r = np.random.Generator(np.random.PCG64(9))
    M, K, N = 37, 24, 16
    a, w = r.standard_normal((M, K)), r.standard_normal((4 * N, K)) / 5
    b, s1, s2, q = r.standard_normal(4 * N), r.standard_normal((M, N)), r.standard_normal((M, N)), r.random(N) + .5
    x = torch.from_numpy(a).T.reshape(1, K, M, 1)
    conv = lambda wt, bt: F.conv2d(x, torch.from_numpy(wt)[:, :, None, None], torch.from_numpy(bt))[0, :, :, 0].T
    wsilu = lambda t: t * torch.sigmoid(4 * t)
    y4 = wsilu(conv(w, b))                                                # [M, 4N], channel 4n + j
    quad = sum(y4[:, j::4] for j in range(4)).numpy()
    np.testing.assert_allclose(O.ref_gemm_epilogue(a, w, b, "quad"), quad, rtol=1e-12, atol=1e-12)
    w1, b1 = w[:N], b[:N]
    y = conv(w1, b1).numpy()
    np.testing.assert_allclose(O.ref_gemm_epilogue(a, w1, b1, "bias"), y, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.ref_gemm_epilogue(a, w1, b1, "wsilu"), wsilu(torch.from_numpy(y)).numpy(),
                               rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.ref_gemm_epilogue(a, w1, b1, "shortcut2", s1, s2), y + s1 + s2, rtol=1e-12)
    np.testing.assert_allclose(O.ref_gemm_epilogue(a, w1, b1, "shortcut_quant", s1, None, q), (y + s1) * q,
                               rtol=1e-12)


def test_dw3x3_reference_against_torch():
    r = np.random.Generator(np.random.PCG64(4))
    H, W, C = 7, 9, 5
    x, w = r.standard_normal((H, W, C)), r.standard_normal((C, 3, 3))
    t = F.conv2d(torch.from_numpy(x).permute(2, 0, 1)[None], torch.from_numpy(w)[:, None], padding=1, groups=C)
    np.testing.assert_allclose(O.ref_dw3x3(x, w), t[0].permute(1, 2, 0).numpy(), rtol=1e-12, atol=1e-12)


def test_oracle_hts_encode_decode_round_trip():
    """The oracle decoder, fed the encoder's compacted symbols, rebuilds the
    encoder's y_hat and reconstruction exactly, and asks the entropy decoder
    for exactly the encoder's CDF indexes."""
    sd = {k: v.float() for k, v in weights.to_torch(weights.round_to_half(weights.init_state_dict("hts", 1))).items()}
    H = W = 64
    frames = synthetic.motion_sequence_yuv420(H, W, 9, seed=2)
    ref = torch.from_numpy(synthetic.frame_to_input(frames[0]))[None]
    x = torch.from_numpy(synthetic.chunk_to_input(frames[1:9]))[None]
    st_e, st_d = O.HTSState(), O.HTSState()
    with torch.inference_mode():
        O.hts_add_ref_feature_from_frame(sd, st_e, ref, True)
        O.hts_add_ref_feature_from_frame(sd, st_d, ref, False)
        r = O.hts_compress(sd, st_e, x, 40, False, skip_thres=0.15, recon=True)
        seen = {}

        def decode_y(idx):
            seen["idx"] = idx
            return (r.sym16.astype(np.int64) >> 8).astype(np.int8)

        x_hat, y_hat = O.hts_decompress(sd, st_d, r.z_int8, decode_y, 40, H, W, False, skip_thres=0.15)
    assert np.array_equal(seen["idx"], (r.sym16.astype(np.int64) & 0xFF).astype(np.uint8))
    assert torch.equal(y_hat, r.y_hat)
    for a, b in zip(x_hat, r.x_hat):
        assert torch.equal(a, b)
    assert torch.equal(st_d.feature_p, st_e.feature_p)


def test_oracle_htl_encode_decode_round_trip():
    """HT-L: four entropy round trips.  The oracle decoder, fed each step's
    compacted symbols, asks for exactly the encoder's CDF indexes in every
    step and rebuilds y_hat, the reconstruction and the context exactly; the
    four steps partition y (each element coded once)."""
    sd = {k: v.float() for k, v in weights.to_torch(weights.round_to_half(weights.init_state_dict("htl", 2))).items()}
    H = W = 64
    frames = synthetic.motion_sequence_yuv420(H, W, 9, seed=3)
    ref = torch.from_numpy(synthetic.frame_to_input(frames[0]))[None]
    x = torch.from_numpy(synthetic.chunk_to_input(frames[1:9]))[None]
    st_e, st_d = O.HTSState(), O.HTSState()
    with torch.inference_mode():
        O.htl_add_ref_feature_from_frame(sd, st_e, ref, True)
        O.htl_add_ref_feature_from_frame(sd, st_d, ref, False)
        r = O.htl_compress(sd, st_e, x, 21, False, skip_thres=0.15, recon=True)
        seen = []

        def decode_y(step, idx):
            seen.append(idx)
            return (r.steps_sym16[step].astype(np.int64) >> 8).astype(np.int8)

        x_hat, y_hat = O.htl_decompress(sd, st_d, r.z_int8, decode_y, 21, H, W, False, skip_thres=0.15)
    assert len(seen) == 4
    for step in range(4):
        assert np.array_equal(seen[step], (r.steps_sym16[step].astype(np.int64) & 0xFF).astype(np.uint8))
        assert r.steps_dense[step].size == r.y.numel() // 4
    assert torch.equal(y_hat, r.y_hat)
    for a, b in zip(x_hat, r.x_hat):
        assert torch.equal(a, b)
    assert torch.equal(st_d.feature_p, st_e.feature_p)
    # y_hat / q_dec - y is the quantisation residual: |.| <= 1/2 where the
    # element was coded (scale > skip threshold) and not clamped
    qd = torch.clamp_min(r.common_params[:, :256], 0.5)
    res = (r.y_hat / qd - r.y).abs()
    assert float(res.median()) < 0.5


def test_htl_fp32_step_restatement_agrees_with_torch_steps():
    """HT-L form: the float32 quarter-step restatement (NHWC, y / max(q, .5),
    per-step scales) and the torch NCHW masked steps with single_part_for_
    writing_4x give the same quarter symbols and conditions, and the same
    final y_hat = (y_hat_3 + y_hat_so_far) * max(q, .5)."""
    r = np.random.Generator(np.random.PCG64(8))
    C, H, W = 8, 4, 6
    P = H * W
    y = (r.standard_normal((P, C)) * 3).astype(np.float16)
    qd = (r.random((P, C)) * 1.5).astype(np.float16)           # some below the 0.5 clamp
    thr = 0.15
    to_nchw = lambda a: torch.from_numpy(a.astype(np.float32).reshape(1, H, W, C)).permute(0, 3, 1, 2)
    masks = O.mask_4x(C, H, W)
    q32 = np.maximum(qd.astype(np.float32), np.float32(0.5))
    ys_t = to_nchw((y.astype(np.float32) * (np.float32(1) / q32)).astype(np.float32))
    ysf = np.full((P, C), 7.0, np.float16)
    so_far = torch.zeros(1, C, H, W)
    for step in range(4):
        sc = np.exp(r.uniform(np.log(0.05), np.log(20), (P, C))).astype(np.float16)
        mn = r.standard_normal((P, C)).astype(np.float16)
        ysf, yout, sym, cond = O.htl_quant_step_f32(y, sc, mn, qd, ysf, step, thr, C, H, W)
        y_q, yh = O.quantize_masked(ys_t, to_nchw(mn), to_nchw(sc), masks[step], thr)
        sym_t, _, cond_t = O.build_index(O._quarter(y_q, masks[step]), O._quarter(to_nchw(sc), masks[step]), thr)
        assert np.array_equal(sym, sym_t), f"step {step}"
        assert np.array_equal(cond.astype(bool), cond_t), f"step {step}"
        if step < 3:
            so_far = so_far + yh
        else:
            ref = ((yh + so_far) * to_nchw(q32)).permute(0, 2, 3, 1).reshape(P, C).numpy()
            np.testing.assert_allclose(yout.astype(np.float32), ref, rtol=2 ** -10, atol=2 ** -14)


def test_oracle_ld_encode_decode_round_trip():
    """Low delay: one entropy round trip over all of y.  The oracle decoder,
    fed the encoder's compacted symbols, asks for exactly the encoder's
    indexes and rebuilds y_hat, the reconstruction and the context (incl.
    the temporal params) exactly, over two frames with a reset between."""
    sd = {k: v.float() for k, v in weights.to_torch(weights.round_to_half(weights.init_state_dict("ld", 3))).items()}
    H = W = 64
    frames = synthetic.motion_sequence_yuv420(H, W, 3, seed=4)
    ref = torch.from_numpy(synthetic.frame_to_input(frames[0]))[None]
    st_e, st_d = O.LDState(), O.LDState()
    with torch.inference_mode():
        O.ld_add_ref_feature_from_frame(sd, st_e, ref, True)
        O.ld_add_ref_feature_from_frame(sd, st_d, ref, False)
        for t, reset in ((1, True), (2, False)):
            x = torch.from_numpy(synthetic.frame_to_input(frames[t]))[None]
            r = O.ld_compress(sd, st_e, x, 50, reset, skip_thres=0.15, recon=True)
            seen = {}

            def decode_y(idx):
                seen["idx"] = idx
                return (r.sym16.astype(np.int64) >> 8).astype(np.int8)

            x_hat, y_hat = O.ld_decompress(sd, st_d, r.z_int8, decode_y, 50, H, W, reset, skip_thres=0.15)
            assert np.array_equal(seen["idx"], (r.sym16.astype(np.int64) & 0xFF).astype(np.uint8))
            assert torch.equal(y_hat, r.y_hat)
            assert torch.equal(x_hat, r.x_hat)
            assert torch.equal(st_d.feature_p, st_e.feature_p)
    # the decoder's next-frame context equals the encoder's (applied lazily on the decoder side)
    st_d.memory = O.ld_feature_adaptor_m(sd, st_d.memory, st_d.feature_p)
    assert torch.equal(st_d.memory, st_e.memory)
    assert torch.equal(O.ld_temporal_prior(sd, st_d.memory), st_e.temporal)


def test_ld_fp32_step_restatement_agrees_with_torch_steps():
    """LD form: the float32 two-half restatement and the torch NCHW masked
    halves give the same symbols, conditions and y_hat * max(q, .5)."""
    r = np.random.Generator(np.random.PCG64(12))
    C, H, W = 8, 5, 6
    P = H * W
    y = (r.standard_normal((P, C)) * 3).astype(np.float16)
    qd = (r.random((P, C)) * 1.5).astype(np.float16)
    sc = np.exp(r.uniform(np.log(0.05), np.log(20), (P, C))).astype(np.float16)
    thr = 0.15
    to_nchw = lambda a: torch.from_numpy(a.astype(np.float32).reshape(1, H, W, C)).permute(0, 3, 1, 2)
    masks = O.mask_2x(C, H, W)
    q32 = np.maximum(qd.astype(np.float32), np.float32(0.5))
    ys_t = to_nchw((y.astype(np.float32) * (np.float32(1) / q32)).astype(np.float32))
    yh = np.zeros((P, C), np.float16)
    sym_all = np.zeros((P, C), np.int16)
    yq_t = torch.zeros(1, C, H, W)
    yh_t = torch.zeros(1, C, H, W)
    for step in range(2):
        mn = r.standard_normal((P, C)).astype(np.float16)
        yh, sym, cond, act = O.ld_quant_step_f32(y, sc, mn, qd, yh, step, thr, C, H, W)
        sym_all[act] = sym[act]
        q_t, h_t = O.quantize_masked(ys_t, to_nchw(mn), to_nchw(sc), masks[step], thr)
        yq_t, yh_t = yq_t + q_t, yh_t + h_t
    sym_t, _, cond_t = O.build_index(yq_t, to_nchw(sc), thr)
    assert np.array_equal(sym_all.reshape(-1), sym_t)
    assert np.array_equal(cond.reshape(-1).astype(bool), cond_t)
    ref = (yh_t * to_nchw(q32)).permute(0, 2, 3, 1).reshape(P, C).numpy()
    np.testing.assert_allclose(yh.astype(np.float32), ref, rtol=2 ** -10, atol=2 ** -14)


def _intra_round_trip(sd, frame, qp, thr):
    r = O.dmci_compress(sd, frame, qp, skip_thres=thr)

    def decode_y(step, idx):
        assert np.array_equal(idx, (r.steps_sym16[step].astype(np.int64) & 0xFF).astype(np.uint8)), step
        return (r.steps_sym16[step].astype(np.int64) >> 8).astype(np.int8)

    H, W = frame.shape[2], frame.shape[3]
    x_hat, y_hat, seen = O.dmci_decompress(sd, r.z_int8, decode_y, qp, H, W, skip_thres=thr)
    return r, x_hat, y_hat, seen


@pytest.mark.parametrize("H,W,thr", [(64, 64, 0.0), (80, 112, 0.15)])
def test_oracle_dmci_encode_decode_round_trip(H, W, thr):
    """DMCI: four entropy round trips over quarter tensors.  The oracle
    decoder asks for exactly the encoder's CDF indexes in each step (steps
    0..3 after z, the reverse of the encoder's write order) and rebuilds
    y_hat and the reconstruction exactly; the four steps' streams partition
    y (C/4 channels per pixel each), in FP32 and in fp16-storage mode."""
    sd = {k: v.float() for k, v in weights.to_torch(weights.round_to_half(weights.init_state_dict("dmci", 3))).items()}
    frame = torch.from_numpy(synthetic.frame_to_input(synthetic.texture_yuv420(H, W, 4)))[None]
    frame = F.pad(frame, (0, (16 - W % 16) % 16, 0, (16 - H % 16) % 16), mode="replicate")
    for fp16 in (False, True):
        with torch.inference_mode(), O.storage_mode(fp16=fp16):
            r, x_hat, y_hat, seen = _intra_round_trip(sd, frame, 27, thr)
        assert len(seen) == 4
        assert sum(len(s) for s in seen) == sum(len(s) for s in r.steps_sym16)
        for step in range(4):
            assert r.steps_dense[step].size == r.y.numel() // 4
        assert torch.equal(y_hat, r.y_hat)
        assert torch.equal(x_hat, r.x_hat)
        if fp16:
            assert torch.equal(y_hat, y_hat.half().float()) and torch.equal(x_hat, x_hat.half().float())


def test_fp16_storage_mode_is_the_fp32_restatement_without_rounding(monkeypatch):
    """storage_mode(fp16=True) with st16 made the identity reproduces the
    FP32 oracle bit for bit: the mode changes only where values are stored
    (and the depthwise-bias fold, an exact identity in real arithmetic,
    which moves the fp32 result by rounding only)."""
    sd = {k: v.float() for k, v in weights.to_torch(weights.round_to_half(weights.init_state_dict("hts", 1))).items()}
    x = torch.from_numpy(np.random.Generator(np.random.PCG64(5)).standard_normal((1, 512, 6, 10)).astype(np.float32))
    with torch.inference_mode():
        ref = O.dcb(sd, "feature_extractor.conv.0.", x)
        monkeypatch.setattr(O, "st16", lambda t: t)
        with O.storage_mode(fp16=True):
            got = O.dcb(sd, "feature_extractor.conv.0.", x)
        assert torch.allclose(got, ref, rtol=1e-5, atol=1e-5)
        monkeypatch.undo()
        with O.storage_mode(fp16=True):
            got16 = O.dcb(sd, "feature_extractor.conv.0.", x)
    # every stored value is an fp16 value, and the rounding moves the block
    # by about the fp16 resolution, not more
    assert torch.equal(got16, got16.half().float())
    rel = float((got16 - ref).abs().max() / ref.abs().max())
    assert 1e-5 < rel < 2e-2, rel
