"""GPU parity of every kernel-level entry point (include/dcvc_ops.h) against
the CPU oracle.  Floating-point GEMM/conv outputs are compared to float64
within the fp16 output rounding plus fp32 accumulation error; every integer,
index and layout kernel must match bit-for-bit."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import dcvc_oracle as O

pytestmark = pytest.mark.gpu


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def f16(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a).astype(np.float16)).to(dev)


def ulp16(x):
    """fp16 unit in the last place of |x| (subnormal floor 2^-24)."""
    a = np.abs(np.asarray(x, np.float64))
    e = np.floor(np.log2(np.maximum(a, 2.0 ** -14)))
    return 2.0 ** (e - 10)


def assert_fp16_close(out, ref, absacc, what=""):
    """|out - ref| <= 1 fp16 ulp(ref) + 2e-6 * sum|a||w| (fp32 accumulation)."""
    out = out.detach().float().cpu().numpy().astype(np.float64)
    err = np.abs(out - ref)
    bound = ulp16(ref) + 2e-6 * absacc + 1e-7
    bad = err > bound
    assert not bad.any(), (f"{what}: {bad.sum()} of {bad.size} outside bound; worst "
                           f"err {err[bad].max():.3g} at ref {ref[bad][np.argmax(err[bad])]:.4g}")


# ------------------------------------------------------------------ GEMMs

GEMM_SHAPES = [(1000, 256, 192), (4096, 512, 256), (257, 128, 512), (8160, 768, 768), (130, 2048, 64)]


@pytest.mark.parametrize("M,K,N", GEMM_SHAPES)
@pytest.mark.parametrize("epi", ["bias", "wsilu", "shortcut", "shortcut2", "shortcut_quant", "quant"])
def test_conv1x1_epilogues(gpu, M, K, N, epi):
    from dcvc_amd import kernels
    r = rng(M + K + N)
    a = r.standard_normal((M, K)).astype(np.float16)
    w = (r.standard_normal((N, K)) / np.sqrt(K)).astype(np.float16)
    b = (r.standard_normal(N) * 0.1).astype(np.float32)
    s1 = r.standard_normal((M, N)).astype(np.float16)
    s2 = r.standard_normal((M, N)).astype(np.float16)
    qtab = (1 + 0.5 * r.random((4, N))).astype(np.float16)
    qp = torch.tensor([2], dtype=torch.int32, device=gpu)
    ref = O.ref_gemm_epilogue(a, w, b, epi, s1, s2, qtab[2])
    absacc = np.abs(a.astype(np.float64)) @ np.abs(w.astype(np.float64)).T
    if epi in ("shortcut_quant", "quant"):
        absacc = absacc * np.abs(qtab[2].astype(np.float64))[None, :]
    out = kernels.conv1x1(f16(a, gpu), f16(w, gpu), torch.from_numpy(b).to(gpu), epi=epi,
                          s1=f16(s1, gpu), s2=f16(s2, gpu), q=f16(qtab, gpu), qp=qp)
    torch.cuda.synchronize()
    assert_fp16_close(out, ref, absacc, f"{epi} {M}x{K}x{N}")


def test_conv1x1_strided_views_and_inplace(gpu):
    """Channel slices of a concat buffer as A / D, and D aliasing the shortcut."""
    from dcvc_amd import kernels
    r = rng(7)
    M, K, N = 1500, 256, 256
    cat = f16(r.standard_normal((M, 3 * K)), gpu)      # A = columns [K, 2K)
    a = cat[:, K:2 * K]
    w = f16(r.standard_normal((N, K)) / 16, gpu)
    b = torch.from_numpy((r.standard_normal(N) * 0.1).astype(np.float32)).to(gpu)
    sbuf = f16(r.standard_normal((M, 2 * N)), gpu)
    left = sbuf[:, :N].clone()
    s = sbuf[:, N:]
    s_ref = s.float().cpu().numpy().astype(np.float64)
    ref = O.ref_gemm_epilogue(a.cpu().numpy(), w.cpu().numpy(), b.cpu().numpy(), "shortcut", s_ref)
    kernels.conv1x1(a, w, b, epi="shortcut", s1=s, out=s)   # in place over the shortcut
    torch.cuda.synchronize()
    absacc = np.abs(a.cpu().numpy().astype(np.float64)) @ np.abs(w.cpu().numpy().astype(np.float64)).T
    assert_fp16_close(s, ref, absacc, "in-place shortcut")
    assert torch.equal(sbuf[:, :N], left)  # the other half of the buffer is untouched


@pytest.mark.parametrize("M,K,N", [(1000, 256, 128), (4096, 512, 256), (200, 768, 768), (520, 384, 192)])
def test_quad_gemm(gpu, M, K, N):
    from dcvc_amd import kernels
    r = rng(M * 3 + N)
    a = r.standard_normal((M, K)).astype(np.float16)
    w = (r.standard_normal((4 * N, K)) / np.sqrt(K)).astype(np.float16)
    b = (r.standard_normal(4 * N) * 0.1).astype(np.float32)
    ref = O.ref_gemm_epilogue(a, w, b, "quad")
    acc = np.abs(a.astype(np.float64)) @ np.abs(w.astype(np.float64)).T
    absacc = acc[:, 0::4] + acc[:, 1::4] + acc[:, 2::4] + acc[:, 3::4]
    out = kernels.conv1x1(f16(a, gpu), f16(w, gpu), torch.from_numpy(b).to(gpu), epi="quad")
    torch.cuda.synchronize()
    # four wsilu terms are summed in fp32 before the single fp16 rounding
    assert_fp16_close(out, ref, absacc + 4 * np.abs(ref) * 1e-6, f"quad {M}x{K}x{N}")


def test_quad_equals_wide_gemm_plus_chunk_add(gpu):
    """Closed form of layers.py:114-125: quad == wsilu(4N-wide GEMM) summed over x[:, j::4]."""
    from dcvc_amd import kernels
    r = rng(11)
    M, K, N = 777, 256, 64
    a = f16(r.standard_normal((M, K)), gpu)
    w = f16(r.standard_normal((4 * N, K)) / 16, gpu)
    b = torch.from_numpy((r.standard_normal(4 * N) * 0.1).astype(np.float32)).to(gpu)
    wide = kernels.conv1x1(a, w, b, epi="wsilu").float()
    want = wide[:, 0::4] + wide[:, 1::4] + wide[:, 2::4] + wide[:, 3::4]
    got = kernels.conv1x1(a, w, b, epi="quad").float()
    torch.cuda.synchronize()
    # wide path rounds each term to fp16 first: allow 4 half-ulps of the terms
    tol = 4 * torch.from_numpy(ulp16(wide.cpu().numpy()).reshape(M, 4 * N)).to(gpu).float().view(M, N, 4).max(-1).values
    assert ((got - want).abs() <= tol + 1e-3).all()


def test_gemm_batch_invariance(gpu):
    """Rows are bit-identical whether computed alone or inside a larger M (any tile placement)."""
    from dcvc_amd import kernels
    r = rng(3)
    K, N = 512, 256
    big = f16(r.standard_normal((5000, K)), gpu)
    w = f16(r.standard_normal((N, K)) / 20, gpu)
    b = torch.zeros(N, device=gpu)
    full = kernels.conv1x1(big, w, b, epi="wsilu")
    part = kernels.conv1x1(big[1234:1234 + 77].contiguous(), w, b, epi="wsilu")
    small = kernels.conv1x1(big[:100], w, b, epi="wsilu")  # different tile config (M <= 2048)
    torch.cuda.synchronize()
    assert torch.equal(full[1234:1234 + 77], part)
    assert torch.equal(full[:100], small)


@pytest.mark.parametrize("epi", ["bias", "wsilu", "shortcut2", "shortcut_quant", "shortcut_then_quant", "quad", "s2k3",
                                 "s1k3", "subpel", "subpel3"])
def test_gemm_tile_configs_bit_identical(gpu, epi):
    """Every tile configuration gives the same bits (same per-element accumulation
    order), including ragged M / N tails, implicit convs and the sub-pixel store.
    Configurations 8-11 and 14 (64-deep K tiles, 128-byte LDS rows) serve
    plain 1x1 products; the other launches fall back to the automatic choice
    (14: to 12).  12 and 13 are the small-grid tiles (8 waves on 128 x 128,
    and 64-row tiles)."""
    from dcvc_amd import _lib, kernels
    lib = _lib.load()
    r = rng(21)
    M, K, N = 2999, 512, 200
    a = f16(r.standard_normal((M, K)), gpu)
    s1 = f16(r.standard_normal((M, N)), gpu)
    s2 = f16(r.standard_normal((M, N)), gpu)
    qtab = f16(1 + 0.5 * r.random((3, N)), gpu)
    qp = torch.tensor([1], dtype=torch.int32, device=gpu)
    H, W, cin = 37, 53, 128

    def run():
        if epi == "quad":
            w = f16(r0.standard_normal((4 * N, K)) / 20, gpu)
            return kernels.conv1x1(a, w, torch.zeros(4 * N, device=gpu), epi="quad")
        if epi in ("s2k3", "s1k3"):
            x = f16(r0.standard_normal((H * W, cin)), gpu)
            w = f16(r0.standard_normal((N, 9 * cin)) / 30, gpu)
            return kernels.conv_bias(x, 1, H, W, w, torch.ones(N, device=gpu), 3, 2 if epi == "s2k3" else 1)
        if epi == "subpel":
            x = f16(r0.standard_normal((H * W, cin)), gpu)
            w = f16(r0.standard_normal((4 * 64, cin)) / 10, gpu)
            return kernels.transposed_conv(x, 1, H, W, w, 64, torch.ones(4 * 64, device=gpu))
        if epi == "subpel3":
            x = f16(r0.standard_normal((H * W, cin)), gpu)
            w = f16(r0.standard_normal((4 * 64, 9 * cin)) / 30, gpu)
            return kernels.subpel_conv3x3(x, 1, H, W, w, 64, torch.ones(4 * 64, device=gpu))
        w = f16(r0.standard_normal((N, K)) / 20, gpu)
        return kernels.conv1x1(a, w, torch.full((N,), 0.1, device=gpu), epi=epi, s1=s1, s2=s2, q=qtab, qp=qp)

    outs = []
    try:
        for policy in range(1, 15):
            assert lib.dcvc_gemm_set_tile_policy(policy) == 0
            r0 = rng(5)
            outs.append(run())
    finally:
        lib.dcvc_gemm_set_tile_policy(0)
    torch.cuda.synchronize()
    assert torch.isfinite(outs[0].float()).all()
    for o in outs[1:]:
        assert torch.equal(o.view(torch.int16), outs[0].view(torch.int16))


@pytest.mark.parametrize("epi", ["wsilu", "shortcut", "quad"])
def test_bf16_tile_configs_bit_identical(gpu, epi):
    """The BF16 products give the same bits in every tile configuration,
    the 64-deep K tiles included."""
    from dcvc_amd import _lib, kernels
    lib = _lib.load()
    r = rng(23)
    M, K, N = 3001, 768, 136
    rows = 4 * N if epi == "quad" else N
    a = f16(r.standard_normal((M, K)), gpu)
    w = f16(r.standard_normal((rows, K)) / 25, gpu)
    s1 = f16(r.standard_normal((M, N)), gpu)
    wq, scale = kernels.weights_to_lowp(w, "bf16")
    bias = torch.full((rows,), 0.05, device=gpu)
    outs = []
    try:
        for policy in range(1, 15):
            assert lib.dcvc_gemm_set_tile_policy(policy) == 0
            outs.append(kernels.gemm_lowp(a, wq, scale, "bf16", bias, epi=epi, s1=s1 if epi == "shortcut" else None))
    finally:
        lib.dcvc_gemm_set_tile_policy(0)
    torch.cuda.synchronize()
    assert torch.isfinite(outs[0].float()).all()
    for o in outs[1:]:
        assert torch.equal(o.view(torch.int16), outs[0].view(torch.int16))


# ------------------------------------------------------------------ convs

@pytest.mark.parametrize("kernel,stride,H,W,cin,n,n_img", [
    (2, 2, 16, 24, 128, 256, 1), (2, 2, 34, 60, 256, 128, 2), (3, 2, 17, 30, 128, 64, 1),
    (3, 2, 34, 60, 384, 256, 1), (3, 1, 12, 20, 128, 192, 2)])
def test_conv_bias(gpu, kernel, stride, H, W, cin, n, n_img):
    from dcvc_amd import kernels
    r = rng(H * W + cin)
    x = r.standard_normal((n_img, H, W, cin)).astype(np.float16)
    wt = (r.standard_normal((n, cin, kernel, kernel)) / np.sqrt(cin * kernel * kernel)).astype(np.float16)
    b = (r.standard_normal(n) * 0.1).astype(np.float32)
    pad = 0 if kernel == 2 else 1
    xt = torch.from_numpy(x.astype(np.float64)).permute(0, 3, 1, 2)
    ref = F.conv2d(xt, torch.from_numpy(wt.astype(np.float64)), torch.from_numpy(b.astype(np.float64)),
                   stride=stride, padding=pad).permute(0, 2, 3, 1).reshape(-1, n).numpy()
    acc = F.conv2d(xt.abs(), torch.from_numpy(np.abs(wt.astype(np.float64))), None, stride=stride,
                   padding=pad).permute(0, 2, 3, 1).reshape(-1, n).numpy()
    w_ohwi = np.ascontiguousarray(wt.transpose(0, 2, 3, 1).reshape(n, -1))
    out = kernels.conv_bias(f16(x.reshape(-1, cin), gpu), n_img, H, W, f16(w_ohwi, gpu),
                            torch.from_numpy(b).to(gpu), kernel, stride)
    torch.cuda.synchronize()
    assert_fp16_close(out, ref, acc, f"conv k{kernel}s{stride}")


def test_rbs2_weight_transform_matches_unshuffle_1x1(gpu):
    """2x2 s2 conv with the pixel_shuffle'd 1x1 weight == 1x1 conv on pixel_unshuffle(x, 2)
    (layers.py:185-186 vs layers_proxy.cpp:264)."""
    from dcvc_amd import kernels
    r = rng(5)
    H, W, cin, n = 20, 28, 128, 128
    x = r.standard_normal((1, H, W, cin)).astype(np.float16)
    w1 = (r.standard_normal((n, 4 * cin)) / 20).astype(np.float16)   # down.weight [n, 4cin]
    b = np.zeros(n, np.float32)
    xt = torch.from_numpy(x.astype(np.float64)).permute(0, 3, 1, 2)
    ref = F.conv2d(F.pixel_unshuffle(xt, 2), torch.from_numpy(w1.astype(np.float64))[:, :, None, None])
    ref = ref.permute(0, 2, 3, 1).reshape(-1, n).numpy()
    # tap-major weight: W'[o, (dy*2+dx)*cin + c] = W[o, c*4 + dy*2 + dx]
    w_tap = np.ascontiguousarray(w1.reshape(n, cin, 4).transpose(0, 2, 1).reshape(n, 4 * cin))
    out = kernels.conv_bias(f16(x.reshape(-1, cin), gpu), 1, H, W, f16(w_tap, gpu),
                            torch.from_numpy(b).to(gpu), 2, 2)
    torch.cuda.synchronize()
    acc = np.abs(F.pixel_unshuffle(xt.abs(), 2).permute(0, 2, 3, 1).reshape(-1, 4 * cin).numpy()) @ np.abs(
        w1.astype(np.float64)).T
    assert_fp16_close(out, ref, acc, "rbs2")


@pytest.mark.parametrize("H,W,cin,cout,n_img", [(17, 30, 128, 256, 1), (8, 12, 256, 512, 2), (34, 60, 256, 128, 1)])
def test_transposed_conv(gpu, H, W, cin, cout, n_img):
    """SubpelConv2x(k=1): pixel_shuffle(conv1x1(x), 2) (layers.py:92-103)."""
    from dcvc_amd import kernels
    r = rng(H + cin)
    x = r.standard_normal((n_img, H, W, cin)).astype(np.float16)
    w = (r.standard_normal((4 * cout, cin)) / np.sqrt(cin)).astype(np.float16)  # up.conv.0.weight
    xt = torch.from_numpy(x.astype(np.float64)).permute(0, 3, 1, 2)
    ref = F.pixel_shuffle(F.conv2d(xt, torch.from_numpy(w.astype(np.float64))[:, :, None, None]), 2)
    ref = ref.permute(0, 2, 3, 1).reshape(-1, cout).numpy()
    acc = F.pixel_shuffle(F.conv2d(xt.abs(), torch.from_numpy(np.abs(w.astype(np.float64)))[:, :, None, None]), 2)
    acc = acc.permute(0, 2, 3, 1).reshape(-1, cout).numpy()
    w_sub = np.ascontiguousarray(w.reshape(cout, 4, cin).transpose(1, 0, 2).reshape(4 * cout, cin))
    out = kernels.transposed_conv(f16(x.reshape(-1, cin), gpu), n_img, H, W, f16(w_sub, gpu), cout)
    torch.cuda.synchronize()
    assert_fp16_close(out, ref, acc, "transposed")


@pytest.mark.parametrize("H,W,cin,cout,n_img", [(17, 30, 256, 512, 1), (8, 12, 128, 64, 2), (1, 3, 64, 32, 1)])
def test_subpel_conv3x3(gpu, H, W, cin, cout, n_img):
    """SubpelConv2x(k=3, padding=1, bias) (layers.py:92-103): pixel_shuffle(conv3x3(x) + b, 2),
    the HT-L decoder up conv, against a float64 torch reference."""
    from dcvc_amd import kernels
    r = rng(H * 3 + cin)
    x = r.standard_normal((n_img, H, W, cin)).astype(np.float16)
    w = (r.standard_normal((4 * cout, cin, 3, 3)) / np.sqrt(9 * cin)).astype(np.float16)  # up.conv.0.weight
    b = (r.standard_normal(4 * cout) * 0.1).astype(np.float16).astype(np.float32)
    xt = torch.from_numpy(x.astype(np.float64)).permute(0, 3, 1, 2)
    ref = F.pixel_shuffle(F.conv2d(xt, torch.from_numpy(w.astype(np.float64)), torch.from_numpy(b.astype(np.float64)),
                                   padding=1), 2)
    ref = ref.permute(0, 2, 3, 1).reshape(-1, cout).numpy()
    acc = F.pixel_shuffle(F.conv2d(xt.abs(), torch.from_numpy(np.abs(w.astype(np.float64))), padding=1), 2)
    acc = acc.permute(0, 2, 3, 1).reshape(-1, cout).numpy()
    # GEMM row sub*Cout + c <- conv channel c*4 + sub; K = tap*Cin + ci
    w_sub = w.reshape(cout, 4, cin, 9).transpose(1, 0, 3, 2).reshape(4 * cout, 9 * cin)
    b_sub = b.reshape(cout, 4).T.reshape(-1)
    out = kernels.subpel_conv3x3(f16(np.ascontiguousarray(x.reshape(-1, cin)), gpu), n_img, H, W,
                                 f16(np.ascontiguousarray(w_sub), gpu), cout, torch.from_numpy(b_sub.copy()).to(gpu))
    torch.cuda.synchronize()
    assert_fp16_close(out, ref, acc, "subpel3x3")


def test_shortcut_then_quant_equals_separate_multiply(gpu):
    """The HT-L block tail fused in one GEMM equals the reference's two
    kernels, conv1x1_bias_shortcut then multiply_with_broadcast, bit for bit;
    the single-rounding fused form (HT-S) differs on some elements."""
    from dcvc_amd import kernels
    r = rng(17)
    M, K, N = 3001, 512, 512
    a = f16(r.standard_normal((M, K)), gpu)
    w = f16(r.standard_normal((N, K)) / 20, gpu)
    b = torch.from_numpy((r.standard_normal(N) * 0.1).astype(np.float32)).to(gpu)
    s1 = f16(r.standard_normal((M, N)), gpu)
    qtab = f16(0.5 + r.random((64, N)), gpu)
    qp = torch.tensor([37], dtype=torch.int32, device=gpu)
    fused = kernels.conv1x1(a, w, b, epi="shortcut_then_quant", s1=s1, q=qtab, qp=qp)
    two = kernels.multiply_with_broadcast(kernels.conv1x1(a, w, b, epi="shortcut", s1=s1), qtab, qp)
    single = kernels.conv1x1(a, w, b, epi="shortcut_quant", s1=s1, q=qtab, qp=qp)
    torch.cuda.synchronize()
    assert torch.equal(fused.view(torch.int16), two.view(torch.int16))
    assert not torch.equal(single.view(torch.int16), two.view(torch.int16))


@pytest.mark.parametrize("H,W,C,n_img", [(136, 240, 256, 1), (17, 30, 128, 2), (1, 1, 64, 1), (68, 120, 768, 1)])
def test_d3x3(gpu, H, W, C, n_img):
    from dcvc_amd import kernels
    r = rng(H * C)
    x = r.standard_normal((n_img, H, W, C)).astype(np.float16)
    w = (r.standard_normal((C, 3, 3)) * 0.3).astype(np.float16)
    ref = np.concatenate([O.ref_dw3x3(x[i], w).reshape(-1, C) for i in range(n_img)])
    acc = np.concatenate([O.ref_dw3x3(np.abs(x[i]), np.abs(w)).reshape(-1, C) for i in range(n_img)])
    wt9 = np.ascontiguousarray(w.reshape(C, 9).T)
    out = kernels.d3x3(f16(x.reshape(-1, C), gpu), n_img, H, W, f16(wt9, gpu))
    torch.cuda.synchronize()
    assert_fp16_close(out, ref, acc, "d3x3")


# 4-row strips (the first three: fewer than 480 blocks of 8-row strips) and
# 8-row strips (the last two); ragged strip ends and odd widths
@pytest.mark.parametrize("H,W,C,n_img", [(68, 120, 256, 1), (34, 60, 256, 2), (9, 5, 64, 3), (136, 240, 256, 1),
                                         (135, 240, 512, 1)])
def test_d3x3_bit_exact_sequential_fp32(gpu, H, W, C, n_img):
    """Each output is the fp32 sum of its nine taps in (ky, kx) order,
    starting from 0, rounded to fp16 once.  An fp16 x fp16 product is exact
    in fp32, so numpy's float32 adds reproduce the kernel's fused
    multiply-adds bit for bit, whichever strip length the launch uses."""
    from dcvc_amd import kernels
    r = rng(H * W + C)
    x = r.standard_normal((n_img, H, W, C)).astype(np.float16)
    w = (r.standard_normal((C, 3, 3)) * 0.3).astype(np.float16)
    xp = np.pad(x.astype(np.float32), ((0, 0), (1, 1), (1, 1), (0, 0)))
    s = np.zeros((n_img, H, W, C), np.float32)
    for ky in range(3):
        for kx in range(3):
            s = s + xp[:, ky:ky + H, kx:kx + W, :] * w[:, ky, kx].astype(np.float32)
    wt9 = np.ascontiguousarray(w.reshape(C, 9).T)
    out = kernels.d3x3(f16(x.reshape(-1, C), gpu), n_img, H, W, f16(wt9, gpu))
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint16)
    want = s.astype(np.float16).reshape(-1, C).view(np.uint16)
    assert np.array_equal(got, want), int((got != want).sum())


# ------------------------------------------------------------------ layout (bit-exact)

# f = 8 with 3 or 24 packed channels and W * cin % 8 == 0 takes the row-segment
# kernel (its right-padding segments included); the rest the per-element one
@pytest.mark.parametrize("H,W,cin,f,pb,pr,n_img", [(64, 96, 3, 8, 0, 0, 1), (60, 90, 3, 8, 4, 6, 1),
                                                   (1080, 1920, 24, 8, 8, 0, 1), (32, 48, 64, 2, 0, 0, 1),
                                                   (60, 90, 24, 8, 4, 6, 2), (1080, 1920, 3, 8, 8, 0, 1),
                                                   (20, 36, 3, 8, 4, 4, 3)])
def test_pad_and_unshuffle(gpu, H, W, cin, f, pb, pr, n_img):
    from dcvc_amd import kernels
    x = torch.from_numpy(rng(H + n_img).standard_normal((n_img, cin, H, W)).astype(np.float16))
    ref = F.pixel_unshuffle(F.pad(x.float(), (0, pr, 0, pb), mode="replicate"), f).half()
    ref = ref.permute(0, 2, 3, 1).reshape(-1, cin * f * f)
    xg = x.permute(0, 2, 3, 1).contiguous().reshape(-1, cin).to(gpu)
    out = kernels.pad_and_unshuffle(xg, n_img, H, W, cin, f, pb, pr)
    torch.cuda.synchronize()
    assert torch.equal(out.cpu(), ref)


# C = 3 with packed output takes the 8-pixel-group kernel; C = 4 the per-pixel one
@pytest.mark.parametrize("H8,W8,C,n_img", [(17, 30, 3, 1), (136, 240, 3, 1), (5, 7, 3, 2), (17, 30, 4, 1)])
def test_pixel_shuffle_8_clamp(gpu, H8, W8, C, n_img):
    from dcvc_amd import kernels
    x = torch.from_numpy(rng(2 + H8).standard_normal((n_img, 64 * C, H8, W8)).astype(np.float16))
    ref = torch.clamp(F.pixel_shuffle(x.float(), 8), -0.5, 0.5).half().permute(0, 2, 3, 1).reshape(-1, C)
    out = kernels.pixel_shuffle_8(x.permute(0, 2, 3, 1).contiguous().reshape(-1, 64 * C).to(gpu), n_img, H8, W8, C)
    torch.cuda.synchronize()
    assert torch.equal(out.cpu(), ref)


def test_multiply_round_int8_pad(gpu):
    from dcvc_amd import kernels
    r = rng(9)
    x = r.standard_normal((999, 512)).astype(np.float16)
    q = (0.5 + r.random((64, 512))).astype(np.float16)
    qp = torch.tensor([37], dtype=torch.int32, device=gpu)
    out = kernels.multiply_with_broadcast(f16(x, gpu), f16(q, gpu), qp)
    ref = (x.astype(np.float32) * q[37].astype(np.float32)[None]).astype(np.float16)
    assert np.array_equal(out.cpu().numpy(), ref)
    z = (r.standard_normal((510, 128)) * 40).astype(np.float16)
    z[0, :4] = [0.5, -0.5, 2.5, -70.0]
    zhat, zi8 = kernels.round_z(f16(z, gpu))
    want = np.clip(O.round_half_away(z.astype(np.float32)), -64, 63)
    assert np.array_equal(zhat.cpu().numpy().astype(np.float32), want)
    assert np.array_equal(zi8.cpu().numpy(), want.reshape(-1).astype(np.int8))
    back = kernels.int8_to_half(zi8, 510, 128)
    assert np.array_equal(back.cpu().numpy().astype(np.float32), want)
    xs = f16(r.standard_normal((2 * 30 * 54, 64)), gpu)
    padded = kernels.replicate_pad(xs, 2, 30, 54, 32, 56)
    ref = F.pad(xs.cpu().float().view(2, 30, 54, 64).permute(0, 3, 1, 2), (0, 2, 0, 2), mode="replicate")
    assert torch.equal(padded.cpu().float(), ref.permute(0, 2, 3, 1).reshape(-1, 64))
    torch.cuda.synchronize()


# ------------------------------------------------------------------ quantisation (bit-exact)

def _rand_prior(r, P, C, heavy_ties=True):
    y = (r.standard_normal((P, C)) * 3).astype(np.float16)
    if heavy_ties:  # exact .5 residuals exercise the half-away rounding
        y[:, ::7] = (np.round(r.standard_normal((P, (C + 6) // 7)) * 4) + 0.5).astype(np.float16)
    scales = np.abs(r.standard_normal((P, C)) * 4).astype(np.float16)
    scales[:, ::5] = -scales[:, ::5]
    means = (r.standard_normal((P, C)) * 2).astype(np.float16)
    means[:, ::7] = 0
    qdec = (0.25 + r.random((P, C)) * 2).astype(np.float16)
    return y, scales, means, qdec


@pytest.mark.parametrize("H,W,n_img,thr", [(17, 30, 1, 0.0), (16, 24, 2, 0.15), (68, 120, 1, 0.0)])
def test_hts_quant_and_restore_steps(gpu, H, W, n_img, thr):
    from dcvc_amd import kernels
    r = rng(H * 31 + n_img)
    C, P = 256, n_img * H * W
    y, scales, mean0, qdec = _rand_prior(r, P, C)
    means_steps = [mean0] + [(r.standard_normal((P, C)) * 2).astype(np.float16) for _ in range(3)]
    yhat_g = torch.full((P, C), 7.0, dtype=torch.float16, device=gpu)
    sym_g = torch.zeros(P * C, dtype=torch.int16, device=gpu)
    cond_g = torch.zeros(P * C, dtype=torch.uint8, device=gpu)
    yhat_r = np.full((P, C), 7.0, np.float16)
    sym_r = np.zeros((P, C), np.int16)
    cond_r = np.zeros((P, C), np.uint8)
    for step in range(4):
        kernels.hts_quant_step(f16(y, gpu), n_img, H, W, f16(scales, gpu), f16(means_steps[step], gpu),
                               f16(qdec, gpu), yhat_g, sym_g, cond_g, thr, step)
        yhat_r, s, c, act = O.hts_quant_step_f32(y, scales, means_steps[step], qdec, yhat_r, step, thr, C, H, W, n_img)
        sym_r[act] = s[act]
        cond_r[act] = c[act]
    torch.cuda.synchronize()
    assert np.array_equal(yhat_g.cpu().numpy().view(np.uint16), yhat_r.view(np.uint16))
    assert np.array_equal(sym_g.cpu().numpy(), sym_r.reshape(-1))
    assert np.array_equal(cond_g.cpu().numpy(), cond_r.reshape(-1))
    # decoder: index table + restore from the encoder's symbols reproduces y_hat bit-exactly
    idx, cond = kernels.hts_build_index_dec(f16(scales, gpu), n_img, H, W, thr)
    assert np.array_equal(cond.cpu().numpy(), cond_r.reshape(-1))
    assert np.array_equal(idx.cpu().numpy(), (sym_r.reshape(-1).astype(np.int64) & 0xFF).astype(np.uint8))
    yq = torch.from_numpy((sym_r.reshape(-1).astype(np.int32) >> 8).astype(np.int8)).to(gpu)
    yq = torch.where(cond.bool(), yq, torch.zeros_like(yq))
    yh_dec = torch.full((P, C), -3.0, dtype=torch.float16, device=gpu)
    for step in range(4):
        kernels.hts_restore_step(yq, n_img, H, W, C, f16(means_steps[step], gpu), f16(qdec, gpu), yh_dec, step)
    torch.cuda.synchronize()
    assert np.array_equal(yh_dec.cpu().numpy().view(np.uint16), yhat_r.view(np.uint16))


@pytest.mark.parametrize("H,W,n_img,thr", [(16, 16, 1, 0.0), (17, 30, 2, 0.15)])
def test_intra_quant_and_restore_steps(gpu, H, W, n_img, thr):
    from dcvc_amd import kernels
    r = rng(H * 13 + n_img)
    C, P = 256, n_img * H * W
    y, _, _, _ = _rand_prior(r, P, C)
    sc_steps = [np.abs(r.standard_normal((P, C)) * 3).astype(np.float16) for _ in range(4)]
    mn_steps = [(r.standard_normal((P, C)) * 2).astype(np.float16) for _ in range(4)]
    q = (0.5 + r.random((64, C))).astype(np.float16)
    qp = torch.tensor([5], dtype=torch.int32, device=gpu)
    ysf_g = torch.full((P, C), 9.0, dtype=torch.float16, device=gpu)
    yout_g = torch.zeros((P, C), dtype=torch.float16, device=gpu)
    ysf_r = np.full((P, C), 9.0, np.float16)
    syms = []
    for step in range(4):
        sym_g = torch.zeros(P * C // 4, dtype=torch.int16, device=gpu)
        cond_g = torch.zeros(P * C // 4, dtype=torch.uint8, device=gpu)
        kernels.intra_quant_step(f16(y, gpu), n_img, H, W, f16(sc_steps[step], gpu), f16(mn_steps[step], gpu),
                                 ysf_g, f16(q, gpu), qp, yout_g, sym_g, cond_g, thr, step)
        ysf_r, yout_r, s, c = O.intra_quant_step_f32(y, sc_steps[step], mn_steps[step], ysf_r, q[5], step, thr,
                                                     C, H, W, n_img)
        torch.cuda.synchronize()
        assert np.array_equal(sym_g.cpu().numpy(), s), f"step {step}"
        assert np.array_equal(cond_g.cpu().numpy(), c), f"step {step}"
        syms.append((s, c))
        idx, cnd = kernels.intra_build_index_dec(f16(sc_steps[step], gpu), n_img, H, W, thr, step)
        assert np.array_equal(cnd.cpu().numpy(), c)
        assert np.array_equal(idx.cpu().numpy(), (s.astype(np.int64) & 0xFF).astype(np.uint8))
    assert np.array_equal(yout_g.cpu().numpy().view(np.uint16), yout_r.view(np.uint16))
    # decoder restore over the same quarters
    ysf_d = torch.full((P, C), -1.0, dtype=torch.float16, device=gpu)
    yout_d = torch.zeros((P, C), dtype=torch.float16, device=gpu)
    for step in range(4):
        s, c = syms[step]
        yq = np.where(c > 0, (s.astype(np.int32) >> 8), 0).astype(np.int8)
        kernels.intra_restore_step(torch.from_numpy(yq).to(gpu), n_img, H, W, C, f16(mn_steps[step], gpu), ysf_d,
                                   f16(q, gpu), qp, yout_d, step)
    torch.cuda.synchronize()
    assert np.array_equal(yout_d.cpu().numpy().view(np.uint16), yout_r.view(np.uint16))


@pytest.mark.parametrize("H,W,n_img,thr", [(17, 30, 1, 0.0), (16, 24, 2, 0.15), (68, 120, 1, 0.15)])
def test_ld_quant_and_restore_steps(gpu, H, W, n_img, thr):
    """LD boundary: 2x checkerboard over channel halves, the common scales in
    both halves, y_hat * max(q_dec, .5) after the second; bit-exact against the
    float32 restatement, and the decoder's indexes and restore reproduce
    y_hat from the encoder's symbols."""
    from dcvc_amd import kernels
    r = rng(H * 5 + n_img)
    C, P = 128, n_img * H * W
    y, scales, mean0, qdec = _rand_prior(r, P, C)
    means_steps = [mean0, (r.standard_normal((P, C)) * 2).astype(np.float16)]
    yhat_g = torch.full((P, C), 7.0, dtype=torch.float16, device=gpu)
    sym_g = torch.zeros(P * C, dtype=torch.int16, device=gpu)
    cond_g = torch.zeros(P * C, dtype=torch.uint8, device=gpu)
    yhat_r = np.full((P, C), 7.0, np.float16)
    sym_r = np.zeros((P, C), np.int16)
    cond_r = np.zeros((P, C), np.uint8)
    for step in range(2):
        kernels.ld_quant_step(f16(y, gpu), n_img, H, W, f16(scales, gpu), f16(means_steps[step], gpu),
                              f16(qdec, gpu), yhat_g, sym_g, cond_g, thr, step)
        yhat_r, s, c, act = O.ld_quant_step_f32(y, scales, means_steps[step], qdec, yhat_r, step, thr, C, H, W, n_img)
        sym_r[act] = s[act]
        cond_r[act] = c[act]
    torch.cuda.synchronize()
    assert np.array_equal(yhat_g.cpu().numpy().view(np.uint16), yhat_r.view(np.uint16))
    assert np.array_equal(sym_g.cpu().numpy(), sym_r.reshape(-1))
    assert np.array_equal(cond_g.cpu().numpy(), cond_r.reshape(-1))
    idx, cond = kernels.hts_build_index_dec(f16(scales, gpu), n_img, H, W, thr)
    assert np.array_equal(cond.cpu().numpy(), cond_r.reshape(-1))
    yq = torch.from_numpy((sym_r.reshape(-1).astype(np.int32) >> 8).astype(np.int8)).to(gpu)
    yq = torch.where(cond.bool(), yq, torch.zeros_like(yq))
    yh_dec = torch.full((P, C), -3.0, dtype=torch.float16, device=gpu)
    for step in range(2):
        kernels.ld_restore_step(yq, n_img, H, W, C, f16(means_steps[step], gpu), f16(qdec, gpu), yh_dec, step)
    torch.cuda.synchronize()
    assert np.array_equal(yh_dec.cpu().numpy().view(np.uint16), yhat_r.view(np.uint16))


@pytest.mark.parametrize("H,W,n_img,thr", [(16, 16, 1, 0.0), (17, 30, 2, 0.15), (68, 120, 1, 0.15)])
def test_htl_quant_and_restore_steps(gpu, H, W, n_img, thr):
    """HT-L boundary: per-step scales and means, y / max(q_dec, .5) quantised,
    (y_hat_3 + y_hat_so_far) * max(q_dec, .5) at the end; bit-exact against the
    float32 restatement, and the decoder's indexes and restore reproduce the
    encoder's quarters and y_hat from its symbols."""
    from dcvc_amd import kernels
    r = rng(H * 7 + n_img)
    C, P = 256, n_img * H * W
    y, _, _, qdec = _rand_prior(r, P, C)
    qdec[:, ::9] = np.float16(0.125)                            # below the 0.5 clamp
    sc_steps = [np.abs(r.standard_normal((P, C)) * 3).astype(np.float16) for _ in range(4)]
    for sc in sc_steps:
        sc[:, ::11] = np.float16(thr)                           # exactly at the skip threshold
    mn_steps = [(r.standard_normal((P, C)) * 2).astype(np.float16) for _ in range(4)]
    ysf_g = torch.full((P, C), 9.0, dtype=torch.float16, device=gpu)
    yout_g = torch.zeros((P, C), dtype=torch.float16, device=gpu)
    ysf_r = np.full((P, C), 9.0, np.float16)
    syms = []
    for step in range(4):
        sym_g = torch.zeros(P * C // 4, dtype=torch.int16, device=gpu)
        cond_g = torch.zeros(P * C // 4, dtype=torch.uint8, device=gpu)
        kernels.htl_quant_step(f16(y, gpu), n_img, H, W, f16(sc_steps[step], gpu), f16(mn_steps[step], gpu),
                               f16(qdec, gpu), ysf_g, yout_g, sym_g, cond_g, thr, step)
        ysf_r, yout_r, s, c = O.htl_quant_step_f32(y, sc_steps[step], mn_steps[step], qdec, ysf_r, step, thr,
                                                   C, H, W, n_img)
        torch.cuda.synchronize()
        assert np.array_equal(sym_g.cpu().numpy(), s), f"step {step}"
        assert np.array_equal(cond_g.cpu().numpy(), c), f"step {step}"
        if step < 3:
            assert np.array_equal(ysf_g.cpu().numpy().view(np.uint16), ysf_r.view(np.uint16)), f"step {step}"
        syms.append((s, c))
        idx, cnd = kernels.intra_build_index_dec(f16(sc_steps[step], gpu), n_img, H, W, thr, step)
        assert np.array_equal(cnd.cpu().numpy(), c)
        assert np.array_equal(idx.cpu().numpy(), (s.astype(np.int64) & 0xFF).astype(np.uint8))
    assert np.array_equal(yout_g.cpu().numpy().view(np.uint16), yout_r.view(np.uint16))
    ysf_d = torch.full((P, C), -1.0, dtype=torch.float16, device=gpu)
    yout_d = torch.zeros((P, C), dtype=torch.float16, device=gpu)
    ysf_rr = np.full((P, C), -1.0, np.float16)
    for step in range(4):
        s, c = syms[step]
        yq = np.where(c > 0, (s.astype(np.int32) >> 8), 0).astype(np.int8)
        kernels.htl_restore_step(torch.from_numpy(yq).to(gpu), n_img, H, W, C, f16(mn_steps[step], gpu),
                                 f16(qdec, gpu), ysf_d, yout_d, step)
        ysf_rr, yo = O.htl_restore_step_f32(yq, mn_steps[step], qdec, ysf_rr, step, C, H, W, n_img)
    torch.cuda.synchronize()
    assert np.array_equal(yo.view(np.uint16), yout_r.view(np.uint16))
    assert np.array_equal(yout_d.cpu().numpy().view(np.uint16), yout_r.view(np.uint16))


@pytest.mark.parametrize("N,p", [(0, 0.5), (1, 1.0), (4095, 0.5), (4097, 0.3), (2088960, 0.5), (100000, 0.0),
                                 (100000, 1.0), (8355840, 0.7)])
def test_conditional_index_and_recover(gpu, N, p):
    from dcvc_amd import kernels
    r = rng(N + 1)
    vals = r.integers(-32768, 32767, size=N, dtype=np.int16)
    cond = (r.random(N) < p).astype(np.uint8)
    out, n, offs = kernels.conditional_index(torch.from_numpy(vals).to(gpu), torch.from_numpy(cond).to(gpu))
    want = vals[cond.astype(bool)]
    assert n == len(want)
    assert np.array_equal(out.cpu().numpy(), want)
    v8 = r.integers(0, 255, size=N, dtype=np.uint8)
    out8, n8, offs8 = kernels.conditional_index(torch.from_numpy(v8).to(gpu), torch.from_numpy(cond).to(gpu))
    assert np.array_equal(out8.cpu().numpy(), v8[cond.astype(bool)])
    dec = r.integers(-128, 127, size=max(n, 1), dtype=np.int8)
    rec = kernels.conditional_recover(torch.from_numpy(dec).to(gpu), torch.from_numpy(cond).to(gpu), offs8)
    torch.cuda.synchronize()
    assert np.array_equal(rec.cpu().numpy(), O.recover(dec[:n], cond, N))


@pytest.mark.parametrize("epi,a_div", [("bias", 1), ("wsilu", 2), ("shortcut", 1), ("shortcut2", 8), ("quad", 1),
                                       ("quad", 4)])
def test_conv1x1_grouped_equals_per_problem(gpu, epi, a_div):
    """Grouped launches (the recon heads' per-frame branches) give every
    problem the bits of its own single-problem launch, with shared inputs."""
    from dcvc_amd import kernels
    r = rng(31 + a_div)
    nb, M, K, N = 8, 3001, 256, 192
    rows = 4 * N if epi == "quad" else N
    a = f16(r.standard_normal((nb // a_div, M, K)), gpu)
    w = f16(r.standard_normal((nb, rows, K)) / 16, gpu)
    b = torch.from_numpy((r.standard_normal((nb, rows)) * 0.1).astype(np.float32)).to(gpu)
    s1 = f16(r.standard_normal((nb, M, N)), gpu)
    s2 = f16(r.standard_normal((nb, M, N)), gpu)
    got = kernels.conv1x1_grouped(a, w, b, epi=epi, s1=s1, s2=s2, a_div=a_div)
    for i in range(nb):
        ref = kernels.conv1x1(a[i // a_div], w[i], b[i], epi=epi, s1=s1[i], s2=s2[i])
        torch.cuda.synchronize()
        assert torch.equal(got[i].view(torch.int16), ref.view(torch.int16)), f"problem {i}"


@pytest.mark.parametrize("prec,k", [(0, 16), (1, 16), (2, 64)])
def test_mfma_calibrate_closed_form(gpu, prec, k):
    """The calibration kernel's MFMAs on all-ones operands: every element of
    each of a wave's 16 accumulators gains K per MFMA, so every lane's sink is
    exactly 16 * 16 * K * iters (fp16 / bf16 K = 16, e4m3 K = 64), and all-zero
    operands give zero."""
    from dcvc_amd import _lib
    blocks, iters = 6, 25
    one = {0: torch.float16, 1: torch.bfloat16, 2: torch.float8_e4m3fn}[prec]
    esize = torch.tensor([], dtype=one).element_size()
    ops = torch.ones(blocks * 64 * 1024 // esize, dtype=one).view(torch.uint8).to(gpu)
    sink = torch.full((blocks * 256,), -1.0, dtype=torch.float32, device=gpu)
    st = torch.cuda.current_stream().cuda_stream
    _lib.call("dcvc_mfma_calibrate", ops.data_ptr(), sink.data_ptr(), blocks, iters, prec, st)
    torch.cuda.synchronize()
    assert torch.all(sink == float(16 * 16 * k * iters))
    ops.zero_()
    _lib.call("dcvc_mfma_calibrate", ops.data_ptr(), sink.data_ptr(), blocks, iters, prec, st)
    torch.cuda.synchronize()
    assert torch.all(sink == 0.0)
