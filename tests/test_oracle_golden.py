"""Pins the CPU oracle's network wiring and the host CDF builders against
golden vectors recorded from the reference's own Python modules on this
project's seeded weights (tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from dcvc_amd import entropy, weights
from oracle import dcvc_oracle as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_modules.npz"))


def g(key):
    return torch.from_numpy(G[key].astype(np.float32))


def close(a, b, rtol=2e-4, atol=2e-5):
    a = a.detach().float()
    b = b.detach().float()
    err = (a - b).abs().max().item()
    scale = b.abs().max().item()
    assert err <= atol + rtol * scale, f"max err {err:.3g} (scale {scale:.3g})"


@pytest.fixture(scope="module")
def sd_dmci():
    sd = weights.round_to_half(weights.perturb_biases(weights.init_state_dict("dmci", 0), 100))
    return weights.to_torch(sd)


@pytest.fixture(scope="module")
def sd_hts():
    sd = weights.round_to_half(weights.perturb_biases(weights.init_state_dict("hts", 1), 101))
    return weights.to_torch(sd)


def test_schema_param_counts():
    # HT-S: 81.2 M parameters published for DCVC-UF HT-S (README complexity figure)
    assert round(weights.param_count("hts") / 1e6, 1) == 81.2


def test_dmci_modules(sd_dmci):
    sd = sd_dmci
    q = sd["q_scale_enc"][5]
    y = O.dmci_encoder(sd, F.pixel_unshuffle(g("dmci.enc.in"), 8), q)
    close(y, g("dmci.enc.out"))
    close(O.dmci_hyper_enc(sd, g("dmci.enc.out")), g("dmci.hyper_enc.out"))
    close(O.dmci_hyper_dec(sd, g("dmci.hyper_dec.in")), g("dmci.hyper_dec.out"))
    close(O.dmci_prior_fusion(sd, g("dmci.hyper_dec.out")), g("dmci.y_prior_fusion.out"))
    close(O.conv(sd, "y_spatial_prior_reduction.", g("dmci.y_prior_fusion.out")), g("dmci.reduction.out"))
    for i in (1, 2, 3):
        close(O.dmci_spatial_prior(sd, i, g("dmci.sp.cat_input")), g(f"dmci.sp{i}.out"))
    # the proxy clamps the shuffled reconstruction to [-0.5, 0.5] (shuffle.cu:237-240)
    close(O.dmci_decoder(sd, g("dmci.dec.in"), sd["q_scale_dec"][7]), torch.clamp(g("dmci.dec.out"), -0.5, 0.5))


def test_hts_modules(sd_hts):
    sd = sd_hts
    mem = O.hts_feature_adaptor_i(sd, g("hts.fa_i.in"))
    close(mem, g("hts.fa_i.out"))
    close(O.hts_feature_adaptor_m(sd, g("hts.fa_i.out"), g("hts.fa_m.in_feature")), g("hts.fa_m.out"))
    ctx = O.hts_feature_extractor(sd, g("hts.fa_i.out"))
    close(ctx, g("hts.fe.out"))
    close(O.hts_encoder(sd, g("hts.encoder.in"), g("hts.fe.out"), 9), g("hts.encoder.out"))
    close(O.hts_hyper_encoder(sd, g("hts.encoder.out")), g("hts.hyper_encoder.out"))
    close(O.hts_hyper_decoder(sd, g("hts.hyper_decoder.in")), g("hts.hyper_decoder.out"))
    close(O.hts_temporal_prior(sd, g("hts.fa_i.out"), 9), g("hts.tpe.out"))
    params = O._hts_prior(sd, None, g("hts.hyper_decoder.in"), g("hts.fa_i.out"), 9)
    close(params, g("hts.prior_fusion.out"))
    red = O.conv(sd, "y_spatial_prior_reduction.", g("hts.prior_fusion.out"))
    close(red, g("hts.reduction.out"))
    for i in (1, 2, 3):
        close(O._hts_spatial_prior(sd, i, g("hts.sp.in_yhat"), g("hts.reduction.out")), g(f"hts.sp{i}.out"))
    feat = O._hts_decoder(sd, g("hts.decoder.in"), g("hts.fe.out"), 9)
    close(feat, g("hts.decoder.out"))
    outs, head = O.hts_recon_head(sd, g("hts.decoder.out"))
    for i in range(8):
        close(outs[i], torch.clamp(g(f"hts.recon.{i}"), -0.5, 0.5))
    _, reset_head = O.hts_recon_head(sd, g("hts.decoder.out"), only_reset=True)
    close(reset_head, g("hts.recon.reset"))
    close(head, g("hts.recon.reset"))


def test_masks_match_reference():
    ours = O.mask_4x(16, 5, 6)
    for i in range(4):
        assert torch.equal(ours[i].float(), g(f"mask4x.{i}"))


def test_cdf_tables_match_reference():
    table = entropy.gaussian_scale_table()
    assert torch.allclose(table, g("gaussian.scale_table"), rtol=0, atol=0)
    cdf, length = entropy.gaussian_cdf_table()
    assert np.array_equal(cdf, G["cdf.gaussian.quantized_cdf"])
    assert np.array_equal(length, G["cdf.gaussian.cdf_length"])
    for model, seed, key in (("dmci", 0, "dmci_z"), ("hts", 1, "hts_z")):
        sd = weights.round_to_half(weights.perturb_biases(weights.init_state_dict(model, seed), 100 + seed))
        zc, zl = entropy.bit_estimator_cdf_table(torch.from_numpy(sd["bit_estimator_z.h"]),
                                                 torch.from_numpy(sd["bit_estimator_z.b"]),
                                                 torch.from_numpy(sd["bit_estimator_z.a"]))
        assert np.array_equal(zl, G[f"cdf.{key}.cdf_length"])
        assert np.array_equal(zc, G[f"cdf.{key}.quantized_cdf"])


GL = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_htl.npz"))


def gl(key):
    return torch.from_numpy(GL[key].astype(np.float32))


@pytest.fixture(scope="module")
def sd_htl():
    sd = weights.round_to_half(weights.perturb_biases(weights.init_state_dict("htl", 2), 102))
    return weights.to_torch(sd)


def test_htl_modules(sd_htl):
    """Every HT-L submodule of the oracle against the reference modules
    (video_model_ht.py, ``not is_hts`` branches) on the same weights."""
    sd = sd_htl
    mem = O.htl_feature_adaptor_i(sd, gl("htl.fa_i.in"))
    close(mem, gl("htl.fa_i.out"))
    close(O.htl_feature_adaptor_m(sd, gl("htl.fa_i.out"), gl("htl.fa_m.in_feature")), gl("htl.fa_m.out"))
    close(O.htl_feature_extractor(sd, gl("htl.fa_i.out")), gl("htl.fe.out"))
    close(O.htl_encoder(sd, gl("htl.encoder.in"), gl("htl.fe.out"), 11), gl("htl.encoder.out"))
    close(O.htl_hyper_encoder(sd, gl("htl.encoder.out")), gl("htl.hyper_encoder.out"))
    close(O.htl_hyper_decoder(sd, gl("htl.hyper_decoder.in")), gl("htl.hyper_decoder.out"))
    close(O.htl_temporal_prior(sd, gl("htl.fa_i.out"), 11), gl("htl.tpe.out"))
    close(O._htl_prior(sd, gl("htl.hyper_decoder.in"), gl("htl.fa_i.out"), 11), gl("htl.prior_fusion.out"))
    close(O.conv(sd, "y_spatial_prior_reduction.", gl("htl.prior_fusion.out")), gl("htl.reduction.out"))
    for i in (1, 2, 3):
        sc, mn = O._htl_spatial_prior(sd, i, gl("htl.sp.in_yhat"), gl("htl.reduction.out"))
        close(torch.cat([sc, mn], 1), gl(f"htl.sp{i}.out"))
    close(O._htl_decoder(sd, gl("htl.decoder.in"), gl("htl.fe.out"), 11), gl("htl.decoder.out"))
    outs, head = O.htl_recon_head(sd, gl("htl.decoder.out"))
    for i in range(8):
        close(outs[i], torch.clamp(gl(f"htl.recon.{i}"), -0.5, 0.5))
    _, reset_head = O.htl_recon_head(sd, gl("htl.decoder.out"), only_reset=True)
    close(reset_head, gl("htl.recon.reset"))
    close(head, gl("htl.recon.reset"))


GD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_ld.npz"))


def gd(key):
    return torch.from_numpy(GD[key].astype(np.float32))


def test_ld_modules():
    """Every low-delay submodule of the oracle against the reference modules
    (video_model_ld.py) on the same weights, and the 2x checkerboard masks."""
    sd = weights.to_torch(weights.round_to_half(weights.perturb_biases(weights.init_state_dict("ld", 3), 103)))
    close(O.ld_feature_adaptor_i(sd, gd("ld.fa_i.in")), gd("ld.fa_i.out"))
    close(O.ld_feature_adaptor_m(sd, gd("ld.fa_i.out"), gd("ld.fa_m.in_feature")), gd("ld.fa_m.out"))
    close(O.ld_feature_extractor(sd, gd("ld.fa_i.out")), gd("ld.fe.out"))
    close(O.ld_encoder(sd, gd("ld.encoder.in"), gd("ld.fe.out"), 13), gd("ld.encoder.out"))
    close(O.ld_hyper_encoder(sd, gd("ld.encoder.out")), gd("ld.hyper_encoder.out"))
    close(O.ld_hyper_decoder(sd, gd("ld.hyper_decoder.in")), gd("ld.hyper_decoder.out"))
    close(O.ld_temporal_prior(sd, gd("ld.fa_i.out")), gd("ld.tpe.out"))
    close(O._ld_prior(sd, gd("ld.hyper_decoder.in"), gd("ld.tpe.out"), 13), gd("ld.prior_fusion.out"))
    close(O.ld_spatial_prior(sd, gd("ld.sp.in_yhat"), gd("ld.prior_fusion.out")), gd("ld.sp.out"))
    close(O._ld_decoder(sd, gd("ld.decoder.in"), gd("ld.fe.out"), 13), gd("ld.decoder.out"))
    x_hat, head = O.ld_recon_head(sd, gd("ld.decoder.out"))
    close(x_hat, torch.clamp(gd("ld.recon"), -0.5, 0.5))
    close(head, gd("ld.recon.reset"))
    ours = O.mask_2x(8, 5, 6)
    for i in range(2):
        assert torch.equal(ours[i].float(), gd(f"mask2x.{i}"))
