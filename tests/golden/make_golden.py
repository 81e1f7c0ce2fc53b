"""Writes the golden vectors in tests/golden/ from the reference's own Python
model code (DCVC-UF, src/models/*.py, src/layers/layers.py).

Run in the build container only (the reference tree does not exist where the
GPU tests run):  python tests/golden/make_golden.py [--reference /root/reference]

What is recorded, on this project's seeded weights (loaded with
strict=True, which also checks that the state-dict schema matches):
  * every DMCI and HT-S submodule the codec proxies run, on seeded inputs at a
    tiny resolution, input and output (float32);
  * the reference's checkerboard masks (get_mask_4x) and scale table;
  * the CDF tables of GaussianEncoder.update / BitEstimator.update, built
    with the reference's own entropy coder compiled into oracle/_ref.
These pin the oracle's network wiring and the host CDF builders; they are
data only, and the reference itself is never imported by the tests.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from dcvc_amd import weights  # noqa: E402


def seeded(shape, seed, scale=1.0):
    """Seeded input, rounded to fp16 values so it is stored exactly as float16."""
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(shape, generator=g) * scale).half().float()


def load(model, sd):
    missing, unexpected = model.load_state_dict({k: torch.from_numpy(v) for k, v in sd.items()}, strict=True)
    assert not missing and not unexpected
    model.eval()
    return model


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default=os.environ.get("DCVC_REFERENCE", "/root/reference"))
    ap.add_argument("--parts", default="base,htl,ld",
                    help="comma list: base (DMCI, HT-S, masks, CDFs), htl (HT-L modules), ld (low-delay modules)")
    args = ap.parse_args()
    sys.path.insert(0, args.reference)
    sys.path.insert(0, os.path.join(REPO, "oracle", "_ref"))
    from src.models.image_model import DMCI
    from src.models.video_model_ht import DMC
    from src.utils.common import ModelStructure

    parts = args.parts.split(",")
    if "base" in parts:
        save(base_vectors(DMCI, DMC, ModelStructure), "reference_modules.npz")
    if "htl" in parts:
        save(htl_vectors(DMC, ModelStructure), "reference_htl.npz")
    if "ld" in parts:
        from src.models.video_model_ld import DMC as DMCLD
        save(ld_vectors(DMCLD), "reference_ld.npz")


def base_vectors(DMCI, DMC, ModelStructure):
    """DMCI and HT-S submodules, masks, scale table and CDF tables."""
    torch.manual_seed(0)
    out = {}
    with torch.no_grad():
        # ---------------------------------------------------------- DMCI
        sd = weights.round_to_half(weights.perturb_biases(weights.init_state_dict("dmci", 0), 100))
        m = load(DMCI(), sd)
        x = seeded((1, 3, 64, 64), 1, 0.25)
        q = torch.from_numpy(sd["q_scale_enc"][5])[None, :, None, None]
        out["dmci.enc.in"] = x
        out["dmci.enc.out"] = y = m.enc(x, q)
        out["dmci.hyper_enc.out"] = z = m.hyper_enc(y)
        z_hat = torch.round(z * 3)
        out["dmci.hyper_dec.in"] = z_hat
        out["dmci.hyper_dec.out"] = h = m.hyper_dec(z_hat)
        out["dmci.y_prior_fusion.out"] = p = m.y_prior_fusion(h)
        out["dmci.reduction.out"] = red = m.y_spatial_prior_reduction(p)
        yh = seeded((1, 256, 4, 4), 2)
        out["dmci.sp.cat_input"] = cat = torch.cat([yh, red], 1)
        for i in (1, 2, 3):
            a = getattr(m, f"y_spatial_prior_adaptor_{i}")(cat)
            out[f"dmci.sp{i}.out"] = m.y_spatial_prior(a)
        qd = torch.from_numpy(sd["q_scale_dec"][7])[None, :, None, None]
        out["dmci.dec.in"] = yd = seeded((1, 256, 4, 4), 3)
        out["dmci.dec.out"] = m.dec(yd, qd)
        masks = m.get_mask_4x(1, 16, 5, 6, "cpu")
        for i, mk in enumerate(masks):
            out[f"mask4x.{i}"] = mk.float()
        out["gaussian.scale_table"] = m.gaussian_encoder.get_scale_table().float()
        # CDF tables through the reference's update() and its own C++ coder
        m.update(0.0)
        qc, ql = m.gaussian_encoder.get_cdf_info()
        out["cdf.gaussian.quantized_cdf"] = torch.from_numpy(qc)
        out["cdf.gaussian.cdf_length"] = torch.from_numpy(ql)
        zc, zl = m.bit_estimator_z.get_cdf_info()
        out["cdf.dmci_z.quantized_cdf"] = torch.from_numpy(zc)
        out["cdf.dmci_z.cdf_length"] = torch.from_numpy(zl)

        # ---------------------------------------------------------- HT-S
        sd = weights.round_to_half(weights.perturb_biases(weights.init_state_dict("hts", 1), 101))
        m = load(DMC(ModelStructure.HTS), sd)
        H8 = W8 = 8
        qe = torch.from_numpy(sd["q_encoder"][9])[None, :, None, None]
        qdd = torch.from_numpy(sd["q_decoder"][9])[None, :, None, None]
        qf = torch.from_numpy(sd["q_feature"][9])[None, :, None, None]
        fi = seeded((1, 192, H8, W8), 10, 0.3)
        out["hts.fa_i.in"] = fi
        out["hts.fa_i.out"] = mem = m.feature_adaptor_i(fi)
        fp = seeded((1, 512, H8, W8), 11, 0.5)
        out["hts.fa_m.in_feature"] = fp
        out["hts.fa_m.out"] = m.feature_adaptor_m(mem, fp)
        out["hts.fe.out"] = ctx = m.feature_extractor(mem)
        x = seeded((1, 24, 64, 64), 12, 0.25)
        out["hts.encoder.in"] = x
        out["hts.encoder.out"] = y = m.encoder(x, ctx, qe)
        out["hts.hyper_encoder.out"] = z = m.hyper_encoder(y)
        z_hat = torch.round(z * 3)
        out["hts.hyper_decoder.in"] = z_hat
        out["hts.hyper_decoder.out"] = hp = m.hyper_decoder(z_hat)
        out["hts.tpe.out"] = tp = m.temporal_prior_encoder(mem, qf)
        out["hts.prior_fusion.out"] = params = m.y_prior_fusion(hp, tp)
        out["hts.reduction.out"] = red = m.y_spatial_prior_reduction(params)
        yh = seeded((1, 256, 4, 4), 13)
        out["hts.sp.in_yhat"] = yh
        for i in (1, 2, 3):
            a = getattr(m, f"y_spatial_prior_adaptor_{i}")(yh, red)
            out[f"hts.sp{i}.out"] = m.y_spatial_prior(a)
        yd = seeded((1, 256, 4, 4), 14)
        out["hts.decoder.in"] = yd
        out["hts.decoder.out"] = feat = m.decoder(yd, ctx, qdd)
        recon = m.recon_head(feat)
        for i, r in enumerate(recon):
            out[f"hts.recon.{i}"] = r
        out["hts.recon.reset"] = m.recon_head(feat, for_reset=True)
        m.update(0.0)
        zc, zl = m.bit_estimator_z.get_cdf_info()
        out["cdf.hts_z.quantized_cdf"] = torch.from_numpy(zc)
        out["cdf.hts_z.cdf_length"] = torch.from_numpy(zl)

    return out


def ld_vectors(DMCLD):
    """Every low-delay submodule the proxy runs (video_model_ld.py), on seeded
    inputs at 64 x 64 pixels, plus the reference's 2x checkerboard masks."""
    out = {}
    with torch.no_grad():
        sd = weights.round_to_half(weights.perturb_biases(weights.init_state_dict("ld", 3), 103))
        m = load(DMCLD(), sd)
        H8 = W8 = 8
        qe = torch.from_numpy(sd["q_encoder"][13])[None, :, None, None]
        qdd = torch.from_numpy(sd["q_decoder"][13])[None, :, None, None]
        qf = torch.from_numpy(sd["q_feature"][13])[None, :, None, None]
        fi = seeded((1, 192, H8, W8), 30, 0.3)
        out["ld.fa_i.in"] = fi
        out["ld.fa_i.out"] = mem = m.feature_adaptor_i(fi)
        fp = seeded((1, 256, H8, W8), 31, 0.5)
        out["ld.fa_m.in_feature"] = fp
        out["ld.fa_m.out"] = m.feature_adaptor_m(mem, fp)
        out["ld.fe.out"] = ctx = m.feature_extractor(mem)
        x = seeded((1, 3, 64, 64), 32, 0.25)
        out["ld.encoder.in"] = x
        out["ld.encoder.out"] = y = m.encoder(x, ctx, qe)
        out["ld.hyper_encoder.out"] = z = m.hyper_encoder(y)
        z_hat = torch.round(z * 3)
        out["ld.hyper_decoder.in"] = z_hat
        out["ld.hyper_decoder.out"] = hp = m.hyper_decoder(z_hat)
        out["ld.tpe.out"] = tp = m.temporal_prior_encoder(mem)
        out["ld.prior_fusion.out"] = params = m.y_prior_fusion(hp, tp, qf)
        yh = seeded((1, 128, 4, 4), 33)
        out["ld.sp.in_yhat"] = yh
        out["ld.sp.out"] = m.y_spatial_prior(yh, params)
        yd = seeded((1, 128, 4, 4), 34)
        out["ld.decoder.in"] = yd
        out["ld.decoder.out"] = feat = m.decoder(yd, ctx, qdd)
        out["ld.recon"] = m.recon_head(feat)
        out["ld.recon.reset"] = m.recon_head(feat, for_reset=True)
        for i, mk in enumerate(m.get_mask_2x(1, 8, 5, 6, "cpu")):
            out[f"mask2x.{i}"] = mk.float()
    return out


def save(out, name):
    arrays = {}
    for k, v in out.items():
        a = v.detach().numpy() if isinstance(v, torch.Tensor) else np.asarray(v)
        if ".in" in k:  # inputs are fp16 values: store exactly as float16
            assert np.array_equal(a.astype(np.float16).astype(a.dtype), a)
            a = a.astype(np.float16)
        arrays[k] = a
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"wrote {path}: {len(arrays)} arrays, {os.path.getsize(path) / 1e6:.2f} MB")


def htl_vectors(DMC, ModelStructure):
    """Every HT-L submodule the proxy runs (the ``not is_hts`` branches of
    video_model_ht.py), on seeded inputs at 64 x 64 pixels."""
    out = {}
    with torch.no_grad():
        sd = weights.round_to_half(weights.perturb_biases(weights.init_state_dict("htl", 2), 102))
        m = load(DMC(ModelStructure.HTL), sd)
        H8 = W8 = 8
        qe = torch.from_numpy(sd["q_encoder"][11])[None, :, None, None]
        qdd = torch.from_numpy(sd["q_decoder"][11])[None, :, None, None]
        qf = torch.from_numpy(sd["q_feature"][11])[None, :, None, None]
        fi = seeded((1, 192, H8, W8), 20, 0.3)
        out["htl.fa_i.in"] = fi
        out["htl.fa_i.out"] = mem = m.feature_adaptor_i(fi)
        fp = seeded((1, 512, H8, W8), 21, 0.5)
        out["htl.fa_m.in_feature"] = fp
        out["htl.fa_m.out"] = m.feature_adaptor_m(mem, fp)
        out["htl.fe.out"] = ctx = m.feature_extractor(mem)
        x = seeded((1, 24, 64, 64), 22, 0.25)
        out["htl.encoder.in"] = x
        out["htl.encoder.out"] = y = m.encoder(x, ctx, qe)
        out["htl.hyper_encoder.out"] = z = m.hyper_encoder(y)
        z_hat = torch.round(z * 3)
        out["htl.hyper_decoder.in"] = z_hat
        out["htl.hyper_decoder.out"] = hp = m.hyper_decoder(z_hat)
        out["htl.tpe.out"] = tp = m.temporal_prior_encoder(mem, qf)
        out["htl.prior_fusion.out"] = params = m.y_prior_fusion(hp, tp)
        out["htl.reduction.out"] = red = m.y_spatial_prior_reduction(params)
        yh = seeded((1, 256, 4, 4), 23)
        out["htl.sp.in_yhat"] = yh
        cat = torch.cat((yh, red), dim=1)
        for i in (1, 2, 3):
            a = getattr(m, f"y_spatial_prior_adaptor_{i}")(cat)
            out[f"htl.sp{i}.out"] = m.y_spatial_prior(a)
        yd = seeded((1, 256, 4, 4), 24)
        out["htl.decoder.in"] = yd
        out["htl.decoder.out"] = feat = m.decoder(yd, ctx, qdd)
        recon = m.recon_head(feat)
        for i, r in enumerate(recon):
            out[f"htl.recon.{i}"] = r
        out["htl.recon.reset"] = m.recon_head(feat, for_reset=True)
    return out


if __name__ == "__main__":
    main()
