#!/usr/bin/env python
"""1080p DCVC-UF HT-S chunk encode+decode throughput on one MI355X.

Workload: one stream of synthetic 1920x1080 YUV420 video (texture + integer
motion + 10 % occlusion): an intra frame, then chunks of 4 frames, each padded
to the model's native 8-frame chunk by repeating its last frame, qp 32,
seeded random-init weights.  The stream holds 8 distinct chunks (32 frames),
coded in order and then again from the first, so the temporal context is
always the previous chunk's.  A step is one chunk compress (encoder proxy)
followed by its decompress (decoder proxy), including the CPU entropy coder.
`value` counts the 4 real frames per step.

Also reported: p50/p99 step latency over the timed steps, the native
8-frame rate, the HT-L chunk model and the low-delay model on the same
stream, the per-launch layer table of one encode + decode (algorithmic
FLOPs / bytes, each launch timed with HIP events on its own stream), the
roofline of the kernel class that takes the most time, and the FP32 CPU
oracle timed on this host.

--profile-step runs the warm-up, then --profile-steps timed steps between
two marker launches (torch's spin kernel, used nowhere else), then a 256 MiB
device copy (the calibration kernel for the FETCH_SIZE / WRITE_SIZE
counters), and prints nothing else: the program to put under rocprofv3.
`python -m dcvc_amd.profiling` keeps only the dispatches between the last
two markers, so the summaries cover the timed steps alone.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "1080p encode+decode frames/s on 1 MI355X, BF16/FP8 vs FP16 plain; p99 latency"
PEAK_FP16_TFLOPS = 2500.0   # MI355X dense FP16/BF16 MFMA
PEAK_HBM_GBS = 8000.0
EPI_NAMES = {0: "bias", 1: "wsilu", 2: "shortcut", 3: "shortcut2", 4: "shortcut_quant", 5: "quant", 6: "quad"}
KIND_NAMES = {0: "conv1x1", 1: "conv_s2", 2: "subpel", 3: "d3x3", 4: "ffn"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def layer_trace(enc, dec, x, qp, pad_b, pad_r, H, W):
    """One encode and one decode with graphs off and the timed launch trace
    on; returns the (n, 8) record array of both sides."""
    import ctypes as C
    import torch
    from dcvc_amd import _lib
    lib = _lib.load()
    buf = np.zeros((50000, 15), np.int64)
    ptr = buf.ctypes.data_as(C.c_void_p)
    enc.proxy.set_use_graphs(False)
    dec.proxy.set_use_graphs(False)
    try:
        _lib.call("dcvc_trace_begin", 1)
        bits, ec = enc.proxy.compress(x, qp, False, pad_b, pad_r)
        torch.cuda.synchronize()
        n = lib.dcvc_trace_end(ptr, len(buf))
        rec_enc = buf[:n].copy()
        _lib.call("dcvc_trace_begin", 1)
        dec.proxy.decompress(bits, qp, H, W, ec, False)
        torch.cuda.synchronize()
        n = lib.dcvc_trace_end(ptr, len(buf))
        rec_dec = buf[:n].copy()
    finally:
        enc.proxy.set_use_graphs(True)
        dec.proxy.set_use_graphs(True)
    return rec_enc, rec_dec


def kernel_classes(recs):
    """Aggregate trace records per kernel class (kind, epilogue)."""
    out = {}
    for kind in np.unique(recs[:, 0]):
        for epi in np.unique(recs[recs[:, 0] == kind][:, 1]):
            r = recs[(recs[:, 0] == kind) & (recs[:, 1] == epi)]
            name = KIND_NAMES[int(kind)] + ("" if kind == 3 else "_" + EPI_NAMES[int(epi)])
            t = float(r[:, 7].sum()) * 1e-9
            out[name] = {"launches": int(len(r)), "time_ms": round(t * 1e3, 4),
                         "tflop": round(float(r[:, 5].sum()) / 1e12, 5), "gbyte": round(float(r[:, 6].sum()) / 1e9, 5),
                         "tflops": round(float(r[:, 5].sum()) / t / 1e12, 2) if t > 0 else None,
                         "gbs": round(float(r[:, 6].sum()) / t / 1e9, 1) if t > 0 else None,
                         "flops_per_launch": int(r[:, 5].sum() // len(r)),
                         "bytes_per_launch": int(r[:, 6].sum() // len(r)),
                         "avg_launch_ms": round(t * 1e3 / len(r), 5)}
    return dict(sorted(out.items(), key=lambda kv: -kv[1]["time_ms"]))


# tile configuration id -> Tile<BM, BN, WM, WN, STAGES[, K tile 64]> of the
# fp16 / bf16 dispatch (gemm.hip); the 64-deep K tiles (8-11, 14) serve plain
# 1x1 products only, and other launches take `TILES_FALLBACK`'s configuration
TILES_QUAD = {1: (256, 256, 2, 2, 5), 2: (128, 256, 2, 2, 5), 3: (128, 128, 4, 1, 5), 4: (128, 256, 2, 2, 3),
              5: (256, 256, 2, 2, 4), 6: (256, 256, 4, 2, 4), 7: (256, 256, 4, 2, 3),
              8: (256, 256, 2, 2, 2, 64), 9: (256, 256, 4, 2, 2, 64), 10: (128, 256, 2, 2, 3, 64),
              11: (128, 128, 4, 1, 2, 64), 12: (128, 256, 4, 2, 5), 13: (64, 256, 2, 2, 5),
              14: (128, 256, 4, 2, 3, 64)}
TILES = {1: (256, 256, 2, 2, 5), 2: (256, 128, 2, 2, 5), 3: (128, 128, 2, 2, 5), 4: (256, 128, 2, 2, 3),
         5: (128, 256, 2, 2, 3), 6: (256, 256, 4, 2, 4), 7: (256, 256, 4, 2, 3),
         8: (256, 256, 2, 2, 2, 64), 9: (256, 256, 4, 2, 2, 64), 10: (256, 128, 2, 2, 3, 64),
         11: (128, 128, 2, 2, 2, 64), 12: (128, 128, 4, 2, 5), 13: (64, 128, 2, 2, 5),
         14: (128, 128, 4, 2, 4, 64)}
TILES_FALLBACK = {8: 3, 9: 3, 10: 3, 11: 3, 14: 12}


def instance_name(r):
    """Kernel instantiation of a trace record, named as the rocprof summaries
    name it (dcvc_amd.profiling.short_name)."""
    kind, epi, amode, dmode, prec, cfg = (int(r[i]) for i in (0, 1, 8, 9, 10, 11))
    if kind == 3:
        return f"dw3x3_kernel<{cfg}>"
    if kind == 4 or prec != 0:
        return f"{KIND_NAMES[kind]}_{EPI_NAMES[epi]}"
    if (amode, dmode) != (0, 0):
        cfg = TILES_FALLBACK.get(cfg, cfg)
    t = (TILES_QUAD if epi == 6 else TILES)[cfg]
    tag = ",k64" if len(t) == 6 else ""
    return "gemm_kernel<" + ",".join(str(v) for v in (*t[:5], epi, amode, dmode, prec)) + tag + ">"


def kernel_instances(recs):
    """Aggregate trace records per kernel instantiation."""
    out = {}
    names = [instance_name(r) for r in recs]
    for n in sorted(set(names)):
        r = recs[np.array([x == n for x in names])]
        t = float(r[:, 7].sum()) * 1e-9
        k0 = r[0]
        out[n] = {"class": KIND_NAMES[int(k0[0])] + ("" if k0[0] == 3 else "_" + EPI_NAMES[int(k0[1])]),
                  "launches": int(len(r)), "time_ms": round(t * 1e3, 4),
                  "tflops": round(float(r[:, 5].sum()) / t / 1e12, 2) if t > 0 else None,
                  "gbs": round(float(r[:, 6].sum()) / t / 1e9, 1) if t > 0 else None,
                  "flops_per_launch": int(r[:, 5].sum() // len(r)), "bytes_per_launch": int(r[:, 6].sum() // len(r)),
                  "avg_launch_ms": round(t * 1e3 / len(r), 5)}
    return dict(sorted(out.items(), key=lambda kv: -kv[1]["time_ms"]))


def roofline_of(instances, classes):
    """Roofline object for the kernel instantiation with the largest total
    time: MFMA-bound when its algorithmic intensity is above the FP16 ridge,
    HBM-bound below; the HBM-bound classes are listed beside it."""
    name, c = next(iter(instances.items()))
    ridge = PEAK_FP16_TFLOPS * 1e12 / (PEAK_HBM_GBS * 1e9)
    intensity = c["flops_per_launch"] / max(c["bytes_per_launch"], 1)
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        if t.get("kernel") == name:
            traffic = t.get("bytes_per_launch")
    common = {"kernel": name, "kernel_class": c["class"], "traffic": traffic, "avg_launch_ms": c["avg_launch_ms"],
              "launches_per_step": c["launches"], "flops_per_launch": c["flops_per_launch"],
              "bytes_per_launch": c["bytes_per_launch"]}
    if intensity >= ridge and c["tflops"] is not None:
        out = {"bound": "mfma", "achieved": c["tflops"], "peak": PEAK_FP16_TFLOPS, "unit": "TFLOP/s",
               "frac": round(c["tflops"] / PEAK_FP16_TFLOPS, 4), **common}
    else:
        out = {"bound": "hbm", "achieved": c["gbs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
               "frac": round(c["gbs"] / PEAK_HBM_GBS, 4) if c["gbs"] else None, **common}
    out["hbm_bound_classes"] = {
        n: {"achieved_gbs": v["gbs"], "frac": round(v["gbs"] / PEAK_HBM_GBS, 4), "time_ms": v["time_ms"]}
        for n, v in classes.items()
        if v["gbs"] and v["flops_per_launch"] / max(v["bytes_per_launch"], 1) < ridge}
    return out


def mfma_ceiling(dev, prec="fp16", blocks=512, iters=2000):
    """Sustained MFMA rate on operands held in registers (dcvc_mfma_calibrate,
    2 waves per SIMD as in the GEMMs), for the operand type of `prec`: random
    N(0, 1) operands, the rate the matrix cores hold on real data, and
    all-zero operands.  TFLOP/s, best of 3 launches."""
    import torch
    from dcvc_amd import _lib
    code, dtype, k = {"fp16": (0, torch.float16, 16), "bf16": (1, torch.bfloat16, 16),
                      "fp8": (2, torch.float8_e4m3fn, 64)}[prec]
    nbytes = blocks * 64 * 1024
    gen = torch.Generator(device="cpu").manual_seed(0)
    esize = torch.tensor([], dtype=dtype).element_size()
    rnd = torch.randn(nbytes // esize, generator=gen).to(dtype).view(torch.uint8).to(dev)
    ops = {"random": rnd, "zero": torch.zeros(nbytes, dtype=torch.uint8, device=dev)}
    sink = torch.empty(blocks * 256, dtype=torch.float32, device=dev)
    st = torch.cuda.current_stream()
    flop = blocks * 4 * 16 * iters * 2.0 * 32 * 32 * k
    out = {}
    for name, t in ops.items():
        _lib.call("dcvc_mfma_calibrate", t.data_ptr(), sink.data_ptr(), blocks, 50, code, st.cuda_stream)
        best = 1e9
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            _lib.call("dcvc_mfma_calibrate", t.data_ptr(), sink.data_ptr(), blocks, iters, code, st.cuda_stream)
            e1.record(st)
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        out[name + "_operands_tflops"] = round(flop / (best * 1e-3) / 1e12, 1)
    return out


def copy_bandwidth(dev, mib=1024):
    """Device-to-device copy of `mib` MiB (torch's copy kernel), read + write
    bytes over the best of 5 event-timed copies: the rate a plain streaming
    kernel reaches, against which the HBM-bound classes can also be read."""
    import torch
    n = mib << 20
    a = torch.empty(n, dtype=torch.uint8, device=dev).fill_(1)
    b = torch.empty_like(a)
    st = torch.cuda.current_stream()
    b.copy_(a)
    best = 1e9
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        b.copy_(a)
        e1.record(st)
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del a, b
    return round(2.0 * n / (best * 1e-3) / 1e9, 1)


def gemm_layers(sd):
    """Every 1x1-conv GEMM layer of a state dict with the epilogue it runs
    with (dc.0: WSiLU, ffn.0: the quad GEMM, the rest: bias; shortcut adds
    are exact fp32 adds after the product)."""
    out = []
    for k, v in sd.items():
        if not k.endswith(".weight") or v.dim() != 4 or v.shape[2] != 1 or v.shape[3] != 1:
            continue
        name = k[:-len(".weight")]
        epi = "wsilu" if name.endswith(".dc.0") else ("quad" if name.endswith(".ffn.0") else "bias")
        out.append((name, epi))
    return out


def per_layer_errors(sd, dev, rows=256, seed=0):
    """Every GEMM layer of the model in FP16 / BF16 / FP8 on the GPU against
    a float64 host product: the layer's own fp16 weights and bias, `rows`
    N(0,1) activation rows.  Per precision: mean and max over the layers of
    the max and mean error relative to the reference's max / mean magnitude."""
    import torch
    from dcvc_amd import kernels
    from oracle import dcvc_oracle as O
    r = np.random.Generator(np.random.PCG64(seed))
    errs = {p: [] for p in ("fp16", "bf16", "fp8")}
    for name, epi in gemm_layers(sd):
        w = sd[name + ".weight"].float()[:, :, 0, 0].numpy().astype(np.float16)
        b = (sd[name + ".bias"].float().numpy() if (name + ".bias") in sd else np.zeros(w.shape[0])).astype(np.float32)
        a = r.standard_normal((rows, w.shape[1])).astype(np.float16)
        ref = O.ref_gemm_epilogue(a, w, b, epi)
        wt = torch.from_numpy(w).to(dev)
        at, bt = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
        for prec in errs:
            wq, sc = (wt, None) if prec == "fp16" else kernels.weights_to_lowp(wt, prec)
            got = kernels.gemm_lowp(at, wq, sc, prec, bt, epi=epi).float().cpu().numpy().astype(np.float64)
            err = np.abs(got - ref)
            errs[prec].append((name, float(err.max() / np.abs(ref).max()), float(err.mean() / np.abs(ref).mean())))
    out = {}
    for prec, e in errs.items():
        worst = max(e, key=lambda t: t[1])
        out[prec] = {"layers": len(e), "mean_of_mean_rel": float(np.mean([t[2] for t in e])),
                     "max_of_mean_rel": float(max(t[2] for t in e)), "mean_of_max_rel": float(np.mean([t[1] for t in e])),
                     "max_of_max_rel": worst[1], "worst_layer": worst[0]}
    return out


def precision_forms(sdp, p_enc_ref, x_hat_i, chunks, H, W, pad_b, pad_r, skip, dev, steps=32):
    """BF16 / FP8 1x1-conv products against the FP16 path.

    * frames/s at qp 32 on the main stream's weights, at equal coded load:
      every precision codes the same 4 chunks from the intra anchor, again
      and again (the anchor is re-applied every 4 chunks, inside the timing).
      With these random-init weights a continuing FP16 / BF16 stream falls
      into a state where every y symbol is skipped after about 8 chunks,
      while an FP8 stream keeps coding ~1 M per chunk, so a continuing
      stream would compare host-coder loads, not the products
      (`stream_y_symbols`: per-chunk counts of a continuing stream);
    * FP8 with e4m3 activations stored between products (the default) and
      with fp16 activations converted inside each product (same bits);
    * per qp in {0, 21, 42, 63}, with per-QP scale ladders in the weights
      (weights.qp_ladder; the reference's init has all-ones rows, which would
      make the qp rows identical), the 8-chunk stream coded in each
      precision: PSNR of the 4 real frames of each chunk, bpp, symbol match
      against FP16;
    * the first chunk's symbols (qp 32) against the FP32 CPU oracle, beside
      the fp16-storage floor (the oracle with fp16 storage against FP32);
    * per-layer error against float64 for every GEMM layer of the model."""
    import torch
    from dcvc_amd import codec, entropy, kernels, weights
    from oracle import dcvc_oracle as O
    out = {"fps": {}, "fps_loop_y_symbols": {}, "fps_loop_bytes_per_chunk": {}, "traced_kernel_ms": {},
           "stream_y_symbols": {}, "per_qp": {}}
    n_chunks = len(chunks)
    P16 = ((H + 15) // 16) * ((W + 15) // 16)

    def make(prec, ref=p_enc_ref):
        e = codec.VideoCodecHTS(share_from=ref)
        d = codec.VideoCodecHTS(share_from=ref)
        e.proxy.set_precision(prec)
        d.proxy.set_precision(prec)
        e.add_ref_feature_from_frame(x_hat_i)
        d.add_ref_feature_from_frame(x_hat_i, apply_feature_adaptor=False)
        return e, d

    def equal_load(prec):
        e, d = make(prec)
        nsym, nbytes = [], []

        def run(k):
            for i in range(k):
                if i % 4 == 0:
                    e.add_ref_feature_from_frame(x_hat_i)
                    d.add_ref_feature_from_frame(x_hat_i, apply_feature_adaptor=False)
                bits, ec = e.proxy.compress(chunks[i % 4], 32, False, pad_b, pad_r)
                d.proxy.decompress(bits, 32, H, W, ec, False)
                nsym.append(e.proxy.last_symbol_count())
                nbytes.append(len(bits))
        run(4)
        torch.cuda.synchronize()
        nsym.clear()
        nbytes.clear()
        t0 = time.perf_counter()
        run(steps)
        torch.cuda.synchronize()
        fps = round(4.0 * steps / (time.perf_counter() - t0), 2)
        return e, d, fps, int(np.mean(nsym)), int(np.mean(nbytes))

    for prec in ("fp16", "bf16", "fp8"):
        e, d, fps, ns, nb = equal_load(prec)
        out["fps"][prec] = fps
        out["fps_loop_y_symbols"][prec] = ns
        out["fps_loop_bytes_per_chunk"][prec] = nb
        rec_e, rec_d = layer_trace(e, d, chunks[0], 32, pad_b, pad_r, H, W)
        recs = np.concatenate([rec_e, rec_d])
        out["traced_kernel_ms"][prec] = round(float(recs[:, 7].sum()) * 1e-6, 3)
        # quad GEMMs (ffn.0): algorithmic FLOPs over event-timed duration,
        # against the dense MFMA peak of the operand type
        quad = (recs[:, 0] == 0) & (recs[:, 1] == 6)
        if quad.any():
            tf = float(recs[quad, 5].sum()) / float(recs[quad, 7].sum()) * 1e-3  # FLOP/ns -> TFLOP/s
            peak = {"fp16": 2500.0, "bf16": 2500.0, "fp8": 5000.0}[prec]
            sus = mfma_ceiling(dev, prec)["random_operands_tflops"]
            out.setdefault("quad_gemm", {})[prec] = {
                "launches": int(quad.sum()), "achieved_tflops": round(tf, 1), "peak_tflops": peak,
                "frac": round(tf / peak, 4), "sustained_tflops": sus, "frac_of_sustained": round(tf / sus, 4),
                "operand_prec_codes": sorted({int(v) for v in recs[quad, 10]})}
        del e, d
        # a continuing stream from the anchor: per-chunk coded y symbols
        e, d = make(prec)
        counts = []
        for i in range(12):
            bits, ec = e.proxy.compress(chunks[i % n_chunks], 32, False, pad_b, pad_r)
            d.proxy.decompress(bits, 32, H, W, ec, False)
            counts.append(e.proxy.last_symbol_count())
        out["stream_y_symbols"][prec] = counts
        del e, d
    kernels.set_fp8_activations(False)
    try:
        e, d, fps, ns, _ = equal_load("fp8")
        out["fps_fp8_in_loop_conversion"] = fps
        del e, d
    finally:
        kernels.set_fp8_activations(True)
    out["fps_vs_fp16"] = {p: round(out["fps"][p] / out["fps"]["fp16"], 3) for p in ("bf16", "fp8")}
    # qp sweep with per-QP ladders
    sdl = entropy.add_cdf_to_state_dict(weights.to_torch(weights.round_to_half(
        weights.qp_ladder(weights.init_state_dict("hts", 1)))))
    ref_l = codec.VideoCodecHTS(sdl, skip)
    for qp in (0, 21, 42, 63):
        res = {}
        ref_syms = None
        for prec in ("fp16", "bf16", "fp8"):
            e, d = make(prec, ref_l)
            nbytes, psnrs, syms, first = 0, [], [], None
            for c in range(n_chunks):
                bits, ec = e.proxy.compress(chunks[c], qp, False, pad_b, pad_r)
                syms.append(e.proxy.debug_copy(5, P16 * 256 * 2, np.int16))
                rec = d.proxy.decompress(bits, qp, H, W, ec, False)
                torch.cuda.synchronize()
                nbytes += len(bits)
                src = chunks[c].float()[:, :, :H, :W]
                for f in range(4):
                    x = rec[f].float()[:, :, :H, :W]
                    mse = float(((x - src[:, 3 * f:3 * f + 3]) ** 2).mean())
                    psnrs.append(10 * np.log10(1.0 / max(mse, 1e-12)))
                if c == 0:
                    first = (len(bits), float(np.mean(psnrs)))
            if prec == "fp16":
                ref_syms = syms
            match = float(np.mean([np.mean(a == b) for a, b in zip(syms, ref_syms)]))
            res[prec] = {"bpp": round(8.0 * nbytes / (4 * n_chunks * H * W), 5),
                         "psnr_db": round(float(np.mean(psnrs)), 4), "symbol_match_vs_fp16": round(match, 5),
                         "first_chunk_bpp": round(8.0 * first[0] / (4 * H * W), 5),
                         "first_chunk_psnr_db": round(first[1], 4)}
            del e, d
        for prec in ("bf16", "fp8"):
            res[prec]["psnr_delta_db"] = round(res[prec]["psnr_db"] - res["fp16"]["psnr_db"], 4)
            res[prec]["bpp_delta_pct"] = round(100 * (res[prec]["bpp"] / res["fp16"]["bpp"] - 1), 3)
            # one chunk from the common anchor: the products' own error, before
            # the stream's temporal state carries it into the next chunks
            res[prec]["first_chunk_psnr_delta_db"] = round(res[prec]["first_chunk_psnr_db"]
                                                           - res["fp16"]["first_chunk_psnr_db"], 4)
            res[prec]["first_chunk_bpp_delta_pct"] = round(100 * (res[prec]["first_chunk_bpp"]
                                                                  / res["fp16"]["first_chunk_bpp"] - 1), 3)
        out["per_qp"][str(qp)] = res
    out["qp_weights"] = "per-QP ladders (weights.qp_ladder: encoder side 0.5 -> 2 over qp, decoder side inverse)"
    # first chunk of the stream at qp 32 against the FP32 oracle (skip 0.15)
    sdo = {k: v.float() for k, v in sdp.items()}  # the headline codec's weights
    ref = x_hat_i.float().cpu().contiguous()
    xin = chunks[0].float().cpu().contiguous(memory_format=torch.contiguous_format)
    syms = {}
    for prec in ("fp16", "bf16", "fp8"):
        e, _ = make(prec)
        e.proxy.compress(chunks[0], 32, False, pad_b, pad_r)
        syms[prec] = e.proxy.debug_copy(5, P16 * 256 * 2, np.int16).astype(np.int64)
        del e
    oracle = {}
    with torch.inference_mode():
        for mode, kw in (("fp32", {}), ("fp16_storage", {"fp16": True})):
            with O.storage_mode(**kw):
                st = O.HTSState()
                O.hts_add_ref_feature_from_frame(sdo, st, ref, True)
                oracle[mode] = O.hts_compress(sdo, st, xin, 32, False, skip_thres=skip).sym16_dense.astype(np.int64)
    agree = lambda a, b: [round(float(np.mean((a >> 8) == (b >> 8))), 5), round(float(np.mean((a & 255) == (b & 255))), 5)]
    out["vs_fp32_oracle"] = {p: {"symbols_indexes": agree(v, oracle["fp32"])} for p, v in syms.items()}
    out["vs_fp32_oracle"]["fp16_storage_floor"] = {"symbols_indexes": agree(oracle["fp16_storage"], oracle["fp32"])}
    out["per_layer_vs_fp64"] = per_layer_errors(sdo, dev)
    return out


def htl_form(x_hat_e, x_hat_d, chunks, H, W, pad_b, pad_r, skip, qp, steps=40, warmup=3):
    """The HT-L chunk model (the second DCVC-UF structure: ratio-1 blocks,
    four entropy round trips per decode) on the same stream: frames/s,
    p50 / p99 step latency, and the GEMM work and top kernel classes of one
    encode + decode."""
    import torch
    from dcvc_amd import codec
    _, sdl = codec.make_state_dicts(seed=0, video="htl")
    e = codec.VideoCodecHTL(sdl, skip)
    d = codec.VideoCodecHTL(share_from=e)
    e.add_ref_feature_from_frame(x_hat_e)
    d.add_ref_feature_from_frame(x_hat_d, apply_feature_adaptor=False)
    n = len(chunks)
    lat, nbytes = [], []
    for i in range(warmup + steps):
        t0 = time.perf_counter()
        bits, ec = e.proxy.compress(chunks[i % n], qp, False, pad_b, pad_r)
        d.proxy.decompress(bits, qp, H, W, ec, False)
        torch.cuda.synchronize()
        if i >= warmup:
            lat.append(time.perf_counter() - t0)
            nbytes.append(len(bits))
    total = float(np.sum(lat))
    out = {"frames_per_s": round(4.0 * steps / total, 3), "ms_per_step": round(1000 * total / steps, 3),
           "p50_ms": round(1000 * float(np.percentile(lat, 50)), 3),
           "p99_ms": round(1000 * float(np.percentile(lat, 99)), 3), "steps": steps,
           "bpp": round(8 * float(np.mean(nbytes)) / (4 * H * W), 5), "precision": "fp16"}
    rec_e, rec_d = layer_trace(e, d, chunks[0], qp, pad_b, pad_r, H, W)
    out["encode_gemm_tflop"] = round(float(rec_e[rec_e[:, 0] <= 2][:, 5].sum()) / 1e12, 4)
    out["decode_gemm_tflop"] = round(float(rec_d[rec_d[:, 0] <= 2][:, 5].sum()) / 1e12, 4)
    classes = kernel_classes(np.concatenate([rec_e, rec_d]))
    out["top_kernel_classes"] = {k: {f: v[f] for f in ("launches", "time_ms", "tflops", "gbs")}
                                 for k, v in list(classes.items())[:4]}
    return out


def tuning_ab(sdp, x_hat_e, x_hat_d, chunks, H, W, pad_b, pad_r, skip, qp, steps=40, warmup=4, rounds=2):
    """The same HT-S step with the autotuner's tile database and with the
    built-in heuristics, on fresh codec pairs (graphs are captured with the
    table in force at their first use), alternated `rounds` times."""
    import torch
    from dcvc_amd import codec, tuning
    if not tuning.DB_PATH.exists():
        return {"database": None}
    with open(tuning.DB_PATH) as f:
        db = json.load(f)
    res = {"tuned": [], "heuristic": []}
    n = len(chunks)
    try:
        for _ in range(rounds):
            for mode in ("heuristic", "tuned"):
                if mode == "tuned":
                    tuning.install(db)
                else:
                    tuning.clear()
                e = codec.VideoCodecHTS(sdp, skip)
                d = codec.VideoCodecHTS(share_from=e)
                e.add_ref_feature_from_frame(x_hat_e)
                d.add_ref_feature_from_frame(x_hat_d, apply_feature_adaptor=False)
                for i in range(warmup):
                    bits, ec = e.proxy.compress(chunks[i % n], qp, False, pad_b, pad_r)
                    d.proxy.decompress(bits, qp, H, W, ec, False)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for i in range(steps):
                    bits, ec = e.proxy.compress(chunks[i % n], qp, False, pad_b, pad_r)
                    d.proxy.decompress(bits, qp, H, W, ec, False)
                torch.cuda.synchronize()
                res[mode].append(round(1000 * (time.perf_counter() - t0) / steps, 3))
                del e, d
    finally:
        tuning.install(db)
    summ = tuning.summary(db)
    return {"database": os.path.relpath(str(tuning.DB_PATH), ROOT), "gemm_shapes": summ["gemm"]["shapes"],
            "gemm_changed": summ["gemm"]["changed"], "d3x3_shapes": summ["d3x3"]["shapes"],
            "d3x3_changed": summ["d3x3"]["changed"], "ms_per_step_tuned": res["tuned"],
            "ms_per_step_heuristic": res["heuristic"],
            "speedup": round(float(np.mean(res["heuristic"])) / float(np.mean(res["tuned"])), 4)}


def skip0_form(sdp, x_hat_e, x_hat_d, chunks, H, W, pad_b, pad_r, qp, steps=40, warmup=3):
    """HT-S on the same stream with skip_thres 0, the reference's default:
    every y symbol is coded, so the encoder hands all of y to the CPU coder
    (8 sub-streams) and the decoder waits on its y decode."""
    import torch
    from dcvc_amd import codec
    e = codec.VideoCodecHTS(sdp, 0.0)
    d = codec.VideoCodecHTS(share_from=e)
    e.add_ref_feature_from_frame(x_hat_e)
    d.add_ref_feature_from_frame(x_hat_d, apply_feature_adaptor=False)
    n = len(chunks)
    lat, nbytes, nsym, ecs = [], [], [], []
    for i in range(warmup + steps):
        t0 = time.perf_counter()
        bits, ec = e.proxy.compress(chunks[i % n], qp, False, pad_b, pad_r)
        d.proxy.decompress(bits, qp, H, W, ec, False)
        torch.cuda.synchronize()
        if i >= warmup:
            lat.append(time.perf_counter() - t0)
            nbytes.append(len(bits))
            nsym.append(e.proxy.last_symbol_count())
            ecs.append(ec)
    total = float(np.sum(lat))
    return {"frames_per_s": round(4.0 * steps / total, 3), "ms_per_step": round(1000 * total / steps, 3),
            "p50_ms": round(1000 * float(np.percentile(lat, 50)), 3),
            "p99_ms": round(1000 * float(np.percentile(lat, 99)), 3), "steps": steps,
            "y_symbols_coded": int(np.mean(nsym)), "ec_parallel": int(np.median(ecs)),
            "bytes_per_chunk": int(np.mean(nbytes)), "bpp": round(8 * float(np.mean(nbytes)) / (4 * H * W), 5),
            "skip_thres": 0.0, "precision": "fp16"}


def ld_form(x_hat_e, x_hat_d, chunks, H, W, pad_b, pad_r, skip, qp, steps=60, warmup=4):
    """The low-delay model (one frame per call, one entropy round trip) on
    the same stream's frames: frames/s, p50 / p99 per-frame latency, and the
    GEMM work of one encode + decode."""
    import torch
    from dcvc_amd import codec
    _, sdl = codec.make_state_dicts(seed=0, video="ld")
    e = codec.VideoCodecLD(sdl, skip)
    d = codec.VideoCodecLD(share_from=e)
    e.add_ref_feature_from_frame(x_hat_e)
    d.add_ref_feature_from_frame(x_hat_d, apply_feature_adaptor=False)
    frames = [c[:, 3 * f:3 * f + 3].contiguous(memory_format=torch.channels_last) for c in chunks for f in range(4)]
    lat, nbytes = [], []
    for i in range(warmup + steps):
        t0 = time.perf_counter()
        bits, ec = e.proxy.compress(frames[i % len(frames)], qp, False, pad_b, pad_r)
        d.proxy.decompress(bits, qp, H, W, ec, False)
        torch.cuda.synchronize()
        if i >= warmup:
            lat.append(time.perf_counter() - t0)
            nbytes.append(len(bits))
    total = float(np.sum(lat))
    out = {"frames_per_s": round(steps / total, 3), "ms_per_frame": round(1000 * total / steps, 3),
           "p50_ms": round(1000 * float(np.percentile(lat, 50)), 3),
           "p99_ms": round(1000 * float(np.percentile(lat, 99)), 3), "steps": steps,
           "bpp": round(8 * float(np.mean(nbytes)) / (H * W), 5), "precision": "fp16"}
    rec_e, rec_d = layer_trace(e, d, frames[0], qp, pad_b, pad_r, H, W)
    out["encode_gemm_tflop"] = round(float(rec_e[rec_e[:, 0] <= 2][:, 5].sum()) / 1e12, 4)
    out["decode_gemm_tflop"] = round(float(rec_d[rec_d[:, 0] <= 2][:, 5].sum()) / 1e12, 4)
    out["traced_kernel_ms"] = round(float(np.concatenate([rec_e, rec_d])[:, 7].sum()) * 1e-6, 3)
    return out


def batching_form(sdi, sdp, rounds=2):
    """16 concurrent streams of uneven resolution (4x 854x480, 6x 1280x720,
    4x 1920x1080, 2x 3840x2160), one 4-frame chunk each per round at a random
    qp.  The plain path (FP16, one stream at a time) against all streams at
    once in FP16, in BF16 (the batching form's precision) and in FP8.  Also the
    bound batching could reach: every stream at the per-pixel rate of the
    best-utilised resolution in the plain path."""
    import torch
    from dcvc_amd import batching
    n = len(batching.MIX_16)
    frames = 4 * n * rounds
    out = {"streams": "4x 854x480, 6x 1280x720, 4x 1920x1080, 2x 3840x2160", "rounds": rounds,
           "frames_counted": frames}
    setup, _ = batching.build_streams(sdi, sdp, batching.MIX_16, n_chunks=rounds + 1, seed=0)
    batching.code_concurrently(setup.contexts, [c[0] for c in setup.chunks], setup.qps)  # graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per_res = {}
    for r in range(rounds):
        for res in batching.code_sequentially(setup.contexts, [c[1 + r] for c in setup.chunks], setup.qps):
            ctx = setup.contexts[res.stream]
            per_res.setdefault((ctx.H, ctx.W), []).append(res.encode_s + res.decode_s)
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    for r in range(rounds):
        batching.code_concurrently(setup.contexts, [c[1 + r] for c in setup.chunks], setup.qps)
    t_con16 = time.perf_counter() - t0
    del setup
    setup2, _ = batching.build_streams(sdi, sdp, batching.MIX_16, n_chunks=rounds + 1, seed=0, precision="bf16")
    batching.code_concurrently(setup2.contexts, [c[0] for c in setup2.chunks], setup2.qps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(rounds):
        batching.code_concurrently(setup2.contexts, [c[1 + r] for c in setup2.chunks], setup2.qps)
    t_con = time.perf_counter() - t0
    del setup2
    # FP8 products on e4m3 activations (the fastest precision of this build)
    setup3, _ = batching.build_streams(sdi, sdp, batching.MIX_16, n_chunks=rounds + 1, seed=0, precision="fp8")
    batching.code_concurrently(setup3.contexts, [c[0] for c in setup3.chunks], setup3.qps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(rounds):
        batching.code_concurrently(setup3.contexts, [c[1 + r] for c in setup3.chunks], setup3.qps)
    t_con8 = time.perf_counter() - t0
    del setup3

    # one batch proxy pair per resolution (one launch per layer over the
    # group's streams), the four groups on concurrent host threads
    def time_groups(precision):
        gs, _ = batching.build_groups(sdi, sdp, batching.MIX_16, n_chunks=rounds + 1, seed=0, precision=precision)
        batching.code_groups_concurrently(gs.groups, [ch[0] for ch in gs.chunks], gs.qps)  # graph capture
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in range(rounds):
            batching.code_groups_concurrently(gs.groups, [ch[1 + r] for ch in gs.chunks], gs.qps)
        dt = time.perf_counter() - t0
        del gs
        return dt
    t_grp = {p: time_groups(p) for p in ("fp16", "bf16", "fp8")}

    # 8 streams of 1920x1080 (the headline's size) as two batches of 4 on
    # two host threads: aggregate 1080p frames/s against the one-stream path
    def time_1080p(precision, n=8, batch=4):
        gs, _ = batching.build_groups(sdi, sdp, [(1080, 1920)] * n, n_chunks=rounds + 1, seed=1,
                                      precision=precision, max_batch=batch)
        batching.code_groups_concurrently(gs.groups, [ch[0] for ch in gs.chunks], gs.qps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in range(rounds):
            batching.code_groups_concurrently(gs.groups, [ch[1 + r] for ch in gs.chunks], gs.qps)
        dt = time.perf_counter() - t0
        del gs
        return round(4 * n * rounds / dt, 2)
    out["aggregate_1080p"] = {"streams": "8 x 1920x1080, two batches of 4 on two host threads, random qp",
                              "frames_per_s": {p: time_1080p(p) for p in ("fp16", "fp8")}}
    # per-pixel cost of each resolution alone (plain path), and the bound
    px = {k: ((k[0] + 15) // 16 * 16) * ((k[1] + 15) // 16 * 16) for k in per_res}
    us_per_mpx = {f"{k[1]}x{k[0]}": round(1e6 * float(np.mean(v)) / (px[k] / 1e6), 1) for k, v in per_res.items()}
    best = min(float(np.mean(v)) / px[k] for k, v in per_res.items())
    t_bound = rounds * sum(best * ((H + 15) // 16 * 16) * ((W + 15) // 16 * 16) for H, W in batching.MIX_16)
    out.update({"plain_fp16_sequential_frames_per_s": round(frames / t_seq, 2),
                "concurrent_fp16_frames_per_s": round(frames / t_con16, 2),
                "concurrent_bf16_frames_per_s": round(frames / t_con, 2),
                "concurrent_fp8_frames_per_s": round(frames / t_con8, 2),
                "speedup_bf16_vs_plain": round(t_seq / t_con, 3), "speedup_fp16_vs_plain": round(t_seq / t_con16, 3),
                "speedup_fp8_vs_plain": round(t_seq / t_con8, 3),
                "grouped_frames_per_s": {p: round(frames / t, 2) for p, t in t_grp.items()},
                "grouped_speedup_vs_plain": {p: round(t_seq / t, 3) for p, t in t_grp.items()},
                "groups": "one batch proxy pair per resolution: 4 x 854x480, 6 x 1280x720, 4 x 1920x1080, "
                          "2 x 3840x2160; every layer one launch over the group",
                "target_speedup": 3.0,
                "plain_us_per_megapixel_step": us_per_mpx,
                "bound_frames_per_s_at_best_per_pixel_rate": round(frames / t_bound, 2),
                "bound_speedup_vs_plain": round(t_seq / t_bound, 3), "precision": "bf16 (fp16 plain baseline)"})
    return out


def serving_form(sdi, sdp, duration=60.0, rate=120.0, n_streams=64, slots=8, sweep=(200.0, 280.0, 360.0),
                 sweep_seconds=10.0, p99_bound_ms=100.0, sweep_fp8=(280.0, 360.0, 440.0)):
    """Poisson arrivals (encode or decode of one 4-frame chunk of one of 64
    persistent streams, resolutions 480p 30 %, 720p 40 %, 1080p 25 %,
    2160p 5 %) served by the continuous-batching loop, HBM budget 200 GB,
    every stream's contexts allocated and graphs captured before the trace
    (warm start).  Then the same load on the plain path (one request at a
    time), and the capacity: the arrival rates of `sweep` for
    `sweep_seconds` each, the highest one whose p99 stays under
    `p99_bound_ms` and whose requests are all served; then the load and a
    capacity sweep with FP8 products."""
    from dcvc_amd import serving
    res = serving.resolution_mix(n_streams, seed=0)
    ex = serving.GpuExecutor(sdi, sdp, res, chunks_per_stream=4, seed=0, keep_results=False)
    t0 = time.perf_counter()
    ex.prewarm()
    warm_s = time.perf_counter() - t0

    def serve(rate_, dur, slots_, seed):
        reqs = serving.poisson_trace(rate_, dur, n_streams, seed=seed)
        srv = serving.Server(ex, slots=slots_, hbm_budget=200 << 30)
        srv.run(reqs, realtime=True)
        st = serving.Server.stats(reqs)
        st.update({"rate_per_s": rate_, "duration_s": dur, "slots": slots_, "max_running": srv.max_running,
                   "resident_gb": round(srv.max_resident_bytes / 2 ** 30, 2)})
        return st

    out = serve(rate, duration, slots, 0)
    out.update({"streams": n_streams, "resolution_mix": serving.MIX_64, "prewarm_s": round(warm_s, 2)})
    out["plain_one_at_a_time"] = serve(rate, min(duration, 20.0), 1, 1)
    cap = []
    for r in sweep:
        st = serve(r, sweep_seconds, slots, 2)
        cap.append({k: st[k] for k in ("rate_per_s", "served", "requests", "p50_ms", "p99_ms", "frames_per_s")})
    ok = [c["rate_per_s"] for c in [{"rate_per_s": rate, "p99_ms": out["p99_ms"], "served": out["served"],
                                     "requests": out["requests"]}] + cap
          if c["p99_ms"] is not None and c["p99_ms"] <= p99_bound_ms and c["served"] == c["requests"]]
    out["capacity"] = {"p99_bound_ms": p99_bound_ms, "sweep": cap,
                       "max_rate_within_bound_per_s": max(ok) if ok else None}
    # the same streams with FP8 products (e4m3 activations): the load at
    # `rate` and the capacity sweep, on a fresh executor
    del ex
    ex = serving.GpuExecutor(sdi, sdp, res, chunks_per_stream=4, seed=0, keep_results=False, precision="fp8")
    ex.prewarm()
    f8 = serve(rate, sweep_seconds, slots, 3)
    cap8 = []
    for r in sweep_fp8:
        st = serve(r, sweep_seconds, slots, 4)
        cap8.append({k: st[k] for k in ("rate_per_s", "served", "requests", "p50_ms", "p99_ms", "frames_per_s")})
    ok8 = [c["rate_per_s"] for c in [f8] + cap8
           if c["p99_ms"] is not None and c["p99_ms"] <= p99_bound_ms and c["served"] == c["requests"]]
    out["fp8"] = {"at_rate": {k: f8[k] for k in ("rate_per_s", "duration_s", "served", "requests", "p50_ms", "p99_ms",
                                                 "frames_per_s")},
                  "capacity": {"sweep": cap8, "max_rate_within_bound_per_s": max(ok8) if ok8 else None}}
    return out


def cpu_baseline():
    """FP32 CPU oracle on this host: HT-S encode (with its context update) and
    decode of one full chunk of the same synthetic 1920x1080 stream (4 real
    frames padded to 8, qp 32, skip 0.15), timed once each; plus the 256x256
    intra frame at all threads and at 1 thread, and the C++ rANS coder in
    symbols/s."""
    import torch
    from dcvc_amd import entropy, synthetic, weights
    from oracle import dcvc_oracle as O
    threads = torch.get_num_threads()
    H, W = 1080, 1920
    sdp = {k: v.float() for k, v in weights.to_torch(weights.round_to_half(weights.init_state_dict("hts", 1))).items()}
    frames = synthetic.motion_sequence_yuv420(H, W, 5, seed=0)
    ref = O.pad_to(torch.from_numpy(synthetic.frame_to_input(frames[0]))[None], 16)
    x = torch.from_numpy(synthetic.chunk_to_input(frames[1:5], 8))[None]
    del frames
    st_e, st_d = O.HTSState(), O.HTSState()
    with torch.inference_mode():
        O.hts_add_ref_feature_from_frame(sdp, st_e, ref, True)
        O.hts_add_ref_feature_from_frame(sdp, st_d, ref, False)
        t0 = time.perf_counter()
        r = O.hts_compress(sdp, st_e, x, 32, False, skip_thres=0.15)
        t_enc = time.perf_counter() - t0
        sym = r.sym16_dense
        t0 = time.perf_counter()
        O.hts_decompress(sdp, st_d, r.z_int8, lambda idx: (sym[r.cond_dense] >> 8).astype(np.int8), 32, H, W,
                         False, skip_thres=0.15)
        t_dec = time.perf_counter() - t0
    out = {"value": round(4.0 / (t_enc + t_dec), 4), "unit": "frames/s", "cores": threads, "kind": "port",
           "sample": f"FP32 torch oracle (the port of the reference's algorithm), one full HT-S chunk of the same "
                     f"1920x1080 stream (4 frames padded to 8, padded to 1088x1920), encode {t_enc:.2f} s + "
                     f"decode {t_dec:.2f} s on {threads} threads; 4 frames counted",
           "encode_s": round(t_enc, 3), "decode_s": round(t_dec, 3)}
    # 256x256 intra frame (encode incl. reconstruction), all threads and 1 thread
    sdi = {k: v.float() for k, v in weights.to_torch(weights.round_to_half(weights.init_state_dict("dmci", 0))).items()}
    xi = torch.from_numpy(synthetic.frame_to_input(synthetic.texture_yuv420(256, 256, seed=0)))[None]
    intra = {}
    with torch.inference_mode():
        for n_thr in (threads, 1):
            torch.set_num_threads(n_thr)
            O.dmci_compress(sdi, xi, 32)
            t0 = time.perf_counter()
            O.dmci_compress(sdi, xi, 32)
            intra[f"threads_{n_thr}"] = round(1.0 / (time.perf_counter() - t0), 3)
    torch.set_num_threads(threads)
    out["intra_256x256_frames_per_s"] = intra
    # C++ rANS coder: symbols/s at ec_parallel 1 and 8 on the Gaussian tables
    g_cdf, g_len = entropy.gaussian_cdf_table()
    rng = np.random.Generator(np.random.PCG64(0))
    n = 2_000_000
    idx = rng.integers(0, 128, n)
    sym = np.clip(np.round(rng.standard_normal(n) * np.exp(np.linspace(np.log(0.11), np.log(16), 128))[idx] / 4),
                  -8, 8).astype(np.int64)
    sym16 = ((sym << 8) + idx).astype(np.int16)
    rans = {}
    for ec in (1, 8):
        e = entropy.RansEncoder()
        e.set_cdf(g_cdf, g_len, 1)
        e.set_entropy_coder_parallel(ec)
        e.reset()
        t0 = time.perf_counter()
        e.encode_y(sym16)
        e.flush()
        t_enc = time.perf_counter() - t0
        stream = e.get_encoded_stream()
        d = entropy.RansDecoder()
        d.set_cdf(g_cdf, g_len, 1)
        d.set_entropy_coder_parallel(ec)
        d.set_stream(stream)
        t0 = time.perf_counter()
        d.decode_y(idx.astype(np.uint8))
        d.get_decoded_tensor()
        t_dec = time.perf_counter() - t0
        rans[f"ec_parallel_{ec}"] = {"encode_msym_per_s": round(n / t_enc / 1e6, 2),
                                     "decode_msym_per_s": round(n / t_dec / 1e6, 2)}
    out["rans"] = rans
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--qp", type=int, default=32)
    ap.add_argument("--skip-thres", type=float, default=0.15)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--profile-step", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-precision-forms", action="store_true")
    ap.add_argument("--no-tuning-ab", action="store_true", help="skip the tuned-vs-heuristic tile comparison")
    ap.add_argument("--no-batching", action="store_true")
    ap.add_argument("--no-htl", action="store_true", help="skip the HT-L chunk-model section")
    ap.add_argument("--no-ld", action="store_true", help="skip the low-delay model section")
    ap.add_argument("--no-skip0", action="store_true", help="skip the HT-S skip_thres 0 section")
    ap.add_argument("--precision", default="fp16", choices=("fp16", "bf16", "fp8"),
                    help="product precision of the headline stream: fp16 is the measured workload, the others "
                         "are for profiling the precision form (--profile-step)")
    ap.add_argument("--serving-seconds", type=float, default=60.0,
                    help="length of the Poisson serving trace (0 skips it)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (requested {max(args.gpus, world)}); streams shard by stream id "
              f"across GPUs with no collectives, so multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    import torch
    from dcvc_amd import codec, synthetic
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    H, W = args.height, args.width
    pad_r, pad_b = synthetic.get_padding_size(H, W, 16)
    log("building models")
    sdi, sdp = codec.make_state_dicts(seed=0)
    i_enc = codec.IntraCodec(sdi, args.skip_thres)
    i_dec = codec.IntraCodec(sdi, args.skip_thres)
    p_enc = codec.VideoCodecHTS(sdp, args.skip_thres)
    p_dec = codec.VideoCodecHTS(share_from=p_enc)
    if args.precision != "fp16":
        p_enc.proxy.set_precision(args.precision)
        p_dec.proxy.set_precision(args.precision)
    log("synthesising frames")
    n_chunks = 8
    frames = synthetic.motion_sequence_yuv420(H, W, 1 + 4 * n_chunks, seed=0)
    x0 = codec.to_input(synthetic.frame_to_input(frames[0]), dev)
    chunks = [codec.to_input(synthetic.chunk_to_input(frames[1 + 4 * c:5 + 4 * c], 8), dev) for c in range(n_chunks)]
    del frames
    enc_i = i_enc.compress(x0, args.qp, pad_b, pad_r)
    p_enc.add_ref_feature_from_frame(enc_i["x_hat"])
    dec_i = i_dec.decompress(enc_i["bit_stream"], {"height": H, "width": W}, args.qp, enc_i["ec_parallel"])
    p_dec.add_ref_feature_from_frame(dec_i["x_hat"], apply_feature_adaptor=False)
    torch.cuda.synchronize()

    counter = [0]

    def step():
        xc = chunks[counter[0] % n_chunks]
        counter[0] += 1
        t0 = time.perf_counter()
        bits, ec = p_enc.proxy.compress(xc, args.qp, False, pad_b, pad_r)
        t1 = time.perf_counter()
        p_dec.proxy.decompress(bits, args.qp, H, W, ec, False)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, len(bits)

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    if args.profile_step:
        torch.cuda.synchronize()
        torch.cuda._sleep(1000)  # window marker
        for _ in range(args.profile_steps):
            step()
        torch.cuda._sleep(1000)
        torch.cuda.synchronize()
        a = torch.empty(64 << 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)  # 256 MiB read + 256 MiB write: counter calibration
        torch.cuda.synchronize()
        log("profile step done")
        return
    log(f"timing {args.steps} steps")
    lat, enc_t, dec_t, nbytes = [], [], [], []
    t_start = time.perf_counter()
    for i in range(args.steps):
        te, td, nb = step()
        enc_t.append(te)
        dec_t.append(td)
        lat.append(te + td)
        nbytes.append(nb)
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start
    ms_per_step = 1000.0 * total / args.steps
    fps = 4.0 * args.steps / total
    log(f"{fps:.2f} frames/s ({ms_per_step:.2f} ms/step)")

    out = {
        "metric": METRIC, "value": round(fps, 3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp16" if args.precision == "fp16" else f"{args.precision} products, fp16 storage",
        "data": "synthetic",
        "config": {"workload": "DCVC-UF HT-S chunk encode+decode, one stream, 4 frames 1920x1080 YUV420 "
                               "padded to the native 8-frame chunk, qp 32", "height": H, "width": W,
                   "chunks_in_stream": n_chunks,
                   "frames_counted_per_step": 4, "frames_coded_per_step": 8, "qp": args.qp,
                   "skip_thres": args.skip_thres, "precision": "fp16 storage, fp32 accumulate (plain path)",
                   "weights": "seeded random init"},
        "fps_native_8frame": round(8.0 * args.steps / total, 3),
        "p50_ms": round(1000 * float(np.percentile(lat, 50)), 3),
        "p99_ms": round(1000 * float(np.percentile(lat, 99)), 3),
        "encode_ms_mean": round(1000 * float(np.mean(enc_t)), 3),
        "decode_ms_mean": round(1000 * float(np.mean(dec_t)), 3),
        "bytes_per_chunk": int(np.mean(nbytes)),
        "bpp": round(8 * float(np.mean(nbytes)) / (4 * H * W), 5),
    }

    if not args.no_roofline:
        rec_e, rec_d = layer_trace(p_enc, p_dec, chunks[counter[0] % n_chunks], args.qp, pad_b, pad_r, H, W)
        recs = np.concatenate([rec_e, rec_d])
        gemm = recs[recs[:, 0] <= 2]
        dw = recs[recs[:, 0] == 3]
        out["layer_table"] = {
            "encode_gemm_tflop": round(float(rec_e[rec_e[:, 0] <= 2][:, 5].sum()) / 1e12, 4),
            "decode_gemm_tflop": round(float(rec_d[rec_d[:, 0] <= 2][:, 5].sum()) / 1e12, 4),
            "depthwise_gb": round(float(dw[:, 6].sum()) / 1e9, 4),
            "gemm_launches": int(len(gemm)), "depthwise_launches": int(len(dw)),
            "traced_kernel_ms": round(float(recs[:, 7].sum()) * 1e-6, 3),
        }
        classes = kernel_classes(recs)
        out["kernel_classes"] = classes
        instances = kernel_instances(recs)
        out["kernel_instances"] = dict(list(instances.items())[:8])
        out["roofline"] = roofline_of(instances, classes)
        if out["roofline"]["bound"] == "mfma":
            # the nominal peak is quoted at minimum switching activity; on
            # real operands the matrix cores sustain less
            mc = mfma_ceiling(dev)
            out["roofline"]["mfma_sustained"] = mc
            out["roofline"]["frac_of_sustained"] = round(out["roofline"]["achieved"] / mc["random_operands_tflops"], 4)
        # a plain 1 GiB device copy: the streaming rate the HBM-bound classes
        # can be read against
        cb = copy_bandwidth(dev)
        out["roofline"]["copy_gbs"] = cb
        for v in out["roofline"]["hbm_bound_classes"].values():
            v["frac_of_copy"] = round(v["achieved_gbs"] / cb, 4)
    if not args.no_tuning_ab:
        log("tuned vs heuristic tiles")
        out["tuning"] = tuning_ab(sdp, enc_i["x_hat"], dec_i["x_hat"], chunks, H, W, pad_b, pad_r, args.skip_thres,
                                  args.qp)
    if not args.no_precision_forms:
        log("precision forms")
        out["precision_forms"] = precision_forms(sdp, p_enc, enc_i["x_hat"], chunks, H, W, pad_b, pad_r,
                                                 args.skip_thres, dev)
    if not args.no_htl:
        log("HT-L form")
        out["htl_form"] = htl_form(enc_i["x_hat"], dec_i["x_hat"], chunks, H, W, pad_b, pad_r, args.skip_thres,
                                   args.qp)
    if not args.no_skip0:
        log("skip_thres 0 form")
        out["skip0_form"] = skip0_form(sdp, enc_i["x_hat"], dec_i["x_hat"], chunks, H, W, pad_b, pad_r, args.qp)
    if not args.no_ld:
        log("LD form")
        out["ld_form"] = ld_form(enc_i["x_hat"], dec_i["x_hat"], chunks, H, W, pad_b, pad_r, args.skip_thres, args.qp)
    if not args.no_batching:
        log("batching form")
        out["batching_form"] = batching_form(sdi, sdp)
    if args.serving_seconds > 0:
        log("serving form")
        out["serving_form"] = serving_form(sdi, sdp, duration=args.serving_seconds)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
