#!/usr/bin/env python
"""1080p DCVC-UF HT-S chunk encode+decode throughput on one MI355X.

Workload: one stream, one chunk of 4 synthetic 1920x1080 YUV420 frames
(texture + integer motion + 10 % occlusion), padded to the model's native
8-frame chunk by repeating the last frame, qp 32, seeded random-init
weights.  A step is one chunk compress (encoder proxy) followed by its
decompress (decoder proxy), including the CPU entropy coder.  `value` counts
the 4 real frames per step.

Also reported: p50/p99 step latency, the native 8-frame rate, the layer-table
work per chunk, the dominant kernel's roofline (timed with HIP events on its
launch stream) and the FP32 CPU oracle timed on this host.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def layer_table(enc, dec, x, qp, pad_b, pad_r, H, W):
    """One encode and one decode with graphs off and the launch trace on;
    returns the (n, 7) record arrays of both sides."""
    import ctypes as C
    import torch
    from dcvc_amd import _lib
    lib = _lib.load()
    buf = np.zeros((50000, 7), np.int64)
    ptr = buf.ctypes.data_as(C.c_void_p)
    enc.proxy.set_use_graphs(False)
    dec.proxy.set_use_graphs(False)
    lib.dcvc_trace_begin()
    bits, ec = enc.proxy.compress(x, qp, False, pad_b, pad_r)
    torch.cuda.synchronize()
    rec_enc = buf[:lib.dcvc_trace_end(ptr, len(buf))].copy()
    lib.dcvc_trace_begin()
    dec.proxy.decompress(bits, qp, H, W, ec, False)
    torch.cuda.synchronize()
    rec_dec = buf[:lib.dcvc_trace_end(ptr, len(buf))].copy()
    enc.proxy.set_use_graphs(True)
    dec.proxy.set_use_graphs(True)
    return rec_enc, rec_dec


def time_quad_kernel(records, dev, reps=5):
    """Replays every quad-GEMM launch of one chunk at its real shape on the
    current stream; returns (avg ms per launch, flops per launch)."""
    import torch
    from dcvc_amd import kernels
    quad = records[(records[:, 0] == 0) & (records[:, 1] == 6)]
    bufs = []
    for M, N, K in quad[:, 2:5]:
        a = torch.randn((int(M), int(K)), device=dev, dtype=torch.float16)
        w = torch.randn((4 * int(N), int(K)), device=dev, dtype=torch.float16) / np.sqrt(K)
        b = torch.zeros(4 * int(N), device=dev)
        o = torch.empty((int(M), int(N)), device=dev, dtype=torch.float16)
        bufs.append((a, w, b, o))
    for a, w, b, o in bufs:
        kernels.conv1x1(a, w, b, epi="quad", out=o)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        for a, w, b, o in bufs:
            kernels.conv1x1(a, w, b, epi="quad", out=o)
    e.record()
    torch.cuda.synchronize()
    n = reps * len(bufs)
    return s.elapsed_time(e) / n, float(quad[:, 5].sum()) / len(quad), len(quad)


def cpu_baseline(budget_s=12.0):
    """FP32 CPU oracle: HT-S encode (+context update) and decode of a cropped
    sample of the same synthetic chunk; frames/s scaled to 1088x1920."""
    import torch
    from dcvc_amd import synthetic, weights
    from oracle import dcvc_oracle as O
    threads = torch.get_num_threads()
    Hs, Ws = 256, 448
    sdp = {k: v.float() for k, v in weights.to_torch(weights.round_to_half(weights.init_state_dict("hts", 1))).items()}
    frames = synthetic.motion_sequence_yuv420(Hs, Ws, 5, seed=0)
    ref = torch.from_numpy(synthetic.frame_to_input(frames[0]))[None]
    x = torch.from_numpy(synthetic.chunk_to_input(frames[1:5]))[None]
    st_e, st_d = O.HTSState(), O.HTSState()
    O.hts_add_ref_feature_from_frame(sdp, st_e, ref, True)
    O.hts_add_ref_feature_from_frame(sdp, st_d, ref, False)
    times = []
    t_all = time.perf_counter()
    with torch.inference_mode():
        while True:
            t0 = time.perf_counter()
            r = O.hts_compress(sdp, st_e, x, 32, False, skip_thres=0.15)
            sym = r.sym16_dense
            O.hts_decompress(sdp, st_d, r.z_int8, lambda idx: (sym[r.cond_dense] >> 8).astype(np.int8), 32, Hs, Ws,
                             False, skip_thres=0.15)
            times.append(time.perf_counter() - t0)
            if time.perf_counter() - t_all > budget_s or len(times) >= 5:
                break
    t = float(np.median(times))
    scale = (1088 * 1920) / (Hs * Ws)
    return {"value": round(4.0 / (t * scale), 4), "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"FP32 torch oracle, HT-S encode+decode of a {Hs}x{Ws} crop of the same 4-frame chunk "
                      f"(padded to 8), {len(times)} run(s), median {t:.2f} s, scaled by area x{scale:.2f} "
                      f"to 1088x1920"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--qp", type=int, default=32)
    ap.add_argument("--skip-thres", type=float, default=0.15)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
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
    log("synthesising frames")
    frames = synthetic.motion_sequence_yuv420(H, W, 5, seed=0)
    x0 = codec.to_input(synthetic.frame_to_input(frames[0]), dev)
    xc = codec.to_input(synthetic.chunk_to_input(frames[1:5], 8), dev)
    enc_i = i_enc.compress(x0, args.qp, pad_b, pad_r)
    p_enc.add_ref_feature_from_frame(enc_i["x_hat"])
    dec_i = i_dec.decompress(enc_i["bit_stream"], {"height": H, "width": W}, args.qp, enc_i["ec_parallel"])
    p_dec.add_ref_feature_from_frame(dec_i["x_hat"], apply_feature_adaptor=False)
    torch.cuda.synchronize()

    def step():
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
        "vs_baseline": None, "dtype": "fp16", "data": "synthetic",
        "config": {"workload": "DCVC-UF HT-S chunk encode+decode, one stream, 4 frames 1920x1080 YUV420 "
                               "padded to the native 8-frame chunk, qp 32", "height": H, "width": W,
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
        rec_e, rec_d = layer_table(p_enc, p_dec, xc, args.qp, pad_b, pad_r, H, W)
        recs = np.concatenate([rec_e, rec_d])
        gemm = recs[recs[:, 0] <= 2]
        dw = recs[recs[:, 0] == 3]
        out["layer_table"] = {
            "encode_gemm_tflop": round(float(rec_e[rec_e[:, 0] <= 2][:, 5].sum()) / 1e12, 4),
            "decode_gemm_tflop": round(float(rec_d[rec_d[:, 0] <= 2][:, 5].sum()) / 1e12, 4),
            "depthwise_gb": round(float(dw[:, 6].sum()) / 1e9, 4),
            "gemm_launches": int(len(gemm)), "depthwise_launches": int(len(dw)),
        }
        avg_ms, flops, n_quad = time_quad_kernel(recs, dev)
        achieved = flops / (avg_ms * 1e-3) / 1e12
        out["roofline"] = {"bound": "mfma", "kernel": "gemm_kernel<EPI_QUAD> (ffn.0 quad GEMM)",
                           "achieved": round(achieved, 2), "peak": PEAK_FP16_TFLOPS, "unit": "TFLOP/s",
                           "frac": round(achieved / PEAK_FP16_TFLOPS, 4), "traffic": None,
                           "avg_launch_ms": round(avg_ms, 5), "launches_per_step": n_quad,
                           "flops_per_launch": int(flops)}
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
