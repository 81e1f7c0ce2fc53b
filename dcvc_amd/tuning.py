"""Tile autotuner and its database.

The reference launches every CUTLASS kernel with a tile picked from per-shape
tuned maps (HINT_MAP_BIAS, cutlass/conv1x1_bias.cu:8-522; HINT_MAP_D3x3,
cutlass/d3x3.cu:7-441; dispatch in cutlass/cutlass_kernel.h:233-249), tuned
for 3840x2160, 1920x1080, 1280x720, 832x480 and 416x240 (README.md:202).

Here the tuner runs the codecs themselves: for each model and resolution it
traces one intra frame plus one chunk encode and decode with per-launch HIP
event timing, once under every candidate GEMM tile configuration and
depthwise strip length, and keeps, per launch shape, the candidate with the
least total time.  The result is a JSON database (``tuning_db.json`` next to
this file) that the drop-in proxies install at ``set_param``
(``ensure_loaded``), before any graph is captured.  Tile choice never
changes results, so the database only moves time.

The HT-S batches of the batching form (streams of one resolution coded as
one batch, M = streams x pixels) are tuned the same way.

Usage on a GPU:  python -m dcvc_amd.tuning [--out PATH] [--models hts,htl,ld]
                                            [--res 1080x1920,...] [--reps 3]
                                            [--batches 4x1080x1920,...]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import threading
import time
from pathlib import Path

import numpy as np

from . import _lib

DB_PATH = Path(__file__).resolve().with_name("tuning_db.json")
RESOLUTIONS = [(2160, 3840), (1080, 1920), (720, 1280), (480, 832), (480, 854), (240, 416)]
# HT-S batches (streams, H, W): the resolution groups of the batching form's
# 16-stream mix (batching.MIX_16), each layer one launch over the group
BATCHES = [(4, 480, 854), (6, 720, 1280), (4, 1080, 1920), (2, 2160, 3840)]
# 6 and 7 hold one block per CU: faster alone on some shapes, slower inside
# the codecs, whose concurrent streams need a second block to share the CU;
# 8-11 (64-deep K tiles on large tiles) were no faster there.  12-14 are the
# small-grid tiles.
GEMM_CANDIDATES = (1, 2, 3, 4, 5, 12, 13, 14)
DW_CANDIDATES = (4, 8)
COLS = 15  # trace record width (include/dcvc_codec.h)

_loaded = None
_load_lock = threading.Lock()


def install(db: dict) -> tuple[int, int]:
    """Installs a database dict; returns (GEMM entries, depthwise entries)."""
    g = [[e["prec"], e["epi"], e["amode"], e["dmode"], e["M"], e["N"], e["K"], e["cfg"]] for e in db.get("gemm", [])]
    d = [[e["n_img"], e["H"], e["W"], e["C"], e["rows"]] for e in db.get("d3x3", [])]
    ga = np.asarray(g, dtype=np.int64).reshape(-1, 8)
    da = np.asarray(d, dtype=np.int64).reshape(-1, 5)
    _lib.call("dcvc_gemm_tune_set", ga.ctypes.data if len(ga) else None, len(ga))
    _lib.call("dcvc_d3x3_tune_set", da.ctypes.data if len(da) else None, len(da))
    return len(ga), len(da)


def clear():
    _lib.call("dcvc_gemm_tune_set", None, 0)
    _lib.call("dcvc_d3x3_tune_set", None, 0)


def ensure_loaded(path: str | os.PathLike | None = None):
    """Installs the tuning database once per process (the proxies call this at
    set_param).  DCVC_TUNING_DB names another file, or 'off' for the built-in
    heuristics.  A missing file means the heuristics."""
    global _loaded
    with _load_lock:
        if _loaded is not None:
            return _loaded
        env = os.environ.get("DCVC_TUNING_DB", "")
        p = Path(path or env or DB_PATH)
        if env.lower() == "off" or not p.exists():
            _loaded = (0, 0)
        else:
            with open(p) as f:
                _loaded = install(json.load(f))
        return _loaded


# ----------------------------------------------------------------- tuner

def _trace(fn):
    lib = _lib.load()
    buf = np.zeros((200000, COLS), np.int64)
    _lib.call("dcvc_trace_begin", 1)
    try:
        fn()
        import torch
        torch.cuda.synchronize()
    finally:
        n = lib.dcvc_trace_end(buf.ctypes.data_as(C.c_void_p), len(buf))
    if n < 0:
        raise _lib.DcvcError("dcvc_trace_end", -n)
    return buf[:n].copy()


def _pass_fn(model, H, W, prec, seed=0):
    """One intra frame + one chunk encode + decode of `model` at H x W, graphs
    off; returns a callable that reruns it."""
    import torch
    from . import codec, synthetic
    dev = torch.device("cuda:0")
    sdi, sdp = codec.make_state_dicts(seed=seed, video=model)
    cls = codec.video_codec(model)
    ie, idec = codec.IntraCodec(sdi, 0.15), codec.IntraCodec(sdi, 0.15)
    pe = cls(sdp, 0.15)
    pd = cls(share_from=pe)
    for o in (ie, idec, pe, pd):
        o.proxy.set_use_graphs(False)
    if prec != "fp16":
        for o in (pe, pd):
            o.proxy.set_precision(prec)
    pad_r, pad_b = synthetic.get_padding_size(H, W, 16)
    n = 1 + cls.FRAME_DELAY
    frames = synthetic.motion_sequence_yuv420(H, W, n, seed=seed + 1)
    x0 = codec.to_input(synthetic.frame_to_input(frames[0]), dev)
    if cls.FRAME_DELAY == 1:
        xp = codec.to_input(synthetic.frame_to_input(frames[1]), dev)
    else:
        xp = codec.to_input(synthetic.chunk_to_input(frames[1:5], cls.FRAME_DELAY), dev)
    del frames

    def run():
        e = ie.compress(x0, 32, pad_b, pad_r)
        pe.add_ref_feature_from_frame(e["x_hat"])
        d = idec.decompress(e["bit_stream"], {"height": H, "width": W}, 32, e["ec_parallel"])
        pd.add_ref_feature_from_frame(d["x_hat"], apply_feature_adaptor=False)
        c = pe.compress(xp, 32, 0, pad_b, pad_r)
        pd.decompress(c["bit_stream"], {"height": H, "width": W}, 32, c["ec_parallel"], 0)

    return run


def _batch_pass_fn(n, H, W, prec, seed=0):
    """One chunk encode + decode of `n` HT-S streams of H x W coded as one
    batch (batching.BatchGroup: every layer one launch over the n pictures,
    M = n x pixels), graphs off.  The streams' I frames are coded once, before
    the timed passes; returns a callable that codes the next chunk."""
    import torch
    from . import batching, codec, synthetic
    dev = torch.device("cuda:0")
    sdi, sdp = codec.make_state_dicts(seed=seed, video="hts")
    shared = codec.VideoCodecHTS(sdp, 0.15)
    group = batching.BatchGroup(range(n), H, W, shared, prec)
    group.enc.set_use_graphs(False)
    group.dec.set_use_graphs(False)
    ie, idec = codec.IntraCodec(sdi, 0.15), codec.IntraCodec(sdi, 0.15)
    firsts, xs = [], []
    for i in range(n):
        frames = synthetic.motion_sequence_yuv420(H, W, 5, seed=seed + 1 + 17 * i)
        firsts.append(codec.to_input(synthetic.frame_to_input(frames[0]), dev))
        xs.append(codec.to_input(synthetic.chunk_to_input(frames[1:5], 8), dev))
    group.start(ie, idec, firsts)
    qps = [32] * n

    def run():
        group.code_chunk(xs, qps)

    return run


def tune_one(model, H, W, prec="fp16", reps=3, batch=1):
    """Per-shape timings of one (model, resolution), or with batch > 1 of an
    HT-S batch of that many streams: returns (gemm, d3x3) dicts key ->
    {candidate: total microseconds}, plus the heuristic's choice."""
    import torch
    run = _pass_fn(model, H, W, prec) if batch == 1 else _batch_pass_fn(batch, H, W, prec)
    run()  # allocation, first-launch costs
    torch.cuda.synchronize()
    clear()

    def timed(policy):
        _lib.call("dcvc_gemm_set_tile_policy", policy)
        runs = [_trace(run) for _ in range(reps)]
        base = runs[0]
        for r in runs[1:]:
            if r.shape != base.shape or not np.array_equal(r[:, :5], base[:, :5]):
                raise RuntimeError("launch sequence changed between passes")
        return base, np.median(np.stack([r[:, 7] for r in runs]), axis=0)

    try:
        heur, _ = timed(0)
        ref_keys = heur[:, :5]
        gmask = heur[:, 0] <= 2
        gemm = {}
        for pol in GEMM_CANDIDATES:
            recs, t = timed(pol)
            if not np.array_equal(recs[:, :5], ref_keys):
                raise RuntimeError("launch sequence changed between policies")
            for i in np.nonzero(gmask)[0]:
                r = heur[i]
                key = (int(r[10]), int(r[1]), int(r[8]), int(r[9]), int(r[2]), int(r[3]), int(r[4]))
                e = gemm.setdefault(key, {"heuristic": int(r[11]), "us": {}})
                e["us"][pol] = e["us"].get(pol, 0.0) + float(t[i]) / 1e3
        _lib.call("dcvc_gemm_set_tile_policy", 0)
        # depthwise: force each strip length for every traced shape
        dmask = heur[:, 0] == 3
        d3 = {}
        keys = sorted({(int(r[12]), int(r[13]), int(r[14]), int(r[3])) for r in heur[dmask]})
        for rows in DW_CANDIDATES:
            ent = np.asarray([[*k, rows] for k in keys], np.int64).reshape(-1, 5)
            _lib.call("dcvc_d3x3_tune_set", ent.ctypes.data if len(ent) else None, len(ent))
            runs = [_trace(run) for _ in range(reps)]
            t = np.median(np.stack([r[:, 7] for r in runs]), axis=0)
            for i in np.nonzero(dmask)[0]:
                r = heur[i]
                e = d3.setdefault((int(r[12]), int(r[13]), int(r[14]), int(r[3])), {"heuristic": int(r[11]), "us": {}})
                e["us"][rows] = e["us"].get(rows, 0.0) + float(t[i]) / 1e3
    finally:
        _lib.call("dcvc_gemm_set_tile_policy", 0)
        clear()
    return gemm, d3


def choose(entry, margin=0.03):
    """The fastest candidate, unless it beats the heuristic's choice by less
    than `margin` (timing noise): then the heuristic's choice."""
    us = entry["us"]
    best = min(us, key=us.get)
    h = entry["heuristic"]
    if h in us and us[best] > (1 - margin) * us[h]:
        return h
    return best


def tune(models=("hts", "htl", "ld"), resolutions=RESOLUTIONS, precs=("fp16",), reps=3, batches=BATCHES, log=print):
    """Tunes every (model, precision, resolution), then the HT-S batches
    (streams, H, W) of the batching form."""
    gemm, d3 = {}, {}
    t_start = time.time()
    jobs = [(model, prec, H, W, 1) for model in models for prec in precs for (H, W) in resolutions]
    jobs += [("hts", prec, H, W, n) for prec in precs for (n, H, W) in batches]
    for model, prec, H, W, n in jobs:
        t0 = time.time()
        g, d = tune_one(model, H, W, prec=prec, reps=reps, batch=n)
        for k, e in g.items():
            acc = gemm.setdefault(k, {"heuristic": e["heuristic"], "us": {}})
            for c, v in e["us"].items():
                acc["us"][c] = acc["us"].get(c, 0.0) + v
        for k, e in d.items():
            acc = d3.setdefault(k, {"heuristic": e["heuristic"], "us": {}})
            for c, v in e["us"].items():
                acc["us"][c] = acc["us"].get(c, 0.0) + v
        what = f"{model} {prec} {H}x{W}" + (f" batch of {n}" if n > 1 else "")
        log(f"tuned {what}: {len(g)} GEMM shapes, {len(d)} depthwise shapes, {time.time() - t0:.1f} s", flush=True)
    import torch
    db = {"device": torch.cuda.get_device_name(0), "created_by": "python -m dcvc_amd.tuning",
          "models": list(models), "precisions": list(precs),
          "resolutions": [f"{W}x{H}" for H, W in resolutions],
          "hts_batches": [f"{n}x{W}x{H}" for n, H, W in batches],
          "tuning_seconds": round(time.time() - t_start, 1), "gemm": [], "d3x3": []}
    for (prec, epi, amode, dmode, M, N, K), e in sorted(gemm.items()):
        db["gemm"].append({"prec": prec, "epi": epi, "amode": amode, "dmode": dmode, "M": M, "N": N, "K": K,
                           "cfg": int(choose(e)), "heuristic": e["heuristic"],
                           "us": {str(c): round(v, 2) for c, v in sorted(e["us"].items())}})
    for (n_img, H, W, Cc), e in sorted(d3.items()):
        db["d3x3"].append({"n_img": n_img, "H": H, "W": W, "C": Cc, "rows": int(choose(e)),
                           "heuristic": e["heuristic"],
                           "us": {str(c): round(v, 2) for c, v in sorted(e["us"].items())}})
    return db


def summary(db):
    """Total traced microseconds under the heuristic and under the database's choices."""
    out = {}
    for name in ("gemm", "d3x3"):
        h = sum(e["us"].get(str(e["heuristic"]), 0.0) for e in db[name])
        t = sum(e["us"][str(e["cfg" if name == "gemm" else "rows"])] for e in db[name])
        changed = sum(1 for e in db[name] if e["cfg" if name == "gemm" else "rows"] != e["heuristic"])
        out[name] = {"shapes": len(db[name]), "changed": changed, "heuristic_us": round(h, 1), "tuned_us": round(t, 1)}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(DB_PATH))
    ap.add_argument("--models", default="hts,htl,ld")
    ap.add_argument("--precs", default="fp16")
    ap.add_argument("--res", default=",".join(f"{h}x{w}" for h, w in RESOLUTIONS), help="HxW list")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--batches", default=",".join(f"{n}x{h}x{w}" for n, h, w in BATCHES),
                    help="HT-S batches, NxHxW list ('' for none)")
    a = ap.parse_args(argv)
    import torch
    torch.cuda.set_device(0)
    res = [tuple(int(v) for v in r.split("x")) for r in a.res.split(",") if r]
    batches = [tuple(int(v) for v in b.split("x")) for b in a.batches.split(",") if b]
    db = tune(models=a.models.split(","), resolutions=res, precs=a.precs.split(","), reps=a.reps, batches=batches)
    with open(a.out, "w") as f:
        json.dump(db, f, indent=1)
    print(json.dumps(summary(db)))
    print(f"wrote {a.out}")


if __name__ == "__main__":
    sys.exit(main())
