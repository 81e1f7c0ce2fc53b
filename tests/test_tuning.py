"""The tile autotuner's database and the tuning table behind it.

CPU: the table entry points (host-only) install, look up and clear entries
and reject malformed ones; the committed database parses and installs.
GPU: one tuning pass over a small HT-S picture times every candidate for
every traced shape, and the codec's bitstream is the same with the tuned
table as with the heuristic (tile choice never changes results)."""
import json

import numpy as np
import pytest

from dcvc_amd import _lib, tuning

F16, BF16 = 0, 1
EPI_BIAS, EPI_SHORTCUT, EPI_QUAD = 0, 2, 6


def _set(entries, fn="dcvc_gemm_tune_set", width=8):
    a = np.asarray(entries, np.int64).reshape(-1, width)
    return getattr(_lib.load(), fn)(a.ctypes.data if len(a) else None, len(a))


def test_gemm_table_overrides_the_heuristic_and_clears():
    lib = _lib.load()
    try:
        h = lib.dcvc_gemm_tile_for(F16, EPI_SHORTCUT, 0, 0, 32640, 512, 256)
        assert 1 <= h <= 14
        other = 5 if h != 5 else 2
        assert _set([[F16, EPI_SHORTCUT, 0, 0, 32640, 512, 256, other]]) == 0
        assert lib.dcvc_gemm_tile_for(F16, EPI_SHORTCUT, 0, 0, 32640, 512, 256) == other
        # the key is the whole shape: other shapes, epilogues, precisions keep the heuristic
        assert lib.dcvc_gemm_tile_for(F16, EPI_SHORTCUT, 0, 0, 32640, 512, 512) == \
            lib.dcvc_gemm_tile_for(F16, EPI_SHORTCUT, 0, 0, 32640, 512, 512)
        assert lib.dcvc_gemm_tile_for(BF16, EPI_SHORTCUT, 0, 0, 32640, 512, 256) == h
        assert lib.dcvc_gemm_tile_for(F16, EPI_BIAS, 0, 0, 32640, 512, 256) != 0
        # a forced policy wins over the table
        assert lib.dcvc_gemm_set_tile_policy(3) == 0
        assert lib.dcvc_gemm_tile_for(F16, EPI_SHORTCUT, 0, 0, 32640, 512, 256) == 3
        assert lib.dcvc_gemm_set_tile_policy(0) == 0
        assert _set([]) == 0
        assert lib.dcvc_gemm_tile_for(F16, EPI_SHORTCUT, 0, 0, 32640, 512, 256) == h
    finally:
        lib.dcvc_gemm_set_tile_policy(0)
        _set([])


def test_heuristic_tile_choice():
    """The automatic choice by grid size (args: prec, epi, amode, dmode, M, N,
    K): long-K 1/8-resolution layers 256 x 256; large grids 2 blocks per CU;
    smaller grids the 8-wave 128 x 128 tile; grids of fewer than 256 128 x 128
    tiles (1/16 and 1/64 resolution) 64 x 128; quads 128 x 256 or 128 x 128."""
    lib = _lib.load()
    tile = lib.dcvc_gemm_tile_for
    assert tile(F16, EPI_BIAS, 0, 0, 32640, 512, 2048) == 1      # K >= 1024, 128 x 2 256-tiles
    assert tile(F16, EPI_SHORTCUT, 0, 0, 32640, 512, 512) == 4   # 128 x 4 = 512 256 x 128 tiles
    assert tile(F16, EPI_SHORTCUT, 0, 0, 32640, 256, 256) == 12  # 255 x 2 = 510 128 x 128 tiles
    assert tile(F16, EPI_SHORTCUT, 0, 0, 8160, 256, 256) == 13   # 64 x 2 = 128 128 x 128 tiles
    assert tile(F16, EPI_BIAS, 0, 0, 510, 1024, 128) == 13
    assert tile(F16, EPI_QUAD, 0, 0, 32640, 256, 256) == 4
    assert tile(F16, EPI_QUAD, 0, 0, 8160, 256, 256) == 3
    # the same rule for bf16 products
    assert tile(BF16, EPI_SHORTCUT, 0, 0, 8160, 256, 256) == 13


def test_gemm_table_rejects_malformed_entries():
    lib = _lib.load()
    h = lib.dcvc_gemm_tile_for(F16, EPI_QUAD, 0, 0, 8160, 768, 768)
    for bad in ([F16, EPI_QUAD, 0, 0, 8160, 768, 768, 0], [F16, EPI_QUAD, 0, 0, 8160, 768, 768, 15],
                [4, EPI_QUAD, 0, 0, 8160, 768, 768, 1], [F16, EPI_QUAD, 0, 0, 0, 768, 768, 1],
                [F16, EPI_QUAD, 0, 0, 1 << 24, 768, 768, 1]):
        assert _set([[F16, EPI_QUAD, 0, 0, 8160, 768, 768, 2], bad]) == 1
    # a rejected table leaves the previous one in place
    assert lib.dcvc_gemm_tile_for(F16, EPI_QUAD, 0, 0, 8160, 768, 768) == h
    assert lib.dcvc_gemm_tile_for(F16, 9, 0, 0, 8160, 768, 768) < 0


def test_depthwise_table():
    lib = _lib.load()
    try:
        h = lib.dcvc_d3x3_strip_for(1, 136, 240, 256)
        assert h in (4, 8)
        other = 12 - h
        assert _set([[1, 136, 240, 256, other]], "dcvc_d3x3_tune_set", 5) == 0
        assert lib.dcvc_d3x3_strip_for(1, 136, 240, 256) == other
        assert lib.dcvc_d3x3_strip_for(1, 68, 120, 256) in (4, 8)
        assert _set([[1, 136, 240, 256, 6]], "dcvc_d3x3_tune_set", 5) == 1
    finally:
        _set([], "dcvc_d3x3_tune_set", 5)
    assert lib.dcvc_d3x3_strip_for(1, 136, 240, 256) == h


def test_committed_database_installs():
    if not tuning.DB_PATH.exists():
        pytest.skip("no tuning database committed")
    db = json.loads(tuning.DB_PATH.read_text())
    assert db["gemm"] and db["d3x3"]
    assert {"3840x2160", "1920x1080", "1280x720", "832x480", "416x240"} <= set(db["resolutions"])
    for e in db["gemm"]:
        assert e["cfg"] in tuning.GEMM_CANDIDATES and str(e["cfg"]) in e["us"]
        assert e["us"][str(e["cfg"])] <= e["us"][str(e["heuristic"])] * 1.0001 or e["cfg"] == e["heuristic"]
    try:
        assert tuning.install(db) == (len(db["gemm"]), len(db["d3x3"]))
        e = db["gemm"][0]
        assert _lib.load().dcvc_gemm_tile_for(e["prec"], e["epi"], e["amode"], e["dmode"], e["M"], e["N"],
                                              e["K"]) == e["cfg"]
    finally:
        tuning.clear()


@pytest.mark.gpu
def test_tuning_pass_and_bitstreams_independent_of_the_table(gpu):
    import torch
    from dcvc_amd import codec, synthetic
    g, d = tuning.tune_one("hts", 128, 192, reps=1)
    assert g and d
    for e in g.values():
        assert set(e["us"]) == set(tuning.GEMM_CANDIDATES) and all(v > 0 for v in e["us"].values())
    for e in d.values():
        assert set(e["us"]) == set(tuning.DW_CANDIDATES)
    # a table that moves every shape off its heuristic choice: same bits
    entries = [[*k, next(c for c in tuning.GEMM_CANDIDATES if c != e["heuristic"])] for k, e in g.items()]
    H, W = 128, 192
    sdi, sdp = codec.make_state_dicts(seed=0)
    frames = synthetic.motion_sequence_yuv420(H, W, 9, seed=2)
    x0 = codec.to_input(synthetic.frame_to_input(frames[0]), gpu)
    xc = codec.to_input(synthetic.chunk_to_input(frames[1:9]), gpu)
    out = []
    for table in ([], entries):
        tuning.install({"gemm": [dict(zip(("prec", "epi", "amode", "dmode", "M", "N", "K", "cfg"), t))
                                 for t in table], "d3x3": []})
        try:
            ie = codec.IntraCodec(sdi)
            pe = codec.VideoCodecHTS(sdp)
            e = ie.compress(x0, 32, 0, 0)
            pe.add_ref_feature_from_frame(e["x_hat"])
            out.append((e["bit_stream"], pe.compress(xc, 32, 0, 0, 0)["bit_stream"]))
            torch.cuda.synchronize()
        finally:
            tuning.clear()
    assert out[0] == out[1]


@pytest.mark.gpu
def test_tuning_pass_of_an_hts_batch(gpu):
    """The tuner's batch pass (streams of one resolution coded as one batch)
    times the batched launch shapes: the 1/8-resolution GEMMs carry M = streams
    x pixels."""
    n, H, W = 2, 64, 96
    g, d = tuning.tune_one("hts", H, W, reps=1, batch=n)
    assert g and d
    ms = {k[4] for k in g}
    assert n * (H // 8) * (W // 8) in ms
    for e in g.values():
        assert set(e["us"]) == set(tuning.GEMM_CANDIDATES) and all(v > 0 for v in e["us"].values())
