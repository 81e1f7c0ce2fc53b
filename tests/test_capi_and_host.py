"""CPU checks of the boundary and the host-side pieces: the C ABI library
loads and exports every symbol include/*.h declares (no GPU calls), the
bitstream container round-trips with the reference byte layout, the
synthetic inputs are deterministic, and host argument validation rejects
bad calls before anything reaches a GPU."""
import io

import numpy as np
import pytest

from dcvc_amd import _lib, stream_helper, synthetic


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _lib.declared_symbols()
    assert len(declared) >= 60
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    # every declared symbol also has a ctypes signature
    assert not [n for n in declared if n not in _lib.SIGNATURES]


def test_version_and_host_only_entry_points():
    lib = _lib.load()
    assert lib.dcvc_version() >= 100
    assert lib.dcvc_compute_ec_parallel(10 ** 7) == 8
    assert lib.dcvc_compact_blocks(4097) == 2


def test_invalid_arguments_are_rejected_on_the_host():
    lib = _lib.load()
    # K not a multiple of 64, null pointers: rejected before any launch
    assert lib.dcvc_conv1x1_bias(None, 64, 10, 60, None, 8, None, None, 8, None) == 1
    assert lib.dcvc_d3x3(None, 8, 1, 4, 4, 8, None, None, 8, None) == 1
    assert lib.dcvc_conv_bias(None, 64, 1, 5, 5, 64, None, 8, None, 5, 2, None, 8, None) == 1
    assert lib.dcvc_conditional_index(None, 3, None, 10, None, None, None, None) == 1
    assert lib.dcvc_hts_set_param(None, 0, None, None, None, None, 0.0) == 1
    # calibration kernel: null or misaligned operands, bad sizes, unknown operand type
    assert lib.dcvc_mfma_calibrate(None, None, 4, 10, 0, None) == 1
    assert lib.dcvc_mfma_calibrate(32, 64, 0, 10, 0, None) == 1
    assert lib.dcvc_mfma_calibrate(32, 64, 4, 0, 0, None) == 1
    assert lib.dcvc_mfma_calibrate(32, 64, 4, 10, 3, None) == 1
    assert lib.dcvc_mfma_calibrate(48, 64, 4, 10, 0, None) == 1


def test_stream_container_layout_and_round_trip():
    f = io.BytesIO()
    sps = {"sps_id": 0, "height": 1080, "width": 1920}
    n = stream_helper.write_sps(f, sps)
    n += stream_helper.write_ip(f, True, 0, 32, 1, 0, b"\x01\x02\x03")
    n += stream_helper.write_ip(f, False, 0, 21, 8, 1, bytes(range(200)))
    data = f.getvalue()
    assert len(data) == n
    # SPS byte: nal type 0 << 4 | id 0; 1080 = 0x438 -> 2-byte form 0x84 0x38
    assert data[:5] == bytes([0x00, 0x84, 0x38, 0x87, 0x80])
    # I NAL: (1 << 4) | 0, qp, ec_part << 1 | reset, length, payload
    assert data[5:12] == bytes([0x10, 32, 2, 3, 1, 2, 3])
    r = io.BytesIO(data)
    h = stream_helper.read_header(r)
    assert h["nal_type"] == stream_helper.NalType.NAL_SPS
    assert stream_helper.read_sps_remaining(r, h["sps_id"]) == sps
    h = stream_helper.read_header(r)
    assert h["nal_type"] == stream_helper.NalType.NAL_I
    assert stream_helper.read_ip_remaining(r) == (32, 1, 0, b"\x01\x02\x03")
    h = stream_helper.read_header(r)
    assert h["nal_type"] == stream_helper.NalType.NAL_P
    assert stream_helper.read_ip_remaining(r) == (21, 8, 1, bytes(range(200)))
    for v in (0, 127, 128, 16383, 16384, (1 << 30) - 1):
        b = io.BytesIO()
        stream_helper.write_uint_adaptive(b, v)
        assert stream_helper.read_uint_adaptive(io.BytesIO(b.getvalue())) == v
    with pytest.raises(ValueError):
        stream_helper.read_ip_remaining(io.BytesIO(bytes([32, 2, 10, 1])))  # truncated payload


def test_synthetic_inputs_are_deterministic_and_shaped():
    a = synthetic.motion_sequence_yuv420(72, 96, 5, seed=7)
    b = synthetic.motion_sequence_yuv420(72, 96, 5, seed=7)
    for fa, fb in zip(a, b):
        for pa, pb in zip(fa, fb):
            assert np.array_equal(pa, pb)
    y, u, v = a[0]
    assert y.shape == (72, 96) and u.shape == (36, 48) and y.dtype == np.uint8
    x = synthetic.chunk_to_input(a[1:5])
    assert x.shape == (24, 72, 96) and x.min() >= -0.5 and x.max() <= 0.5
    # padding to 8 frames repeats the last frame (test_video.py:104-110)
    assert np.array_equal(x[9:12], x[21:24])
    # occlusion re-noises ~10 % of luma pixels relative to a pure shift
    assert synthetic.get_padding_size(1080, 1920, 16) == (0, 8)


def test_profile_kernel_names():
    """rocprofv3 kernel names map to the short names and classes the profile
    summaries and bench.py use, with or without the K tile depth parameter."""
    from dcvc_amd import profiling as P
    quad = "void dcvc::gemm_kernel<dcvc::Tile<128, 256, 2, 2, 3, 6, 0, 0, 0, 32> >(dcvc::GemmArgs, int)"
    assert P.short_name(quad) == "gemm_kernel<128,256,2,2,3,6,0,0,0> [conv1x1_quad]"
    assert P.short_name(quad.replace(", 32>", ">")) == P.short_name(quad)
    k64 = "void dcvc::gemm_kernel<dcvc::Tile<256, 256, 4, 2, 2, 1, 0, 0, 1, 64> >(dcvc::GemmArgs, int)"
    assert P.short_name(k64) == "gemm_kernel<256,256,4,2,2,1,0,0,1,k64> [conv1x1_wsilu]"
    conv = "void dcvc::gemm_kernel<dcvc::Tile<256, 128, 2, 2, 3, 0, 1, 0, 0, 32> >(dcvc::GemmArgs, int)"
    assert P.gemm_class(conv) == "conv_s2_bias"


def test_bench_tile_tables_follow_the_gemm_dispatch():
    """bench.py names each traced GEMM launch by its tile configuration; its
    tables must list, for every configuration id, the Tile the fp16 / bf16
    dispatch of gemm.hip launches (64-deep K tiles with their fallback for
    launches that are not plain 1x1 products)."""
    import importlib.util
    import re
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    spec = importlib.util.spec_from_file_location("bench_mod", root / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    src = (root / "dcvc_amd" / "csrc" / "gemm.hip").read_text()
    body = src[src.index("static hipError_t dispatch_tiles"):src.index("static hipError_t dispatch_plain")]
    quad_part = body[body.index("} else if constexpr (EPI == EPI_QUAD) {"):]
    quad_part, plain_part = quad_part[:quad_part.index("    } else {")], quad_part[quad_part.index("    } else {"):]
    # configuration 3 is the switch's default
    pat = re.compile(r"(?:case (\d+)|default):( if constexpr \(kPlain\))? return launch_tile<(\d+), (\d+), (\d+), "
                     r"(\d+), (\d+), EPI, AMODE, DMODE, PREC(?:, (\d+))?>")
    n_policies = int(re.search(r"constexpr int kPolicies = (\d+);", src).group(1))
    for part, table in ((quad_part, bench.TILES_QUAD), (plain_part, bench.TILES)):
        found = {}
        for m in pat.finditer(part):
            t = tuple(int(v) for v in m.group(3, 4, 5, 6, 7))
            if m.group(8) and m.group(8) != "32":
                t += (int(m.group(8)),)
            found[int(m.group(1) or 3)] = (t, bool(m.group(2)))
        assert set(found) == set(range(1, n_policies + 1)) == set(table)
        for cfg, (t, plain_only) in found.items():
            assert table[cfg] == t, cfg
            assert (cfg in bench.TILES_FALLBACK) == plain_only, cfg
    # the fallback of a plain-only configuration is the configuration the
    # dispatch falls through to (the default tile equals configuration 3's)
    for cfg, fb in bench.TILES_FALLBACK.items():
        assert len(bench.TILES[cfg]) == 6 and len(bench.TILES[fb]) == 5
