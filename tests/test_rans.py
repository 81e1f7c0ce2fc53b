"""CPU entropy coder: byte-exact against the reference's MLCodec_extensions_cpp
(built from /root/reference sources into oracle/_ref by oracle/build_ref.sh),
round trips, escape symbols, stream merging for every parallelism, and
failure handling on truncated / corrupted streams."""
import math
import os
import sys

import numpy as np
import pytest
import torch

from dcvc_amd import entropy

REF_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")


def ref_module():
    if REF_DIR not in sys.path:
        sys.path.insert(0, REF_DIR)
    try:
        import MLCodec_extensions_cpp as m  # reference coder, test oracle only
    except ImportError:
        pytest.skip("oracle/_ref reference coder not built (oracle/build_ref.sh)")
    return m


def rng(s):
    return np.random.Generator(np.random.PCG64(s))


@pytest.fixture(scope="module")
def tables():
    g_cdf, g_len = entropy.gaussian_cdf_table()
    r = rng(5)
    h = torch.from_numpy(r.standard_normal((64, 128, 4)).astype(np.float32) * 0.5)
    b = torch.from_numpy(r.standard_normal((64, 128, 4)).astype(np.float32) * 0.5)
    a = torch.from_numpy(r.standard_normal((64, 128, 3)).astype(np.float32) * 0.5)
    z_cdf, z_len = entropy.bit_estimator_cdf_table(h, b, a)
    return g_cdf, g_len, z_cdf, z_len


def random_symbols(r, n, heavy=False):
    idx = r.integers(0, 128, size=n)
    scale = np.exp(np.log(0.11) + idx / 127.0 * (np.log(16) - np.log(0.11)))
    s = np.clip(np.round(r.standard_normal(n) * scale * (3 if heavy else 1)), -128, 127).astype(np.int64)
    if heavy:
        s[r.integers(0, n, size=max(1, n // 50))] = r.integers(-128, 128, size=max(1, n // 50))
    return ((s << 8) + idx).astype(np.int16)


def test_pmf_to_quantized_cdf_matches_reference():
    m = ref_module()
    r = rng(0)
    for trial in range(300):
        n = int(r.integers(2, 40))
        p = r.random(n).astype(np.float32)
        if trial % 3 == 0:
            p[r.integers(0, n, size=n // 2)] = 1e-9  # forces frequency stealing
        p = p / p.sum()
        assert entropy.pmf_to_quantized_cdf(p) == list(m.pmf_to_quantized_cdf(p.tolist()))


def test_gaussian_pmf_closed_form(tables):
    """The Gaussian table's pmf is N(0, s) mass of [k-0.5, k+0.5]; checked
    against the erfc closed form in float64 through the first nonzero row."""
    g_cdf, g_len, _, _ = tables
    scales = entropy.gaussian_scale_table().double().numpy()
    for i in (0, 40, 90, 127):
        L = int(g_len[i]) - 2
        half = (L - 1) // 2
        k = np.arange(-half, half + 1)
        pmf = 0.5 * (np.vectorize(math.erfc)(-(k + 0.5) / (scales[i] * math.sqrt(2)))
                     - np.vectorize(math.erfc)(-(k - 0.5) / (scales[i] * math.sqrt(2))))
        cdf = g_cdf[i, : L + 2].astype(np.int64)
        freq = np.diff(cdf) / 65536.0
        order = [half] + [v for j in range(1, half + 1) for v in (half + j, half - j)]
        # quantisation: rounding (1/65536 each) plus min-frequency steals ((L+1)/65536 at most)
        assert np.allclose(freq[:L], pmf[order], atol=(L + 3) / 65536.0)
        assert cdf[-1] == 65536 and np.all(np.diff(cdf) > 0)


@pytest.mark.parametrize("parallel", [1, 2, 3, 4, 5, 7, 8])
def test_stream_bytes_match_reference(tables, parallel):
    m = ref_module()
    g_cdf, g_len, z_cdf, z_len = tables
    r = rng(parallel)
    y = random_symbols(r, 50000 + parallel * 777, heavy=True)
    z = np.clip(np.round(r.standard_normal(128 * 17 * 30) * 3), -64, 63).astype(np.int8)
    qp = 37
    ours = entropy.RansEncoder()
    theirs = m.RansEncoder()
    for enc in (ours, theirs):
        enc.set_cdf(z_cdf, z_len, 0)
        enc.set_cdf(g_cdf, g_len, 1)
        enc.reset()
        enc.set_entropy_coder_parallel(parallel)
        enc.encode_y(y[: len(y) // 2])
        enc.encode_y(y[len(y) // 2:])
        enc.encode_z(z, qp * 128, 128)
        enc.flush()
    a = ours.get_encoded_stream()
    b = np.asarray(theirs.get_encoded_stream())
    assert a.tobytes() == b.tobytes()
    # decode with ours
    dec = entropy.RansDecoder()
    dec.set_cdf(z_cdf, z_len, 0)
    dec.set_cdf(g_cdf, g_len, 1)
    dec.set_entropy_coder_parallel(parallel)
    dec.set_stream(a)
    assert np.array_equal(dec.decode_z(z.size, qp * 128, 128), z)
    idx = (y.astype(np.int64) & 0xFF).astype(np.uint8)
    sym = (y.astype(np.int64) >> 8).astype(np.int8)
    h = len(y) // 2
    # rANS is LIFO: batches come back in the reverse of their encode order
    assert np.array_equal(dec.decode_y(idx[h:]), sym[h:])
    assert np.array_equal(dec.decode_y(idx[:h]), sym[:h])


def test_round_trip_without_reference(tables):
    g_cdf, g_len, z_cdf, z_len = tables
    r = rng(99)
    for parallel in (1, 2, 6, 8):
        y = random_symbols(r, 20000, heavy=True)
        enc = entropy.RansEncoder()
        enc.set_cdf(g_cdf, g_len, 1)
        enc.set_cdf(z_cdf, z_len, 0)
        enc.reset()
        enc.set_entropy_coder_parallel(parallel)
        enc.encode_y(y)
        enc.flush()
        s = enc.get_encoded_stream()
        dec = entropy.RansDecoder()
        dec.set_cdf(g_cdf, g_len, 1)
        dec.set_cdf(z_cdf, z_len, 0)
        dec.set_entropy_coder_parallel(parallel)
        dec.set_stream(s)
        out = dec.decode_y((y.astype(np.int64) & 0xFF).astype(np.uint8))
        assert np.array_equal(out, (y.astype(np.int64) >> 8).astype(np.int8))


def test_empty_and_tiny_inputs(tables):
    g_cdf, g_len, _, _ = tables
    for n, parallel in [(0, 1), (0, 3), (1, 1), (2, 8), (7, 8)]:
        y = random_symbols(rng(n), n)
        enc = entropy.RansEncoder()
        enc.set_cdf(g_cdf, g_len, 1)
        enc.reset()
        enc.set_entropy_coder_parallel(parallel)
        enc.encode_y(y)
        enc.flush()
        s = enc.get_encoded_stream()
        dec = entropy.RansDecoder()
        dec.set_cdf(g_cdf, g_len, 1)
        dec.set_entropy_coder_parallel(parallel)
        dec.set_stream(s)
        out = dec.decode_y((y.astype(np.int64) & 0xFF).astype(np.uint8))
        assert np.array_equal(out, (y.astype(np.int64) >> 8).astype(np.int8))


def test_truncated_and_corrupted_streams_raise(tables):
    g_cdf, g_len, _, _ = tables
    y = random_symbols(rng(3), 40000)
    idx = (y.astype(np.int64) & 0xFF).astype(np.uint8)
    for parallel in (1, 4):
        enc = entropy.RansEncoder()
        enc.set_cdf(g_cdf, g_len, 1)
        enc.reset()
        enc.set_entropy_coder_parallel(parallel)
        enc.encode_y(y)
        enc.flush()
        s = enc.get_encoded_stream()
        dec = entropy.RansDecoder()
        dec.set_cdf(g_cdf, g_len, 1)
        dec.set_entropy_coder_parallel(parallel)
        with pytest.raises(Exception):
            dec.set_stream(s[: len(s) // 3])
            dec.decode_y(idx)
        with pytest.raises(Exception):
            dec.set_stream(s[:2])
        bad = s.copy()
        bad[: 4 * (parallel // 2)] = 0xFF  # corrupt header offsets
        if parallel > 2:
            with pytest.raises(Exception):
                dec.set_stream(bad)


def test_bad_arguments_raise(tables):
    g_cdf, g_len, _, _ = tables
    enc = entropy.RansEncoder()
    with pytest.raises(Exception):
        enc.set_entropy_coder_parallel(9)
    with pytest.raises(Exception):
        enc.encode_y(np.zeros(4, np.int16))  # cdf table not set
    enc.set_cdf(g_cdf, g_len, 1)
    enc.reset()
    enc.encode_y(np.array([(1 << 8) + 200], np.int16))  # cdf index 200 >= 128
    with pytest.raises(Exception):
        enc.flush()


def test_compute_ec_parallel():
    assert entropy.compute_ec_parallel(0) == 1
    assert entropy.compute_ec_parallel(3 * 32768 + 5) == 3
    assert entropy.compute_ec_parallel(2088960) == 8


def wide_table(rows=4, half=159):
    """Laplace-shaped rows of 2 * half + 1 symbols, cdf_length 2 * half + 3:
    the reference holds cdf_length - 2 as int8 (py_rans.cpp:266), so 321
    wraps to a maximum symbol of 63 and larger values take the escape path."""
    n = 2 * half + 1
    k = torch.arange(n, dtype=torch.float64) - half
    pmf = torch.stack([torch.exp(-k.abs() / (2.0 + 8 * r)) for r in range(rows)])
    pmf = (pmf / pmf.sum(1, keepdim=True) * 0.999).float()
    tail = torch.full((rows, 1), 0.001)
    lengths = torch.full((rows,), n, dtype=torch.int32)
    return entropy.pmf_to_cdf(pmf, tail, lengths, n).numpy(), (lengths + 2).numpy().astype(np.int32)


def test_wide_rows_round_trip_and_match_reference():
    cdf, lens = wide_table()
    assert int(lens.max()) - 2 > 256
    for bad_len in (131, 257):  # negative int8 maximum: rejected
        with pytest.raises(Exception):
            entropy.RansDecoder().set_cdf(cdf, np.full_like(lens, bad_len), 1)
    r = rng(11)
    n = 30000
    idx = r.integers(0, len(lens), size=n)
    v = np.clip(np.round(r.laplace(0, 6 + 20 * idx)), -128, 127).astype(np.int64)
    y = ((v << 8) + idx).astype(np.int16)
    for parallel in (1, 3):
        enc = entropy.RansEncoder()
        enc.set_cdf(cdf, lens, 1)
        enc.reset()
        enc.set_entropy_coder_parallel(parallel)
        enc.encode_y(y)
        enc.flush()
        s = enc.get_encoded_stream()
        try:
            m = ref_module()
        except pytest.skip.Exception:
            m = None
        if m is not None:
            ref = m.RansEncoder()
            ref.set_cdf(cdf, lens, 1)
            ref.reset()
            ref.set_entropy_coder_parallel(parallel)
            ref.encode_y(y)
            ref.flush()
            assert s.tobytes() == np.asarray(ref.get_encoded_stream()).tobytes()
        dec = entropy.RansDecoder()
        dec.set_cdf(cdf, lens, 1)
        dec.set_entropy_coder_parallel(parallel)
        dec.set_stream(s)
        assert np.array_equal(dec.decode_y(idx.astype(np.uint8)), v.astype(np.int8))


def test_garbage_streams_raise_or_decode_in_bounds(tables):
    """Random bytes as a stream: the decoder either raises or returns symbols,
    never reads outside the stream (the reader clamps at its end byte)."""
    g_cdf, g_len, _, _ = tables
    w_cdf, w_len = wide_table()
    r = rng(12)
    for trial in range(200):
        cdf, lens = (g_cdf, g_len) if trial % 2 else (w_cdf, w_len)
        dec = entropy.RansDecoder()
        dec.set_cdf(cdf, lens, 1)
        dec.set_entropy_coder_parallel(int(r.integers(1, 9)))
        data = r.integers(0, 256, size=int(r.integers(0, 200)), dtype=np.uint8)
        idx = r.integers(0, len(lens), size=int(r.integers(1, 400))).astype(np.uint8)
        try:
            dec.set_stream(data)
            out = dec.decode_y(idx)
        except Exception:
            continue
        assert out.shape == idx.shape


def test_escape_only_stream_outgrows_the_first_buffer():
    """Every symbol escapes through a near-zero tail frequency (16 bits) plus
    up to 12 bits of bypass bins, about 3.5 bytes a symbol: more than the
    encoder's first output buffer (2 bytes a symbol), so it must grow."""
    row = entropy.pmf_to_quantized_cdf(np.array([0.7, 0.3 - 1e-7, 1e-7], np.float32))
    assert row[3] - row[2] == 1  # the escape symbol (max_value 2) at the minimum frequency
    cdf = np.array([row] * 4, np.int32)
    lens = np.full(4, 4, np.int32)
    r = rng(21)
    n = 20000
    v = r.choice(np.array([-128, -100, 90, 127]), size=n)
    y = ((v.astype(np.int64) << 8) + r.integers(0, 4, size=n)).astype(np.int16)
    for parallel in (1, 2):
        enc = entropy.RansEncoder()
        enc.set_cdf(cdf, lens, 1)
        enc.reset()
        enc.set_entropy_coder_parallel(parallel)
        enc.encode_y(y)
        enc.flush()
        s = enc.get_encoded_stream()
        assert s.size > 3 * n
        try:
            m = ref_module()
        except pytest.skip.Exception:
            m = None
        if m is not None:
            ref = m.RansEncoder()
            ref.set_cdf(cdf, lens, 1)
            ref.reset()
            ref.set_entropy_coder_parallel(parallel)
            ref.encode_y(y)
            ref.flush()
            assert s.tobytes() == np.asarray(ref.get_encoded_stream()).tobytes()
        dec = entropy.RansDecoder()
        dec.set_cdf(cdf, lens, 1)
        dec.set_entropy_coder_parallel(parallel)
        dec.set_stream(s)
        assert np.array_equal(dec.decode_y((y.astype(np.int64) & 0xFF).astype(np.uint8)), v.astype(np.int8))


def test_zero_frequency_symbol_raises():
    """A symbol whose cdf row gives it no frequency cannot be coded: the
    encoder raises instead of renormalising forever."""
    cdf = np.array([[0, 65535, 65535, 65536]] * 2, np.int32)  # value 1 has no frequency
    lens = np.full(2, 4, np.int32)
    enc = entropy.RansEncoder()
    enc.set_cdf(cdf, lens, 1)
    enc.reset()
    enc.encode_y(np.array([(0 << 8) + 1, (-5 << 8) + 0], np.int16))  # codable: 0, and -5 as an escape
    enc.flush()
    enc.reset()
    enc.encode_y(np.array([(0 << 8) + 1, (1 << 8) + 0], np.int16))  # +1 codes value 1
    with pytest.raises(Exception):
        enc.flush()
