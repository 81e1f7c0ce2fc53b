"""CPU pins of the oracle's integer boundary (scale->index table, rounding,
masks) and of the C++ table used by the kernels."""
import numpy as np
import pytest
import torch

from oracle import dcvc_oracle as O


def test_threshold_table_matches_float64_formula_on_every_fp16():
    """Every finite positive fp16 scale (the GPU stores scales in fp16) indexes
    identically through the table and through the float64 formula."""
    bits = np.arange(0, 0x7C00, dtype=np.uint16)
    s = bits.view(np.float16).astype(np.float32)
    assert np.array_equal(O.scale_to_index(s), O.scale_index_exact(s))
    neg = -s
    assert np.all(O.scale_to_index(neg) == 0)


def test_threshold_table_random_float32_and_edges():
    r = np.random.Generator(np.random.PCG64(0))
    s = np.exp(r.uniform(np.log(0.05), np.log(20.0), 200000)).astype(np.float32)
    assert np.array_equal(O.scale_to_index(s), O.scale_index_exact(s))
    th = O.scale_index_thresholds()
    for k in range(1, 127):
        if np.isinf(th[k]):
            break
        below = np.nextafter(th[k], np.float32(0))
        assert O.scale_index_exact(th[k]) >= k > O.scale_index_exact(below)
    # closed-form endpoints: ln(0.11) ~ LOG_SCALE_MIN -> 0; ln(16) < LOG_SCALE_MAX -> 126
    assert O.scale_to_index(np.float32(0.11)) == 0
    assert O.scale_to_index(np.float32(16.0)) == 126
    assert O.scale_to_index(np.float32(1e9)) == 126


def test_threshold_table_consistent_with_scale_table():
    """The entropy model's scale table exp(linspace(ln .11, ln 16, 128))
    (entropy_models.py:179-182): scale_table[k] falls in bin k or k-1 (the
    truncated constants shift bins by < 1 step)."""
    table = np.exp(np.linspace(np.log(0.11), np.log(16.0), 128)).astype(np.float32)
    idx = O.scale_to_index(table)
    k = np.arange(128)
    assert np.all((idx == k) | (idx == k - 1) | ((k == 127) & (idx == 126)))


def test_cpp_table_equals_oracle_table():
    from dcvc_amd import kernels
    th_cpp = kernels.host_thresholds()
    th_py = O.scale_index_thresholds()
    assert np.array_equal(th_cpp[1:], th_py[1:])


def test_round_half_away():
    x = np.array([0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 0.49999997, -0.49999997, 3.0], np.float32)
    assert np.array_equal(O.round_half_away(x), [1, 2, 3, -1, -2, -3, 0, 0, 3])
    t = torch.from_numpy(x)
    assert torch.equal(O.round_half_away(t), torch.tensor([1., 2, 3, -1, -2, -3, 0, 0, 3]))


def test_mask4_partition_and_pattern():
    """Each element is active in exactly one of the 4 steps; step 0 of group 0
    is the (0,0) position (common_model.py:180-185)."""
    C, H, W = 8, 5, 7
    ms = torch.stack(O.mask_4x(C, H, W)).int().sum(0)
    assert torch.all(ms == 1)
    m0 = O.mask_4x(C, H, W)[0][0]
    assert bool(m0[0, 0, 0]) and not bool(m0[0, 0, 1]) and not bool(m0[0, 1, 0])
    # group g in step k uses position g ^ {0,3,2,1}[k] (the kernels' closed form)
    for k, x in enumerate([0, 3, 2, 1]):
        mk = O.mask_4x(C, H, W)[k][0]
        for g in range(4):
            pos = g ^ x
            dy, dx = pos >> 1, pos & 1
            assert bool(mk[g * 2, dy, dx]) and int(mk[g * 2].sum()) == len(range(dy, H, 2)) * len(range(dx, W, 2))


def test_compute_ec_parallel():
    assert O.compute_ec_parallel(0) == 1
    assert O.compute_ec_parallel(32767) == 1
    assert O.compute_ec_parallel(65536) == 2
    assert O.compute_ec_parallel(10 ** 8) == 8
