"""The host E4M3 model used by the fp8-activation GPU tests, pinned by
values whose encodings follow from the format and by torch's own
float8_e4m3fn conversion over every finite fp16 value."""
import numpy as np
import torch

from e4m3_model import e4m3_rne_bytes


def test_e4m3_model_spot_values():
    """The host model on values whose encodings follow from the format."""
    vals = np.array([0.0, 1.0, -1.0, 448.0, 500.0, -1e4, 2.0 ** -6, 2.0 ** -9, 2.0 ** -10, 1.0625, 1.1875,
                     3 * 2.0 ** -10, -2.0 ** -11], np.float16)
    want = np.array([0x00, 0x38, 0xB8, 0x7E, 0x7E, 0xFE, 0x08, 0x01, 0x00, 0x38, 0x3A, 0x02, 0x80], np.uint8)
    # 1.0625 is the tie between 1.0 (mantissa 0, even) and 1.125: rounds to 1.0;
    # 1.1875 is the tie between 1.125 (odd) and 1.25 (even): rounds to 1.25;
    # 2^-10 is the tie between 0 and 2^-9: to 0; 3 * 2^-10 between 2^-9 and 2^-8: to 2^-8
    assert np.array_equal(e4m3_rne_bytes(vals), want)


def test_e4m3_model_equals_torch_float8_on_every_fp16():
    v = np.arange(65536, dtype=np.uint16).view(np.float16)
    v = v[np.isfinite(v)]
    t = torch.from_numpy(np.clip(v.astype(np.float32), -448, 448)).to(torch.float8_e4m3fn).view(torch.uint8)
    assert np.array_equal(e4m3_rne_bytes(v), t.numpy())
