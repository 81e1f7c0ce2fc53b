"""Host model of the fp16 -> OCP FP8 E4M3 conversion the library applies to
activations (test helper)."""
import numpy as np


def e4m3_rne_bytes(x16):
    """fp16 values -> OCP E4M3 bytes: saturate at +-448, then round to
    nearest, ties to even mantissa (subnormals m / 8 * 2^-6).  Written from
    the format's definition, independent of the library's conversion."""
    x = np.asarray(x16, np.float16).astype(np.float64)
    x = np.clip(x, -448.0, 448.0)
    a = np.abs(x)
    e = np.floor(np.log2(np.maximum(a, 2.0 ** -6)))
    step = 2.0 ** (e - 3)
    q = a / step
    r = np.floor(q)
    frac = q - r
    r = np.where(frac > 0.5, r + 1, np.where(frac < 0.5, r, r + (r % 2)))
    v = r * step  # rounded magnitude, exactly representable
    # encode: normal if v >= 2^-6
    out = np.zeros(x.shape, np.int64)
    normal = v >= 2.0 ** -6
    ev = np.floor(np.log2(np.where(normal, v, 1.0)))
    mant = np.where(normal, v / 2.0 ** ev - 1.0, 0.0) * 8
    out = np.where(normal, ((ev + 7).astype(np.int64) << 3) | np.round(mant).astype(np.int64),
                   np.round(v / 2.0 ** -9).astype(np.int64))
    sign = np.signbit(x)  # kept on zeros as well: -0 and negatives that round to 0 give 0x80
    return (out | (sign.astype(np.int64) << 7)).astype(np.uint8)
