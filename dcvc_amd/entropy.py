"""Host-side entropy model: the CPU rANS coder (C++ in libdcvc_amd.so) with the
interface of the reference's MLCodec_extensions_cpp module
(src/cpp/py_rans/bind.cpp:14-40), and the CDF tables the codec proxies load.

CDF tables follow src/models/entropy_models.py:
* ``gaussian_cdf_table``      GaussianEncoder.update (:184-217)
* ``bit_estimator_cdf_table`` BitEstimator.update    (:113-149)
Both reorder each pmf to 0, +1, -1, +2, ... (:45-57), append the tail mass
and quantise it with ``pmf_to_quantized_cdf`` (py_rans.cpp:35-94).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib

MAX_ENTROPY_CODING_VALUE = 8          # entropy_models.py:12
SCALE_MIN, SCALE_MAX, SCALE_LEVEL = 0.11, 16.0, 128


def pmf_to_quantized_cdf(pmf):
    pmf = np.ascontiguousarray(np.asarray(pmf, dtype=np.float32))
    out = np.zeros(pmf.size + 1, dtype=np.uint32)
    _lib.call("dcvc_pmf_to_quantized_cdf", pmf.ctypes.data, pmf.size, out.ctypes.data)
    return out.tolist()


def compute_ec_parallel(symbol_count):
    return int(_lib.load().dcvc_compute_ec_parallel(int(symbol_count)))


class RansEncoder:
    """Same methods as MLCodec_extensions_cpp.RansEncoder."""

    def __init__(self):
        # the library handle is kept on the instance: at interpreter exit the
        # module globals may already be gone when __del__ runs
        self._lib = _lib.load()
        self._h = self._lib.dcvc_rans_encoder_create()
        if not self._h:
            raise _lib.DcvcError("dcvc_rans_encoder_create", 2)
        self._keep = []

    def __del__(self):
        h = getattr(self, "_h", None)
        lib = getattr(self, "_lib", None)
        if h and lib is not None:
            lib.dcvc_rans_encoder_destroy(h)
            self._h = None

    def set_cdf(self, cdfs, cdf_sizes, index):
        cdfs = np.ascontiguousarray(np.asarray(cdfs, dtype=np.int32))
        sizes = np.ascontiguousarray(np.asarray(cdf_sizes, dtype=np.int32).reshape(-1))
        n = sizes.size
        row = cdfs.size // n
        _lib.call("dcvc_rans_encoder_set_cdf", self._h, cdfs.ctypes.data, n, row, sizes.ctypes.data, index)

    def set_entropy_coder_parallel(self, n):
        _lib.call("dcvc_rans_encoder_set_parallel", self._h, int(n))

    def reset(self):
        self._keep = []
        _lib.call("dcvc_rans_encoder_reset", self._h)

    def encode_y(self, symbols):
        s = np.ascontiguousarray(np.asarray(symbols, dtype=np.int16).reshape(-1))
        self._keep.append(s)
        _lib.call("dcvc_rans_encoder_encode_y", self._h, s.ctypes.data, s.size)

    def encode_z(self, symbols, cdf_offset, ch):
        s = np.ascontiguousarray(np.asarray(symbols, dtype=np.int8).reshape(-1))
        self._keep.append(s)
        _lib.call("dcvc_rans_encoder_encode_z", self._h, s.ctypes.data, s.size, int(cdf_offset), int(ch))

    def flush(self):
        _lib.call("dcvc_rans_encoder_flush", self._h)
        self._keep = []

    def get_encoded_stream(self):
        p = C.POINTER(C.c_uint8)()
        n = C.c_long()
        _lib.call("dcvc_rans_encoder_stream", self._h, C.byref(p), C.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint8)


class RansDecoder:
    """Same methods as MLCodec_extensions_cpp.RansDecoder (decode returns the array)."""

    def __init__(self):
        self._lib = _lib.load()
        self._h = self._lib.dcvc_rans_decoder_create()
        if not self._h:
            raise _lib.DcvcError("dcvc_rans_decoder_create", 2)
        self._last = np.zeros(0, np.int8)

    def __del__(self):
        h = getattr(self, "_h", None)
        lib = getattr(self, "_lib", None)
        if h and lib is not None:
            lib.dcvc_rans_decoder_destroy(h)
            self._h = None

    def set_cdf(self, cdfs, cdf_sizes, index):
        cdfs = np.ascontiguousarray(np.asarray(cdfs, dtype=np.int32))
        sizes = np.ascontiguousarray(np.asarray(cdf_sizes, dtype=np.int32).reshape(-1))
        n = sizes.size
        _lib.call("dcvc_rans_decoder_set_cdf", self._h, cdfs.ctypes.data, n, cdfs.size // n, sizes.ctypes.data,
                  index)

    def set_entropy_coder_parallel(self, n):
        _lib.call("dcvc_rans_decoder_set_parallel", self._h, int(n))

    def set_stream(self, data):
        d = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8) if isinstance(data, (bytes, bytearray))
                                 else np.asarray(data, dtype=np.uint8))
        self._stream = d
        _lib.call("dcvc_rans_decoder_set_stream", self._h, d.ctypes.data, d.size)

    def decode_z(self, total_size, cdf_offset, ch):
        out = np.zeros(int(total_size), np.int8)
        _lib.call("dcvc_rans_decoder_decode_z", self._h, out.size, int(cdf_offset), int(ch), out.ctypes.data)
        self._last = out
        return out

    def decode_y(self, indexes):
        idx = np.ascontiguousarray(np.asarray(indexes, dtype=np.uint8).reshape(-1))
        out = np.zeros(idx.size, np.int8)
        _lib.call("dcvc_rans_decoder_decode_y", self._h, idx.ctypes.data, idx.size, out.ctypes.data)
        self._last = out
        return out

    def get_decoded_tensor(self):
        return self._last


# ---------------------------------------------------------------------------
# CDF tables
# ---------------------------------------------------------------------------

def reorder_prob(prob):
    """pmf (symmetric around the centre, tail last) -> order 0, +1, -1, +2, -2, ..."""
    length = prob.shape[0]
    out = prob.clone()
    center = (length - 1) // 2
    out[0] = prob[center]
    for i in range(1, center + 1):
        out[2 * i - 1] = prob[center + i]
        out[2 * i] = prob[center - i]
    return out


def pmf_to_cdf(pmf, tail_mass, pmf_length, max_length):
    cdf = torch.zeros((len(pmf_length), max_length + 2), dtype=torch.int32)
    for i, p in enumerate(pmf):
        prob = torch.cat((p[: int(pmf_length[i])], tail_mass[i]), dim=0)
        q = pmf_to_quantized_cdf(reorder_prob(prob).numpy())
        cdf[i, : len(q)] = torch.tensor(q, dtype=torch.int32)
    return cdf


def _normal_cdf(x, scale):
    # torch.distributions.Normal(0, scale).cdf
    return 0.5 * (1 + torch.erf(x * scale.reciprocal() / math.sqrt(2)))


def gaussian_scale_table():
    return torch.exp(torch.linspace(math.log(SCALE_MIN), math.log(SCALE_MAX), SCALE_LEVEL))


def gaussian_cdf_table():
    """(quantized_cdf int32 [128, 19], cdf_length int32 [128])."""
    scales = gaussian_scale_table()
    zeros = torch.zeros_like(scales)
    sym_range = zeros + MAX_ENTROPY_CODING_VALUE
    for i in range(MAX_ENTROPY_CODING_VALUE, 1, -1):
        probs = _normal_cdf(zeros + i, scales)
        sym_range = torch.where(probs > 0.999, torch.full_like(sym_range, i), sym_range)
    sym_range = sym_range.int()
    pmf_length = 2 * sym_range + 1
    max_length = 2 * MAX_ENTROPY_CODING_VALUE + 1
    samples = (torch.arange(max_length) - sym_range[:, None]).float()
    sc = scales[:, None]
    upper = _normal_cdf(samples + 0.5, sc)
    lower = _normal_cdf(samples - 0.5, sc)
    pmf = upper - lower
    tail_mass = 2 * lower[:, :1]
    cdf = pmf_to_cdf(pmf, tail_mass, pmf_length, max_length)
    return cdf.numpy(), (pmf_length + 2).numpy().astype(np.int32)


def _bit_estimator_prob(x, h, b, a):
    # layers.py:13-19
    for i in range(4):
        x = x * torch.nn.functional.softplus(h[:, :, i:i + 1, None]) + b[:, :, i:i + 1, None]
        if i != 3:
            x = x + torch.tanh(x) * torch.tanh(a[:, :, i:i + 1, None])
    return torch.sigmoid(x)


def bit_estimator_cdf_table(h, b, a):
    """h, b [64, C, 4], a [64, C, 3] float tensors -> (cdf [64*C, 19], cdf_length [64*C])."""
    h, b, a = (torch.as_tensor(t, dtype=torch.float32) for t in (h, b, a))
    qp_num, ch = h.shape[0], h.shape[1]
    zeros = torch.zeros((qp_num, ch, 1, 1))
    sym_range = zeros + MAX_ENTROPY_CODING_VALUE
    for i in range(MAX_ENTROPY_CODING_VALUE, 1, -1):
        neg = _bit_estimator_prob(zeros - i, h, b, a)
        pos = _bit_estimator_prob(zeros + i, h, b, a)
        sym_range = torch.where(torch.logical_and(neg < 0.001, pos > 0.999), torch.full_like(sym_range, i),
                                sym_range)
    sym_range = sym_range.int()
    pmf_length = sym_range * 2 + 1
    max_length = MAX_ENTROPY_CODING_VALUE * 2 + 1
    samples = torch.arange(max_length)[None, None, None, :] - sym_range
    lower = _bit_estimator_prob(samples - 0.5, h, b, a)
    upper = _bit_estimator_prob(samples + 0.5, h, b, a)
    pmf = (upper - lower)[:, :, 0, :]
    upper = _bit_estimator_prob(sym_range.float(), h, b, a)
    tail_mass = lower[:, :, 0, :1] + (1.0 - upper[:, :, 0, -1:])
    pmf = pmf.reshape([-1, max_length])
    tail_mass = tail_mass.reshape([-1, 1])
    pmf_length = pmf_length.reshape([-1])
    cdf = pmf_to_cdf(pmf, tail_mass, pmf_length, max_length)
    return cdf.numpy(), (pmf_length + 2).numpy().astype(np.int32)


def add_cdf_to_state_dict(sd):
    """CompressionModel.add_cdf_to_state_dict (common_model.py:64-70) for a
    state dict holding bit_estimator_z.{h,b,a}."""
    g_cdf, g_len = gaussian_cdf_table()
    z_cdf, z_len = bit_estimator_cdf_table(sd["bit_estimator_z.h"], sd["bit_estimator_z.b"],
                                           sd["bit_estimator_z.a"])
    sd = dict(sd)
    sd["gaussian_encoder.quantized_cdf"] = torch.from_numpy(g_cdf)
    sd["gaussian_encoder.cdf_length"] = torch.from_numpy(g_len)
    sd["bit_estimator_z.quantized_cdf"] = torch.from_numpy(z_cdf)
    sd["bit_estimator_z.cdf_length"] = torch.from_numpy(z_len)
    return sd
