"""Bitstream container with the reference's byte layout
(src/utils/stream_helper.py:37-160): NAL header byte (type << 4 | sps id),
SPS = adaptive-length height and width, I/P NAL = qp byte, (ec_part << 1 |
reset) byte, adaptive-length payload size, payload."""
from __future__ import annotations

import enum
import io
import struct


class NalType(enum.IntEnum):
    NAL_SPS = 0
    NAL_I = 1
    NAL_P = 2


def write_uint_adaptive(f, a):
    """1 byte below 2^7, 2 bytes (10xxxxxx) below 2^14, else 4 bytes (11xxxxxx)."""
    if a < (1 << 7):
        f.write(bytes([a]))
        return 1
    if a < (1 << 14):
        f.write(bytes([0x80 | (a >> 8), a & 0xFF]))
        return 2
    if a >= (1 << 30):
        raise ValueError("value too large for the adaptive uint field")
    f.write(bytes([0xC0 | (a >> 24), (a >> 16) & 0xFF, (a >> 8) & 0xFF, a & 0xFF]))
    return 4


def _read(f, n):
    b = f.read(n)
    if len(b) != n:
        raise ValueError("truncated bitstream")
    return b


def read_uint_adaptive(f):
    a3 = _read(f, 1)[0]
    if (a3 >> 7) == 0:
        return a3
    a2 = _read(f, 1)[0]
    if (a3 >> 6) == 0x02:
        return ((a3 & 0x3F) << 8) + a2
    a1, a0 = _read(f, 2)
    return ((a3 & 0x3F) << 24) + (a2 << 16) + (a1 << 8) + a0


def write_sps(f, sps):
    if not 0 <= sps["sps_id"] < 16:
        raise ValueError("sps_id must fit in 4 bits")
    n = f.write(bytes([(int(NalType.NAL_SPS) << 4) + sps["sps_id"]]))
    n += write_uint_adaptive(f, sps["height"])
    n += write_uint_adaptive(f, sps["width"])
    return n


def write_ip(f, is_i_frame, sps_id, qp, ec_part, reset_feature_memory, bit_stream):
    if not 0 <= qp < 256 or not 0 <= ec_part < 128:
        raise ValueError("qp / ec_part out of range")
    t = NalType.NAL_I if is_i_frame else NalType.NAL_P
    n = f.write(bytes([(int(t) << 4) + sps_id, qp, (ec_part << 1) + int(bool(reset_feature_memory))]))
    n += write_uint_adaptive(f, len(bit_stream))
    n += f.write(bytes(bit_stream))
    return n


def read_header(f):
    flag = _read(f, 1)[0]
    nal_type = flag >> 4
    if nal_type >= 3:
        raise ValueError(f"unsupported NAL type {nal_type}")
    return {"nal_type": NalType(nal_type), "sps_id": flag & 0x0F}


def read_sps_remaining(f, sps_id):
    return {"sps_id": sps_id, "height": read_uint_adaptive(f), "width": read_uint_adaptive(f)}


def read_ip_remaining(f):
    qp = _read(f, 1)[0]
    flag = _read(f, 1)[0]
    n = read_uint_adaptive(f)
    return qp, (flag >> 1) & 0x7F, flag & 0x01, _read(f, n)


class SPSHelper:
    def __init__(self):
        self.spss = []

    def add_sps_by_id(self, sps):
        for i, s in enumerate(self.spss):
            if s["sps_id"] == sps["sps_id"]:
                self.spss[i] = dict(sps)
                return
        self.spss.append(dict(sps))

    def get_sps_by_id(self, sps_id):
        for s in self.spss:
            if s["sps_id"] == sps_id:
                return s
        return None

    def get_sps_id(self, target):
        max_id = -1
        for s in self.spss:
            if s["height"] == target["height"] and s["width"] == target["width"]:
                return s["sps_id"], False
            max_id = max(max_id, s["sps_id"])
        if max_id >= 15:
            raise ValueError("too many SPS")
        s = dict(target)
        s["sps_id"] = max_id + 1
        self.spss.append(s)
        return s["sps_id"], True


def new_buffer():
    return io.BytesIO()
