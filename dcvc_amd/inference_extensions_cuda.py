"""Drop-in replacement of the reference's ``inference_extensions_cuda`` module
(src/layers/extensions/inference/bind.cpp:11-39): the same four proxy
classes with the same methods and argument meaning, backed by the MI355X
C ABI in libdcvc_amd.so (include/dcvc_codec.h).

Model code imports it as ``from inference_extensions_cuda import DMCHTSProxy``
(image_model.py:197, video_model_ht.py:419-424) once this package directory
is on ``sys.path`` under that name (``dcvc_amd.install_module_alias()``).

Tensors are the caller's torch HIP tensors: pictures are fp16 [1, C, H, W]
in channels_last memory (test_video.py:118-122).  Returned reconstructions
are views of proxy-owned buffers, overwritten by the next call, as in the
reference (dmc_hts_proxy.cpp:347, dmci_proxy.cpp:418).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib, tuning


_PREC = {"fp16": 0, "bf16": 1, "fp8": 2}


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _nhwc(t: torch.Tensor) -> torch.Tensor:
    """fp16 [1, C, H, W] -> contiguous NHWC storage (a view when channels_last)."""
    if t.dtype != torch.float16:
        t = t.half()
    if not t.is_cuda:
        raise ValueError("inputs must be on the GPU")
    if not t.is_contiguous(memory_format=torch.channels_last):
        t = t.contiguous(memory_format=torch.channels_last)
    return t


class _StateDictArgs:
    """Marshals a state dict into the (names, float32 data, shapes, ndims) arrays."""

    def __init__(self, state_dict):
        self.keep = []
        names, ptrs, shapes, ndims = [], [], [], []
        for k, v in state_dict.items():
            a = torch.as_tensor(v).detach().to("cpu", torch.float32).contiguous().numpy()
            if a.ndim > 4:
                raise ValueError(f"{k}: more than 4 dimensions")
            self.keep.append(a)
            names.append(k.encode())
            ptrs.append(a.ctypes.data)
            shp = list(a.shape) + [1] * (4 - a.ndim)
            shapes.extend(shp)
            ndims.append(a.ndim)
        n = len(names)
        self.n = n
        self.names = (C.c_char_p * n)(*names)
        self.ptrs = (C.c_void_p * n)(*ptrs)
        self.shapes = (C.c_int64 * (4 * n))(*shapes)
        self.ndims = (C.c_int * n)(*ndims)

    def args(self):
        return (self.n, C.cast(self.names, C.c_void_p), C.cast(self.ptrs, C.c_void_p),
                C.cast(self.shapes, C.c_void_p), C.cast(self.ndims, C.c_void_p))


def _bits_to_numpy(p, n):
    if n.value == 0:
        return np.zeros(0, np.uint8)
    return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n.value,)).copy()


class DMCIProxy:
    """DMCIProxy (dmci_proxy.h:132-254)."""

    def __init__(self):
        self._lib = _lib.load()
        self._h = self._lib.dcvc_dmci_create()
        if not self._h:
            raise _lib.DcvcError("dcvc_dmci_create", 2)
        self._x_hat = None
        self._size = None
        self._device = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.dcvc_dmci_destroy(h)
            self._h = None

    def set_param(self, state_dict, skip_threshold):
        dev = state_dict["q_scale_dec"].device if isinstance(state_dict["q_scale_dec"], torch.Tensor) else None
        self._device = dev if dev is not None and dev.type == "cuda" else torch.device("cuda", torch.cuda.current_device())
        tuning.ensure_loaded()  # tile choices, before any graph is captured
        a = _StateDictArgs(state_dict)
        _lib.call("dcvc_dmci_set_param", self._h, *a.args(), float(skip_threshold))

    def set_use_graphs(self, on):
        _lib.call("dcvc_dmci_set_use_graphs", self._h, int(bool(on)))

    def set_precision(self, prec):
        """'fp16' (default), 'bf16' or 'fp8': precision of every 1x1-conv product."""
        _lib.call("dcvc_dmci_set_precision", self._h, _PREC[prec])

    def _prepare(self, height, width):
        Hp, Wp = (height + 15) // 16 * 16, (width + 15) // 16 * 16
        if self._size != (Hp, Wp):
            self._x_hat = torch.empty((1, 3, Hp, Wp), dtype=torch.float16, device=self._device,
                                      memory_format=torch.channels_last)
            _lib.call("dcvc_dmci_prepare", self._h, Hp, Wp, self._x_hat.data_ptr(), 3)
            self._size = (Hp, Wp)

    def compress(self, x, qp, padding_b, padding_r):
        """-> (np.uint8 bitstream, x_hat [1,3,Hp,Wp], ec_parallel)  (dmci_proxy.cpp:296-421)"""
        x = _nhwc(x)
        _, c, H, W = x.shape
        if c != 3:
            raise ValueError("DMCI input must have 3 channels")
        self._prepare(H + padding_b, W + padding_r)
        bits = C.c_void_p()
        n = C.c_long()
        ec = C.c_int()
        _lib.call("dcvc_dmci_compress", self._h, x.data_ptr(), 3, H, W, int(qp), int(padding_b), int(padding_r),
                  _stream(), C.byref(bits), C.byref(n), C.byref(ec))
        return _bits_to_numpy(bits, n), self._x_hat, ec.value

    def decompress(self, bit_stream, qp, height, width, entropy_coder_parallel):
        """-> x_hat [1,3,Hp,Wp]  (dmci_proxy.cpp:423-602)"""
        self._prepare(height, width)
        b = np.ascontiguousarray(np.asarray(bit_stream, dtype=np.uint8))
        _lib.call("dcvc_dmci_decompress", self._h, b.ctypes.data, b.size, int(qp), int(height), int(width),
                  int(entropy_coder_parallel), _stream())
        return self._x_hat

    def step_symbol_count(self, step):
        return int(self._lib.dcvc_dmci_step_symbol_count(self._h, step))

    def debug_copy(self, what, nbytes, dtype):
        out = np.zeros(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        _lib.call("dcvc_dmci_debug_copy", self._h, int(what), out.ctypes.data, out.nbytes)
        return out


class _ChunkProxy:
    """Shared body of the 8-frame chunk proxies (HT-S, HT-L): the same methods
    and argument meaning as the reference classes, over the C ABI entry
    points named dcvc_<prefix>_*."""
    FRAMES = 8
    _PREFIX = ""
    _NAME = ""

    def __init__(self):
        self._lib = _lib.load()
        self._h = getattr(self._lib, f"dcvc_{self._PREFIX}_create")()
        if not self._h:
            raise _lib.DcvcError(f"dcvc_{self._PREFIX}_create", 2)
        self._frames = None
        self._size = None
        self._device = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            getattr(self._lib, f"dcvc_{self._PREFIX}_destroy")(h)
            self._h = None

    def _call(self, fn, *args):
        return _lib.call(f"dcvc_{self._PREFIX}_{fn}", self._h, *args)

    def set_param(self, state_dict, skip_threshold):
        q = state_dict["q_decoder"]
        dev = q.device if isinstance(q, torch.Tensor) else None
        self._device = dev if dev is not None and dev.type == "cuda" else torch.device("cuda", torch.cuda.current_device())
        tuning.ensure_loaded()  # tile choices, before any graph is captured
        a = _StateDictArgs(state_dict)
        self._call("set_param", *a.args(), float(skip_threshold))

    def share_weights_from(self, other):
        """Use `other`'s weights (one HBM copy for many stream contexts)."""
        if not isinstance(other, _ChunkProxy) or other._PREFIX != self._PREFIX:
            raise TypeError("weights can only be shared between proxies of the same model structure")
        self._device = other._device
        self._call("share_weights", other._h)

    def set_use_graphs(self, on):
        self._call("set_use_graphs", int(bool(on)))

    def set_precision(self, prec):
        """'fp16' (default), 'bf16' or 'fp8': precision of every 1x1-conv product."""
        self._call("set_precision", _PREC[prec])

    def device_bytes(self):
        """HBM held by this context's activation buffers (weights excluded)."""
        return int(getattr(self._lib, f"dcvc_{self._PREFIX}_device_bytes")(self._h))

    def _prepare(self, Hp, Wp):
        if self._size != (Hp, Wp):
            self._frames = [torch.empty((1, 3, Hp, Wp), dtype=torch.float16, device=self._device,
                                        memory_format=torch.channels_last) for _ in range(self.FRAMES)]
            ptrs = (C.c_void_p * self.FRAMES)(*[f.data_ptr() for f in self._frames])
            self._call("prepare", Hp, Wp, C.cast(ptrs, C.c_void_p), 3)
            self._size = (Hp, Wp)

    def add_ref_feature_from_frame(self, frame, apply_adaptor=True):
        """frame: [1,3,Hp,Wp], a decoded I frame."""
        frame = _nhwc(frame)
        _, c, Hp, Wp = frame.shape
        if c != 3 or Hp % 16 or Wp % 16:
            raise ValueError("reference frame must be [1,3,H,W] with H, W multiples of 16")
        self._prepare(Hp, Wp)
        self._call("add_ref_feature_from_frame", frame.data_ptr(), 3, Hp, Wp, int(bool(apply_adaptor)), _stream())

    def compress(self, x, qp, reset_feature_memory, padding_b, padding_r):
        """-> (np.uint8 bitstream, ec_parallel); x: [1,24,H,W]"""
        x = _nhwc(x)
        _, c, H, W = x.shape
        if c != 24:
            raise ValueError(f"{self._NAME} chunk input must have 24 channels (8 frames)")
        bits = C.c_void_p()
        n = C.c_long()
        ec = C.c_int()
        self._call("compress", x.data_ptr(), 24, H, W, int(qp), int(bool(reset_feature_memory)), int(padding_b),
                   int(padding_r), _stream(), C.byref(bits), C.byref(n), C.byref(ec))
        return _bits_to_numpy(bits, n), ec.value

    def decompress(self, bit_stream, qp, height, width, entropy_coder_parallel, reset_feature_memory):
        """-> list of 8 tensors [1,3,Hp,Wp]"""
        b = np.ascontiguousarray(np.asarray(bit_stream, dtype=np.uint8))
        self._call("decompress", b.ctypes.data, b.size, int(qp), int(height), int(width), int(entropy_coder_parallel),
                   int(bool(reset_feature_memory)), _stream())
        return list(self._frames)

    # --- test hooks (not in the reference interface) ---
    def debug_copy(self, what, nbytes, dtype):
        out = np.zeros(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        self._call("debug_copy", int(what), out.ctypes.data, out.nbytes)
        return out


class DMCHTSProxy(_ChunkProxy):
    """DMCHTSProxy (dmc_hts_proxy.h:215-355; add_ref_feature_from_frame
    dmc_hts_proxy.cpp:492-502, compress :504-585, decompress :587-710)."""
    _PREFIX = "hts"
    _NAME = "HT-S"

    def last_symbol_count(self):
        return int(self._lib.dcvc_hts_last_symbol_count(self._h))


class DMCHTSBatchProxy(DMCHTSProxy):
    """n HT-S streams of one resolution coded as one batch (batching form; no
    reference counterpart: the reference holds one DMCHTSProxy per stream,
    dmc_hts_proxy.cpp:892-1000).  Every layer is one launch over the n
    pictures; each stream keeps its own QP, temporal context and bitstream,
    and gets the same bits as from a DMCHTSProxy of its own.  The streams move
    in lockstep: one chunk of every stream per compress_batch /
    decompress_batch call."""
    MAX_STREAMS = 16

    def __init__(self, n):
        if not 1 <= int(n) <= self.MAX_STREAMS:
            raise ValueError(f"1 to {self.MAX_STREAMS} streams per batch")
        super().__init__()
        self.n = int(n)

    def _prepare(self, Hp, Wp):
        if self._size != (Hp, Wp):
            self._frames = [torch.empty((self.n, 3, Hp, Wp), dtype=torch.float16, device=self._device,
                                        memory_format=torch.channels_last) for _ in range(self.FRAMES)]
            ptrs = (C.c_void_p * self.FRAMES)(*[f.data_ptr() for f in self._frames])
            _lib.call("dcvc_hts_prepare_batch", self._h, self.n, Hp, Wp, C.cast(ptrs, C.c_void_p), 3)
            self._size = (Hp, Wp)

    def add_ref_feature_from_frame(self, frames, apply_adaptor=True):
        """frames: [n,3,Hp,Wp], one decoded I frame per stream."""
        frames = _nhwc(frames)
        n, c, Hp, Wp = frames.shape
        if n != self.n or c != 3 or Hp % 16 or Wp % 16:
            raise ValueError(f"reference frames must be [{self.n},3,H,W] with H, W multiples of 16")
        self._prepare(Hp, Wp)
        self._call("add_ref_feature_from_frame", frames.data_ptr(), 3, Hp, Wp, int(bool(apply_adaptor)), _stream())

    def compress(self, *args, **kwargs):
        raise TypeError("a batch proxy codes all its streams at once: use compress_batch")

    def decompress(self, *args, **kwargs):
        raise TypeError("a batch proxy codes all its streams at once: use decompress_batch")

    def compress_batch(self, x, qps, reset_feature_memory, padding_b, padding_r):
        """x: [n,24,H,W], one chunk per stream; qps: n QPs.
        -> (n np.uint8 bitstreams, n ec_parallel values)"""
        x = _nhwc(x)
        n, c, H, W = x.shape
        if n != self.n or c != 24:
            raise ValueError(f"chunk input must be [{self.n},24,H,W]")
        if len(qps) != n:
            raise ValueError("one qp per stream")
        q = (C.c_int * n)(*[int(v) for v in qps])
        bits = (C.c_void_p * n)()
        nb = (C.c_long * n)()
        ec = (C.c_int * n)()
        _lib.call("dcvc_hts_compress_batch", self._h, x.data_ptr(), 24, H, W, q, int(bool(reset_feature_memory)),
                  int(padding_b), int(padding_r), _stream(), bits, nb, ec)
        return [_bits_to_numpy(C.c_void_p(bits[i]), C.c_long(nb[i])) for i in range(n)], list(ec)

    def decompress_batch(self, bit_streams, qps, height, width, entropy_coder_parallel, reset_feature_memory):
        """-> list of 8 tensors [n,3,Hp,Wp] (frame i of every stream)"""
        n = self.n
        if len(bit_streams) != n or len(qps) != n or len(entropy_coder_parallel) != n:
            raise ValueError("one bitstream, qp and ec_parallel per stream")
        arrs = [np.ascontiguousarray(np.asarray(b, dtype=np.uint8)) for b in bit_streams]
        ptrs = (C.c_void_p * n)(*[a.ctypes.data if a.size else None for a in arrs])
        sizes = (C.c_long * n)(*[a.size for a in arrs])
        q = (C.c_int * n)(*[int(v) for v in qps])
        ec = (C.c_int * n)(*[int(v) for v in entropy_coder_parallel])
        _lib.call("dcvc_hts_decompress_batch", self._h, ptrs, sizes, q, int(height), int(width), ec,
                  int(bool(reset_feature_memory)), _stream())
        return list(self._frames)

    def last_symbol_count(self, stream=0):
        return int(self._lib.dcvc_hts_batch_symbol_count(self._h, int(stream)))


class DMCHTLProxy(_ChunkProxy):
    """DMCHTLProxy (dmc_htl_proxy.h; add_ref_feature_from_frame
    dmc_htl_proxy.cpp:583-593, compress :595-716, decompress :718-915)."""
    _PREFIX = "htl"
    _NAME = "HT-L"

    def step_symbol_count(self, step):
        return int(self._lib.dcvc_htl_step_symbol_count(self._h, step))


class DMCLDProxy(_ChunkProxy):
    """DMCLDProxy (dmc_ld_proxy.h; add_ref_feature_from_frame
    dmc_ld_proxy.cpp:407-418, compress :420-485, decompress :487-593): the
    low-delay model, one frame per call; decompress returns one tensor."""
    FRAMES = 1
    _PREFIX = "ld"
    _NAME = "LD"

    def _prepare(self, Hp, Wp):
        if self._size != (Hp, Wp):
            self._frames = [torch.empty((1, 3, Hp, Wp), dtype=torch.float16, device=self._device,
                                        memory_format=torch.channels_last)]
            self._call("prepare", Hp, Wp, self._frames[0].data_ptr(), 3)
            self._size = (Hp, Wp)

    def compress(self, x, qp, reset_feature_memory, padding_b, padding_r):
        """-> (np.uint8 bitstream, ec_parallel); x: [1,3,H,W], one frame"""
        x = _nhwc(x)
        _, c, H, W = x.shape
        if c != 3:
            raise ValueError("LD input must be one frame [1,3,H,W]")
        bits = C.c_void_p()
        n = C.c_long()
        ec = C.c_int()
        self._call("compress", x.data_ptr(), 3, H, W, int(qp), int(bool(reset_feature_memory)), int(padding_b),
                   int(padding_r), _stream(), C.byref(bits), C.byref(n), C.byref(ec))
        return _bits_to_numpy(bits, n), ec.value

    def decompress(self, bit_stream, qp, height, width, entropy_coder_parallel, reset_feature_memory):
        """-> x_hat [1,3,Hp,Wp]"""
        return super().decompress(bit_stream, qp, height, width, entropy_coder_parallel, reset_feature_memory)[0]

    def last_symbol_count(self):
        return int(self._lib.dcvc_ld_last_symbol_count(self._h))
