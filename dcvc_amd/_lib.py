"""ctypes binding of libdcvc_amd.so (the C ABI in include/dcvc_ops.h and
include/dcvc_codec.h).  Loading fails loudly: there is no CPU fallback."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libdcvc_amd.so"

_lib = None

P = C.c_void_p
I = C.c_int
L = C.c_long
F = C.c_float

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "dcvc_version": [],
    "dcvc_scale_index_thresholds": [P],
    "dcvc_gemm_set_tile_policy": [I],
    "dcvc_gemm_tune_set": [P, I],
    "dcvc_gemm_tile_for": [I, I, I, I, L, I, I],
    "dcvc_d3x3_tune_set": [P, I],
    "dcvc_d3x3_strip_for": [I, I, I, I],
    "dcvc_mfma_calibrate": [P, P, I, I, I, P],
    "dcvc_conv1x1_bias": [P, I, I, I, P, I, P, P, I, P],
    "dcvc_conv1x1_bias_wsilu": [P, I, I, I, P, I, P, P, I, P],
    "dcvc_conv1x1_bias_shortcut": [P, I, I, I, P, I, P, P, I, P, I, P],
    "dcvc_conv1x1_bias_shortcut2": [P, I, I, I, P, I, P, P, I, P, I, P, I, P],
    "dcvc_conv1x1_bias_shortcut_with_quant": [P, I, I, I, P, I, P, P, I, P, I, P, P, I, P],
    "dcvc_conv1x1_bias_shortcut_then_quant": [P, I, I, I, P, I, P, P, I, P, I, P, P, I, P],
    "dcvc_conv1x1_bias_with_quant": [P, I, I, I, P, I, P, P, I, P, P, I, P],
    "dcvc_conv1x1_bias_wsilu_chunk_add": [P, I, I, I, P, I, P, P, I, P],
    "dcvc_depth_conv_ffn": [P, I, I, I, I, P, P, P, P, I, P, I, P, I, P, I, P, I, P],
    "dcvc_set_ffn_fusion": [I],
    "dcvc_conv1x1_grouped": [P, I, L, I, I, I, P, L, I, P, L, I, P, I, L, P, I, L, P, I, L, I, P],
    "dcvc_conv_bias": [P, I, I, I, I, I, P, I, P, I, I, P, I, P],
    "dcvc_weights_to_lowp": [P, I, I, I, P, P, P],
    "dcvc_gemm_lowp": [P, I, I, I, P, I, P, I, I, P, P, I, P, I, P, I, P, P, I, P],
    "dcvc_gemm_fp8_act": [P, I, I, I, I, P, I, P, I, I, P, P, I, P, I, P, I, P, P, I, P, I, P],
    "dcvc_d3x3_fp8_out": [P, I, I, I, I, I, P, P, I, P, I, P],
    "dcvc_set_fp8_activations": [I],
    "dcvc_get_fp8_activations": [],
    "dcvc_transposed_conv": [P, I, I, I, I, I, P, I, P, P, I, P],
    "dcvc_subpel_conv3x3": [P, I, I, I, I, I, P, I, P, P, I, P],
    "dcvc_d3x3": [P, I, I, I, I, I, P, P, I, P],
    "dcvc_pad_and_unshuffle": [P, I, I, I, I, I, I, I, I, P, I, P],
    "dcvc_pixel_shuffle_8": [P, I, I, I, I, I, P, I, P],
    "dcvc_multiply_with_broadcast": [P, I, L, I, P, I, P, P, I, P],
    "dcvc_round_z": [P, I, L, I, P, I, P, P],
    "dcvc_int8_to_half": [P, L, I, P, I, P],
    "dcvc_replicate_pad": [P, I, I, I, I, I, I, I, P, I, P],
    "dcvc_hts_quant_step": [P, I, I, I, I, I, P, I, P, I, P, I, P, I, P, P, P, F, I, P],
    "dcvc_hts_build_index_dec": [P, I, I, I, I, I, P, F, P, P, P],
    "dcvc_hts_restore_step": [P, I, I, I, I, P, I, P, I, P, I, I, P],
    "dcvc_intra_quant_step": [P, I, I, I, I, I, P, I, P, I, P, I, P, I, P, P, I, P, P, P, F, I, P],
    "dcvc_intra_build_index_dec": [P, I, I, I, I, I, P, F, I, P, P, P],
    "dcvc_intra_restore_step": [P, I, I, I, I, P, I, P, I, P, I, P, P, I, I, P],
    "dcvc_ld_quant_step": [P, I, I, I, I, I, P, I, P, I, P, I, P, I, P, P, P, F, I, P],
    "dcvc_ld_restore_step": [P, I, I, I, I, P, I, P, I, P, I, I, P],
    "dcvc_htl_quant_step": [P, I, I, I, I, I, P, I, P, I, P, I, P, I, P, I, P, P, P, F, I, P],
    "dcvc_htl_restore_step": [P, I, I, I, I, P, I, P, I, P, I, P, I, I, P],
    "dcvc_compact_blocks": [L],
    "dcvc_conditional_index": [P, I, P, L, P, P, P, P],
    "dcvc_conditional_index_noalloc": [P, I, P, L, P, P, P],
    "dcvc_conditional_recover": [P, P, L, P, P, P],
    # entropy coder
    "dcvc_last_error": [],
    "dcvc_pmf_to_quantized_cdf": [P, I, P],
    "dcvc_compute_ec_parallel": [L],
    "dcvc_rans_encoder_create": [],
    "dcvc_rans_encoder_destroy": [P],
    "dcvc_rans_encoder_set_cdf": [P, P, I, I, P, I],
    "dcvc_rans_encoder_reset": [P],
    "dcvc_rans_encoder_set_parallel": [P, I],
    "dcvc_rans_encoder_encode_y": [P, P, L],
    "dcvc_rans_encoder_encode_z": [P, P, L, I, I],
    "dcvc_rans_encoder_flush": [P],
    "dcvc_rans_encoder_stream": [P, P, P],
    "dcvc_rans_decoder_create": [],
    "dcvc_rans_decoder_destroy": [P],
    "dcvc_rans_decoder_set_cdf": [P, P, I, I, P, I],
    "dcvc_rans_decoder_set_parallel": [P, I],
    "dcvc_rans_decoder_set_stream": [P, P, L],
    "dcvc_rans_decoder_decode_z": [P, L, I, I, P],
    "dcvc_rans_decoder_decode_y": [P, P, L, P],
    # codec proxies
    "dcvc_dmci_create": [],
    "dcvc_dmci_destroy": [P],
    "dcvc_dmci_set_param": [P, I, P, P, P, P, F],
    "dcvc_dmci_prepare": [P, I, I, P, I],
    "dcvc_dmci_compress": [P, P, I, I, I, I, I, I, P, P, P, P],
    "dcvc_dmci_decompress": [P, P, L, I, I, I, I, P],
    "dcvc_dmci_set_use_graphs": [P, I],
    "dcvc_dmci_set_precision": [P, I],
    "dcvc_dmci_step_symbol_count": [P, I],
    "dcvc_dmci_debug_copy": [P, I, P, L],
    "dcvc_hts_create": [],
    "dcvc_hts_destroy": [P],
    "dcvc_hts_set_param": [P, I, P, P, P, P, F],
    "dcvc_hts_share_weights": [P, P],
    "dcvc_hts_prepare": [P, I, I, P, I],
    "dcvc_hts_add_ref_feature_from_frame": [P, P, I, I, I, I, P],
    "dcvc_hts_compress": [P, P, I, I, I, I, I, I, I, P, P, P, P],
    "dcvc_hts_decompress": [P, P, L, I, I, I, I, I, P],
    "dcvc_hts_set_use_graphs": [P, I],
    "dcvc_hts_set_precision": [P, I],
    "dcvc_hts_device_bytes": [P],
    "dcvc_hts_debug_copy": [P, I, P, L],
    "dcvc_hts_last_symbol_count": [P],
    "dcvc_hts_prepare_batch": [P, I, I, I, P, I],
    "dcvc_hts_compress_batch": [P, P, I, I, I, P, I, I, I, P, P, P, P],
    "dcvc_hts_decompress_batch": [P, P, P, P, I, I, P, I, P],
    "dcvc_hts_batch_symbol_count": [P, I],
    "dcvc_htl_create": [],
    "dcvc_htl_destroy": [P],
    "dcvc_htl_set_param": [P, I, P, P, P, P, F],
    "dcvc_htl_share_weights": [P, P],
    "dcvc_htl_prepare": [P, I, I, P, I],
    "dcvc_htl_add_ref_feature_from_frame": [P, P, I, I, I, I, P],
    "dcvc_htl_compress": [P, P, I, I, I, I, I, I, I, P, P, P, P],
    "dcvc_htl_decompress": [P, P, L, I, I, I, I, I, P],
    "dcvc_htl_set_use_graphs": [P, I],
    "dcvc_htl_set_precision": [P, I],
    "dcvc_htl_device_bytes": [P],
    "dcvc_htl_debug_copy": [P, I, P, L],
    "dcvc_htl_step_symbol_count": [P, I],
    "dcvc_ld_create": [],
    "dcvc_ld_destroy": [P],
    "dcvc_ld_set_param": [P, I, P, P, P, P, F],
    "dcvc_ld_share_weights": [P, P],
    "dcvc_ld_prepare": [P, I, I, P, I],
    "dcvc_ld_add_ref_feature_from_frame": [P, P, I, I, I, I, P],
    "dcvc_ld_compress": [P, P, I, I, I, I, I, I, I, P, P, P, P],
    "dcvc_ld_decompress": [P, P, L, I, I, I, I, I, P],
    "dcvc_ld_set_use_graphs": [P, I],
    "dcvc_ld_set_precision": [P, I],
    "dcvc_ld_device_bytes": [P],
    "dcvc_ld_debug_copy": [P, I, P, L],
    "dcvc_ld_last_symbol_count": [P],
    "dcvc_trace_begin": [I],
    "dcvc_trace_end": [P, I],
}
_RESTYPE = {"dcvc_last_error": C.c_char_p, "dcvc_rans_encoder_create": C.c_void_p,
            "dcvc_rans_decoder_create": C.c_void_p, "dcvc_rans_encoder_destroy": None,
            "dcvc_rans_decoder_destroy": None, "dcvc_dmci_create": C.c_void_p, "dcvc_hts_create": C.c_void_p,
            "dcvc_dmci_destroy": None, "dcvc_hts_destroy": None, "dcvc_hts_last_symbol_count": C.c_long,
            "dcvc_hts_device_bytes": C.c_long, "dcvc_hts_batch_symbol_count": C.c_long,
            "dcvc_dmci_step_symbol_count": C.c_long, "dcvc_htl_create": C.c_void_p, "dcvc_htl_destroy": None,
            "dcvc_htl_device_bytes": C.c_long, "dcvc_htl_step_symbol_count": C.c_long,
            "dcvc_ld_create": C.c_void_p, "dcvc_ld_destroy": None, "dcvc_ld_device_bytes": C.c_long,
            "dcvc_ld_last_symbol_count": C.c_long}


def declared_symbols():
    """Every function name declared in include/*.h (parsed from the headers)."""
    import re
    names = []
    for h in sorted((_PKG.parent / "include").glob("*.h")):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"\b(dcvc_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def load(path: str | os.PathLike | None = None):
    """Load the library (once) and attach argtypes.  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run `python -m dcvc_amd.build` (no CPU fallback exists)")
    lib = C.CDLL(str(p), mode=C.RTLD_GLOBAL)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


class DcvcError(RuntimeError):
    CODES = {1: "invalid argument", 2: "HIP runtime failure", 3: "malformed bitstream",
             4: "entropy coder overflow", 5: "call order violated", 6: "byte budget exceeded"}

    def __init__(self, fn, code):
        super().__init__(f"{fn} failed: {self.CODES.get(code, 'error')} (status {code})")
        self.code = code


def last_error():
    m = load().dcvc_last_error()
    return m.decode() if m else ""


def call(name, *args):
    lib = load()
    r = getattr(lib, name)(*args)
    if r != 0:
        err = DcvcError(name, r)
        msg = last_error()
        if msg:
            err.args = (f"{err.args[0]}: {msg}",)
        raise err
    return r
