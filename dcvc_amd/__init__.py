"""MI355X-native DCVC-UF inference hot path.

``dcvc_amd.inference_extensions_cuda`` is the drop-in for the reference's
CUDA proxy module; ``dcvc_amd.entropy`` is the CPU entropy coder and CDF
tables; ``dcvc_amd.kernels`` exposes the individual HIP kernels.
"""
import sys

__version__ = "0.1.0"


def install_module_alias():
    """Make ``import inference_extensions_cuda`` and ``import
    MLCodec_extensions_cpp`` resolve to this package's drop-in modules, the
    names the reference model code imports (image_model.py:197,
    video_model_ht.py:419-424, entropy_models.py:34-40)."""
    from . import entropy, inference_extensions_cuda as m
    sys.modules.setdefault("inference_extensions_cuda", m)
    sys.modules.setdefault("MLCodec_extensions_cpp", entropy)
    return m
