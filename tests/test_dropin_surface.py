"""The drop-in surface on the CPU: after install_module_alias(), the module
names the reference model code imports resolve to this package, and the
four proxy classes expose the methods pybind binds (bind.cpp:11-39) with the
arity of the reference's C++ declarations (dmci_proxy.h:142-145,
dmc_hts_proxy.h:229-236, dmc_htl_proxy.h:239-246, dmc_ld_proxy.h:216-222).
No GPU call is made: classes are inspected, not constructed."""
import importlib
import inspect
import sys

# class -> method -> parameter names after self, from the reference headers
REFERENCE_SURFACE = {
    "DMCIProxy": {
        "set_param": ["state_dict", "skip_threshold"],
        "compress": ["x", "qp", "padding_b", "padding_r"],
        "decompress": ["bit_stream", "qp", "height", "width", "entropy_coder_parallel"],
    },
    "DMCHTSProxy": {
        "set_param": ["state_dict", "skip_threshold"],
        "add_ref_feature_from_frame": ["frame", "apply_adaptor"],
        "compress": ["x", "qp", "reset_feature_memory", "padding_b", "padding_r"],
        "decompress": ["bit_stream", "qp", "height", "width", "entropy_coder_parallel", "reset_feature_memory"],
    },
}
REFERENCE_SURFACE["DMCHTLProxy"] = REFERENCE_SURFACE["DMCHTSProxy"]
REFERENCE_SURFACE["DMCLDProxy"] = REFERENCE_SURFACE["DMCHTSProxy"]

# MLCodec_extensions_cpp (src/cpp/py_rans/bind.cpp:14-40)
CODER_SURFACE = {
    "RansEncoder": ["encode_y", "encode_z", "flush", "get_encoded_stream", "reset", "set_cdf",
                    "set_entropy_coder_parallel"],
    "RansDecoder": ["set_stream", "decode_y", "decode_z", "set_cdf", "set_entropy_coder_parallel"],
}


def test_install_module_alias_resolves_reference_imports():
    saved = {k: sys.modules.pop(k, None) for k in ("inference_extensions_cuda", "MLCodec_extensions_cpp")}
    try:
        import dcvc_amd
        m = dcvc_amd.install_module_alias()
        ext = importlib.import_module("inference_extensions_cuda")
        assert ext is m
        from inference_extensions_cuda import DMCHTLProxy, DMCHTSProxy, DMCIProxy, DMCLDProxy  # noqa: F401
        coder = importlib.import_module("MLCodec_extensions_cpp")
        from dcvc_amd import entropy
        assert coder is entropy
    finally:
        for k, v in saved.items():
            if v is None:
                sys.modules.pop(k, None)
            else:
                sys.modules[k] = v


def test_proxy_classes_match_the_reference_bindings():
    from dcvc_amd import inference_extensions_cuda as ext
    for cls_name, methods in REFERENCE_SURFACE.items():
        cls = getattr(ext, cls_name)
        assert inspect.signature(cls).parameters == {} or not [
            p for p in inspect.signature(cls).parameters.values() if p.default is inspect.Parameter.empty], cls_name
        for meth, params in methods.items():
            fn = getattr(cls, meth)
            got = [p for p in inspect.signature(fn).parameters.values()][1:]
            required = [p for p in got if p.default is inspect.Parameter.empty]
            assert len(required) <= len(params) <= len(got), (cls_name, meth, [p.name for p in got])


def test_entropy_module_matches_the_coder_bindings():
    from dcvc_amd import entropy
    for cls_name, methods in CODER_SURFACE.items():
        cls = getattr(entropy, cls_name)
        missing = [m for m in methods if not callable(getattr(cls, m, None))]
        assert not missing, (cls_name, missing)
    assert callable(entropy.pmf_to_quantized_cdf)
