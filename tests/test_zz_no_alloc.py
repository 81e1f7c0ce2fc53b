"""The allocation-free compaction entry point (dcvc_conditional_index_noalloc)
against dcvc_conditional_index, which takes its block offsets from a
caller-allocated workspace.  Two calls in a row share the library's
per-GPU scratch array: the second, smaller call must not see the first
call's offsets."""
import ctypes

import numpy as np
import pytest
import torch

from dcvc_amd import _lib

MAX_N = 16384 * 4096


def test_noalloc_rejects_oversized_input_on_the_host():
    lib = _lib.load()
    buf = ctypes.create_string_buffer(16)
    p = ctypes.cast(buf, ctypes.c_void_p)
    # rejected before any device work: no GPU is touched
    assert lib.dcvc_conditional_index_noalloc(p, 2, p, MAX_N + 1, p, p, None) == 1
    assert "exceeds the maximum" in _lib.last_error()
    assert lib.dcvc_conditional_index_noalloc(p, 3, p, 10, p, p, None) == 1
    assert lib.dcvc_conditional_index_noalloc(None, 2, None, 10, None, None, None) == 1


@pytest.mark.gpu
def test_noalloc_matches_workspace_entry_point(gpu):
    from dcvc_amd import kernels
    r = np.random.Generator(np.random.PCG64(1234))
    # 2160p HT-S y size, then a smaller ragged size
    for call, (N, p) in enumerate([(8355840, 0.6), (100001, 0.3)]):
        for dt, lo, hi in ((np.int16, -32768, 32767), (np.uint8, 0, 255)):
            print(f"noalloc call {call}: N={N} dtype={dt.__name__}", flush=True)
            vals = torch.from_numpy(r.integers(lo, hi, size=N, dtype=dt)).to(gpu)
            cond = torch.from_numpy((r.random(N) < p).astype(np.uint8)).to(gpu)
            want, n_want, _ = kernels.conditional_index(vals, cond)
            out = torch.empty_like(vals)
            count = torch.empty(1, dtype=torch.int32, device=gpu)
            got, n_got = kernels.conditional_index_noalloc(vals, cond, out, count)
            torch.cuda.synchronize()
            assert n_got == n_want
            assert torch.equal(got, want)
            host = vals.cpu().numpy()[cond.cpu().numpy().astype(bool)]
            assert np.array_equal(got.cpu().numpy(), host)
