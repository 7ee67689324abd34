"""The allocation-free scan entry point (mfem_scan_i32_noalloc, for callers
that cannot allocate device memory) against the workspace entry point
mfem_scan_i32 through the C ABI: bit-exact (integer work) on two calls in a
row, the second with a smaller input, so the shared scratch's leftover
contents from the first call must not leak into the second; sizes above the
maximum are rejected with an error and launch nothing."""
import ctypes

import numpy as np
import pytest
import torch

from mfem_amd import _ffi

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _scan_ws(lib, a, inclusive, dev):
    n = a.numel()
    out = torch.empty_like(a)
    tot = torch.zeros(1, dtype=torch.int32, device=dev)
    nb = int(lib.mfem_scan_workspace_bytes(n))
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
    rc = lib.mfem_scan_i32(_p(a), _p(out), n, int(inclusive), _p(tot), _p(ws), ctypes.c_size_t(nb), _stream(dev))
    assert rc == 0, lib.mfem_collide_last_error()
    return out, tot


def _scan_noalloc(lib, a, inclusive, dev, out=None, tot=None):
    n = a.numel()
    out = torch.full_like(a, -7) if out is None else out
    tot = torch.full((1,), -7, dtype=torch.int32, device=dev) if tot is None else tot
    rc = lib.mfem_scan_i32_noalloc(_p(a), _p(out), n, int(inclusive), _p(tot), _stream(dev))
    assert rc == 0, lib.mfem_collide_last_error()
    return out, tot


@pytest.mark.parametrize("inclusive", [False, True])
def test_noalloc_scan_matches_workspace_scan_twice(gpu, inclusive):
    lib = _ffi.load()
    rng = np.random.default_rng(5 + int(inclusive))
    # first call: near the maximum (two levels of tile sums in the scratch);
    # second call: smaller (one level), after the first left its sums there
    for n in (_ffi.MFEM_SCAN_NOALLOC_MAX - 12345, 3 * 1024 * 1024 + 17, 5000, 1000):
        a = torch.from_numpy(rng.integers(0, 7, n).astype(np.int32)).to(gpu)
        ref, rtot = _scan_ws(lib, a, inclusive, gpu)
        out, tot = _scan_noalloc(lib, a, inclusive, gpu)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), n
        assert int(tot.item()) == int(rtot.item()) == int(a.sum().item()), n
        print("n", n, "total", int(tot.item()), flush=True)


def test_noalloc_scan_rejects_sizes_above_the_maximum(gpu):
    lib = _ffi.load()
    n = _ffi.MFEM_SCAN_NOALLOC_MAX + 1
    # pointers are never dereferenced: the size check comes first
    a = torch.zeros(8, dtype=torch.int32, device=gpu)
    rc = lib.mfem_scan_i32_noalloc(_p(a), _p(a), n, 0, None, _stream(gpu))
    assert rc == -3  # MFEM_ECAPACITY
    assert b"exceeds the maximum" in lib.mfem_collide_last_error()
    assert lib.mfem_scan_i32_noalloc(_p(a), _p(a), -1, 0, None, _stream(gpu)) == -1


def test_noalloc_scan_empty_and_single_tile(gpu):
    lib = _ffi.load()
    a = torch.zeros(0, dtype=torch.int32, device=gpu)
    tot = torch.full((1,), -7, dtype=torch.int32, device=gpu)
    assert lib.mfem_scan_i32_noalloc(_p(a), _p(a), 0, 0, _p(tot), _stream(gpu)) == 0
    torch.cuda.synchronize()
    assert int(tot.item()) == 0
    b = torch.arange(1, 1025, dtype=torch.int32, device=gpu)
    out, t = _scan_noalloc(lib, b, False, gpu)
    torch.cuda.synchronize()
    assert torch.equal(out.cpu(), torch.from_numpy(np.cumsum(np.arange(1, 1025)) - np.arange(1, 1025)).int())
    assert int(t.item()) == 1024 * 1025 // 2
