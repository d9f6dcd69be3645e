"""The allocation-free split GEMM (phpc_dgemm_split_device_noalloc) through the C ABI: the
same bits as the allocating entry point (phpc_dgemm_split_device, CRT method) on two calls
in a row, the second smaller; no device memory is taken; sizes beyond its static scratch
are refused with C left unchanged."""
import ctypes

import numpy as np
import pytest

from conftest import phpc

pytestmark = pytest.mark.gpu


def _noalloc(A, B, C, slices=0, target=1e-14):
    info = phpc.SplitInfo()
    m, k = A.shape
    n = B.shape[1]
    rc = phpc.lib().phpc_dgemm_split_device_noalloc(phpc.dptr(A), A.stride(0), phpc.dptr(B), B.stride(0),
                                                    phpc.dptr(C), C.stride(0), m, k, n, slices, target, None,
                                                    ctypes.byref(info))
    return rc, info


def _operands(m, k, n, dist, seed):
    import torch
    A = torch.empty((m, k), dtype=torch.float64, device="cuda")
    B = torch.empty((k, n), dtype=torch.float64, device="cuda")
    C0 = torch.empty((m, n), dtype=torch.float64, device="cuda")
    phpc.fill_device(A, dist, seed)
    phpc.fill_device(B, dist, seed + 1)
    phpc.fill_device(C0, phpc.DIST_UNIFORM, seed + 2)
    return A, B, C0


@pytest.mark.parametrize("slices", [0, 12])
def test_noalloc_equals_allocating_entry_point(slices):
    import torch
    # the largest product the scratch holds at 20 moduli, then a smaller ragged one
    for m, k, n, dist, seed in ((1024, 1024, 1024, phpc.DIST_UNIFORM, 1), (300, 257, 200, phpc.DIST_WIDE, 7)):
        print(f"noalloc m={m} k={k} n={n} dist={dist} slices={slices}", flush=True)
        A, B, C0 = _operands(m, k, n, dist, seed)
        want = C0.clone()
        ref = phpc.dgemm_split_device(A, B, want, slices=slices, target=1e-14, method=phpc.SPLIT_CRT)
        got = C0.clone()
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        rc, info = _noalloc(A, B, got, slices=slices)
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        assert rc == 0, phpc.lib().phpc_last_error_message()
        assert (info.slices, info.int_bits, info.method) == (ref.slices, ref.int_bits, 1)
        assert torch.equal(got.view(torch.int64), want.view(torch.int64))
        assert free1 >= free0, (free0 - free1)  # nothing allocated on the device


def test_noalloc_refuses_sizes_beyond_its_scratch():
    import torch
    A, B, C0 = _operands(1200, 1024, 1024, phpc.DIST_UNIFORM, 3)
    got = C0.clone()
    rc, _ = _noalloc(A, B, got)
    assert rc == phpc.PHPC_ERR_ARG
    assert "allocation-free" in phpc.lib().phpc_last_error_message().decode()
    torch.cuda.synchronize()
    assert torch.equal(got.view(torch.int64), C0.view(torch.int64))
    # with fewer moduli fixed by the caller the same product fits
    rc, info = _noalloc(A, B, got, slices=14)
    assert rc == 0 and info.slices == 14
