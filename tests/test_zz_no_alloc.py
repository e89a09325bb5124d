"""The allocation-free scan (dmgs_inclusive_scan_u32_noalloc) gives the
same results as dmgs_inclusive_scan_u32 with a caller workspace, through the
C-ABI, on two calls in a row (the second smaller, so it reuses the static
scratch the first one left), and rejects sizes beyond its maximum."""
import numpy as np
import pytest
import torch

from dmgs_amd import _native as N

pytestmark = pytest.mark.gpu


def _scan_with_workspace(lib, x):
    n = x.numel()
    out = torch.empty_like(x)
    ws = torch.empty(lib.dmgs_scan_scratch_size(n), dtype=torch.uint8, device=x.device)
    N.check(lib.dmgs_inclusive_scan_u32(N.ptr(x), N.ptr(out), n, N.ptr(ws), N.stream_handle()), "scan")
    return out


def test_noalloc_scan_matches_workspace_scan():
    lib = N.load()
    rng = np.random.default_rng(17)
    # large first (many block sums), then a smaller input that reuses the scratch
    for n in (5_000_003, 70_001, 4096, 1):
        print(f"case: allocation-free scan n={n}", flush=True)
        host = rng.integers(0, 60, size=n, dtype=np.uint32)
        x = torch.from_numpy(host.view(np.int32)).cuda()
        ref = _scan_with_workspace(lib, x)
        out = torch.full_like(x, -1)
        N.check(lib.dmgs_inclusive_scan_u32_noalloc(N.ptr(x), N.ptr(out), n, N.stream_handle()), "noalloc scan")
        torch.cuda.synchronize()
        assert torch.equal(out, ref), f"n={n}: allocation-free scan differs from the workspace scan"
        assert np.array_equal(out.cpu().numpy().view(np.uint32), np.cumsum(host, dtype=np.uint64).astype(np.uint32))
    # in place, as the workspace scan allows
    host = rng.integers(0, 9, size=123_457, dtype=np.uint32)
    x = torch.from_numpy(host.view(np.int32)).cuda()
    ref = _scan_with_workspace(lib, x)
    N.check(lib.dmgs_inclusive_scan_u32_noalloc(N.ptr(x), N.ptr(x), x.numel(), N.stream_handle()), "in place")
    torch.cuda.synchronize()
    assert torch.equal(x, ref)
    # beyond the maximum: rejected before any launch, with the limit in the message
    assert lib.dmgs_inclusive_scan_u32_noalloc(N.ptr(x), N.ptr(x), (1 << 30) + 1, N.stream_handle()) == -1
    assert "maximum" in N.last_error()
