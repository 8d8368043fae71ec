"""trax_prefix_sum_noalloc (scratch in a fixed device array, no allocation)
against trax_prefix_sum through the C ABI: two calls in a row, the second
with a smaller input, and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _scan_both(ctx, a):
    from trax_amd import runtime
    want = torch.from_numpy(a.view(np.int32).copy()).cuda()
    runtime.prefix_sum(ctx, want)
    got = torch.from_numpy(a.view(np.int32).copy()).cuda()
    rc = ctx.lib.trax_prefix_sum_noalloc(C.c_void_p(got.data_ptr()), got.numel(), runtime.stream_ptr())
    assert rc == 0
    torch.cuda.synchronize()
    return got.cpu().numpy().view(np.uint32), want.cpu().numpy().view(np.uint32)


def test_noalloc_scan_equals_the_context_scan(ctx):
    rng = np.random.default_rng(17)
    # large-array form (chunk partials in the fixed scratch), then a smaller
    # single-pass input reusing the same scratch
    for n in (5_000_003, 300_001, 1, 8192 * 3):
        a = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        print(f"no-alloc scan n={n}", flush=True)
        got, want = _scan_both(ctx, a)
        assert np.array_equal(got, want), n
        assert np.array_equal(want[: min(n, 1000)], np.concatenate([[0], np.cumsum(a[: min(n, 1000) - 1],
                                                                                    dtype=np.uint32)]))


def test_noalloc_scan_rejects_sizes_beyond_its_scratch(ctx):
    from trax_amd import runtime
    t = torch.zeros(64, dtype=torch.int32, device="cuda")
    rc = ctx.lib.trax_prefix_sum_noalloc(C.c_void_p(t.data_ptr()), (1 << 30) + 1, runtime.stream_ptr())
    assert rc == -1
    torch.cuda.synchronize()
    assert int(t.abs().sum()) == 0  # nothing was enqueued
    assert ctx.lib.trax_prefix_sum_noalloc(None, 5, runtime.stream_ptr()) == -1
