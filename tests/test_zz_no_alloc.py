"""mray_hip_moment_index_noalloc (the allocation-free index scan) against
mray_hip_moment_index through the C ABI: identical index arrays (integer
work, bit-exact) on two calls in a row, the second with a smaller image, so
the shared scratch array keeps the first call's contents when the second
runs; and the size limit is an error, not a launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_index_matches_allocating_entry_point():
    import torch

    from mray_amd import hip

    hip.init(0)
    lib = hip.load_library()
    stream = hip._stream(None)
    gen = torch.Generator(device="cpu").manual_seed(11)
    for W, H in ((1500, 900), (37, 23)):
        print(f"[noalloc] index scan W={W} H={H}", flush=True)
        counts = torch.randint(0, 101, (W * H,), dtype=torch.int32, generator=gen).cuda()
        ref = torch.empty(W * H + 1, dtype=torch.int32, device="cuda")
        got = torch.full((W * H + 1,), -1, dtype=torch.int32, device="cuda")
        assert lib.mray_hip_moment_index(hip._ptr(counts), W, H, hip._ptr(ref), stream) == 0
        assert lib.mray_hip_moment_index_noalloc(hip._ptr(counts), W, H, hip._ptr(got), stream) == 0
        torch.cuda.synchronize()
        r, g = ref.cpu().numpy().view(np.uint32), got.cpu().numpy().view(np.uint32)
        assert np.array_equal(r, g)
        c = counts.cpu().numpy().astype(np.uint64)
        assert np.array_equal(g.astype(np.uint64), np.concatenate(([0], np.cumsum(c))))
    # in place (counts and index the same array, as generate's compaction uses it)
    W, H = 640, 480
    buf = torch.randint(0, 50, (W * H + 1,), dtype=torch.int32, generator=gen).cuda()
    expect = np.concatenate(([0], np.cumsum(buf.cpu().numpy()[:-1].astype(np.uint64))))
    assert lib.mray_hip_moment_index_noalloc(hip._ptr(buf), W, H, hip._ptr(buf), stream) == 0
    torch.cuda.synchronize()
    assert np.array_equal(buf.cpu().numpy().view(np.uint32).astype(np.uint64), expect)


def test_noalloc_index_rejects_sizes_beyond_its_scratch():
    import torch

    from mray_amd import hip

    hip.init(0)
    lib = hip.load_library()
    small = torch.zeros(16, dtype=torch.int32, device="cuda")
    # 8193 x 8192 pixels exceed the 67,108,864 the scratch array covers: the
    # host rejects the size before any launch
    rc = lib.mray_hip_moment_index_noalloc(hip._ptr(small), 8193, 8192, hip._ptr(small), hip._stream(None))
    assert rc == -1
    assert b"scratch limit" in lib.mray_hip_last_error()
