"""GPU quantization of compact moment images (mray_hip_quantize /
mray_hip_dequantize) against the compiled reference's streams in
tests/golden/golden_quant.npz and against the oracle's restatement of
moment_quantization (src/moment_quantization.cpp:33-163, 244-360).

Bar: bit-identical streams and dequantized values; at full size, closed
forms that hold without the oracle (32-bit words are the clamped floats,
re-quantizing the decoded image gives the same stream)."""
import pathlib

import numpy as np
import pytest

from oracle_ffi import Oracle, quant_bits, stream_bytes
from test_oracle_golden import QUANT_CASES

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def hip():
    from mray_amd import hip as h

    h.init(0)
    return h


def dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run(hip, W, H, data, index, pc, bits):
    import torch

    q = hip.quantize(dev(data), dev(index), W, H, pc, bits)
    torch.cuda.synchronize()
    return q.cpu().numpy()


def random_image(rng, W, H, max_m, pc, zero_frac=0.2):
    counts = rng.integers(1, max_m + 1, W * H).astype(np.uint32)
    counts[rng.random(W * H) < zero_frac] = 0
    index = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    n = int(index[-1])
    lo, hi = (-0.05, 1.05) if pc else (-0.34, 0.34)
    data = rng.uniform(lo, hi, n).astype(np.float32)
    data[::29] = 0.0
    data[::31] = -0.0
    return counts, index, data


@pytest.mark.parametrize("case", QUANT_CASES)
def test_quantize_matches_golden(hip, case):
    g = dict(np.load(GOLD / "golden_quant.npz"))
    index, data, pc, bits = g[f"{case}_index"], g[f"{case}_data"], int(g[f"{case}_pc"]), g[f"{case}_bits"]
    W, H = 19, 11
    print(f"[quantize] golden case={case}", flush=True)
    stream = run(hip, W, H, data, index, pc, bits)
    assert np.array_equal(stream, g[f"{case}_stream"])
    dec = hip.dequantize(dev(stream), dev(index), W, H, pc, bits).cpu().numpy()
    expect = g[f"{case}_decoded"] if f"{case}_decoded" in g else Oracle().dequantize(W, H, index, pc, bits, stream)
    assert np.array_equal(dec.view(np.uint32), expect.view(np.uint32))


@pytest.mark.parametrize("pc, widths", [(1, 10), (1, "table"), (0, 8), (0, 13), (0, 1), (1, 32), (0, 31)])
def test_quantize_matches_oracle(hip, pc, widths):
    """Images of several tiles (the scan's tile sums and carries), pixels with
    up to 100 moments, the reference's widths and a descending table."""
    rng = np.random.default_rng(11)
    W, H = 173, 61
    counts, index, data = random_image(rng, W, H, 100, pc)
    if widths == "table":
        bits = quant_bits(256, pc, table=np.clip(16 - np.arange(100) // 8, 3, 16))
    else:
        bits = quant_bits(256, pc, widths)
    print(f"[quantize] oracle pc={pc} widths={widths} W={W} H={H} words={len(data)}", flush=True)
    O = Oracle()
    stream = run(hip, W, H, data, index, pc, bits)
    ref = O.quantize(W, H, data, index, pc, bits)
    assert len(stream) == stream_bytes(index, bits)
    assert np.array_equal(stream, ref)
    dec = hip.dequantize(dev(stream), dev(index), W, H, pc, bits).cpu().numpy()
    assert np.array_equal(dec.view(np.uint32), O.dequantize(W, H, index, pc, bits, ref).view(np.uint32))


def test_quantize_round_trip_full_size(hip):
    """1024^2 pixels of up to 50 moments (the mixing-layer image's size):
    decoding and re-encoding gives the same stream, and every decoded value
    is within one step of its clamped input."""
    import torch

    rng = np.random.default_rng(5)
    W = H = 1024
    counts, index, data = random_image(rng, W, H, 50, 1, zero_frac=0.05)
    bits = quant_bits(256, 1, 10)
    print(f"[quantize] round trip W={W} H={H} words={len(data)}", flush=True)
    d_index = dev(index)
    s1 = hip.quantize(dev(data), d_index, W, H, True, bits)
    dec = hip.dequantize(s1, d_index, W, H, True, bits)
    s2 = hip.quantize(dec, d_index, W, H, True, bits)
    torch.cuda.synchronize()
    assert s1.numel() == stream_bytes(index, bits)
    assert torch.equal(s1, s2)
    ls = torch.from_numpy(np.concatenate([np.arange(c) for c in counts]).astype(np.int64)).cuda()
    step = torch.where(ls == 0, 1.0 / 65535, 1.0 / 1023).to(torch.float32)
    err = (dec - dev(data).clamp(0, 1)).abs()
    # the last word of a stream that ends inside a byte loses its last bits
    assert bool((err[:-1] <= step[:-1] * 1.001).all())


def test_quantize_64bit_offsets(hip):
    """A stream beyond 2^32 bits (2048 x 1400 pixels of 100 moments at 32
    bits, 9.2 Gbit): the 64-bit offsets place every word, and 32-bit words
    are the clamped floats themselves."""
    import torch

    W, H, m = 2048, 1400, 100
    n = W * H * m
    print(f"[quantize] 64-bit offsets W={W} H={H} m={m} words={n}", flush=True)
    index = torch.arange(0, (W * H + 1) * m, m, dtype=torch.int64, device="cuda").to(torch.int32)
    g = torch.Generator(device="cuda").manual_seed(3)
    data = torch.rand(n, generator=g, device="cuda") * 1.2 - 0.1
    data[data.abs() < 1e-3] = 0.5
    bits = quant_bits(256, 1, table=[32] * 256)  # moment 0 too (the default table gives it 16 bits)
    s = hip.quantize(data, index, W, H, True, bits)
    torch.cuda.synchronize()
    assert s.numel() == n * 4
    assert torch.equal(s.view(torch.int32), data.clamp(0, 1).view(torch.int32))
    del s
    torch.cuda.empty_cache()


def test_quantize_empty_and_errors(hip):
    import torch

    W, H = 8, 4
    index = np.zeros(W * H + 1, np.uint32)
    bits = quant_bits(256, 1, 10)
    s = run(hip, W, H, np.zeros(1, np.float32), index, 1, bits)
    assert s.size == 0
    dec = hip.dequantize(torch.empty(0, dtype=torch.uint8, device="cuda"), dev(index), W, H, 1, bits)
    assert dec.numel() == 0
    counts = np.full(W * H, 5, np.uint32)
    index = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    data = dev(np.full(int(index[-1]), 0.5, np.float32))
    with pytest.raises(hip.MrayError, match="at most 4 moments"):
        hip.quantize(data, dev(index), W, H, True, bits[:4])
    with pytest.raises(hip.MrayError, match="widths must be 1..32"):
        hip.quantize(data, dev(index), W, H, True, [0] * 5)
    stream = hip.quantize(data, dev(index), W, H, True, bits)
    with pytest.raises(hip.MrayError, match="whole byte"):
        hip.dequantize(stream[:-1], dev(index), W, H, True, bits)


@pytest.mark.parametrize("case", ["ml", "mix", "mix64"])
def test_curve_table_matches_reference(hip, case):
    """The quantization-curve search on the GPU (mray_hip_quantization_table)
    picks the compiled reference's tables (tests/golden/golden_curve.npz) for
    start widths 8, 10 and 14."""
    g = dict(np.load(GOLD / "golden_curve.npz"))
    W, H, m = (int(v) for v in g[f"{case}_shape"])
    d_data, d_index = dev(g[f"{case}_data"]), dev(g[f"{case}_index"])
    d_params = dev(g[f"{case}_params"].view(np.int16))
    for sb in (8, 10, 14):
        print(f"[quantize] curve case={case} W={W} H={H} m={m} start_bits={sb}", flush=True)
        t = hip.curve_table(d_data, d_index, W, H, m, d_params, sb)
        assert t == g[f"{case}_table{sb}"].tolist()
