"""Entropy coding and moment-image files on the GPU, through the C ABI.

* mray_hip_entropy_encode / mray_hip_entropy_decode against the compiled
  reference's entropy_encode / entropy_decode (golden_entropy.npz) byte for
  byte, including a stream long enough for the model's count halving, and
  against the oracle's coder on larger images (1024 x 1024 with 50 moments,
  the reference eval's 10-bit setting); decode(encode(x)) is the quantized x;
* mray_hip_quantize_dense / mray_hip_dequantize_dense against the oracle's
  quantize / dequantize value for value;
* every moment-image file variant the reference writes (LZ4 sections,
  packed quantized moments, entropy-coded moments, the quantization table,
  dense images) written by the C++ host library byte for byte against the
  compiled reference's write_moment_image (golden_files.npz), in the order the
  reference wrote them (its quantization table is process state), and read
  back: the values its reader returns, or for the dense quantized LZ4 files
  it cannot read, the quantized values.
"""
import hashlib
import pathlib

import numpy as np
import pytest

from oracle_ffi import Oracle
from test_files_cpu import ENTROPY_CASES, file_image, halving_data, requantized

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
FILE_SPECS = {"c_f32": ("A", 0, 32, 0), "c_q10": ("A", 0, 10, 0), "c_lz4_f32": ("A", 1, 32, 0),
              "c_lz4_q10": ("A", 1, 10, 0), "c_ent8": ("A", 0, 8, 1), "c_lz4_ent10": ("A", 1, 10, 1),
              "c_eb_lz4": ("E", 1, 32, 0), "d_f32": ("D", 0, 32, 0), "d_q10_raw": ("D", 0, 10, 0),
              "d_lz4_f32": ("D", 1, 32, 0), "d_lz4_q12": ("D", 1, 12, 0), "d_lz4_q16": ("D", 1, 16, 0),
              "pc_q10": ("P", 0, 10, 0), "pc_lz4_q10": ("P", 1, 10, 0), "pc_tab_q10": ("P", 0, 10, 0),
              "pc_tab_lz4_ent10": ("P", 1, 10, 1), "c_tab_lz4_ent10": ("A", 1, 10, 1)}


@pytest.fixture(scope="module")
def hipmod():
    from mray_amd import hip

    hip.init(0)
    return hip


def dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("case", ENTROPY_CASES)
def test_entropy_matches_reference_golden(hipmod, case):
    ent = np.load(GOLD / "golden_entropy.npz")
    W, H, m, pc = (int(v) for v in ent[f"{case}_shape"])
    index, data, bits = ent[f"{case}_index"], ent[f"{case}_data"], ent[f"{case}_bits"]
    print(f"[entropy] {case} {W}x{H} m={m} bits={list(bits)}", flush=True)
    sec = hipmod.entropy_encode(dev(data), dev(index.view(np.int32)), W, H, m, pc, bits)
    assert np.array_equal(sec, ent[f"{case}_section"])
    dec = hipmod.entropy_decode(ent[f"{case}_section"], dev(index.view(np.int32)), W, H, m, pc, bits)
    assert np.array_equal(dec.cpu().numpy().view(np.uint32), ent[f"{case}_decoded"].view(np.uint32))


def test_entropy_model_halving(hipmod):
    ent = np.load(GOLD / "golden_entropy.npz")
    W, H, m, pc = (int(v) for v in ent["halving_shape"])
    index = (np.arange(W * H + 1) * 2).astype(np.uint32)
    data = halving_data(W, H, ent["halving_levels"], ent["halving_level"])
    print(f"[entropy] halving {W}x{H}", flush=True)
    sec = hipmod.entropy_encode(dev(data), dev(index.view(np.int32)), W, H, m, pc, ent["halving_bits"])
    assert len(sec) == int(ent["halving_section_len"])
    assert hashlib.sha256(sec.tobytes()).digest() == ent["halving_section_sha256"].tobytes()
    dec = hipmod.entropy_decode(sec, dev(index.view(np.int32)), W, H, m, pc, ent["halving_bits"]).cpu().numpy()
    level = ent["halving_level"]
    for k, v in enumerate(ent["halving_decoded_levels"]):
        if np.any(level == k):
            assert np.all(dec[1::2][level == k] == v)


@pytest.mark.parametrize("W,H,m,pc,width,empty", [(1024, 1024, 50, 0, 10, 0.1), (512, 384, 100, 1, 12, 0.3),
                                                  (300, 200, 255, 0, 14, 0.0), (257, 129, 3, 0, 1, 0.5)])
def test_entropy_matches_oracle(hipmod, W, H, m, pc, width, empty):
    rng = np.random.default_rng(W + m)
    counts = rng.integers(1, m + 1, W * H).astype(np.uint32)
    counts[rng.random(W * H) < empty] = 0
    index = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    n = int(index[-1])
    # moments decaying with their order, as generate's are
    l = np.concatenate([np.arange(c) for c in counts]).astype(np.float32)
    data = (rng.random(n) if pc else rng.normal(0.0, 0.3, n) / (1.0 + l)).astype(np.float32)
    bits = np.array([16] + [width] * (m - 1), np.uint8)
    print(f"[entropy] oracle {W}x{H} m={m} pc={pc} width={width}", flush=True)
    O = Oracle()
    ref = O.entropy_encode(W, H, m, data, index, pc, bits)
    sec = hipmod.entropy_encode(dev(data), dev(index.view(np.int32)), W, H, m, pc, bits)
    assert len(sec) == len(ref) and np.array_equal(sec, ref)
    dec = hipmod.entropy_decode(sec, dev(index.view(np.int32)), W, H, m, pc, bits).cpu().numpy()
    assert np.array_equal(dec.view(np.uint32), O.entropy_decode(W, H, m, ref, index, pc, bits).view(np.uint32))
    if n < 200000:
        assert np.array_equal(dec.view(np.uint32), requantized(O, data, index, pc, bits).view(np.uint32))


def test_entropy_rejects_bad_widths(hipmod):
    index = np.array([0, 2, 4], np.uint32)
    data = np.zeros(4, np.float32)
    with pytest.raises(Exception, match="widths"):
        hipmod.entropy_encode(dev(data), dev(index.view(np.int32)), 2, 1, 2, 0, [10, 15])
    with pytest.raises(Exception, match="at most"):
        hipmod.entropy_encode(dev(data), dev(index.view(np.int32)), 2, 1, 1, 0, [10])


@pytest.mark.parametrize("m,bits", [(12, 10), (7, 16), (20, 3), (5, 24), (9, 32)])
def test_dense_quantization_matches_oracle(hipmod, m, bits):
    import ctypes as C

    rng = np.random.default_rng(m * bits)
    npix = 333
    data = rng.normal(0.0, 0.2, npix * m).astype(np.float32)
    data[::m] = rng.uniform(-0.1, 1.1, npix)
    q = hipmod.quantize_dense(dev(data), m, bits).cpu().numpy()
    nb = (bits + 7) // 8
    assert q.size == npix * m * nb
    O = Oracle()
    lib = O.lib
    lib.orc_quantize_value.restype = C.c_uint32
    lib.orc_quantize_value.argtypes = [C.c_float, C.c_int, C.c_float, C.c_float]
    inv_pi = float(np.float32(1.0 / np.pi))
    words = np.zeros(npix * m, np.uint32)
    for i in range(npix * m):
        lo, hi = (0.0, 1.0) if i % m == 0 else (-inv_pi, inv_pi)
        words[i] = lib.orc_quantize_value(float(data[i]), bits, lo, hi)
    got = np.zeros(npix * m, np.uint32)
    for b in range(nb):
        got |= q[b::nb].astype(np.uint32) << (8 * b)
    assert np.array_equal(got, words)
    back = hipmod.dequantize_dense(dev(q), npix * m, m, bits).cpu().numpy()
    lib.orc_dequantize_value.restype = C.c_float
    lib.orc_dequantize_value.argtypes = [C.c_uint32, C.c_int, C.c_float, C.c_float]
    exp = np.array([lib.orc_dequantize_value(int(words[i]), bits, *((0.0, 1.0) if i % m == 0 else (-inv_pi, inv_pi)))
                    for i in range(npix * m)], np.float32)
    assert np.array_equal(back.view(np.uint32), exp.view(np.uint32))


def test_files_match_reference(hipmod, tmp_path):
    """Every file variant, written in the reference's order by one process."""
    from mray_amd import moment_file

    files = np.load(GOLD / "golden_files.npz")
    order = [str(s) for s in files["order"]]
    assert order == list(FILE_SPECS)
    assert moment_file.quantization_table() is None, "the quantization table was set by an earlier test"
    O = Oracle()
    for name in order:
        key, compress, bits, entropy = FILE_SPECS[name]
        img = file_image(files, key)
        table = files["table"] if name.startswith("pc_tab") and moment_file.quantization_table() is None else None
        print(f"[files] {name}: compress={compress} bits={bits} entropy={entropy}", flush=True)
        path = tmp_path / f"{name}.bin"
        moment_file.write_moment_image(path, img, compress, bits, entropy, table=table)
        got = np.frombuffer(path.read_bytes(), np.uint8)
        want = files[f"{name}_file"]
        assert len(got) == len(want), (name, len(got), len(want))
        assert np.array_equal(got, want), name
        back = moment_file.load_moment_image(path)
        if f"{name}_read" in files.files:
            assert np.array_equal(back["data"].view(np.uint32), files[f"{name}_read"].view(np.uint32)), name
        else:
            # dense quantized LZ4: the quantized values, value for value
            m = img["m"]
            exp = requantized(O, img["data"], (np.arange(img["W"] * img["H"] + 1) * m).astype(np.uint32), 0,
                              np.full(m, bits, np.uint8))
            assert np.array_equal(back["data"].view(np.uint32), exp.view(np.uint32)), name
    assert np.array_equal(moment_file.quantization_table()[:len(files["table"])], files["table"])
