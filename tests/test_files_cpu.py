"""Entropy coding and moment-image files without a GPU.

* the oracle's FastAC coder (orc_entropy_encode / orc_entropy_decode)
  against the compiled reference's entropy_encode / entropy_decode sections
  and decoded values (golden_entropy.npz), including a stream long enough for
  the adaptive model's count halving, and live against the compiled reference
  where it is built;
* the round trip the coder must satisfy whatever its bytes: decode(encode(x))
  is quantize-then-dequantize of x, value for value;
* the host library's files of float moments (LZ4 sections, error bounds,
  dense images; no quantization, so no GPU) byte for byte against the files
  the compiled reference's write_moment_image wrote (golden_files.npz), and
  what its reader returns for them.
"""
import ctypes as C
import hashlib
import pathlib

import numpy as np
import pytest

from oracle_ffi import REF_SO, Oracle, Reference

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
ENTROPY_CASES = ["plain8", "pc_table", "w14", "w1", "pc_w32_zeroth"]
# files of float moments: no quantization or entropy coding
FLOAT_FILES = ["c_f32", "c_lz4_f32", "c_eb_lz4", "d_f32", "d_q10_raw", "d_lz4_f32"]


@pytest.fixture(scope="module")
def ent():
    return np.load(GOLD / "golden_entropy.npz")


@pytest.fixture(scope="module")
def files():
    return np.load(GOLD / "golden_files.npz")


def halving_data(W, H, levels, level):
    data = np.empty(2 * W * H, np.float32)
    data[0::2] = (np.arange(W * H) % 97) / np.float32(97.0)
    data[1::2] = levels[level]
    return data


def requantized(O, data, index, pc, bits):
    """dequantize(quantize(x)) of every value of a compact image, with the
    16-bit truncation of moment 0 the section's u16 words apply."""
    out = np.empty_like(data)
    lib = O.lib
    lib.orc_quantize_value.restype = C.c_uint32
    lib.orc_quantize_value.argtypes = [C.c_float, C.c_int, C.c_float, C.c_float]
    lib.orc_dequantize_value.restype = C.c_float
    lib.orc_dequantize_value.argtypes = [C.c_uint32, C.c_int, C.c_float, C.c_float]
    inv_pi = float(np.float32(1.0 / np.pi))
    for p in range(len(index) - 1):
        for l in range(int(index[p + 1] - index[p])):
            lo, hi = (0.0, 1.0) if (l == 0 or pc) else (-inv_pi, inv_pi)
            q = lib.orc_quantize_value(float(data[index[p] + l]), int(bits[l]), lo, hi)
            if l == 0:
                q &= 0xFFFF
            out[index[p] + l] = lib.orc_dequantize_value(q, int(bits[l]), lo, hi)
    return out


@pytest.mark.parametrize("case", ENTROPY_CASES)
def test_oracle_entropy_matches_reference_golden(ent, case):
    W, H, m, pc = (int(v) for v in ent[f"{case}_shape"])
    index, data, bits = ent[f"{case}_index"], ent[f"{case}_data"], ent[f"{case}_bits"]
    O = Oracle()
    sec = O.entropy_encode(W, H, m, data, index, pc, bits)
    assert np.array_equal(sec, ent[f"{case}_section"])
    dec = O.entropy_decode(W, H, m, ent[f"{case}_section"], index, pc, bits)
    assert np.array_equal(dec.view(np.uint32), ent[f"{case}_decoded"].view(np.uint32))


def test_oracle_entropy_model_halving(ent):
    W, H, m, pc = (int(v) for v in ent["halving_shape"])
    index = (np.arange(W * H + 1) * 2).astype(np.uint32)
    data = halving_data(W, H, ent["halving_levels"], ent["halving_level"])
    O = Oracle()
    sec = O.entropy_encode(W, H, m, data, index, pc, ent["halving_bits"])
    assert len(sec) == int(ent["halving_section_len"])
    assert np.array_equal(sec[-64:], ent["halving_stream1_tail"])
    assert hashlib.sha256(sec.tobytes()).digest() == ent["halving_section_sha256"].tobytes()
    dec = O.entropy_decode(W, H, m, sec, index, pc, ent["halving_bits"])
    level = ent["halving_level"]
    got = dec[1::2]
    for k, v in enumerate(ent["halving_decoded_levels"]):
        if np.any(level == k):
            assert np.all(got[level == k] == v)


@pytest.mark.parametrize("W,H,m,pc,width", [(33, 21, 12, 0, 10), (17, 40, 9, 1, 6), (25, 25, 5, 0, 14),
                                           (12, 9, 30, 1, 3)])
def test_oracle_entropy_round_trip(W, H, m, pc, width):
    """Lossless coding: the decoded values are the quantized values."""
    rng = np.random.default_rng(W * 1000 + m)
    counts = rng.integers(0, m + 1, W * H).astype(np.uint32)
    index = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    n = int(index[-1])
    data = (rng.random(n) if pc else rng.normal(0, 0.2, n)).astype(np.float32)
    bits = np.array([16] + [width] * (m - 1), np.uint8)
    O = Oracle()
    sec = O.entropy_encode(W, H, m, data, index, pc, bits)
    dec = O.entropy_decode(W, H, m, sec, index, pc, bits)
    assert np.array_equal(dec.view(np.uint32), requantized(O, data, index, pc, bits).view(np.uint32))
    # the layout: u32 count of non-empty pixels, their u16 words, then m-1
    # streams each behind its u32 length, nothing after the last
    nz = int(np.count_nonzero(counts))
    assert int(np.frombuffer(sec[:4].tobytes(), np.uint32)[0]) == nz
    pos = 4 + 2 * nz
    for _ in range(m - 1):
        pos += 4 + int(np.frombuffer(sec[pos:pos + 4].tobytes(), np.uint32)[0])
    assert pos == len(sec)


@pytest.mark.skipif(not REF_SO.exists(), reason="compiled reference not built (build container only)")
@pytest.mark.parametrize("W,H,m,pc,width", [(64, 48, 20, 0, 10), (40, 40, 16, 1, 12), (128, 96, 4, 0, 14),
                                           (50, 30, 40, 0, 2)])
def test_oracle_entropy_matches_live_reference(W, H, m, pc, width):
    rng = np.random.default_rng(m * 7 + width)
    counts = rng.integers(0, m + 1, W * H).astype(np.uint32)
    index = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    n = int(index[-1])
    data = (rng.random(n) if pc else rng.normal(0, 0.1, n)).astype(np.float32)
    bits = np.array([16 if pc else width] + [width] * (m - 1), np.uint8)
    O, R = Oracle(), Reference()
    a = O.entropy_encode(W, H, m, data, index, pc, bits)
    b = R.entropy_encode(W, H, m, data, index, pc, bits)
    assert np.array_equal(a, b)
    assert np.array_equal(O.entropy_decode(W, H, m, a, index, pc, bits).view(np.uint32),
                          R.entropy_decode(W, H, m, b, index, pc, bits).view(np.uint32))


def file_image(files, key):
    W, H, m, pc = (int(v) for v in files[f"img{key}_shape"])
    img = dict(W=W, H=H, m=m, prediction_coded=pc, warp=0, dmin=(-1.0, -1.0, -1.0), dmax=(1.0, 1.0, 1.0),
               data=files[f"img{key}_data"])
    for k in ("bounds", "index", "coding_params", "error_bounds"):
        if f"img{key}_{k}" in files.files:
            img[k] = files[f"img{key}_{k}"]
    return img


def case_spec(name):
    spec = {"c_f32": ("A", 0, 32, 0), "c_lz4_f32": ("A", 1, 32, 0), "c_eb_lz4": ("E", 1, 32, 0),
            "d_f32": ("D", 0, 32, 0), "d_q10_raw": ("D", 0, 10, 0), "d_lz4_f32": ("D", 1, 32, 0)}
    return spec[name]


@pytest.mark.parametrize("name", FLOAT_FILES)
def test_float_files_match_reference(files, name, tmp_path):
    from mray_amd import moment_file

    if not moment_file.LIB_PATH.exists():
        pytest.skip("libmray_host.so not built")
    key, compress, bits, entropy = case_spec(name)
    img = file_image(files, key)
    path = tmp_path / f"{name}.bin"
    moment_file.write_moment_image(path, img, compress, bits, entropy)
    assert np.array_equal(np.frombuffer(path.read_bytes(), np.uint8), files[f"{name}_file"])
    # the reference's file, read back: the data its reader returns
    ref_path = tmp_path / f"{name}_ref.bin"
    ref_path.write_bytes(files[f"{name}_file"].tobytes())
    got = moment_file.load_moment_image(ref_path)
    assert np.array_equal(got["data"].view(np.uint32), files[f"{name}_read"].view(np.uint32))
    if "bounds" in img:
        assert np.array_equal(got["bounds"], img["bounds"])
    if img.get("error_bounds") is not None:
        assert np.array_equal(got["error_bounds"], img["error_bounds"])
    if img.get("index") is not None:
        assert np.array_equal(got["index"], img["index"])
