"""Reader for moment-image files (the uncompressed variants of
src/moment_image_io.cpp: float moments, or the quantized stream of a compact
image, with or without its quantization table) used by the CLI tests."""
import numpy as np


def read_moment_image(path):
    b = open(path, "rb").read()
    w, h = np.frombuffer(b, np.uint16, 2, 0)
    num_moments, warp, flags, qbits = b[4], b[5], b[6], b[7]
    dmin = np.frombuffer(b, np.float32, 3, 8)
    dmax = np.frombuffer(b, np.float32, 3, 20)
    assert not (flags & 1) and not (qbits & 0x40)
    off = 32
    npix = int(w) * int(h)
    out = dict(width=int(w), height=int(h), num_moments=int(num_moments), coding_warp=int(warp),
               compact=bool(flags & 2), prediction_code=bool(flags & 4), domain=(dmin, dmax),
               quantization_bits=int(qbits & 0x3F))
    if qbits & 0x80:  # the quantization table follows the header
        out["quant_table"] = np.frombuffer(b, np.uint8, int(num_moments), off).copy()
        off += int(num_moments)
    if out["compact"]:
        counts = np.frombuffer(b, np.uint8, npix, off).astype(np.uint32)
        off += npix
        index = np.zeros(npix + 1, np.uint32)
        index[1:] = np.cumsum(counts)
        out["index"] = index
    if flags & 16:
        out["bounds"] = np.frombuffer(b, np.uint32, npix, off).copy()
        off += 4 * npix
    if flags & 32:
        off += 2 * npix
    if flags & 128:
        out["coding_params"] = np.frombuffer(b, np.uint16, 2 * (int(num_moments) - 1), off).copy()
        off += 4 * (int(num_moments) - 1)
    n = int(out["index"][-1]) if out["compact"] else npix * int(num_moments)
    if out["quantization_bits"] < 32:
        out["stream"] = np.frombuffer(b, np.uint8, len(b) - off, off).copy()
        return out
    out["data"] = np.frombuffer(b, np.float32, n, off).copy()
    assert off + 4 * n == len(b)
    return out
