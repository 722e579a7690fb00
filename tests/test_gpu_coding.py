"""Compaction and prediction coding on the GPU (through the C ABI) against the
oracle and the reference's golden outputs.

Bar: bit-identical. The kernels repeat the reference's FP32 operations in its
order with the host C library's sinf/cosf, so moment counts, index offsets,
compact data, prediction-coded words, quantized coding parameters and the
multipliers decoded from a coded image must equal the reference CPU path's
bit for bit. Inputs are the oracle's generated moments (bit-identical to the
reference's, pinned by the CPU tests), so every difference is the kernels'.
"""
import pathlib

import numpy as np
import pytest

from mray_amd.types import GenParams, RecParams
from oracle_ffi import Oracle
from test_gpu_parity import CASES, case_inputs

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def hip():
    from mray_amd import hip as H

    H.init(0)
    return H


@pytest.fixture(scope="module")
def O():
    return Oracle()


def dev(a):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).cuda()


def host_u32(t):
    return t.detach().cpu().numpy().view(np.uint32)


def host_u16(t):
    return t.detach().cpu().numpy().view(np.uint16)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_num_moments_on_reference_vectors(hip, O):
    """The compaction vectors of the reference's tests at six thresholds
    (tests/test_moment_compaction.cpp), counts from the compiled reference."""
    g = dict(np.load(GOLD / "golden_coding.npz"))
    mo = g["compaction_moments"]
    n, m = mo.shape
    d = dev(mo)
    for i, th in enumerate(g["compaction_thr"]):
        c = host_u32(hip.compute_num_moments(d, n, 1, m, float(th)))
        print(f"[num_moments] threshold {th:g}: counts {np.bincount(c).nonzero()[0].tolist()}")
        assert np.array_equal(c.astype(np.int32), g["compaction_counts"][i])


@pytest.mark.parametrize("m", [16, 64, 100, 255])
def test_num_moments_search_paths(hip, O, m):
    """The count search against the oracle where its binary search takes
    every path: moments of smooth densities (searches that step up, as the
    large frames' do, and the Levinson recursion resumes from the kept
    state) and random moments outside the moment space (the |u|^2 >= 1 guard
    fires and the resumed recursion keeps it), at three thresholds."""
    rng = np.random.default_rng(m)
    n = 384
    # smooth: moments of sums of two bumps on [-pi, 0] (the bounded moments
    # of a density in [0, 1]); random: uniform values
    phi = np.linspace(-np.pi, 0.0, 513)
    smooth = np.empty((n // 2, m), np.float64)
    for i in range(n // 2):
        c, w = rng.uniform(-3.0, -0.2, 2), rng.uniform(0.05, 0.6, 2)
        f = np.clip(0.6 * np.exp(-((phi - c[0]) / w[0]) ** 2) + 0.4 * np.exp(-((phi - c[1]) / w[1]) ** 2), 0, 1)
        for j in range(m):
            smooth[i, j] = np.trapezoid(f * np.cos(j * phi), phi) / np.pi
    smooth[:, 0] = np.clip(smooth[:, 0], 0.02, 0.98)
    rand = np.column_stack([rng.uniform(0.1, 0.9, n // 2), rng.uniform(-0.3, 0.3, (n // 2, m - 1))])
    mo = np.ascontiguousarray(np.vstack([smooth, rand]).astype(np.float32))
    d = dev(mo)
    for th in (1e-2, 1e-3, 1e-4):
        print(f"[num_moments] m={m} threshold {th:g}", flush=True)
        ref = O.num_moments(n, 1, m, mo.reshape(-1), None, th)
        got = host_u32(hip.compute_num_moments(d, n, 1, m, th))
        assert np.array_equal(got, ref), (m, th, np.flatnonzero(got != ref)[:8])


def test_coded_pipeline_matches_reference_outputs(hip, O):
    """generate (oracle) -> compaction -> prediction coding -> coded preparation ->
    PLAIN march, against the compiled reference's own outputs for the same
    frame (golden_images.npz, case mix_coded)."""
    import torch

    g = dict(np.load(GOLD / "golden_images.npz"))
    W, H, m = 16, 12, 16
    from mray_amd import synth

    cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=16 / 12)
    gp = GenParams(step_size=1.0 / 48, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=1,
                   prediction_coding=1, error_threshold=0.01, bias=1e-4)
    rp = RecParams(step_size=1.0 / 48, bias=1e-4, early_ray_termination=0.98, medium_emission=1.0)
    mo, b, _ = O.generate(g["mix_coded_volume"], (0, 0, 0), (1, 1, 1), cam, W, H, m, gp)
    data, index = hip.compact_image(dev(mo), W, H, m, gp.error_threshold, bounds=dev(b))
    print(f"[coded] index total {host_u32(index)[-1]} (reference {g['mix_coded_index'][-1]})", flush=True)
    assert np.array_equal(host_u32(index), g["mix_coded_index"])
    q = hip.prediction_encode(data, index, W, H, m, gp.coding_warp, gp.bias)
    torch.cuda.synchronize()
    assert np.array_equal(host_u16(q), g["mix_coded_coding_params"])
    assert np.array_equal(bits(data.cpu().numpy()), bits(g["mix_coded_data"]))
    lag = hip.prepare_coded(data, W, H, m, gp.coding_warp, q, index=index)
    table = torch.from_numpy(O.tf(g["tf_rgba8"], rp.step_size))
    rgba, _ = hip.reconstruct(cam, (0, 0, 0), (1, 1, 1), lag, W, H, m, table, rp, form="plain", index=index,
                              bounds=dev(b))
    torch.cuda.synchronize()
    same = (bits(rgba.cpu().numpy()) == bits(g["mix_coded_rgba"])).all(axis=-1)
    print(f"[coded] bit-identical rays {same.sum()}/{W * H}")
    assert same.all()


@pytest.mark.parametrize("threshold", [0.0, 1e-3, 1e-2])
@pytest.mark.parametrize("name", ["ml_ortho_m30", "mix_persp_m50", "mixw_persp_m50", "u8_ortho_m16",
                                  "ml_nobounds_m7", "tiny_m1", "tiny_m2"])
def test_compaction_and_coding_match_oracle(hip, O, name, threshold):
    import torch

    vol, dmin, dmax, cam, W, H, m, gp = case_inputs(name)
    mo, b, _ = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    bnd = b if gp.transform_to_bounds else None
    print(f"[coding] case={name} threshold={threshold}", flush=True)
    counts_ref = O.num_moments(W, H, m, mo.reshape(-1), bnd, threshold)
    index_ref, comp_ref = O.compact(W, H, m, mo.reshape(-1), counts_ref)

    d_mo = dev(mo)
    counts = hip.compute_num_moments(d_mo, W, H, m, threshold, bounds=dev(bnd) if bnd is not None else None)
    assert np.array_equal(host_u32(counts), counts_ref)
    index = hip.moment_index(counts, W, H)
    assert np.array_equal(host_u32(index), index_ref)
    data = hip.compact(d_mo, index, W, H, m)
    assert np.array_equal(bits(data.cpu().numpy()), bits(comp_ref))

    for warp in (0, 1):
        coded_ref, q_ref = O.prediction_encode(W, H, m, comp_ref, index_ref, warp, gp.bias)
        coded = torch.empty(max(data.numel(), 1), device="cuda")[: data.numel()]
        coded.copy_(data)
        q = hip.prediction_encode(coded, index, W, H, m, warp, gp.bias)
        torch.cuda.synchronize()
        coded_h = coded.cpu().numpy()
        diff = np.count_nonzero(bits(coded_h) != bits(coded_ref))
        print(f"  warp {warp}: coded words differing {diff}/{len(coded_ref)}")
        assert diff == 0
        params_f = np.zeros(max(2 * (m - 1), 1), np.float32)
        if q is not None:
            assert np.array_equal(host_u16(q), q_ref)
            params_f = (q_ref.astype(np.float32) / np.float32(65535.0)).astype(np.float32)
        lag = hip.prepare_coded(coded, W, H, m, warp, q, index=index).cpu().numpy()
        ref = np.zeros_like(coded_ref)
        for p in range(W * H):
            s, e = index_ref[p], index_ref[p + 1]
            if e > s:
                ref[s:e] = O.prepare_pred(int(e - s), coded_ref[s:e], warp, params_f)
        nd = np.count_nonzero(bits(lag) != bits(ref))
        print(f"  warp {warp}: decoded multipliers differing {nd}/{len(ref)}")
        assert nd == 0


def test_index_scan_large_and_in_place(hip):
    """The exclusive scan across many tiles (4096 counts each), in place."""
    import torch

    rng = np.random.default_rng(3)
    W, H = 1031, 517  # not a multiple of the tile
    c = rng.integers(0, 101, W * H).astype(np.uint32)
    ref = np.zeros(W * H + 1, np.uint64)
    ref[1:] = np.cumsum(c, dtype=np.uint64)
    buf = torch.zeros(W * H + 1, dtype=torch.int32, device="cuda")
    buf[: W * H] = dev(c)
    hip.moment_index(buf, W, H, out=buf)
    assert np.array_equal(host_u32(buf).astype(np.uint64), ref)


def test_coding_error_paths(hip):
    import torch

    from mray_amd.hip import MrayError

    x = torch.zeros(16, dtype=torch.float32, device="cuda")
    idx = torch.zeros(5, dtype=torch.int32, device="cuda")
    with pytest.raises(MrayError):
        hip.prediction_encode(x, idx, 2, 2, 4, coding_warp=7)
    with pytest.raises(MrayError):
        hip.compute_num_moments(x, 2, 2, 0, 1e-3)
