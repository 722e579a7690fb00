"""Adjoint of the reconstruct march on the GPU (mray_hip_reconstruct_backward)
against the FP64 oracle adjoint, against central finite differences of the
GPU forward, and in use: transfer-function fitting through
torch.autograd (the differentiation workload, scaled down).

Bars: the GPU forward is FP32 with an FP64 series, the oracle is FP64
throughout; the two forwards differ by ~1e-5 relative, so gradients must agree
to 2e-3 normwise (table) and per pixel (multipliers, norm floored at 1% of the
mean pixel norm). Rays where the two forwards take different discrete
branches are compared with zero upstream gradient: another step count (early
ray termination is discontinuous in tau) or images more than 1e-4 apart (the
reference's lookup takes its table cell from the front sample, so a density
within rounding of a cell edge can land in neighbouring cells in FP32 and
FP64; the image jumps there). The table gradient is compared on the cases
whose densities keep clear of cell edges: on ml_ortho_m40 many samples sit
within the FP32/FP64 density difference (~1e-5) of an edge, so their table
gradient lands one entry over while the image barely changes; that case's
table gradient is checked against finite differences of the GPU forward
instead (test_adjoint_finite_differences_on_gpu). Finite differences of the
FP32 forward with a relative step of 1e-3 (the configuration's eps) resolve a
directional derivative to ~1e-4: bar 2e-3."""
import numpy as np
import pytest

from mray_amd import synth, tf as tfmod
from mray_amd.types import GenParams, RecParams
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from mray_amd import hip as H

    H.init(0)
    return H


CASES = {
    "ml_ortho_m40": (lambda: synth.marschner_lobb(48), (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 40, 32, 40,
                     2.0 / 96, 4.0, 5e-5),
    "mixw_persp_m30": (lambda: synth.mixing_layer(40, width=0.05), (0, 0, 0), (1, 1, 1),
                       synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=36 / 30), 36, 30, 30, 0.5 / 40,
                       40.0, 1e-3),
    "ml_persp_m16": (lambda: synth.marschner_lobb(32), (-1, -1, -1), (1, 1, 1),
                     synth.perspective_camera(position=(0.15, -0.1, -2.6), fov_half_deg=24.0, aspect=1.0), 24, 24,
                     16, 2.0 / 64, 4.0, 5e-5),
    # beyond 64 moments: two (m <= 128) or four lanes per ray in the reverse
    "ml_ortho_m100": (lambda: synth.marschner_lobb(32), (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 22, 18, 100,
                      2.0 / 64, 4.0, 1e-3),
    "ml_persp_m200": (lambda: synth.marschner_lobb(32), (-1, -1, -1), (1, 1, 1),
                      synth.perspective_camera(position=(0.15, -0.1, -2.6), fov_half_deg=24.0, aspect=1.0), 18, 18,
                      200, 2.0 / 64, 4.0, 1e-3),
    # enough tile rows for the workgroups' dealing out to the XCDs to remap
    # (136 x 200 rays: 107 workgroups of 4 tiles, 80 of them remapped)
    "mixw_persp_m16_rows": (lambda: synth.mixing_layer(40, width=0.05), (0, 0, 0), (1, 1, 1),
                            synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=136 / 200), 136, 200, 16,
                            0.5 / 40, 40.0, 1e-3),
}


def setup(O, name, ert=0.99):
    f, dmin, dmax, cam, W, H, m, step, factor, bias = CASES[name]
    vol = f()
    gp = GenParams(step_size=step, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    mo, b, _ = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    rp = RecParams(step_size=step, bias=bias, early_ray_termination=ert, medium_emission=1.0)
    lag = np.stack([O.prepare_std(m, mo.reshape(-1, m)[p], rp.bias) for p in range(W * H)]).astype(np.float32)
    table = tfmod.table_from_rgba8(synth.tf_ramps(0), step * factor)
    return dmin, dmax, cam, W, H, m, rp, lag, table, b


@pytest.mark.parametrize("name", list(CASES))
def test_adjoint_matches_oracle(hip, name):
    import torch

    O = Oracle()
    dmin, dmax, cam, W, H, m, rp, lag, table, b = setup(O, name)
    rng = np.random.default_rng(5)
    g = rng.normal(size=(H, W, 4)).astype(np.float32)
    rgba64, st64 = O.reconstruct_lagrange_f64(cam, dmin, dmax, W, H, m, lag, table, rp, bounds=b)
    lag_d = torch.from_numpy(lag).cuda()
    b_d = torch.from_numpy(b.view(np.int32)).cuda()
    tab_d = torch.from_numpy(table).cuda()
    print(f"[adjoint] case={name}", flush=True)
    rgba, st = hip.reconstruct(cam, dmin, dmax, lag_d, W, H, m, tab_d, rp, form="mixed", bounds=b_d,
                               want_steps=True)
    torch.cuda.synchronize()
    jump = np.abs(rgba.cpu().numpy() - rgba64).max(axis=-1).reshape(-1) > 1e-4
    same = (st.cpu().numpy() == st64) & ~jump
    g[~same.reshape(H, W)] = 0.0
    gl_ref, gt_ref = O.reconstruct_adjoint_f64(cam, dmin, dmax, W, H, m, lag.reshape(-1), table, rp,
                                               g.astype(np.float64), bounds=b)
    g_lag, g_tf = hip.reconstruct_backward(cam, dmin, dmax, lag_d, W, H, m, tab_d, rp, rgba,
                                           torch.from_numpy(g).cuda(), bounds=b_d)
    torch.cuda.synchronize()
    g_lag = g_lag.cpu().numpy().reshape(-1, m).astype(np.float64)
    gl_ref = gl_ref.reshape(-1, m)
    norms = np.linalg.norm(gl_ref, axis=1)
    floor = 0.01 * norms[same].mean()
    rel_lag = np.linalg.norm(g_lag - gl_ref, axis=1) / np.maximum(norms, floor)
    g_tf = g_tf.cpu().numpy().astype(np.float64)
    rel_tf = np.linalg.norm(g_tf - gt_ref) / np.linalg.norm(gt_ref)
    print(f"  rays on other branches: step count {np.count_nonzero(st.cpu().numpy() != st64)}, image jump "
          f"{np.count_nonzero(jump)} of {W * H}; multipliers max rel {rel_lag[same].max():.2e}; "
          f"table rel {rel_tf:.2e}")
    dif = np.abs(g_tf - gt_ref).reshape(-1)
    for i in np.argsort(dif)[::-1][:3]:
        print(f"    entry {i // 4} ch {i % 4}: gpu {g_tf.reshape(-1)[i]:.6e} fp64 {gt_ref.reshape(-1)[i]:.6e}")
    assert np.count_nonzero(~same) <= max(2, 0.02 * W * H)
    assert rel_lag[same].max() <= 2e-3
    if name not in ("ml_ortho_m40", "ml_ortho_m100"):
        assert rel_tf <= 2e-3


def test_table_gradient_scale_invariance(hip):
    """The table gradient is summed per workgroup in 64-bit fixed point at a
    scale taken from a bound on each ray's contribution: scaling dL/drgba by
    2^-40 or 2^40 (exact in float) scales dL/dtable by the same factor, to
    1e-6 normwise; a zero incoming gradient gives exactly zero, and a
    non-finite one gives a non-finite table gradient."""
    import torch

    O = Oracle()
    name = list(CASES)[0]
    dmin, dmax, cam, W, H, m, rp, lag, table, b = setup(O, name)
    rng = np.random.default_rng(11)
    g = torch.from_numpy(rng.normal(size=(H, W, 4)).astype(np.float32)).cuda()
    lag_d = torch.from_numpy(lag).cuda()
    b_d = torch.from_numpy(b.view(np.int32)).cuda()
    tab_d = torch.from_numpy(table).cuda()
    print(f"[table gradient scale] case={name}", flush=True)
    rgba, _ = hip.reconstruct(cam, dmin, dmax, lag_d, W, H, m, tab_d, rp, form="mixed", bounds=b_d)
    ref = None
    for e in (0, -40, 40):
        _, g_tf = hip.reconstruct_backward(cam, dmin, dmax, lag_d, W, H, m, tab_d, rp, rgba, g * 2.0 ** e,
                                           bounds=b_d, want_lagrange=False)
        v = g_tf.cpu().numpy().astype(np.float64) * 2.0 ** (-e)
        if ref is None:
            ref = v
            assert np.linalg.norm(ref) > 0
        else:
            rel = np.linalg.norm(v - ref) / np.linalg.norm(ref)
            print(f"  scale 2^{e}: rel {rel:.2e}")
            assert rel <= 1e-6
    _, g0 = hip.reconstruct_backward(cam, dmin, dmax, lag_d, W, H, m, tab_d, rp, rgba, torch.zeros_like(g),
                                     bounds=b_d, want_lagrange=False)
    assert torch.count_nonzero(g0).item() == 0
    # a non-finite incoming gradient is not dropped by the fixed-point sums:
    # that workgroup adds floats, and the table gradient is non-finite
    gi = g.clone()
    gi[H // 2, W // 2, 3] = float("inf")
    _, gt_inf = hip.reconstruct_backward(cam, dmin, dmax, lag_d, W, H, m, tab_d, rp, rgba, gi, bounds=b_d,
                                         want_lagrange=False)
    assert not torch.isfinite(gt_inf).all()


def test_reconstruct_adjoint_compact_matches_dense(hip):
    """Compact multipliers (index with 0, 5 and m entries per pixel): the
    reverse pass with the index equals the dense one on the zero-padded
    multipliers (the series is the same), and pixels without moments add
    their incoming gradient to table entry 0 (the zero ray's colour); both
    paths sum the table gradient in fixed point. Bar 1e-5 normwise (float
    atomics across workgroups)."""
    import torch

    O = Oracle()
    name = list(CASES)[0]
    dmin, dmax, cam, W, H, m, rp, lag, table, b = setup(O, name)
    rng = np.random.default_rng(13)
    npix = W * H
    nm = np.where(np.arange(npix) % 3 == 0, 0, np.where(np.arange(npix) % 3 == 1, 5, m)).astype(np.int64)
    index = np.concatenate([[0], np.cumsum(nm)]).astype(np.uint32)
    lag2 = lag.reshape(npix, m)
    compact = np.concatenate([lag2[p, :nm[p]] for p in range(npix)]).astype(np.float32)
    dense = np.zeros_like(lag2)
    for p in range(npix):
        dense[p, :nm[p]] = lag2[p, :nm[p]]
    g = rng.normal(size=(H, W, 4)).astype(np.float32)
    g_dense = g.copy()
    g_dense.reshape(npix, 4)[nm == 0] = 0.0
    print(f"[reconstruct adjoint compact] case={name} W={W} H={H} m={m}", flush=True)
    tab_d = torch.from_numpy(table).cuda()
    b_d = torch.from_numpy(b.view(np.int32)).cuda()
    c_d = torch.from_numpy(compact).cuda()
    i_d = torch.from_numpy(index.view(np.int32)).cuda()
    d_d = torch.from_numpy(dense.reshape(-1)).cuda()
    rgba_c, _ = hip.reconstruct(cam, dmin, dmax, c_d, W, H, m, tab_d, rp, form="mixed", index=i_d, bounds=b_d)
    rgba_d, _ = hip.reconstruct(cam, dmin, dmax, d_d, W, H, m, tab_d, rp, form="mixed", bounds=b_d)
    gl_c, gt_c = hip.reconstruct_backward(cam, dmin, dmax, c_d, W, H, m, tab_d, rp, rgba_c,
                                          torch.from_numpy(g).cuda(), index=i_d, bounds=b_d)
    gl_d, gt_d = hip.reconstruct_backward(cam, dmin, dmax, d_d, W, H, m, tab_d, rp, rgba_d,
                                          torch.from_numpy(g_dense).cuda(), bounds=b_d)
    gl_c = gl_c.cpu().numpy()
    gl_d = gl_d.cpu().numpy().reshape(npix, m)
    expect = np.concatenate([gl_d[p, :nm[p]] for p in range(npix)])
    rel_l = np.linalg.norm(gl_c - expect) / np.linalg.norm(expect)
    zero_rays = g.reshape(npix, 4)[nm == 0].astype(np.float64).sum(axis=0)
    gt_expect = gt_d.cpu().numpy().astype(np.float64)
    gt_expect[0] += zero_rays
    rel_t = np.linalg.norm(gt_c.cpu().numpy() - gt_expect) / np.linalg.norm(gt_expect)
    print(f"  multipliers rel {rel_l:.2e}, table rel {rel_t:.2e}")
    assert rel_l <= 1e-5
    assert rel_t <= 1e-5


def test_adjoint_finite_differences_on_gpu(hip):
    """Directional derivatives of L = sum(g * rgba) w.r.t. the table, central
    differences with eps = 1e-3 of the GPU forward, no early termination:
    along the gradient itself (a large derivative, resolved to ~1e-4) and
    along two random directions. Bar: 2e-3 relative plus the FP32 forward's
    rounding noise divided by 2 eps."""
    import torch

    O = Oracle()
    dmin, dmax, cam, W, H, m, rp, lag, table, b = setup(O, "ml_ortho_m40", ert=1.5)
    rng = np.random.default_rng(8)
    g = torch.from_numpy(rng.normal(size=(H, W, 4)).astype(np.float32)).cuda()
    lag_d = torch.from_numpy(lag).cuda()
    b_d = torch.from_numpy(b.view(np.int32)).cuda()
    tab = torch.from_numpy(table).cuda()

    def loss(t):
        rgba, _ = hip.reconstruct(cam, dmin, dmax, lag_d, W, H, m, t, rp, form="mixed", bounds=b_d)
        return float((rgba.double() * g.double()).sum())

    rgba, _ = hip.reconstruct(cam, dmin, dmax, lag_d, W, H, m, tab, rp, form="mixed", bounds=b_d)
    _, g_tf = hip.reconstruct_backward(cam, dmin, dmax, lag_d, W, H, m, tab, rp, rgba, g, bounds=b_d,
                                       want_lagrange=False)
    # rounding noise of the FP32 forward in L: ~16 ulps per output value
    # (about a hundred rounded compositing steps), independent between the
    # two evaluations; divided by 2 eps it bounds the noise of the difference
    r = rgba.cpu().numpy().astype(np.float64)
    gn = g.cpu().numpy().astype(np.float64)
    noise = 16.0 * np.sqrt(2.0 * ((np.spacing(np.abs(r).astype(np.float32)) * gn) ** 2).sum())
    dirs = [g_tf * tab.abs().clamp_min(1e-4) ** 2 / (g_tf * tab.abs().clamp_min(1e-4)).abs().max()]
    dirs += [torch.from_numpy(rng.normal(size=(256, 4)).astype(np.float32)).cuda() * tab.abs().clamp_min(1e-4)
             for _ in range(2)]
    for d in dirs:
        eps = 1e-3
        fd = (loss(tab + eps * d) - loss(tab - eps * d)) / (2 * eps)
        an = float((g_tf.double() * d.double()).sum())
        bar = 2e-3 * abs(an) + noise / (2 * eps)
        print(f"[adjoint fd] table: fd {fd:.6e} adjoint {an:.6e} bar {bar:.1e}")
        assert abs(fd - an) <= bar


def test_transfer_function_fitting(hip):
    """A perturbed transfer function (target + N(0, 0.05)) is fitted to the
    target image by gradient descent through the adjoint; the L2 image loss
    must fall by at least 10x."""
    import torch

    from mray_amd import autograd as ag

    O = Oracle()
    vol = synth.marschner_lobb(64)
    dmin, dmax, cam, W, H, m = (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 64, 64, 40
    step = 2.0 / 128
    gp = GenParams(step_size=step, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    v = hip.Volume(torch.from_numpy(vol).cuda(), dmin, dmax)
    mo, b, _ = hip.generate(v, cam, gp, W, H, m, form="mixed")
    rp = RecParams(step_size=step, bias=5e-5, early_ray_termination=1.5, medium_emission=1.0)
    lag = hip.prepare(mo, W, H, m, rp.bias, form="mixed")
    rgba8 = synth.tf_ramps(0)
    target_tf = torch.from_numpy(rgba8.astype(np.float32) / 255.0).cuda()
    start = torch.from_numpy(synth.tf_perturbed(rgba8, seed=3).astype(np.float32)).cuda()
    s = step * 4.0
    with torch.no_grad():
        target = ag.reconstruct(lag, ag.transfer_table(target_tf, s), cam, dmin, dmax, W, H, m, rp, bounds=b)
    x = start.clone().requires_grad_(True)
    opt = torch.optim.Adam([x], lr=0.01)
    losses = []
    for it in range(60):
        opt.zero_grad()
        img = ag.reconstruct(lag, ag.transfer_table(x, s), cam, dmin, dmax, W, H, m, rp, bounds=b)
        loss = ((img - target) ** 2).mean()
        loss.backward()
        opt.step()
        with torch.no_grad():
            x.clamp_(0.0, 1.0)
        losses.append(float(loss))
    print(f"[fit] loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert losses[-1] <= 0.1 * losses[0]


@pytest.mark.parametrize("name", ["ml_ortho_m40", "mixw_persp_m30", "ml_persp_m16"])
def test_prepare_adjoint_matches_fp64_finite_differences(hip, name):
    """dL/dmoments of prepare against central differences of the FP64 oracle
    preparation (relative step 1e-6 of the moment scale, FP64 throughout):
    per pixel L = sum_k gbar_k lambda_k, one random direction per pixel.
    Bar: 1e-5 relative to the sum of the magnitudes of the terms (the
    conditioning of the Levinson recursion amplifies FP64 rounding)."""
    import torch

    O = Oracle()
    dmin, dmax, cam, W, H, m, rp, lag, table, b = setup(O, name)
    f, _, _, _, _, _, _, step, _, _ = CASES[name]
    gp = GenParams(step_size=step, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    mo, _, _ = O.generate(f(), dmin, dmax, cam, W, H, m, gp)
    mo = mo.reshape(-1, m)
    rng = np.random.default_rng(21)
    gbar = rng.normal(size=(W * H, m)).astype(np.float32)
    gm = hip.prepare_backward(torch.from_numpy(np.ascontiguousarray(mo)).cuda(), W, H, m, rp.bias,
                              torch.from_numpy(gbar).cuda()).cpu().numpy().astype(np.float64)
    worst = 0.0
    for p in range(0, W * H, max(1, W * H // 200)):
        x = mo[p].astype(np.float64)
        d = rng.normal(size=m) * 0.1
        eps = 1e-6
        lp = O.prepare_std_f64(m, x + eps * d, rp.bias)
        lm = O.prepare_std_f64(m, x - eps * d, rp.bias)
        fd = float(gbar[p].astype(np.float64) @ (lp - lm)) / (2 * eps)
        an = float(gm[p] @ d)
        scale = float(np.abs(gm[p] * d).sum()) + 1e-12
        worst = max(worst, abs(fd - an) / scale)
    print(f"[prepare adjoint] case={name}: worst |fd - adjoint| / sum|terms| {worst:.2e}")
    assert worst <= 1e-5


@pytest.mark.parametrize("m", [72, 100])
def test_prepare_adjoint_large_m(hip, m):
    """Beyond 64 moments the preparation reverse runs one thread per pixel
    with a global scratch: dL/dmoments against central differences of the
    FP64 oracle preparation, as test_prepare_adjoint_matches_fp64_finite_differences
    (moments from the oracle's generate)."""
    import torch

    O = Oracle()
    vol = synth.marschner_lobb(32)
    dmin, dmax = (-1, -1, -1), (1, 1, 1)
    cam = synth.ortho_camera()
    W, H = 10, 8
    gp = GenParams(step_size=2.0 / 64, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    print(f"[prepare adjoint large m] m={m} W={W} H={H}", flush=True)
    mo, _, _ = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    mo = mo.reshape(-1, m)
    bias = 5e-5
    rng = np.random.default_rng(23)
    gbar = rng.normal(size=(W * H, m)).astype(np.float32)
    gm = hip.prepare_backward(torch.from_numpy(np.ascontiguousarray(mo)).cuda(), W, H, m, bias,
                              torch.from_numpy(gbar).cuda()).cpu().numpy().astype(np.float64)
    worst = 0.0
    for p in range(0, W * H, 7):
        x = mo[p].astype(np.float64)
        d = rng.normal(size=m) * 0.1
        eps = 1e-6
        lp = O.prepare_std_f64(m, x + eps * d, bias)
        lm = O.prepare_std_f64(m, x - eps * d, bias)
        fd = float(gbar[p].astype(np.float64) @ (lp - lm)) / (2 * eps)
        an = float(gm[p] @ d)
        scale = float(np.abs(gm[p] * d).sum()) + 1e-12
        worst = max(worst, abs(fd - an) / scale)
    print(f"  worst |fd - adjoint| / sum|terms| {worst:.2e}")
    assert worst <= 1e-5


def test_pipeline_adjoint_reconstruct_prepare(hip):
    """dL/dmoments through the march and the preparation (GPU adjoints
    chained) against central differences of the FP64 pipeline (oracle
    preparation in FP64, FP64 march) along random directions, without early
    termination. Bar: 2e-3 relative (the GPU forward is FP32 with an FP64
    series; its image differs from FP64 by ~1e-5 relative)."""
    import torch

    O = Oracle()
    dmin, dmax, cam, W, H, m, rp, lag, table, b = setup(O, "ml_persp_m16", ert=1.5)
    f, _, _, _, _, _, _, step, _, _ = CASES["ml_persp_m16"]
    gp = GenParams(step_size=step, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    mo, _, _ = O.generate(f(), dmin, dmax, cam, W, H, m, gp)
    rng = np.random.default_rng(31)
    g = rng.normal(size=(H, W, 4)).astype(np.float32)
    mo_d = torch.from_numpy(np.ascontiguousarray(mo)).cuda()
    b_d = torch.from_numpy(b.view(np.int32)).cuda()
    tab_d = torch.from_numpy(table).cuda()
    lag_d = hip.prepare(mo_d, W, H, m, rp.bias, form="mixed")
    rgba, _ = hip.reconstruct(cam, dmin, dmax, lag_d, W, H, m, tab_d, rp, form="mixed", bounds=b_d)
    g_lag, _ = hip.reconstruct_backward(cam, dmin, dmax, lag_d, W, H, m, tab_d, rp, rgba,
                                        torch.from_numpy(g).cuda(), bounds=b_d, want_tf=False)
    g_mo = hip.prepare_backward(mo_d, W, H, m, rp.bias, g_lag).cpu().numpy().astype(np.float64).reshape(-1)

    mo64 = mo.astype(np.float64).reshape(-1, m)
    tf64 = table.astype(np.float64)

    def loss(x):
        lam = np.stack([O.prepare_std_f64(m, x[p], rp.bias) for p in range(W * H)])
        return float((g.astype(np.float64) * O.reconstruct_lagrange64_f64(cam, dmin, dmax, W, H, m,
                                                                          lam.reshape(-1), tf64, rp,
                                                                          bounds=b)).sum())

    for trial in range(2):
        d = rng.normal(size=mo64.shape) * 0.01
        eps = 1e-6
        fd = (loss(mo64 + eps * d) - loss(mo64 - eps * d)) / (2 * eps)
        an = float(g_mo @ d.reshape(-1))
        print(f"[pipeline adjoint] fd {fd:.6e} adjoint {an:.6e}")
        assert abs(fd - an) <= 2e-3 * abs(fd) + 1e-6


@pytest.mark.parametrize("persp,m,W,H", [(False, 16, 18, 14), (True, 40, 20, 20), (False, 100, 18, 14),
                                         (True, 200, 20, 20), (True, 16, 136, 96)])
def test_generate_adjoint_matches_fp64_finite_differences(hip, persp, m, W, H):
    """dL/dvolume of generate (L = sum gbar * moments) against central
    differences of the FP64 oracle's generate. Without bound normalisation
    and with a data range that never clamps, the moments are linear in the
    voxels, so central differences are exact up to the float storage of the
    perturbed volume (relative 1e-5 at eps = 1e-2). Bar: 2e-4 relative.
    136 x 96 rays are 51 workgroups of 4 tiles, 40 of them remapped when the
    tile rows are dealt out to the XCDs."""
    import torch

    O = Oracle()
    rng = np.random.default_rng(41)
    if persp:
        vol, dmin, dmax = synth.marschner_lobb(24), (-1, -1, -1), (1, 1, 1)
        cam = synth.perspective_camera(position=(0.1, 0.05, -2.6), fov_half_deg=22.0, aspect=W / H)
    else:
        vol, dmin, dmax = synth.marschner_lobb(20), (-1, -1, -1), (1, 1, 1)
        cam = synth.ortho_camera()
    gp = GenParams(step_size=2.0 / 48, transform_to_bounds=0, data_min=-0.1, data_max=1.1, compact_image=0,
                   prediction_coding=0)
    gbar = rng.normal(size=(H, W, m))
    v = hip.Volume(torch.from_numpy(vol).cuda(), dmin, dmax)
    gv = hip.generate_backward(v, cam, gp, W, H, m, torch.from_numpy(gbar.astype(np.float32)).cuda())
    gv = gv.cpu().numpy().astype(np.float64)

    def loss(vv):
        mo, _ = O.generate(np.ascontiguousarray(vv, np.float32), dmin, dmax, cam, W, H, m, gp, f64=True)
        return float((gbar * mo).sum())

    for trial in range(3):
        d = rng.normal(size=vol.shape)
        eps = 1e-2
        fd = (loss(vol + eps * d) - loss(vol - eps * d)) / (2 * eps)
        an = float((gv * d).sum())
        print(f"[generate adjoint] persp={persp} m={m}: fd {fd:.9e} adjoint {an:.9e}")
        assert abs(fd - an) <= 2e-4 * np.abs(gv * d).sum()


def test_generate_adjoint_with_forward_ray_range(hip):
    """The forward's per-ray normalisation (ray_range) is what it packed into
    the 16-bit bounds, and the reverse pass given it equals the reverse pass
    that recomputes it (float atomics: the same sums in another order, bar
    1e-5 of the largest gradient)."""
    import torch

    rng = np.random.default_rng(5)
    vol = synth.marschner_lobb(24)
    dmin, dmax = (-1, -1, -1), (1, 1, 1)
    cam = synth.perspective_camera(position=(0.1, 0.05, -2.6), fov_half_deg=22.0, aspect=1.0)
    W = H = 20
    m = 24
    gp = GenParams(step_size=2.0 / 48, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    v = hip.Volume(torch.from_numpy(vol).cuda(), dmin, dmax)
    rr = torch.empty((H * W, 2), dtype=torch.float32, device="cuda")
    _, b, _ = hip.generate(v, cam, gp, W, H, m, form="mixed", ray_range=rr)
    r = rr.cpu().numpy()
    bb = b.cpu().numpy().astype(np.uint32)
    assert np.array_equal(bb & 0xFFFF, (r[:, 0] * np.float32(65535.0)).astype(np.uint16))
    assert np.array_equal(bb >> 16, (r[:, 1] * np.float32(65535.0)).astype(np.uint16))
    gbar = torch.from_numpy(rng.normal(size=(H, W, m)).astype(np.float32)).cuda()
    g1 = hip.generate_backward(v, cam, gp, W, H, m, gbar, ray_range=rr).cpu().numpy()
    g2 = hip.generate_backward(v, cam, gp, W, H, m, gbar).cpu().numpy()
    assert np.abs(g1 - g2).max() <= 1e-5 * np.abs(g2).max()


def test_generate_adjoint_scale_invariance(hip):
    """The workgroup's voxel window sums in 64-bit fixed point at a scale
    from a bound on each ray's contribution: dL/dmoments scaled by 2^-40 or
    2^40 (exact in float) scales dL/dvoxels by the same factor (1e-6
    normwise), a zero incoming gradient gives exactly zero, and a non-finite
    one gives a non-finite gradient (not a silently dropped summand)."""
    import torch

    rng = np.random.default_rng(9)
    vol = synth.marschner_lobb(24)
    dmin, dmax = (-1, -1, -1), (1, 1, 1)
    cam = synth.perspective_camera(position=(0.1, 0.05, -2.6), fov_half_deg=22.0, aspect=1.0)
    W = H = 20
    m = 24
    gp = GenParams(step_size=2.0 / 48, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    print(f"[generate adjoint scale] W={W} H={H} m={m}", flush=True)
    v = hip.Volume(torch.from_numpy(vol).cuda(), dmin, dmax)
    rr = torch.empty((H * W, 2), dtype=torch.float32, device="cuda")
    hip.generate(v, cam, gp, W, H, m, form="mixed", ray_range=rr)
    gbar = torch.from_numpy(rng.normal(size=(H, W, m)).astype(np.float32)).cuda()
    ref = hip.generate_backward(v, cam, gp, W, H, m, gbar, ray_range=rr).cpu().numpy().astype(np.float64)
    assert np.linalg.norm(ref) > 0
    for e in (-40, 40):
        ge = hip.generate_backward(v, cam, gp, W, H, m, gbar * 2.0 ** e, ray_range=rr).cpu().numpy()
        rel = np.linalg.norm(ge.astype(np.float64) * 2.0 ** (-e) - ref) / np.linalg.norm(ref)
        print(f"  scale 2^{e}: rel {rel:.2e}")
        assert rel <= 1e-6
    g0 = hip.generate_backward(v, cam, gp, W, H, m, torch.zeros_like(gbar), ray_range=rr)
    assert torch.count_nonzero(g0).item() == 0
    # a non-finite incoming gradient reaches the voxels (float atomics then)
    gi = gbar.clone()
    gi[H // 2, W // 2, 1] = float("inf")
    g_inf = hip.generate_backward(v, cam, gp, W, H, m, gi, ray_range=rr)
    assert not torch.isfinite(g_inf).all()


def test_render_autograd_chain(hip):
    """torch.autograd through generate -> prepare -> reconstruct: the volume
    gradient equals the GPU adjoints applied by hand in sequence, and a
    short step against it lowers the L2 distance to a target image."""
    import torch

    from mray_amd import autograd as ag

    dmin, dmax, cam, W, H, m = (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 32, 32, 24
    gp = GenParams(step_size=2.0 / 64, transform_to_bounds=0, data_min=-0.1, data_max=1.1, compact_image=0,
                   prediction_coding=0)
    rp = RecParams(step_size=2.0 / 64, bias=1e-3, early_ray_termination=1.5, medium_emission=1.0)
    table = torch.from_numpy(tfmod.table_from_rgba8(synth.tf_ramps(0), rp.step_size * 4.0)).cuda()
    target_vol = torch.from_numpy(synth.marschner_lobb(32)).cuda()
    with torch.no_grad():
        target = ag.render(target_vol, dmin, dmax, cam, gp, W, H, m, table, rp)
    x = (target_vol * 0.8 + 0.1).clone().requires_grad_(True)
    img = ag.render(x, dmin, dmax, cam, gp, W, H, m, table, rp)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    # the same chain by hand
    vol = hip.Volume(x.detach(), dmin, dmax)
    mo, _, _ = hip.generate(vol, cam, gp, W, H, m, form="mixed")
    lag = hip.prepare(mo, W, H, m, rp.bias, form="mixed")
    rgba, _ = hip.reconstruct(cam, dmin, dmax, lag, W, H, m, table, rp, form="mixed")
    g_lag, _ = hip.reconstruct_backward(cam, dmin, dmax, lag, W, H, m, table, rp, rgba, 2 * (rgba - target),
                                        want_tf=False)
    g_mo = hip.prepare_backward(mo, W, H, m, rp.bias, g_lag)
    g_vol = hip.generate_backward(vol, cam, gp, W, H, m, g_mo)
    torch.cuda.synchronize()
    rel = float((x.grad - g_vol).norm() / g_vol.norm())
    print(f"[render autograd] |autograd - manual| / |manual| = {rel:.2e}")
    assert rel <= 1e-5
    # a short step against the gradient (largest voxel change 1e-3) lowers the loss
    with torch.no_grad():
        x2 = x - (1e-3 / float(x.grad.abs().max())) * x.grad
        l2 = float(((ag.render(x2, dmin, dmax, cam, gp, W, H, m, table, rp) - target) ** 2).sum())
    print(f"[render autograd] loss {float(loss):.6e} -> {l2:.6e} after a step of 1e-3 against the gradient")
    assert l2 < float(loss)
