"""Moment-image benchmark on one MI355X.

Workload (the precision-study configuration): synthetic 1024^3 fp32
mixing-layer volume, perspective camera, 1024x1024 rays, 50 moments,
step 0.5 voxel. One step = generate + prepare + reconstruct of one frame
with the volume already resident in HBM.

Prints ONE JSON line: Mrays/s of the MIXED form (value), the PLAIN form
beside it, per-kernel times, the roofline of the dominant kernel and a CPU
baseline (the C oracle, FP32 reference arithmetic, OpenMP) timed on a bounded
row sample of the same frame.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP32_VALU = 157.3e12  # MI355X FP32 vector (packed) FLOP/s
PEAK_HBM = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024, help="volume edge")
    ap.add_argument("--res", type=int, default=1024, help="image edge")
    ap.add_argument("--moments", type=int, default=50)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-row-step", type=int, default=16, help="CPU baseline: every k-th row")
    ap.add_argument("--cpu-runs", type=int, default=3, help="CPU baseline: runs, the median is reported")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision-row-step", type=int, default=64,
                    help="image RMSE against the FP64 oracle on every k-th row")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward measurement")
    ap.add_argument("--no-dvr", action="store_true", help="skip the reference direct volume renderer")
    ap.add_argument("--no-quant", action="store_true", help="skip the moment-image quantization")
    ap.add_argument("--no-large", action="store_true", help="skip the large noise-and-blobs volume")
    ap.add_argument("--large-only", action="store_true", help="time only the large noise-and-blobs volume")
    ap.add_argument("--no-particles", action="store_true", help="skip the 20 M particle workload")
    ap.add_argument("--particles-only", action="store_true", help="time only the 20 M particle workload")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="time only the adjoint workload and print its figures (for kernel profiles)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (HBM counter calibration)")
    return ap.parse_args()


def flop_generate(ng, m):
    # contract formula, generate part: N_g (8(m-1) + 30) per ray
    return float(np.sum(ng.astype(np.float64))) * (8 * (m - 1) + 30)


def flop_reconstruct(nr, m):
    return float(np.sum(nr.astype(np.float64))) * (8 * (m - 1) + 48)


def flop_prepare(npix, m):
    return float(npix) * 20 * m * m


KERNEL_SYMBOL = {"generate": "generate_mixed_kernel", "prepare": "prepare_mixed_kernel",
                 "reconstruct": "reconstruct_mixed_kernel"}


def measured_traffic(kernel, n, res, m):
    """HBM bytes per dispatch of a mixed-form kernel on the default workload,
    from the calibrated FETCH_SIZE / WRITE_SIZE counter passes summarised in
    profiles/ (counters cannot be read from inside the profiled program)."""
    if (n, res, m) != (1024, 1024, 50):
        return None
    f = ROOT / "profiles" / "hbm_traffic_mixing_layer_1024.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic" and parts[1] == KERNEL_SYMBOL[kernel]:
            return int(parts[2])
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU sharding is described in DESIGN.md "
                                   "and not run", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from mray_amd import hip, synth, tf as tfmod
    from mray_amd.types import GenParams, RecParams

    hip.init(0)
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_ratio(args.steps)}), flush=True)
        return
    if args.large_only:
        print(json.dumps({"large_volume": large_volume(args.steps)}), flush=True)
        return
    if args.particles_only:
        print(json.dumps({"particles": particles_frame(args.steps)}), flush=True)
        return
    n, W, H, m = args.n, args.res, args.res, args.moments
    dmin, dmax = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), fov_half_deg=25.0, znear=0.01, aspect=W / H)
    step = 0.5 / n
    gp = GenParams(step_size=step, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0, bias=1e-4)
    # MESE bias 1e-3: at 50 moments the sharp mixing layer needs it for float
    # moment images to be well posed (lerp toward uniform-density moments)
    rp = RecParams(step_size=step, bias=1e-3, early_ray_termination=0.99, medium_emission=1.0)
    # opacity scale (the reference's preintegration_step_size_factor): the ramps
    # of the seed-0 transfer function only cover the thin interface, so a large
    # factor gives a partly opaque image (mean alpha ~0.2)
    tf_factor = 256.0
    table = torch.from_numpy(tfmod.table_from_rgba8(synth.tf_ramps(0), rp.step_size * tf_factor)).cuda()

    print(f"[bench] generating {n}^3 mixing layer on the GPU", file=sys.stderr, flush=True)
    vox = hip.synth_mixing_layer(n, seed=1)
    vol = hip.Volume(vox, dmin, dmax)
    torch.cuda.synchronize()

    if args.copy_mib:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters
        # calibrate the per-dispatch HBM traffic read for the kernels below
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    mo = torch.empty((H, W, m), dtype=torch.float32, device="cuda")
    lag = torch.empty_like(mo)
    # the mixed form keeps its multipliers in FP64 between prepare and march
    lag64 = torch.empty((H, W, m), dtype=torch.float64, device="cuda")
    bnd = torch.empty(H * W, dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")

    def prep(form, bias, out_f32=None):
        if form == "mixed":
            return hip.prepare(mo, W, H, m, bias, form="mixed64", out=lag64)
        return hip.prepare(mo, W, H, m, bias, form=form, out=lag if out_f32 is None else out_f32)

    # untimed pass: per-ray sample counts for the algorithmic FLOP count
    _, _, ng = hip.generate(vol, cam, gp, W, H, m, form="mixed", moments=mo, bounds=bnd, want_samples=True)
    prep("mixed", rp.bias)
    _, nr = hip.reconstruct(cam, dmin, dmax, lag64, W, H, m, table, rp, form="mixed", bounds=bnd, rgba=rgba,
                            want_steps=True)
    torch.cuda.synchronize()
    ng, nr = ng.cpu().numpy(), nr.cpu().numpy()
    alpha_mean = float(rgba[..., 3].mean())

    def run(form, steps, warmup):
        ev = {k: [] for k in ("gen", "prep", "rec")}
        for i in range(warmup + steps):
            e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
            e[0].record()
            hip.generate(vol, cam, gp, W, H, m, form=form, moments=mo, bounds=bnd)
            e[1].record()
            lg = prep(form, rp.bias)
            e[2].record()
            hip.reconstruct(cam, dmin, dmax, lg, W, H, m, table, rp, form=form, bounds=bnd, rgba=rgba)
            e[3].record()
            if i >= warmup:
                ev["gen"].append(e)
        torch.cuda.synchronize()
        return ev

    print(f"[bench] mixed: {args.warmup} warmup + {args.steps} timed steps", file=sys.stderr, flush=True)
    run("mixed", 0, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev = run("mixed", args.steps, 0)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_gen = np.mean([e[0].elapsed_time(e[1]) for e in ev["gen"]])
    ms_prep = np.mean([e[1].elapsed_time(e[2]) for e in ev["gen"]])
    ms_rec = np.mean([e[2].elapsed_time(e[3]) for e in ev["gen"]])
    ms_step = wall * 1e3 / args.steps
    mrays = W * H / (ms_step * 1e-3) / 1e6

    plain = None
    if args.plain_steps > 0:
        print(f"[bench] plain: 1 warmup + {args.plain_steps} timed steps", file=sys.stderr, flush=True)
        run("plain", 0, 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        evp = run("plain", args.plain_steps, 0)
        torch.cuda.synchronize()
        wall_p = time.perf_counter() - t0
        ms_step_p = wall_p * 1e3 / args.plain_steps
        plain = {"value": round(W * H / (ms_step_p * 1e-3) / 1e6, 3), "ms_per_step": round(ms_step_p, 3),
                 "ms": {k: round(float(np.mean([e[i].elapsed_time(e[i + 1]) for e in evp["gen"]])), 3)
                        for i, k in enumerate(("generate", "prepare", "reconstruct"))}}

    f_gen, f_rec, f_prep = flop_generate(ng, m), flop_reconstruct(nr, m), flop_prepare(W * H, m)
    kernels = {"generate": (ms_gen, f_gen), "reconstruct": (ms_rec, f_rec), "prepare": (ms_prep, f_prep)}
    dom = max(kernels, key=lambda k: kernels[k][0])
    ms_dom, f_dom = kernels[dom]
    achieved = f_dom / (ms_dom * 1e-3)
    total_flop = f_gen + f_rec + f_prep

    adjoint = None if args.no_adjoint else adjoint_ratio(args.steps)
    adjoint_c1 = None
    if not args.no_adjoint:
        adjoint_c1 = adjoint_frame(vol, vox, cam, dmin, dmax, gp, rp, table, W, H, m, args.steps,
                                   f"the timed frame ({n}^3 mixing layer, {W}x{H} rays, {m} moments)")

    # the reference direct volume renderer (the ground truth of the moment
    # images) on the same volume, camera, table and step: one frame per step
    dvr = None
    if not args.no_dvr:
        from mray_amd.types import RefParams

        refp = RefParams(step_size=step, medium_emission=rp.medium_emission,
                         early_ray_termination=rp.early_ray_termination)
        dvr_rgba = torch.empty_like(rgba)
        hip.reference(vol, cam, table, refp, W, H, rgba=dvr_rgba)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            hip.reference(vol, cam, table, refp, W, H, rgba=dvr_rgba)
        e1.record()
        torch.cuda.synchronize()
        ms_dvr = e0.elapsed_time(e1) / args.steps
        dvr = {"workload": "reference DVR, same volume, camera, table and step", "ms_per_frame": round(ms_dvr, 3),
               "Mrays_per_s": round(W * H / (ms_dvr * 1e-3) / 1e6, 3)}

    # quantization of the frame's moment image for a moment-image file (10-bit
    # words, the [-1/pi, 1/pi] range of moments 1..m-1): encode and decode per
    # call, the bit-offset scan and the length read-back included
    quant = None
    if not args.no_quant:
        qindex = torch.arange(0, (W * H + 1) * m, m, dtype=torch.int64, device="cuda").to(torch.int32)
        qbits = hip.quantization_table(m, False, 10)
        qs = hip.quantize(mo, qindex, W, H, False, qbits)
        dq = hip.dequantize(qs, qindex, W, H, False, qbits)
        torch.cuda.synchronize()
        times = {}
        for name, fn in (("encode", lambda: hip.quantize(mo, qindex, W, H, False, qbits)),
                         ("decode", lambda: hip.dequantize(qs, qindex, W, H, False, qbits))):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(args.steps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            times[name] = e0.elapsed_time(e1) / args.steps
        # algorithmic bytes: the floats, the stream and the index (one pass each)
        qbytes = W * H * m * 4 + qs.numel() + (W * H + 1) * 4
        quant = {"workload": f"the {W}x{H} x {m} moment image, 10-bit words", "stream_bytes": int(qs.numel()),
                 "ms": {k: round(v, 3) for k, v in times.items()},
                 "GB_per_s": {k: round(qbytes / (v * 1e-3) / 1e9, 1) for k, v in times.items()}}
        del qs, dq, qindex
        # the quantization-curve search of generate with prediction coding and
        # 10 bits: the frame compacted (error threshold 1e-3) and coded, every
        # 16th row and column (the CLI's 64x64 sample)
        cdata, cindex = hip.compact_image(mo, W, H, m, 1e-3, bounds=bnd)
        cparams = hip.prediction_encode(cdata, cindex, W, H, m, 0, 1e-4)
        torch.cuda.synchronize()
        hi, hd = cindex.cpu().numpy().view(np.uint32).astype(np.int64), cdata.cpu().numpy()
        sw, f = 64, W // 64
        pix = [(y * f) * W + x * f for y in range(sw) for x in range(sw)]
        s_index = np.concatenate([[0], np.cumsum(hi[1:][pix] - hi[:-1][pix])]).astype(np.uint32)
        s_data = np.concatenate([hd[hi[p]:hi[p + 1]] for p in pix]).astype(np.float32)
        d_si, d_sd = torch.from_numpy(s_index.view(np.int32)).cuda(), torch.from_numpy(s_data).cuda()
        t0 = time.perf_counter()
        curve = hip.curve_table(d_sd, d_si, sw, sw, m, cparams, 10)
        quant["curve_search"] = {"workload": f"{sw}x{sw} sample of the coded frame, {int(s_index[-1])} moments",
                                 "ms": round((time.perf_counter() - t0) * 1e3, 2), "table": curve}
        # the arithmetic coder of the reference eval's file options (compact,
        # prediction-coded, 10-bit curve table, entropy_coding=1): the section
        # from device moments to host bytes and back, per call
        sec = hip.entropy_encode(cdata, cindex, W, H, m, True, curve)
        times = {}
        for name, fn in (("encode", lambda: hip.entropy_encode(cdata, cindex, W, H, m, True, curve)),
                         ("decode", lambda: hip.entropy_decode(sec, cindex, W, H, m, True, curve))):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(max(1, args.steps // 2)):
                fn()
            torch.cuda.synchronize()
            times[name] = (time.perf_counter() - t0) * 1e3 / max(1, args.steps // 2)
        nmom = int(hi[-1])
        quant["entropy_coding"] = {"workload": f"the compacted, prediction-coded frame ({nmom} moments), "
                                               f"curve table, one stream per moment",
                                   "section_bytes": int(sec.size), "bits_per_moment": round(8 * sec.size / nmom, 3),
                                   "ms": {k: round(v, 2) for k, v in times.items()},
                                   "Msymbols_per_s": {k: round(nmom / (v * 1e-3) / 1e6, 1) for k, v in times.items()}}
        del cdata, cindex

    # images of both forms at the bench's MESE bias and at the reference
    # eval's 5e-5 (eval/marschner_lobb/marschner_lobb.json), for the precision
    # figures against the FP64 oracle
    gpu_images = {}
    if not args.no_precision:
        for form in ("mixed", "plain"):
            hip.generate(vol, cam, gp, W, H, m, form=form, moments=mo, bounds=bnd)
            for bias in (rp.bias, 5e-5):
                lg = prep(form, bias)
                img, st = hip.reconstruct(cam, dmin, dmax, lg, W, H, m, table,
                                          RecParams(step_size=step, bias=bias, early_ray_termination=0.99,
                                                    medium_emission=1.0), form=form, bounds=bnd, want_steps=True)
                gpu_images[(form, bias)] = (img.cpu().numpy(), st.cpu().numpy())
        torch.cuda.synchronize()

    large = None if args.no_large else large_volume(args.steps)
    particles = None if args.no_particles else particles_frame(args.steps)

    cpu, fp32_rows = None, None
    vol_np = None
    if not args.no_cpu or not args.no_precision:
        print("[bench] copying the volume to the host", file=sys.stderr, flush=True)
        vol_np = vox.cpu().numpy()
    if not args.no_cpu:
        cpu, rows32 = cpu_baseline(vol_np, dmin, dmax, cam, W, H, m, gp, rp, args.cpu_row_step, tf_factor,
                                   runs=args.cpu_runs)
        if args.precision_row_step % args.cpu_row_step == 0:
            fp32_rows = rows32[:2]
    precision = None
    if not args.no_precision:
        sys.path.insert(0, str(ROOT / "tests"))
        from oracle_ffi import Oracle

        O = Oracle()
        if fp32_rows is None:
            mo32, b32, _ = O.generate(vol_np, dmin, dmax, cam, W, H, m, gp, row0=0, row_step=args.precision_row_step)
            fp32_rows = (mo32, b32)
        table_np = tfmod.table_from_rgba8(synth.tf_ramps(0), rp.step_size * tf_factor)
        precision = precision_vs_fp64(O, vol_np, dmin, dmax, cam, W, H, m, gp, rp, table_np, gpu_images, fp32_rows,
                                      args.precision_row_step)

    line = {
        "metric": "Mrays/s gen+reconstruct, mixed vs plain form at RMSE<=1e-4; adjoint/forward",
        "value": round(mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"mixing layer {n}^3 f32 (resident in HBM), perspective {W}x{H} rays, {m} moments, "
                               f"step 0.5 voxel; generate + prepare + reconstruct, mixed form (FP64 multipliers)",
                   "volume": [n, n, n], "image": [W, H], "moments": m, "form": "mixed"},
        "plain": plain,
        "mixed_over_plain": round(mrays / plain["value"], 3) if plain else None,
        "ms": {"generate": round(ms_gen, 3), "prepare": round(ms_prep, 3), "reconstruct": round(ms_rec, 3)},
        "samples_per_ray": {"generate": round(float(ng.mean()), 1), "reconstruct": round(float(nr.mean()), 1)},
        "image_alpha_mean": round(alpha_mean, 3),
        "roofline": {"bound": "valu", "kernel": dom, "achieved": round(achieved / 1e12, 3),
                     "peak": PEAK_FP32_VALU / 1e12, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_VALU, 4),
                     "traffic": measured_traffic(dom, n, W, m)},
        "roofline_frame": {"achieved_TFLOPs": round(total_flop / (ms_step * 1e-3) / 1e12, 3),
                           "frac": round(total_flop / (ms_step * 1e-3) / PEAK_FP32_VALU, 4)},
        "adjoint": adjoint,
        "adjoint_frame": adjoint_c1,
        "reference_dvr": dvr,
        "large_volume": large,
        "particles": particles,
        "quantization": quant,
        "rmse_vs_fp64": precision,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def adjoint_ratio(steps):
    """Adjoint / forward time on the differentiation configuration: Marschner-Lobb
    256^3 f32, orthographic 512x512 rays, 40 moments, step 0.5 voxel. Forward =
    generate + prepare + reconstruct (mixed); adjoint = the reverse of all three
    (gradient to the voxels and to the 256 x 4 table); the table-only adjoint
    (what transfer-function fitting needs) is the reconstruct reverse alone."""
    import torch

    from mray_amd import hip, synth, tf as tfmod
    from mray_amd.types import GenParams, RecParams

    n, W, H, m = 256, 512, 512, 40
    dmin, dmax = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    cam = synth.ortho_camera()
    step = 1.0 / n
    gp = GenParams(step_size=step, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    rp = RecParams(step_size=step, bias=5e-5, early_ray_termination=0.99, medium_emission=1.0)
    table = torch.from_numpy(tfmod.table_from_rgba8(synth.tf_ramps(0), step * 4.0)).cuda()
    print("[bench] adjoint: Marschner-Lobb 256^3, 512x512, 40 moments", file=sys.stderr, flush=True)
    vox = hip.synth_marschner_lobb(n)
    vol = hip.Volume(vox, dmin, dmax)
    g = torch.randn((H, W, 4), device="cuda", generator=torch.Generator(device="cuda").manual_seed(0))
    gvol = torch.zeros_like(vox)
    # per-ray normalisation written by the forward and read by the reverse
    # (what the autograd wrapper keeps between the two)
    rr = torch.empty((H * W, 2), dtype=torch.float32, device="cuda")

    def fwd():
        mo, b, _ = hip.generate(vol, cam, gp, W, H, m, form="mixed", ray_range=rr)
        lag = hip.prepare(mo, W, H, m, rp.bias, form="mixed")
        rgba, _ = hip.reconstruct(cam, dmin, dmax, lag, W, H, m, table, rp, form="mixed", bounds=b)
        return mo, b, lag, rgba

    def adj(mo, b, lag, rgba, tf_only):
        g_lag, g_tf = hip.reconstruct_backward(cam, dmin, dmax, lag, W, H, m, table, rp, rgba, g, bounds=b,
                                               want_lagrange=not tf_only)
        if tf_only:
            return
        g_mo = hip.prepare_backward(mo, W, H, m, rp.bias, g_lag)
        gvol.zero_()
        hip.generate_backward(vol, cam, gp, W, H, m, g_mo, out=gvol, ray_range=rr)

    def timed(fn, k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / k

    state = fwd()
    adj(*state, False)
    adj(*state, True)
    k = max(2, steps)
    ms_fwd = timed(fwd, k)
    ms_adj = timed(lambda: adj(*state, False), k)
    ms_tf = timed(lambda: adj(*state, True), k)
    return {"workload": "Marschner-Lobb 256^3 f32, ortho 512x512, 40 moments, step 0.5 voxel",
            "forward_ms": round(ms_fwd, 3), "adjoint_ms": round(ms_adj, 3), "adjoint_over_forward": round(
                ms_adj / ms_fwd, 3), "table_adjoint_ms": round(ms_tf, 3),
            "table_adjoint_over_forward": round(ms_tf / ms_fwd, 3)}


def host_info():
    """CPU model (lscpu) and the cores this process may run on."""
    import subprocess

    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return {"cpu_model": model, "nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count()}


def large_volume(steps, plain_steps=1):
    """The large-volume workload: 2048 x 2048 x 1920 uint8 noise-and-blobs
    volume (7.5 GB, resident), perspective 2048 x 2048 rays, step 1 voxel, 100
    moments generated, per-pixel counts from the compaction at error
    threshold 1e-3 clamped to [8, 100], compact preparation and
    reconstruction. One frame = generate + compaction (counts, clamp, index,
    compact copy) + prepare + reconstruct."""
    import torch

    from mray_amd import hip, synth, tf as tfmod
    from mray_amd.types import GenParams, RecParams

    nz, ny, nx = synth.blob_volume_shape()
    W = H = 2048
    m = 100
    dmin, dmax = (0.0, 0.0, 0.0), (1.0, ny / nx, nz / nx)
    cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), fov_half_deg=25.0, znear=0.01, aspect=1.0)
    step = 1.0 / nx
    gp = GenParams(step_size=step, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=1,
                   prediction_coding=0, error_threshold=1e-3, bias=1e-4)
    rp = RecParams(step_size=step, bias=1e-3, early_ray_termination=0.99, medium_emission=1.0)
    table = torch.from_numpy(tfmod.table_from_rgba8(synth.tf_ramps(0), step * 128.0)).cuda()
    print(f"[bench] large volume: generating {nx}x{ny}x{nz} uint8 on the GPU", file=sys.stderr, flush=True)
    vox = hip.synth_noise_blobs(nx, ny, nz, seed=2)
    vol = hip.Volume(vox, dmin, dmax)
    mo = torch.empty((H, W, m), dtype=torch.float32, device="cuda")
    bnd = torch.empty(H * W, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()

    def frame(form, ev=None, want=False):
        rec = lambda i: ev[i].record() if ev else None
        rec(0)
        _, _, ng = hip.generate(vol, cam, gp, W, H, m, form=form, moments=mo, bounds=bnd, want_samples=want)
        rec(1)
        data, index = hip.compact_image(mo, W, H, m, gp.error_threshold, bounds=bnd, count_range=(8, 100))
        rec(2)
        lag = hip.prepare(data, W, H, m, rp.bias, form="mixed64" if form == "mixed" else form, index=index)
        rec(3)
        img, nr = hip.reconstruct(cam, dmin, dmax, lag, W, H, m, table, rp, form=form, index=index, bounds=bnd,
                                  want_steps=want)
        rec(4)
        return ng, nr, index, img

    ng, nr, index, img = frame("mixed", want=True)
    torch.cuda.synchronize()
    ng, nr = ng.cpu().numpy().astype(np.float64), nr.cpu().numpy().astype(np.float64)
    counts = np.diff(index.cpu().numpy().view(np.uint32).astype(np.int64)).astype(np.float64)
    alpha_mean = float(img[..., 3].mean())
    del index, img

    def timed(form, k):
        evs = []
        t0 = time.perf_counter()
        for _ in range(k):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
            frame(form, ev)
            evs.append(ev)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / k
        parts = {name: round(float(np.mean([e[i].elapsed_time(e[i + 1]) for e in evs])), 3)
                 for i, name in enumerate(("generate", "compaction", "prepare", "reconstruct"))}
        return wall, parts

    print("[bench] large volume: mixed", file=sys.stderr, flush=True)
    frame("mixed")
    ms_mixed, parts = timed("mixed", max(2, steps // 2))
    out = {"workload": f"noise-and-blobs {nx}x{ny}x{nz} uint8 (resident), perspective {W}x{H} rays, step 1 voxel, "
                       f"{m} moments generated, compaction 1e-3 clamped to [8, 100], compact prepare + reconstruct",
           "Mrays_per_s": round(W * H / (ms_mixed * 1e-3) / 1e6, 3), "ms_per_frame": round(ms_mixed, 3),
           "ms": parts, "mean_moments": round(float(counts[counts > 0].mean()), 2),
           "empty_pixels": int(np.count_nonzero(counts == 0)),
           "samples_per_ray": {"generate": round(float(ng.mean()), 1), "reconstruct": round(float(nr.mean()), 1)},
           "image_alpha_mean": round(alpha_mean, 3)}
    # contract formula: generate at m, compaction 6 * 20 m^2 per pixel, then
    # each pixel's own count in prepare and reconstruct
    flop = {"generate": float(ng.sum()) * (8 * (m - 1) + 30), "compaction": W * H * 6.0 * 20 * m * m,
            "prepare": float((20.0 * counts * counts).sum()),
            "reconstruct": float((nr * (8 * (np.maximum(counts, 1) - 1) + 48)).sum())}
    out["roofline"] = {k: round(flop[k] / (parts[k] * 1e-3) / PEAK_FP32_VALU, 4) for k in flop}
    out["roofline_frame"] = round(sum(flop.values()) / (ms_mixed * 1e-3) / PEAK_FP32_VALU, 4)
    if plain_steps > 0:
        print("[bench] large volume: plain", file=sys.stderr, flush=True)
        ms_plain, parts_p = timed("plain", plain_steps)
        out["plain"] = {"Mrays_per_s": round(W * H / (ms_plain * 1e-3) / 1e6, 3), "ms_per_frame": round(ms_plain, 3),
                        "ms": parts_p}
        out["mixed_over_plain"] = round(ms_plain / ms_mixed, 3)
    del vol, vox, mo
    torch.cuda.empty_cache()
    return out


def particles_frame(steps, n_particles=20_000_000):
    """The particle workload: 20 M particles from a mixture of 32 anisotropic
    Gaussians in [0,1] x [0,1] x [0,0.5] (seed 5, uniform mass), smoothing
    length 0.004, gathered onto a 512 x 512 x 256 grid with the reference's
    Particles::sample, then a mixed frame at 1024 x 1024 rays, 40 moments,
    step 0.5 voxel, and its adjoint back to the grid values."""
    import torch

    from mray_amd import hip, synth, tf as tfmod
    from mray_amd.types import GenParams, RecParams

    nx, ny, nz = 512, 512, 256
    W = H = 1024
    m = 40
    h = 0.004
    dmin, dmax = (0.0, 0.0, 0.0), (1.0, 1.0, 0.5)
    print(f"[bench] particles: {n_particles} particles (host generator)", file=sys.stderr, flush=True)
    x, y, z, v = synth.particle_mixture(n_particles, seed=5, n_gauss=32, box=(1.0, 1.0, 0.5))
    dx, dy, dz, dv = (torch.from_numpy(a).cuda() for a in (x, y, z, v))
    del x, y, z, v
    grid = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
    hip.particles_to_grid(dx, dy, dz, dv, h, dmin, dmax, nx, ny, nz, out=grid)
    torch.cuda.synchronize()
    print(f"[bench] particles: gather onto {nx}x{ny}x{nz}", file=sys.stderr, flush=True)
    k = max(2, steps // 2)
    t0 = time.perf_counter()
    for _ in range(k):
        hip.particles_to_grid(dx, dy, dz, dv, h, dmin, dmax, nx, ny, nz, out=grid)
    torch.cuda.synchronize()
    ms_splat = (time.perf_counter() - t0) * 1e3 / k
    occupied = float((grid > 0).float().mean())
    del dx, dy, dz, dv
    vol = hip.Volume(grid, dmin, dmax)
    step = 0.5 / nx
    cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), fov_half_deg=25.0, znear=0.01, aspect=1.0)
    gp = GenParams(step_size=step, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0, bias=1e-4)
    rp = RecParams(step_size=step, bias=1e-3, early_ray_termination=0.99, medium_emission=1.0)
    table = torch.from_numpy(tfmod.table_from_rgba8(synth.tf_ramps(0), step * 64.0)).cuda()
    adj = adjoint_frame(vol, grid, cam, dmin, dmax, gp, rp, table, W, H, m, steps,
                        f"{W}x{H} rays, {m} moments, step 0.5 voxel on the gathered grid")
    out = {"workload": f"{n_particles} particles (32-Gaussian mixture, h = {h}) gathered onto {nx}x{ny}x{nz} f32, "
                       f"perspective {W}x{H} rays, {m} moments, step 0.5 voxel; mixed forward and its adjoint "
                       f"to the grid values",
           "gather_ms": round(ms_splat, 3), "gather_Mvoxels_per_s": round(nx * ny * nz / (ms_splat * 1e-3) / 1e6, 1),
           "occupied_voxels": round(occupied, 4), "forward_ms": adj["forward_ms"], "adjoint_ms": adj["adjoint_ms"],
           "adjoint_over_forward": adj["adjoint_over_forward"],
           "Mrays_per_s": round(W * H / (adj["forward_ms"] * 1e-3) / 1e6, 3)}
    del vol, grid
    torch.cuda.empty_cache()
    return out


def adjoint_frame(vol, vox, cam, dmin, dmax, gp, rp, table, W, H, m, steps, workload):
    """Adjoint / forward time on a frame of the timed workload: forward =
    generate + prepare + reconstruct (mixed), adjoint = reconstruct reverse +
    prepare reverse + generate reverse to the voxels (a gradient volume of the
    volume's size, zeroed per pass)."""
    import torch

    from mray_amd import hip

    print(f"[bench] adjoint on {workload}", file=sys.stderr, flush=True)
    g = torch.randn((H, W, 4), device="cuda", generator=torch.Generator(device="cuda").manual_seed(0))
    gvol = torch.zeros_like(vox)
    rr = torch.empty((H * W, 2), dtype=torch.float32, device="cuda")

    def fwd():
        mo, b, _ = hip.generate(vol, cam, gp, W, H, m, form="mixed", ray_range=rr)
        lag = hip.prepare(mo, W, H, m, rp.bias, form="mixed")
        rgba, _ = hip.reconstruct(cam, dmin, dmax, lag, W, H, m, table, rp, form="mixed", bounds=b)
        return mo, b, lag, rgba

    def adj(mo, b, lag, rgba):
        g_lag, _ = hip.reconstruct_backward(cam, dmin, dmax, lag, W, H, m, table, rp, rgba, g, bounds=b)
        g_mo = hip.prepare_backward(mo, W, H, m, rp.bias, g_lag)
        gvol.zero_()
        hip.generate_backward(vol, cam, gp, W, H, m, g_mo, out=gvol, ray_range=rr)

    def timed(fn, k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / k

    state = fwd()
    adj(*state)
    k = max(2, steps)
    ms_fwd = timed(fwd, k)
    ms_adj = timed(lambda: adj(*state), k)
    del gvol
    return {"workload": workload, "forward_ms": round(ms_fwd, 3), "adjoint_ms": round(ms_adj, 3),
            "adjoint_over_forward": round(ms_adj / ms_fwd, 3)}


def cpu_baseline(vol, dmin, dmax, cam, W, H, m, gp, rp, row_step, tf_factor, runs=3):
    """The C oracle (the reference CPU path's loops and FP32 arithmetic, OpenMP
    over rows with the default static schedule) on every row_step-th row of the
    same frame, extrapolated to Mrays/s: the median of `runs` runs. Threads:
    OMP_NUM_THREADS, the host cores this job is given (else every core it may
    run on). Returns the baseline and the FP32 moments, bounds and image of the
    sampled rows (for the precision figures)."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_ffi import Oracle
    from mray_amd import synth, tf as tfmod

    O = Oracle()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    os.environ["MRAY_ORACLE_THREADS"] = str(threads)
    info = host_info()
    print(f"[bench] cpu baseline: {threads} threads ({info['cpu_model']}), every {row_step}th row, {runs} runs",
          file=sys.stderr, flush=True)
    table = tfmod.table_from_rgba8(synth.tf_ramps(0), rp.step_size * tf_factor)
    t_gen, t_rec = [], []
    for r in range(runs):
        t0 = time.perf_counter()
        mo, b, _ = O.generate(vol, dmin, dmax, cam, W, H, m, gp, row0=0, row_step=row_step)
        t1 = time.perf_counter()
        rgba, st = O.reconstruct(cam, dmin, dmax, W, H, m, mo.reshape(-1), table, rp, bounds=b, row0=0,
                                 row_step=row_step)
        t2 = time.perf_counter()
        t_gen.append(t1 - t0)
        t_rec.append(t2 - t1)
        print(f"[bench] cpu baseline run {r + 1}/{runs}: generate {t1 - t0:.2f} s, prepare+reconstruct "
              f"{t2 - t1:.2f} s", file=sys.stderr, flush=True)
    tot = [a + b_ for a, b_ in zip(t_gen, t_rec)]
    k = int(np.argsort(tot)[len(tot) // 2])
    rows = (H + row_step - 1) // row_step
    rays = rows * W
    base = {"value": round(rays / tot[k] / 1e6, 6), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"rows 0,{row_step},... ({rays} of {W * H} rays) of the same frame, median of {runs} runs "
                      f"(generate {t_gen[k]:.2f} s + prepare/reconstruct {t_rec[k]:.2f} s); "
                      f"{threads} OpenMP threads",
            "omp_max_threads": int(O.lib.orc_omp_max_threads()), **info}
    return base, (mo, b, rgba, st)


def precision_vs_fp64(O, vol, dmin, dmax, cam, W, H, m, gp, rp, table_np, gpu_images, fp32, row_step):
    """Float RGBA RMSE against the FP64 oracle (the reference's discretisation
    in FP64 with exact functions) on every row_step-th row of the frame, for
    the GPU images given per (form, bias) and for the FP32 oracle; rays whose
    early-termination step count differs from the FP64 march (a flip on
    rounding at the threshold) are counted apart."""
    from mray_amd.types import RecParams

    rows = np.arange(0, H, row_step)
    print(f"[bench] precision: FP64 oracle on {len(rows)} rows", file=sys.stderr, flush=True)
    mo64, b64 = O.generate(vol, dmin, dmax, cam, W, H, m, gp, f64=True, row0=0, row_step=row_step)
    mo32, b32 = fp32
    out = {"rows": f"0,{row_step},... ({len(rows) * W} rays)"}
    for bias in sorted({k[1] for k in gpu_images}, reverse=True):
        rpb = RecParams(step_size=rp.step_size, bias=bias, early_ray_termination=rp.early_ray_termination,
                        medium_emission=rp.medium_emission)
        r64, s64 = O.reconstruct_f64(cam, dmin, dmax, W, H, m, mo64, table_np, rpb, bounds=b64, row0=0,
                                     row_step=row_step)
        r32, s32 = O.reconstruct(cam, dmin, dmax, W, H, m, mo32.reshape(-1), table_np, rpb, bounds=b32, row0=0,
                                 row_step=row_step)
        ref = r64[rows].reshape(-1, 4)
        st64 = s64.reshape(H, W)[rows].reshape(-1)
        entry = {}
        imgs = dict(gpu_images)
        imgs[("fp32_oracle", bias)] = (r32, s32)
        for (form, bb), (img, st) in imgs.items():
            if bb != bias:
                continue
            d = img[rows].reshape(-1, 4).astype(np.float64) - ref
            same = st.reshape(H, W)[rows].reshape(-1) == st64
            entry[form] = {"rmse": float(f"{np.sqrt((d[same] ** 2).mean()):.3e}"),
                           "max": float(f"{np.abs(d[same]).max():.3e}"),
                           "rmse_all_rays": float(f"{np.sqrt((d ** 2).mean()):.3e}"),
                           "ert_flips": int(np.count_nonzero(~same))}
        out[f"bias_{bias:g}"] = entry
    return out


if __name__ == "__main__":
    main()
