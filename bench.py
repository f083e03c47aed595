#!/usr/bin/env python
"""Benchmark of the hot path: two-pass V-BM3D (BM3Dv2, radius=1) at 1080p.

Workload (one step): GRAYS float32 1920x1080, 30-frame synthetic video
(frame 0 = sinusoids + rectangles, later frames shifted by integer motion
U{-3..3}, fresh N(0, 15/255) noise per frame, seed 1), parameters sigma=15,
block_step=4, bm_range=9, ps_num=2, ps_range=4, radius=1:
    basic = BM3Dv2(src, radius=1)            -> bm3d_execute_vagg
    final = BM3Dv2(src, ref=basic, radius=1) -> bm3d_execute_vagg
(each pass runs the BM3D evaluations over a ring of frames and the
VAggregate average straight from their accumulation regions; the per-frame
stacks of the two-stage form are never written -- the outputs are identical).
Launch shapes (tiles, waves, frames per launch) come from the plans'
autotuner, run on the same clip before the timed region (--no-autotune:
the plans' defaults). The clip is resident in HBM before the timed region;
value = frames/s of
the whole two-pass job over K steps, for the mixed-precision path
(--precision, default mixed: FP16 block-matching pre-scoring with exact FP32
rescoring of the best candidates). The plain FP32 path is timed the same way
beside it (fp32_value, mixed_speedup = value / fp32_value).

roofline: the block matching (dominant) of the reported path, bound by VALU
issue (FP16 dot2 pre-scores, FP32 rescoring; no MFMA). On the mixed path it
is two kernels per evaluation batch, the spatial search by strips
(bm3d_match_dx_spatial) then the temporal search (bm3d_match_dx), timed as
one launch. Algorithmic FLOPs per launch = 192 x (in-plane candidates
evaluated in that launch), the candidate count taken from the kernel's own
counter; duration from HIP events bracketing each launch on its stream
inside the timed region; peak = MI355X's published vector peak, 157.3
TFLOP/s. roofline_filter: the same
for the filtering kernel, with the reference's DCT flop counts per group.

quality: PSNR of both paths against the clean clip; share of group lists
the mixed path shares with the FP32 path; and, on a 192x160 crop of the
first 5 frames, max-abs error and PSNR delta of both GPU paths against the
FP64 CPU oracle running the same pipeline.

cpu_baseline: the CPU oracle (a C restatement of the reference algorithm,
scalar), frame-parallel on host threads (one per core, at most 16: a GPU
box's CPU share), on a bounded sample: the basic and the final BM3D window
(radius 1) of 8 whole 1080p frames as 16 independent jobs (about a minute of
CPU work, a few seconds of wall time), the CPU model stated.

batch_64_clips (main line; --no-batch skips it): the 64-clip mix below,
per-clip FP32 plans against one autotuned mixed-precision batch graph.

host workload (--workload host): the same clip in pageable host memory,
through the plugin-level BM3Dv2 filters' host-frame route (whole-clip calls
and per-frame calls from 4 threads), beside the same filters on device
frames. Prints its own JSON line (PCIe-inclusive frames/s).

batch workload (--workload batch): 64 seeded GRAYS clips of 8 frames, each
drawing its size (360p..2160p), radius (0..2) and sigma (5..25); two-pass
BM3Dv2 through one captured batch graph, against the same clips run one by
one through FP32 plans. Prints its own JSON line (aggregate Mpixel/s).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision mixed|fp32]
                [--workload vbm3d|batch|host]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "vapoursynth-bm3dcuda_amd"))

H, W, N = 1080, 1920, 30
PARAMS = dict(sigma=15.0, block_step=4, bm_range=9, radius=1, ps_num=2, ps_range=4)
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 peak


def profiled_traffic(kernel, precision):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summaries (lines `traffic <kernel> <precision> <bytes>` under profiles/),
    or None when no summary covers this kernel and precision."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            f = line.split()
            if len(f) == 4 and f[0] == "traffic" and f[1] == kernel and f[2] == precision:
                return int(f[3])
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(noisy, basic, threads, frames=(4, 9, 14, 19, 24, 27, 10, 21)):
    """Oracle timing on host cores (see module docstring): whole 1080p
    windows, basic and final pass of each frame as independent jobs."""
    import numpy as np
    from oracle import oracle as O

    r = PARAMS["radius"]
    sig_b = [O.scale_sigma(PARAMS["sigma"], False)]
    sig_f = [O.scale_sigma(PARAMS["sigma"], True)]
    jobs = []
    for n in frames:
        win = np.ascontiguousarray(np.transpose(noisy[n - r:n + r + 1], (1, 0, 2, 3)))
        bwin = np.ascontiguousarray(np.transpose(basic[n - r:n + r + 1], (1, 0, 2, 3)))
        jobs.append((win, None))
        jobs.append((win, bwin))

    def work(job):
        s, b = job
        if b is None:
            O.bm3d_window(s, None, sig_b, PARAMS["block_step"], PARAMS["bm_range"], r,
                          PARAMS["ps_num"], PARAMS["ps_range"], False, False)
        else:
            O.bm3d_window(s, b, sig_f, PARAMS["block_step"], PARAMS["bm_range"], r,
                          PARAMS["ps_num"], PARAMS["ps_range"], False, True)

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(work, jobs))
    dt = time.perf_counter() - t0
    return {"value": round(len(frames) / dt, 4), "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C, -O2, scalar) two-pass BM3D windows (basic and final, r=1) of "
                      f"{len(frames)} whole 1080p frames, {len(jobs)} jobs on {threads} threads of "
                      f"{cpu_model()} ({os.cpu_count()} logical CPUs visible; a GPU box's share is 16); "
                      f"{dt:.2f} s wall"}


class Pipeline:
    """Two-pass BM3Dv2 (radius 1) over a device-resident clip."""

    def __init__(self, core, src, precision, spatial_mfma=False):
        self.core = core
        self.src = src
        n, _, h, w = src.shape
        self.n = n
        p = PARAMS
        self.pb = core.Plan(w, h, p["sigma"], p["block_step"], p["bm_range"], p["radius"],
                            p["ps_num"], p["ps_range"], final=False, precision=precision)
        self.pf = core.Plan(w, h, p["sigma"], p["block_step"], p["bm_range"], p["radius"],
                            p["ps_num"], p["ps_range"], final=True, precision=precision)
        if spatial_mfma and precision == "mixed":
            # the opt-in matcher whose spatial search runs on the matrix cores
            self.pb.select_kernels("mixed-mfma", "mixed")
            self.pf.select_kernels("mixed-mfma", "mixed")
        import torch
        self.basic = torch.empty_like(src)
        self.final = torch.empty_like(src)

    def autotune(self):
        """Launch shapes for both passes from the plans' autotuner (untimed)."""
        self.pb.execute_vagg(self.src, out=self.basic)
        tb = self.pb.autotune(self.src)
        tf = self.pf.autotune(self.src, self.basic)
        return {"basic": tb, "final": tf}

    def step(self, stream=None):
        # BM3Dv2 = BM3D + VAggregate, fused: the stacks are not materialised
        self.pb.execute_vagg(self.src, out=self.basic, stream=stream)
        self.pf.execute_vagg(self.src, self.basic, out=self.final, stream=stream)


def time_path(core, capi, src, precision, steps, warmup, autotune, spatial_mfma=False):
    import torch
    pipe = Pipeline(core, src, precision, spatial_mfma)
    tuning = pipe.autotune() if autotune else {"basic": pipe.pb.tuning, "final": pipe.pf.tuning}
    log(f"{precision}: launch shapes {tuning}")
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        pipe.step(stream)
    torch.cuda.synchronize()
    capi.profile_reset()
    capi.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pipe.step(stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    capi.profile_enable(False)
    match_ms, match_n = capi.profile_read(capi.KERNEL_MATCH)
    filt_ms, filt_n = capi.profile_read(capi.KERNEL_FILTER)
    capi.profile_reset()
    ncand = pipe.pb.count_candidates(src) + pipe.pf.count_candidates(src, pipe.basic)
    flops_step = 192.0 * ncand
    avg_ms = match_ms / match_n
    flops_launch = flops_step / (match_n / steps)
    achieved = flops_launch / (avg_ms * 1e-3) / 1e12
    # filter: the reference's DCT counts per group (hard 192 (44 + 51), Wiener
    # 192 (2 44 + 51)) plus 6 per coefficient for shrinkage and aggregation
    nblk = pipe.pb.num_blocks
    f_flops_step = src.shape[0] * nblk * (192.0 * (44 + 51) + 192.0 * (2 * 44 + 51) + 2 * 512 * 6)
    f_avg_ms = filt_ms / filt_n
    f_flops_launch = f_flops_step / (filt_n / steps)
    log(f"{precision}: {steps} steps in {dt:.3f}s, match {match_ms / steps:.2f} ms/step, "
        f"filter {filt_ms / steps:.2f} ms/step")
    return {
        "fps": src.shape[0] * steps / dt, "ms_per_step": dt / steps * 1e3,
        "match_ms_step": match_ms / steps, "filter_ms_step": filt_ms / steps,
        "avg_launch_ms": avg_ms, "flops_launch": flops_launch, "achieved_tflops": achieved,
        "ncand": ncand, "pipe": pipe, "tuning": tuning,
        "two_launch_match": precision == "mixed" and pipe.pb.uses_strip_search(),
        "filter_avg_launch_ms": f_avg_ms, "filter_flops_launch": f_flops_launch,
        "filter_achieved_tflops": f_flops_launch / (f_avg_ms * 1e-3) / 1e12,
    }


def oracle_crop_errors(core, noisy, clean):
    """Both GPU paths vs the FP64 oracle on a 192x160 crop of 5 frames."""
    import numpy as np
    import torch
    from bm3d_amd import synth
    from oracle import oracle as O

    crop_n = np.ascontiguousarray(noisy[:5, :, 200:360, 400:592])
    crop_c = np.ascontiguousarray(clean[:5, :, 200:360, 400:592])
    p = PARAMS
    op = O.Params(sigma=(p["sigma"],), block_step=p["block_step"], bm_range=p["bm_range"],
                  radius=p["radius"], ps_num=p["ps_num"], ps_range=p["ps_range"])
    _, f64 = O.two_pass(crop_n.astype(np.float64), op, dtype=np.float64)
    ref_psnr = synth.psnr(f64, crop_c)
    out = {"oracle_fp64_psnr_db": round(ref_psnr, 4)}
    src = torch.from_numpy(crop_n).cuda()
    for prec in ("fp32", "mixed"):
        pipe = Pipeline(core, src, prec)
        pipe.step()
        fin = pipe.final.cpu().numpy()
        out[prec] = {"max_abs_vs_fp64": float(np.abs(fin - f64).max()),
                     "delta_psnr_db": round(synth.psnr(fin, crop_c) - ref_psnr, 5)}
    return out


def batch_workload(n_clips=64, frames=8):
    """The 64-clip serving mix (bm3d_amd.synth.batch_clips): clip k draws its
    resolution from 360p..2160p, radius from {0, 1, 2} and sigma from
    U[5, 25] with seed k; content and integer motion as the main workload,
    noise sd = sigma / 255."""
    from bm3d_amd import synth
    return synth.batch_clips(n_clips, frames)


def run_batch(args, main_line=False):
    """Aggregate Mpixel/s of the 64-clip mix: one batched graph (mixed
    precision, and FP32) vs per-clip FP32 plan launches, two passes each.
    main_line: only per-clip FP32 and the mixed graph, returned as a dict
    for the main JSON line instead of printed."""
    import torch
    from bm3d_amd import core, synth

    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    clips = batch_workload()
    log(f"generated {len(clips)} clips in {time.perf_counter() - t0:.1f}s")
    srcs = [torch.from_numpy(c["noisy"]).cuda() for c in clips]
    outs = [torch.empty_like(s) for s in srcs]
    basics = [torch.empty_like(s) for s in srcs]
    px = sum(s.numel() for s in srcs)  # output pixels of one run (GRAYS)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    # per-clip FP32 launches: the two passes of each clip through its plans
    plans = []
    for c, s in zip(clips, srcs):
        _, _, h, w = s.shape
        kw = dict(synth.BATCH_PARAMS, radius=c["radius"], precision="fp32")
        plans.append((core.Plan(w, h, c["sigma"], **kw), core.Plan(w, h, c["sigma"], final=True, **kw)))

    def per_clip():
        for (pb, pf), s, o, b in zip(plans, srcs, outs, basics):
            pb.execute_vagg(s, out=b)
            pf.execute_vagg(s, b, out=o)

    t_clip = timed(per_clip, args.steps)
    log(f"per-clip FP32: {t_clip * 1e3:.1f} ms per run")
    ref_basic = [b.clone() for b in basics]
    del plans
    res = {"per_clip_fp32_mpixel_s": round(px / t_clip / 1e6, 1)}
    for prec in (("mixed",) if main_line else ("fp32", "mixed")):
        jobs = [dict(synth.BATCH_PARAMS, src=s, out=o, basic=b, passes=2, radius=c["radius"], sigma=[c["sigma"]],
                     precision=prec) for c, s, o, b in zip(clips, srcs, outs, basics)]
        batch = core.Batch(jobs)
        if not args.no_autotune:
            batch.autotune()
        t_b = timed(batch.execute, args.steps)
        info = batch.info
        log(f"batch {prec}: {t_b * 1e3:.1f} ms per run, {info}")
        res[f"batch_{prec}_mpixel_s"] = round(px / t_b / 1e6, 1)
        res[f"batch_{prec}_ms"] = round(t_b * 1e3, 3)
        res[f"batch_{prec}_graph"] = dict(info, launch_shapes=batch.tuning)
        if prec == "fp32":
            # The batched FP32 graph reproduces the per-clip FP32 basic
            # estimates up to the order of the float weight atomics (the
            # final pass matches on the basic estimate, so an ulp there can
            # flip a near-tie group; with extractor_exp both are bit-exact,
            # see tests/test_gpu_batch.py).
            res["batch_fp32_basic_max_abs_vs_per_clip"] = max(float((b - r).abs().max())
                                                              for b, r in zip(basics, ref_basic))
        del batch
    if main_line:
        return {"workload": "64 GRAYS clips x 8 frames, 360p..2160p, radius U{0,1,2}, sigma U[5,25], "
                            "two passes, one captured graph (autotuned)", "pixels_per_run": px,
                "mixed_mpixel_s": res["batch_mixed_mpixel_s"], "mixed_ms": res["batch_mixed_ms"],
                "per_clip_fp32_mpixel_s": res["per_clip_fp32_mpixel_s"],
                "speedup_vs_per_clip_fp32": round(res["batch_mixed_mpixel_s"] / res["per_clip_fp32_mpixel_s"], 3),
                "graph": res["batch_mixed_graph"]}
    out = {
        "metric": "aggregate Mpixel/s, 64-clip batch of two-pass BM3Dv2 (360p..2160p, radius 0..2)",
        "value": res["batch_mixed_mpixel_s"],
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "higher_is_better": True,
        "dtype": "f16+f32",
        "data": "synthetic (64 seeded clips, 8 frames each, sizes/radius/sigma drawn per clip)",
        "config": {"workload": "64 GRAYS clips x 8 frames, resolution in {640x360, 1280x720, 1920x1080, "
                               "3840x2160}, radius U{0,1,2}, sigma U[5,25], block_step=4, bm_range=9, "
                               "ps_num=2, ps_range=4, two passes", "pixels_per_run": px},
        "speedup_vs_per_clip_fp32": round(res["batch_mixed_mpixel_s"] / res["per_clip_fp32_mpixel_s"], 3),
        **res,
    }
    print(json.dumps(out), flush=True)


def run_host(args):
    """The VapourSynth route: the 1080p clip in pageable host memory, through
    the plugin-level BM3Dv2 filters (bm3d_filter_get_frames_host), two
    passes. Whole-clip calls (the filter pipelines upload, compute and
    download in chunks), per-frame calls from 4 threads (as VapourSynth's
    GetFrame; each call carries its own temporal window), and the same
    filters on device-resident frames, for comparison. PCIe-inclusive:
    every input frame crosses to the device and every output frame back."""
    import ctypes
    import threading

    import numpy as np
    import torch
    from bm3d_amd import _capi as C, core, synth, vs

    torch.cuda.set_device(0)
    nf = args.frames
    clean, noisy = synth.make_clip(H, W, nf, 1, noise_sd=15.0, seed=1, motion=True)
    log(f"generated {nf} frames sha256={synth.sha256(noisy)[:16]}")
    prec = args.precision
    if not args.no_autotune:  # fills the process tuning cache the filters' plans read
        Pipeline(core, torch.from_numpy(noisy).cuda(), prec).autotune()
    kw = dict(sigma=[PARAMS["sigma"]], block_step=PARAMS["block_step"], bm_range=PARAMS["bm_range"],
              radius=PARAMS["radius"], ps_num=PARAMS["ps_num"], ps_range=PARAMS["ps_range"], chroma=None,
              device_id=None, fast=None, extractor_exp=None, zero_init=None, precision=prec)
    a = vs._args(**kw)
    host_clip = vs.Clip.from_array(noisy, "GRAY", device="cpu")
    vi = host_clip.video_info()
    hb, hf = ctypes.c_void_p(), ctypes.c_void_p()
    C.check(C.lib().bm3dv2_filter_create(ctypes.byref(a), ctypes.byref(vi), None, ctypes.byref(hb)))
    C.check(C.lib().bm3dv2_filter_create(ctypes.byref(a), ctypes.byref(vi), ctypes.byref(vi), ctypes.byref(hf)))
    basic = np.empty_like(noisy)
    final = np.empty_like(noisy)
    fr_src = host_clip.host_frames(0, nf - 1)
    fr_basic = vs.Clip.from_array(torch.from_numpy(basic), "GRAY", device="cpu")
    basic_t = torch.from_numpy(basic)
    final_t = torch.from_numpy(final)
    basic_c = vs.Clip([basic_t[:, 0]], "GRAY", base=basic_t)
    final_c = vs.Clip([final_t[:, 0]], "GRAY", base=final_t)
    del fr_basic
    fb, ff = basic_c.host_frames(0, nf - 1), final_c.host_frames(0, nf - 1)

    def whole():
        C.check(C.lib().bm3d_filter_get_frames_host(hb, fr_src, None, 0, fb, 0, nf))
        C.check(C.lib().bm3d_filter_get_frames_host(hf, fr_src, fb, 0, ff, 0, nf))

    def per_frame(threads=4):
        errs = []

        def run(h, ref, out):
            def work(t):
                for n in range(t, nf, threads):
                    lo, hi = ctypes.c_int(), ctypes.c_int()
                    C.lib().bm3d_filter_host_range(h, n, 1, ctypes.byref(lo), ctypes.byref(hi))
                    src = ctypes.cast(ctypes.byref(fr_src, lo.value * ctypes.sizeof(C.HostFrame)),
                                      ctypes.POINTER(C.HostFrame))
                    rf = None if ref is None else ctypes.cast(
                        ctypes.byref(ref, lo.value * ctypes.sizeof(C.HostFrame)), ctypes.POINTER(C.HostFrame))
                    o = ctypes.cast(ctypes.byref(out, n * ctypes.sizeof(C.HostFrame)), ctypes.POINTER(C.HostFrame))
                    if C.lib().bm3d_filter_get_frames_host(h, src, rf, lo.value, o, n, 1):
                        errs.append(C.lib().bm3d_last_error().decode())
            ts = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
            for t in ts:
                t.start()
            for t in ts:
                t.join()

        run(hb, None, fb)  # every basic frame before the final pass reads them
        run(hf, fb, ff)
        if errs:
            raise RuntimeError(errs[0])

    def timed(fn, reps):
        fn()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t) / reps

    t_whole = timed(whole, args.steps)
    out_whole = final.copy()
    log(f"host whole-clip: {t_whole * 1e3:.1f} ms per two-pass run")
    t_pf = timed(per_frame, args.steps)
    log(f"host per-frame x4 threads: {t_pf * 1e3:.1f} ms per two-pass run")
    pf_max_abs = float(np.abs(final - out_whole).max())
    # the same filters on device frames (no PCIe in the loop)
    src_d = vs.Clip.from_array(noisy, "GRAY")
    b_d = vs._alloc_like(vi, "GRAY", (0, 0), "cuda")
    f_d = vs._alloc_like(vi, "GRAY", (0, 0), "cuda")
    frs, frb, frf = src_d.frames(), b_d.frames(), f_d.frames()
    stream = core._stream_ptr(torch.cuda.current_stream())

    def device():
        C.check(C.lib().bm3d_filter_get_frames(hb, ctypes.byref(frs), None, ctypes.byref(frb), 0, nf, stream))
        C.check(C.lib().bm3d_filter_get_frames(hf, ctypes.byref(frs), ctypes.byref(frb), ctypes.byref(frf), 0, nf,
                                               stream))
        torch.cuda.synchronize()

    t_dev = timed(device, args.steps)
    dev_max_abs = float(np.abs(f_d.numpy() - out_whole).max())
    C.lib().bm3d_filter_destroy(hb)
    C.lib().bm3d_filter_destroy(hf)
    r = PARAMS["radius"]
    frame_mb = H * W * 4 / 1e6
    out = {
        "metric": "1080p frames/s, two-pass BM3Dv2 radius=1 from host frames (PCIe-inclusive)",
        "value": round(nf / t_whole, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "higher_is_better": True,
        "dtype": "f32" if prec == "fp32" else "f16+f32",
        "data": "synthetic (the vbm3d workload's clip, in pageable host memory)",
        "config": {"workload": "GRAYS 1920x1080 x30 frames, two-pass BM3Dv2 radius=1 through "
                               "bm3d_filter_get_frames_host", "frames": nf, "precision": prec},
        "whole_clip_fps": round(nf / t_whole, 3),
        "per_frame_4_threads_fps": round(nf / t_pf, 3),
        "device_frames_fps": round(nf / t_dev, 3),
        "per_frame_max_abs_vs_whole_clip": pf_max_abs,
        "device_max_abs_vs_host": dev_max_abs,
        "pcie_mb_per_output_frame": {
            # basic: clip up, basic down; final: clip and basic up, final down
            "whole_clip": round(5 * frame_mb, 2),
            # each call uploads its window of 4r + 1 frames (twice in the final pass)
            "per_frame": round(((4 * r + 1) * 3 + 2) * frame_mb, 2),
            # source/source.cpp:316-366: 2r + 1 window slots up (twice in the
            # final pass) and a 2(2r + 1)-frame stack down, per pass
            "reference_per_frame": round(((2 * r + 1) * (1 + 2) + 2 * 2 * (2 * r + 1)) * frame_mb, 2),
        },
        "psnr_final_db": round(synth.psnr(out_whole, clean), 4),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="mixed", choices=["fp32", "mixed"],
                    help="path reported as value (the other is timed beside it)")
    ap.add_argument("--frames", type=int, default=N)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the run, fill and copy a buffer of this many MiB (a known-size "
                         "reference for HBM byte counters)")
    ap.add_argument("--no-compare", action="store_true", help="time only --precision")
    ap.add_argument("--spatial-mfma", action="store_true",
                    help="mixed path: the matcher's spatial search on the matrix cores (opt-in kernel)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the plans' default launch shapes instead of autotuned ones")
    ap.add_argument("--workload", default="vbm3d", choices=["vbm3d", "batch", "host"],
                    help="vbm3d: the 1080p two-pass V-BM3D clip (default); batch: the 64-clip mix; "
                         "host: the vbm3d clip from host memory through the plugin-level filters")
    ap.add_argument("--cpu-threads", type=int, default=min(os.cpu_count() or 1, 16))
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 64-clip batch figures in the main line")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the two-pass V-BM3D path shards by frame range, and "
              "this library runs and reports the single-GPU share only (use --gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from bm3d_amd import _capi, core, synth

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "host":
        run_host(args)
        return
    torch.cuda.set_device(0)
    nf = args.frames
    t0 = time.perf_counter()
    clean, noisy = synth.make_clip(H, W, nf, 1, noise_sd=15.0, seed=1, motion=True)
    log(f"generated {nf} frames in {time.perf_counter() - t0:.1f}s sha256={synth.sha256(noisy)[:16]}")
    src = torch.from_numpy(noisy).cuda()

    paths = [args.precision] if args.no_compare else ["fp32", "mixed"]
    res = {p: time_path(core, _capi, src, p, args.steps, args.warmup, not args.no_autotune, args.spatial_mfma)
           for p in paths}
    main_r = res[args.precision]

    basic_host = main_r["pipe"].basic.cpu().numpy()
    quality = {}
    for p, r in res.items():
        quality[f"psnr_final_{p}_db"] = round(synth.psnr(r["pipe"].final.cpu().numpy(), clean), 4)
    quality["psnr_noisy_db"] = round(synth.psnr(noisy, clean), 4)
    if "fp32" in res and "mixed" in res:
        g32 = res["fp32"]["pipe"].pb.match(src, first=0, count=4).cpu().numpy()
        gmx = res["mixed"]["pipe"].pb.match(src, first=0, count=4).cpu().numpy()
        quality["mixed_group_agreement_basic"] = float(np.all(g32 == gmx, axis=-1).mean())
        b32 = res["fp32"]["pipe"].basic
        g32f = res["fp32"]["pipe"].pf.match(src, b32, first=0, count=4).cpu().numpy()
        gmxf = res["mixed"]["pipe"].pf.match(src, b32, first=0, count=4).cpu().numpy()
        quality["mixed_group_agreement_final"] = float(np.all(g32f == gmxf, axis=-1).mean())
        log("oracle FP64 crop comparison ...")
        quality["vs_fp64_oracle_crop"] = oracle_crop_errors(core, noisy, clean)

    out = {
        "metric": "1080p frames/s, two-pass BM3Dv2 radius=1 (V-BM3D)",
        "value": round(main_r["fps"], 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f16+f32",
        "data": "synthetic (seeded sinusoids+rectangles, integer motion, N(0,15/255) noise)",
        "config": {"workload": "GRAYS 1920x1080 x30 frames, two-pass BM3Dv2 radius=1, sigma=15, "
                               "block_step=4, bm_range=9, ps_num=2, ps_range=4",
                   "frames": nf, "precision": args.precision,
                   "launch_shapes": res[args.precision]["tuning"]},
        "mpixel_per_s": round(main_r["fps"] * W * H / 1e6, 2),
        "roofline": {
            "bound": "valu",
            "compute_pipe": "VALU: FP16 v_dot2 pre-scores (mixed) / FP32 FMA (fp32); peak = MI355X "
                            "published vector peak (FP32 with packed math, = FP16 dot2 rate); no MFMA",
            # the mixed path's matching is two launches per evaluation batch
            # (the spatial search by strips, then the temporal search, each
            # a one-job multi-job kernel reading its arguments from a table);
            # a "launch" here is the pair, timed together
            "kernel": ("bm3d_match_dx_spatial_multi+bm3d_match_dx_temporal_multi" if main_r["two_launch_match"] else
                       "bm3d_match_dx" if args.precision == "mixed" else "bm3d_match"),
            "achieved": round(main_r["achieved_tflops"], 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(main_r["achieved_tflops"] / PEAK_FP32_TFLOPS, 4),
            "traffic": profiled_traffic("bm3d_match_dx" if args.precision == "mixed" else "bm3d_match",
                                        args.precision),
            "flops_per_launch": main_r["flops_launch"],
            "avg_launch_ms": round(main_r["avg_launch_ms"], 4),
            "candidates_per_step": main_r["ncand"],
        },
        "roofline_filter": {
            "bound": "valu",
            "compute_pipe": "VALU + LDS (FP16 MFMA for the 3-D transforms in the mixed kernel); achieved "
                            "counts the reference's DCT flops, not the MFMA products",
            "kernel": "bm3d_filter_mx" if args.precision == "mixed" else "bm3d_filter",
            "achieved": round(main_r["filter_achieved_tflops"], 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(main_r["filter_achieved_tflops"] / PEAK_FP32_TFLOPS, 4),
            "traffic": profiled_traffic("bm3d_filter_mx" if args.precision == "mixed" else "bm3d_filter",
                                        args.precision),
            "flops_per_launch": main_r["filter_flops_launch"],
            "avg_launch_ms": round(main_r["filter_avg_launch_ms"], 4),
        },
        "kernel_ms_per_step": {p: {"bm3d_match": round(r["match_ms_step"], 3),
                                   "bm3d_filter": round(r["filter_ms_step"], 3)}
                               for p, r in res.items()},
        "quality": quality,
    }
    if "fp32" in res and "mixed" in res:
        out["fp32_value"] = round(res["fp32"]["fps"], 3)
        out["mixed_value"] = round(res["mixed"]["fps"], 3)
        out["mixed_speedup"] = round(res["mixed"]["fps"] / res["fp32"]["fps"], 4)
    if not args.no_batch:
        del res
        torch.cuda.empty_cache()
        out["batch_64_clips"] = run_batch(args, main_line=True)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(noisy, basic_host, args.cpu_threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    if args.copy_mb > 0:
        a = torch.empty(args.copy_mb << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        a.fill_(1.0)   # writes copy_mb MiB
        b.copy_(a)     # reads and writes copy_mb MiB
        torch.cuda.synchronize()
        log(f"copy reference: fill + copy of {args.copy_mb} MiB done")


if __name__ == "__main__":
    main()
