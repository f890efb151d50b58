#!/usr/bin/env python3
"""Throughput of the E-field propagation path on one MI355X.

Workload (default, id 2): 1024^2 pupil, 9 wavelengths (600-660 nm) x 3 point
sources (0, 2, 6 lambda/D) through aperture/apodizer, phase aberration, a
217-hex phase FPM (standard 8-ring layout), Lyot stop and a 128^2 image:
27 fields per step, each carried pupil -> FPM -> Lyot -> image plus |E|^2 and
the 3-10 lambda/D dark-hole mean.  Inputs are synthetic (Philox, seed
20261203 + id) and resident in HBM before timing; the calibration
(unocculted) propagation is computed once and amortised.

One JSON line: value = propagations/s of the mixed-precision path that meets the
1e-11 contrast gate on this workload (bf16x6); the FP64 plain path is timed on the
same workload beside it, with the contrast error of each mixed precision against
it (on this workload and on a dark-hole chain), other precisions (bf16x3), the
roofline of the contraction kernel and a CPU baseline (the oracle restatement of
csim, OpenMP, timed on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAKS = {  # MI355X dense peaks (TFLOP/s)
    "fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0, "bf16x6": 2500.0,
}
PRODUCTS = {"fp64": 1, "fp32": 1, "bf16x3": 3, "bf16x6": 6}


def build_chain(ctx, w, precision):
    from csim_amd import workloads as W
    from csim_amd.chain import Chain
    ch = Chain(ctx, w.script)
    W.apply(ch, w)
    ch.set_precision(precision)
    return ch


def time_chain(ch, steps, warmup, prof_steps=5):
    """Timed steps with every stage replayed from one HIP graph (the production path),
    then a separate pass of prof_steps eager steps with HIP events around each
    contraction launch pair (the per-kernel timer turns graph replay off)."""
    import torch
    ch.set_graph(True)
    for _ in range(max(3, warmup)):                # eager run, graph capture, first replay
        ch.contrast()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        c, _ = ch.contrast()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ch.set_profiling(True)
    for _ in range(prof_steps):
        ch.contrast()
    torch.cuda.synchronize()
    prof = ch.profile()
    ch.set_profiling(False)
    return dt, c, prof


def traffic_profile(key):
    return os.path.join("profiles", f"{key}_workload2_hbm_traffic.md")


def measured_traffic(key):
    """HBM bytes per launch of kernel `key` from the committed FETCH_SIZE / WRITE_SIZE
    summary of this workload (rocprofv3 --pmc passes of `bench.py --pmc-probe`)."""
    import re
    path = os.path.join(HERE, traffic_profile(key))
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch\[" + re.escape(key) + r"\]: (\d+)", txt, re.M)
    return int(m.group(1)) if m else None


def gemm_roofline(prof, precision, with_traffic=True):
    """Roofline of the contraction kernel from the in-process timer; `traffic` is the
    bench workload's measured FETCH/WRITE bytes per launch (with_traffic: the bench
    workload itself, not the single-field legs)."""
    key = f"gemm_{precision}"
    st = prof.get(key)
    if not st or st["launches"] == 0:
        return None
    per_launch_flops = st["flops_total"] / st["launches"]
    avg_s = st["ms_total"] / st["launches"] / 1e3
    achieved = per_launch_flops / avg_s / 1e12
    peak = PEAKS[precision]
    traffic = measured_traffic(key) if with_traffic else None
    return {"bound": "mfma", "kernel": key, "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic,
            "traffic_source": traffic_profile(key) if traffic else None,
            "algorithmic_bytes_per_launch": st["bytes_total"] / st["launches"],
            "avg_launch_ms": round(st["ms_total"] / st["launches"], 4),
            "flops_per_launch": per_launch_flops, "mfma_products": PRODUCTS[precision],
            "issued_frac": round(achieved * PRODUCTS[precision] / peak, 4)}


def region_mask(ch):
    """3-10 lambda/D annulus of the chain's output grid (regionContrast block)."""
    inf = ch.info()
    n, d, loD = inf["out_rows"], inf["out_pixel_size"], inf["lambda_over_d"]
    x = -(n * d - d) / 2 + np.arange(n) * d
    r = np.sqrt(x[None, :] ** 2 + x[:, None] ** 2) / loD
    return (r >= 3) & (r <= 10)


def host_cpu():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, budget_s=12.0, cand_sags=None, what="pupil->FPM->Lyot->image"):
    """The oracle (FP64 CPU restatement of csim, Bluestein zoomFft, OpenMP over columns,
    the oracle's own FFT: FFTW is not on this host) on a bounded sample of the same
    workload's fields."""
    import oracle
    from oracle.chain import OracleChain
    oc = OracleChain(w.script, w.arrays, sags=w.meta.get("sags"))
    nl = len(oc.m.lambdas)
    nc = 1 if cand_sags is None else len(cand_sags)
    n = 0
    t0 = time.perf_counter()
    while True:
        k = n % (oc.m.nsrc * nl * nc)
        cand, r = divmod(k, oc.m.nsrc * nl)
        s, li = divmod(r, nl)
        oc.field(s, li, cand_sags=None if cand_sags is None else cand_sags[cand])
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": round(n / dt, 4), "unit": "propagations/s", "cores": oracle.num_threads(), "kind": "port",
            "host_cpu": host_cpu(), "fft": "the oracle's radix-2 FFT (FFTW absent)",
            "sample": f"{n} propagations ({what}) cycling over the workload's {oc.m.nsrc * nl * nc} fields, "
                      f"Bluestein chirp-z zoomFft, {dt:.1f} s"}


def leg_single(ctx, w, precision, steps, what):
    """One-field configurations (C0, C1): FP64 plain and the mixed precision beside it,
    propagations/s, the contraction roofline, the per-stage timer, and the mixed
    path's contrast error against FP64 on the workload's own region."""
    out = {}
    res = {}
    for p in ("fp64", precision):
        ch = build_chain(ctx, w, p)
        dt, c, prof = time_chain(ch, steps, 3)
        res[p] = (c, ch.normalized_image(), ch)
        out[p] = {"value": round(w.meta["fields"] * steps / dt, 3), "unit": "propagations/s",
                  "ms_per_propagation": round(dt / steps * 1e3, 4), "roofline": gemm_roofline(prof, p, False),
                  "stages": {k: {"launches": v["launches"], "ms_per_launch": round(v["ms_total"] / max(1, v["launches"]), 4),
                                 "bytes_per_launch": v["bytes_total"] / max(1, v["launches"])}
                             for k, v in prof.items()}}
    c64, img64, ch64 = res["fp64"]
    cm, imgm, chm = res[precision]
    sel = region_mask(ch64)
    out["speedup_vs_fp64"] = round(out[precision]["value"] / out["fp64"]["value"], 3)
    out["contrast_error"] = {"region_mean_fp64": float(c64[0]), "region_mean_abs": float(abs(cm[0] - c64[0])),
                             "max_pixel_abs": float(np.max(np.abs(imgm - img64)[sel])),
                             "pass_on_timed_workload": float(abs(cm[0] - c64[0])) <= 1e-11}
    out["workload"] = what
    for _, _, ch in res.values():
        ch.close()
    return out


def leg_c1_fresnel(leg):
    """The 2048^2 chain's Fresnel step against HBM.  Its three FFT passes carry the
    elements around it: the first row pass generates the on-axis point-source field and
    applies the pupil mask, the aberration and the first PIAA sag as it reads, the last
    row pass applies the second PIAA sag as it writes.  Compulsory bytes of the fused
    step: the four FP64 maps read once and the field written once, 2048^2 x (4 x 8 + 16) B
    = 201,326,592 B.  Each pass also reads and writes the whole field (row transforms,
    column transforms with the transfer function between them, row transforms), which
    puts the three-pass floor at 2048^2 x (3 x 8 + 16 + 2 x 16 + 16 + 8 + 16) B =
    469,762,048 B."""
    st = leg["fp64"]["stages"]
    if "fresnel_as" not in st:
        return None
    t = st["fresnel_as"]["ms_per_launch"] * 1e-3
    comp, three = 201326592, 469762048
    return {"bound": "hbm", "peak": 8000.0, "unit": "GB/s", "step_ms": round(t * 1e3, 4),
            "compulsory_bytes": comp, "achieved": round(comp / t / 1e9, 1), "frac": round(comp / t / 8e12, 4),
            "three_pass_bytes": three, "three_pass_achieved": round(three / t / 1e9, 1),
            "three_pass_frac": round(three / t / 8e12, 4)}


def de_group(ctx, precision):
    """The DE generation with uneven pupils: 32 candidates (random thicknesses, Philox
    seed 20261203 + 3) on a 512^2 and 32 on a 1024^2 chain of workload 3, 9 wavelengths
    each, merged into one chain group.  Returns (group, chains, [(workload, sags)])."""
    from csim_amd import workloads as W
    from csim_amd.chain import ChainGroup
    rng = np.random.Generator(np.random.Philox(20261203 + 3))
    chains, members = [], []
    for n in (512, 1024):
        w = W.workload3(n=n, n_lambda=9)
        t = rng.random((32, w.meta["n_hex"])) * 1e-6
        ch = build_chain(ctx, w, precision)
        ch.set_candidates("fpm", W.thickness_to_sag(t))
        ch.contrast()                        # compile, tune, calibrate
        chains.append(ch)
        members.append((w, W.thickness_to_sag(t)))
    return ChainGroup(chains), chains, members


def de_batch_uneven(ctx, precision, steps=5, cpu=False):
    """DE generation with uneven pupils: 64 candidates x 9 wavelengths, 32 on 512^2
    and 32 on 1024^2 pupils (radii halved), one grouped launch per contraction
    stage; beside it, the two chains run one after the other."""
    import torch
    g, chains, members = de_group(ctx, precision)
    g.contrast()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.contrast()
    torch.cuda.synchronize()
    dt_g = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        for ch in chains:
            ch.contrast()
    torch.cuda.synchronize()
    dt_s = (time.perf_counter() - t0) / steps
    fields = 2 * 32 * 9
    # csim's work per propagation: 1024^2 6,576,668,672 flop, 512^2 1,946,157,056 (8 per complex
    # MAC); the candidates share their pupil-plane fields, whose first contraction runs once
    flops = 32 * 9 * (6576668672 + 1946157056)
    res = {"fields_per_generation": fields, "ms_per_generation": round(dt_g * 1e3, 3),
           "propagations_per_s": round(fields / dt_g, 1), "effective_tflops": round(flops / dt_g / 1e12, 1),
           "grouped_mft_stages_per_generation": g.merged_launches(),
           "ms_per_generation_two_chains_separately": round(dt_s * 1e3, 3)}
    g.close()
    for ch in chains:
        ch.close()
    if cpu:
        # the oracle on candidates of each pupil size; a generation is 288 fields of each
        cb = [cpu_baseline(wm, 3.0, cand_sags=sg[:4]) for wm, sg in members]
        t_gen = 288 / cb[0]["value"] + 288 / cb[1]["value"]
        res["cpu_baseline"] = {"value": round(fields / t_gen, 4), "unit": "propagations/s", "cores": cb[0]["cores"],
                               "kind": "port", "host_cpu": cb[0]["host_cpu"], "fft": cb[0]["fft"],
                               "sample": f"512^2: {cb[0]['sample']}; 1024^2: {cb[1]['sample']}; rate of a "
                                         "generation (288 fields of each size) from the two"}
    return res


def pmc_probe(precision, workload="w2"):
    """Short run for hardware-counter and kernel-trace passes (rocprofv3): a plain
    device copy of known size (512 MiB read, 512 MiB written) that calibrates the byte
    counters, then the workload's steps with graph replay off (every kernel its own
    dispatch).  A one-element fill (`probe_marker`, a torch elementwise kernel) is
    launched right before the last step, so the last step's dispatches are the ones
    that follow it in the trace."""
    import torch
    from csim_amd import workloads as W
    from csim_amd.ops import Context
    x = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(2):
        y.copy_(x)
    torch.cuda.synchronize()
    del x, y
    ctx = Context(0)
    if workload == "c3":
        ch, chains, _ = de_group(ctx, precision)
        for c in chains:
            c.set_graph(False)
    else:
        w = {"w2": W.workload2, "c0": W.workload0, "c1": W.workload1}[workload]()
        ch = build_chain(ctx, w, precision)
        ch.set_graph(False)
        chains = [ch]
    for _ in range(2):
        ch.contrast()
    torch.cuda.synchronize()
    marker = torch.zeros(1, dtype=torch.int32, device="cuda")
    marker.add_(7)
    torch.cuda.synchronize()
    ch.contrast()
    torch.cuda.synchronize()
    if workload == "c3":
        ch.close()
    for c in chains:
        c.close()
    print(f"probe done: workload {workload}, precision {precision}", flush=True)
    return 0


def contrast_errors(ctx, w, precision, c_mix, img_mix, c64, img64, sel):
    """Error of a mixed-precision run against the FP64 plain path in contrast units, on
    the timed workload itself (the 3-10 lambda/D region mean of regionContrast, and the
    per-pixel maximum there) and on a real dark hole (the same pupil and wavelengths,
    on-axis source, 3 lambda/D occulter, mean C ~ 3e-6)."""
    from csim_amd import workloads as W
    d = float(abs(c_mix[0] - c64[0]))
    err = {"gate": 1e-11, "metric": "absolute error of the 3-10 lambda/D region mean of I / max(I_cal)",
           "region_mean_fp64": float(c64[0]), "region_mean_abs": d, "region_mean_rel": d / float(c64[0]),
           "max_pixel_abs": float(np.max(np.abs(img_mix - img64)[sel])),
           "pass_on_timed_workload": d <= 1e-11}
    wd = W.workload2_dark()
    dk64, dkm = build_chain(ctx, wd, "fp64"), build_chain(ctx, wd, precision)
    cd64, _ = dk64.contrast()
    cdm, _ = dkm.contrast()
    seld = region_mask(dk64)
    err.update({"dark_hole_mean_fp64": float(cd64[0]), "dark_hole_mean_abs": float(abs(cdm[0] - cd64[0])),
                "dark_hole_max_pixel_abs": float(np.max(np.abs(dkm.normalized_image() - dk64.normalized_image())[seld]))})
    err["pass_on_dark_hole"] = err["dark_hole_mean_abs"] <= 1e-11
    for ch in (dk64, dkm):
        ch.close()
    return err


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="bf16x6", choices=["fp64", "fp32", "bf16x3", "bf16x6"],
                    help="headline precision (bf16x6: the one that meets the 1e-11 gate on this workload)")
    ap.add_argument("--also", default="bf16x3", help="comma list of further precisions timed beside it ('' none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="skip the C0 and C1 single-field legs")
    ap.add_argument("--pmc-probe", action="store_true", help="short run for rocprofv3 counter and trace passes")
    ap.add_argument("--probe-workload", default="w2", choices=["w2", "c0", "c1", "c3"],
                    help="workload of --pmc-probe: w2 the 27-field bench workload, c0 one 1024^2 field, "
                         "c1 the 2048^2 PIAA + Fresnel chain, c3 the DE generation with uneven pupils")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print("bench.py measures one GPU: every field shards independently, so the whole-job rate at N GPUs is "
              "N times this one-GPU figure; multi-GPU runs are not part of this build.", file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 1
    if a.pmc_probe:
        return pmc_probe(a.precision, a.probe_workload)
    if a.precision == "fp64":
        print("--precision fp64 is for --pmc-probe only: the bench times FP64 beside the mixed path", file=sys.stderr)
        return 2
    from csim_amd import _lib
    from csim_amd import workloads as W
    from csim_amd.ops import Context
    ctx = Context(0)
    w = W.workload2()
    fields = w.meta["fields"]

    mix = build_chain(ctx, w, a.precision)
    dt_mix, c_mix, prof_mix = time_chain(mix, a.steps, a.warmup)
    img_mix = mix.normalized_image()
    out = {
        "metric": "propagations/s at 1024^2 pupil: mixed vs FP64 plain, contrast err <= 1e-11",
        "value": round(fields * a.steps / dt_mix, 3),
        "unit": "propagations/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(dt_mix / a.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": a.precision, "data": "synthetic (Philox seed 20261203+2)",
        "config": {"workload": "1024^2 pupil, 9 lambda 600-660 nm x 3 sources (0,2,6 l/D), 217-hex FPM, "
                               "Lyot 0.9D, 128^2 image over +-25 l/D; 27 fields per step",
                   "fields_per_step": fields, "precision": a.precision,
                   "timing": "HIP-graph replay of the compiled chain; per-kernel HIP events in a separate pass"},
    }
    rf = gemm_roofline(prof_mix, a.precision)
    if rf:
        out["roofline"] = rf
    out["kernels_mixed"] = prof_mix
    if not a.no_fp64:
        plain = build_chain(ctx, w, "fp64")
        steps64 = max(3, a.steps // 2)
        dt64, c64, prof64 = time_chain(plain, steps64, 1)
        img64 = plain.normalized_image()
        sel = region_mask(plain)
        out["fp64_plain"] = {"value": round(fields * steps64 / dt64, 3), "unit": "propagations/s",
                             "ms_per_step": round(dt64 / steps64 * 1e3, 4),
                             "roofline": gemm_roofline(prof64, "fp64")}
        out["speedup_vs_fp64"] = round(out["value"] / out["fp64_plain"]["value"], 3)
        out["contrast_error"] = contrast_errors(ctx, w, a.precision, c_mix, img_mix, c64, img64, sel)
        others = {}
        for p in [q for q in a.also.split(",") if q and q != a.precision]:
            ch = build_chain(ctx, w, p)
            dt, c, prof = time_chain(ch, a.steps, a.warmup)
            others[p] = {"value": round(fields * a.steps / dt, 3), "unit": "propagations/s",
                         "ms_per_step": round(dt / a.steps * 1e3, 4),
                         "speedup_vs_fp64": round(fields * a.steps / dt / out["fp64_plain"]["value"], 3),
                         "roofline": gemm_roofline(prof, p),
                         "contrast_error": contrast_errors(ctx, w, p, c, ch.normalized_image(), c64, img64, sel)}
            ch.close()
        if others:
            out["other_precisions"] = others
        plain.close()
    if not a.no_uneven:
        out["de_batch_uneven_512_1024"] = de_batch_uneven(ctx, a.precision, cpu=not a.no_cpu)
    if not a.no_legs:
        out["C0_mono_1024"] = leg_single(ctx, W.workload0(), a.precision, max(10, a.steps), "C0: 1024^2 pupil, "
                                         "650 nm, occulter 3 l/D, Lyot 0.9D, 128^2 image; one field per step")
        l1 = leg_single(ctx, W.workload1(), a.precision, max(5, a.steps // 2), "C1: 2048^2 field, two PIAA "
                        "sags around a 1 m angular-spectrum step at 650 nm, downsample x2, then the C0 chain")
        l1["fresnel_roofline"] = leg_c1_fresnel(l1)
        out["C1_piaa_fresnel_2048"] = l1
        if not a.no_cpu:
            out["C0_mono_1024"]["cpu_baseline"] = cpu_baseline(W.workload0(), 4.0)
            out["C1_piaa_fresnel_2048"]["cpu_baseline"] = cpu_baseline(
                W.workload1(), 6.0, what="PIAA sags, Fresnel AS, downsample, pupil->FPM->Lyot->image")
    out["tuner"] = {k: v for k, v in _lib.tune_report().items() if k != "choices"}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
