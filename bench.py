#!/usr/bin/env python3
"""Benchmark of the point-cloud hologram path on one MI355X.

Headline workload (one "step"): the 1,000,000-point cloud with complex
amplitudes (x, y uniform in +-8 mm, z in [0.4, 0.6] m, |a| in [0.1, 1],
arg a in [0, 2 pi), SplitMix64 seed 20261202+3) rendered with Fresnel
diffraction at 3840x2160, 3.74 um pitch, 532 nm on the BF16 3-term
matrix-core kernel (H = U diag(a) V^T), then phase-encoded and normalised to
8 bits -- everything resident in HBM before the timed region starts.

value = point-pixel pairs per second (G/s) over the whole step.
Reported beside it:
  * field relative L2 against the library's FP64 probe on sampled pixels
    (the probe is pinned to the extended-precision oracle by the test suite
    at every workload geometry), per diffraction mode;
  * p50/p99 ms per hologram, per diffraction mode;
  * the headline kernel's roofline (matrix-core FLOPs, HIP events on the
    context's stream; HBM traffic from rocprofv3 counters);
  * the same workload on the FP32 direct kernel and the other matrix-core
    variants (pairs/s, rel-L2, speed-up over the direct kernel);
  * the 100k-point RGB RS + Fresnel workload (3840x2160, FP32 direct), the
    10k-point RS workload (1920x1080, FP32 and FP64) and one step of the
    2M-point RGB workload (7680x4320, RS FP32 direct + Fresnel FP16 3-term);
  * RS on the matrix cores (BF16 1- and 3-term) beside the FP32 direct kernel
    on the 100k-point and the 2M-point RGB workloads;
  * encode/normalise HBM rates, the WRP generator, the serving stream;
  * the CPU baseline (the reference's OpenMP CPU path, restated in oracle/,
    timed on a bounded sample on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# FP32 FLOPs per (point, pixel) pair in the direct kernels' separable pair
# loops (FMA = 2): RS 6 FMA (the cross term folded into the column factor),
# Fresnel 4 FMA.  The per-lane column factor (a few FMAs and one v_sin/v_cos
# per point, shared by 8 rows) is not counted.
FLOP_PER_PAIR = {0: 12, 1: 8}
PEAK_FP32_TFLOPS = 157.3          # published MI355X FP32 vector peak
# FP64 mode, pair loop of the separable form (FMA = 2): RS series correction
# + complex products, Fresnel one complex multiply-add
FLOP_PER_PAIR_FP64 = {0: 37, 1: 8}
PEAK_FP64_TFLOPS = 78.6           # published MI355X FP64 vector peak
PEAK_BF16_TFLOPS = 2500.0         # published dense BF16/FP16 matrix peak
MODE_GPU, MODE_FLOAT = 1, 2
KERNELS = {"direct": MODE_GPU | MODE_FLOAT, "fp64": MODE_GPU, "bf16x1": MODE_GPU | MODE_FLOAT | 0x10,
           "bf16x3": MODE_GPU | MODE_FLOAT | 0x20, "fp16x3": MODE_GPU | MODE_FLOAT | 0x40}
MFMA_TERMS = {"bf16x1": 1, "bf16x3": 3, "fp16x3": 3}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=3, choices=[1, 2, 3, 5])
    ap.add_argument("--points", type=int, default=0, help="override point count (0 = workload's)")
    ap.add_argument("--kernel", default="bf16x3", choices=list(KERNELS))
    ap.add_argument("--fresnel-kernel", default=None, choices=list(KERNELS),
                    help="kernel for the Fresnel part of the step (default: --kernel)")
    ap.add_argument("--rs-kernel", default=None, choices=list(KERNELS),
                    help="kernel for the RS part of the step (default: --kernel if direct / fp64, else direct)")
    ap.add_argument("--variants", type=int, default=1,
                    help="also time the 1M-point Fresnel workload on every Fresnel kernel (0 = off)")
    ap.add_argument("--blocks", default="rgb2160,rs2160,rs1080,large4320,rs4320",
                    help="other workloads to time beside the headline (comma list, '' = none)")
    ap.add_argument("--serving", type=int, default=1,
                    help="also run the mixed Poisson request stream on the serving engine (0 = off)")
    ap.add_argument("--requests", type=int, default=120)
    ap.add_argument("--only-serving-burst", action="store_true",
                    help="run only the serving stream's burst pass and print its figures (for kernel traces)")
    ap.add_argument("--serving-rs-kernel", default="bf16x3", choices=["direct", "bf16x1", "bf16x3"],
                    help="kernel of the RS requests in the serving stream (bf16x3: matrix cores for the points "
                         "whose tile expansion holds, FP32 direct for the rest)")
    ap.add_argument("--wrp", type=int, default=1, help="also time the WRP generator (0 = off)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run ohx_field_accumulate over two 3840x2160 FP64 fields (reads 265.4 MB, "
                         "writes 132.7 MB), the known-size kernel that scales FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--probe-pixels", type=int, default=512)
    return ap.parse_args()


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, target_s: float):
    """The reference's OpenMP CPU path (oracle/ restatement), bounded sample:
    a leading subset of the workload's points on the full pixel grid, all of
    the workload's channels and diffraction modes.  The subset is sized so
    that one run takes about target_s / 5; three runs give the median, and a
    single-thread run on 1/threads of the points gives the 1-thread rate.
    The leading points are an unbiased sample of the cloud (i.i.d. draws), so
    pairs/s carries over to the full workload."""
    from oracle import oracle as orc
    from openholo_amd import synth
    omp_env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(omp_env or 0) or (os.cpu_count() or 1)
    g = orc.make_geom(w.pn, w.pp)

    def run(n, nthr):
        v = synth.points(w, n)
        t0 = time.perf_counter()
        for mode in w.modes:
            orc.gen_restate(g, v, w.lambdas, mode, n_threads=nthr)
        return time.perf_counter() - t0

    per_pt = lambda n: w.n_pixels * len(w.lambdas) * len(w.modes) * n
    # the reference parallelises over points: start at one point per thread
    n = min(w.n_points, threads)
    while True:
        t = run(n, threads)
        if t >= target_s / 5 or n >= w.n_points:
            break
        n = min(w.n_points, max(n + threads, int(n * target_s / 5 / max(t, 1e-3))))
    times = sorted([t] + [run(n, threads) for _ in range(2)])
    t_med = times[1]
    n1 = max(1, n // threads)
    t1 = run(n1, 1)
    return {"value": per_pt(n) / t_med / 1e9, "unit": "Gpairs/s", "cores": threads, "kind": "port",
            "sample": f"first {n} of {w.n_points} points x {w.pn[0]}x{w.pn[1]} px x "
                      f"{len(w.lambdas)} ch x modes {list(w.modes)}: {per_pt(n):.3e} pairs, median of 3 runs "
                      f"{t_med:.2f} s (runs {', '.join(f'{x:.2f}' for x in times)} s) "
                      f"(oracle/ restatement of genCghPointCloudCPU, g++ -O3 -fopenmp)",
            "one_thread_Gpairs_per_s": per_pt(n1) / t1 / 1e9,
            "one_thread_sample": f"first {n1} points, {t1:.2f} s",
            "host": {"cpu_model": _cpu_model(), "nproc": os.cpu_count(), "OMP_NUM_THREADS": omp_env or None}}


# HBM bytes per launch of the headline kernel (k_pc_prep + k_fresnel_gemm +
# k_gemm_reduce) at the 1M-point 3840x2160 workload: rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE summed over the launch's dispatches, each scaled by
# the factor the same counter reads on a copy kernel of known size;
# profiles/traffic_1m_2160p.md gives the counts, the factors and the commit.
# k_fresnel_gemm<BF16x3> alone (128 x 256 tiles, 64-byte point records):
# 883.8 MB read (the per-stage point records re-read per tile, mostly L2
# hits) + 1.062 GB written (the split-K planes); the whole launch (prep,
# records, GEMM, split-K reduce): 3.59 GB.
TRAFFIC = {"bf16x3": 1.95e9}
TRAFFIC_LAUNCH = {"bf16x3": 3.59e9}
# Speed of light of the direct kernels' separable pair loops (per lane and point,
# 8 rows: column factor + 8 pairs) on this
# hardware (openholo_amd/bin/valu_rates); profiles/valu_mfma_rates_gfx950.md.
MIX_SOL_PAIRS_PER_S = {0: 8.27e12, 1: 11.43e12}


def wrp_block(ctx, capi, synth, Event, launches=3):
    """The WRP generator (ohx_wrp_generate) on the reference's WRP example
    geometry (Colab/Config/Generation_WRP (RGB).xml: 1920x1080 at 10 um, WRP
    2.7 mm in front of the deepest point, hologram 0.5 m behind it) with
    10k and 100k synthetic points, three channels: ms per hologram (HIP events
    on the launch stream) and window terms per second."""
    out = {"geometry": "1920x1080 at 10 um, RGB 638/520/450 nm, dz 2.7 mm, d 0.5 m"}
    pn, pp, lams = (1920, 1080), (1e-5, 1e-5), (638e-9, 520e-9, 450e-9)
    stream = ctx.stream
    for n in (10_000, 100_000):
        v = synth.random_cloud(20261202 + 6, n, 1.5e-3, (0.0, 0.3e-3))
        zmax = float(v["pos"][:, 2].max())
        dv = ctx.to_device(v)
        fld = ctx.alloc(16 * pn[0] * pn[1])
        ps = [capi.wrp_params(pn, pp, lam, zmax + 2.7e-3, 0.5, channel=ch) for ch, lam in enumerate(lams)]
        for p in ps:
            ctx.wrp_generate(dv, n, p, fld)
        ctx.sync()
        terms = 0
        ev = []
        for _ in range(launches):
            a, b = Event(), Event()
            a.record(stream)
            for p in ps:
                ctx.wrp_generate(dv, n, p, fld)
                terms += ctx.last_stats().pairs_nominal
            b.record(stream)
            ev.append((a, b))
        ctx.sync()
        ms = float(np.median([a.elapsed_ms(b) for a, b in ev]))
        out[f"points_{n}"] = {"ms_per_hologram": round(ms, 3),
                              "window_terms_per_s": float(f"{terms / launches / (ms / 1e3):.4g}")}
        dv.free()
        fld.free()
    return out


def fresnel_variants(ctx, capi, synth, Event, launches=2):
    """The 1M-point complex-amplitude Fresnel workload (3840x2160, 532 nm) on
    the FP32 direct kernel and every matrix-core variant: pairs/s (HIP events
    on the launch stream) and field rel-L2 against the FP64 probe."""
    w = synth.WORKLOADS[3]
    n, npx = w.n_points, w.n_pixels
    pts = synth.points(w, n)
    dv = ctx.to_device(pts)
    fld = ctx.alloc(16 * npx)
    p = capi.params(w.pn, w.pp, w.lambdas[0], use_phase=True)
    pix = np.random.default_rng(99).choice(npx, 256, replace=False)
    ref = ctx.probe_f64(dv, n, p, 1, pix)
    out = {"workload": w.name, "points": n, "resolution": list(w.pn), "probe_pixels": len(pix)}
    for name in ("direct", "bf16x1", "bf16x3", "fp16x3"):
        mode = KERNELS[name]
        ctx.field_zero(fld, npx)
        ctx.diffract(dv, n, fld, p, 1, mode)          # warm-up, also the accuracy sample
        ctx.sync()
        f = fld.download(np.complex128, npx)[pix]
        rel = float(np.linalg.norm(f - ref) / np.linalg.norm(ref))
        ts = []
        for _ in range(launches):
            a, b = Event(), Event()
            a.record(ctx.stream)
            ctx.diffract(dv, n, fld, p, 1, mode)
            b.record(ctx.stream)
            ts.append((a, b))
        ctx.sync()
        ms = float(np.mean([a.elapsed_ms(b) for a, b in ts]))
        gps = n * npx / (ms / 1e3) / 1e9
        entry = {"ms": round(ms, 2), "Gpairs_per_s": round(gps, 1), "rel_l2_vs_fp64": float(f"{rel:.3e}")}
        if name in MFMA_TERMS:
            tf = n * npx * 8 * MFMA_TERMS[name] / (ms / 1e3) / 1e12
            entry["mfma_tflops"] = round(tf, 1)
            entry["mfma_frac"] = round(tf / PEAK_BF16_TFLOPS, 4)
        out[name] = entry
        print(f"# variant {name}: {ms:.1f} ms, {gps:.0f} Gpairs/s, rel-L2 {rel:.2e}", file=sys.stderr, flush=True)
    d = out["direct"]["Gpairs_per_s"]
    for name in MFMA_TERMS:
        out[name]["speedup_vs_direct"] = round(out[name]["Gpairs_per_s"] / d, 2)
    dv.free()
    fld.free()
    return out


def rs_variants(ctx, capi, synth, Event, wid, kernels, launches=2, direct_ms=None):
    """RS on the matrix cores (rank-2 tile expansion of the RS term; points
    outside its bounds run the FP32 direct kernel in the same launch) beside
    the FP32 direct kernel: every channel of workload wid, pairs/s from HIP
    events on the launch stream, field rel-L2 against the FP64 probe, and the
    share of points that ran on the matrix cores.  direct_ms: the direct
    kernel's ms per channel from another block, when "direct" is not timed
    here."""
    w = synth.WORKLOADS[wid]
    n, npx, nch = w.n_points, w.n_pixels, len(w.lambdas)
    dv = ctx.to_device(synth.points(w, n))
    fld = ctx.alloc(16 * npx)
    params = [capi.params(w.pn, w.pp, lam, channel=ch) for ch, lam in enumerate(w.lambdas)]
    pix = np.random.default_rng(98).choice(npx, 256, replace=False)
    refs = [ctx.probe_f64(dv, n, params[ch], 0, pix) for ch in range(nch)]
    out = {"workload": w.name, "points": n, "resolution": list(w.pn), "channels": nch, "probe_pixels": len(pix)}
    for name in kernels:
        num = den = 0.0
        on_mfma = 0
        ts = []
        for ch in range(nch):
            ctx.field_zero(fld, npx)
            ctx.diffract(dv, n, fld, params[ch], 0, KERNELS[name])    # warm-up, also the accuracy sample
            ctx.sync()
            on_mfma += ctx.last_stats().gemm_points
            f = fld.download(np.complex128, npx)[pix]
            num += float(np.sum(np.abs(f - refs[ch]) ** 2))
            den += float(np.sum(np.abs(refs[ch]) ** 2))
            for _ in range(launches):
                a, b = Event(), Event()
                a.record(ctx.stream)
                ctx.diffract(dv, n, fld, params[ch], 0, KERNELS[name])
                b.record(ctx.stream)
                ts.append((a, b))
            ctx.sync()
        ms = float(np.mean([a.elapsed_ms(b) for a, b in ts]))
        entry = {"ms_per_channel": round(ms, 2), "Gpairs_per_s": round(n * npx / (ms / 1e3) / 1e9, 1),
                 "rel_l2_vs_fp64": float(f"{np.sqrt(num / den):.3e}")}
        if name in MFMA_TERMS:
            # two K entries per matrix-core point
            tf = on_mfma / nch * npx * 16 * MFMA_TERMS[name] / (ms / 1e3) / 1e12
            entry["matrix_core_point_fraction"] = round(on_mfma / (n * nch), 4)
            entry["mfma_tflops"] = round(tf, 1)
            entry["mfma_frac"] = round(tf / PEAK_BF16_TFLOPS, 4)
        out[name] = entry
        print(f"# rs {w.name} {name}: {ms:.1f} ms/ch, {entry['Gpairs_per_s']:.0f} Gpairs/s, "
              f"rel-L2 {entry['rel_l2_vs_fp64']:.2e}", file=sys.stderr, flush=True)
    d_ms = out["direct"]["ms_per_channel"] if "direct" in out else direct_ms
    if d_ms:
        if "direct" not in out:
            out["direct_ms_per_channel"] = d_ms
        for name in kernels:
            if name in MFMA_TERMS:
                out[name]["speedup_vs_direct"] = round(d_ms / out[name]["ms_per_channel"], 2)
    dv.free()
    fld.free()
    return out


def serving_burst(capi, synth, n_requests: int, rs_kernel: str = "bf16x3"):
    """Only the serving stream's burst pass (every request submitted at once,
    after one warm-up request of each kind), for a kernel trace of the
    server's stream: the GPU's busy fraction while work is queued."""
    reqs = synth.serving_requests(n_requests)
    clouds = [synth.random_cloud(r["seed"], r["n"], 4e-3, (0.49, 0.51)) for r in reqs]
    modes = {1: KERNELS["bf16x3"], 0: KERNELS[rs_kernel]}

    def submit(srv, i):
        r = reqs[i]
        return srv.submit(clouds[i], capi.params(r["pn"], r["pp"], r["lambda"]), r["diff"], modes[r["diff"]],
                          want_field=False, want_u8=True)

    pairs = sum(r["n"] * r["pn"][0] * r["pn"][1] for r in reqs)
    with capi.Server() as srv:
        for i in range(min(4, len(reqs))):
            srv.wait(submit(srv, i))
        t0 = time.perf_counter()
        ids = [submit(srv, i) for i in range(len(reqs))]
        for j in ids:
            st, *_ = srv.wait(j)
            assert st == capi.OK
        burst_s = time.perf_counter() - t0
        c = srv.counters()
    return {"serving_burst": {"requests": len(reqs), "seconds": round(burst_s, 4),
                              "Gpairs_per_s": round(pairs / burst_s / 1e9, 1),
                              "iterations": int(c.iterations), "kernels": {"RS": rs_kernel, "Fresnel": "bf16x3"}}}


def serving_stream(capi, synth, n_requests: int, rs_kernel: str = "bf16x3", server_kw=None):
    """Mixed request stream on the continuous-batching server.  Pass 1 runs
    every request alone (submit, wait) for its alone latency; pass 2 submits
    every request at once to measure capacity (jobs/s); pass 3 replays the
    same requests with Poisson arrivals at 70% of that rate and reports
    per-request latency (completion - arrival), latency per diffraction mode,
    slowdown against the alone latency, and throughput."""
    import threading
    reqs = synth.serving_requests(n_requests)
    clouds = [synth.random_cloud(r["seed"], r["n"], 4e-3, (0.49, 0.51)) for r in reqs]
    fresnel_mode = KERNELS["bf16x3"]
    rs_mode = KERNELS[rs_kernel]

    def submit(srv, i):
        r = reqs[i]
        p = capi.params(r["pn"], r["pp"], r["lambda"])
        return srv.submit(clouds[i], p, r["diff"], fresnel_mode if r["diff"] == 1 else rs_mode,
                          want_field=False, want_u8=True)

    pairs = sum(r["n"] * r["pn"][0] * r["pn"][1] for r in reqs)
    with capi.Server(**(server_kw or {})) as srv:
        # warm-up: one request of each kind
        for i in range(min(4, len(reqs))):
            srv.wait(submit(srv, i))
        alone = np.zeros(len(reqs))
        for i in range(len(reqs)):
            st, stt, _, _ = srv.wait(submit(srv, i))
            assert st == capi.OK
            alone[i] = stt.end_ms - stt.submit_ms
        t0 = time.perf_counter()
        ids = [submit(srv, i) for i in range(len(reqs))]
        for j in ids:
            st, *_ = srv.wait(j)
            assert st == capi.OK
        burst_s = time.perf_counter() - t0
        capacity = len(reqs) / burst_s
        rate = 0.7 * capacity
        gaps = -np.log(1.0 - synth.uniform01(20261202 + 44, len(reqs))) / rate
        ids = [None] * len(reqs)

        def producer():
            t = time.perf_counter()
            for i in range(len(reqs)):
                t += gaps[i]
                d = t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[i] = submit(srv, i)

        th = threading.Thread(target=producer)
        th.start()
        lat, first_submit, last_end, phases = [], None, None, []
        for i in range(len(reqs)):
            while ids[i] is None:
                time.sleep(0.0005)
            st, stt, _, _ = srv.wait(ids[i])
            assert st == capi.OK
            lat.append(stt.end_ms - stt.submit_ms)
            phases.append((stt.admit_ms - stt.submit_ms, stt.start_ms - stt.admit_ms, stt.end_ms - stt.start_ms,
                           stt.iterations))
            first_submit = stt.submit_ms if first_submit is None else min(first_submit, stt.submit_ms)
            last_end = stt.end_ms if last_end is None else max(last_end, stt.end_ms)
        th.join()
        c = srv.counters()
    span_s = (last_end - first_submit) / 1e3
    lat = np.array(lat)
    slow = lat / alone
    diff = np.array([r["diff"] for r in reqs])
    per_mode = {}
    for d, name in ((1, "Fresnel"), (0, "RS")):
        m = diff == d
        if m.any():
            per_mode[name] = {"requests": int(m.sum()), "p50_ms": _pctl(lat[m], 50), "p99_ms": _pctl(lat[m], 99),
                              "alone_p50_ms": _pctl(alone[m], 50), "alone_p99_ms": _pctl(alone[m], 99),
                              "slowdown_p50": _pctl(slow[m], 50), "slowdown_p99": _pctl(slow[m], 99)}
    worst = [{"mode": "Fresnel" if reqs[i]["diff"] == 1 else "RS", "points": reqs[i]["n"],
              "resolution": list(reqs[i]["pn"]), "alone_ms": round(float(alone[i]), 2),
              "latency_ms": round(float(lat[i]), 2), "queued_ms": round(phases[i][0], 2),
              "admitted_to_first_slice_ms": round(phases[i][1], 2), "slices_ms": round(phases[i][2], 2),
              "iterations": int(phases[i][3])}
             for i in np.argsort(-slow)[:3]]
    return {
        "requests": len(reqs), "pairs": float(pairs), "worst_slowdown_jobs": worst,
        "burst_capacity_jobs_per_s": round(capacity, 2), "poisson_rate_jobs_per_s": round(rate, 2),
        "throughput_Gpairs_per_s": round(pairs / span_s / 1e9, 1), "burst_Gpairs_per_s": round(pairs / burst_s / 1e9, 1),
        "p50_ms": _pctl(lat, 50), "p99_ms": _pctl(lat, 99), "mean_ms": round(float(lat.mean()), 2),
        "alone_p50_ms": _pctl(alone, 50), "alone_p99_ms": _pctl(alone, 99),
        "slowdown_p50": _pctl(slow, 50), "slowdown_p99": _pctl(slow, 99),
        "per_mode": per_mode,
        "iterations": int(c.iterations), "graph_replays": int(c.graph_replays),
        "peak_reserved_GB": round(c.peak_reserved_bytes / 1e9, 2),
        "kernels": {"RS": rs_kernel, "Fresnel": "bf16x3"},
    }


MODE_NAMES = {0: "RS", 1: "Fresnel"}


def _pctl(xs, q):
    return round(float(np.percentile(xs, q)), 3) if len(xs) else None


def run_workload(ctx, capi, synth, Event, w, kernel_of, steps, warmup, probe_pixels, n=None, label="",
                 accuracy_channels=None):
    """Time `steps` steps of workload w (every diffraction mode of w, every
    channel: zero, diffract, then phase-encode and normalise the channels),
    resident in HBM, with HIP events on the context's stream.  kernel_of maps
    a diffraction flag to a KERNELS name.  Returns the measurements plus the
    device buffers' host-side handles for further legs (caller frees).
    accuracy_channels: channels whose field is checked against the probe
    after the timed region (default all)."""
    n = n or w.n_points
    npx, nch = w.n_pixels, len(w.lambdas)
    stream = ctx.stream
    pts = synth.points(w, n)
    dv = ctx.to_device(pts)
    fields = [ctx.alloc(16 * npx) for _ in range(nch)]
    encs = [ctx.alloc(8 * npx) for _ in range(nch)]
    u8s = [ctx.alloc(npx) for _ in range(nch)]
    params = [capi.params(w.pn, w.pp, lam, channel=ch, use_phase=w.complex_amplitude)
              for ch, lam in enumerate(w.lambdas)]
    kev, holo_ev, enc_ev = {}, {}, []

    def hologram(diff, timed):
        e0 = Event()
        e0.record(stream)
        for ch in range(nch):
            ctx.field_zero(fields[ch], npx)
            a, b = Event(), Event()
            a.record(stream)
            ctx.diffract(dv, n, fields[ch], params[ch], diff, KERNELS[kernel_of[diff]])
            b.record(stream)
            if timed:
                kev.setdefault(diff, []).append((a, b))
        em = Event()
        em.record(stream)
        ctx.encode_normalize_u8(fields, npx, capi.ENCODE_PHASE, encs, u8s)
        e1 = Event()
        e1.record(stream)
        if timed:
            holo_ev.setdefault(diff, []).append((e0, e1))
            enc_ev.append((em, e1))

    for i in range(warmup):
        for diff in w.modes:
            hologram(diff, False)
        ctx.sync()
        print(f"# {label}warmup step {i + 1}/{warmup} done", file=sys.stderr, flush=True)
    ctx.sync()
    t_start, t_stop = Event(), Event()
    wall0 = time.perf_counter()
    t_start.record(stream)
    for i in range(steps):
        for diff in w.modes:
            hologram(diff, True)
    t_stop.record(stream)
    ctx.sync()
    wall = time.perf_counter() - wall0
    ms_step = t_start.elapsed_ms(t_stop) / steps
    pairs_step = n * npx * nch * len(w.modes)
    print(f"# {label}timed {steps} steps: {ms_step:.1f} ms/step (wall {wall:.2f} s)", file=sys.stderr, flush=True)

    launch_ms = {d: [a.elapsed_ms(b) for a, b in ev] for d, ev in kev.items()}
    holo_ms = {d: [a.elapsed_ms(b) for a, b in ev] for d, ev in holo_ev.items()}
    enc_ms = float(np.mean([a.elapsed_ms(b) for a, b in enc_ev])) if enc_ev else float("nan")

    # accuracy: the field vs the library's FP64 probe on sampled pixels, and the
    # share of nominal pairs inside the points' anti-aliasing boxes
    rng = np.random.default_rng(1234)
    pix = rng.choice(npx, min(probe_pixels, npx), replace=False)
    rel, in_box = {}, {}
    for diff in w.modes:
        num = den = 0.0
        boxed = nominal = 0
        for ch in (range(nch) if accuracy_channels is None else accuracy_channels):
            ctx.field_zero(fields[ch], npx)
            ctx.diffract(dv, n, fields[ch], params[ch], diff, KERNELS[kernel_of[diff]])
            st = ctx.last_stats()
            boxed += st.pairs_in_box
            nominal += st.pairs_nominal
            f = fields[ch].download(np.complex128, npx)[pix]
            ref = ctx.probe_f64(dv, n, params[ch], diff, pix)
            num += float(np.sum(np.abs(f - ref) ** 2))
            den += float(np.sum(np.abs(ref) ** 2))
        rel[MODE_NAMES[diff]] = float(f"{np.sqrt(num / den):.3e}")
        in_box[MODE_NAMES[diff]] = round(boxed / nominal, 6) if nominal else None
    per_mode = {}
    for diff in w.modes:
        ms_l = float(np.mean(launch_ms[diff]))
        per_mode[MODE_NAMES[diff]] = {
            "kernel": kernel_of[diff], "ms_per_channel": round(ms_l, 3),
            "Gpairs_per_s": round(n * npx / (ms_l / 1e3) / 1e9, 1),
            "p50_ms_per_hologram": _pctl(holo_ms[diff], 50), "p99_ms_per_hologram": _pctl(holo_ms[diff], 99),
            "holograms_timed": len(holo_ms[diff]),
            "rel_l2_vs_fp64": rel[MODE_NAMES[diff]], "pairs_in_box_fraction": in_box[MODE_NAMES[diff]]}
    res = {"workload": w.name, "points": n, "resolution": list(w.pn), "channels": nch,
           "modes": [MODE_NAMES[m] for m in w.modes], "steps": steps, "warmup": warmup,
           "ms_per_step": round(ms_step, 3), "Gpairs_per_s": round(pairs_step / (ms_step / 1e3) / 1e9, 1),
           "per_mode": per_mode, "encode_normalize_ms_per_hologram": round(enc_ms, 4),
           "probe_pixels": int(len(pix)),
           "accuracy_channels": list(range(nch)) if accuracy_channels is None else list(accuracy_channels)}
    bufs = {"dv": dv, "fields": fields, "encs": encs, "u8s": u8s, "params": params, "pts": pts, "n": n,
            "launch_ms": launch_ms, "enc_ms": enc_ms, "pairs_step": pairs_step}
    return res, bufs


def free_bufs(bufs):
    for b in [bufs["dv"]] + bufs["fields"] + bufs["encs"] + bufs["u8s"]:
        b.free()


def direct_roofline(w, n, diff, launch_ms):
    """Roofline of an FP32 direct kernel: FP32 FLOPs of the pair loop against the
    published FP32 vector peak, and against the loop's own speed of light."""
    k_ms = float(np.mean(launch_ms))
    pps = n * w.n_pixels / (k_ms / 1e3)
    fl = FLOP_PER_PAIR[diff]
    achieved = pps * fl / 1e12
    return {"bound": "valu", "achieved": round(achieved, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "kernel": f"k_pc_direct<{MODE_NAMES[diff]},FP32>", "avg_launch_ms": round(k_ms, 3),
            "model": f"{fl} FP32 FLOP/pair in the separable pair loop (FMA=2); the per-lane column factor "
                     f"(v_sin/v_cos once per point, shared by 8 rows) is not counted",
            "mix_sol_pairs_per_s": MIX_SOL_PAIRS_PER_S[diff],
            "frac_of_mix_sol": round(pps / MIX_SOL_PAIRS_PER_S[diff], 4)}


def host_handover(ctx, capi, w, bufs, kernel_of):
    """PCIe-inclusive rate of the drop-in (ophPointCloud::generateHologram):
    vertices from the host, every channel's FP64 field handed back into
    page-locked complex_H planes, each channel's copy ordered after its own
    launch so it overlaps the next channel; one step on the wall clock
    (never the value)."""
    n, npx, nch = bufs["n"], w.n_pixels, len(w.lambdas)
    fields, params = bufs["fields"], bufs["params"]
    host_bufs = [capi.HostBuffer(16 * npx) for _ in range(nch)]
    host_fields = [b.array(np.complex128, npx) for b in host_bufs]
    ctx.sync()
    t0 = time.perf_counter()
    dv_h = ctx.to_device(bufs["pts"])
    for diff in w.modes:
        for ch in range(nch):
            ctx.field_zero(fields[ch], npx)
            ctx.diffract(dv_h, n, fields[ch], params[ch], diff, KERNELS[kernel_of[diff]])
            fields[ch].download_async_after(host_fields[ch])
        ctx.copy_sync()
    ctx.sync()
    host_ms = (time.perf_counter() - t0) * 1e3
    dv_h.free()
    t0 = time.perf_counter()
    assert capi.lib().ohx_memcpy_d2h(ctx.h, host_fields[0].ctypes.data, fields[0].ptr, 16 * npx) == capi.OK
    d2h_s = time.perf_counter() - t0
    del host_fields
    for b in host_bufs:
        b.free()
    out = {"ms_per_step": round(host_ms, 2), "Gpairs_per_s": round(bufs["pairs_step"] / (host_ms / 1e3) / 1e9, 1),
           "field_d2h_GB_per_s": round(16 * npx / d2h_s / 1e9, 2),
           "what": "vertex upload + per channel and mode: zero, diffract, FP64 field to a page-locked host "
                   "plane (copy ordered after its own channel, overlapping the next); wall clock"}
    print(f"# host hand-over: {host_ms:.1f} ms/step, field D2H {out['field_d2h_GB_per_s']} GB/s",
          file=sys.stderr, flush=True)
    return out


def encode_block(ctx, capi, w, bufs, Event):
    """Encode + normalise (HBM-bound) and SSB encoding of the resident fields."""
    npx, nch = w.n_pixels, len(w.lambdas)
    # per pixel and channel: the encode pass reads the FP64 field (16 B), writes
    # the encoded plane (8 B) and takes the min/max on the way; the 8-bit pass
    # reads the encoded plane (8 B) and writes 1 B: 33 B algorithmic.  Timed
    # on its own (ten calls after one untimed), as the two-call form below;
    # in_step_ms is the same call timed inside the steps, right after each
    # diffraction launch.
    ev = []
    for rep in range(11):
        a, b = Event(), Event()
        a.record(ctx.stream)
        ctx.encode_normalize_u8(bufs["fields"], npx, capi.ENCODE_PHASE, bufs["encs"], bufs["u8s"])
        b.record(ctx.stream)
        if rep:
            ev.append((a, b))
    ctx.sync()
    enc_ms = float(np.mean([a.elapsed_ms(b) for a, b in ev]))
    enc_bytes = 33 * npx * nch
    out = {"ms_per_hologram": round(enc_ms, 4), "bytes": enc_bytes,
           "GB_per_s": round(enc_bytes / (enc_ms / 1e3) / 1e9, 1), "peak_GB_per_s": 8000.0,
           "frac": round(enc_bytes / (enc_ms / 1e3) / 1e9 / 8000.0, 4),
           "in_step_ms": round(bufs["enc_ms"], 4),
           "what": "ohx_encode_normalize_u8 (k_encode_mm<PHASE> over all channels + k_mm_final + k_to_u8), "
                   "HIP events"}
    # the same work as two calls (ohx_encode per channel, then ohx_normalize_u8:
    # a separate min/max pass, 41 B)
    two = []
    for rep in range(4):
        a, b = Event(), Event()
        a.record(ctx.stream)
        for ch in range(nch):
            ctx.encode(bufs["fields"][ch], npx, bufs["encs"][ch], 0)
        ctx.normalize_u8(bufs["encs"], npx, bufs["u8s"])
        b.record(ctx.stream)
        if rep:
            two.append((a, b))
    ctx.sync()
    two_ms = float(np.mean([a.elapsed_ms(b) for a, b in two]))
    out["two_calls"] = {"ms_per_hologram": round(two_ms, 4),
                        "GB_per_s_of_41_B": round(41 * npx * nch / (two_ms / 1e3) / 1e9, 1)}
    ssb_ev = []
    for rep in range(4):
        a, b = Event(), Event()
        a.record(ctx.stream)
        for ch in range(nch):
            ctx.encode_ssb(bufs["fields"][ch], w.pn, w.pp, bufs["encs"][ch])
        b.record(ctx.stream)
        if rep:
            ssb_ev.append((a, b))
    ctx.sync()
    ssb_ms = float(np.mean([a.elapsed_ms(b) for a, b in ssb_ev]))
    ssb_bytes = 24 * npx * nch
    out["ssb"] = {"ms_per_hologram": round(ssb_ms, 4), "bytes": ssb_bytes,
                  "GB_per_s": round(ssb_bytes / (ssb_ms / 1e3) / 1e9, 1),
                  "what": "k_encode_ssb per channel (encodeHologram, default spectrum shift), HIP events"}
    return out


def workload_block(ctx, capi, synth, Event, wid, kernel_of, steps, warmup, probe_pixels, extra=None,
                   accuracy_channels=None):
    """One of the other workloads, timed as the headline is."""
    w = synth.WORKLOADS[wid]
    res, bufs = run_workload(ctx, capi, synth, Event, w, kernel_of, steps, warmup, probe_pixels,
                             label=f"[{w.name}] ", accuracy_channels=accuracy_channels)
    if extra:
        extra(res, bufs, w)
    free_bufs(bufs)
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not implemented", file=sys.stderr)
        sys.exit(2)

    from openholo_amd import capi, synth
    from openholo_amd.hip_events import Event
    if args.only_serving_burst:
        print(json.dumps(serving_burst(capi, synth, args.requests, args.serving_rs_kernel)), flush=True)
        return

    w = synth.WORKLOADS[args.workload]
    kernel_of = {0: args.rs_kernel or (args.kernel if args.kernel in ("direct", "fp64") else "direct"),
                 1: args.fresnel_kernel or args.kernel}
    ctx = capi.Context(0)
    if args.copy_calibration:
        npx_c = 3840 * 2160
        a, b = ctx.alloc(16 * npx_c), ctx.alloc(16 * npx_c)
        for _ in range(3):
            ctx.accumulate(a, b, npx_c)      # k_field_acc: 32 B read + 16 B written per pixel
        ctx.sync()
        a.free()
        b.free()
    res, bufs = run_workload(ctx, capi, synth, Event, w, kernel_of, args.steps, args.warmup,
                             args.probe_pixels, n=args.points or None)
    n = bufs["n"]
    dom = 0 if 0 in w.modes else 1          # dominant kernel: RS when present
    launch_ms = bufs["launch_ms"][dom]
    kname = kernel_of[dom]
    if kname in MFMA_TERMS:
        k_ms = float(np.mean(launch_ms))
        achieved = n * w.n_pixels / (k_ms / 1e3) * 8 * MFMA_TERMS[kname] / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_BF16_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / PEAK_BF16_TFLOPS, 4),
                    "traffic": TRAFFIC.get(kname) if (args.workload == 3 and n == w.n_points) else None,
                    "traffic_launch": TRAFFIC_LAUNCH.get(kname) if (args.workload == 3 and n == w.n_points) else None,
                    "traffic_source": "rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors from k_field_acc "
                                      "of known size), profiles/traffic_1m_2160p.md",
                    "kernel": f"k_fresnel_gemm<{kname}>", "avg_launch_ms": round(k_ms, 3),
                    "model": f"8 FLOP per complex MAC x {MFMA_TERMS[kname]} term(s); the events bracket "
                             f"k_pc_prep + k_fresnel_gemm + k_gemm_reduce of each launch"}
    elif kname == "direct":
        roofline = direct_roofline(w, n, dom, launch_ms)
        roofline["traffic"] = None
    else:
        k_ms = float(np.mean(launch_ms))
        fl = FLOP_PER_PAIR_FP64[dom]
        achieved = n * w.n_pixels / (k_ms / 1e3) * fl / 1e12
        roofline = {"bound": "valu", "achieved": round(achieved, 2), "peak": PEAK_FP64_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
                    "kernel": f"k_pc_direct64<{MODE_NAMES[dom]}>", "avg_launch_ms": round(k_ms, 3),
                    "model": f"{fl} FP64 FLOP/pair in the pair loop (FMA=2); the per-lane column factor "
                             f"(one FP64 sincospi per point, shared by 8 rows) is not counted"}
    encode_hbm = encode_block(ctx, capi, w, bufs, Event)
    handover = host_handover(ctx, capi, w, bufs, kernel_of)
    free_bufs(bufs)

    var = fresnel_variants(ctx, capi, synth, Event) if args.variants else None
    blocks = {}
    wanted = [b for b in args.blocks.split(",") if b]
    if "rgb2160" in wanted:
        def rs_roofline(r, b, wl):
            r["roofline_rs"] = direct_roofline(wl, b["n"], 0, b["launch_ms"][0])
        blocks["rgb_2160p"] = workload_block(ctx, capi, synth, Event, 2, {0: "direct", 1: "direct"}, 3, 1,
                                             args.probe_pixels, extra=rs_roofline)
    if "rs2160" in wanted:
        blocks["rs_2160p_matrix_core"] = rs_variants(ctx, capi, synth, Event, 2, ("direct", "bf16x1", "bf16x3"))
    if "rs1080" in wanted:
        blocks["rs_1080p"] = workload_block(ctx, capi, synth, Event, 1, {0: "direct"}, 20, 3, args.probe_pixels)
        blocks["rs_1080p_fp64"] = workload_block(ctx, capi, synth, Event, 1, {0: "fp64"}, 5, 1,
                                                 args.probe_pixels)
    if "large4320" in wanted:
        # one step (3 channels x RS + Fresnel, ~40 s); accuracy and box share on
        # channel 0 only, to keep the run within minutes
        blocks["large_4320p"] = workload_block(ctx, capi, synth, Event, 5, {0: "direct", 1: "fp16x3"}, 1, 0, 256,
                                               accuracy_channels=[0])
    if "rs4320" in wanted:
        d_ms = blocks["large_4320p"]["per_mode"]["RS"]["ms_per_channel"] if "large_4320p" in blocks else None
        blocks["rs_4320p_matrix_core"] = rs_variants(ctx, capi, synth, Event, 5, ("bf16x1", "bf16x3"), launches=1,
                                                     direct_ms=d_ms)
    wrp = wrp_block(ctx, capi, synth, Event) if args.wrp else None
    if wrp:
        print(f"# wrp: {wrp}", file=sys.stderr, flush=True)
    ctx.close()
    serving = None
    if args.serving:
        serving = serving_stream(capi, synth, args.requests, args.serving_rs_kernel)
        print(f"# serving: p50 {serving['p50_ms']} ms, p99 {serving['p99_ms']} ms, "
              f"{serving['throughput_Gpairs_per_s']} Gpairs/s", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(w, args.cpu_seconds)
        print(f"# cpu baseline: {cpu['value']:.4f} Gpairs/s on {cpu['cores']} threads", file=sys.stderr, flush=True)

    pm = res["per_mode"]
    line = {
        "metric": "point-pixel pairs/s (G/s) + field rel-L2 vs FP64; p50/p99 ms per hologram",
        "value": res["Gpairs_per_s"], "unit": "Gpairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if kernel_of[dom] == "fp64" else ("bf16" if kernel_of[dom].startswith("bf16") else
                                                          "f16" if kernel_of[dom] == "fp16x3" else "f32"),
        "data": "synthetic (SplitMix64 point cloud)",
        "config": {"workload": w.name, "points": n, "resolution": list(w.pn), "pitch_m": list(w.pp),
                   "lambdas_m": list(w.lambdas), "modes": res["modes"],
                   "complex_amplitude": w.complex_amplitude,
                   "kernel": {m: v["kernel"] for m, v in pm.items()}, "parallelism": "single GPU"},
        "field_rel_l2_vs_fp64": {m: v["rel_l2_vs_fp64"] for m, v in pm.items()},
        "pairs_in_box_fraction": {m: v["pairs_in_box_fraction"] for m, v in pm.items()},
        "kernel_ms_per_channel": {m: v["ms_per_channel"] for m, v in pm.items()},
        "latency_ms_per_hologram": {m: {"p50": v["p50_ms_per_hologram"], "p99": v["p99_ms_per_hologram"],
                                        "holograms": v["holograms_timed"]} for m, v in pm.items()},
        "probe_pixels": res["probe_pixels"],
        "roofline": roofline,
        "cpu_baseline": cpu,
        "host_handover": handover,
        "encode_hbm": encode_hbm,
    }
    if len(pm) == 1:
        only = next(iter(pm.values()))
        line["p50_ms_per_hologram"] = only["p50_ms_per_hologram"]
        line["p99_ms_per_hologram"] = only["p99_ms_per_hologram"]
    if var:
        line["variants"] = var
    if blocks:
        line["workloads"] = blocks
    if wrp:
        line["wrp"] = wrp
    if serving:
        line["serving"] = serving
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
