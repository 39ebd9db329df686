"""Benchmark: output Mpix/s of the ImageWorsener pixel pipeline on one MI355X.

Workload (one step = one pass of the pixel pipeline over one image, inputs
already resident in HBM):
  4096x4096 RGBA 16-bit (synthetic, seed 2) -> 1000x1000, Mitchell cubic
  (B=C=1/3) widened by the 4.096x scale, alpha premultiplied, 8-bit sRGB
  output, Floyd-Steinberg error diffusion on all four channels.

Prints ONE JSON line.  `value` is the chosen precision form's output Mpix/s
(the plain FP64 form when the fast form is not requested or not available);
the plain form's number, per-kernel times from HIP events, p50/p99 step
latency, the roofline of the dominant kernel and the reference CPU build on
this host's cores are reported beside it.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "output Mpix/s on 1 GPU, chosen vs FP64 plain at <=1 LSB err; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
MFMA_F16_PEAK_TFLOPS = 2500.0
# Per-kernel slots of iwamd_job_kernel_ms: the vertical pass (or the whole
# fused resample: band kernel or the FP16-MFMA kernel), the horizontal pass
# (0 when fused), error diffusion.
KERNELS = ["resample", "hpass", "errdiff"]


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--precision", choices=["plain", "fast", "both"], default="both")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--size", type=int, default=4096)
    p.add_argument("--out", type=int, default=1000)
    p.add_argument("--workload", choices=["c1", "c3", "c4"], default="c1",
                   help="c1: the headline single image; c3: the 4096-image batch; c4: the request loop")
    p.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                   help="after timing, run one plain device-to-device copy of MIB MiB (calibrates HBM counters)")
    return p.parse_args()


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


def device_sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def measure(lib, job, steps, warmup):
    dj = lib.device_job(job)
    try:
        for _ in range(warmup):
            dj.run()
        dj.sync()
        device_sync()
        lat, kern = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            s0 = time.perf_counter()
            dj.run()
            dj.sync()
            lat.append((time.perf_counter() - s0) * 1e3)
            kern.append(dj.kernel_ms())
        device_sync()
        wall = time.perf_counter() - t0
        res = dj.finish()
    except Exception:
        dj.close()
        raise
    kern_avg = [sum(k[i] for k in kern) / len(kern) for i in range(len(kern[0]))]
    lat.sort()
    return {"ms_per_step": wall * 1e3 / steps, "kernels_ms": dict(zip(KERNELS, kern_avg)),
            "p50_ms": lat[len(lat) // 2], "p99_ms": lat[min(len(lat) - 1, int(0.99 * len(lat)))],
            "out_pixels": dj.out_pixels, "alg_bytes": dj.alg_bytes, "flops": dj.flops, "result": res}


def mfma_use(flops, ms):
    """Useful resampling FLOP rate of the chosen form's band contraction
    against MI355X's dense FP16 matrix peak.  Useful = a multiply and an add
    per listed tap of both passes; the split hi/lo operands execute three
    FP16 products per useful one, so the issued rate is three times this."""
    if not flops or not ms:
        return None
    t = flops / (ms * 1e-3) / 1e12
    return {"useful_tflops": round(t, 3), "peak": MFMA_F16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(t / MFMA_F16_PEAK_TFLOPS, 6), "split_products": 3,
            "issued_frac": round(3 * t / MFMA_F16_PEAK_TFLOPS, 6), "useful_flops_per_launch": flops}


def lsb_error(plain, fast):
    from imageworsener_amd import api
    return api.code_error(plain, fast)


def measured_traffic(kernel, summary="c1_plain_hbm_counters.txt"):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated against a plain copy
    kernel); None when absent."""
    path = os.path.join(ROOT, "profiles", summary)
    try:
        with open(path) as f:
            for line in f:
                t = line.split()
                if len(t) == 3 and t[0] == "traffic_bytes_per_launch" and t[1] == kernel:
                    return int(t[2])
    except OSError:
        pass
    return None


def cpu_baseline(job, out_pixels):
    """The reference core (oracle/_ref/libiwref.so, built from the reference's
    own sources) on one image, single thread; else the C restatement."""
    from imageworsener_amd import api
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libiwref.so")
    if os.path.exists(ref_path):
        lib, kind, ex = api.Library(ref_path, extensions=False), "reference", None
    else:
        oracle = ctypes.CDLL(os.path.join(ROOT, "oracle", "_build", "libiwo.so"), mode=os.RTLD_LOCAL)
        oracle.iwo_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib, kind = api.Library(), "port"
        ex = lambda plan, img2: bool(oracle.iwo_run(ctypes.addressof(plan), ctypes.addressof(img2)))
    t0 = time.perf_counter()
    r = lib.run(job, executor=ex)
    dt = time.perf_counter() - t0
    return {"value": round(out_pixels / 1e6 / dt, 4), "unit": "Mpix/s", "cores": 1, "kind": kind,
            "sample": "one full workload image through iw_process_image on one host thread (%.1f s)" % dt,
            "ok": r["ok"]}, r


def run_batch_timed(lib, jobs, steps, warmup):
    b = lib.batch(jobs)
    try:
        out_px, alg_bytes, n = b.sizes()
        flops = b.flops()
        for _ in range(warmup):
            b.run()
        b.sync()
        lat, kern = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            s0 = time.perf_counter()
            b.run()
            b.sync()
            lat.append((time.perf_counter() - s0) * 1e3)
            kern.append(b.kernel_ms())
        wall = time.perf_counter() - t0
        res = b.finish()
    finally:
        b.close()
    return {"ms": wall * 1e3 / steps, "kernel_ms": sum(kern) / len(kern), "lat": sorted(lat), "out_px": out_px,
            "alg_bytes": alg_bytes, "n": n, "flops": flops, "res": res}


def bench_c3(a):
    """4096 RGBA8 images (sides log-uniform 64..1024, factors 0.25..2 per axis,
    triangle / Catmull-Rom / Lanczos2 / Lanczos3, 8-bit, no dither) as one
    batch: one device block, one hipGraph replay per step."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from imageworsener_amd import api
    import cases
    lib = api.Library()
    log("generating the C3 batch (4096 images)")
    t0 = time.perf_counter()
    jobs = cases.c3_jobs(4096)
    log("generated in %.1f s" % (time.perf_counter() - t0))
    forms = {}
    want = ["plain", "fast"] if a.precision == "both" else [a.precision]
    for form in want:
        prec = api.PRECISION_PLAIN if form == "plain" else api.PRECISION_FAST
        log("C3 %s: staging, uploading, %d warmup + %d timed replays" % (form, a.warmup, a.steps))
        forms[form] = run_batch_timed(lib, [dict(j, precision=prec) for j in jobs], a.steps, a.warmup)
        log("C3 %s: %.3f ms/step" % (form, forms[form]["ms"]))
    chosen = "fast" if "fast" in forms else "plain"
    m = forms[chosen]
    err = None
    if "fast" in forms and "plain" in forms:
        per = [lsb_error(x, y) for x, y in zip(forms["plain"]["res"], forms["fast"]["res"])]
        per = [e for e in per if e and e.get("comparable", True)]
        w = [r["pixels"].size for r in forms["plain"]["res"]]
        err = {"images": len(per),
               "max_lsb": max(e["max_lsb"] for e in per),
               "mean_lsb": round(sum(e["mean_lsb"] * n for e, n in zip(per, w)) / max(1, sum(w)), 5),
               "max_linear": max(e["max_linear"] for e in per),
               "images_over_1_lsb": sum(1 for e in per if e["max_lsb"] > 1)}
    base = None
    if not a.no_cpu_baseline and "plain" in forms:
        ref_path = os.path.join(ROOT, "oracle", "_ref", "libiwref.so")
        if os.path.exists(ref_path):
            ref = api.Library(ref_path, extensions=False)
            res = forms["plain"]["res"]
            done_px, t_cpu, same, k = 0.0, 0.0, True, 0
            while k < len(jobs) and t_cpu < 10.0:
                c0 = time.perf_counter()
                r = ref.run(jobs[k])
                t_cpu += time.perf_counter() - c0
                done_px += r["width"] * r["height"]
                same &= bool(r["ok"] and res[k]["ok"] and (r["pixels"] == res[k]["pixels"]).all())
                k += 1
            base = {"value": round(done_px / 1e6 / t_cpu, 4), "unit": "Mpix/s", "cores": 1, "kind": "reference",
                    "sample": "first %d of the 4096 images through iw_process_image on one host thread (%.1f s)"
                              % (k, t_cpu), "plain_bytes_identical": same}
    ach = m["alg_bytes"] / (m["kernel_ms"] * 1e-3) / 1e9
    line = {"metric": METRIC, "value": round(m["out_px"] / 1e6 / (m["ms"] / 1e3), 3), "unit": "Mpix/s",
            "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(m["ms"], 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f16-mfma/f32" if chosen == "fast" else "f64", "data": "synthetic",
            "config": {"workload": "C3: 4096 RGBA8 images, sides 64..1024, x0.25..2 per axis, "
                                   "triangle/catrom/lanczos2/lanczos3, 8-bit, no dither; one batch",
                       "global_batch": m["n"], "precision": chosen},
            "roofline": {"kernel": "%s (persistent, hipGraph step)" % (
                             "fast_kernel, both LDS classes" if chosen == "fast" else "band_kernel"),
                         "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4),
                         "traffic": measured_traffic("c3_fast", "c3_fast_hbm_counters.txt") if chosen == "fast" else None,
                         "alg_bytes_per_launch": m["alg_bytes"]},
            "cpu_baseline": base,
            "mfma": mfma_use(forms["fast"]["flops"], forms["fast"]["kernel_ms"]) if "fast" in forms else None,
            "latency_ms": {"p50": round(m["lat"][len(m["lat"]) // 2], 4), "p99": round(m["lat"][-1], 4)},
            "fast_vs_plain": err,
            "forms": {f: {"mpix_s": round(v["out_px"] / 1e6 / (v["ms"] / 1e3), 3), "ms_per_step": round(v["ms"], 4),
                          "kernel_ms": round(v["kernel_ms"], 4)} for f, v in forms.items()},
            "out_mpix": round(m["out_px"] / 1e6, 3)}
    print(json.dumps(line), flush=True)


def bench_c4(a):
    """Serving: requests (RGB 8/16-bit, sides log-uniform 256..2048, one scale
    0.1..1.5, Lanczos3 or Mitchell, 8-bit output with Floyd-Steinberg) through
    the request loop.  Saturation throughput first (all submitted at once),
    then Poisson arrivals at 70% of it for p50/p99 latency."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from imageworsener_amd import api
    import cases
    lib = api.Library()
    n = max(a.steps, 8) * 12
    log("generating %d serving requests" % n)
    reqs = cases.c4_requests(count=n, lo_side=256, hi_side=2048)
    prec = api.PRECISION_FAST if a.precision == "fast" else api.PRECISION_PLAIN
    reqs = [(dict(j, precision=prec), g) for j, g in reqs]
    out_px = sum(j["canvas"][0] * j["canvas"][1] for j, _ in reqs)

    def serve(rate):
        srv = api.Server(lib, hbm_budget=32 << 30, streams=8)
        try:
            t0 = time.perf_counter()
            for job, gap in reqs:
                if rate:
                    due = t0 + gap / rate
                    time.sleep(max(0.0, due - time.perf_counter()))
                    t0 = due
                srv.submit(job)
            res = srv.results()
        finally:
            srv.close()
        lat = sorted(r["times_ms"][2] - r["times_ms"][0] for r in res if "times_ms" in r)
        span = max(r["times_ms"][2] for r in res) - min(r["times_ms"][0] for r in res)
        return res, lat, span

    log("saturation run (all %d requests at once)" % n)
    res, lat_sat, span = serve(0.0)
    sat_rate = n / (span / 1e3)
    log("saturation: %.1f requests/s, %.1f Mpix/s" % (sat_rate, out_px / 1e6 / (span / 1e3)))
    rate = 0.7 * sat_rate
    log("Poisson arrivals at %.1f requests/s" % rate)
    res2, lat, span2 = serve(rate)
    ok = all(r["ok"] for r in res) and all(r["ok"] for r in res2)
    same = all((x["pixels"] == y["pixels"]).all() for x, y in zip(res, res2))
    line = {"metric": METRIC, "value": round(out_px / 1e6 / (span / 1e3), 3), "unit": "Mpix/s", "n_gpus": 1,
            "steps": n, "warmup": 0, "ms_per_step": round(span / n, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == api.PRECISION_PLAIN else "f16-mfma/f32",
            "data": "synthetic",
            "config": {"workload": "C4 (reduced sides 256..2048): %d requests, RGB 8/16-bit, x0.1..1.5, "
                                   "lanczos3/mitchell, 8-bit FS; request loop with continuous batching" % n,
                       "global_batch": n, "precision": "plain" if prec == api.PRECISION_PLAIN else "fast"},
            "latency_ms": {"rate_per_s": round(rate, 2), "p50": round(lat[len(lat) // 2], 3),
                           "p99": round(lat[min(len(lat) - 1, int(0.99 * len(lat)))], 3), "max": round(lat[-1], 3)},
            "saturation": {"requests_per_s": round(sat_rate, 2), "p50": round(lat_sat[len(lat_sat) // 2], 3)},
            "all_ok": ok, "same_bytes_in_both_runs": bool(same), "out_mpix": round(out_px / 1e6, 3)}
    print(json.dumps(line), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (shard by image or "
              "strip across processes; see DESIGN.md)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.workload == "c3":
        return bench_c3(a)
    if a.workload == "c4":
        return bench_c4(a)
    from imageworsener_amd import api, synth
    lib = api.Library()
    log("generating %dx%d RGBA16 input" % (a.size, a.size))
    img = synth.c1(a.size)
    job = api.image_job(synth.be16(img), a.size, a.size, api.IMGTYPE["RGBA"], 16, canvas=(a.out, a.out),
                        resize=api.resize_both("mitchell"), dither=[(api.CH_ALL, 2, 0)])
    forms = {}
    want = ["plain", "fast"] if a.precision == "both" else [a.precision]
    for form in want:
        j = dict(job)
        j["precision"] = api.PRECISION_PLAIN if form == "plain" else api.PRECISION_FAST
        try:
            log("measuring %s: %d warmup + %d timed steps" % (form, a.warmup, a.steps))
            forms[form] = measure(lib, j, a.steps, a.warmup)
            log("%s: %.3f ms/step, kernels %s" % (form, forms[form]["ms_per_step"], forms[form]["kernels_ms"]))
        except RuntimeError as e:
            log("%s unavailable: %s" % (form, e))
    chosen = "fast" if "fast" in forms else "plain"
    m = forms[chosen]
    value = m["out_pixels"] / 1e6 / (m["ms_per_step"] / 1e3)
    err = None
    if "fast" in forms and "plain" in forms:
        log("error of the fast form against the plain form (with the dither, and without it)")
        nodither = dict(job, dither=[(api.CH_ALL, 0, 0)])
        err = {"with_fs": lsb_error(forms["plain"]["result"], forms["fast"]["result"]),
               "pre_dither": lsb_error(lib.run(nodither), lib.run(dict(nodither, precision=api.PRECISION_FAST)))}

    # Roofline of the dominant (longest) kernel, in algorithmic bytes per
    # launch: vpass reads the input image and writes the FP32 intermediate;
    # hpass reads the intermediate and writes the FP64 samples that error
    # diffusion consumes; errdiff reads those samples and writes the codes.
    # errdiff is one dependency chain per plane (latency-bound): its fraction
    # of HBM peak is reported as measured, not as its figure of merit.
    k = m["kernels_ms"]
    dom = max(k, key=k.get)
    H_in, W_in, C = a.size, a.size, 4
    inter_bytes = C * W_in * a.out * 4
    out_bytes = a.out * a.out * C
    in_bytes = H_in * W_in * C * 2
    samp_bytes = C * a.out * a.out * 8
    fused = k.get("hpass", 0.0) == 0.0
    alg_of = {"resample": in_bytes + (samp_bytes if fused else inter_bytes), "hpass": inter_bytes + samp_bytes,
              "errdiff": samp_bytes + out_bytes}
    alg = alg_of[dom]
    achieved = alg / (k[dom] * 1e-3) / 1e9
    traffic = measured_traffic({"resample": "vpass"}.get(dom, dom)) if (
        (a.size, a.out, chosen) == (4096, 1000, "plain") and not fused) or dom == "errdiff" else None
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "alg_bytes_per_launch": alg,
            "note": "errdiff is a serial chain per plane; step time %.1f ns/pixel" % (
                k["errdiff"] * 1e6 / (a.out * a.out)) if dom == "errdiff" else ""}
    # The data-parallel part of the same step (input image in, codes out).
    rs_ms = k.get("resample", 0.0) + k.get("hpass", 0.0)
    resample = {"kernels": "fused" if fused else "vpass+hpass", "ms": round(rs_ms, 4),
                "alg_bytes": in_bytes + out_bytes,
                "achieved_GBs": round((in_bytes + out_bytes) / (rs_ms * 1e-3) / 1e9, 1) if rs_ms > 0 else None}

    base = None
    if not a.no_cpu_baseline:
        log("timing the CPU baseline (one image, one thread)")
        j = dict(job)
        base, ref_res = cpu_baseline(j, m["out_pixels"])
        same = (ref_res["ok"] and forms.get("plain") and
                bool((ref_res["pixels"] == forms["plain"]["result"]["pixels"]).all()))
        base["plain_bytes_identical"] = bool(same)
        base.pop("ok", None)

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpix/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(m["ms_per_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "C1: %dx%d RGBA16 -> %dx%d Mitchell (B=C=1/3), premultiplied alpha, "
                               "8-bit sRGB, Floyd-Steinberg on all channels" % (a.size, a.size, a.out, a.out),
                   "global_batch": 1, "precision": chosen},
        "roofline": roof,
        "resample": resample,
        "mfma": mfma_use(forms["fast"]["flops"], forms["fast"]["kernels_ms"].get("resample"))
                if "fast" in forms else None,
        "cpu_baseline": base,
        "latency_ms": {"p50": round(m["p50_ms"], 4), "p99": round(m["p99_ms"], 4)},
        "fast_vs_plain": err,
        "forms": {f: {"mpix_s": round(v["out_pixels"] / 1e6 / (v["ms_per_step"] / 1e3), 3),
                      "ms_per_step": round(v["ms_per_step"], 4),
                      "kernels_ms": {kk: round(vv, 4) for kk, vv in v["kernels_ms"].items()}}
                  for f, v in forms.items()},
    }
    print(json.dumps(line), flush=True)
    if a.copy_probe:
        n = a.copy_probe << 20
        t = ctypes.c_double()
        lib.dll.iwamd_copy_probe.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_double)]
        if not lib.dll.iwamd_copy_probe(n, ctypes.byref(t)):
            raise RuntimeError("copy probe failed")
        log("copy probe: %d bytes in %.3f ms (%.1f GB/s read+write)" % (n, t.value, 2 * n / t.value / 1e6))


if __name__ == "__main__":
    main()
