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
KERNELS = ["vpass", "hpass", "errdiff"]


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--precision", choices=["plain", "fast", "both"], default="both")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--size", type=int, default=4096)
    p.add_argument("--out", type=int, default=1000)
    p.add_argument("--workload", choices=["c1", "c3"], default="c1",
                   help="c1: the headline single image; c3: the 4096-image batch")
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
            "out_pixels": dj.out_pixels, "alg_bytes": dj.alg_bytes, "result": res}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated against a plain copy
    kernel); None when absent."""
    path = os.path.join(ROOT, "profiles", "c1_plain_hbm_counters.txt")
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


def bench_c3(a):
    """4096 RGBA8 images (sides log-uniform 64..1024, factors 0.25..2 per axis,
    triangle / Catmull-Rom / Lanczos2 / Lanczos3, 8-bit, no dither) as one
    batch: one device block, one hipGraph replay per step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from imageworsener_amd import api
    import cases
    lib = api.Library()
    log("generating the C3 batch (4096 images)")
    t0 = time.perf_counter()
    jobs = cases.c3_jobs(4096)
    log("generated in %.1f s; staging and uploading" % (time.perf_counter() - t0))
    b = lib.batch(jobs)
    try:
        out_px, alg_bytes, n = b.sizes()
        for _ in range(a.warmup):
            b.run()
        b.sync()
        lat, kern = [], []
        t0 = time.perf_counter()
        for _ in range(a.steps):
            s0 = time.perf_counter()
            b.run()
            b.sync()
            lat.append((time.perf_counter() - s0) * 1e3)
            kern.append(b.kernel_ms())
        wall = time.perf_counter() - t0
        res = b.finish()
    finally:
        b.close()
    ms = wall * 1e3 / a.steps
    kms = sum(kern) / len(kern)
    base = None
    if not a.no_cpu_baseline:
        ref_path = os.path.join(ROOT, "oracle", "_ref", "libiwref.so")
        kind = "reference" if os.path.exists(ref_path) else None
        if kind:
            ref = api.Library(ref_path, extensions=False)
            done_px, t_cpu, same, k = 0.0, 0.0, True, 0
            while k < len(jobs) and t_cpu < 10.0:
                c0 = time.perf_counter()
                r = ref.run(jobs[k])
                t_cpu += time.perf_counter() - c0
                done_px += r["width"] * r["height"]
                same &= bool(r["ok"] and res[k]["ok"] and (r["pixels"] == res[k]["pixels"]).all())
                k += 1
            base = {"value": round(done_px / 1e6 / t_cpu, 4), "unit": "Mpix/s", "cores": 1, "kind": kind,
                    "sample": "first %d of the 4096 images through iw_process_image on one host thread (%.1f s)"
                              % (k, t_cpu), "plain_bytes_identical": same}
    line = {"metric": METRIC, "value": round(out_px / 1e6 / (ms / 1e3), 3), "unit": "Mpix/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "C3: 4096 RGBA8 images, sides 64..1024, x0.25..2 per axis, "
                                   "triangle/catrom/lanczos2/lanczos3, 8-bit, no dither; one batch",
                       "global_batch": n, "precision": "plain"},
            "roofline": {"kernel": "band (persistent) + graph", "bound": "hbm",
                         "achieved": round(alg_bytes / (kms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(alg_bytes / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "traffic": None, "alg_bytes_per_launch": alg_bytes},
            "cpu_baseline": base,
            "latency_ms": {"p50": round(sorted(lat)[len(lat) // 2], 4), "p99": round(max(lat), 4)},
            "kernel_ms": round(kms, 4), "out_mpix": round(out_px / 1e6, 3)}
    print(json.dumps(line), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (shard by image or "
              "strip across processes; see DESIGN.md)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.workload == "c3":
        return bench_c3(a)
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
    alg_of = {"vpass": in_bytes + inter_bytes, "hpass": inter_bytes + samp_bytes, "errdiff": samp_bytes + out_bytes}
    alg = alg_of[dom]
    achieved = alg / (k[dom] * 1e-3) / 1e9
    traffic = measured_traffic(dom) if (a.size, a.out, chosen) == (4096, 1000, "plain") else None
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "alg_bytes_per_launch": alg,
            "note": "errdiff is a serial chain per plane; step time %.1f ns/pixel" % (
                k["errdiff"] * 1e6 / (a.out * a.out)) if dom == "errdiff" else ""}
    # The data-parallel part of the same step (input image in, codes out).
    rs_ms = k.get("vpass", 0.0) + k.get("hpass", 0.0)
    resample = {"kernels": "vpass+hpass", "ms": round(rs_ms, 4),
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
        "cpu_baseline": base,
        "latency_ms": {"p50": round(m["p50_ms"], 4), "p99": round(m["p99_ms"], 4)},
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
