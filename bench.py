"""Benchmark: output Mpix/s of the ImageWorsener pixel pipeline on one MI355X.

Default workload (the metric's configuration, C3): 4096 RGBA8 images (seed 4;
sides log-uniform 64..1024; per-axis factors 0.25..2; triangle, Catmull-Rom,
Lanczos2 or Lanczos3; 8-bit sRGB output, no dither) as ONE batch: one device
block, a descriptor table, persistent kernels, one hipGraph replay per step.
Inputs are resident in HBM before the timed region.  Both forms run:

  plain  FP64 resampling, byte-identical to the reference core
  fast   the chosen precision form: FP32 operands and sums on MFMA
         (v_mfma_f32_16x16x4f32) for the resampling, FP64 from the resampled
         value on

`value` is the chosen form's output Mpix/s; it is reported as the headline only
when its error against the plain form is <= 1 LSB on every one of the 4096
images (else the plain form's number is the headline).  The reference core
(oracle/_ref/libiwref.so, compiled from the reference's own sources) is timed
on this host's cores on a bounded sample of the same images, one thread and
all cores (independent processes), and its bytes are compared with both forms.

Other workloads: --workload c1 (one 4096^2 RGBA16 image -> 1000^2, Mitchell,
Floyd-Steinberg on all channels), c2 (2048^2 Gray16 linear -> 3072^2,
Catmull-Rom, 4 levels; ordered, random seed 3, Jarvis-Judice-Ninke), c4 (the
request loop: Poisson arrivals, p50/p99 latency).  The default run also
reports C1's plain and chosen forms under the key "c1".

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import json
import multiprocessing as mp
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "output Mpix/s on 1 GPU, chosen vs FP64 plain at <=1 LSB err; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
MFMA_F32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix peak
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libiwref.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "libiwo.so")
# Arithmetic of each form's resampling (what the line's dtype names).
DTYPE = {"plain": "f64", "fast": "f32"}
# Per-kernel slots of iwamd_job_kernel_ms: the vertical pass (or the whole
# fused resample: band kernel or the MFMA kernel), the horizontal pass
# (0 when fused), error diffusion.
KERNELS = ["resample", "hpass", "errdiff"]


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--precision", choices=["plain", "fast", "both"], default="both")
    p.add_argument("--workload", choices=["c3", "c1", "c2", "c4"], default="c3")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-c1", action="store_true", help="c3: leave out the C1 side measurement")
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="CPU baseline: seconds of work per leg (one thread; all cores)")
    p.add_argument("--images", type=int, default=4096, help="c3: images in the batch")
    p.add_argument("--c4-max-side", type=int, default=8192, help="c4: largest input side (the configuration's 8192)")
    p.add_argument("--c4-requests", type=int, default=1000, help="c4: requests per run (the configuration's 1000)")
    p.add_argument("--c4-streams", type=int, default=8, help="c4: the request loop's streams (and HIP hardware queues)")
    p.add_argument("--size", type=int, default=4096, help="c1: input side")
    p.add_argument("--out", type=int, default=1000, help="c1: output side")
    p.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                   help="after timing, run one plain device-to-device copy of MIB MiB (calibrates HBM counters)")
    return p.parse_args()


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


def host_procs():
    """Worker processes for host-side work: this process's CPU share, at most
    16 (the share of one GPU on a shared host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        libc = subprocess.run(["ldd", "--version"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                              text=True, timeout=10).stdout.splitlines()[0].strip()
    except Exception:
        libc = None
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"cpu_model": model, "host_cpus": os.cpu_count(), "cpus_available": affinity, "libc": libc}


# ---- synthetic C3 images, generated in worker processes ---------------------
# Workers are forked before this process touches the GPU.

def _c3_image(args):
    from imageworsener_amd import synth
    w, h, seed = args
    return synth.c3_image(w, h, seed)


def c3_jobs(count):
    from imageworsener_amd import api, synth
    lst = synth.c3_list(count)
    procs = host_procs()
    with mp.get_context("fork").Pool(procs) as pool:
        imgs = pool.map(_c3_image, [(w, h, s) for (w, h, _, _, _, s) in lst], chunksize=8)
    return [api.image_job(img, w, h, api.IMGTYPE["RGBA"], 8, canvas=(ow, oh), resize=api.resize_both(filt))
            for img, (w, h, ow, oh, filt, _) in zip(imgs, lst)]


# ---- CPU baseline -------------------------------------------------------------

_SHARED = {}


def _ref_leg(args):
    """One worker of the all-cores leg: the reference core over its own
    images (indices) until `seconds` of work are done."""
    from imageworsener_amd import api
    idx, seconds = args
    lib = api.Library(REF_LIB, extensions=False)
    jobs = _SHARED["jobs"]
    px, n, t0 = 0.0, 0, time.perf_counter()
    for i in idx:
        r = lib.run(jobs[i])
        if r["ok"]:
            px += r["width"] * r["height"]
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    return px, time.perf_counter() - t0, n


def cpu_baseline_batch(jobs, seconds):
    """The reference core timed on this host before the GPU is used: (i) one
    thread, the images in order; (ii) every available core as independent
    processes over further, disjoint images.  Returns the baseline object and
    the one-thread leg's results (for the byte comparison)."""
    from imageworsener_amd import api
    if not os.path.exists(REF_LIB):
        log("reference core not built (%s): no CPU baseline" % REF_LIB)
        return None, []
    lib = api.Library(REF_LIB, extensions=False)
    log("CPU baseline, one thread: reference core for %.0f s" % seconds)
    res, px, t = [], 0.0, 0.0
    while len(res) < len(jobs) and t < seconds:
        t0 = time.perf_counter()
        r = lib.run(jobs[len(res)])
        t += time.perf_counter() - t0
        px += r["width"] * r["height"] if r["ok"] else 0
        res.append(r)
    one = {"value": round(px / 1e6 / t, 4), "cores": 1, "images": len(res), "seconds": round(t, 2)}
    procs = host_procs()
    k0 = len(res)
    rest = list(range(k0, len(jobs)))
    log("CPU baseline, %d processes: reference core for %.0f s each" % (procs, seconds))
    _SHARED["jobs"] = jobs
    with mp.get_context("fork").Pool(procs) as pool:
        legs = pool.map(_ref_leg, [(rest[w::procs], seconds) for w in range(procs)], chunksize=1)
    _SHARED.clear()
    apx = sum(l[0] for l in legs)
    wall = max(l[1] for l in legs)
    allc = {"value": round(apx / 1e6 / wall, 4), "cores": procs, "images": sum(l[2] for l in legs),
            "seconds": round(wall, 2)}
    base = {"value": allc["value"], "unit": "Mpix/s", "cores": procs, "kind": "reference",
            "sample": "reference core (oracle/_ref/libiwref.so, gcc -O3, no -march) through iw_process_image: "
                      "%d processes on disjoint images of the same batch, %.0f s each; one thread: the first %d "
                      "images" % (procs, seconds, one["images"]),
            "one_thread": one, "all_cores": allc}
    base.update(host_info())
    return base, res


# ---- GPU ---------------------------------------------------------------------

def run_batch_timed(lib, jobs, steps, warmup):
    t0 = time.perf_counter()
    b = lib.batch(jobs)
    t_up = time.perf_counter() - t0
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
            "alg_bytes": alg_bytes, "n": n, "flops": flops, "res": res, "stage_upload_s": t_up}


def measure_job(lib, job, steps, warmup):
    dj = lib.device_job(job)
    try:
        for _ in range(warmup):
            dj.run()
        dj.sync()
        lat, kern = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            s0 = time.perf_counter()
            dj.run()
            dj.sync()
            lat.append((time.perf_counter() - s0) * 1e3)
            kern.append(dj.kernel_ms())
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
    against MI355X's dense FP32 matrix peak.  Useful = a multiply and an add
    per listed tap of both passes (the dense band blocks also multiply the
    zeros around each list, which this leaves out)."""
    if not flops or not ms:
        return None
    t = flops / (ms * 1e-3) / 1e12
    return {"useful_tflops": round(t, 3), "peak": MFMA_F32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(t / MFMA_F32_PEAK_TFLOPS, 6), "useful_flops_per_launch": flops}


def measured_traffic(kernel, summary):
    """HBM bytes per step of `kernel` on this workload, from a committed
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


def error_over(pairs, weights=None):
    """Max/mean code error over (a, b) result pairs, each pair one image."""
    from imageworsener_amd import api
    per = [api.code_error(a, b) for a, b in pairs]
    keep = [(e, a["pixels"].size) for e, (a, _) in zip(per, pairs) if e and e.get("comparable", True)]
    if not keep:
        return None
    tot = sum(n for _, n in keep)
    return {"images": len(keep), "images_not_comparable": len(pairs) - len(keep),
            "max_lsb": max(e["max_lsb"] for e, _ in keep),
            "mean_lsb": round(sum(e["mean_lsb"] * n for e, n in keep) / max(1, tot), 6),
            "max_linear": max(e["max_linear"] for e, _ in keep),
            "images_over_1_lsb": sum(1 for e, _ in keep if e["max_lsb"] > 1)}


def same_bytes(a, b):
    return bool(a["ok"] and b["ok"] and a["pixels"].size == b["pixels"].size and (a["pixels"] == b["pixels"]).all()
                and a["imgtype"] == b["imgtype"])


def bench_c3(a):
    from imageworsener_amd import api
    log("generating the C3 batch (%d images, %d processes)" % (a.images, host_procs()))
    t0 = time.perf_counter()
    jobs = c3_jobs(a.images)
    log("generated in %.1f s" % (time.perf_counter() - t0))
    base, ref_res = (None, []) if a.no_cpu_baseline else cpu_baseline_batch(jobs, a.cpu_seconds)

    lib = api.Library()
    forms = {}
    want = ["plain", "fast"] if a.precision == "both" else [a.precision]
    for form in want:
        prec = api.PRECISION_PLAIN if form == "plain" else api.PRECISION_FAST
        log("C3 %s: staging and upload, %d warmup + %d timed replays" % (form, a.warmup, a.steps))
        forms[form] = run_batch_timed(lib, [dict(j, precision=prec) for j in jobs], a.steps, a.warmup)
        log("C3 %s: %.3f ms/step (staging and upload %.1f s)" % (form, forms[form]["ms"],
                                                                  forms[form]["stage_upload_s"]))
    err = None
    if "fast" in forms and "plain" in forms:
        log("error of the chosen form against the plain form over all %d images" % len(jobs))
        err = error_over(list(zip(forms["plain"]["res"], forms["fast"]["res"])))
    fast_ok = err is not None and err["max_lsb"] <= 1
    chosen = "fast" if "fast" in forms and (fast_ok or "plain" not in forms) else "plain"
    m = forms[chosen]
    if base is not None:
        k = len(ref_res)
        if "plain" in forms:
            base["plain_bytes_identical"] = all(same_bytes(r, g) for r, g in zip(ref_res, forms["plain"]["res"][:k]))
        if "fast" in forms:
            base["fast_vs_reference"] = error_over(list(zip(ref_res, forms["fast"]["res"][:k])))
    ach = m["alg_bytes"] / (m["kernel_ms"] * 1e-3) / 1e9
    kname = ("fast_kernel, every plan variant" if chosen == "fast" else "band_kernel, all band heights and LDS classes")
    line = {"metric": METRIC, "value": round(m["out_px"] / 1e6 / (m["ms"] / 1e3), 3), "unit": "Mpix/s",
            "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(m["ms"], 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": DTYPE[chosen], "data": "synthetic",
            "config": {"workload": "C3: %d RGBA8 images (seed 4), sides log-uniform 64..1024, x0.25..2 per "
                                   "axis, triangle/catrom/lanczos2/lanczos3, 8-bit sRGB, no dither; one batch, "
                                   "one hipGraph replay per step" % m["n"],
                       "global_batch": m["n"], "precision": chosen,
                       "dtype_note": "fast: FP32 operands and sums on MFMA, FP64 from the resampled value "
                                     "on; plain: FP64 throughout"},
            "roofline": {"kernel": "%s (persistent launches of one hipGraph step)" % kname,
                         "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 5),
                         "traffic": measured_traffic("c3_" + chosen, "c3_%s_hbm_counters.txt" % chosen),
                         "alg_bytes_per_launch": int(m["alg_bytes"]),
                         "kernel_ms": round(m["kernel_ms"], 4)},
            "cpu_baseline": base,
            "mfma": mfma_use(forms["fast"]["flops"], forms["fast"]["kernel_ms"]) if "fast" in forms else None,
            "latency_ms": {"p50": round(m["lat"][len(m["lat"]) // 2], 4), "p99": round(m["lat"][-1], 4)},
            "fast_vs_plain": err,
            "forms": {f: {"dtype": DTYPE[f], "mpix_s": round(v["out_px"] / 1e6 / (v["ms"] / 1e3), 3),
                          "ms_per_step": round(v["ms"], 4), "kernel_ms": round(v["kernel_ms"], 4),
                          "achieved_GBs": round(v["alg_bytes"] / (v["kernel_ms"] * 1e-3) / 1e9, 1)}
                      for f, v in forms.items()},
            "out_mpix": round(m["out_px"] / 1e6, 3)}
    for f in forms.values():
        f["res"] = None
    if not a.no_c1:
        line["c1"] = c1_side(a, lib)
    print(json.dumps(line), flush=True)
    copy_probe(lib, a.copy_probe)


def c1_job(size, out):
    from imageworsener_amd import api, synth
    img = synth.c1(size)
    return api.image_job(synth.be16(img), size, size, api.IMGTYPE["RGBA"], 16, canvas=(out, out),
                         resize=api.resize_both("mitchell"), dither=[(api.CH_ALL, 2, 0)])


def c1_forms(a, lib, job):
    from imageworsener_amd import api
    forms = {}
    want = ["plain", "fast"] if a.precision == "both" else [a.precision]
    for form in want:
        j = dict(job, precision=api.PRECISION_PLAIN if form == "plain" else api.PRECISION_FAST)
        log("C1 %s: %d warmup + %d timed steps" % (form, a.warmup, a.steps))
        forms[form] = measure_job(lib, j, a.steps, a.warmup)
        log("C1 %s: %.3f ms/step, kernels %s" % (form, forms[form]["ms_per_step"], forms[form]["kernels_ms"]))
    return forms


def c1_alg(a, fused):
    """Algorithmic bytes per launch of each C1 kernel: the resample reads the
    input image and writes the FP32 intermediate (two-pass) or the FP64
    samples error diffusion consumes (fused); hpass reads the intermediate
    and writes those samples; errdiff reads them and writes the codes."""
    C = 4
    in_bytes = a.size * a.size * C * 2
    inter_bytes = C * a.size * a.out * 4
    samp_bytes = C * a.out * a.out * 8
    out_bytes = a.out * a.out * C
    return {"resample": in_bytes + (samp_bytes if fused else inter_bytes), "hpass": inter_bytes + samp_bytes,
            "errdiff": samp_bytes + out_bytes}, in_bytes + out_bytes


def c1_side(a, lib):
    """C1's plain and chosen forms, reported beside the headline."""
    from imageworsener_amd import api
    log("C1: generating %dx%d RGBA16 input" % (a.size, a.size))
    job = c1_job(a.size, a.out)
    forms = c1_forms(a, lib, job)
    out = {"workload": "C1: %dx%d RGBA16 -> %dx%d Mitchell (B=C=1/3), premultiplied alpha, 8-bit sRGB, "
                       "Floyd-Steinberg on all channels" % (a.size, a.size, a.out, a.out), "forms": {}}
    for f, v in forms.items():
        k = v["kernels_ms"]
        fused = k.get("hpass", 0.0) == 0.0
        rs = k["resample"] + k.get("hpass", 0.0)
        _, b_alg = c1_alg(a, fused)
        out["forms"][f] = {"dtype": DTYPE[f], "mpix_s": round(v["out_pixels"] / 1e6 / (v["ms_per_step"] / 1e3), 3),
                           "ms_per_step": round(v["ms_per_step"], 4),
                           "kernels_ms": {kk: round(vv, 4) for kk, vv in k.items()},
                           "resample_kernels": "fused (one kernel)" if fused else "vpass+hpass",
                           "resample_achieved_GBs": round(b_alg / (rs * 1e-3) / 1e9, 1),
                           "errdiff_ns_per_pixel": round(k["errdiff"] * 1e6 / (a.out * a.out), 2)}
    if "fast" in forms and "plain" in forms:
        nod = dict(job, dither=[(api.CH_ALL, 0, 0)])
        out["fast_vs_plain"] = {
            "with_fs": api.code_error(forms["plain"]["result"], forms["fast"]["result"]),
            "pre_dither": api.code_error(lib.run(nod), lib.run(dict(nod, precision=api.PRECISION_FAST)))}
    if not a.no_cpu_baseline and os.path.exists(REF_LIB):
        log("C1: reference core, one image, one thread")
        ref = api.Library(REF_LIB, extensions=False)
        t0 = time.perf_counter()
        r = ref.run(job)
        dt = time.perf_counter() - t0
        out["cpu_reference_one_thread"] = {"mpix_s": round(a.out * a.out / 1e6 / dt, 4), "seconds": round(dt, 2)}
        if "plain" in forms:
            out["plain_bytes_identical"] = same_bytes(r, forms["plain"]["result"])
    return out


def copy_probe(lib, mib):
    """One plain device-to-device copy of mib MiB after the timed work: the
    reference point that calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    if not mib:
        return
    n = mib << 20
    t = ctypes.c_double()
    lib.dll.iwamd_copy_probe.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_double)]
    if not lib.dll.iwamd_copy_probe(n, ctypes.byref(t)):
        raise RuntimeError("copy probe failed")
    log("copy probe: %d bytes in %.3f ms (%.1f GB/s read+write)" % (n, t.value, 2 * n / t.value / 1e6))


def bench_c1(a):
    """C1 on its own: one line whose value is the plain FP64 form (byte-
    identical to the reference; the chosen form differs by more than 1 LSB
    under error diffusion)."""
    from imageworsener_amd import api
    lib = api.Library()
    side = c1_side(a, lib)
    f = "plain" if "plain" in side["forms"] else "fast"
    m = side["forms"][f]
    k = m["kernels_ms"]
    dom = max(k, key=k.get)
    alg, _ = c1_alg(a, m["resample_kernels"].startswith("fused"))
    ach = alg[dom] / (k[dom] * 1e-3) / 1e9
    traffic = measured_traffic(dom if dom != "resample" else "vpass", "c1_plain_hbm_counters.txt") \
        if f == "plain" and (a.size, a.out) == (4096, 1000) else None
    line = {"metric": METRIC, "value": m["mpix_s"], "unit": "Mpix/s", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": DTYPE[f], "data": "synthetic",
            "config": {"workload": side["workload"], "global_batch": 1, "precision": f},
            "roofline": {"kernel": dom, "bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 6), "traffic": traffic,
                         "alg_bytes_per_launch": alg[dom],
                         "note": "errdiff is one serial chain per plane (latency-bound)" if dom == "errdiff" else ""},
            "cpu_baseline": {"value": side["cpu_reference_one_thread"]["mpix_s"], "unit": "Mpix/s", "cores": 1,
                             "kind": "reference", "sample": "one full image through iw_process_image, one thread"}
            if "cpu_reference_one_thread" in side else None,
            "c1": side}
    print(json.dumps(line), flush=True)
    copy_probe(lib, a.copy_probe)


def bench_c2(a):
    """C2: one 2048^2 Gray16 linear image -> 3072^2, Catmull-Rom, 4 grey
    levels, once per dither mode (ordered, random seed 3, JJN); plain FP64
    only (the byte-exact configuration).  Value: output Mpix/s over the three
    modes (a step is one pass of each)."""
    from imageworsener_amd import api
    import cases
    lib = api.Library()
    modes = {}
    ref = api.Library(REF_LIB, extensions=False) if (os.path.exists(REF_LIB) and not a.no_cpu_baseline) else None
    for d in ("o", "r", "jjn"):
        job = dict(cases.c2_job(dither=d), precision=api.PRECISION_PLAIN)
        log("C2 %s: %d warmup + %d timed steps" % (d, a.warmup, a.steps))
        m = measure_job(lib, job, a.steps, a.warmup)
        k = m["kernels_ms"]
        mode = {"ms_per_step": round(m["ms_per_step"], 4), "kernels_ms": {kk: round(vv, 4) for kk, vv in k.items()},
                "mpix_s": round(m["out_pixels"] / 1e6 / (m["ms_per_step"] / 1e3), 3)}
        if d == "jjn":
            mode["errdiff_ns_per_pixel"] = round(k["errdiff"] * 1e6 / m["out_pixels"], 2)
        if ref is not None:
            t0 = time.perf_counter()
            r = ref.run(job)
            dt = time.perf_counter() - t0
            mode["cpu_reference_one_thread_s"] = round(dt, 3)
            mode["plain_bytes_identical"] = same_bytes(r, m["result"])
        modes[d] = (m, mode)
    tot_px = sum(m["out_pixels"] for m, _ in modes.values())
    tot_ms = sum(m["ms_per_step"] for m, _ in modes.values())
    # Roofline: the JJN run's longest kernel.
    mj = modes["jjn"][0]
    k = mj["kernels_ms"]
    dom = max(k, key=k.get)
    in_b, inter_b, samp_b, out_b = 2048 * 2048 * 2, 2048 * 3072 * 4, 3072 * 3072 * 8, 3072 * 3072
    alg = {"resample": in_b + inter_b, "hpass": inter_b + samp_b, "errdiff": samp_b + out_b}[dom]
    ach = alg / (k[dom] * 1e-3) / 1e9
    cpu = None
    if ref is not None:
        t = sum(mode["cpu_reference_one_thread_s"] for _, mode in modes.values())
        cpu = {"value": round(tot_px / 1e6 / t, 4), "unit": "Mpix/s", "cores": 1, "kind": "reference",
               "sample": "the three full C2 runs through iw_process_image, one thread (%.1f s)" % t}
    line = {"metric": METRIC, "value": round(tot_px / 1e6 / (tot_ms / 1e3), 3), "unit": "Mpix/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(tot_ms, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "C2: 2048x2048 Gray16 linear -> 3072x3072 Catmull-Rom, 4 grey levels; ordered, "
                                   "random (seed 3) and Jarvis-Judice-Ninke; a step is one run of each",
                       "global_batch": 3, "precision": "plain"},
            "roofline": {"kernel": "%s (jjn run)" % dom, "bound": "hbm", "achieved": round(ach, 2),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 6), "traffic": None,
                         "alg_bytes_per_launch": alg,
                         "note": "errdiff is one serial chain (latency-bound)" if dom == "errdiff" else ""},
            "cpu_baseline": cpu, "modes": {d: mode for d, (_, mode) in modes.items()}}
    print(json.dumps(line), flush=True)


def _c4_image(args):
    from imageworsener_amd import synth
    w, h, depth, seed = args
    img = synth.gratings_rgb(w, h, depth, seed)
    return img if depth == 8 else synth.be16(img)


def c4_requests(n, max_side):
    """The serving requests (cases.c4_requests), images generated in forked
    worker processes before the GPU is used."""
    from imageworsener_amd import api, synth
    lst = synth.c4_list(n, 5, 256, max_side)
    with mp.get_context("fork").Pool(host_procs()) as pool:
        imgs = pool.map(_c4_image, [(w, h, d, s) for (w, h, d, _, _, _, s, _) in lst], chunksize=1)
    return [(api.image_job(pix, w, h, api.IMGTYPE["RGB"], d, canvas=(ow, oh), resize=api.resize_both(f),
                           dither=[(api.CH_ALL, 2, 0)], depth=8), gap)
            for pix, (w, h, d, ow, oh, f, _, gap) in zip(imgs, lst)]


def bench_c4(a):
    """Serving: requests (RGB 8/16-bit, sides log-uniform 256..8192, one scale
    0.1..1.5, Lanczos3 or Mitchell, 8-bit output with Floyd-Steinberg) through
    the request loop.  Saturation throughput first (all submitted at once),
    then Poisson arrivals at 70% of it for p50/p99 latency."""
    from imageworsener_amd import api
    n = max(1, a.c4_requests)
    # One HIP hardware queue per stream of the loop: with fewer (HIP's default
    # is 4), streams share a queue and a flight's long error-diffusion kernels
    # hold back the flights queued behind them.  Set before the HIP runtime
    # starts (the library is loaded below).
    queues = min(32, max(1, a.c4_streams))
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < queues:
        os.environ["GPU_MAX_HW_QUEUES"] = str(queues)
    log("generating %d serving requests (sides 256..%d, %d processes)" % (n, a.c4_max_side, host_procs()))
    t0 = time.perf_counter()
    reqs = c4_requests(n, a.c4_max_side)
    log("generated in %.1f s" % (time.perf_counter() - t0))
    lib = api.Library()
    prec = api.PRECISION_FAST if a.precision == "fast" else api.PRECISION_PLAIN
    form = "fast" if prec == api.PRECISION_FAST else "plain"
    reqs = [(dict(j, precision=prec), g) for j, g in reqs]
    out_px = sum(j["canvas"][0] * j["canvas"][1] for j, _ in reqs)

    def serve(rate):
        srv = api.Server(lib, hbm_budget=32 << 30, streams=a.c4_streams, timeout_s=120.0)
        try:
            t0 = time.perf_counter()
            for job, gap in reqs:
                if rate:
                    due = t0 + gap / rate
                    time.sleep(max(0.0, due - time.perf_counter()))
                    t0 = due
                srv.submit(job)
            res = srv.results()
            st = srv.stats()
        finally:
            srv.close()
        lat = sorted(r["times_ms"][2] - r["times_ms"][0] for r in res if "times_ms" in r)
        span = max(r["times_ms"][2] for r in res) - min(r["times_ms"][0] for r in res)
        return res, lat, span, st

    log("saturation run (all %d requests at once)" % n)
    res, lat_sat, span, st1 = serve(0.0)
    sat_rate = n / (span / 1e3)
    log("saturation: %.1f requests/s, %.1f Mpix/s" % (sat_rate, out_px / 1e6 / (span / 1e3)))
    rate = 0.7 * sat_rate
    log("Poisson arrivals at %.1f requests/s" % rate)
    res2, lat, span2, st2 = serve(rate)
    log("Poisson run done")
    ok = all(r["ok"] for r in res) and all(r["ok"] for r in res2)
    same = all(same_bytes(x, y) for x, y in zip(res, res2))
    line = {"metric": METRIC, "value": round(out_px / 1e6 / (span / 1e3), 3), "unit": "Mpix/s", "n_gpus": 1,
            "steps": n, "warmup": 0, "ms_per_step": round(span / n, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": DTYPE[form], "data": "synthetic",
            "config": {"workload": "C4 (sides 256..%d): %d requests, RGB 8/16-bit, x0.1..1.5, "
                                   "lanczos3/mitchell, 8-bit FS; request loop with continuous batching"
                                   % (a.c4_max_side, n),
                       "global_batch": n, "precision": form, "streams": a.c4_streams,
                       "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4)},
            "latency_ms": {"rate_per_s": round(rate, 2), "p50": round(lat[len(lat) // 2], 3),
                           "p99": round(lat[min(len(lat) - 1, int(0.99 * len(lat)))], 3), "max": round(lat[-1], 3)},
            "saturation": {"requests_per_s": round(sat_rate, 2), "p50": round(lat_sat[len(lat_sat) // 2], 3)},
            "loop": {"saturation": st1, "poisson": st2},
            "all_ok": ok, "same_bytes_in_both_runs": bool(same), "out_mpix": round(out_px / 1e6, 3)}
    print(json.dumps(line), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this library runs on one GPU; --gpus %d is not supported (the path shards by image, "
              "plane or output strip across processes; nothing here runs on more than one GPU)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    {"c3": bench_c3, "c1": bench_c1, "c2": bench_c2, "c4": bench_c4}[a.workload](a)


if __name__ == "__main__":
    main()
