"""Parity jobs shared by the CPU (oracle vs reference) and GPU (library vs
reference) tests: the benchmark configurations at full or reduced size, and
option sweeps the reference's own tests do not reach at these sizes."""
import numpy as np

from imageworsener_amd import api, synth

RGB, RGBA, GRAY, GRAYA = (api.IMGTYPE[k] for k in ("RGB", "RGBA", "GRAY", "GRAYA"))
V_CVT_GRAY, V_EDGE_X, V_EDGE_Y, V_INTCLAMP, V_NOGAMMA = 11, 17, 18, 15, 13
V_TRANSLATE_X, V_TRANSLATE_Y = 42, 43


def c0_job(w=640, h=480, ow=320, oh=240):
    img = synth.c0(w, h)
    return api.image_job(img, w, h, RGB, 8, canvas=(ow, oh), resize=api.resize_both("lanczos3"))


def c1_job(size=4096, out=1000):
    img = synth.c1(size)
    return api.image_job(synth.be16(img), size, size, RGBA, 16, canvas=(out, out),
                         resize=api.resize_both("mitchell"), dither=[(api.CH_ALL, 2, 0)])


def c2_job(size=2048, out=3072, dither="o", seed=3):
    img = synth.c2(size)
    fam, sub = api.DITHER[dither]
    job = api.image_job(synth.be16(img), size, size, GRAY, 16, canvas=(out, out),
                        resize=api.resize_both("catrom"), input_cs=(api.CS["linear"], 0.0),
                        color_count=[(api.CH_ALL, 4)], dither=[(api.CH_ALL, fam, sub)])
    if dither in ("r", "r2"):
        job["random_seed"] = seed
    return job


def sweep_jobs():
    """(name, job) pairs covering filters, dithers, depths, alpha and
    colour-space options on small synthetic images."""
    jobs = []
    rgb = synth.c0(97, 61)
    rgba = synth.c3_image(83, 71, 7)
    g16 = synth.c2(64)
    for filt in ["triangle", "catrom", "mitchell", "lanczos2", "lanczos3", "hermite", "gaussian", "box",
                 "boxavg", "mix", "nearest", "quadratic"]:
        jobs.append(("rgb-down-" + filt, api.image_job(rgb, 97, 61, RGB, 8, canvas=(41, 29),
                                                        resize=api.resize_both(filt))))
        jobs.append(("rgba-up-" + filt, api.image_job(rgba, 83, 71, RGBA, 8, canvas=(131, 107),
                                                       resize=api.resize_both(filt))))
    for name, (fam, sub) in api.DITHER.items():
        for cc in (0, 3, 7):
            job = api.image_job(rgba, 83, 71, RGBA, 8, canvas=(57, 49), resize=api.resize_both("catrom"),
                                dither=[(api.CH_ALL, fam, sub)], random_seed=11)
            if cc:
                job["color_count"] = [(api.CH_ALL, cc)]
            jobs.append(("dither-%s-cc%d" % (name, cc), job))
    for d in ("none", "f", "o", "r"):
        fam, sub = api.DITHER[d]
        jobs.append(("g16-depth16-" + d, api.image_job(synth.be16(g16), 64, 64, GRAY, 16, canvas=(45, 77), depth=16,
                                                       resize=api.resize_both("lanczos3"),
                                                       dither=[(api.CH_ALL, fam, sub)], random_seed=-5)))
    # Error diffusion across rows wider than one LDS chunk, and 16-bit
    # output (tables too large for LDS).
    wide = synth.c2(160)
    for d in ("f", "jjn", "atkinson", "sierra"):
        fam, sub = api.DITHER[d]
        jobs.append(("wide-ed-" + d, api.image_job(synth.be16(wide), 160, 160, GRAY, 16, canvas=(2600, 23),
                                                   resize=api.resize_both("catrom"),
                                                   dither=[(api.CH_ALL, fam, sub)], color_count=[(api.CH_ALL, 5)])))
    jobs.append(("wide-ed-rgba16", api.image_job(rgba, 83, 71, RGBA, 8, canvas=(1300, 17), depth=16,
                                                 resize=api.resize_both("lanczos2"), dither=[(api.CH_ALL, 2, 0)])))
    jobs.append(("rgba-gray", api.image_job(rgba, 83, 71, RGBA, 8, canvas=(50, 50), resize=api.resize_both("catrom"),
                                            values=[(V_CVT_GRAY, 1)])))
    jobs.append(("rgba-bkgd", api.image_job(rgba, 83, 71, RGBA, 8, canvas=(50, 50), resize=api.resize_both("catrom"),
                                            bkgd=(0.2, 0.5, 0.9, 1.0))))
    jobs.append(("rgba-checker", api.image_job(rgba, 83, 71, RGBA, 8, canvas=(50, 50),
                                               resize=api.resize_both("lanczos3"), bkgd=(0.2, 0.5, 0.9, 0.5),
                                               checker=(5, (0.9, 0.1, 0.1, 0.7)), checker_origin=(2, 3))))
    jobs.append(("rgb-edge-t", api.image_job(rgb, 97, 61, RGB, 8, canvas=(60, 60), resize=api.resize_both("mitchell"),
                                             values=[(V_EDGE_X, 3), (V_EDGE_Y, 1)],
                                             values_dbl=[(V_TRANSLATE_X, 2.5), (V_TRANSLATE_Y, -1.25)])))
    jobs.append(("rgb-intclamp", api.image_job(rgb, 97, 61, RGB, 8, canvas=(150, 150),
                                               resize=api.resize_both("lanczos3"), values=[(V_INTCLAMP, 1)])))
    jobs.append(("rgb-nogamma", api.image_job(rgb, 97, 61, RGB, 8, canvas=(150, 150), resize=api.resize_both("catrom"),
                                              values=[(V_NOGAMMA, 1)])))
    jobs.append(("rgb-rec709", api.image_job(rgb, 97, 61, RGB, 8, canvas=(33, 44), resize=api.resize_both("catrom"),
                                             output_cs=(api.CS["rec709"], 0.0))))
    jobs.append(("rgb-gamma", api.image_job(rgb, 97, 61, RGB, 8, canvas=(33, 44), resize=api.resize_both("catrom"),
                                            output_cs=(api.CS["gamma"], 1.8), input_cs=(api.CS["gamma"], 2.2))))
    jobs.append(("rgb16-out", api.image_job(rgb, 97, 61, RGB, 8, canvas=(70, 30), depth=16,
                                            resize=api.resize_both("lanczos3"))))
    return jobs


def c3_jobs(count=4096, seed=4, max_side=None):
    """The batched configuration: RGBA8 images of log-uniform sides, each
    resized by its own factors with its own filter, 8-bit, no dither.
    max_side keeps only images whose sides are at most that (test subsets)."""
    jobs = []
    for (w, h, ow, oh, filt, iseed) in synth.c3_list(count, seed):
        if max_side and (w > max_side or h > max_side):
            continue
        img = synth.c3_image(w, h, iseed)
        jobs.append(api.image_job(img, w, h, RGBA, 8, canvas=(ow, oh), resize=api.resize_both(filt)))
    return jobs


def c4_requests(count=1000, seed=5, lo_side=256, hi_side=8192):
    """Serving requests: (job, gap) with RGB 8/16-bit input, one scale per
    request, Lanczos3 or Mitchell, 8-bit output with Floyd-Steinberg."""
    out = []
    for (w, h, depth, ow, oh, filt, iseed, gap) in synth.c4_list(count, seed, lo_side, hi_side):
        img = synth.gratings_rgb(w, h, depth, iseed)
        pix = img if depth == 8 else synth.be16(img)
        job = api.image_job(pix, w, h, RGB, depth, canvas=(ow, oh), resize=api.resize_both(filt),
                            dither=[(api.CH_ALL, 2, 0)], depth=8)
        out.append((job, gap))
    return out


def assert_same(a, b, what=""):
    assert a["ok"], (what, a["error"])
    assert b["ok"], (what, b["error"])
    for k in ("width", "height", "imgtype", "bit_depth", "bpr", "has_colorkey", "colorkey", "palette"):
        assert a[k] == b[k], (what, k, a[k], b[k])
    if not np.array_equal(a["pixels"], b["pixels"]):
        d = np.nonzero(a["pixels"] != b["pixels"])[0]
        raise AssertionError("%s: %d bytes differ, first at %d (%d vs %d)" %
                             (what, len(d), d[0], a["pixels"][d[0]], b["pixels"][d[0]]))
