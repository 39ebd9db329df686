"""GPU paths that need their own cases, each compared byte for byte with the
reference core (oracle/_ref/libiwref.so) on the same inputs:

- error diffusion on output rows wider than 4096 pixels, where the error rows
  live in HBM instead of LDS (errdiff_kernel<*, *, false>): Floyd-Steinberg
  and the x+2 matrices, 8-bit and 5-level outputs (decision records in LDS),
  4- and 2-level outputs (records in registers) and 16-bit outputs (tables in
  HBM) -- the reference's scatter is src/imagew-main.c:497-567 and its row
  rotation :1124-1136;
- the full 2048^2 -> 3072^2 Gray16 image with ordered, random (seed 3) and
  Jarvis-Judice-Ninke dithering to 4 levels;
- floating-point output (IW_PROFILE_HDRI, put_raw_sample_flt32,
  src/imagew-main.c:419-425, :632-636, :1117-1118);
- floating-point input in a linear colour space (get_raw_sample_flt32,
  src/imagew-main.c:129-134, :267-272), and the input depths other than 32 that
  the reference reads as 0.0;
- floating-point input in sRGB, Rec.709 and gamma colour spaces, which the
  reference converts per sample with libm (src/imagew-main.c:53-66,
  :355-364) and the host converts to linear doubles before the upload;
- reduced input and output maximum colour codes (iw_set_input_max_color_code
  and iw_set_output_max_color_code, src/imagew-api.c:384, :442, reaching
  src/imagew-main.c:1982-2007 and :2145-2183);
- the chosen-precision form's error taken directly against the reference.
"""
import sys

import numpy as np
import pytest

import cases
from imageworsener_amd import api, synth

pytestmark = pytest.mark.gpu

RGB, RGBA, GRAY, GRAYA = cases.RGB, cases.RGBA, cases.GRAY, cases.GRAYA
IW_VAL_CVT_TO_GRAYSCALE = 11  # include/imagew.h
PROFILE_HDRI = 0x4000
SAMPLETYPE_FLOAT = 1


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# ---- error diffusion on rows wider than 4096 ------------------------------

def _wide_jobs():
    g16 = synth.be16(synth.c2(160)[:40])          # 160 x 40 Gray16
    rgb16 = synth.be16(synth.gratings_rgb(200, 30, 16, 77))
    rgba8 = synth.c3_image(120, 25, 9)
    jobs = []
    for d in ("f", "jjn", "stucki", "atkinson", "sierralite"):
        fam, sub = api.DITHER[d]
        jobs.append(("g16-5000x9-" + d, api.image_job(g16, 160, 40, GRAY, 16, canvas=(5000, 9),
                                                      resize=api.resize_both("catrom"),
                                                      dither=[(api.CH_ALL, fam, sub)])))
        # 5 levels: records in LDS; 4 and 2 levels: records in registers.
        for cc in (5, 4, 2):
            jobs.append(("g16-5000x9-cc%d-%s" % (cc, d), api.image_job(g16, 160, 40, GRAY, 16, canvas=(5000, 9),
                                                                      resize=api.resize_both("catrom"),
                                                                      dither=[(api.CH_ALL, fam, sub)],
                                                                      color_count=[(api.CH_ALL, cc)])))
    for d in ("f", "jjn"):
        fam, sub = api.DITHER[d]
        jobs.append(("rgb16-4600x6-" + d, api.image_job(rgb16, 200, 30, RGB, 16, canvas=(4600, 6),
                                                        resize=api.resize_both("lanczos3"),
                                                        dither=[(api.CH_ALL, fam, sub)])))
        # 16-bit output: 65536-level decision tables stay in HBM.
        jobs.append(("rgb16-4600x6-d16-" + d, api.image_job(rgb16, 200, 30, RGB, 16, canvas=(4600, 6), depth=16,
                                                            resize=api.resize_both("mitchell"),
                                                            dither=[(api.CH_ALL, fam, sub)])))
        jobs.append(("rgba8-4099x7-" + d, api.image_job(rgba8, 120, 25, RGBA, 8, canvas=(4099, 7),
                                                        resize=api.resize_both("lanczos2"),
                                                        dither=[(api.CH_ALL, fam, sub)])))
    return jobs


@pytest.mark.parametrize("path", ["band", "twopass"])
@pytest.mark.parametrize("name,job", _wide_jobs(), ids=[n for n, _ in _wide_jobs()])
def test_errdiff_rows_over_4096(prod, ref, name, job, path):
    say("%s via %s" % (name, path))
    cases.assert_same(ref.run(job), prod.run(dict(job, options=[("amd:resample", path)])), name)


def test_errdiff_rows_over_4096_fast_form_runs(prod):
    """The chosen form hands the same wide planes to the same error-diffusion
    kernels; its result has the plan's shape and codes."""
    name, job = _wide_jobs()[0]
    a = prod.run(dict(job, precision=api.PRECISION_FAST))
    assert a["ok"], a["error"]
    assert (a["width"], a["height"]) == (5000, 9)


# ---- full C2 --------------------------------------------------------------

@pytest.mark.parametrize("dither", ["o", "r", "jjn"])
def test_c2_full(prod, ref, dither):
    job = cases.c2_job(dither=dither)
    say("C2 2048^2 G16 linear -> 3072^2 catrom cc4 %s (reference on CPU ~3 s)" % dither)
    cases.assert_same(ref.run(job), prod.run(job), "C2 " + dither)


# ---- floating-point output ------------------------------------------------

def _float_out_jobs():
    rgba = synth.c3_image(83, 71, 7)
    rgb = synth.c0(97, 61)
    g16 = synth.be16(synth.c2(64))
    prof = api.PROFILE_PNG | PROFILE_HDRI
    return [
        ("rgba8-float-down", api.image_job(rgba, 83, 71, RGBA, 8, canvas=(41, 29), profile=prof,
                                           resize=api.resize_both("lanczos3"))),
        ("rgba8-float-up", api.image_job(rgba, 83, 71, RGBA, 8, canvas=(131, 107), profile=prof,
                                         resize=api.resize_both("catrom"))),
        ("rgb8-float-linear-out", api.image_job(rgb, 97, 61, RGB, 8, canvas=(60, 40), profile=prof,
                                                resize=api.resize_both("mitchell"),
                                                output_cs=(api.CS["linear"], 0.0))),
        ("g16-float-dither-ignored", api.image_job(g16, 64, 64, GRAY, 16, canvas=(45, 77), profile=prof,
                                                   resize=api.resize_both("lanczos2"),
                                                   dither=[(api.CH_ALL, 2, 0)], color_count=[(api.CH_ALL, 4)])),
        ("rgba8-float-bkgd", api.image_job(rgba, 83, 71, RGBA, 8, canvas=(50, 50), profile=prof,
                                           resize=api.resize_both("catrom"), bkgd=(0.2, 0.5, 0.9, 1.0))),
        ("rgb8-float-wide-rows", api.image_job(rgb, 97, 61, RGB, 8, canvas=(2300, 13), profile=prof,
                                               resize=api.resize_both("triangle"))),
    ]


@pytest.mark.parametrize("path", ["band", "twopass"])
@pytest.mark.parametrize("name,job", _float_out_jobs(), ids=[n for n, _ in _float_out_jobs()])
def test_float_output(prod, ref, name, job, path):
    want = ref.run(job)
    got = prod.run(dict(job, options=[("amd:resample", path)]))
    assert want["ok"] and want["sampletype"] == SAMPLETYPE_FLOAT and want["bit_depth"] == 32, want
    cases.assert_same(want, got, name)
    assert got["sampletype"] == SAMPLETYPE_FLOAT


# ---- floating-point input -------------------------------------------------

def float_job(arr, imgtype, canvas, depth_bits=32, **settings):
    """Interleaved float32 samples (shape (h, w, nch)) as a job; the stored
    depth may be other than 32 to reach the reference's read-as-zero case."""
    h, w = arr.shape[:2]
    nch = api.NCH[imgtype]
    raw = np.ascontiguousarray(arr, dtype="<f4").view(np.uint8).reshape(-1)
    bpr = (w * depth_bits * nch + 7) // 8
    if depth_bits != 32:
        raw = np.zeros(bpr * h, np.uint8)
    job = {"image": {"pixels": raw, "width": w, "height": h, "imgtype": imgtype, "bit_depth": depth_bits,
                     "bpr": bpr, "sampletype": SAMPLETYPE_FLOAT}, "canvas": canvas}
    job.update(settings)
    return job


def _float_in_jobs():
    rng = np.random.default_rng(21)
    rgba = rng.random((53, 67, 4), dtype=np.float64).astype(np.float32)
    rgba[..., 3][rgba[..., 3] < 0.1] = 0.0
    rgb = (rng.random((40, 70, 3)) * 1.2 - 0.1).astype(np.float32)   # out-of-range samples too
    gray = rng.random((64, 48, 1)).astype(np.float32)
    lin = (api.CS["linear"], 0.0)
    return [
        ("rgba-f32-lin-8bit", float_job(rgba, RGBA, (30, 25), input_cs=lin, resize=api.resize_both("lanczos3"))),
        ("rgba-f32-lin-16bit-up", float_job(rgba, RGBA, (101, 90), input_cs=lin, depth=16,
                                            resize=api.resize_both("catrom"))),
        ("rgb-f32-lin-range-fs", float_job(rgb, RGB, (50, 33), input_cs=lin, resize=api.resize_both("mitchell"),
                                           dither=[(api.CH_ALL, 2, 0)])),
        ("rgb-f32-lin-float-out", float_job(rgb, RGB, (50, 33), input_cs=lin, profile=api.PROFILE_PNG | PROFILE_HDRI,
                                            resize=api.resize_both("lanczos2"))),
        ("gray-f32-lin-ordered", float_job(gray, GRAY, (31, 77), input_cs=lin, resize=api.resize_both("triangle"),
                                           dither=[(api.CH_ALL, 1, 0)], color_count=[(api.CH_ALL, 6)])),
        ("rgba-f32-lin-crop-orient", dict(float_job(rgba, RGBA, (33, 29), input_cs=lin,
                                                    resize=api.resize_both("catrom")),
                                          crop=(3, 5, 50, 40), reorient=5)),
        ("rgb-f16-reads-zero", float_job(rgb, RGB, (20, 10), depth_bits=16, input_cs=lin,
                                         resize=api.resize_both("catrom"))),
        ("rgb-f16-srgb-reads-zero", float_job(rgb, RGB, (20, 10), depth_bits=16,
                                              resize=api.resize_both("catrom"))),
    ]


@pytest.mark.parametrize("path", ["band", "twopass"])
@pytest.mark.parametrize("name,job", _float_in_jobs(), ids=[n for n, _ in _float_in_jobs()])
def test_float_input(prod, ref, name, job, path):
    cases.assert_same(ref.run(job), prod.run(dict(job, options=[("amd:resample", path)])), name)


def _float_nonlinear_jobs():
    rng = np.random.default_rng(3)
    rgb = (rng.random((20, 30, 3)) * 1.1 - 0.05).astype(np.float32)   # out-of-range samples too
    rgba = rng.random((26, 19, 4)).astype(np.float32)
    rgba[..., 3][rgba[..., 3] < 0.15] = 0.0
    return [
        ("rgb-f32-srgb", float_job(rgb, RGB, (15, 10), resize=api.resize_both("catrom"))),   # default input cs: sRGB
        ("rgb-f32-rec709-fs", float_job(rgb, RGB, (41, 23), input_cs=(api.CS["rec709"], 0.0),
                                        resize=api.resize_both("lanczos3"), dither=[(api.CH_ALL, 2, 0)])),
        ("rgba-f32-gamma", float_job(rgba, RGBA, (33, 40), input_cs=(api.CS["gamma"], 1.8),
                                     resize=api.resize_both("mitchell"))),
        ("rgba-f32-srgb-gray-crop-orient", dict(float_job(rgba, RGBA, (12, 14), resize=api.resize_both("triangle"),
                                                          values=[(IW_VAL_CVT_TO_GRAYSCALE, 1)]),
                                                crop=(2, 3, 15, 20), reorient=6)),
    ]


@pytest.mark.parametrize("path", ["band", "twopass"])
@pytest.mark.parametrize("name,job", _float_nonlinear_jobs(), ids=[n for n, _ in _float_nonlinear_jobs()])
def test_float_input_nonlinear(prod, ref, name, job, path):
    """32-bit float input in a non-linear colour space: the reference converts
    each sample with libm's pow (src/imagew-main.c:53-66, :355-364); the host
    converts the image to linear doubles with the same expressions before the
    upload, so the bytes are the reference's."""
    want = ref.run(job)
    assert want["ok"], want["error"]
    cases.assert_same(want, prod.run(dict(job, options=[("amd:resample", path)])), name)


# ---- reduced maximum colour codes -----------------------------------------

def _maxcc_jobs():
    rgb = synth.c0(97, 61)
    rgba = synth.c3_image(83, 71, 7)
    g16 = synth.be16(synth.c2(64))
    nonalpha, alpha, gray = api.CH_NONALPHA, api.CH_ALPHA, 4
    return [
        ("in-maxcc-200-ch0", api.image_job(rgb, 97, 61, RGB, 8, canvas=(41, 29), resize=api.resize_both("catrom"),
                                           input_maxcc=[(0, 200)])),
        ("in-maxcc-all-rgba", api.image_job(rgba, 83, 71, RGBA, 8, canvas=(60, 50), resize=api.resize_both("lanczos3"),
                                            input_maxcc=[(0, 250), (1, 255), (2, 127), (3, 250)])),
        ("in-maxcc-g16-1000", api.image_job(g16, 64, 64, GRAY, 16, canvas=(45, 77), depth=16,
                                            resize=api.resize_both("mitchell"), input_maxcc=[(0, 1000)])),
        ("out-maxcc-31-rgb", api.image_job(rgb, 97, 61, RGB, 8, canvas=(41, 29), resize=api.resize_both("catrom"),
                                           output_maxcc=[(0, 31), (1, 63), (2, 31)])),
        ("out-maxcc-rgba-fs", api.image_job(rgba, 83, 71, RGBA, 8, canvas=(57, 49), resize=api.resize_both("catrom"),
                                            output_maxcc=[(0, 15), (1, 15), (2, 15), (alpha, 3)],
                                            dither=[(api.CH_ALL, 2, 0)])),
        # A reduced red code on a grey image makes the reference write RGB.
        ("out-maxcc-gray-red-ordered", api.image_job(g16, 64, 64, GRAY, 16, canvas=(45, 77),
                                                     resize=api.resize_both("lanczos2"), output_maxcc=[(0, 5)],
                                                     dither=[(api.CH_ALL, 1, 0)])),
        ("out-maxcc-gray-type-ignored", api.image_job(g16, 64, 64, GRAY, 16, canvas=(45, 77),
                                                      resize=api.resize_both("lanczos2"), output_maxcc=[(gray, 5)])),
        ("out-maxcc-16bit-1023", api.image_job(g16, 64, 64, GRAY, 16, canvas=(45, 77), depth=16,
                                               resize=api.resize_both("lanczos3"), output_maxcc=[(1, 1023)],
                                               dither=[(api.CH_ALL, 3, 0)], random_seed=4)),
        ("out-maxcc-cc", api.image_job(rgb, 97, 61, RGB, 8, canvas=(41, 29), resize=api.resize_both("catrom"),
                                       output_maxcc=[(0, 100), (1, 100), (2, 100)], color_count=[(api.CH_ALL, 6)])),
        ("both-maxcc", api.image_job(rgba, 83, 71, RGBA, 8, canvas=(41, 90), resize=api.resize_both("mitchell"),
                                     input_maxcc=[(3, 200)], output_maxcc=[(alpha, 200), (nonalpha, 0)],
                                     dither=[(api.CH_ALL, 2, 1)])),
    ]


@pytest.mark.parametrize("path", ["band", "twopass"])
@pytest.mark.parametrize("name,job", _maxcc_jobs(), ids=[n for n, _ in _maxcc_jobs()])
def test_max_color_codes(prod, ref, name, job, path):
    want = ref.run(job)
    if name.startswith(("out-", "both")) and "ignored" not in name:
        # The reduced codes must reach the output description, or the case
        # pins nothing.
        assert want["ok"] and want["reduced_maxcolors"], name
    cases.assert_same(want, prod.run(dict(job, options=[("amd:resample", path)])), name)
    got = prod.run(job)
    assert got["reduced_maxcolors"] == want["reduced_maxcolors"]
    assert got["maxcolorcode"] == want["maxcolorcode"]


# ---- chosen form against the reference -------------------------------------

NO_OPT = [(o, 0) for o in (1, 2, 3, 4, 5)]


def _chosen_jobs():
    jobs = [("C0", cases.c0_job())]
    jobs += [("C3-%d" % i, j) for i, j in enumerate(cases.c3_jobs(count=4096, max_side=256)[:40])]
    jobs += [(n, j) for n, j in cases.sweep_jobs() if n.startswith("rgb-down") or n.startswith("rgba-up")]
    return jobs


@pytest.mark.parametrize("name,job", _chosen_jobs(), ids=[n for n, _ in _chosen_jobs()])
def test_chosen_form_vs_reference(prod, ref, name, job):
    """The chosen form (FP32 MFMA resampling) against the
    reference core's bytes: at most 1 LSB on 8-bit outputs without dither."""
    want = ref.run(dict(job, allow_opt=NO_OPT))
    got = prod.run(dict(job, allow_opt=NO_OPT, precision=api.PRECISION_FAST))
    assert want["ok"] and got["ok"], (want["error"], got["error"])
    e = api.code_error(want, got)
    assert e["max_lsb"] <= 1, (name, e)
