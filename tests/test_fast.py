"""The chosen-precision form (IWAMD_PRECISION_FAST: FP32 MFMA resampling,
FP32 sums, FP64 quantisation) against the plain FP64 form on the GPU, whose
bytes equal the reference's.  Bar: at most 1 LSB on 8-bit outputs without
error diffusion (ordered and random dither included: a decision can move by
one level); with error diffusion a small difference reroutes the error, so
only the mean is bounded there.  The colour of a pixel whose alpha is 0 in
either result is not compared (the output optimiser blackens fully
transparent pixels)."""
import sys

import numpy as np
import pytest

import cases
from imageworsener_amd import api

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# Both forms keep the plan's output format (no optimiser reductions), so their
# samples line up.
NO_OPT = [(o, 0) for o in (1, 2, 3, 4, 5)]


def plain(job):
    return dict(job, allow_opt=NO_OPT)


def fast(job):
    return dict(job, allow_opt=NO_OPT, precision=api.PRECISION_FAST)


def lsb_stats(a, b):
    """Max and mean code difference (colour of fully transparent pixels left
    out, see api.code_error)."""
    assert a["ok"] and b["ok"], (a["error"], b["error"])
    assert (a["width"], a["height"], a["bit_depth"]) == (b["width"], b["height"], b["bit_depth"])
    e = api.code_error(a, b)
    return e["max_lsb"], e["mean_lsb"]


def test_fast_c0(prod):
    job = cases.c0_job()
    mx, mean = lsb_stats(prod.run(plain(job)), prod.run(fast(job)))
    say("C0 fast vs plain: max %d LSB, mean %.5f LSB" % (mx, mean))
    assert mx <= 1 and mean < 0.05


@pytest.mark.parametrize("name,job", [(n, j) for n, j in cases.sweep_jobs()
                                      if j.get("depth", 8) == 8 and "-f" not in n and "jjn" not in n
                                      and "stucki" not in n and "burkes" not in n and "sierra" not in n
                                      and "atkinson" not in n and "wide-ed" not in n],
                         ids=lambda v: v if isinstance(v, str) else "")
def test_fast_sweep(prod, name, job):
    """At most one output level apart: 1 LSB, or one posterised step
    (255/(n-1) codes) with a colour count n."""
    mx, mean = lsb_stats(prod.run(plain(job)), prod.run(fast(job)))
    cc = dict(job.get("color_count", [])).get(10, 0)
    step = -(-255 // (cc - 1)) if cc else 1
    assert mx <= step, (name, mx, mean)


def _jobs16():
    """16-bit outputs: the sweep's 16-bit cases without error diffusion, and
    RGBA16 -> RGBA16 down and up (premultiplied alpha, sRGB)."""
    jobs = [(n, j) for n, j in cases.sweep_jobs() if j.get("depth", 8) == 16 and "wide-ed" not in n and not n.endswith("-f")]
    img = cases.synth.be16(cases.synth.c1(512))
    jobs.append(("rgba16-down", api.image_job(img, 512, 512, api.IMGTYPE["RGBA"], 16, canvas=(125, 125), depth=16,
                                              resize=api.resize_both("mitchell"))))
    jobs.append(("rgba16-up", api.image_job(img, 512, 512, api.IMGTYPE["RGBA"], 16, canvas=(700, 610), depth=16,
                                            resize=api.resize_both("lanczos3"))))
    return jobs


@pytest.mark.parametrize("name,job", _jobs16(), ids=lambda v: v if isinstance(v, str) else "")
def test_fast_16bit(prod, name, job):
    """16-bit outputs: FP32 operands and sums keep the resampled value within
    a small fraction of a 16-bit step, so codes differ by at most 1 LSB."""
    a, b = prod.run(plain(job)), prod.run(fast(job))
    assert a["bit_depth"] == 16
    mx, mean = lsb_stats(a, b)
    say("%s fast vs plain (16-bit): max %d LSB, mean %.5f LSB" % (name, mx, mean))
    assert mx <= 1 and mean < 0.05, (name, mx, mean)


def test_fast_c3_batch(prod):
    jobs = cases.c3_jobs(count=300, max_side=256)
    plain_res = prod.run_batch([plain(j) for j in jobs])
    quick = prod.run_batch([fast(j) for j in jobs])
    worst = 0
    for k, (a, b) in enumerate(zip(plain_res, quick)):
        mx, _ = lsb_stats(a, b)
        worst = max(worst, mx)
    say("C3 subset (%d jobs) fast vs plain: max %d LSB" % (len(jobs), worst))
    assert worst <= 1


def test_fast_error_diffusion_mean(prod):
    job = cases.c1_job(size=512, out=125)
    mx, mean = lsb_stats(prod.run(plain(job)), prod.run(fast(job)))
    say("C1/8 FS fast vs plain: max %d LSB, mean %.4f LSB" % (mx, mean))
    assert mean < 0.5


def test_fast_alone_equals_mixed(prod):
    """A tile's sums depend only on its own plan: a job's bytes alone (work
    items of single tiles) and inside a batch large enough that work items
    are runs of up to eight tiles of a band are the same."""
    jobs = [fast(j) for j in cases.c3_jobs(count=400, max_side=512)]
    mixed = prod.run_batch(jobs)
    for k in range(0, len(jobs), 67):
        alone = prod.run(jobs[k])
        cases.assert_same(alone, mixed[k], "fast alone/mixed %d" % k)
    say("fast form: %d jobs alone equal to the batch of %d" % (len(range(0, len(jobs), 67)), len(jobs)))
