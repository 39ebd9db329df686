"""Useful-FLOP counts (iwamd_job_flops, iwamd_batch_flops) that the bench
divides by the resampling time.  A nearest-neighbour enlargement by exactly 2
puts every output sample a quarter pixel from one input sample and farther
than half a pixel from every other, so each pass lists one tap per sample:
2*C*(out_h*in_w + out_w*out_h).  A batch counts the sum of its jobs."""
import sys

import numpy as np
import pytest

import cases
from imageworsener_amd import api

pytestmark = pytest.mark.gpu


def _nearest2x(w, h, seed):
    rng = np.random.default_rng(seed)
    img = rng.integers(0, 256, size=(h, w, 3), dtype=np.uint8)
    return api.image_job(img, w, h, api.IMGTYPE["RGB"], 8, canvas=(2 * w, 2 * h),
                         resize=api.resize_both("nearest"))


def _job_flops(prod, job):
    dj = prod.device_job(job)
    try:
        return dj.flops
    finally:
        dj.close()


@pytest.mark.parametrize("w,h", [(64, 48), (37, 91)])
def test_nearest_double_flops(prod, w, h):
    print("nearest x2 %dx%d" % (w, h), flush=True)
    got = _job_flops(prod, _nearest2x(w, h, w * h))
    assert got == 2 * 3 * ((2 * h) * w + (2 * w) * (2 * h))


def test_batch_flops_is_sum(prod):
    jobs = cases.c3_jobs(count=24, max_side=200)
    each = [_job_flops(prod, j) for j in jobs]
    assert all(f > 0 for f in each)
    b = prod.batch(jobs)
    try:
        total = b.flops()
    finally:
        b.close()
    assert total == pytest.approx(sum(each), rel=1e-12)
    sys.stdout.flush()
