"""Orientation and crop through the direct colour-in: the band kernel and the
chosen-precision kernel read input pixels through an affine byte map built
from the same logical-to-physical function as the generic reader
(src/imagew-main.c:129-276; the settings: src/imagew-api.c:366, :660).  Every orientation of
iw_reorient_image, with and without a crop, for RGBA8 (one 32-bit load per
pixel), RGB8 (byte loads) and RGBA16 inputs: the plain form is byte-identical
to the reference core, the chosen form within 1 LSB of the plain form."""
import numpy as np
import pytest

import cases
from imageworsener_amd import api

pytestmark = pytest.mark.gpu

FORMATS = ["rgba8", "rgb8", "rgba16"]


def _job(fmt, orient, crop):
    w, h = 90, 61
    rgba = cases.synth.c3_image(w, h, 7 + orient)
    if fmt == "rgba8":
        pix, t, depth = rgba, api.IMGTYPE["RGBA"], 8
    elif fmt == "rgb8":
        pix, t, depth = np.ascontiguousarray(rgba[..., :3]), api.IMGTYPE["RGB"], 8
    else:
        pix, t, depth = cases.synth.be16(rgba.astype(np.uint16) * 257), api.IMGTYPE["RGBA"], 16
    job = api.image_job(pix, w, h, t, depth, canvas=(47, 70), resize=api.resize_both("lanczos3"),
                        reorient=orient)
    if crop:
        job["crop"] = (3, 2, 80, 55)
    return job


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("crop", [False, True])
@pytest.mark.parametrize("orient", range(8))
def test_orientation(prod, ref, fmt, crop, orient):
    print("orient %d crop %s %s" % (orient, crop, fmt), flush=True)
    job = _job(fmt, orient, crop)
    want = ref.run(job)
    got = prod.run(job)
    cases.assert_same(want, got, "plain orient %d crop %s %s" % (orient, crop, fmt))
    # Both forms keep the plan's output format (no optimiser reductions).
    no_opt = [(o, 0) for o in (1, 2, 3, 4, 5)]
    a = prod.run(dict(job, allow_opt=no_opt))
    b = prod.run(dict(job, allow_opt=no_opt, precision=api.PRECISION_FAST))
    assert a["ok"] and b["ok"], (a["error"], b["error"])
    assert api.code_error(a, b)["max_lsb"] <= 1, (orient, crop, fmt)
