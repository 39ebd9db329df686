"""The output optimiser with its scans on the device (src/imagew-opt.c):
transparent pixels blackened, colour / transparency / 16-bit precision
flags, colour-key candidates for binary transparency and the palette's
colour set are formed on the GPU, the host builds the new image.  Images are
built so that each decision of the optimiser goes both ways -- a colour key
other than 192, no free key, palettes of 2..256 colours spread over many
workgroups, 257 colours met only when the workgroups' sets are merged, 16-bit
images with and without 16-bit precision -- and every result must equal the
reference core's (type, depth, palette, colour key, bytes) through the
single-job path, a batch and the request loop."""
import sys

import numpy as np
import pytest

import cases
from imageworsener_amd import api

RGB, RGBA, GRAY, GRAYA = cases.RGB, cases.RGBA, cases.GRAY, cases.GRAYA
LINEAR = (1, 0.0)  # IW_CSTYPE_LINEAR: samples pass through unchanged


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def be16(a):
    return a.astype(">u2").view(np.uint8)


def job_of(arr, imgtype, depth, **kw):
    h, w = arr.shape[:2]
    pix = be16(arr) if depth == 16 else arr.astype(np.uint8)
    j = api.image_job(pix.reshape(-1), w, h, imgtype, depth, input_cs=LINEAR, output_cs=LINEAR,
                      canvas=(w, h), resize=api.resize_both("mix"), depth=depth)
    j.update(kw)
    return j


def binary_alpha(rng, h, w, full):
    return np.where(rng.random((h, w)) < 0.3, 0, full)


def rgba8_keys(rng, h, w, used_r):
    """Many colours, alpha 0 or 255, and pixels (r, 192, 192) for r in used_r."""
    a = rng.integers(0, 256, (h, w, 4))
    a[..., 3] = binary_alpha(rng, h, w, 255)
    # Keep random colours off the key plane, then place the key pixels.
    a[..., 1] = np.where(a[..., 1] == 192, 191, a[..., 1])
    idx = rng.choice(h * w, size=len(used_r) * 3, replace=False)
    for k, r in enumerate(np.repeat(used_r, 3)):
        y, x = divmod(int(idx[k]), w)
        a[y, x] = (r, 192, 192, 255)
    return a


def rgba16_keys(rng, h, w, used_lo):
    a = rng.integers(0, 65536, (h, w, 4))
    a[..., 3] = binary_alpha(rng, h, w, 65535)
    a[..., 1] = np.where(a[..., 1] == 192 * 257, 1, a[..., 1])
    idx = rng.choice(h * w, size=len(used_lo) * 2, replace=False)
    for k, lo in enumerate(np.repeat(used_lo, 2)):
        y, x = divmod(int(idx[k]), w)
        a[y, x] = (192 * 256 + lo, 192 * 257, 192 * 257, 65535)
    return a


def graya(rng, h, w, values, full, depth):
    a = np.zeros((h, w, 2), np.int64)
    a[..., 0] = rng.choice(np.asarray(values), (h, w))
    a[..., 1] = binary_alpha(rng, h, w, full)
    # Every value appears on an opaque pixel at least once.
    for k, v in enumerate(values):
        a[k // w, k % w] = (v, full)
    return a


def banded_colours(h, w, n, alpha=True):
    """n colours, each over a run of pixels: a workgroup sees one or two of
    them, the image all n."""
    run = max(1, (h * w) // n)
    k = np.minimum(np.arange(h * w) // run, n - 1).reshape(h, w)
    c = np.stack([(k * 37) % 256, (k * 101) % 256, (k * 13 + k // 256) % 256, np.full_like(k, 255)], -1)
    if not alpha:
        return c[..., :3]
    return c


def named_jobs():
    rng = np.random.default_rng(20240611)
    out = []
    # Binary transparency, RGB 8-bit: key moves away from 192; no free key.
    out.append(("rgba8 key 231", job_of(rgba8_keys(rng, 200, 300, np.arange(150, 231)), RGBA, 8)))
    out.append(("rgba8 key 192", job_of(rgba8_keys(rng, 120, 170, np.array([0, 5, 191, 193])), RGBA, 8)))
    out.append(("rgba8 no key", job_of(rgba8_keys(rng, 200, 300, np.arange(256)), RGBA, 8)))
    # RGB 16-bit with 16-bit precision.
    out.append(("rgba16 key", job_of(rgba16_keys(rng, 150, 210, np.arange(100, 200)), RGBA, 16)))
    out.append(("rgba16 no key", job_of(rgba16_keys(rng, 150, 210, np.arange(256)), RGBA, 16)))
    # Gray with binary alpha: 255 values + transparent -> gray8 key.
    vals = [v for v in range(256) if v != 77]
    out.append(("graya8 key 77", job_of(graya(rng, 160, 240, vals, 255, 8), GRAYA, 8)))
    out.append(("graya8 full", job_of(graya(rng, 160, 240, list(range(256)), 255, 8), GRAYA, 8)))
    lo = [192 * 256 + v for v in range(256) if v != 3]
    g16 = graya(rng, 128, 128, lo + [5, 1000, 65535], 65535, 16)
    out.append(("graya16 key 3", job_of(g16, GRAYA, 16)))
    # Colour images without transparency that are gray: RGB -> GRAY.
    g = rng.integers(0, 256, (90, 130))
    out.append(("rgb8 gray", job_of(np.stack([g, g, g], -1), RGB, 8)))
    # Palettes: few colours spread over many workgroups.
    out.append(("rgba8 250 colours", job_of(banded_colours(400, 300, 250), RGBA, 8)))
    out.append(("rgba8 256 colours", job_of(banded_colours(400, 300, 256), RGBA, 8)))
    out.append(("rgba8 257 colours", job_of(banded_colours(400, 300, 257), RGBA, 8)))
    out.append(("rgb8 200 colours", job_of(banded_colours(333, 517, 200, alpha=False), RGB, 8)))
    # Many colours within the first workgroup.
    out.append(("rgb8 noise", job_of(rng.integers(0, 256, (64, 4096, 3)), RGB, 8)))
    # Translucent palette with cleared transparent colours.
    c = banded_colours(300, 300, 120)
    c[..., 3] = np.where((np.arange(300)[:, None] + np.arange(300)[None, :]) % 7 == 0, 0,
                         np.where(np.arange(300)[None, :] % 5 == 0, 128, 255))
    out.append(("rgba8 translucent palette", job_of(c, RGBA, 8)))
    # Low-depth gray palettes.
    out.append(("gray8 2 levels", job_of(rng.choice([0, 255], (200, 211)), GRAY, 8)))
    out.append(("gray8 4 levels", job_of(rng.choice([0, 85, 170, 255], (200, 211)), GRAY, 8)))
    out.append(("gray8 16 levels", job_of(rng.choice(np.arange(16) * 17, (200, 211)), GRAY, 8)))
    out.append(("gray8 17 levels", job_of(rng.choice(np.arange(17) * 15, (200, 211)), GRAY, 8)))
    # 16-bit without 16-bit precision: reduced to 8 bits, then a palette.
    out.append(("rgb16 reducible palette", job_of(banded_colours(200, 260, 180, alpha=False) * 257, RGB, 16)))
    out.append(("rgba16 reducible", job_of(banded_colours(210, 190, 300) * 257, RGBA, 16)))
    # 16-bit gray with and without precision.
    out.append(("gray16 precise", job_of(rng.integers(0, 65536, (100, 170)), GRAY, 16)))
    out.append(("gray16 reducible", job_of(rng.integers(0, 256, (100, 170)) * 257, GRAY, 16)))
    # Options that switch parts of the optimiser off.
    out.append(("rgba8 key, no binary trns", job_of(rgba8_keys(rng, 80, 90, np.arange(150, 200)), RGBA, 8,
                                                     allow_opt=[(5, 0)])))
    out.append(("rgba8 palette, no palette", job_of(banded_colours(100, 120, 60), RGBA, 8,
                                                     allow_opt=[(2, 0)])))
    out.append(("rgba8 PAM profile", job_of(banded_colours(100, 120, 60), RGBA, 8, profile=api.PROFILE_PAM)))
    return out


JOBS = named_jobs()


@pytest.mark.gpu
@pytest.mark.parametrize("name,job", JOBS, ids=[n for n, _ in JOBS])
def test_opt_single(prod, ref, name, job):
    say("opt scan single: " + name)
    cases.assert_same(ref.run(job), prod.run(job), name)


@pytest.mark.gpu
def test_opt_batch(prod, ref):
    say("opt scan batch of %d" % len(JOBS))
    got = prod.run_batch([j for _, j in JOBS])
    for (name, job), g in zip(JOBS, got):
        cases.assert_same(ref.run(job), g, "batch " + name)


@pytest.mark.gpu
def test_opt_server(prod, ref):
    say("opt scan request loop, %d requests" % len(JOBS))
    srv = api.Server(prod, streams=3)
    try:
        for _, job in JOBS:
            srv.submit(job)
        got = srv.results()
    finally:
        srv.close()
    for (name, job), g in zip(JOBS, got):
        cases.assert_same(ref.run(job), g, "server " + name)


def test_opt_outcomes_cover_each_branch(ref):
    """The reference takes every branch these images were built for (on
    the CPU, so the cases stay meaningful whatever the device does)."""
    res = {n: ref.run(j) for n, j in JOBS}
    pal = api.IMGTYPE.get("PALETTE", 0x1000)
    assert res["rgba8 key 231"]["imgtype"] == RGB and res["rgba8 key 231"]["colorkey"] == (231, 192, 192)
    assert res["rgba8 key 192"]["colorkey"] == (192, 192, 192)
    assert not res["rgba8 no key"]["has_colorkey"] and res["rgba8 no key"]["imgtype"] == RGBA
    assert res["rgba16 key"]["has_colorkey"] and res["rgba16 key"]["bit_depth"] == 16
    assert res["rgba16 key"]["colorkey"][0] == 192 * 256 + 200
    assert not res["rgba16 no key"]["has_colorkey"] and res["rgba16 no key"]["imgtype"] == RGBA
    assert res["graya8 key 77"]["colorkey"][0] == 77 and res["graya8 key 77"]["imgtype"] == GRAY
    assert res["graya8 full"]["imgtype"] == GRAYA
    assert res["graya16 key 3"]["colorkey"][0] == 192 * 256 + 3 and res["graya16 key 3"]["bit_depth"] == 16
    assert res["rgb8 gray"]["imgtype"] == GRAY
    for name, n in (("rgba8 250 colours", 250), ("rgba8 256 colours", 256), ("rgb8 200 colours", 200),
                    ("gray8 2 levels", 2), ("gray8 4 levels", 4), ("gray8 16 levels", 16),
                    ("rgb16 reducible palette", 180)):
        assert res[name]["imgtype"] == pal and len(res[name]["palette"]) == n, name
    assert res["rgba8 translucent palette"]["imgtype"] == pal
    assert any(e[3] == 0 for e in res["rgba8 translucent palette"]["palette"])
    assert any(0 < e[3] < 255 for e in res["rgba8 translucent palette"]["palette"])
    for name in ("rgba8 257 colours", "rgb8 noise", "rgba16 reducible", "rgba8 palette, no palette",
                 "rgba8 PAM profile"):
        assert res[name]["imgtype"] == RGB, name
    assert res["gray8 17 levels"]["imgtype"] == GRAY
    assert res["rgba16 reducible"]["bit_depth"] == 8
    assert res["gray16 precise"]["bit_depth"] == 16
    assert res["gray16 reducible"]["bit_depth"] == 8
    assert res["rgba8 key, no binary trns"]["imgtype"] == RGBA


def test_opt_host_scans_on_oracle(prod, ref, oracle_exec):
    """The same images with the pixels from the CPU oracle and the scans on
    the host (the path a context without a device scan takes)."""
    for name, job in JOBS:
        cases.assert_same(ref.run(job), prod.run(job, executor=oracle_exec), "host " + name)
