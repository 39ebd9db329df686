"""GPU parity: the HIP path through the C-ABI against the reference core,
byte for byte, and the reference's own regression cases through the
reference CLI linked against this library."""
import os
import sys

import numpy as np
import pytest

import cases
import fixtures

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_device_present(prod):
    assert prod.dll.iwamd_device_count() >= 1


@pytest.mark.skipif(not os.path.exists(os.path.join(fixtures.REF_DIR, "imagew-amd")),
                    reason="oracle/_ref not built")
def test_golden_fixtures_drop_in_cli():
    """The reference's CLI and codecs, unmodified, linked against this
    library, reproduce the reference's expected outputs."""
    exp = fixtures.load_expected()
    bad = []
    for c in fixtures.load_cases():
        ok, msg = fixtures.compare_case("imagew-amd", c, exp)
        if not ok:
            bad.append("%s: %s" % (c["name"], msg))
    say("golden cases failing: %d" % len(bad))
    assert not bad, "\n".join(bad)


def test_c0_full(prod, ref):
    job = cases.c0_job()
    say("C0 640x480 RGB8 -> 320x240 lanczos3")
    cases.assert_same(ref.run(job), prod.run(job), "C0")


def test_c1_reduced(prod, ref):
    job = cases.c1_job(size=1024, out=250)
    say("C1/4 1024^2 RGBA16 -> 250^2 mitchell FS")
    cases.assert_same(ref.run(job), prod.run(job), "C1/4")


@pytest.mark.parametrize("dither", ["o", "r", "jjn"])
def test_c2_reduced(prod, ref, dither):
    job = cases.c2_job(size=512, out=768, dither=dither)
    say("C2/4 512^2 G16 -> 768^2 catrom cc4 " + dither)
    cases.assert_same(ref.run(job), prod.run(job), "C2/4 " + dither)


@pytest.mark.parametrize("path", ["band", "twopass"])
@pytest.mark.parametrize("name,job", cases.sweep_jobs(), ids=[n for n, _ in cases.sweep_jobs()])
def test_sweep(prod, ref, name, job, path):
    """Every sweep case through the fused band kernel and through the
    two-pass kernels."""
    j = dict(job, options=[("amd:resample", path)])
    cases.assert_same(ref.run(job), prod.run(j), name + " " + path)


@pytest.mark.parametrize("path", ["band", "twopass"])
def test_c1_reduced_paths(prod, ref, path):
    job = cases.c1_job(size=1024, out=250)
    say("C1/4 via " + path)
    cases.assert_same(ref.run(job), prod.run(dict(job, options=[("amd:resample", path)])), "C1/4 " + path)


def test_c1_full(prod, ref):
    job = cases.c1_job()
    say("C1 4096^2 RGBA16 -> 1000^2 mitchell FS (reference on CPU ~6 s)")
    cases.assert_same(ref.run(job), prod.run(job), "C1")


def test_device_job_reruns_identical(prod, ref):
    """A resident job re-run gives the same bytes (error-diffusion state and
    outputs are re-initialised by every run)."""
    job = cases.c2_job(size=256, out=384, dither="f")
    dj = prod.device_job(job)
    for _ in range(3):
        dj.run()
    dj.sync()
    res = dj.finish()
    cases.assert_same(ref.run(job), res, "rerun")
