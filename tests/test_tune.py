"""The autotuner (iwamd_tune_entries): band height and tile width change how
the work is cut, not the bytes, so every shape gives the same output, and a
batch fills the table with the fastest timed candidate of each class."""
import sys

import pytest

import cases
from imageworsener_amd import api

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _jobs():
    rgba = cases.synth.c3_image(203, 157, 9)
    jobs = [("c0", cases.c0_job())]
    for filt, canvas in (("lanczos3", (90, 61)), ("catrom", (377, 290)), ("triangle", (150, 400))):
        jobs.append(("rgba-%s" % filt, api.image_job(rgba, 203, 157, api.IMGTYPE["RGBA"], 8, canvas=canvas,
                                                     resize=api.resize_both(filt))))
    return jobs


def _with(job, **opts):
    j = dict(job)
    j["options"] = list(job.get("options", [])) + [(k, str(v)) for k, v in opts.items()]
    return j


@pytest.mark.parametrize("name,job", _jobs(), ids=lambda v: v if isinstance(v, str) else "")
def test_band_rows_same_bytes(prod, name, job):
    base = prod.run(_with(job, **{"amd:resample": "band"}))
    assert base["ok"], base["error"]
    for rows in (1, 2, 4):
        got = prod.run(_with(job, **{"amd:resample": "band", "amd:band_rows": rows}))
        cases.assert_same(base, got, "%s rows=%d" % (name, rows))


@pytest.mark.parametrize("name,job", _jobs(), ids=lambda v: v if isinstance(v, str) else "")
def test_tile_width_same_bytes(prod, name, job):
    fast = dict(job, precision=api.PRECISION_FAST)
    a = prod.run(_with(fast, **{"amd:tile": 16}))
    b = prod.run(_with(fast, **{"amd:tile": 32}))
    cases.assert_same(a, b, name)


def test_batch_tunes_classes(prod):
    jobs = cases.c3_jobs(count=120, max_side=320)
    plain = prod.run_batch(jobs)
    fast = prod.run_batch([dict(j, precision=api.PRECISION_FAST) for j in jobs])
    table = prod.tune_entries()
    say("tuned classes: %d" % len(table))
    for e in table[:12]:
        say("  %s" % e)
    assert any(e["form"] == "band" for e in table) and any(e["form"] == "tile" for e in table)
    for e in table:
        c = e["candidates"]
        assert len(c) >= 2 and e["shape"] in c
        assert c[e["shape"]] == min(c.values())
    # Tuned shapes leave the bytes as the job alone gives them.
    for k in (0, len(jobs) // 2, len(jobs) - 1):
        cases.assert_same(prod.run(jobs[k]), plain[k], "job %d" % k)
    assert all(r["ok"] for r in fast)
