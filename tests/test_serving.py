"""The request loop (iwamd_server_*): requests submitted at Poisson arrival
times while others run; each result equals the same job run alone (and so
the reference's bytes), and requests overlap on the device."""
import sys
import time

import pytest

import cases

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_serving_mixed_equals_alone(prod):
    from imageworsener_amd import api
    reqs = cases.c4_requests(count=24, lo_side=128, hi_side=640)
    say("serving: %d requests" % len(reqs))
    srv = api.Server(prod, hbm_budget=256 << 20, streams=4)
    try:
        for job, gap in reqs:
            srv.submit(job)
            time.sleep(0.002 * gap)
        got = srv.results()
    finally:
        srv.close()
    for k, ((job, _), g) in enumerate(zip(reqs, got)):
        cases.assert_same(prod.run(job), g, "request %d" % k)
    spans = [g["times_ms"][1:] for g in got]
    overlap = max(sum(1 for s, d in spans if s <= t < d) for t, _ in spans)
    say("max requests in flight together: %d" % overlap)
    assert overlap >= 2


def test_serving_matches_reference(prod, ref):
    from imageworsener_amd import api
    reqs = cases.c4_requests(count=6, lo_side=96, hi_side=256)
    srv = api.Server(prod, streams=2)
    try:
        for job, _ in reqs:
            srv.submit(job)
        got = srv.results()
    finally:
        srv.close()
    for k, ((job, _), g) in enumerate(zip(reqs, got)):
        cases.assert_same(ref.run(job), g, "request %d" % k)
