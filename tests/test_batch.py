"""Batches (iwamd_batch_*): many jobs in one device block, one hipGraph with
persistent band kernels.  Every job's bytes equal the reference core's on
the same job, whatever else is in the batch, and re-runs repeat them."""
import sys

import pytest

import cases

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_batch_mixed_options(prod, ref):
    """The option sweep (all filters, dithers incl. error diffusion and the
    reference PRNG, 16-bit output, backgrounds, edges) as one batch."""
    named = cases.sweep_jobs()
    say("batch of %d sweep jobs" % len(named))
    got = prod.run_batch([j for _, j in named])
    for (name, job), g in zip(named, got):
        cases.assert_same(ref.run(job), g, "batch " + name)


def test_batch_c3_subset_and_rerun(prod, ref):
    """C3-shaped jobs (sides <= 200) in one batch, run twice."""
    jobs = cases.c3_jobs(count=400, max_side=200)
    say("C3 subset: %d jobs" % len(jobs))
    b = prod.batch(jobs)
    try:
        b.run()
        b.run()
        b.sync()
        got = b.finish()
    finally:
        b.close()
    for k, (job, g) in enumerate(zip(jobs, got)):
        cases.assert_same(ref.run(job), g, "c3[%d]" % k)


def test_batch_alone_equals_mixed(prod):
    """A job's bytes alone and inside a larger batch are the same."""
    jobs = cases.c3_jobs(count=200, max_side=160)
    alone = [prod.run_batch([j])[0] for j in jobs[:6]]
    mixed = prod.run_batch(jobs)
    for k in range(6):
        cases.assert_same(alone[k], mixed[k], "alone/mixed %d" % k)
