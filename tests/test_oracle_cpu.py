"""CPU tests: pin the oracle and the library's host layer.

 - the reference's own regression outputs (tests/golden, 128 cases) through
   the reference core and through "library decisions + oracle pixels";
 - byte equality of oracle vs reference core on the benchmark configurations
   (reduced sizes) and an option sweep;
 - closed forms and brute force for the pieces the GPU path reformulates:
   PRNG jump-ahead, nearest-colour search, quantisation thresholds.
"""
import ctypes
import math
import os
import struct

import numpy as np
import pytest

import cases
import fixtures

pytestmark = []


def ref_cli_available():
    return os.path.exists(os.path.join(fixtures.REF_DIR, "imagew"))


# ---- golden regression fixtures ------------------------------------------

GOLDEN = fixtures.load_cases()


@pytest.mark.skipif(not ref_cli_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("binary", ["imagew", "imagew-oracle"])
def test_golden_fixtures(binary):
    exp = fixtures.load_expected()
    bad = []
    for c in GOLDEN:
        ok, msg = fixtures.compare_case(binary, c, exp)
        if not ok:
            bad.append("%s: %s" % (c["name"], msg))
    assert not bad, "\n".join(bad)


# ---- oracle vs reference core, in process ---------------------------------

def test_c0_oracle_equals_reference(prod, ref, oracle_exec):
    job = cases.c0_job()
    cases.assert_same(ref.run(job), prod.run(job, executor=oracle_exec), "C0")


def test_c1_reduced_oracle_equals_reference(prod, ref, oracle_exec):
    job = cases.c1_job(size=512, out=125)
    cases.assert_same(ref.run(job), prod.run(job, executor=oracle_exec), "C1/8")


@pytest.mark.parametrize("dither", ["o", "r", "jjn"])
def test_c2_reduced_oracle_equals_reference(prod, ref, oracle_exec, dither):
    job = cases.c2_job(size=128, out=192, dither=dither)
    cases.assert_same(ref.run(job), prod.run(job, executor=oracle_exec), "C2/16 " + dither)


@pytest.mark.parametrize("name,job", cases.sweep_jobs(), ids=[n for n, _ in cases.sweep_jobs()])
def test_sweep_oracle_equals_reference(prod, ref, oracle_exec, name, job):
    cases.assert_same(ref.run(job), prod.run(job, executor=oracle_exec), name)


# ---- closed forms ----------------------------------------------------------

A, M = 0xfff0bf23, 0xfff0bf2300000001


def test_c3_c4_shapes_oracle_equals_reference(prod, ref, oracle_exec):
    """Small jobs of the batched and serving configurations (RGBA8 with
    per-axis factors; RGB 8/16-bit with Floyd-Steinberg)."""
    jobs = cases.c3_jobs(count=300, max_side=120)
    jobs += [j for j, _ in cases.c4_requests(count=6, lo_side=64, hi_side=160)]
    for k, job in enumerate(jobs):
        cases.assert_same(ref.run(job), prod.run(job, executor=oracle_exec), "shape %d" % k)


@pytest.mark.parametrize("seed", [0, 3, 12345, -7, 0x7fffffff])
def test_prng_jump_ahead_closed_form(oracle, seed):
    """Q_n = a*x_n + c_n + 1 satisfies b*Q_n == Q_(n-1) (mod m), so
    x_n = (Q_1 * b^-(n-1) mod m - 1) div a; compare with the oracle's
    sequential generator at scattered n."""
    x0 = (0x03333333 + seed) & 0xffffffff
    c0 = (0x05555555 + seed) & 0xffffffff
    t = A * x0 + c0
    x1, c1 = 0xffffffff - (t & 0xffffffff), t >> 32
    q1 = A * x1 + c1 + 1
    binv = pow(1 << 32, -1, M)
    for n in [1, 2, 3, 10, 997, 4096, 65537]:
        q = q1 * pow(binv, n - 1, M) % M
        assert (q - 1) // A == oracle.iwo_prng_nth(seed, n), n


def _nc_reference_search(tbl, s):
    # The reference's fixed-step search over 255 midpoints
    # (src/imagew-main.c:638-666), stated as a loop over its steps.
    x, d = 127, 64
    while True:
        x = x - d if (x > 254 or tbl[x] > s) else x + d
        if d == 1:
            return x if (x > 254 or tbl[x] > s) else x + 1
        d //= 2


def test_nc_search_is_count_of_midpoints():
    """The GPU counts midpoints <= s; the reference's fixed-step search gives
    the same index on sorted tables (with ties and out-of-range s)."""
    rng = np.random.default_rng(1)
    for trial in range(300):
        tbl = np.sort(rng.random(255) if trial % 3 else rng.integers(0, 5, 255) / 4.0)
        probes = np.concatenate([rng.random(40), tbl[rng.integers(0, 255, 20)], [-1.0, 0.0, 1.0, 2.0]])
        for s in probes:
            assert _nc_reference_search(tbl, s) == int(np.searchsorted(tbl, s, side="right"))


def _from_linear(v, cstype, gamma):
    # Transfer functions by their published definitions (IEC 61966-2-1 sRGB,
    # ITU-R BT.709, pure gamma), with the reference's 0.999999999 shortcut.
    if v > 0.999999999:
        return 1.0
    if cstype == 1:
        return v
    if cstype == 2:
        return math.pow(v, 1.0 / gamma)
    if cstype == 3:
        return 4.5 * v if v < 0.020 else 1.099 * math.pow(v, 0.45) - 0.099
    return 12.92 * v if v <= 0.0031308 else 1.055 * math.pow(v, 1.0 / 2.4) - 0.055


def _level(s, cstype, gamma, P):
    e = _from_linear(s, cstype, gamma) * P
    return min(max(e, 0.0), float(P))


def _prev(d):
    return struct.unpack("<d", struct.pack("<q", struct.unpack("<q", struct.pack("<d", d))[0] - 1))[0]


def _next(d):
    return struct.unpack("<d", struct.pack("<q", struct.unpack("<q", struct.pack("<d", d))[0] + 1))[0]


@pytest.mark.parametrize("cstype,gamma,P", [(0, 0.0, 255), (0, 0.0, 3), (1, 0.0, 255), (2, 1.8, 255),
                                            (3, 0.0, 255), (0, 0.0, 1), (0, 0.0, 65535)])
def test_quant_thresholds_against_libm(prod, cstype, gamma, P):
    """Each threshold is where libm's level crosses an integer: e(T_k) >= k >
    e(prev T_k), and e(U_k) > k >= e(prev U_k).  For 8-bit tables also scan
    +-64 doubles around every threshold for a non-monotone level."""
    T = (ctypes.c_double * P)()
    U = (ctypes.c_double * (P + 1))()
    prod.dll.iwamd_quant_thresholds.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                ctypes.c_void_p, ctypes.c_void_p]
    assert prod.dll.iwamd_quant_thresholds(cstype, gamma, P, T, U)
    ks = range(1, P + 1) if P < 1000 else list(range(1, 200)) + list(range(P - 200, P + 1))
    for k in ks:
        t = T[k - 1]
        assert _level(t, cstype, gamma, P) >= k
        assert _level(_prev(t), cstype, gamma, P) < k
    for k in (range(0, P) if P < 1000 else list(range(0, 200)) + list(range(P - 200, P))):
        u = U[k]
        assert _level(u, cstype, gamma, P) > k
        assert _level(_prev(u), cstype, gamma, P) <= k
    assert math.isinf(U[P])
    if P <= 255:
        for k in range(1, P + 1):
            s = T[k - 1]
            for _ in range(64):
                s = _prev(s)
                assert _level(s, cstype, gamma, P) < k
            s = T[k - 1]
            for _ in range(64):
                assert _level(s, cstype, gamma, P) >= k
                s = _next(s)


@pytest.mark.parametrize("family,p1,p2", [(0x25, 0.0, 0.5), (0x25, 1 / 3, 1 / 3), (0x25, 1.0, 0.0)])
def test_cubic_filter_closed_forms(oracle, family, p1, p2):
    """Mitchell-Netravali: k(0) = (6-2B)/6, k(1) = B/6, k(2) = 0, and the
    integer-shift weights sum to 1."""
    f = lambda x: oracle.iwo_filter(family, p1, p2, x, 10, 20)
    assert f(0.0) == pytest.approx((6 - 2 * p1) / 6, abs=1e-15)
    assert f(1.0) == pytest.approx(p1 / 6, abs=1e-15)
    assert f(2.0) == 0.0
    for x in np.linspace(0, 1, 17):
        assert f(x) + f(1 - x) + f(1 + x) + f(2 - x) == pytest.approx(1.0, abs=1e-12)


def test_lanczos_closed_form(oracle):
    for x in [0.0, 0.25, 0.5, 1.0, 1.5, 2.9]:
        want = 1.0 if x == 0 else (math.sin(math.pi * x) / (math.pi * x)) * (math.sin(math.pi * x / 3) / (math.pi * x / 3))
        assert oracle.iwo_filter(0x31, 3.0, 0.0, x, 10, 20) == pytest.approx(want, abs=1e-15)
    assert oracle.iwo_filter(0x31, 3.0, 0.0, 3.0, 10, 20) == 0.0


def test_srgb_transfer_closed_form(oracle):
    for v in [0.0, 0.01, 0.04045, 0.5, 0.999]:
        want = v / 12.92 if v <= 0.04045 else ((v + 0.055) / 1.055) ** 2.4
        assert oracle.iwo_to_linear(v, 0, 0.0) == want
        # The standard's toe thresholds (0.04045 / 0.0031308) do not meet
        # exactly, so the round trip is close, not exact, near the toe.
        assert oracle.iwo_from_linear(oracle.iwo_to_linear(v, 0, 0.0), 0, 0.0) == pytest.approx(v, abs=1e-7)


def test_catrom_odd_upscale_reproduces_samples(prod, ref):
    """Catmull-Rom at an odd integer upscale passes through the source
    samples at aligned centres (its taps at |x|=1,2 are exactly 0)."""
    from imageworsener_amd import api
    g = (np.arange(35 * 9, dtype=np.uint32) * 37 % 256).astype(np.uint8).reshape(9, 35)
    job = api.image_job(g, 35, 9, api.IMGTYPE["GRAY"], 8, canvas=(105, 27), resize=api.resize_both("catrom"),
                        input_cs=(api.CS["linear"], 0.0), output_cs=(api.CS["linear"], 0.0))
    r = ref.run(job)
    out = r["pixels"].reshape(27, 105)
    assert np.array_equal(out[1::3, 1::3], g)
