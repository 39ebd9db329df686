"""The error-diffusion decision data the GPU walk uses in place of the
reference's per-pixel distance comparison (src/imagew-main.c:712-748):

  M[k]  -- least s at which a non-exact sample of level k rounds up, i.e.
           where (lin[k+1] - s) <= (s - lin[k]) first holds in FP64;
  eb[k] -- the largest |error| a sample of level k leaves.

Both are checked against the reference's own formulas evaluated directly on
the same doubles: the boundary of M on adjacent doubles, the decision on
random samples, and eb as an upper bound that some sample attains."""
import ctypes
import random
import struct

import pytest

from imageworsener_amd import api


def _prev(d):
    return struct.unpack("<d", struct.pack("<q", struct.unpack("<q", struct.pack("<d", d))[0] - 1))[0]


def _tables(prod, cstype, gamma, maxcc, cc):
    lay = api._PlanLayout.get()
    och = lay.OChan()
    och.ditherfamily, och.dithersubtype = api.DITHER["f"]
    och.color_count = cc
    och.maxcolorcode_int = maxcc
    och.maxcolorcode_dbl = float(maxcc)
    och.out_cs = api.CsDescr(cstype, 0, gamma)
    P = (cc - 1) if cc else maxcc
    lin = (ctypes.c_double * (P + 1))()
    M = (ctypes.c_double * (P + 1))()
    eb = (ctypes.c_double * (P + 1))()
    prod.dll.iwamd_ed_tables.argtypes = [ctypes.c_void_p] * 4
    prod.dll.iwamd_ed_tables.restype = ctypes.c_int
    assert prod.dll.iwamd_ed_tables(ctypes.byref(och), lin, M, eb) == P
    T = (ctypes.c_double * P)()
    U = (ctypes.c_double * (P + 1))()
    prod.dll.iwamd_quant_thresholds.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                ctypes.c_void_p, ctypes.c_void_p]
    assert prod.dll.iwamd_quant_thresholds(cstype, gamma, P, T, U)
    return P, list(T), list(U), list(lin), list(M), list(eb)


def _ref_error(s, k, U, lin):
    """The reference's choice at level k: (error, rounded up)."""
    if s < U[k]:
        return 0.0, False
    d_floor = s - lin[k]
    d_ceil = lin[k + 1] - s
    if d_ceil <= d_floor:
        return -d_ceil, True
    return d_floor, False


CASES = [("srgb8", api.CS["srgb"], 0.0, 255, 0), ("linear-cc4", api.CS["linear"], 0.0, 255, 4),
         ("gamma-cc7", api.CS["gamma"], 2.2, 255, 7), ("rec709-8", api.CS["rec709"], 0.0, 255, 0),
         ("gamma05-cc5", api.CS["gamma"], 0.5, 255, 5), ("srgb16", api.CS["srgb"], 0.0, 65535, 0)]


@pytest.mark.parametrize("name,cstype,gamma,maxcc,cc", CASES, ids=[c[0] for c in CASES])
def test_ed_tables(prod, name, cstype, gamma, maxcc, cc):
    P, T, U, lin, M, eb = _tables(prod, cstype, gamma, maxcc, cc)
    assert M[P] == 2.0 and eb[P] == 0.0
    rng = random.Random(7)
    levels = range(P) if P < 1000 else list(range(64)) + rng.sample(range(64, P - 64), 200) + list(range(P - 64, P))
    for k in levels:
        lo, hi = lin[k], lin[k + 1]
        up = lambda s: (hi - s) <= (s - lo)
        m = M[k]
        if m == 2.0:
            assert not up(1.0)
        elif m == 0.0:
            assert up(0.0)
        else:
            assert up(m) and not up(_prev(m)), (k, m)
        # Samples of level k that are not exact: [U[k], T[k]).
        a, b = U[k], T[k]
        if not a < b:
            assert eb[k] == 0.0
            continue
        pts = [a, _prev(b)]
        if a < m < b:
            pts += [m, _prev(m)]
        pts += [a + (b - a) * rng.random() for _ in range(40)]
        worst = 0.0
        for s in pts:
            if not a <= s < b:
                continue
            err, ceil_ = _ref_error(s, k, U, lin)
            assert ceil_ == (s >= m), (k, s)
            assert abs(err) <= eb[k], (k, s, err, eb[k])
            worst = max(worst, abs(err))
        assert worst == eb[k], (k, worst, eb[k])


def _next(d):
    return struct.unpack("<d", struct.pack("<q", struct.unpack("<q", struct.pack("<d", d))[0] + 1))[0]


def _level(s, T):
    import bisect
    return bisect.bisect_right(T, s)


@pytest.mark.parametrize("name,cstype,gamma,maxcc,cc", CASES, ids=[c[0] for c in CASES])
def test_ed_fast_form(prod, name, cstype, gamma, maxcc, cc):
    """The fast walk's decision form (one round-up threshold, one exact
    interval per window) gives the reference's error for every sample of
    every window it accepts; windows of one level, and of two levels that
    stay clear of the lower level's exact samples and of a second round-up
    threshold, are accepted."""
    P, T, U, lin, M, eb = _tables(prod, cstype, gamma, maxcc, cc)
    lay = api._PlanLayout.get()
    och = lay.OChan()
    och.ditherfamily, och.dithersubtype = api.DITHER["f"]
    och.color_count = cc
    och.maxcolorcode_int = maxcc
    och.maxcolorcode_dbl = float(maxcc)
    och.out_cs = api.CsDescr(cstype, 0, gamma)
    fn = prod.dll.iwamd_ed_fast_check
    fn.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    ordered = all(M[k] <= T[k] <= M[k + 1] for k in range(P))
    rng = random.Random(11)
    levels = range(P + 1) if P < 1000 else list(range(40)) + rng.sample(range(40, P - 40), 100) + list(range(P - 40, P + 1))
    accepted = windows = 0
    for k in levels:
        lo_edge = T[k - 1] if k > 0 else 0.0
        hi_edge = T[k] if k < P else 1.0
        span = max(hi_edge - lo_edge, 1e-12)
        for _ in range(12):
            # Windows inside level k and windows reaching into level k+1.
            sL = min(max(lo_edge + span * rng.random(), 0.0), 1.0)
            sH = min(sL + span * rng.choice([0.05, 0.3, 0.6, 1.2]) * rng.random(), 1.0)
            pts = [sL, sH] + [sL + (sH - sL) * rng.random() for _ in range(24)]
            for th in ([T[k]] if k < P else []) + [U[k], M[k]] + ([U[k + 1], M[k + 1]] if k < P else []):
                for v in (_prev(th), th, _next(th)):
                    if sL <= v <= sH:
                        pts.append(v)
            n = len(pts)
            s_arr = (ctypes.c_double * n)(*pts)
            e_arr = (ctypes.c_double * n)()
            r = fn(ctypes.byref(och), sL, sH, n, s_arr, e_arr)
            windows += 1
            kl, kh = _level(sL, T), _level(sH, T)
            if kl == kh:
                assert r == 1, (k, sL, sH)
            elif kh == kl + 1 and ordered and sL >= U[kl] and not (sL < M[kl] and sH >= M[kl + 1]):
                assert r == 1, (k, sL, sH)
            else:
                assert r == 0, (k, sL, sH)
            if r != 1:
                continue
            accepted += 1
            for s, got in zip(pts, e_arr):
                want, _ = _ref_error(s, _level(s, T), U, lin)
                assert got == want, (name, k, sL, sH, s, got, want)
    assert accepted > windows // 2


def _ref_code(s, P, T, U, lin, code):
    """The reference's no-dither output code for linear sample s
    (src/imagew-main.c:669-736): the floor level, kept when exact, else the
    nearer of the two candidates in linear light (ties up)."""
    k = _level(s, T)
    if k >= P or s < U[k]:
        return code[k]
    return code[k + 1] if (lin[k + 1] - s) <= (s - lin[k]) else code[k]


@pytest.mark.parametrize("name,cstype,gamma,maxcc,cc", CASES, ids=[c[0] for c in CASES])
def test_quant_fold(prod, name, cstype, gamma, maxcc, cc):
    """The kernels' no-dither quantiser -- one search over the folded
    boundaries D[k] = min(T[k], max(T[k-1], U[k], M[k])) through the bin
    index -- gives the reference's code at every threshold, exactness bound
    and round-up point and the doubles beside them, and at random samples."""
    P, T, U, lin, M, eb = _tables(prod, cstype, gamma, maxcc, cc)
    code = [float(k) if not cc else float(int(0.5000000001 + k * (maxcc / P))) for k in range(P + 1)]
    lay = api._PlanLayout.get()
    och = lay.OChan()
    och.ditherfamily, och.dithersubtype = api.DITHER["none"]
    och.color_count = cc
    och.maxcolorcode_int = maxcc
    och.maxcolorcode_dbl = float(maxcc)
    och.out_cs = api.CsDescr(cstype, 0, gamma)
    fn = prod.dll.iwamd_quant_fold_check
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    rng = random.Random(13)
    levels = range(P + 1) if P < 1000 else list(range(64)) + rng.sample(range(64, P - 64), 300) + list(range(P - 64, P + 1))
    pts = [0.0, 1.0, _prev(1.0)] + [rng.random() for _ in range(4000)]
    for k in levels:
        for th in ([T[k]] if k < P else []) + [U[k], M[k]]:
            if 0.0 <= th <= 1.0:
                pts += [v for v in (_prev(th), th, _next(th)) if 0.0 <= v <= 1.0]
    n = len(pts)
    s_arr = (ctypes.c_double * n)(*pts)
    c_arr = (ctypes.c_double * n)()
    assert fn(ctypes.byref(och), n, s_arr, c_arr) == 1
    bad = [(s, c_arr[i], _ref_code(s, P, T, U, lin, code)) for i, s in enumerate(pts)
           if c_arr[i] != _ref_code(s, P, T, U, lin, code)]
    assert not bad, bad[:5]
