"""The library's own codecs on either side of the pixel path -- PNM/PAM
(csrc/iw_pnm.cpp, reference src/imagew-pnm.c:439-958) and MIFF floating-point
images (csrc/iw_miff.cpp, reference src/imagew-miff.c:193-897) -- checked
against the reference's own codecs (oracle/_ref/libiwref_codecs.so: the
reference core and codec sources compiled by oracle/Makefile).

Each case is a whole file through the CLI's call sequence: read, settings,
process, write.  The reference side reads, processes and writes with its own
code; this library's side reads and writes with its codecs and processes
either with the CPU oracle on its plan (tests that run without a GPU: they
check the codecs) or with the GPU pipeline (marked gpu).  The written files
must be byte-identical.  Inputs are generated here: 8- and 16-bit, reduced
maxval, bitmaps, PAM tuple types, header comments; MIFF with alpha, gray,
64-bit samples, density, rendering intent, background label, a skipped
profile, and samples outside [0, 1] (MIFF carries them unclamped)."""
import struct

import numpy as np
import pytest

from imageworsener_amd import api

FMT_MIFF, FMT_PNM, FMT_PAM = 5, 8, 12
WRITE = {"pnm": ("iw_write_pnm_file", FMT_PNM), "pam": ("iw_write_pam_file", FMT_PAM),
         "miff": ("iw_write_miff_file", FMT_MIFF)}
OUR_WRITE = {"pnm": ("iw_write_file_by_fmt", FMT_PNM), "pam": ("iw_write_file_by_fmt", FMT_PAM),
             "miff": ("iw_write_file_by_fmt", FMT_MIFF)}
READ = {"pnm": "iw_read_pnm_file", "miff": "iw_read_miff_file"}
OUR_READ = {"pnm": ("iw_read_file_by_fmt", FMT_PNM), "miff": ("iw_read_file_by_fmt", FMT_MIFF)}


def _rng(seed):
    return np.random.default_rng(seed)


def pnm(code, w, h, maxval, nch, seed, comment=False):
    r = _rng(seed)
    if code == 4:
        rowbytes = (w + 7) // 8
        body = r.integers(0, 256, size=h * rowbytes, dtype=np.uint8).tobytes()
        head = b"P4\n%d %d\n" % (w, h)
        return head + body
    samples = r.integers(0, maxval + 1, size=h * w * nch)
    body = samples.astype(">u2").tobytes() if maxval > 255 else samples.astype(np.uint8).tobytes()
    if comment:
        head = b"P%d\n# made for a test\n%d %d\n# maxval next\n%d\n" % (code, w, h, maxval)
    else:
        head = b"P%d\n%d %d\n%d\n" % (code, w, h, maxval)
    return head + body


def pam(w, h, maxval, tupltype, nch, seed):
    r = _rng(seed)
    samples = r.integers(0, maxval + 1, size=h * w * nch)
    if tupltype.startswith(b"BLACKANDWHITE"):
        samples = samples % 2
    body = samples.astype(">u2").tobytes() if maxval > 255 else samples.astype(np.uint8).tobytes()
    head = (b"P7\nWIDTH %d\nHEIGHT %d\nDEPTH %d\nMAXVAL %d\nTUPLTYPE %s\nENDHDR\n" %
            (w, h, nch, maxval, tupltype))
    return head + body


def miff(w, h, gray, alpha, depth, seed, extra=b"", cs=b"RGB", gamma=True, lo=-0.2, hi=1.3, profile=0):
    r = _rng(seed)
    nch = (1 if gray else 3) + (1 if alpha else 0)
    v = r.uniform(lo, hi, size=h * w * nch)
    if alpha:
        v.reshape(-1, nch)[:, -1] = r.uniform(0.0, 1.0, size=h * w)
    head = b"id=ImageMagick  version=1.0\nclass=DirectClass  colors=0  matte=%s\n" % (b"True" if alpha else b"False")
    head += b"columns=%d  rows=%d  depth=%d\n" % (w, h, depth)
    head += b"type=Grayscale\ncolorspace=Gray\n" if gray else b"colorspace=%s\n" % cs
    head += b"compression=None  quality=0\n" + extra
    if gamma:
        head += b"gamma=1.0\n"
    if profile:
        head += b"profile:icc=%d\n" % profile
    head += b"quantum:format={floating-point}\n\x0c\x0a\x3a\x1a"
    body = (b"\x5a" * profile) + (v.astype(">f4") if depth == 32 else v.astype(">f8")).tobytes()
    return head + body


def _settings(kind, w, h):
    if kind == "same":
        return {"canvas": (w, h)}
    if kind == "lanczos3":
        return {"canvas": (max(1, w * 2 // 3), max(1, h * 3 // 2)), "resize": api.resize_both("lanczos3")}
    if kind == "fs":
        return {"canvas": (w + 5, max(1, h - 3)), "resize": api.resize_both("catrom"), "dither": [(api.CH_ALL, 2, 0)]}
    if kind == "depth16":
        return {"canvas": (max(1, w // 2), h), "resize": api.resize_both("mitchell"), "depth": 16}
    raise ValueError(kind)


PNM_INPUTS = [
    ("pgm8", lambda: pnm(5, 37, 23, 255, 1, 1), 37, 23),
    ("pgm16", lambda: pnm(5, 29, 31, 65535, 1, 2), 29, 31),
    ("ppm8-comments", lambda: pnm(6, 41, 19, 255, 3, 3, comment=True), 41, 19),
    ("ppm16", lambda: pnm(6, 23, 27, 65535, 3, 4), 23, 27),
    ("ppm-maxval1000", lambda: pnm(6, 30, 20, 1000, 3, 5), 30, 20),
    ("pgm-maxval15", lambda: pnm(5, 33, 21, 15, 1, 6), 33, 21),
    ("pbm", lambda: pnm(4, 45, 17, 1, 1, 7), 45, 17),
    ("pam-rgba8", lambda: pam(26, 22, 255, b"RGB_ALPHA", 4, 8), 26, 22),
    ("pam-graya16", lambda: pam(24, 18, 65535, b"GRAYSCALE_ALPHA", 2, 9), 24, 18),
    ("pam-gray8", lambda: pam(28, 16, 255, b"GRAYSCALE", 1, 10), 28, 16),
    ("pam-bw", lambda: pam(20, 12, 1, b"BLACKANDWHITE", 1, 11), 20, 12),
]

MIFF_INPUTS = [
    ("rgb32", lambda: miff(21, 17, False, False, 32, 21), 21, 17, True),
    ("rgba64", lambda: miff(19, 23, False, True, 64, 22), 19, 23, True),
    ("graya32", lambda: miff(25, 13, True, True, 32, 23), 25, 13, True),
    ("gray32-density-intent-bkgd-profile",
     lambda: miff(17, 15, True, False, 32, 24, profile=37,
                  extra=b"units=PixelsPerCentimeter\nresolution=37.80x42.50\nrendering-intent=Relative\n"
                        b"background-color=#ffff80000000\n"), 17, 15, True),
    # No gamma attribute: the reference assumes sRGB and converts each float
    # sample with libm (the GPU path converts the image on the host first).
    ("rgb32-srgb", lambda: miff(18, 14, False, False, 32, 25, gamma=False, lo=0.0, hi=1.0), 18, 14, False),
]


def _pair(ref_codecs, lib, data, src, dst, job, executor):
    want = ref_codecs.run_file(data, READ[src], WRITE[dst], job)
    got = lib.run_file(data, OUR_READ[src], OUR_WRITE[dst], job, executor=executor)
    return want, got


def _assert_same(name, want, got):
    assert want[0], (name, "reference failed", want[2])
    assert got[0], (name, got[2])
    if want[1] != got[1]:
        a, b = want[1], got[1]
        k = next(i for i in range(min(len(a), len(b))) if a[i] != b[i]) if a[:len(b)] != b[:len(a)] else min(len(a), len(b))
        raise AssertionError("%s: %d vs %d bytes, first difference at %d" % (name, len(a), len(b), k))


CPU_PNM = [(n, f, w, h, s, o) for (n, f, w, h) in PNM_INPUTS for s in ("same", "lanczos3", "fs") for o in ("pnm", "pam")]


@pytest.mark.parametrize("name,make,w,h,kind,out", CPU_PNM, ids=["%s-%s-%s" % (c[0], c[4], c[5]) for c in CPU_PNM])
def test_pnm_through_oracle(prod, ref_codecs, oracle_exec, name, make, w, h, kind, out):
    data = make()
    want, got = _pair(ref_codecs, prod, data, "pnm", out, _settings(kind, w, h), oracle_exec)
    _assert_same(name, want, got)


CPU_MIFF = [(n, f, w, h, s, o) for (n, f, w, h, _) in MIFF_INPUTS for s in ("same", "lanczos3") for o in ("miff", "pnm")]


@pytest.mark.parametrize("name,make,w,h,kind,out", CPU_MIFF, ids=["%s-%s-%s" % (c[0], c[4], c[5]) for c in CPU_MIFF])
def test_miff_through_oracle(prod, ref_codecs, oracle_exec, name, make, w, h, kind, out):
    data = make()
    want, got = _pair(ref_codecs, prod, data, "miff", out, _settings(kind, w, h), oracle_exec)
    _assert_same(name, want, got)


@pytest.mark.parametrize("name,make,w,h", [(n, f, w, h) for (n, f, w, h) in PNM_INPUTS[:4]],
                         ids=[c[0] for c in PNM_INPUTS[:4]])
def test_pnm_to_miff_through_oracle(prod, ref_codecs, oracle_exec, name, make, w, h):
    """Integer input, floating-point MIFF output (the HDRI profile)."""
    want, got = _pair(ref_codecs, prod, make(), "pnm", "miff", _settings("lanczos3", w, h), oracle_exec)
    _assert_same(name, want, got)


def test_miff_reads_back_what_it_writes(prod, ref_codecs, oracle_exec):
    """A MIFF written by this library reads back (with either library) to a
    file that writes out the same bytes again: the float round trip."""
    data = miff(23, 19, False, True, 32, 31)
    job = _settings("same", 23, 19)
    ok, once, err = prod.run_file(data, OUR_READ["miff"], OUR_WRITE["miff"], job, executor=oracle_exec)
    assert ok, err
    ok, twice, err = prod.run_file(once, OUR_READ["miff"], OUR_WRITE["miff"], job, executor=oracle_exec)
    assert ok, err
    ok, ref_twice, err = ref_codecs.run_file(once, READ["miff"], WRITE["miff"], job)
    assert ok, err
    assert once == twice == ref_twice
    # The float samples survive unchanged (same size, no clamping).
    hdr = once.index(b"\x0c\x0a\x3a\x1a") + 4
    body = np.frombuffer(once[hdr:], dtype=">f4")
    src = np.frombuffer(data[data.index(b"\x0c\x0a\x3a\x1a") + 4:], dtype=">f4")
    assert body.shape == src.shape and (body.min() < 0.0 or body.max() > 1.0)


@pytest.mark.parametrize("case", ["zip", "class", "depth16", "format", "truncated"])
def test_miff_errors_like_the_reference(prod, ref_codecs, oracle_exec, case):
    data = miff(8, 6, False, False, 32, 41)
    if case == "zip":
        data = data.replace(b"compression=None", b"compression=Zip")
    elif case == "class":
        data = data.replace(b"class=DirectClass", b"class=PseudoClass")
    elif case == "depth16":
        data = data.replace(b"depth=32", b"depth=16")
    elif case == "format":
        data = data.replace(b"{floating-point}", b"{unsigned}")
    else:
        data = data[:-10]
    want, got = _pair(ref_codecs, prod, data, "miff", "miff", _settings("same", 8, 6), oracle_exec)
    assert not want[0] and not got[0]
    assert want[2] == got[2], (want[2], got[2])


@pytest.mark.parametrize("case", ["maxval0", "maxval70000", "badsig", "truncated"])
def test_pnm_errors_like_the_reference(prod, ref_codecs, oracle_exec, case):
    data = pnm(6, 8, 6, 255, 3, 51)
    if case == "maxval0":
        data = data.replace(b"\n255\n", b"\n0\n")
    elif case == "maxval70000":
        data = data.replace(b"\n255\n", b"\n70000\n")
    elif case == "badsig":
        data = b"Q" + data[1:]
    else:
        data = data[:-7]
    want, got = _pair(ref_codecs, prod, data, "pnm", "pnm", _settings("same", 8, 6), oracle_exec)
    assert want[0] == got[0]
    if not want[0]:
        assert want[2] == got[2], (want[2], got[2])
    else:
        assert want[1] == got[1]


GPU_PNM = [(n, f, w, h, s, o) for (n, f, w, h) in PNM_INPUTS for s in ("lanczos3", "fs", "depth16") for o in ("pnm", "pam")]


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,w,h,kind,out", GPU_PNM, ids=["%s-%s-%s" % (c[0], c[4], c[5]) for c in GPU_PNM])
def test_pnm_through_gpu(prod, ref_codecs, name, make, w, h, kind, out):
    want, got = _pair(ref_codecs, prod, make(), "pnm", out, _settings(kind, w, h), None)
    _assert_same(name, want, got)


GPU_MIFF = [(n, f, w, h, s, o) for (n, f, w, h, _) in MIFF_INPUTS for s in ("same", "lanczos3")
            for o in ("miff", "pnm")]


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,w,h,kind,out", GPU_MIFF, ids=["%s-%s-%s" % (c[0], c[4], c[5]) for c in GPU_MIFF])
def test_miff_through_gpu(prod, ref_codecs, name, make, w, h, kind, out):
    want, got = _pair(ref_codecs, prod, make(), "miff", out, _settings(kind, w, h), None)
    _assert_same(name, want, got)
