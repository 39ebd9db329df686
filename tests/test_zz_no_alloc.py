"""iwamd_process_device_noalloc: the pixel pass with the input and output in
the caller's device memory and every other buffer in the library's fixed
per-GPU array.  Each result is checked byte for byte against iw_process_image
(the existing entry point) on the same job with every optional optimisation
of the output turned off (iw_set_allow_opt), which leaves the pixel
pipeline's image with only the reference's always-on clearing of fully
transparent pixels' colour (src/imagew-opt.c:1200) and the negation where it
is asked for; two calls in a row, the second on a smaller input, share the
array.  A job too large for the array, or given short buffers, fails with an
error that names the sizes."""
import ctypes

import numpy as np
import pytest

import cases
from imageworsener_amd import api, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# IW_OPT_GRAYSCALE, _PALETTE, _16_TO_8, _STRIP_ALPHA, _BINARY_TRNS (src/imagew.h).
NO_OPT = [(k, 0) for k in (1, 2, 3, 4, 5)]
V_NEGATE = 53  # IW_VAL_NEGATE_TARGET


def _fn(lib):
    fn = lib.dll.iwamd_process_device_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                   ctypes.c_void_p, ctypes.POINTER(api.Image)]
    fn.restype = ctypes.c_int
    return fn


def _noalloc(lib, job, out_bytes):
    d = lib.dll
    ctx = d.iw_create_context(None)
    if not ctx:
        raise MemoryError("iw_create_context failed")
    try:
        lib._setup(ctx, job)
        pix = np.ascontiguousarray(job["image"]["pixels"], dtype=np.uint8).reshape(-1)
        d_in = torch.from_numpy(pix.copy()).to("cuda")
        d_out = torch.full((out_bytes,), 0xA5, dtype=torch.uint8, device="cuda")
        st = torch.cuda.current_stream()
        desc = api.Image()
        ok = _fn(lib)(ctx, d_in.data_ptr(), d_in.numel(), d_out.data_ptr(), out_bytes,
                      ctypes.c_void_p(st.cuda_stream), ctypes.byref(desc))
        err = None
        if not ok:
            buf = ctypes.create_string_buffer(300)
            err = d.iw_get_errormsg(ctx, buf, 300).decode(errors="replace")
        st.synchronize()
        return bool(ok), err, d_out.cpu().numpy(), desc
    finally:
        d.iw_destroy_context(ctx)


def _check(lib, name, job):
    job = dict(job, allow_opt=NO_OPT)
    want = lib.run(job)
    assert want["ok"], (name, want["error"])
    n = want["bpr"] * want["height"]
    ok, err, got, desc = _noalloc(lib, job, n + 64)
    assert ok, (name, err)
    assert (desc.width, desc.height, desc.bpr, desc.imgtype, desc.bit_depth) == (
        want["width"], want["height"], want["bpr"], want["imgtype"], want["bit_depth"]), name
    assert not desc.pixels
    assert np.array_equal(got[:n], want["pixels"]), name
    # Nothing past the image is written.
    assert (got[n:] == 0xA5).all(), name


def _jobs():
    RGB, RGBA, GRAY = cases.RGB, cases.RGBA, cases.GRAY
    rgba = synth.c3_image(256, 192, 21)
    g16 = synth.be16(synth.c2(160))
    rgb16 = synth.be16(synth.gratings_rgb(300, 200, 16, 5))
    return [
        ("rgba8-fs", api.image_job(rgba, 256, 192, RGBA, 8, canvas=(200, 150), resize=api.resize_both("mitchell"),
                                   dither=[(api.CH_ALL, 2, 0)])),
        ("rgb8-small-lanczos3-fs", api.image_job(synth.c0(96, 72), 96, 72, RGB, 8, canvas=(50, 37),
                                                 resize=api.resize_both("lanczos3"), dither=[(api.CH_ALL, 2, 0)])),
        ("g16-jjn-cc4", api.image_job(g16, 160, 160, GRAY, 16, canvas=(240, 240), resize=api.resize_both("catrom"),
                                      input_cs=(api.CS["linear"], 0.0), color_count=[(api.CH_ALL, 4)],
                                      dither=[(api.CH_ALL, 2, 1)])),
        ("rgb16-ordered-depth16", api.image_job(rgb16, 300, 200, RGB, 16, canvas=(123, 77), depth=16,
                                                resize=api.resize_both("lanczos2"), dither=[(api.CH_ALL, 1, 0)])),
        ("rgba8-chosen-form", dict(api.image_job(rgba, 256, 192, RGBA, 8, canvas=(97, 211),
                                                 resize=api.resize_both("catrom")), precision=api.PRECISION_FAST)),
        ("rgba8-negate-o", api.image_job(rgba, 256, 192, RGBA, 8, canvas=(130, 90), resize=api.resize_both("triangle"),
                                         dither=[(api.CH_ALL, 1, 0)], values=[(V_NEGATE, 1)])),
        ("rgb16-negate-depth16", api.image_job(rgb16, 300, 200, RGB, 16, canvas=(64, 48), depth=16,
                                               resize=api.resize_both("mitchell"), values=[(V_NEGATE, 1)])),
    ]


def test_noalloc_matches_process_image(prod):
    for name, job in _jobs():
        print(name, flush=True)
        _check(prod, name, job)


def test_noalloc_two_calls_second_smaller(prod):
    """The second call's input and output are smaller than the first's; the
    array keeps what the first call left, and the second must not read it."""
    big = api.image_job(synth.c3_image(300, 280, 3), 300, 280, cases.RGBA, 8, canvas=(420, 390),
                        resize=api.resize_both("lanczos3"), dither=[(api.CH_ALL, 2, 0)])
    small = api.image_job(synth.c3_image(40, 30, 4), 40, 30, cases.RGBA, 8, canvas=(23, 17),
                          resize=api.resize_both("lanczos3"), dither=[(api.CH_ALL, 2, 0)])
    _check(prod, "big", big)
    _check(prod, "small", small)
    _check(prod, "big-again", big)


def test_noalloc_rejects_jobs_over_the_limit(prod):
    limit = prod.dll.iwamd_noalloc_scratch_bytes
    limit.restype = ctypes.c_size_t
    # 3 planes x 3000 x 3000 FP64 error-diffusion samples: far above 64 MiB.
    job = api.image_job(synth.c0(64, 48), 64, 48, cases.RGB, 8, canvas=(3000, 3000),
                        resize=api.resize_both("triangle"), dither=[(api.CH_ALL, 2, 0)])
    ok, err, _, _ = _noalloc(prod, job, 3000 * 3000 * 3)
    assert not ok
    assert "scratch" in err and str(limit()) in err, err


def test_noalloc_rejects_short_buffers(prod):
    job = api.image_job(synth.c0(64, 48), 64, 48, cases.RGB, 8, canvas=(30, 20),
                        resize=api.resize_both("triangle"))
    ok, err, _, _ = _noalloc(prod, job, 30 * 20 * 3 - 1)
    assert not ok and "Device output" in err, err
