"""ctypes binding of the C-ABI (include/imagew.h, include/iwamd.h).

The same binding drives any library exporting the iw_* API: this package's
libimageworsener_amd.so (GPU), and in tests the reference core built under
oracle/_ref.  A job is a plain dict describing the CS2 call sequence
(iw_create_context -> iw_set_input_image -> setters ->
iw_set_output_canvas_size -> iw_process_image -> iw_get_output_image); see
run() for the keys.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libimageworsener_amd.so")

# Constants (include/imagew.h).
IMGTYPE = {"GRAY": 0x0001, "GRAYA": 0x0101, "RGB": 0x0010, "RGBA": 0x0110, "PALETTE": 0x1000}
NCH = {0x0001: 1, 0x0101: 2, 0x0010: 3, 0x0110: 4, 0x1000: 1}
RESIZE = {"null": 0x02, "nearest": 0x10, "mix": 0x11, "box": 0x20, "triangle": 0x21, "quadratic": 0x22,
          "gaussian": 0x23, "hermite": 0x24, "cubic": 0x25, "boxavg": 0x26, "sinc": 0x30, "lanczos": 0x31,
          "hanning": 0x32, "blackman": 0x33}
CH_ALL, CH_NONALPHA, CH_ALPHA = 10, 11, 3
DITHER = {"none": (0, 0), "o": (1, 0), "halftone": (1, 1), "f": (2, 0), "jjn": (2, 1), "stucki": (2, 2),
          "burkes": (2, 3), "sierra": (2, 4), "sierra2": (2, 5), "sierralite": (2, 6), "atkinson": (2, 7),
          "r": (3, 0), "r2": (3, 1)}
CS = {"srgb": 0, "linear": 1, "gamma": 2, "rec709": 3}
PROFILE_PNG = (0x0001 | 0x0002 | 0x0004 | 0x0008 | 0x0010 | 0x0020 | 0x0040 | 0x0100 | 0x0200 | 0x0400 |
               0x0800 | 0x1000 | 0x8000 | 0x10000)
PROFILE_PAM = 0x0040 | 0x8000 | 0x0002 | 0x0008 | 0x0001
PRECISION_PLAIN, PRECISION_FAST = 0, 1

# Named filters as (family, param1, param2), as the reference CLI maps them
# (src/imagew-cmd.c:1622-1735).
FILTERS = {
    "triangle": ("triangle", 0.0, 0.0), "catrom": ("cubic", 0.0, 0.5), "mitchell": ("cubic", 1 / 3, 1 / 3),
    "bspline": ("cubic", 1.0, 0.0), "lanczos2": ("lanczos", 2.0, 0.0), "lanczos3": ("lanczos", 3.0, 0.0),
    "lanczos": ("lanczos", 3.0, 0.0), "hermite": ("hermite", 0.0, 0.0), "gaussian": ("gaussian", 0.0, 0.0),
    "box": ("box", 0.0, 0.0), "mix": ("mix", 0.0, 0.0), "nearest": ("nearest", 0.0, 0.0),
    "quadratic": ("quadratic", 0.0, 0.0), "boxavg": ("boxavg", 0.0, 0.0),
}


class CsDescr(ctypes.Structure):
    _fields_ = [("cstype", ctypes.c_int), ("srgb_intent", ctypes.c_int), ("gamma", ctypes.c_double)]


class Color(ctypes.Structure):
    _fields_ = [("c", ctypes.c_double * 4)]


class Image(ctypes.Structure):
    _fields_ = [("imgtype", ctypes.c_int), ("bit_depth", ctypes.c_int), ("sampletype", ctypes.c_int),
                ("width", ctypes.c_int), ("height", ctypes.c_int), ("pixels", ctypes.c_void_p),
                ("bpr", ctypes.c_size_t), ("orient_transform", ctypes.c_uint),
                ("native_grayscale", ctypes.c_int), ("density_code", ctypes.c_int),
                ("density_x", ctypes.c_double), ("density_y", ctypes.c_double),
                ("has_colorkey_trns", ctypes.c_int), ("colorkey", ctypes.c_uint * 3),
                ("reduced_maxcolors", ctypes.c_int), ("maxcolorcode", ctypes.c_uint * 5),
                ("has_bkgdlabel", ctypes.c_int), ("bkgdlabel", Color), ("rendering_intent", ctypes.c_int)]


class RGBA8(ctypes.Structure):
    _fields_ = [("r", ctypes.c_uint8), ("g", ctypes.c_uint8), ("b", ctypes.c_uint8), ("a", ctypes.c_uint8)]


class Palette(ctypes.Structure):
    _fields_ = [("num_entries", ctypes.c_int), ("entry", RGBA8 * 256)]


_c = ctypes
_SIGS = {
    "iw_create_context": (_c.c_void_p, [_c.c_void_p]),
    "iw_destroy_context": (None, [_c.c_void_p]),
    "iw_process_image": (_c.c_int, [_c.c_void_p]),
    "iw_malloc_large": (_c.c_void_p, [_c.c_void_p, _c.c_size_t, _c.c_size_t]),
    "iw_set_input_image": (None, [_c.c_void_p, _c.POINTER(Image)]),
    "iw_get_output_image": (None, [_c.c_void_p, _c.POINTER(Image)]),
    "iw_get_output_palette": (_c.POINTER(Palette), [_c.c_void_p]),
    "iw_set_output_profile": (None, [_c.c_void_p, _c.c_uint]),
    "iw_set_output_depth": (None, [_c.c_void_p, _c.c_int]),
    "iw_set_output_canvas_size": (None, [_c.c_void_p, _c.c_int, _c.c_int]),
    "iw_set_output_image_size": (None, [_c.c_void_p, _c.c_double, _c.c_double]),
    "iw_set_input_crop": (None, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_int, _c.c_int]),
    "iw_reorient_image": (None, [_c.c_void_p, _c.c_uint]),
    "iw_set_resize_alg": (None, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_double, _c.c_double, _c.c_double]),
    "iw_set_dither_type": (None, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_int]),
    "iw_set_color_count": (None, [_c.c_void_p, _c.c_int, _c.c_int]),
    "iw_set_random_seed": (None, [_c.c_void_p, _c.c_int, _c.c_int]),
    "iw_set_value": (None, [_c.c_void_p, _c.c_int, _c.c_int]),
    "iw_set_value_dbl": (None, [_c.c_void_p, _c.c_int, _c.c_double]),
    "iw_set_input_colorspace": (None, [_c.c_void_p, _c.POINTER(CsDescr)]),
    "iw_set_output_colorspace": (None, [_c.c_void_p, _c.POINTER(CsDescr)]),
    "iw_set_apply_bkgd_2": (None, [_c.c_void_p, _c.POINTER(Color)]),
    "iw_set_bkgd_checkerboard_2": (None, [_c.c_void_p, _c.c_int, _c.POINTER(Color)]),
    "iw_set_bkgd_checkerboard_origin": (None, [_c.c_void_p, _c.c_int, _c.c_int]),
    "iw_set_channel_offset": (None, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_double]),
    "iw_set_allow_opt": (None, [_c.c_void_p, _c.c_int, _c.c_int]),
    "iw_set_max_malloc": (None, [_c.c_void_p, _c.c_size_t]),
    "iw_set_grayscale_weights": (None, [_c.c_void_p, _c.c_double, _c.c_double, _c.c_double]),
    "iw_set_input_max_color_code": (None, [_c.c_void_p, _c.c_int, _c.c_int]),
    "iw_set_output_max_color_code": (None, [_c.c_void_p, _c.c_int, _c.c_int]),
    "iw_get_errorflag": (_c.c_int, [_c.c_void_p]),
    "iw_get_errormsg": (_c.c_char_p, [_c.c_void_p, _c.c_char_p, _c.c_int]),
    "iw_get_version_int": (_c.c_int, []),
    "iw_set_option": (None, [_c.c_void_p, _c.c_char_p, _c.c_char_p]),
}
_EXT_SIGS = {
    "iwamd_prepare": (_c.c_int, [_c.c_void_p, _c.c_void_p]),
    "iwamd_finish": (_c.c_int, [_c.c_void_p, _c.c_void_p]),
    "iwamd_set_precision": (None, [_c.c_void_p, _c.c_int]),
    "iwamd_tune_entries": (_c.c_int, [_c.c_void_p, _c.c_int]),
    "iwamd_plan_size": (_c.c_size_t, []),
    "iwamd_device_count": (_c.c_int, []),
    "iwamd_version": (_c.c_char_p, []),
    "iwamd_last_timings": (_c.c_int, [_c.c_void_p, _c.POINTER(_c.c_double), _c.c_int]),
    "iwamd_job_create": (_c.c_void_p, [_c.c_void_p]),
    "iwamd_job_run": (_c.c_int, [_c.c_void_p]),
    "iwamd_job_sync": (_c.c_int, [_c.c_void_p]),
    "iwamd_job_finish": (_c.c_int, [_c.c_void_p]),
    "iwamd_job_destroy": (None, [_c.c_void_p]),
    "iwamd_job_kernel_ms": (_c.c_int, [_c.c_void_p, _c.POINTER(_c.c_double), _c.c_int]),
    "iwamd_job_sizes": (None, [_c.c_void_p, _c.POINTER(_c.c_double), _c.POINTER(_c.c_double)]),
    "iwamd_job_flops": (_c.c_double, [_c.c_void_p]),
    "iwamd_batch_flops": (_c.c_double, [_c.c_void_p]),
    "iwamd_batch_create": (_c.c_void_p, []),
    "iwamd_batch_add": (_c.c_int, [_c.c_void_p, _c.c_void_p]),
    "iwamd_batch_upload": (_c.c_int, [_c.c_void_p]),
    "iwamd_batch_run": (_c.c_int, [_c.c_void_p]),
    "iwamd_batch_sync": (_c.c_int, [_c.c_void_p]),
    "iwamd_batch_finish": (_c.c_int, [_c.c_void_p]),
    "iwamd_batch_destroy": (None, [_c.c_void_p]),
    "iwamd_batch_kernel_ms": (_c.c_int, [_c.c_void_p, _c.POINTER(_c.c_double), _c.c_int]),
    "iwamd_batch_sizes": (None, [_c.c_void_p, _c.POINTER(_c.c_double), _c.POINTER(_c.c_double),
                                 _c.POINTER(_c.c_int)]),
    "iwamd_server_create": (_c.c_void_p, [_c.c_size_t, _c.c_int]),
    "iwamd_server_submit": (_c.c_int, [_c.c_void_p, _c.c_void_p]),
    "iwamd_server_wait": (_c.c_int, [_c.c_void_p, _c.c_int]),
    "iwamd_server_drain": (_c.c_int, [_c.c_void_p]),
    "iwamd_server_request_ms": (_c.c_int, [_c.c_void_p, _c.c_int, _c.POINTER(_c.c_double),
                                           _c.POINTER(_c.c_double), _c.POINTER(_c.c_double)]),
    "iwamd_server_destroy": (None, [_c.c_void_p]),
    "iwamd_server_set_timeout": (None, [_c.c_void_p, _c.c_double]),
    "iwamd_server_error": (_c.c_char_p, [_c.c_void_p]),
    "iwamd_server_stats": (None, [_c.c_void_p, _c.POINTER(_c.c_size_t), _c.POINTER(_c.c_size_t),
                                  _c.POINTER(_c.c_int)]),
}

# Byte offsets into struct iwamd_plan used by executors driven from Python.
PLAN_OUT_BPR_AND_HEIGHT = None


class TuneEntry(ctypes.Structure):
    _fields_ = [("form", ctypes.c_int), ("channels", ctypes.c_int), ("taps_v_log2", ctypes.c_int),
                ("taps_h_log2", ctypes.c_int), ("width_log2", ctypes.c_int), ("shape", ctypes.c_int),
                ("n", ctypes.c_int), ("cand", ctypes.c_int * 3), ("ms", ctypes.c_double * 3)]


_READ_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                            ctypes.POINTER(ctypes.c_size_t))
_WRITE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t)
_CLOSE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)
_SIZE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64))
_SEEK_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int)
_TELL_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64))


class IODescr(ctypes.Structure):
    """struct iw_iodescr (include/imagew.h)."""
    _fields_ = [("fp", ctypes.c_void_p), ("read_fn", _READ_FN), ("write_fn", _WRITE_FN), ("close_fn", _CLOSE_FN),
                ("getfilesize_fn", _SIZE_FN), ("seek_fn", _SEEK_FN), ("tell_fn", _TELL_FN)]


class MemFile:
    """An in-memory file behind an iw_iodescr: read from `data`, or collect
    what is written in `buf`."""

    def __init__(self, data=b""):
        self.buf = bytearray(data)
        self.pos = 0

        def rd(ctx, io, buf, n, got):
            k = max(0, min(n, len(self.buf) - self.pos))
            ctypes.memmove(buf, bytes(self.buf[self.pos:self.pos + k]), k)
            self.pos += k
            got[0] = k
            return 1

        def wr(ctx, io, buf, n):
            self.buf += ctypes.string_at(buf, n)
            return 1

        def size(ctx, io, p):
            p[0] = len(self.buf)
            return 1

        def seek(ctx, io, off, whence):
            self.pos = off if whence == 0 else (self.pos + off if whence == 1 else len(self.buf) + off)
            return 1

        def tell(ctx, io, p):
            p[0] = self.pos
            return 1

        self._cb = (_READ_FN(rd), _WRITE_FN(wr), _CLOSE_FN(lambda c, i: 1), _SIZE_FN(size), _SEEK_FN(seek),
                    _TELL_FN(tell))
        self.io = IODescr(None, *self._cb)


class Library:
    """A loaded iw_* library.  mode RTLD_LOCAL keeps several such libraries
    (this one and the reference core) apart in one process."""

    def __init__(self, path=LIB_PATH, extensions=True):
        if not os.path.exists(path):
            raise RuntimeError("library not built: %s (run __graft_entry__.build())" % path)
        self.path = path
        self.dll = ctypes.CDLL(path, mode=os.RTLD_LOCAL | os.RTLD_NOW)
        for name, (res, args) in _SIGS.items():
            f = getattr(self.dll, name)
            f.restype = res
            f.argtypes = args
        self.has_ext = False
        if extensions and hasattr(self.dll, "iwamd_prepare"):
            for name, (res, args) in _EXT_SIGS.items():
                f = getattr(self.dll, name)
                f.restype = res
                f.argtypes = args
            self.has_ext = True

    def __getattr__(self, name):
        return getattr(self.dll, name)

    def run(self, job, executor=None):
        """Run one job.  executor: None for iw_process_image, or a callable
        (plan_ptr, img2_ptr) -> bool used between iwamd_prepare and
        iwamd_finish (tests use it to run the CPU oracle on the same plan)."""
        d = self.dll
        ctx = d.iw_create_context(None)
        if not ctx:
            raise MemoryError("iw_create_context failed")
        try:
            return self._run(ctx, job, executor)
        finally:
            d.iw_destroy_context(ctx)

    def tune_entries(self):
        """The autotuner's table (iwamd_tune_entries): one dict per tuned
        class: form ("band" rows per workgroup or "tile" output columns),
        channels, log2 buckets of taps and width, the chosen shape, and the
        timed candidates with their resampling ms."""
        n = self.dll.iwamd_tune_entries(None, 0)
        buf = (TuneEntry * max(n, 1))()
        n = self.dll.iwamd_tune_entries(buf, n)
        out = []
        for e in buf[:n]:
            out.append({"form": "band" if e.form == 0 else "tile", "channels": e.channels,
                        "taps_v_log2": e.taps_v_log2, "taps_h_log2": e.taps_h_log2, "width_log2": e.width_log2,
                        "shape": e.shape, "candidates": {e.cand[k]: round(e.ms[k], 4) for k in range(e.n)}})
        return out

    def batch(self, jobs):
        """A Batch holding `jobs`, uploaded."""
        b = Batch(self)
        try:
            for j in jobs:
                b.add(j)
            b.upload()
        except Exception:
            b.close()
            raise
        return b

    def run_batch(self, jobs):
        """Run `jobs` as one batch; results in order."""
        b = self.batch(jobs)
        try:
            b.run()
            return b.finish()
        finally:
            b.close()

    def device_job(self, job):
        """A DeviceJob: settings applied, input and tables resident on the
        GPU, kernels not yet run."""
        d = self.dll
        ctx = d.iw_create_context(None)
        if not ctx:
            raise MemoryError("iw_create_context failed")
        self._setup(ctx, job)
        h = d.iwamd_job_create(ctx)
        if not h:
            buf = ctypes.create_string_buffer(200)
            msg = d.iw_get_errormsg(ctx, buf, 200).decode(errors="replace")
            d.iw_destroy_context(ctx)
            raise RuntimeError("iwamd_job_create failed: " + msg)
        return DeviceJob(self, ctx, h)

    def _run(self, ctx, job, executor):
        d = self.dll
        self._setup(ctx, job)
        if executor is None:
            ok = d.iw_process_image(ctx)
        else:
            plan = ctypes.create_string_buffer(d.iwamd_plan_size())
            ok = d.iwamd_prepare(ctx, plan)
            if ok:
                out_bpr, out_h = plan_output_size(plan)
                img2 = ctypes.create_string_buffer(out_bpr * out_h + 1)
                ok = executor(plan, img2) and d.iwamd_finish(ctx, img2)
        return self._collect(ctx, ok)

    def _setup(self, ctx, job):
        d = self.dll
        if "image" in job:
            im = job["image"]
            pix = np.ascontiguousarray(im["pixels"], dtype=np.uint8).reshape(-1)
            bpr = im["bpr"]
            mem = d.iw_malloc_large(ctx, bpr, im["height"])
            if not mem:
                raise MemoryError("iw_malloc_large failed")
            ctypes.memmove(mem, pix.ctypes.data, bpr * im["height"])
            img = Image()
            img.imgtype = im["imgtype"]
            img.bit_depth = im["bit_depth"]
            img.sampletype = im.get("sampletype", 0)
            img.width = im["width"]
            img.height = im["height"]
            img.pixels = mem
            img.bpr = bpr
            d.iw_set_input_image(ctx, ctypes.byref(img))
        if "crop" in job:
            d.iw_set_input_crop(ctx, *job["crop"])
        if "reorient" in job:
            d.iw_reorient_image(ctx, job["reorient"])
        if "input_cs" in job:
            cs = CsDescr(job["input_cs"][0], 0, job["input_cs"][1])
            d.iw_set_input_colorspace(ctx, ctypes.byref(cs))
        if "output_cs" in job:
            cs = CsDescr(job["output_cs"][0], 0, job["output_cs"][1])
            d.iw_set_output_colorspace(ctx, ctypes.byref(cs))
        if "profile" in job or "image" in job:
            d.iw_set_output_profile(ctx, job.get("profile", PROFILE_PNG))
        if "depth" in job:
            d.iw_set_output_depth(ctx, job["depth"])
        for dim, (fam, blur, p1, p2) in job.get("resize", {}).items():
            d.iw_set_resize_alg(ctx, dim, RESIZE[fam] if isinstance(fam, str) else fam, blur, p1, p2)
        for ct, fam, sub in job.get("dither", []):
            d.iw_set_dither_type(ctx, ct, fam, sub)
        for ct, n in job.get("color_count", []):
            d.iw_set_color_count(ctx, ct, n)
        for ch, n in job.get("input_maxcc", []):
            d.iw_set_input_max_color_code(ctx, ch, n)
        for ct, n in job.get("output_maxcc", []):
            d.iw_set_output_max_color_code(ctx, ct, n)
        if "random_seed" in job:
            d.iw_set_random_seed(ctx, 0, job["random_seed"])
        for code, v in job.get("values", []):
            d.iw_set_value(ctx, code, v)
        for code, v in job.get("values_dbl", []):
            d.iw_set_value_dbl(ctx, code, v)
        for opt, n in job.get("allow_opt", []):
            d.iw_set_allow_opt(ctx, opt, n)
        if "bkgd" in job:
            d.iw_set_apply_bkgd_2(ctx, ctypes.byref(Color((ctypes.c_double * 4)(*job["bkgd"]))))
        if "checker" in job:
            size, c2 = job["checker"]
            d.iw_set_bkgd_checkerboard_2(ctx, size, ctypes.byref(Color((ctypes.c_double * 4)(*c2))))
        if "checker_origin" in job:
            d.iw_set_bkgd_checkerboard_origin(ctx, *job["checker_origin"])
        if "max_malloc" in job:
            d.iw_set_max_malloc(ctx, job["max_malloc"])
        if self.has_ext and "precision" in job:
            d.iwamd_set_precision(ctx, job["precision"])
        for name, val in job.get("options", []):
            d.iw_set_option(ctx, name.encode(), val.encode())
        if "canvas" in job:
            d.iw_set_output_canvas_size(ctx, job["canvas"][0], job["canvas"][1])

    def run_file(self, data, read, write, job=None, executor=None):
        """A file through the library the way the reference's CLI drives it:
        read `data` with the codec function named `read` (e.g.
        "iw_read_pnm_file", or ("iw_read_file_by_fmt", fmt)), apply `job`'s
        settings (no "image" key; its
        "profile" defaults to the one iw_get_profile_by_fmt gives the output
        format), process (iw_process_image, or `executor` as in run()), and
        write with the codec function named `write`.  `write` is (name, fmt):
        the format is set as the output format first, and passed on to a
        *_by_fmt function.  Returns (ok, bytes, error message)."""
        d = self.dll
        job = dict(job or {})
        ctx = d.iw_create_context(None)
        if not ctx:
            raise MemoryError("iw_create_context failed")
        try:
            src, dst = MemFile(data), MemFile()
            rname, rfmt = read if isinstance(read, tuple) else (read, None)
            rfn = getattr(d, rname)
            rfn.restype = ctypes.c_int
            rfn.argtypes = [ctypes.c_void_p, ctypes.POINTER(IODescr)] + ([] if rfmt is None else [ctypes.c_int])
            wname, fmt = write
            ok = bool(rfn(ctx, ctypes.byref(src.io)) if rfmt is None else rfn(ctx, ctypes.byref(src.io), rfmt))
            if ok:
                if "profile" not in job:
                    d.iw_get_profile_by_fmt.argtypes, d.iw_get_profile_by_fmt.restype = [ctypes.c_int], ctypes.c_uint
                    job["profile"] = d.iw_get_profile_by_fmt(fmt)
                self._setup(ctx, job)
                d.iw_set_value(ctx, 52, fmt)  # IW_VAL_OUTPUT_FORMAT
                if executor is None:
                    ok = bool(d.iw_process_image(ctx))
                else:
                    plan = ctypes.create_string_buffer(d.iwamd_plan_size())
                    ok = bool(d.iwamd_prepare(ctx, plan))
                    if ok:
                        out_bpr, out_h = plan_output_size(plan)
                        img2 = ctypes.create_string_buffer(out_bpr * out_h + 1)
                        ok = bool(executor(plan, img2)) and bool(d.iwamd_finish(ctx, img2))
            if ok:
                wfn = getattr(d, wname)
                wfn.restype = ctypes.c_int
                by_fmt = wname.endswith("_by_fmt")
                wfn.argtypes = [ctypes.c_void_p, ctypes.POINTER(IODescr)] + ([ctypes.c_int] if by_fmt else [])
                ok = bool(wfn(ctx, ctypes.byref(dst.io), fmt) if by_fmt else wfn(ctx, ctypes.byref(dst.io)))
            err = None
            if d.iw_get_errorflag(ctx):
                buf = ctypes.create_string_buffer(200)
                err = d.iw_get_errormsg(ctx, buf, 200).decode(errors="replace")
            return ok and not err, bytes(dst.buf), err
        finally:
            d.iw_destroy_context(ctx)

    def _collect(self, ctx, ok):
        d = self.dll
        res = {"ok": bool(ok) and not d.iw_get_errorflag(ctx), "error": None}
        if d.iw_get_errorflag(ctx):
            buf = ctypes.create_string_buffer(200)
            res["error"] = d.iw_get_errormsg(ctx, buf, 200).decode(errors="replace")
        if not ok:
            return res
        out = Image()
        d.iw_get_output_image(ctx, ctypes.byref(out))
        n = out.bpr * out.height
        res.update(width=out.width, height=out.height, imgtype=out.imgtype, bit_depth=out.bit_depth,
                   sampletype=out.sampletype, bpr=out.bpr,
                   pixels=np.ctypeslib.as_array(ctypes.cast(out.pixels, ctypes.POINTER(ctypes.c_uint8)),
                                                (n,)).copy() if n else np.zeros(0, np.uint8),
                   has_colorkey=out.has_colorkey_trns, colorkey=tuple(out.colorkey),
                   reduced_maxcolors=out.reduced_maxcolors, maxcolorcode=tuple(out.maxcolorcode))
        pal = d.iw_get_output_palette(ctx)
        res["palette"] = ([(e.r, e.g, e.b, e.a) for e in pal.contents.entry[:pal.contents.num_entries]]
                          if pal else None)
        if self.has_ext:
            t = (ctypes.c_double * 8)()
            k = d.iwamd_last_timings(ctx, t, 8)
            res["timings_ms"] = list(t[:k])
        return res


class Batch:
    """Many jobs in one device block and one hipGraph (iwamd_batch_*).
    add() prepares a job; upload() places them all; run() replays the graph;
    finish() returns the results in add order."""

    def __init__(self, lib):
        self.lib = lib
        self.h = lib.dll.iwamd_batch_create()
        if not self.h:
            raise RuntimeError("iwamd_batch_create failed")
        self.ctxs = []
        self.failed = {}

    def add(self, job):
        d = self.lib.dll
        ctx = d.iw_create_context(None)
        if not ctx:
            raise MemoryError("iw_create_context failed")
        self.lib._setup(ctx, job)
        idx = d.iwamd_batch_add(self.h, ctx)
        if idx < 0:
            self.failed[len(self.ctxs)] = True
        self.ctxs.append(ctx)
        return idx

    def upload(self):
        if not self.lib.dll.iwamd_batch_upload(self.h):
            raise RuntimeError("iwamd_batch_upload failed")

    def run(self):
        if not self.lib.dll.iwamd_batch_run(self.h):
            raise RuntimeError("iwamd_batch_run failed")

    def sync(self):
        if not self.lib.dll.iwamd_batch_sync(self.h):
            raise RuntimeError("iwamd_batch_sync failed")

    def kernel_ms(self):
        t = ctypes.c_double()
        return t.value if self.lib.dll.iwamd_batch_kernel_ms(self.h, ctypes.byref(t), 1) else None

    def sizes(self):
        px, nb, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self.lib.dll.iwamd_batch_sizes(self.h, ctypes.byref(px), ctypes.byref(nb), ctypes.byref(n))
        return px.value, nb.value, n.value

    def flops(self):
        """Useful resampling FLOPs of the uploaded batch (multiply and add
        per listed tap, both passes, all channels)."""
        return self.lib.dll.iwamd_batch_flops(self.h)

    def finish(self):
        ok = self.lib.dll.iwamd_batch_finish(self.h)
        return [self.lib._collect(ctx, ok and i not in self.failed) for i, ctx in enumerate(self.ctxs)]

    def close(self):
        if self.h:
            self.lib.dll.iwamd_batch_destroy(self.h)
            self.h = None
        for ctx in self.ctxs:
            self.lib.dll.iw_destroy_context(ctx)
        self.ctxs = []


class Server:
    """The request loop (iwamd_server_*): submit() queues a job and returns at
    once; the library's worker admits jobs by HBM budget and runs them in
    overlapping batches.  results() returns every job's result in submit
    order, with its submit/start/done times (ms)."""

    def __init__(self, lib, hbm_budget=0, streams=4, timeout_s=120.0):
        self.lib = lib
        self.h = lib.dll.iwamd_server_create(hbm_budget, streams)
        if not self.h:
            raise RuntimeError("iwamd_server_create failed")
        lib.dll.iwamd_server_set_timeout(self.h, timeout_s * 1e3)
        self.ctxs, self.ids = [], []

    def submit(self, job):
        d = self.lib.dll
        ctx = d.iw_create_context(None)
        if not ctx:
            raise MemoryError("iw_create_context failed")
        self.lib._setup(ctx, job)
        rid = d.iwamd_server_submit(self.h, ctx)
        self.ctxs.append(ctx)
        self.ids.append(rid)
        return rid

    def drain(self):
        return bool(self.lib.dll.iwamd_server_drain(self.h))

    def error(self):
        e = self.lib.dll.iwamd_server_error(self.h)
        return e.decode(errors="replace") if e else None

    def stats(self):
        held, peak, n = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_int()
        self.lib.dll.iwamd_server_stats(self.h, ctypes.byref(held), ctypes.byref(peak), ctypes.byref(n))
        return {"bytes_held": held.value, "max_bytes_held": peak.value, "flights": n.value}

    def results(self):
        """Every result in submit order; raises if the loop timed out (the
        message names what was pending)."""
        if not self.drain() and self.error():
            raise TimeoutError(self.error())
        out = []
        for ctx, rid in zip(self.ctxs, self.ids):
            ok = rid >= 0 and bool(self.lib.dll.iwamd_server_wait(self.h, rid))
            r = self.lib._collect(ctx, ok)
            t = [ctypes.c_double() for _ in range(3)]
            if rid >= 0 and self.lib.dll.iwamd_server_request_ms(self.h, rid, *[ctypes.byref(v) for v in t]):
                r["times_ms"] = tuple(v.value for v in t)
            out.append(r)
        return out

    def close(self):
        if self.h:
            self.lib.dll.iwamd_server_destroy(self.h)
            self.h = None
        for ctx in self.ctxs:
            self.lib.dll.iw_destroy_context(ctx)
        self.ctxs = []


class DeviceJob:
    """Input and tables resident in HBM; run() launches the kernels again
    over them (asynchronously on the job's stream)."""

    def __init__(self, lib, ctx, handle):
        self.lib, self.ctx, self.h = lib, ctx, handle
        px, nb = ctypes.c_double(), ctypes.c_double()
        lib.dll.iwamd_job_sizes(handle, ctypes.byref(px), ctypes.byref(nb))
        self.out_pixels, self.alg_bytes = px.value, nb.value
        self.flops = lib.dll.iwamd_job_flops(handle)

    def run(self):
        if not self.lib.dll.iwamd_job_run(self.h):
            raise RuntimeError("iwamd_job_run failed")

    def sync(self):
        if not self.lib.dll.iwamd_job_sync(self.h):
            raise RuntimeError("iwamd_job_sync failed")

    def kernel_ms(self):
        t = (ctypes.c_double * 3)()
        n = self.lib.dll.iwamd_job_kernel_ms(self.h, t, 3)
        return list(t[:n])

    def finish(self):
        """Copy the output back, complete the context and return run()'s
        result dict; releases the job."""
        ok = self.lib.dll.iwamd_job_finish(self.h)
        res = self.lib._collect(self.ctx, ok)
        self.close()
        return res

    def close(self):
        if self.h:
            self.lib.dll.iwamd_job_destroy(self.h)
            self.h = None
        if self.ctx:
            self.lib.dll.iw_destroy_context(self.ctx)
            self.ctx = None


def plan_output_size(plan):
    """(out_bpr, out_height) from an iwamd_plan buffer: fields located by the
    plan's C layout, which this parses from the header-defined order."""
    return _PlanLayout.get().out_size(plan)


class _PlanLayout:
    _inst = None

    @classmethod
    def get(cls):
        if cls._inst is None:
            cls._inst = cls()
        return cls._inst

    def __init__(self):
        # Mirror of struct iwamd_plan (include/iwamd.h); only needed to find
        # out_bpr and out_height, but laid out in full so the offsets are
        # the compiler's.
        c = ctypes

        class Resize(c.Structure):
            _fields_ = [("family", c.c_int), ("edge_policy", c.c_int), ("num_in", c.c_int),
                        ("num_out", c.c_int), ("param1", c.c_double), ("param2", c.c_double),
                        ("blur_factor", c.c_double), ("out_true_size", c.c_double), ("translate", c.c_double),
                        ("use_offset", c.c_int), ("channel_offset", c.c_double * 3)]

        class IChan(c.Structure):
            _fields_ = [("channeltype", c.c_int), ("input_channel", c.c_int), ("cvt_to_grayscale", c.c_int),
                        ("output_channel", c.c_int), ("need_unassoc_alpha", c.c_int), ("early_bkgd", c.c_int),
                        ("bkgd_color_lin", c.c_double), ("edge_sample_value", c.c_double * 2),
                        ("in_cs", CsDescr)]

        class OChan(c.Structure):
            _fields_ = [("channeltype", c.c_int), ("ditherfamily", c.c_int), ("dithersubtype", c.c_int),
                        ("color_count", c.c_int), ("maxcolorcode_int", c.c_int), ("maxcolorcode_dbl", c.c_double),
                        ("use_nc_table", c.c_int), ("random_seed", c.c_int), ("bkgd1_color_lin", c.c_double),
                        ("bkgd2_color_lin", c.c_double), ("out_cs", CsDescr)]

        class Plan(c.Structure):
            _fields_ = [("in_pixels", c.c_void_p), ("in_bpr", c.c_size_t), ("in_width", c.c_int),
                        ("in_height", c.c_int), ("in_bit_depth", c.c_int), ("in_sampletype", c.c_int),
                        ("in_imgtype", c.c_int), ("in_nch_physical", c.c_int), ("orient_transform", c.c_uint),
                        ("input_start_x", c.c_int), ("input_start_y", c.c_int), ("input_w", c.c_int),
                        ("input_h", c.c_int), ("in_alpha_index", c.c_int), ("in_has_alpha", c.c_int),
                        ("in_maxcolorcode", c.c_double), ("in_ch_maxcolorcode", c.c_double * 4),
                        ("in_ch_slow", c.c_int * 4), ("grayscale_formula", c.c_int),
                        ("grayscale_weight", c.c_double * 3), ("inter_nch", c.c_int),
                        ("inter_alpha_index", c.c_int), ("inter_width", c.c_int), ("inter_height", c.c_int),
                        ("ich", IChan * 4), ("rs", Resize * 2), ("intclamp", c.c_int), ("out_width", c.c_int),
                        ("out_height", c.c_int), ("out_bit_depth", c.c_int), ("out_sampletype", c.c_int),
                        ("out_imgtype", c.c_int), ("out_nch", c.c_int), ("out_bpr", c.c_size_t),
                        ("och", OChan * 4), ("apply_bkgd", c.c_int), ("bkgd_strategy_late", c.c_int),
                        ("bkgd_checkerboard", c.c_int), ("bkgd_check_size", c.c_int),
                        ("bkgd_check_origin", c.c_int * 2), ("bkgd_has_transparency", c.c_int),
                        ("bkgd1alpha", c.c_double), ("bkgd2alpha", c.c_double), ("precision", c.c_int)]

        self.Plan = Plan
        self.OChan = OChan

    def out_size(self, buf):
        p = self.Plan.from_buffer(buf)
        return p.out_bpr, p.out_height


def code_error(plain, fast):
    """Difference of two 8- or 16-bit results of the same job, in output codes
    (LSBs of the output depth) and in linear light (sRGB codes decoded): max,
    mean, share of samples that differ.  Colour samples of pixels whose alpha is 0 in either result are
    left out: the output optimiser blackens fully transparent pixels
    (src/imagew-opt.c:1147-1200), so a one-level alpha change to 0 clears
    the colour, and such colour carries no visible value."""
    if not (plain["ok"] and fast["ok"]):
        return None
    if (plain["imgtype"], plain["bpr"], plain["height"], plain.get("bit_depth")) != (
            fast["imgtype"], fast["bpr"], fast["height"], fast.get("bit_depth")):
        return {"comparable": False}  # the output optimiser chose different formats
    nch = NCH[plain["imgtype"]]
    depth = plain.get("bit_depth", 8)
    if depth == 16:  # big-endian samples
        a = np.frombuffer(plain["pixels"].tobytes(), dtype=">u2").astype(np.int64).reshape(-1, nch)
        b = np.frombuffer(fast["pixels"].tobytes(), dtype=">u2").astype(np.int64).reshape(-1, nch)
    else:
        a = plain["pixels"].astype(np.int64).reshape(-1, nch)
        b = fast["pixels"].astype(np.int64).reshape(-1, nch)
    maxc = 65535.0 if depth == 16 else 255.0
    d = np.abs(a - b)
    has_alpha = plain["imgtype"] in (IMGTYPE["RGBA"], IMGTYPE["GRAYA"])
    masked = 0
    clear = None
    if has_alpha:
        clear = (a[:, -1] == 0) | (b[:, -1] == 0)
        masked = int(clear.sum())
        d[clear, :-1] = 0
    lin = lambda c: np.where(c / maxc <= 0.04045, c / maxc / 12.92, ((c / maxc + 0.055) / 1.055) ** 2.4)
    dl = np.abs(lin(a.astype(np.float64)) - lin(b.astype(np.float64)))
    if has_alpha:
        dl[clear, :-1] = 0.0
    return {"max_lsb": int(d.max()) if d.size else 0, "mean_lsb": round(float(d.mean()), 5) if d.size else 0.0,
            "frac_differ": round(float((d > 0).mean()), 5) if d.size else 0.0,
            "max_linear": round(float(dl.max()), 6) if dl.size else 0.0,
            "mean_linear": round(float(dl.mean()), 7) if dl.size else 0.0,
            "transparent_pixels_left_out": masked}


def image_job(pixels, width, height, imgtype, bit_depth, **settings):
    """Convenience: a job dict for an interleaved image buffer."""
    nch = NCH[imgtype]
    bpr = (width * bit_depth * nch + 7) // 8
    job = {"image": {"pixels": pixels, "width": width, "height": height, "imgtype": imgtype,
                     "bit_depth": bit_depth, "bpr": bpr}}
    job.update(settings)
    return job


def resize_both(name, blur=1.0):
    fam, p1, p2 = FILTERS[name]
    return {0: (fam, blur, p1, p2), 1: (fam, blur, p1, p2)}
