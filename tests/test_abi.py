"""The C-ABI library loads without a GPU and exports every function the
public headers declare (include/imagew.h, include/iwamd.h)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("imagew.h", "iwamd.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"IW_EXPORT\([^)]*\)\s+(\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_reference_api():
    names = declared_functions()
    assert len(names) >= 113 + 8
    for must in ("iw_create_context", "iw_process_image", "iw_set_resize_alg", "iw_set_output_depth",
                 "iw_set_dither_type", "iw_set_color_count", "iw_set_random_seed", "iw_get_output_image",
                 "iwamd_prepare", "iwamd_finish"):
        assert must in names


def test_library_exports_every_declared_function(prod):
    missing = [n for n in declared_functions() if not hasattr(prod.dll, n)]
    assert not missing, missing


def test_reference_core_exports_match_ours(prod, ref):
    """Every public iw_* symbol of the reference core exists here too."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", ref.path], stdout=subprocess.PIPE, text=True).stdout
    ref_syms = sorted({l.split()[-1] for l in out.splitlines() if " T iw_" in l})
    missing = [n for n in ref_syms if not hasattr(prod.dll, n)]
    assert not missing, missing


def test_process_without_gpu_fails_loudly(prod):
    """No silent CPU fallback: without a HIP device iw_process_image reports
    an error."""
    if prod.dll.iwamd_device_count() > 0:
        pytest.skip("a GPU is present")
    from imageworsener_amd import api
    import numpy as np
    job = api.image_job(np.zeros((4, 4, 3), np.uint8), 4, 4, api.IMGTYPE["RGB"], 8, canvas=(8, 8))
    r = prod.run(job)
    assert not r["ok"] and "HIP" in r["error"]


def test_reprocess_guard_and_errors(prod, ref):
    """Error behaviour matches the reference: missing profile, oversize
    dimensions, second use of a context."""
    from imageworsener_amd import api
    import numpy as np
    img = np.zeros((4, 4, 3), np.uint8)
    for lib in (prod, ref):
        job = api.image_job(img, 4, 4, api.IMGTYPE["RGB"], 8, canvas=(8, 8), profile=0)
        r = lib.run(job)
        assert not r["ok"] and r["error"] == "Output profile not set"
        job = api.image_job(img, 4, 4, api.IMGTYPE["RGB"], 8, canvas=(50000, 8))
        r = lib.run(job)
        assert not r["ok"] and r["error"].startswith("Image dimensions too large")
        job = api.image_job(img, 4, 4, api.IMGTYPE["RGB"], 8, canvas=(40000, 40000), max_malloc=1 << 20)
        r = lib.run(job)
        assert not r["ok"] and r["error"] == "Image too large to process"
