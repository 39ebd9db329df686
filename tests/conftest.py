import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libiwref.so")
REF_CODECS_LIB = os.path.join(ROOT, "oracle", "_ref", "libiwref_codecs.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "libiwo.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def prod():
    from imageworsener_amd import api
    return api.Library()


@pytest.fixture(scope="session")
def ref():
    from imageworsener_amd import api
    if not os.path.exists(REF_LIB):
        pytest.skip("reference core not built (oracle/_ref needs /root/reference at build time)")
    return api.Library(REF_LIB, extensions=False)


@pytest.fixture(scope="session")
def ref_codecs():
    """The reference core with its own codecs (PNM/PAM, MIFF, ...)."""
    from imageworsener_amd import api
    if not os.path.exists(REF_CODECS_LIB):
        pytest.skip("reference codecs not built (oracle/_ref needs /root/reference at build time)")
    return api.Library(REF_CODECS_LIB, extensions=False)


@pytest.fixture(scope="session")
def oracle():
    if not os.path.exists(ORACLE_LIB):
        pytest.skip("oracle not built")
    lib = ctypes.CDLL(ORACLE_LIB, mode=os.RTLD_LOCAL | os.RTLD_NOW)
    lib.iwo_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.iwo_run.restype = ctypes.c_int
    lib.iwo_prng_nth.argtypes = [ctypes.c_int, ctypes.c_uint64]
    lib.iwo_prng_nth.restype = ctypes.c_uint32
    lib.iwo_filter.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               ctypes.c_int, ctypes.c_int]
    lib.iwo_filter.restype = ctypes.c_double
    lib.iwo_to_linear.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double]
    lib.iwo_to_linear.restype = ctypes.c_double
    lib.iwo_from_linear.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double]
    lib.iwo_from_linear.restype = ctypes.c_double
    return lib


@pytest.fixture(scope="session")
def oracle_exec(oracle):
    """Executor for api.Library.run: the CPU oracle on the library's plan."""
    def run(plan, img2):
        return bool(oracle.iwo_run(ctypes.addressof(plan), ctypes.addressof(img2)))
    return run
