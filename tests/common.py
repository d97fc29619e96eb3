"""Shared helpers for the test modules: fixture loading and library handles."""
import functools
import hashlib
import json
import os

import pytest

from ffi import Api, Oracle, PRODUCT_DIR, ORACLE_SO, load_ref

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SEED = 0x584D5353


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def dg(b):
    return hashlib.sha256(bytes(b)).hexdigest()


@functools.lru_cache(None)
def oracle():
    if not os.path.exists(ORACLE_SO):
        pytest.skip("oracle not built (python __graft_entry__.py build)")
    return Oracle()


@functools.lru_cache(None)
def product(core=0):
    """The HIP library; core 0 = plain (libxmss_mi355x.so), 1 = BDS (_fast)."""
    path = os.path.join(PRODUCT_DIR, "libxmss_mi355x_fast.so" if core else "libxmss_mi355x.so")
    assert os.path.exists(path), "library not built: " + path
    api = Api(path, core)
    import ctypes as C
    api.lib.xmss_mi355x_rng_seed.argtypes = [C.c_uint64]
    api.lib.xmss_mi355x_device_ok.restype = C.c_int
    assert api.lib.xmss_mi355x_device_ok() == 1, "no usable GPU for the HIP library"
    return api
