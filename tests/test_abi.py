"""C-ABI checks that need no GPU: the library loads, exports every function
include/*.h declares, and its host-only parts (generators, argument checks)
behave."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s+\*?(sd_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_functions():
    names = declared_functions()
    assert "sd_ctx_step" in names and "sd_radix_sort_pairs_u32" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol(sdlib):
    lib = sdlib.load()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_covers_header(sdlib):
    assert declared_functions() <= set(sdlib.exported_symbols())


def test_version_and_error_strings(sdlib):
    lib = sdlib.load()
    assert b"gfx950" in lib.sd_version()
    assert isinstance(lib.sd_last_error(), bytes)


def test_argument_errors_are_status_codes(sdlib):
    lib = sdlib.load()
    g = sdlib.Grid.make((0, 0, 0), 0.0, (1, 1, 1))  # zero cell size
    st = lib.sd_hash_keys_u32(None, None, None, 4, C.byref(g), None, None, None)
    assert st == sdlib.SD_ERR_ARG
    assert b"cell" in lib.sd_last_error()
    nb = C.c_size_t()
    assert lib.sd_radix_sort_temp_bytes(-1, C.byref(nb)) == sdlib.SD_ERR_ARG
    assert lib.sd_radix_sort_temp_bytes(1000, C.byref(nb)) == 0 and nb.value > 0


def test_generator_deterministic_and_overlap_free(sdlib, oracle):
    from stardust_amd import workloads as W
    w = W.dem_c1(n=3000)
    a = w.generate()
    b = w.generate()
    for k in a:
        assert np.array_equal(a[k], b[k])
    pos4 = np.stack([a["x"], a["y"], a["z"], a["r"]], 1)
    assert len(oracle.brute_pairs(pos4, 0.0)) == 0
    assert np.all(a["x"] >= a["r"]) and np.all(a["x"] <= 1 - a["r"])


def test_generator_lognormal_radii(sdlib):
    from stardust_amd import workloads as W
    w = W.dem_c2(n=20000)
    s = w.generate()
    r = s["r"]
    assert r.min() >= 0.002 and r.max() <= 0.012
    med = float(np.median(r))
    assert abs(med - 0.005) / 0.005 < 0.02
    sig = float(np.std(np.log(r[(r > 0.0021) & (r < 0.0119)])))
    assert abs(sig - 0.25) < 0.02


def test_generator_reports_impossible_packing(sdlib):
    x = np.empty(10, np.float32)
    lo = np.zeros(3, np.float32)
    hi = np.full(3, 0.05, np.float32)
    lib = sdlib.load()
    st = lib.sd_gen_spheres(1, 10, sdlib.ptr(lo), sdlib.ptr(hi), 0.02, 0.0, 0.02, 0.02, sdlib.ptr(x), sdlib.ptr(x),
                            sdlib.ptr(x), sdlib.ptr(x))
    assert st in (sdlib.SD_ERR_CAPACITY, sdlib.SD_ERR_ARG)
