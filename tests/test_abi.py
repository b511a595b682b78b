"""The C ABI library loads and exports every symbol its headers declare
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
from opensubdiv_amd import _lib


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(osd\w+)\s*\(", text)))


def test_hip_library_exports_every_declared_symbol():
    names = declared("osdhip.h")
    assert "osdhip_eval_stencils" in names and "osdhip_eval_patches" in names
    lib = ctypes.CDLL(_lib.HIP_LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the whole ABI
    assert set(names) <= set(_lib.hip_symbols())


def test_version_string_without_device():
    lib = _lib.load_hip()
    assert lib.osdhip_version().decode().startswith("osdhip")


def test_library_is_gfx950_code_object():
    data = open(_lib.HIP_LIB, "rb").read()
    assert b"gfx950" in data


def test_far_library_exports_every_declared_symbol():
    if not os.path.exists(os.path.join(ROOT, "include", "osdfar.h")):
        pytest.skip("no table compiler header")
    names = declared("osdfar.h")
    lib = ctypes.CDLL(_lib.FAR_LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
