"""The C ABI: the built library exports every function the header declares."""
import ctypes
import os
import re

import partmc_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "partmc_coag.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pmc_gpu_\w+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "pmc_gpu_mc_coag" in names
    assert set(names) == set(partmc_amd.ABI), set(names) ^ set(partmc_amd.ABI)


def test_library_loads_and_exports_every_symbol():
    assert os.path.exists(partmc_amd.LIB_PATH), "build first: python -m partmc_amd.build"
    lib = ctypes.CDLL(partmc_amd.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_last_error_is_a_string():
    lib = partmc_amd.load()
    assert isinstance(lib.pmc_gpu_last_error(), bytes)
