"""CPU checks of the C ABI: the library loads and exports every symbol include/readdy_amd.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "readdy_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rdy_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_engine_abi():
    syms = declared_symbols()
    for must in ["rdy_engine_create", "rdy_calculate_forces", "rdy_integrate", "rdy_react", "rdy_run",
                 "rdy_neighbor_pairs", "rdy_reaction_candidates"]:
        assert must in syms


def test_library_exports_every_declared_symbol():
    from readdy_amd import _lib
    lib = _lib.load()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert set(declared_symbols()) <= set(_lib.SIGNATURES), set(declared_symbols()) - set(_lib.SIGNATURES)


def test_create_rejects_invalid_context_without_gpu_work():
    from readdy_amd import _lib as L
    lib = L.load()
    ctx = L.ContextDesc()
    ctx.n_types = 0
    h = ctypes.c_void_p()
    assert lib.rdy_engine_create(ctypes.byref(ctx), None, ctypes.byref(h)) == -1
    assert b"n_types" in lib.rdy_last_error()


def test_python_binding_declares_every_abi_entry_point():
    """readdy_amd/_lib.py gives every entry point of include/readdy_amd.h its ctypes signature."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    header = open(os.path.join(root, "include", "readdy_amd.h")).read()
    binding = open(os.path.join(root, "readdy_amd", "_lib.py")).read()
    declared = set(re.findall(r"\b(rdy_[a-z0-9_]+)\s*\(", header))
    bound = set(re.findall(r'"(rdy_[a-z0-9_]+)"', binding))
    assert declared and declared == bound, (sorted(declared - bound), sorted(bound - declared))
