"""The C ABI library without a GPU: it loads, exports every function include/bxstep.h declares, and
its struct layouts equal the ctypes mirrors in bounce_amd/capi.py (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

from bounce_amd import capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "bxstep.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(bx_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def declared_structs():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"typedef struct (bx_[a-z_]+) \{", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libbxstep.so is not built (run bounce_amd/build.py)")
    return C.CDLL(capi.LIB_PATH)


def test_header_declares_the_step_entry_points():
    names = declared_functions()
    for n in ("bx_context_create", "bx_batch_create", "bx_batch_step", "bx_batch_get_bodies", "bx_batch_sync"):
        assert n in names
    assert len(names) >= 20


def test_library_exports_every_declared_function(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_ctypes_mirrors(lib):
    lib.bx_struct_size.restype = C.c_uint32
    lib.bx_struct_size.argtypes = [C.c_char_p]
    structs = declared_structs()
    assert "bx_body_state" in structs and "bx_contact_state" in structs
    for name in structs:
        size = lib.bx_struct_size(name.encode())
        assert size > 0, name
        assert size == C.sizeof(getattr(capi, name)), name
    assert lib.bx_struct_size(b"no_such_struct") == 0


def test_version_and_errors_without_device(lib):
    lib.bx_version.restype = C.c_char_p
    assert lib.bx_version().startswith(b"bxstep")
    lib.bx_batch_step.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
    assert lib.bx_batch_step(None, None, 1) == capi.BX_E_INVALID
    lib.bx_last_error.restype = C.c_char_p
    assert lib.bx_last_error() != b""


def test_kernel_names(lib):
    lib.bx_kernel_name.restype = C.c_char_p
    lib.bx_kernel_name.argtypes = [C.c_uint32]
    names = [lib.bx_kernel_name(i).decode() for i in range(11)]
    assert "k_solve_velocity" in names and "k_update_contacts" in names
    assert lib.bx_kernel_name(99) == b""
