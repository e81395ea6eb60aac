"""The C ABI library loads and exports every function include/*.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import glob
import os
import re

from barney_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(bnamd_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_api():
    names = declared_functions()
    assert {"bnamd_init", "bnamd_world_build", "bnamd_trace", "bnamd_render_surface",
            "bnamd_volume_create", "bnamd_volume_render"} <= names


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert declared_functions() <= set(capi.SIGNATURES), sorted(declared_functions() - set(capi.SIGNATURES))


def test_errors_without_device_are_reported():
    lib = capi.load()
    # NULL arguments are rejected with a message before any GPU work
    assert lib.bnamd_trace(None, None, None, 1, None) == -1
    assert b"NULL" in lib.bnamd_last_error()
    assert lib.bnamd_world_info_get(None, None) == -1
    assert lib.bnamd_volume_render(None, None, None, None, None, None) == -1
