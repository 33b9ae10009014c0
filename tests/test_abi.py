"""The C ABI library loads and exports every symbol include/emo_api.h
declares (no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

from apc_524_project_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "emo_api.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(emo_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_header_matches_binding_list():
    assert set(declared_functions()) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libemopic.so not built (run python -m apc_524_project_amd.build)")
    so = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_functions():
        assert hasattr(so, name), name


def test_version_string_without_gpu():
    assert b"gfx950" in _lib.lib().emo_version()


def test_struct_layouts_match_header():
    # offsets of the last members pin the ctypes mirrors to the C layout
    assert _lib.EmoConfig.max_particles.offset == ctypes.sizeof(_lib.EmoConfig) - 8
    assert ctypes.sizeof(_lib.EmoLoader) % 8 == 0
