"""The C ABI library loads and exports every symbol include/emo_api.h
declares (no compute calls: these run without a GPU)."""
import ctypes
import os
import subprocess
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


def test_struct_layouts_match_header(tmp_path):
    # the C compiler's sizes and offsets of the header's structs against the
    # ctypes mirrors
    exe = str(tmp_path / "layout")
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "layout_check.c")
    subprocess.run(["gcc", "-std=c99", src, "-o", exe], check=True)
    out = subprocess.run([exe], check=True, stdout=subprocess.PIPE, text=True).stdout.split()
    vals = {out[i]: int(out[i + 1]) for i in range(0, len(out), 2)}
    assert vals["emo_config"] == ctypes.sizeof(_lib.EmoConfig)
    assert vals["max_particles"] == _lib.EmoConfig.max_particles.offset
    assert vals["neutralize"] == _lib.EmoConfig.neutralize.offset
    assert vals["waist"] == _lib.EmoConfig.waist.offset
    assert vals["emo_loader"] == ctypes.sizeof(_lib.EmoLoader)
    assert vals["drift"] == _lib.EmoLoader.drift.offset
    assert vals["seed"] == _lib.EmoLoader.seed.offset
