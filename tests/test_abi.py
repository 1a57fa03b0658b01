"""The C ABI library loads and exports every symbol include/harmgpu.h declares
(no device calls)."""
import ctypes
import os
import re

import pytest

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_symbols():
    txt = open(os.path.join(REPO, "include", "harmgpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(hg_\w+)\s*\(", txt, re.M)))


def test_library_exports_header():
    from harm_finalwham_amd import harmgpu
    L = harmgpu.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert sorted(harmgpu.EXPORTS) == names


def test_context_rejects_bad_config_without_gpu_work():
    from harm_finalwham_amd import harmgpu
    with pytest.raises(harmgpu.HarmError):
        harmgpu.Grid(0, 8, 1)


def test_context_rejects_planes_past_32_bit_offsets_without_gpu_work():
    """The reconstruction kernels address a primitive plane with 32-bit byte
    offsets: a grid whose plane (ghosts included) holds 2^29 zones or more is
    refused before any device allocation."""
    from harm_finalwham_amd import harmgpu
    with pytest.raises(harmgpu.HarmError, match="2\\^29"):
        harmgpu.Grid(2048, 2048, 160, nb=6)  # 2060 x 2060 x 172 zones with ghosts
    with pytest.raises(harmgpu.HarmError, match="2\\^29"):
        harmgpu.Grid(2036, 2036, 116, nb=6)  # exactly 2^29 zones with ghosts
