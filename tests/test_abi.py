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
