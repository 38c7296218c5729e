"""The C ABI: the built library exports every entry point include/dgswe.h
declares, the Python mirror binds them all, and without a GPU the product path
fails loudly (no CPU fallback)."""
import ctypes
import os
import re
import subprocess

import pytest

from dgcom_amd import cases, swe2d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "dgswe.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|uint64_t|const char\*)\s+(dg_\w+)\s*\(", text, re.M)))


@pytest.mark.parametrize("fma", [False, True], ids=["exact", "fma"])
def test_library_exports_every_declared_symbol(fma):
    path = swe2d.library_path(fma)
    assert os.path.exists(path), "build the library first (python -m dgcom_amd.build)"
    names = declared()
    assert len(names) >= 20
    L = ctypes.CDLL(path)
    for n in names:
        assert hasattr(L, n), n
    nm = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (dg_\w+)", nm))
    assert set(names) <= exported
    assert set(swe2d.EXPORTS) == set(names)


def test_no_cpu_fallback_without_gpu():
    import torch  # noqa: F401  (only to ask whether a device is present)

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    c = cases.drying_beach(nx=4, ny=2)
    with pytest.raises(RuntimeError, match="dg_create"):
        swe2d.SWEPhysics2D.from_case(c)


def test_missing_library_raises(tmp_path, monkeypatch):
    monkeypatch.setattr(swe2d, "library_path", lambda fma=False: str(tmp_path / "absent.so"))
    monkeypatch.setattr(swe2d, "_libs", {})
    with pytest.raises(RuntimeError, match="missing"):
        swe2d.load()


def test_host_entry_points_refuse_null_arguments():
    """Entry points that touch no device refuse a null tape by name (no GPU needed)."""
    L = swe2d.load()
    v, w = ctypes.c_int64(), ctypes.c_int64()
    assert L.dg_tape_records_used(None, ctypes.byref(v)) != 0
    assert b"dg_tape_records_used" in L.dg_last_error()
    assert L.dg_tape_stats(None, ctypes.byref(v), ctypes.byref(w)) != 0
    assert b"dg_tape_stats" in L.dg_last_error()
    assert L.dg_revolve_count(40, 5, ctypes.byref(v)) == 0 and v.value == swe2d.revolve_count(40, 5) > 40
